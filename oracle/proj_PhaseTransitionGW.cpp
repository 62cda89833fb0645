// ORACLE -- test infrastructure only (see oracle.hpp).
#include "../sledgehamr_amd/projects/PhaseTransitionGW/PhaseTransitionGW.h"
#include "oracle_capi.hpp"
ORC_DEFINE_API(PhaseTransitionGW::PhaseTransitionGW)
