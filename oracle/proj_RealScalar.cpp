// ORACLE -- test infrastructure only (see oracle.hpp).
#include "../sledgehamr_amd/projects/RealScalar/RealScalar.h"
#include "oracle_capi.hpp"
ORC_DEFINE_API(RealScalar::RealScalar)
