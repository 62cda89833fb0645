// ORACLE -- test infrastructure only (see oracle.hpp).
#include "../sledgehamr_amd/projects/ComplexScalar/ComplexScalar.h"
#include "oracle_capi.hpp"
ORC_DEFINE_API(ComplexScalar::ComplexScalar)
