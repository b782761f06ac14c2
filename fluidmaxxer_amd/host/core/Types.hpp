#pragma once
// Scalar type of parameters, device scalar fields and populations (Types.hpp:5-6 of the
// reference): double by default, float with -DFMX_DATA_TYPE=float.
#include "fmx_abi.h"

#ifndef FMX_DATA_TYPE
#define FMX_DATA_TYPE double
#endif

using DATA_TYPE = FMX_DATA_TYPE;

namespace fmx {
// C ABI precision code of DATA_TYPE
constexpr int kPrec = sizeof(DATA_TYPE) == 8 ? FMX_F64 : FMX_F32;
static_assert(sizeof(DATA_TYPE) == 4 || sizeof(DATA_TYPE) == 8, "DATA_TYPE must be float or double");
}  // namespace fmx
