// coupled_pair.hip -- FP32 instantiation of the two-cells-per-thread fused step (coupled_pair.hpp);
// the FP16 / BF16 instantiations live in coupled_half.hip.
#include "coupled_pair.hpp"

namespace fmx {

template int launch_coupled_pair<FMX_F32>(int, const CoupledPtrs<FMX_F32> &, const CoupledParams<float> &, i64, i64,
                                          hipStream_t);

}  // namespace fmx
