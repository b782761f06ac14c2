// coupled_half.hip -- the fused coupled step for FP16 / BF16 population storage.
//
// The same kernel as coupled.hip (coupled_kernel.hpp), compiled with FMA contraction and using the
// hardware reciprocal (1 ulp) for its divisions: with 16-bit storage the kernel is bound by vector
// ALU issue rather than memory, and these halve the arithmetic's instruction count while changing
// results far below the storage rounding (measured against the FP64 restatement).
#include "coupled_kernel.hpp"
#include "coupled_pair.hpp"

namespace fmx {

template int launch_coupled_step<FMX_F16>(int, const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                          hipStream_t);
template int launch_coupled_step<FMX_BF16>(int, const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64, i64,
                                           hipStream_t);
template int launch_coupled_rc<FMX_F16>(const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                        const RcArgs &, hipStream_t);
template int launch_coupled_rc<FMX_BF16>(const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64, i64,
                                         const RcArgs &, hipStream_t);
template int launch_coupled_split<FMX_F16>(const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                           const SplitArgs &, hipStream_t);
template int launch_coupled_split<FMX_BF16>(const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64, i64,
                                            const SplitArgs &, hipStream_t);
template int launch_coupled_pair<FMX_F16>(int, const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                          hipStream_t);
template int launch_coupled_pair<FMX_BF16>(int, const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64, i64,
                                           hipStream_t);

}  // namespace fmx
