// diagnostics.hip -- FP64 sums over the coupled state: the concentration, the adsorbed load and the
// NS density (sum of every stored population), with a count of non-finite values, for the mass and
// adsorbed-load diagnostics of a run (the reference has none; its only reduction is the host loop of
// OutletCSV.hpp:29-58).
//
// Two passes in a fixed order, so a sum repeats to the bit: block b of a grid sized by n alone reduces
// the contiguous chunk [b*chunk, (b+1)*chunk) -- FP64 per thread (a cell's 19 populations added in
// FP64 too), then a wave64 butterfly of shuffles, then the block's four wave sums in wave order -- into
// four partials; one block then adds the partials in block order the same way.
#include <cmath>
#include <mutex>

#include "coupled_common.hpp"
#include "diagnostics.hpp"

namespace fmx {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The four sums of a 256-thread block in wave order; valid in thread 0.
__device__ __forceinline__ void block_sums(double (&s)[kMassCols], double (&out)[kMassCols]) {
    __shared__ double red[4][kMassCols];
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
#pragma unroll
    for (int k = 0; k < kMassCols; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) red[wave][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) out[k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

template <int P>
__global__ void __launch_bounds__(256)
    k_mass_partial(const typename Storage<P>::C *__restrict__ phi, const typename Storage<P>::C *__restrict__ q,
                   const typename Storage<P>::S *__restrict__ f, i64 n, typename Storage<P>::C fref, i64 chunk,
                   double *__restrict__ partial) {
    using C = typename Storage<P>::C;
    const i64 begin = i64(blockIdx.x) * chunk;
    const i64 end = begin + chunk < n ? begin + chunk : n;
    double s[kMassCols] = {0.0, 0.0, 0.0, 0.0};
    for (i64 c = begin + threadIdx.x; c < end; c += blockDim.x) {
        if (phi) {
            const double v = double(phi[c]);
            s[0] += v;
            s[3] += std::isfinite(v) ? 0.0 : 1.0;
        }
        if (q) {
            const double v = double(q[c]);
            s[1] += v;
            s[3] += std::isfinite(v) ? 0.0 : 1.0;
        }
        if (f) {  // a non-finite population makes the cell's sum non-finite
            double r = 0.0;
#pragma unroll
            for (int k = 0; k < 19; ++k) r += double(ld_pop<P>(f, i64(k) * n + c, w19<C>(k), fref));
            s[2] += r;
            s[3] += std::isfinite(r) ? 0.0 : 1.0;
        }
    }
    double out[kMassCols];
    block_sums(s, out);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) partial[i64(blockIdx.x) * kMassCols + k] = out[k];
}

__global__ void __launch_bounds__(256) k_mass_final(const double *__restrict__ partial, int nblocks,
                                                    double *__restrict__ out) {
    double s[kMassCols] = {0.0, 0.0, 0.0, 0.0};
    for (int b = int(threadIdx.x); b < nblocks; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) s[k] += partial[i64(b) * kMassCols + k];
    double tot[kMassCols];
    block_sums(s, tot);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) out[k] = tot[k];
}

// Density sum over the cells whose link word has no solid bit (out[0]) and their number (out[1]),
// for the deviation reference of 16-bit NS storage; the same fixed-order two-pass reduction.
template <int P>
__global__ void __launch_bounds__(256)
    k_fluid_rho_partial(const typename Storage<P>::S *__restrict__ f, const std::uint32_t *__restrict__ links, i64 n,
                        typename Storage<P>::C fref, i64 chunk, double *__restrict__ partial) {
    using C = typename Storage<P>::C;
    const i64 begin = i64(blockIdx.x) * chunk;
    const i64 end = begin + chunk < n ? begin + chunk : n;
    double s[kMassCols] = {0.0, 0.0, 0.0, 0.0};
    for (i64 c = begin + threadIdx.x; c < end; c += blockDim.x) {
        if (links[c] & FMX_LINK_SOLID) continue;
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < 19; ++k) r += double(ld_pop<P>(f, i64(k) * n + c, w19<C>(k), fref));
        s[0] += r;
        s[1] += 1.0;
    }
    double out[kMassCols];
    block_sums(s, out);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kMassCols; ++k) partial[i64(blockIdx.x) * kMassCols + k] = out[k];
}

int mass_blocks(i64 n) {
    const i64 want = (n + kMassCellsPerBlock - 1) / kMassCellsPerBlock;
    return int(want < 1 ? 1 : (want > kMassMaxBlocks ? kMassMaxBlocks : want));
}

template <int P>
int launch_mass(const void *phi, const void *q, const void *f, i64 n, double fref, double *partial, double *out,
                hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const int nblocks = mass_blocks(n);
    const i64 chunk = (n + nblocks - 1) / nblocks;
    hipLaunchKernelGGL(k_mass_partial<P>, dim3(nblocks), dim3(256), 0, s, static_cast<const C *>(phi),
                       static_cast<const C *>(q), static_cast<const S *>(f), n, C(fref), chunk, partial);
    if (int rc = after_launch("k_mass_partial")) return rc;
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, s, partial, nblocks, out);
    return after_launch("k_mass_final");
}

template int launch_mass<FMX_F32>(const void *, const void *, const void *, i64, double, double *, double *, hipStream_t);
template int launch_mass<FMX_F64>(const void *, const void *, const void *, i64, double, double *, double *, hipStream_t);
template int launch_mass<FMX_F16>(const void *, const void *, const void *, i64, double, double *, double *, hipStream_t);
template int launch_mass<FMX_BF16>(const void *, const void *, const void *, i64, double, double *, double *, hipStream_t);

int launch_mass_any(int prec, const void *phi, const void *q, const void *f, i64 n, double fref, double *partial,
                    double *out, hipStream_t s) {
    switch (prec) {
        case FMX_F32: return launch_mass<FMX_F32>(phi, q, f, n, fref, partial, out, s);
        case FMX_F64: return launch_mass<FMX_F64>(phi, q, f, n, fref, partial, out, s);
        case FMX_F16: return launch_mass<FMX_F16>(phi, q, f, n, fref, partial, out, s);
        case FMX_BF16: return launch_mass<FMX_BF16>(phi, q, f, n, fref, partial, out, s);
        default: return set_error(FMX_ERR_ARG, "fmx_reduce_mass: unknown precision %d", prec);
    }
}

template <int P>
int launch_fluid_rho(const void *f, const std::uint32_t *links, i64 n, double fref, double *partial, double *out,
                     hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const int nblocks = mass_blocks(n);
    const i64 chunk = (n + nblocks - 1) / nblocks;
    hipLaunchKernelGGL(k_fluid_rho_partial<P>, dim3(nblocks), dim3(256), 0, s, static_cast<const S *>(f), links, n,
                       C(fref), chunk, partial);
    if (int rc = after_launch("k_fluid_rho_partial")) return rc;
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, s, partial, nblocks, out);
    return after_launch("k_mass_final");
}

int launch_fluid_rho_any(int prec, const void *f, const std::uint32_t *links, i64 n, double fref, double *partial,
                         double *out, hipStream_t s) {
    switch (prec) {
        case FMX_F32: return launch_fluid_rho<FMX_F32>(f, links, n, fref, partial, out, s);
        case FMX_F64: return launch_fluid_rho<FMX_F64>(f, links, n, fref, partial, out, s);
        case FMX_F16: return launch_fluid_rho<FMX_F16>(f, links, n, fref, partial, out, s);
        case FMX_BF16: return launch_fluid_rho<FMX_BF16>(f, links, n, fref, partial, out, s);
        default: return set_error(FMX_ERR_ARG, "fluid density: unknown precision %d", prec);
    }
}

// ---------------------------------------------------------------- no-allocation variant
// The partials of fmx_reduce_mass_static: the largest partial-sum grid (kMassMaxBlocks blocks), 64 KiB
// per GPU, held until the process ends.  In host code `scratch` names a host-side stand-in, not a device
// address: the kernels get the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kMassMaxBlocks * kMassCols];

constexpr int kMaxDevices = 64;
// cells the static scratch serves: its kMassMaxBlocks blocks of at most 2^24 cells each
constexpr i64 kMassStaticMaxCells = i64(kMassMaxBlocks) << 24;

// Device address of `scratch` on the current device: looked up once per device (with that device
// current, on its first call) and kept.
static int scratch_address(double **out) {
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "fmx_reduce_mass_static: hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(FMX_ERR_UNSUPPORTED, "fmx_reduce_mass_static: device %d beyond the %d supported", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                               "fmx_reduce_mass_static: hipGetSymbolAddress(scratch)"))
            return rc;
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return FMX_OK;
}

}  // namespace fmx

using namespace fmx;

extern "C" {

int64_t fmx_reduce_mass_workspace_bytes(int64_t n) {
    if (n <= 0) return -1;
    return int64_t(mass_blocks(n)) * kMassCols * int64_t(sizeof(double));
}

int fmx_reduce_mass(int prec, const void *phi, const void *q, const void *f, int64_t n, double f_ref, double *out,
                    void *workspace, int64_t workspace_bytes, void *stream) {
    FMX_TRY_BEGIN
    if (n <= 0 || !out) return set_error(FMX_ERR_ARG, "fmx_reduce_mass: bad arguments");
    const int64_t need = fmx_reduce_mass_workspace_bytes(n);
    if (!workspace || workspace_bytes < need)
        return set_error(FMX_ERR_ARG, "fmx_reduce_mass: workspace of %lld bytes given, %lld needed "
                                      "(fmx_reduce_mass_workspace_bytes)",
                         (long long)(workspace ? workspace_bytes : 0), (long long)need);
    return launch_mass_any(prec, phi, q, f, n, f_ref, static_cast<double *>(workspace), out, as_stream(stream));
    FMX_TRY_END
}

int fmx_reduce_mass_static(int prec, const void *phi, const void *q, const void *f, int64_t n, double f_ref,
                           double *out, void *stream) {
    FMX_TRY_BEGIN
    if (n <= 0 || !out) return set_error(FMX_ERR_ARG, "fmx_reduce_mass_static: bad arguments");
    if (n > kMassStaticMaxCells)
        return set_error(FMX_ERR_ARG, "fmx_reduce_mass_static: %lld cells exceed the %lld its static scratch serves "
                                      "(use fmx_reduce_mass with a workspace)",
                         (long long)n, (long long)kMassStaticMaxCells);
    if (mass_blocks(n) * kMassCols > kMassMaxBlocks * kMassCols)
        return set_error(FMX_ERR_ARG, "fmx_reduce_mass_static: partial grid larger than the static scratch");
    double *partial = nullptr;
    if (int rc = scratch_address(&partial)) return rc;
    return launch_mass_any(prec, phi, q, f, n, f_ref, partial, out, as_stream(stream));
    FMX_TRY_END
}

}  // extern "C"
