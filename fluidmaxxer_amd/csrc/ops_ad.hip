// ops_ad.hip -- D3Q7 advection-diffusion operators, LDF sorption sources, layout
// conversion and FP64 diagnostics reductions.
#include "fmx_internal.hpp"

namespace fmx {

// --------------------------------------------------------------- AD init
// k_ad_init_eq_d3q7 (ADBGKCollision.cu:29-38)
template <class T> __global__ void __launch_bounds__(256) k_ad_init_eq(T *__restrict__ g, i64 n, T phi0) {
    const i64 c = gtid();
    if (c >= n) return;
#pragma unroll
    for (int q = 0; q < 7; ++q) g[q * n + c] = phi0 * w7<T>(q);
}

// ------------------------------------------------------ AD collide+stream
// k_ad_bgk_stream_pull_d3q7_source (ADBGKCollision.cu:41-128)
template <class T>
__global__ void __launch_bounds__(256)
    k_ad_collide(const T *__restrict__ g, T *__restrict__ g_new, T *__restrict__ phi_out, const T *__restrict__ tau,
                 const T *__restrict__ ux, const T *__restrict__ uy, const T *__restrict__ uz,
                 const T *__restrict__ S, fmx_grid gr, T phiMin) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= gr.nx || y >= gr.ny) return;
    const i64 n = i64(gr.nx) * gr.ny * gr.nz;
    const i64 cell = cell_index(x, y, z, gr.nx, gr.ny);
    const T omega = T(1) / tau[cell];
    const T vx = ux[cell], vy = uy[cell], vz = uz[cell];
    T gi[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        int xs = x - Q7::cx[q], ys = y - Q7::cy[q], zs = z - Q7::cz[q];
        bool ok = true;
        if (gr.periodic_x) xs = (xs % gr.nx + gr.nx) % gr.nx; else ok = ok && xs >= 0 && xs < gr.nx;
        if (gr.periodic_y) ys = (ys % gr.ny + gr.ny) % gr.ny; else ok = ok && ys >= 0 && ys < gr.ny;
        if (gr.periodic_z) zs = (zs % gr.nz + gr.nz) % gr.nz; else ok = ok && zs >= 0 && zs < gr.nz;
        gi[q] = ok ? g[q * n + cell_index(xs, ys, zs, gr.nx, gr.ny)] : g[q * n + cell];
    }
    T phi = T(0);
#pragma unroll
    for (int q = 0; q < 7; ++q) phi += gi[q];
    if (phi < phiMin) phi = phiMin;
    const T src = S ? S[cell] : T(0);
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const T eu = T(Q7::cx[q]) * vx + T(Q7::cy[q]) * vy + T(Q7::cz[q]) * vz;
        const T geq = w7<T>(q) * phi * (T(1) + T(4) * eu);
        g_new[q * n + cell] = gi[q] + omega * (geq - gi[q]) + w7<T>(q) * src;
    }
    if (phi_out) phi_out[cell] = phi;
}

// k_ad_dirichlet (ADDirichlet.cu:11-30), k_ad_dirichlet_tanh_pulse_d3q7 (ADDirichletTanhPulse.cu:11-31, launched at :91-96)
template <class T>
__global__ void __launch_bounds__(256) k_ad_bc_dirichlet(T *__restrict__ g, const i64 *__restrict__ cells, i64 count,
                                                         i64 n, T phi) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k];
#pragma unroll
    for (int q = 0; q < 7; ++q) g[q * n + c] = phi * w7<T>(q);
}

// ADDirichlet with the value of the step read on the device: the per-step inlet concentration series
// (ADDirichletTanhPulse.cu:66-100 computes phi_in(s) in double on the host and casts it to T)
template <class T>
__global__ void __launch_bounds__(256) k_ad_bc_dirichlet_series(T *__restrict__ g, const i64 *__restrict__ cells,
                                                                i64 count, i64 n, const double *__restrict__ series,
                                                                i64 step) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k];
    const T phi = T(series[step]);
#pragma unroll
    for (int q = 0; q < 7; ++q) g[q * n + c] = phi * w7<T>(q);
}

// OutletCSV::log's statistics (OutletCSV.hpp:29-58) on the device: FP64 mean, min and max of phi over
// the listed cells, in a fixed summation order (one workgroup of 1024).
template <class T>
__global__ void __launch_bounds__(1024) k_outlet_stats(const T *__restrict__ phi, const i64 *__restrict__ cells,
                                                       i64 count, double *__restrict__ out) {
    __shared__ double rs[1024], rmn[1024], rmx[1024];
    const int t = int(threadIdx.x);
    double s = 0.0, mn = __builtin_inf(), mx = -__builtin_inf();
    for (i64 k = t; k < count; k += blockDim.x) {
        const double v = double(phi[cells[k]]);
        s += v;
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    rs[t] = s;
    rmn[t] = mn;
    rmx[t] = mx;
    __syncthreads();
    for (int h = int(blockDim.x) / 2; h > 0; h >>= 1) {
        if (t < h) {
            rs[t] += rs[t + h];
            rmn[t] = rmn[t + h] < rmn[t] ? rmn[t + h] : rmn[t];
            rmx[t] = rmx[t + h] > rmx[t] ? rmx[t + h] : rmx[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = rs[0] / double(count);
        out[1] = rmn[0];
        out[2] = rmx[0];
    }
}

// k_ad_neumann (ADNeumannOutlet.cu:5-31)
template <class T>
__global__ void __launch_bounds__(256) k_ad_bc_neumann(T *__restrict__ g, const i64 *__restrict__ cells,
                                                       const i64 *__restrict__ src, i64 count, i64 n) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k], s = src[k];
#pragma unroll
    for (int q = 0; q < 7; ++q) g[q * n + c] = g[q * n + s];
}

// ------------------------------------------------------------------ LDF
// k_henry_ldf_source (LinearLDF.cu:13-48) and k_langmuir_ldf_source (LangmuirLDF.cu:7-41)
template <class T, bool LANGMUIR>
__global__ void __launch_bounds__(256)
    k_ldf(i64 n, const std::uint8_t *__restrict__ region, const T *__restrict__ phi, T *__restrict__ S,
          T *__restrict__ q, T a, T b, T kLDF, T phiMin) {
    const i64 i = gtid();
    if (i >= n) return;
    if (!region[i]) {
        S[i] = T(0);
        return;
    }
    T c = phi[i];
    if (c < phiMin) c = phiMin;
    const T ce = c - phiMin;
    T qeq;
    if (LANGMUIR) {
        const T bc = b * ce;
        qeq = a * (bc / (T(1) + bc));
    } else {
        qeq = a * ce;
    }
    const T qi = q[i];
    const T dq = kLDF * (qeq - qi);
    q[i] = qi + dq;
    S[i] = -dq;
}

// --------------------------------------------------------------- layout
template <class T>
__global__ void __launch_bounds__(256) k_pack_aos(const T *__restrict__ soa, T *__restrict__ aos, int Q, i64 n) {
    const i64 i = gtid();
    if (i >= n * Q) return;
    const i64 c = i / Q;
    const int q = int(i - c * Q);
    aos[i] = soa[q * n + c];
}
template <class T>
__global__ void __launch_bounds__(256) k_unpack_aos(const T *__restrict__ aos, T *__restrict__ soa, int Q, i64 n) {
    const i64 i = gtid();
    if (i >= n * Q) return;
    const i64 c = i / Q;
    const int q = int(i - c * Q);
    soa[q * n + c] = aos[i];
}

// Byte fill as kernels (a memset node in a captured graph is what these avoid): 32-bit words of the
// repeated byte over the aligned middle, single bytes at the two ends.
__global__ void __launch_bounds__(256) k_fill_words(std::uint32_t *__restrict__ p, i64 words, std::uint32_t pattern) {
    const i64 stride = i64(gridDim.x) * blockDim.x;
    for (i64 i = gtid(); i < words; i += stride) p[i] = pattern;
}
__global__ void k_fill_bytes(unsigned char *__restrict__ p, i64 nbytes, unsigned char value) {
    const i64 i = gtid();
    if (i < nbytes) p[i] = value;
}

// Streaming probes: 16 bytes per lane, grid-stride, non-temporal.  They measure the HBM rate a
// plain stream reaches on this device (the practical ceiling the stencil kernels are compared with).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_stream_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, i64 n16) {
    const i64 stride = i64(gridDim.x) * blockDim.x;
    for (i64 i = gtid(); i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
__global__ void __launch_bounds__(256) k_stream_read(const u32x4 *__restrict__ src, i64 n16, unsigned *__restrict__ sink) {
    const i64 stride = i64(gridDim.x) * blockDim.x;
    unsigned acc = 0;
    for (i64 i = gtid(); i < n16; i += stride) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads live; practically never taken
}

// The population access pattern without the arithmetic: each thread copies one element of K planes
// (read all, then write all), one cell per thread.
template <class T, int K>
__global__ void __launch_bounds__(256) k_plane_copy(const T *__restrict__ src, T *__restrict__ dst, i64 n) {
    const i64 c = gtid();
    if (c >= n) return;
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src[i64(k) * n + c];
#pragma unroll
    for (int k = 0; k < K; ++k) dst[i64(k) * n + c] = v[k];
}

template <class T> static int ad_collide(const void *g, void *gn, void *phi, const void *tau, const void *ux,
                                         const void *uy, const void *uz, const void *S, const fmx_grid &gr,
                                         double minC, hipStream_t s) {
    const dim3 block(64, 4, 1);
    const dim3 grid((gr.nx + 63) / 64, (gr.ny + 3) / 4, gr.nz);
    hipLaunchKernelGGL(k_ad_collide<T>, grid, block, 0, s, static_cast<const T *>(g), static_cast<T *>(gn),
                       static_cast<T *>(phi), static_cast<const T *>(tau), static_cast<const T *>(ux),
                       static_cast<const T *>(uy), static_cast<const T *>(uz), static_cast<const T *>(S), gr, T(minC));
    return after_launch("k_ad_collide");
}

template <class T, bool L>
static int ldf(i64 n, const std::uint8_t *region, const void *phi, void *S, void *q, double a, double b, double k,
               double minC, hipStream_t s) {
    hipLaunchKernelGGL((k_ldf<T, L>), dim3(blocks_for(n)), dim3(256), 0, s, n, region, static_cast<const T *>(phi),
                       static_cast<T *>(S), static_cast<T *>(q), T(a), T(b), T(k), T(minC));
    return after_launch("k_ldf");
}


}  // namespace fmx

using namespace fmx;

extern "C" {

int fmx_ad_init_eq(int prec, void *g, const fmx_grid *grid, double phi0, void *stream) {
    FMX_TRY_BEGIN
    if (!g || !grid || !grid_ok(*grid)) return set_error(FMX_ERR_ARG, "fmx_ad_init_eq: bad arguments");
    const i64 n = grid_cells(*grid);
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ad_init_eq<float>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<float *>(g), n,
                           float(phi0));
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ad_init_eq<double>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<double *>(g), n, phi0);
    else return unsupported_prec("fmx_ad_init_eq", prec);
    return after_launch("k_ad_init_eq");
    FMX_TRY_END
}

int fmx_ad_collide_stream(int prec, const void *g, void *g_new, void *phi_out, const void *tau, const void *ux,
                          const void *uy, const void *uz, const void *S, const fmx_grid *grid, double min_conc,
                          void *stream) {
    FMX_TRY_BEGIN
    if (!g || !g_new || g == g_new || !tau || !ux || !uy || !uz || !grid || !grid_ok(*grid))
        return set_error(FMX_ERR_ARG, "fmx_ad_collide_stream: bad arguments");
    if (prec == FMX_F32) return ad_collide<float>(g, g_new, phi_out, tau, ux, uy, uz, S, *grid, min_conc, as_stream(stream));
    if (prec == FMX_F64) return ad_collide<double>(g, g_new, phi_out, tau, ux, uy, uz, S, *grid, min_conc, as_stream(stream));
    return unsupported_prec("fmx_ad_collide_stream", prec);
    FMX_TRY_END
}

int fmx_ad_bc_dirichlet(int prec, void *g, const int64_t *cells, int64_t count, double phi, int64_t ncells,
                        void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!g || !cells || count < 0 || ncells <= 0) return set_error(FMX_ERR_ARG, "fmx_ad_bc_dirichlet: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ad_bc_dirichlet<float>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<float *>(g),
                           cells, count, ncells, float(phi));
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ad_bc_dirichlet<double>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<double *>(g),
                           cells, count, ncells, phi);
    else return unsupported_prec("fmx_ad_bc_dirichlet", prec);
    return after_launch("k_ad_bc_dirichlet");
    FMX_TRY_END
}

int fmx_ad_bc_dirichlet_series(int prec, void *g, const int64_t *cells, int64_t count, const double *series,
                               int64_t step, int64_t ncells, void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!g || !cells || !series || count < 0 || ncells <= 0 || step < 0)
        return set_error(FMX_ERR_ARG, "fmx_ad_bc_dirichlet_series: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ad_bc_dirichlet_series<float>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<float *>(g), cells, count, ncells, series, step);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ad_bc_dirichlet_series<double>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<double *>(g), cells, count, ncells, series, step);
    else return unsupported_prec("fmx_ad_bc_dirichlet_series", prec);
    return after_launch("k_ad_bc_dirichlet_series");
    FMX_TRY_END
}

int fmx_reduce_outlet_mean(int prec, const void *phi, const int64_t *cells, int64_t count, double *out, void *stream) {
    FMX_TRY_BEGIN
    if (!phi || !cells || count <= 0 || !out) return set_error(FMX_ERR_ARG, "fmx_reduce_outlet_mean: bad arguments");
    hipStream_t s = as_stream(stream);
    // the concentration is in the compute precision: double for FP64 storage, float otherwise
    if (prec == FMX_F64)
        hipLaunchKernelGGL(k_outlet_stats<double>, dim3(1), dim3(1024), 0, s, static_cast<const double *>(phi), cells,
                           count, out);
    else if (prec == FMX_F32 || prec == FMX_F16 || prec == FMX_BF16)
        hipLaunchKernelGGL(k_outlet_stats<float>, dim3(1), dim3(1024), 0, s, static_cast<const float *>(phi), cells,
                           count, out);
    else return set_error(FMX_ERR_ARG, "fmx_reduce_outlet_mean: unknown precision %d", prec);
    return after_launch("k_outlet_stats");
    FMX_TRY_END
}

int fmx_ad_bc_neumann(int prec, void *g_new, const int64_t *cells, const int64_t *src, int64_t count, int64_t ncells,
                      void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!g_new || !cells || !src || count < 0 || ncells <= 0)
        return set_error(FMX_ERR_ARG, "fmx_ad_bc_neumann: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ad_bc_neumann<float>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<float *>(g_new), cells, src, count, ncells);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ad_bc_neumann<double>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<double *>(g_new), cells, src, count, ncells);
    else return unsupported_prec("fmx_ad_bc_neumann", prec);
    return after_launch("k_ad_bc_neumann");
    FMX_TRY_END
}

int fmx_ldf_linear(int prec, int64_t n, const std::uint8_t *region, const void *phi, void *S, void *q, double K,
                   double k_ldf, double min_conc, void *stream) {
    FMX_TRY_BEGIN
    if (n == 0) return FMX_OK;
    if (n < 0 || !region || !phi || !S || !q) return set_error(FMX_ERR_ARG, "fmx_ldf_linear: bad arguments");
    if (prec == FMX_F32) return ldf<float, false>(n, region, phi, S, q, K, 0.0, k_ldf, min_conc, as_stream(stream));
    if (prec == FMX_F64) return ldf<double, false>(n, region, phi, S, q, K, 0.0, k_ldf, min_conc, as_stream(stream));
    return unsupported_prec("fmx_ldf_linear", prec);
    FMX_TRY_END
}

int fmx_ldf_langmuir(int prec, int64_t n, const std::uint8_t *region, const void *phi, void *S, void *q, double q_max,
                     double b, double k_ldf, double min_conc, void *stream) {
    FMX_TRY_BEGIN
    if (n == 0) return FMX_OK;
    if (n < 0 || !region || !phi || !S || !q) return set_error(FMX_ERR_ARG, "fmx_ldf_langmuir: bad arguments");
    if (prec == FMX_F32) return ldf<float, true>(n, region, phi, S, q, q_max, b, k_ldf, min_conc, as_stream(stream));
    if (prec == FMX_F64) return ldf<double, true>(n, region, phi, S, q, q_max, b, k_ldf, min_conc, as_stream(stream));
    return unsupported_prec("fmx_ldf_langmuir", prec);
    FMX_TRY_END
}

int fmx_pack_aos(int prec, const void *soa, void *aos, int Q, int64_t n, void *stream) {
    FMX_TRY_BEGIN
    if (!soa || !aos || n <= 0 || (Q != 19 && Q != 7)) return set_error(FMX_ERR_ARG, "fmx_pack_aos: bad arguments");
    hipStream_t s = as_stream(stream);
    const i64 tot = n * Q;
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_pack_aos<float>, dim3(blocks_for(tot)), dim3(256), 0, s, static_cast<const float *>(soa),
                           static_cast<float *>(aos), Q, n);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_pack_aos<double>, dim3(blocks_for(tot)), dim3(256), 0, s, static_cast<const double *>(soa),
                           static_cast<double *>(aos), Q, n);
    else return unsupported_prec("fmx_pack_aos", prec);
    return after_launch("k_pack_aos");
    FMX_TRY_END
}

int fmx_unpack_aos(int prec, const void *aos, void *soa, int Q, int64_t n, void *stream) {
    FMX_TRY_BEGIN
    if (!soa || !aos || n <= 0 || (Q != 19 && Q != 7)) return set_error(FMX_ERR_ARG, "fmx_unpack_aos: bad arguments");
    hipStream_t s = as_stream(stream);
    const i64 tot = n * Q;
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_unpack_aos<float>, dim3(blocks_for(tot)), dim3(256), 0, s, static_cast<const float *>(aos),
                           static_cast<float *>(soa), Q, n);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_unpack_aos<double>, dim3(blocks_for(tot)), dim3(256), 0, s,
                           static_cast<const double *>(aos), static_cast<double *>(soa), Q, n);
    else return unsupported_prec("fmx_unpack_aos", prec);
    return after_launch("k_unpack_aos");
    FMX_TRY_END
}

static int fill_bytes(void *dst, unsigned char value, size_t bytes, void *stream, const char *who) {
    if (bytes == 0) return FMX_OK;
    if (!dst) return set_error(FMX_ERR_ARG, "%s: null pointer", who);
    hipStream_t s = as_stream(stream);
    unsigned char *p = static_cast<unsigned char *>(dst);
    const i64 head = i64((4 - reinterpret_cast<std::uintptr_t>(p) % 4) % 4) < i64(bytes)
                         ? i64((4 - reinterpret_cast<std::uintptr_t>(p) % 4) % 4)
                         : i64(bytes);
    if (head) {
        hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(4), 0, s, p, head, value);
        if (int rc = after_launch("k_fill_bytes")) return rc;
    }
    const i64 words = (i64(bytes) - head) / 4;
    if (words) {
        const unsigned nb = blocks_for(words) < 65536u ? blocks_for(words) : 65536u;
        hipLaunchKernelGGL(k_fill_words, dim3(nb), dim3(256), 0, s, reinterpret_cast<std::uint32_t *>(p + head), words,
                           std::uint32_t(value) * 0x01010101u);
        if (int rc = after_launch("k_fill_words")) return rc;
    }
    const i64 tail = i64(bytes) - head - 4 * words;
    if (tail) {
        hipLaunchKernelGGL(k_fill_bytes, dim3(1), dim3(4), 0, s, p + head + 4 * words, tail, value);
        if (int rc = after_launch("k_fill_bytes")) return rc;
    }
    return FMX_OK;
}

int fmx_fill_zero(void *dst, size_t bytes, void *stream) {
    FMX_TRY_BEGIN
    return fill_bytes(dst, 0, bytes, stream, "fmx_fill_zero");
    FMX_TRY_END
}

// cudaMemset's contract (the low byte of value in every byte), done by the fill kernels above
int fmx_memset(void *dst, int value, size_t bytes, void *stream) {
    FMX_TRY_BEGIN
    return fill_bytes(dst, static_cast<unsigned char>(value & 0xff), bytes, stream, "fmx_memset");
    FMX_TRY_END
}

int fmx_stream_probe(int mode, void *dst, const void *src, size_t bytes, void *stream) {
    FMX_TRY_BEGIN
    if (!src || bytes < 16 || bytes % 16 || (mode == 0 && !dst))
        return set_error(FMX_ERR_ARG, "fmx_stream_probe: bad arguments");
    const i64 n16 = i64(bytes / 16);
    int dev = 0, cus = 256;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (int rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"))
        return rc;
    const unsigned nb = unsigned(cus) * 8u;
    if (mode >= 2 && mode <= 4) {  // 26-plane copies with 2-, 4- and 8-byte elements
        if (!dst) return set_error(FMX_ERR_ARG, "fmx_stream_probe: bad arguments");
        const i64 esz = mode == 2 ? 2 : (mode == 3 ? 4 : 8);
        const i64 n = i64(bytes) / (26 * esz);
        if (mode == 2)
            hipLaunchKernelGGL((k_plane_copy<unsigned short, 26>), dim3(blocks_for(n)), dim3(256), 0, as_stream(stream),
                               static_cast<const unsigned short *>(src), static_cast<unsigned short *>(dst), n);
        else if (mode == 3)
            hipLaunchKernelGGL((k_plane_copy<unsigned, 26>), dim3(blocks_for(n)), dim3(256), 0, as_stream(stream),
                               static_cast<const unsigned *>(src), static_cast<unsigned *>(dst), n);
        else
            hipLaunchKernelGGL((k_plane_copy<unsigned long long, 26>), dim3(blocks_for(n)), dim3(256), 0,
                               as_stream(stream), static_cast<const unsigned long long *>(src),
                               static_cast<unsigned long long *>(dst), n);
    } else if (mode == 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, as_stream(stream), static_cast<const u32x4 *>(src),
                           static_cast<u32x4 *>(dst), n16);
    else
        hipLaunchKernelGGL(k_stream_read, dim3(nb), dim3(256), 0, as_stream(stream), static_cast<const u32x4 *>(src), n16,
                           static_cast<unsigned *>(dst));
    return after_launch("k_stream_probe");
    FMX_TRY_END
}


}  // extern "C"
