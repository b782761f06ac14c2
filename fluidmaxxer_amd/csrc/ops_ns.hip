// ops_ns.hip -- D3Q19 Navier-Stokes operators, one kernel per reference operator.
//
// These are the drop-in kernels behind the NS operator API (NSInterfaces.hpp:9-22):
// each reproduces one reference kernel's arithmetic on structure-of-arrays
// populations.  The fused coupled step (coupled.hip) is the fast path; these
// serve the general operator API (arbitrary masks, any operator order).
#include "fmx_internal.hpp"

namespace fmx {

// Source coordinates of the pull x - c for every c component in {-1, 0, 1},
// with the reference's periodic wrap / reject rule (BGKCollision.cu:26-41).
struct Pull3 {
    int xs[3], ys[3], zs[3];
    bool vx[3], vy[3], vz[3];
    __device__ __forceinline__ Pull3(int x, int y, int z, const fmx_grid &g) {
        axis(x, g.nx, g.periodic_x != 0, xs, vx);
        axis(y, g.ny, g.periodic_y != 0, ys, vy);
        axis(z, g.nz, g.periodic_z != 0, zs, vz);
    }
    // index d + 1 holds coordinate v + d
    static __device__ __forceinline__ void axis(int v, int n, bool periodic, int *s, bool *ok) {
        s[1] = v;
        ok[1] = true;
        s[0] = v - 1;
        s[2] = v + 1;
        ok[0] = s[0] >= 0;
        ok[2] = s[2] < n;
        if (periodic) {
            if (!ok[0]) s[0] += n;
            if (!ok[2]) s[2] -= n;
            ok[0] = ok[2] = true;
        }
    }
    // pull source for direction (cx, cy, cz): cell x - c
    __device__ __forceinline__ bool valid(int cx, int cy, int cz) const {
        return vx[1 - cx] && vy[1 - cy] && vz[1 - cz];
    }
    __device__ __forceinline__ i64 src(int cx, int cy, int cz, int nx, int ny) const {
        return cell_index(xs[1 - cx], ys[1 - cy], zs[1 - cz], nx, ny);
    }
};

template <class T>
__device__ __forceinline__ T feq19(int q, T rho, T ux, T uy, T uz, T u2) {
    const T eu = T(Q19::cx[q]) * ux + T(Q19::cy[q]) * uy + T(Q19::cz[q]) * uz;
    const T cu = T(3) * eu;
    return w19<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
}

// ------------------------------------------------------------------ init
template <class T> __global__ void __launch_bounds__(256) k_ns_init_eq(T *__restrict__ f, i64 n, T rho0) {
    const i64 c = gtid();
    if (c >= n) return;
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q * n + c] = w19<T>(q) * rho0;
}

// ------------------------------------------------------- collide + stream
// OP: 0 BGK with obstacle bounce-back (BGKCollision.cu:74-177)
//     1 porous BGK, u <- clamp(phi) * j / rho (PorousBGKCollision.cu:64-155)
//     2 partial bounce-back, f_out = (1 - ns) f^c + ns f_opp (Walsh et al. 2009)
template <class T, int OP>
__global__ void __launch_bounds__(256)
    k_ns_collide(const T *__restrict__ f, T *__restrict__ f_new, const std::uint8_t *__restrict__ obst,
                 const T *__restrict__ field, fmx_grid g, T omega, T rho0) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const i64 n = i64(g.nx) * g.ny * g.nz;
    const i64 cell = cell_index(x, y, z, g.nx, g.ny);

    if (OP == 0 && obst && obst[cell]) {  // solid: on-site reverse (BGKCollision.cu:90-97)
#pragma unroll
        for (int q = 0; q < 19; ++q) f_new[q * n + cell] = f[Q19::opp[q] * n + cell];
        return;
    }
    const Pull3 P(x, y, z, g);
    T fi[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        if (!P.valid(Q19::cx[q], Q19::cy[q], Q19::cz[q])) {
            fi[q] = f[q * n + cell];
        } else {
            const i64 src = P.src(Q19::cx[q], Q19::cy[q], Q19::cz[q], g.nx, g.ny);
            if (OP == 0 && obst && obst[src]) fi[q] = f[Q19::opp[q] * n + cell];
            else fi[q] = f[q * n + src];
        }
    }
    T rho = T(0), jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T v = fi[q];
        rho += v;
        jx += v * T(Q19::cx[q]);
        jy += v * T(Q19::cy[q]);
        jz += v * T(Q19::cz[q]);
    }
    T ux, uy, uz;
    if (OP == 1) {  // PorousBGKCollision.cu:118-136
        if (!(rho > T(0))) {
            rho = rho0;
            jx = jy = jz = T(0);
        }
        const T invR = T(1) / rho;
        ux = jx * invR;
        uy = jy * invR;
        uz = jz * invR;
        T phi = field ? field[cell] : T(1);
        if (phi < T(0)) phi = T(0);
        if (phi > T(1)) phi = T(1);
        ux *= phi;
        uy *= phi;
        uz *= phi;
    } else {  // BGKCollision.cu:149-158
        if (!(rho > T(0))) {
            rho = rho0;
            ux = uy = uz = T(0);
        } else {
            const T invR = T(1) / rho;
            ux = jx * invR;
            uy = jy * invR;
            uz = jz * invR;
        }
    }
    const T u2 = ux * ux + uy * uy + uz * uz;
    T ns = T(0);
    if (OP == 2) {
        ns = field ? field[cell] : T(0);
        if (ns < T(0)) ns = T(0);
        if (ns > T(1)) ns = T(1);
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T feq = feq19<T>(q, rho, ux, uy, uz, u2);
        const T fc = fi[q] + omega * (feq - fi[q]);
        if (OP == 2) f_new[q * n + cell] = (T(1) - ns) * fc + ns * fi[Q19::opp[q]];
        else f_new[q * n + cell] = fc;
    }
}

// --------------------------------------------------------------- moments
// k_compute_macros_d3q19 (Macros.cu:20-61)
template <class T>
__global__ void __launch_bounds__(256)
    k_ns_moments(const T *__restrict__ f, T *__restrict__ rho, T *__restrict__ ux, T *__restrict__ uy,
                 T *__restrict__ uz, i64 n) {
    const i64 c = gtid();
    if (c >= n) return;
    T r = T(0), vx = T(0), vy = T(0), vz = T(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T fi = f[q * n + c];
        r += fi;
        vx += fi * T(Q19::cx[q]);
        vy += fi * T(Q19::cy[q]);
        vz += fi * T(Q19::cz[q]);
    }
    const T invR = (r > T(0)) ? (T(1) / r) : T(0);
    rho[c] = r;
    ux[c] = vx * invR;
    uy[c] = vy * invR;
    uz[c] = vz * invR;
}

// ------------------------------------------------------------------- BCs
// k_velocity_dirichlet_neq_pre_d3q19 (VelocityDirichletNEQ.cu:49-96)
template <class T>
__global__ void __launch_bounds__(256)
    k_ns_bc_neq(T *__restrict__ f, const i64 *__restrict__ bnd, const i64 *__restrict__ nei, i64 count, i64 n, T ux,
                T uy, T uz) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 cb = bnd[k], cn = nei[k];
    T fN[19];
    T r = T(0), jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        fN[q] = f[q * n + cn];
        r += fN[q];
        jx += fN[q] * T(Q19::cx[q]);
        jy += fN[q] * T(Q19::cy[q]);
        jz += fN[q] * T(Q19::cz[q]);
    }
    const T inv = (r > T(0)) ? (T(1) / r) : T(0);
    const T uxN = jx * inv, uyN = jy * inv, uzN = jz * inv;
    const T u2N = uxN * uxN + uyN * uyN + uzN * uzN;
    const T u2B = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T feqN = feq19<T>(q, r, uxN, uyN, uzN, u2N);
        const T feqB = feq19<T>(q, r, ux, uy, uz, u2B);
        f[q * n + cb] = feqB + (fN[q] - feqN);
    }
}

// k_velocity_dirichlet_eq_pre_d3q19 (VelocityDirichlet.cu:54-84)
template <class T>
__global__ void __launch_bounds__(256)
    k_ns_bc_eq(T *__restrict__ f, const i64 *__restrict__ cells, i64 count, i64 n, T rho0, T ux, T uy, T uz) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k];
    const T u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q * n + c] = feq19<T>(q, rho0, ux, uy, uz, u2);
}

// k_pressure_dirichlet_outlet_d3q19 (PressureDirichlet.cu:26-96)
template <class T>
__global__ void __launch_bounds__(256)
    k_ns_bc_pressure(T *__restrict__ f, const i64 *__restrict__ cells, const i64 *__restrict__ nei, i64 count, i64 n,
                     T rho_out) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k], cn = nei[k];
    T rhoN = T(0), ux = T(0), uy = T(0), uz = T(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T fi = f[q * n + cn];
        rhoN += fi;
        ux += fi * T(Q19::cx[q]);
        uy += fi * T(Q19::cy[q]);
        uz += fi * T(Q19::cz[q]);
    }
    if (rhoN > T(0)) {
        const T inv = T(1) / rhoN;
        ux *= inv;
        uy *= inv;
        uz *= inv;
    } else {
        ux = uy = uz = T(0);
    }
    const T u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const T eu = T(Q19::cx[q]) * ux + T(Q19::cy[q]) * uy + T(Q19::cz[q]) * uz;
        f[q * n + c] = w19<T>(q) * rho_out * (T(1) + T(3) * eu + T(4.5) * eu * eu - T(1.5) * u2);
    }
}

// BounceBack (BounceBack.cu:12-31) as a true swap of each opposite pair
template <class T>
__global__ void __launch_bounds__(256) k_ns_bc_bounceback(T *__restrict__ f, const i64 *__restrict__ cells, i64 count,
                                                          i64 n) {
    const i64 k = gtid();
    if (k >= count) return;
    const i64 c = cells[k];
    T v[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = f[Q19::opp[q] * n + c];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q * n + c] = v[q];
}

// ------------------------------------------------------------- launchers
template <class T> static int init_eq(void *f, const fmx_grid &g, double rho0, hipStream_t s) {
    const i64 n = grid_cells(g);
    hipLaunchKernelGGL(k_ns_init_eq<T>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<T *>(f), n, T(rho0));
    return after_launch("k_ns_init_eq");
}

template <class T>
static int collide(int op, const void *f, void *fn, const std::uint8_t *obst, const void *field, const fmx_grid &g,
                   double tau, double rho0, hipStream_t s) {
    T t = T(tau);
    if (t < T(0.5001)) t = T(0.5001);  // BGKCollision.cu:203-204, PorousBGKCollision.cu:180
    const T omega = T(1) / t;
    const dim3 block(64, 4, 1);
    const dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    const T *fi = static_cast<const T *>(f);
    T *fo = static_cast<T *>(fn);
    const T *fl = static_cast<const T *>(field);
    switch (op) {
        case FMX_NS_BGK: hipLaunchKernelGGL((k_ns_collide<T, 0>), grid, block, 0, s, fi, fo, obst, fl, g, omega, T(rho0)); break;
        case FMX_NS_POROUS: hipLaunchKernelGGL((k_ns_collide<T, 1>), grid, block, 0, s, fi, fo, obst, fl, g, omega, T(rho0)); break;
        case FMX_NS_PBB: hipLaunchKernelGGL((k_ns_collide<T, 2>), grid, block, 0, s, fi, fo, obst, fl, g, omega, T(rho0)); break;
        default: return set_error(FMX_ERR_ARG, "fmx_ns_collide_stream: unknown operator %d", op);
    }
    return after_launch("k_ns_collide");
}

}  // namespace fmx

using namespace fmx;

extern "C" {

int fmx_ns_init_eq(int prec, void *f, const fmx_grid *grid, double rho0, void *stream) {
    FMX_TRY_BEGIN
    if (!f || !grid || !grid_ok(*grid)) return set_error(FMX_ERR_ARG, "fmx_ns_init_eq: bad arguments");
    if (prec == FMX_F32) return init_eq<float>(f, *grid, rho0, as_stream(stream));
    if (prec == FMX_F64) return init_eq<double>(f, *grid, rho0, as_stream(stream));
    return unsupported_prec("fmx_ns_init_eq", prec);
    FMX_TRY_END
}

int fmx_ns_collide_stream(int prec, int op, const void *f, void *f_new, const std::uint8_t *obstacle, const void *field,
                          const fmx_grid *grid, double tau, double rho0, void *stream) {
    FMX_TRY_BEGIN
    if (!f || !f_new || !grid || !grid_ok(*grid) || f == f_new)
        return set_error(FMX_ERR_ARG, "fmx_ns_collide_stream: bad arguments");
    if (prec == FMX_F32) return collide<float>(op, f, f_new, obstacle, field, *grid, tau, rho0, as_stream(stream));
    if (prec == FMX_F64) return collide<double>(op, f, f_new, obstacle, field, *grid, tau, rho0, as_stream(stream));
    return unsupported_prec("fmx_ns_collide_stream", prec);
    FMX_TRY_END
}

int fmx_ns_moments(int prec, const void *f, void *rho, void *ux, void *uy, void *uz, const fmx_grid *grid,
                   void *stream) {
    FMX_TRY_BEGIN
    if (!f || !rho || !ux || !uy || !uz || !grid || !grid_ok(*grid))
        return set_error(FMX_ERR_ARG, "fmx_ns_moments: bad arguments");
    const i64 n = grid_cells(*grid);
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ns_moments<float>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const float *>(f),
                           static_cast<float *>(rho), static_cast<float *>(ux), static_cast<float *>(uy),
                           static_cast<float *>(uz), n);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ns_moments<double>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const double *>(f),
                           static_cast<double *>(rho), static_cast<double *>(ux), static_cast<double *>(uy),
                           static_cast<double *>(uz), n);
    else return unsupported_prec("fmx_ns_moments", prec);
    return after_launch("k_ns_moments");
    FMX_TRY_END
}

int fmx_ns_bc_velocity_neq(int prec, void *f, const int64_t *bnd, const int64_t *nei, int64_t count, double ux,
                           double uy, double uz, int64_t ncells, void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!f || !bnd || !nei || count < 0 || ncells <= 0)
        return set_error(FMX_ERR_ARG, "fmx_ns_bc_velocity_neq: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ns_bc_neq<float>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<float *>(f), bnd,
                           nei, count, ncells, float(ux), float(uy), float(uz));
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ns_bc_neq<double>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<double *>(f), bnd,
                           nei, count, ncells, ux, uy, uz);
    else return unsupported_prec("fmx_ns_bc_velocity_neq", prec);
    return after_launch("k_ns_bc_neq");
    FMX_TRY_END
}

int fmx_ns_bc_velocity_eq(int prec, void *f, const int64_t *cells, int64_t count, double rho0, double ux, double uy,
                          double uz, int64_t ncells, void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!f || !cells || count < 0 || ncells <= 0) return set_error(FMX_ERR_ARG, "fmx_ns_bc_velocity_eq: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ns_bc_eq<float>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<float *>(f), cells,
                           count, ncells, float(rho0), float(ux), float(uy), float(uz));
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ns_bc_eq<double>, dim3(blocks_for(count)), dim3(256), 0, s, static_cast<double *>(f), cells,
                           count, ncells, rho0, ux, uy, uz);
    else return unsupported_prec("fmx_ns_bc_velocity_eq", prec);
    return after_launch("k_ns_bc_eq");
    FMX_TRY_END
}

int fmx_ns_bc_pressure(int prec, void *f_new, const int64_t *cells, const int64_t *nei, int64_t count, double rho_out,
                       int64_t ncells, void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!f_new || !cells || !nei || count < 0 || ncells <= 0)
        return set_error(FMX_ERR_ARG, "fmx_ns_bc_pressure: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ns_bc_pressure<float>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<float *>(f_new), cells, nei, count, ncells, float(rho_out));
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ns_bc_pressure<double>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<double *>(f_new), cells, nei, count, ncells, rho_out);
    else return unsupported_prec("fmx_ns_bc_pressure", prec);
    return after_launch("k_ns_bc_pressure");
    FMX_TRY_END
}

int fmx_ns_bc_bounceback(int prec, void *f_new, const int64_t *cells, int64_t count, int64_t ncells, void *stream) {
    FMX_TRY_BEGIN
    if (count == 0) return FMX_OK;
    if (!f_new || !cells || count < 0 || ncells <= 0) return set_error(FMX_ERR_ARG, "fmx_ns_bc_bounceback: bad arguments");
    hipStream_t s = as_stream(stream);
    if (prec == FMX_F32)
        hipLaunchKernelGGL(k_ns_bc_bounceback<float>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<float *>(f_new), cells, count, ncells);
    else if (prec == FMX_F64)
        hipLaunchKernelGGL(k_ns_bc_bounceback<double>, dim3(blocks_for(count)), dim3(256), 0, s,
                           static_cast<double *>(f_new), cells, count, ncells);
    else return unsupported_prec("fmx_ns_bc_bounceback", prec);
    return after_launch("k_ns_bc_bounceback");
    FMX_TRY_END
}

}  // extern "C"
