// fmx_device.hpp -- shared device-side definitions for the gfx950 kernels.
//
// Populations live in structure-of-arrays order: population q of cell c is at
// pop[q * N + c], c = x + Nx * (y + Ny * z) (x fastest, the reference's cell
// order, NSLattice.hpp:83-86).  All offsets are 64-bit.
//
// The lattice tables are compile-time constants (D3Q19.hpp:5-50, D3Q7.hpp:5-28);
// every loop over q is fully unrolled, so no table is ever indexed at run time.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fmx_abi.h"

namespace fmx {

using i64 = std::int64_t;

// ------------------------------------------------------------ descriptors
struct Q19 {
    static constexpr int Q = 19;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    // 1 / w_q, exact small integers
    static constexpr int inv_w[19] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36};
};
struct Q7 {
    static constexpr int Q = 7;
    static constexpr int cx[7] = {0, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[7] = {0, 0, 0, 1, -1, 0, 0};
    static constexpr int cz[7] = {0, 0, 0, 0, 0, 1, -1};
    static constexpr int opp[7] = {0, 2, 1, 4, 3, 6, 5};
    static constexpr int inv_w[7] = {4, 8, 8, 8, 8, 8, 8};
};

// w_q evaluated as T(1)/T(n), as the reference tables do.
template <class T> __host__ __device__ constexpr T w19(int q) {
    return T(1) / T(Q19::inv_w[q]);
}
template <class T> __host__ __device__ constexpr T w7(int q) {
    return T(1) / T(Q7::inv_w[q]);
}

// --------------------------------------------------- storage conversions
// Storage S, compute C.  Half formats hold the scaled deviation
// h = p / w_q - ref, so p = w_q * (ref + h): every direction uses the full
// half-precision mantissa regardless of its weight.
template <int P> struct Storage;
template <> struct Storage<FMX_F32> {
    using S = float;
    using C = float;
    static constexpr int bytes = 4;
};
template <> struct Storage<FMX_F64> {
    using S = double;
    using C = double;
    static constexpr int bytes = 8;
};
template <> struct Storage<FMX_F16> {
    using S = _Float16;
    using C = float;
    static constexpr int bytes = 2;
};
template <> struct Storage<FMX_BF16> {
    using S = __bf16;
    using C = float;
    static constexpr int bytes = 2;
};

template <int P> struct IsHalf {
    static constexpr bool value = (P == FMX_F16 || P == FMX_BF16);
};

// load population value; `w` and `ref` only matter for half storage
template <int P>
__device__ __forceinline__ typename Storage<P>::C ld_pop(const typename Storage<P>::S *p, i64 i,
                                                         typename Storage<P>::C w, typename Storage<P>::C ref) {
    using C = typename Storage<P>::C;
    if constexpr (IsHalf<P>::value) return w * (ref + C(p[i]));
    else return p[i];
}

template <int P>
__device__ __forceinline__ void st_pop(typename Storage<P>::S *p, i64 i, typename Storage<P>::C v,
                                       typename Storage<P>::C inv_w, typename Storage<P>::C ref) {
    using S = typename Storage<P>::S;
    if constexpr (IsHalf<P>::value) p[i] = S(v * inv_w - ref);
    else p[i] = v;
}

template <int P>
__device__ __forceinline__ void st_pop_nt(typename Storage<P>::S *p, i64 i, typename Storage<P>::C v,
                                          typename Storage<P>::C inv_w, typename Storage<P>::C ref) {
    using S = typename Storage<P>::S;
    if constexpr (IsHalf<P>::value) {
        S h = S(v * inv_w - ref);
        __builtin_nontemporal_store(h, p + i);
    } else {
        __builtin_nontemporal_store(v, p + i);
    }
}

__device__ __forceinline__ i64 cell_index(int x, int y, int z, int nx, int ny) {
    return i64(x) + i64(nx) * (i64(y) + i64(ny) * i64(z));
}

// 1-D grid-stride helpers for the plane / list kernels
__device__ __forceinline__ i64 gtid() { return i64(blockIdx.x) * blockDim.x + threadIdx.x; }

}  // namespace fmx
