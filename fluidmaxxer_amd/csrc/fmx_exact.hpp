// fmx_exact.hpp -- compile-time expansion of the lattice sums.
//
// The reference evaluates sums such as rho = sum_q f_q, j_x = sum_q f_q c_xq and
// e.u = c_x u_x + c_y u_y + c_z u_z literally, multiplying by the integer velocity
// components.  Without fast-math a compiler may not drop the `v * 0` terms (they would
// matter for infinities and the sign of zero), so this header expands every such sum at
// compile time: zero terms are omitted and +-1 factors become additions/subtractions, in
// the reference's left-to-right order.  For finite values the results are bit-identical to
// the literal evaluation (adding +-0 to a non-zero partial sum is exact, x * 1 = x,
// x + (-1) y = x - y).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

namespace fmx {

template <int Begin, int End, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

// acc + c * v for c in {-1, 0, 1}
template <int C, class T> __device__ __forceinline__ T axpy1(T acc, T v) {
    if constexpr (C == 1) return acc + v;
    else if constexpr (C == -1) return acc - v;
    else return acc;
}

// (A x + B y) + C z with A, B, C in {-1, 0, 1}
template <int A, int B, int Cz, class T> __device__ __forceinline__ T dot3(T x, T y, T z) {
    if constexpr (A != 0) {
        T s = (A > 0) ? x : -x;
        s = axpy1<B>(s, y);
        return axpy1<Cz>(s, z);
    } else if constexpr (B != 0) {
        T s = (B > 0) ? y : -y;
        return axpy1<Cz>(s, z);
    } else if constexpr (Cz != 0) {
        return (Cz > 0) ? z : -z;
    } else {
        return T(0);
    }
}

}  // namespace fmx
