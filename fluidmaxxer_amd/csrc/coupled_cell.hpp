// coupled_cell.hpp -- the per-cell arithmetic of the fused coupled step, shared by the
// one-cell-per-thread and the two-cells-per-thread kernels and by every storage precision.  Each
// function evaluates the reference's expressions in the reference's order; compiled without FMA
// contraction and with FAST = false (FP32 / FP64 storage) every kernel gives the same bits as the
// serial restatement of the reference.  FAST (FP16 / BF16 storage) swaps IEEE division for the
// hardware reciprocal.
#pragma once

#include <type_traits>

#include "coupled_common.hpp"

namespace fmx {

// 1/x: IEEE division, or with FAST the hardware reciprocal (v_rcp_f32, 1 ulp) for float.
template <bool FAST, class C> __device__ __forceinline__ C recip(C x) {
    if constexpr (FAST && std::is_same_v<C, float>) return __builtin_amdgcn_rcpf(x);
    else return C(1) / x;
}

// NS collision of one cell, in place: fo holds the pulled populations on entry and the
// populations the step stores on exit (unchanged for a BGK solid cell).
//   BGK:          BGKCollision.cu:149-173
//   porous BGK:   PorousBGKCollision.cu:118-150, u scaled by the clamped porosity `fld`
//   partial BB:   BGK collision blended with the reversed pulled populations by the clamped
//                 solid fraction `fld` (Walsh et al. 2009)
template <int OP, bool FAST = false, class C>
__device__ __forceinline__ void ns_collide_cell(C (&fo)[19], bool solid, C fld, const CoupledParams<C> &prm) {
    C rho = C(0), jx = C(0), jy = C(0), jz = C(0);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        rho += fo[q];
        jx = axpy1<Q19::cx[q]>(jx, fo[q]);
        jy = axpy1<Q19::cy[q]>(jy, fo[q]);
        jz = axpy1<Q19::cz[q]>(jz, fo[q]);
    });
    C ux, uy, uz;
    if (OP == FMX_NS_POROUS) {  // PorousBGKCollision.cu:118-136
        if (!(rho > C(0))) {
            rho = prm.rho0;
            jx = jy = jz = C(0);
        }
        const C invR = recip<FAST>(rho);
        ux = jx * invR;
        uy = jy * invR;
        uz = jz * invR;
        const C ph = fld < C(0) ? C(0) : (fld > C(1) ? C(1) : fld);
        ux *= ph;
        uy *= ph;
        uz *= ph;
    } else {  // BGKCollision.cu:149-158
        if (!(rho > C(0))) {
            rho = prm.rho0;
            ux = uy = uz = C(0);
        } else {
            const C invR = recip<FAST>(rho);
            ux = jx * invR;
            uy = jy * invR;
            uz = jz * invR;
        }
    }
    const C u2x15 = C(1.5) * (ux * ux + uy * uy + uz * uz);
    C ns = C(0);
    if (OP == FMX_NS_PBB) ns = fld < C(0) ? C(0) : (fld > C(1) ? C(1) : fld);
    C fc[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const C feq = feq_bgk<q>(rho, ux, uy, uz, u2x15);
        fc[q] = fo[q] + prm.omega * (feq - fo[q]);
    });
    if (OP == FMX_NS_PBB) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            fc[q] = (C(1) - ns) * fc[q] + ns * fo[Q19::opp[q]];
        });
    }
    if (!solid) {
#pragma unroll
        for (int q = 0; q < 19; ++q) fo[q] = fc[q];
    }
}

// Velocity of the stored populations (Macros.cu:44-60): the AD coupling and the boundary updates.
template <bool FAST = false, class C>
__device__ __forceinline__ void post_moments(const C (&fo)[19], C &r2, C &ux2, C &uy2, C &uz2) {
    C r = C(0), vx = C(0), vy = C(0), vz = C(0);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        r += fo[q];
        vx = axpy1<Q19::cx[q]>(vx, fo[q]);
        vy = axpy1<Q19::cy[q]>(vy, fo[q]);
        vz = axpy1<Q19::cz[q]>(vz, fo[q]);
    });
    const C inv2 = (r > C(0)) ? recip<FAST>(r) : C(0);
    r2 = r;
    ux2 = vx * inv2;
    uy2 = vy * inv2;
    uz2 = vz * inv2;
}

// feq(rho0, u_BC): next step's VelocityDirichletEq inlet populations (VelocityDirichlet.cu:121-127)
template <int q, class C> __device__ __forceinline__ C inlet_eq(const CoupledParams<C> &prm) {
    return feq_bgk<q>(prm.rho0, prm.uin_x, prm.uin_y, prm.uin_z, C(1.5) * prm.uin2);
}

// Next step's VelocityDirichletNEQ value of (0, y, z) from the (1, y, z) populations f and their
// moments (VelocityDirichletNEQ.cu:81-95).
template <int q, class C>
__device__ __forceinline__ C inlet_neq(C f, C r2, C ux2, C uy2, C uz2, const CoupledParams<C> &prm) {
    const C u2N = C(1.5) * (ux2 * ux2 + uy2 * uy2 + uz2 * uz2);
    const C u2B = C(1.5) * prm.uin2;
    const C feqN = feq_bgk<q>(r2, ux2, uy2, uz2, u2N);
    const C feqB = feq_bgk<q>(r2, prm.uin_x, prm.uin_y, prm.uin_z, u2B);
    return feqB + (f - feqN);
}

// PressureDirichlet value of (nx-1, y, z) from the velocity of (nx-2, y, z) (PressureDirichlet.cu:81-95).
template <int q, class C> __device__ __forceinline__ C outlet_pressure(C ux2, C uy2, C uz2, const CoupledParams<C> &prm) {
    const C u2o = C(1.5) * (ux2 * ux2 + uy2 * uy2 + uz2 * uz2);
    const C eu = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(ux2, uy2, uz2);
    return w19<C>(q) * prm.rho_out * (C(1) + C(3) * eu + C(4.5) * eu * eu - u2o);
}

// LDF sorption of one region cell (LinearLDF.cu:27-47, LangmuirLDF.cu:23-44): the new adsorbed load
// and the source term S = -dq, from the previous step's concentration.
template <bool FAST = false, class C>
__device__ __forceinline__ C ldf_cell(C phi_lag, C q_old, const CoupledParams<C> &prm, C &q_new) {
    C c = phi_lag;
    if (c < prm.min_conc) c = prm.min_conc;
    const C ce = c - prm.min_conc;
    C qeq;
    if (prm.isotherm == FMX_ISO_LANGMUIR) {
        const C bc = prm.iso_b * ce;
        qeq = prm.iso_a * (FAST ? bc * recip<FAST>(C(1) + bc) : bc / (C(1) + bc));
    } else {
        qeq = prm.iso_a * ce;
    }
    const C dq = prm.k_ldf * (qeq - q_old);
    q_new = q_old + dq;
    return -dq;
}

// D3Q7 collision with source (ADBGKCollision.cu:90-127): returns the clamped concentration.
template <class C>
__device__ __forceinline__ C ad_collide_cell(const C (&gi)[7], C src_term, C om, C ax, C ay, C az, C min_conc,
                                             C (&go)[7]) {
    C phi = C(0);
#pragma unroll
    for (int q = 0; q < 7; ++q) phi += gi[q];
    if (phi < min_conc) phi = min_conc;
    static_for<0, 7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const C eu = dot3<Q7::cx[q], Q7::cy[q], Q7::cz[q]>(ax, ay, az);
        const C geq = w7<C>(q) * phi * (C(1) + C(4) * eu);
        go[q] = gi[q] + om * (geq - gi[q]) + w7<C>(q) * src_term;
    });
    return phi;
}

}  // namespace fmx
