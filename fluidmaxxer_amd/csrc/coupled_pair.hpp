// coupled_pair.hpp -- the fused coupled step with two neighbouring cells (x, x+1) per thread.
//
// Same per-cell arithmetic as k_coupled (coupled_cell.hpp), so the same bits; what changes is
// the memory access.  Every population plane is read and written as aligned pairs -- 4 bytes per
// lane with FP16/BF16 storage, 8 bytes with FP32 -- which the memory pipeline moves faster than
// single 2- or 4-byte elements (26-plane copy without arithmetic on MI355X: 3.9, 4.3 and
// 4.8 TB/s for 2-, 4- and 8-byte elements).  A pull along +-x needs the element one cell over:
// the neighbouring lane's pair holds it and hands it over with a lane shuffle; only the first and
// last lane of a wave load that element themselves.
//
// Layout assumptions (checked on the host): nx % 128 == 0, so a wave covers 128 cells of one row
// and rows start at wave boundaries; storage FP32, FP16 or BF16.
#pragma once

#include "coupled_cell.hpp"

namespace fmx {

template <class S> struct PairWord;
template <> struct PairWord<float> {
    using U = unsigned long long;
};
template <> struct PairWord<_Float16> {
    using U = unsigned;
};
template <> struct PairWord<__bf16> {
    using U = unsigned;
};
template <> struct PairWord<std::uint32_t> {
    using U = unsigned long long;
};

// elements elem, elem+1 of a plane (elem even, so the pair is aligned)
template <class S> __device__ __forceinline__ void ld_pair(const S *base, unsigned elem, S &a, S &b) {
    using U = typename PairWord<S>::U;
    const U v = ld_b(reinterpret_cast<const U *>(base), elem * unsigned(sizeof(S)));
    S t[2];
    __builtin_memcpy(t, &v, sizeof(U));
    a = t[0];
    b = t[1];
}
template <class S> __device__ __forceinline__ void st_pair(S *base, unsigned elem, S a, S b) {
    using U = typename PairWord<S>::U;
    const S t[2] = {a, b};
    U v;
    __builtin_memcpy(&v, t, sizeof(U));
    st_b_nt(reinterpret_cast<U *>(base), elem * unsigned(sizeof(S)), v);
}
template <class S> __device__ __forceinline__ S ld_one(const S *base, unsigned elem) {
    return ld_b(base, elem * unsigned(sizeof(S)));
}

// one storage element through a 32-bit lane shuffle
template <class S> __device__ __forceinline__ int to_lane_word(S v) {
    if constexpr (sizeof(S) == 4) {
        int i;
        __builtin_memcpy(&i, &v, 4);
        return i;
    } else {
        unsigned short u;
        __builtin_memcpy(&u, &v, 2);
        return int(u);
    }
}
template <class S> __device__ __forceinline__ S from_lane_word(int i) {
    S v;
    if constexpr (sizeof(S) == 4) {
        __builtin_memcpy(&v, &i, 4);
    } else {
        const unsigned short u = static_cast<unsigned short>(i);
        __builtin_memcpy(&v, &u, 2);
    }
    return v;
}

// The pulled pair of plane `pl` for cells (x0, x0+1) along (cx, cy, cz): the source cell's element,
// or the cell's own element where the source lies outside the box (keep-local).
template <int cx, class S>
__device__ __forceinline__ void pull_pair(const S *pl, bool row_ok, unsigned src_row, unsigned own_row, int x0, int nx,
                                          int lane, S &a, S &b) {
    if (!row_ok || cx == 0) {
        ld_pair(pl, (row_ok ? src_row : own_row) + unsigned(x0), a, b);
    } else if (cx == 1) {  // cell x0+1 pulls x0 (own lane), cell x0 pulls x0-1 (previous lane)
        S e0, e1;
        ld_pair(pl, src_row + unsigned(x0), e0, e1);
        b = e0;
        a = from_lane_word<S>(__shfl_up(to_lane_word(e1), 1, 64));
        if (lane == 0) a = ld_one(pl, x0 == 0 ? own_row : src_row + unsigned(x0) - 1u);
    } else {  // cell x0 pulls x0+1 (own lane), cell x0+1 pulls x0+2 (next lane)
        S e0, e1;
        ld_pair(pl, src_row + unsigned(x0), e0, e1);
        a = e1;
        b = from_lane_word<S>(__shfl_down(to_lane_word(e0), 1, 64));
        if (lane == 63) b = ld_one(pl, x0 + 2 == nx ? own_row + unsigned(nx - 1) : src_row + unsigned(x0) + 2u);
    }
}

template <int P, int OP>
__global__ void __launch_bounds__(256) k_coupled_pair(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm,
                                                      i64 step, i64 acc_row) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    int bx, by, bz;
    block_coords(prm, bx, by, bz);
    const int x0 = 2 * (bx * prm.block_x + int(threadIdx.x));
    const int y = by * prm.block_y + int(threadIdx.y);
    const int z = bz;
    if (y >= prm.ny) return;  // whole waves: a wave lies in one row
    const int lane = int(threadIdx.x) & 63;
    const int nx = prm.nx, ny = prm.ny, nz = prm.nz;
    const i64 n = prm.n;

    int ym = y - 1, yp = y + 1, zm = z - 1, zp = z + 1;
    bool ym_ok = ym >= 0, yp_ok = yp < ny, zm_ok = zm >= 0, zp_ok = zp < nz;
    if (prm.py) {
        if (!ym_ok) ym += ny;
        if (!yp_ok) yp -= ny;
        ym_ok = yp_ok = true;
    }
    if (prm.pz) {
        if (!zm_ok) zm += nz;
        if (!zp_ok) zp -= nz;
        zm_ok = zp_ok = true;
    }
    const unsigned ys[3] = {unsigned(ym), unsigned(y), unsigned(yp)};
    const unsigned zs[3] = {unsigned(zm), unsigned(z), unsigned(zp)};
    const bool yok[3] = {ym_ok, true, yp_ok};
    const bool zok[3] = {zm_ok, true, zp_ok};
    const unsigned own_row = unsigned(nx) * (unsigned(y) + unsigned(ny) * unsigned(z));
    // first cell of row (y + dy, z + dz)
    auto row_of = [&](int dy, int dz) -> unsigned { return unsigned(nx) * (ys[1 + dy] + unsigned(ny) * zs[1 + dz]); };
    const unsigned c0 = own_row + unsigned(x0), c1 = c0 + 1u;

    std::uint32_t lk0, lk1;
    ld_pair(p.links, c0, lk0, lk1);
    const bool solid0 = (OP == FMX_NS_BGK) && (lk0 & FMX_LINK_SOLID);
    const bool solid1 = (OP == FMX_NS_BGK) && (lk1 & FMX_LINK_SOLID);

    // ---------------------------------------------------------------- NS pulls
    C fo0[19], fo1[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
        const bool row_ok = yok[1 - cy] && zok[1 - cz];
        S a, b;
        pull_pair<cx>(plane(p.f, q, n), row_ok, row_ok ? row_of(-cy, -cz) : own_row, own_row, x0, nx, lane, a, b);
        if constexpr (OP == FMX_NS_BGK && q != 0) {  // bounce-back: the cell's own opposite population
            const bool bb0 = (lk0 >> q) & 1u, bb1 = (lk1 >> q) & 1u;
            if (bb0 || bb1) {
                S r0, r1;
                ld_pair(plane(p.f, Q19::opp[q], n), c0, r0, r1);
                if (bb0) a = r0;
                if (bb1) b = r1;
            }
        }
        fo0[q] = decode<P>(a, w19<C>(q), prm.f_ref);
        fo1[q] = decode<P>(b, w19<C>(q), prm.f_ref);
    });

    // ---------------------------------------------------------------- AD pulls
    const C phi_pre = (prm.ad_inlet == 2) ? C(p.inlet_series[step]) : C(0);
    C gi0[7], gi1[7];
    static_for<0, 7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
        const bool row_ok = yok[1 - cy] && zok[1 - cz];
        const unsigned src_row = row_ok ? row_of(-cy, -cz) : own_row;
        S a, b;
        pull_pair<cx>(plane(p.g, q, n), row_ok, src_row, own_row, x0, nx, lane, a, b);
        auto value = [&](int xc, S raw) -> C {
            const bool xin = (cx == 1) ? (xc > 0) : ((cx == -1) ? (xc < nx - 1) : true);
            const bool in = row_ok && xin;
            const int sx = in ? xc - cx : xc;
            if (prm.ad_inlet == 2 && sx == 0) return phi_pre * w7<C>(q);
            if constexpr (IsHalf<P>::value) {
                // deviation about the source cell's lagged concentration (x - 1 cell's on a Neumann outlet)
                const unsigned s = in ? src_row + unsigned(xc - cx) : own_row + unsigned(xc);
                const unsigned sref = (prm.ad_outlet && sx == nx - 1) ? s - 1u : s;
                return decode_g<P>(raw, w7<C>(q), p.phi[sref]);
            } else {
                return C(raw);
            }
        };
        gi0[q] = value(x0, a);
        gi1[q] = value(x0 + 1, b);
    });
    const bool sorbs0 = prm.isotherm && (lk0 & FMX_LINK_REGION);
    const bool sorbs1 = prm.isotherm && (lk1 & FMX_LINK_REGION);
    C phi_lag0 = C(0), phi_lag1 = C(0), q_old0 = C(0), q_old1 = C(0);
    if (sorbs0 || sorbs1) {
        ld_pair(p.phi, c0, phi_lag0, phi_lag1);
        ld_pair(p.q, c0, q_old0, q_old1);
    }

    // ------------------------------------------ NS collide (in place) and moments
    C fld0 = (OP == FMX_NS_POROUS) ? C(1) : C(0), fld1 = fld0;
    if (OP != FMX_NS_BGK && p.field) ld_pair(p.field, c0, fld0, fld1);
    constexpr bool FAST = IsHalf<P>::value;
    ns_collide_cell<OP, FAST>(fo0, solid0, fld0, prm);
    ns_collide_cell<OP, FAST>(fo1, solid1, fld1, prm);
    C r2_0, ux0, uy0, uz0, r2_1, ux1, uy1, uz1;
    post_moments<FAST>(fo0, r2_0, ux0, uy0, uz0);
    post_moments<FAST>(fo1, r2_1, ux1, uy1, uz1);

    // ------------------------------------------------------------- NS stores
    const bool in0 = prm.ns_inlet && x0 == 0;        // cell x0 is on the inlet plane
    const bool out1 = prm.ns_outlet && x0 + 2 == nx;  // cell x0+1 is on the outlet plane
    if (in0 && p.inlet_raw) {
        const unsigned np = unsigned(ny) * unsigned(nz), pc = unsigned(y) + unsigned(ny) * unsigned(z);
#pragma unroll
        for (int q = 0; q < 19; ++q) p.inlet_raw[i64(q) * np + pc] = fo0[q];
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        C v0 = fo0[q], v1 = fo1[q];
        if (in0) v0 = (prm.ns_inlet == 2) ? inlet_eq<q>(prm) : inlet_neq<q>(fo1[q], r2_1, ux1, uy1, uz1, prm);
        if (out1) v1 = outlet_pressure<q>(ux0, uy0, uz0, prm);
        st_pair(plane(p.fn, q, n), c0, encode<P>(v0, C(Q19::inv_w[q]), prm.f_ref),
                encode<P>(v1, C(Q19::inv_w[q]), prm.f_ref));
    });

    // ------------------------------------------------------ sorption and AD collision
    C src0 = C(0), src1 = C(0);
    if (sorbs0) {
        C qn;
        src0 = ldf_cell<FAST>(phi_lag0, q_old0, prm, qn);
        p.q_new[c0] = qn;
    }
    if (sorbs1) {
        C qn;
        src1 = ldf_cell<FAST>(phi_lag1, q_old1, prm, qn);
        p.q_new[c1] = qn;
    }
    C om0 = prm.omega_ad, om1 = prm.omega_ad;
    if (p.tau_field) {
        C t0, t1;
        ld_pair(p.tau_field, c0, t0, t1);
        om0 = recip<FAST>(t0);
        om1 = recip<FAST>(t1);
    }
    const bool cpl = prm.coupling != 0;
    C go0[7], go1[7];
    const C phi0 = ad_collide_cell(gi0, src0, om0, cpl ? ux0 : C(0), cpl ? uy0 : C(0), cpl ? uz0 : C(0),
                                   prm.min_conc, go0);
    const C phi1 = ad_collide_cell(gi1, src1, om1, cpl ? ux1 : C(0), cpl ? uy1 : C(0), cpl ? uz1 : C(0),
                                   prm.min_conc, go1);
    st_pair(p.phi_new, c0, phi0, phi1);

    // ------------------------------------------------------------- AD stores
    const bool ain0 = prm.ad_inlet == 1 && x0 == 0;     // ADDirichlet inlet value
    const bool aout1 = prm.ad_outlet && x0 + 2 == nx;   // Neumann outlet: copy of the x0 cell's populations
    const C is0 = FAST ? recip<FAST>(g_scale(phi0)) : C(0), is1 = FAST ? recip<FAST>(g_scale(phi1)) : C(0);
    static_for<0, 7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const C iw = C(Q7::inv_w[q]);
        const S e0 = encode_g<P>(ain0 ? prm.phi_in * w7<C>(q) : go0[q], iw, phi0, is0);
        const S e1 = aout1 ? encode_g<P>(go0[q], iw, phi0, is0) : encode_g<P>(go1[q], iw, phi1, is1);
        st_pair(plane(p.gn, q, n), c0, e0, e1);
    });

    // ------------------------------------------------- outlet breakthrough sum
    if (p.outlet_acc && x0 + 2 == nx) {
        if (acc_row < 0) {  // a row of the outlet ring, summed later in a fixed order
            p.outlet_acc[unsigned(y) + unsigned(ny) * unsigned(z)] = double(phi1);
        } else {
            const int slot = int((unsigned(y) + unsigned(ny) * unsigned(z)) % FMX_OUTLET_SLOTS);
            unsafeAtomicAdd(p.outlet_acc + acc_row * (FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE) + slot * FMX_OUTLET_STRIDE,
                            double(phi1));
        }
    }
}
// The pair kernel's launch shape for a row of nx cells: nx/2 threads per row in blocks of 256.
inline bool pair_shape(int nx, int ny, int &block_x, int &block_y, int &gx, int &gy) {
    if (nx < 128 || nx % 128 != 0) return false;
    const int tpr = nx / 2;
    block_x = (tpr % 256 == 0) ? 256 : ((tpr % 128 == 0) ? 128 : 64);
    block_y = 256 / block_x;
    gx = tpr / block_x;
    gy = (ny + block_y - 1) / block_y;
    return true;
}

template <int P>
int launch_coupled_pair(int ns_op, const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm,
                        i64 step, i64 acc_row, hipStream_t s) {
    CoupledParams<typename Storage<P>::C> pp = prm;
    if (!pair_shape(prm.nx, prm.ny, pp.block_x, pp.block_y, pp.gx, pp.gy))
        return set_error(FMX_ERR_ARG, "k_coupled_pair: nx = %d is not a multiple of 128", prm.nx);
    pp.gz = prm.nz;
    pp.tile_y = pp.tile_z = 0;
    const dim3 grid(coupled_blocks(pp.gx, pp.gy, pp.gz)), block(pp.block_x, pp.block_y, 1);
    switch (ns_op) {
        case FMX_NS_BGK:
            hipLaunchKernelGGL((k_coupled_pair<P, FMX_NS_BGK>), grid, block, 0, s, ptr, pp, step, acc_row);
            break;
        case FMX_NS_POROUS:
            hipLaunchKernelGGL((k_coupled_pair<P, FMX_NS_POROUS>), grid, block, 0, s, ptr, pp, step, acc_row);
            break;
        case FMX_NS_PBB:
            hipLaunchKernelGGL((k_coupled_pair<P, FMX_NS_PBB>), grid, block, 0, s, ptr, pp, step, acc_row);
            break;
        default: return set_error(FMX_ERR_ARG, "fmx_coupled: unknown ns_op %d", ns_op);
    }
    return after_launch("k_coupled_pair");
}

}  // namespace fmx
