// coupled_kernel.hpp -- the fused coupled step kernel (one cell per thread) and its launcher.
//
// Included by coupled.hip (FP32 / FP64 storage: IEEE division, no FMA contraction, bit-identical
// to the serial restatement of the reference) and coupled_half.hip (FP16 / BF16 storage, compiled
// with FMA contraction and the hardware reciprocal: its accuracy is set by the 16-bit storage).
// Each precision is instantiated in exactly one of the two files.
#pragma once

#include "coupled_cell.hpp"

namespace fmx {

// Byte offsets are 32-bit (a scalar plane base plus a vector offset) unless four population planes
// exceed 4 GiB (WIDE: 64-bit offsets).  FP16 and BF16 storage are built for 7 waves per SIMD (72
// VGPRs): with every pull and concentration gather issued as one run of loads, 8 waves (64 VGPRs)
// spill and run 25% slower, 6 waves 1.5% slower (512^3).  FP64 for 4 (128 VGPRs instead of 130, at
// most 20 bytes of spill; 1% faster than 3); FP32 keeps the compiler's choice.
// FZ: frozen solid cells may be present (FMX_FREEZE=1); without it the frozen link bit is not
// looked at, and the pulls are not wrapped in a branch that is decided lane by lane.  (Ignoring the bit
// is always correct: a frozen cell's two buffers hold its two alternating states, which the full path
// reproduces.)
// RC: row-compacted NS planes (RowSeg; implies FZ): p.f / p.fn hold only the active cells; a frozen
// lane stores no NS population and hands zero velocity to the AD update.
// ADONLY: the AD half of the split step (XA = SplitArgs): no NS work; the velocity of an active cell
// comes from the NS kernel's u, a frozen cell's is zero.
template <int P, int OP, bool WIDE, bool FZ, bool RC, bool ADONLY = false, class XA = RcArgs>
__device__ __forceinline__ void coupled_body(const CoupledPtrs<P> &p, const CoupledParams<typename Storage<P>::C> &prm,
                                             i64 step, i64 acc_row, const XA &rc) {
    static_assert(!RC || (FZ && OP == FMX_NS_BGK), "row-compacted storage is the frozen-solid BGK step");
    static_assert(!ADONLY || (FZ && !RC && OP == FMX_NS_BGK), "the AD half of the split step");
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    using Off = std::conditional_t<WIDE, i64, unsigned>;
    constexpr unsigned SB = sizeof(S);
    constexpr unsigned CB = sizeof(C);
    int bx, by, bz;
    block_coords(prm, bx, by, bz);
    const int x = bx * prm.block_x + int(threadIdx.x);
    // block_x is a multiple of 64, so every wave lies in one row and y is wave-uniform
    const int y = __builtin_amdgcn_readfirstlane(by * prm.block_y + int(threadIdx.y));
    const int z = bz;
    if (x >= prm.nx || y >= prm.ny) return;
    const int nx = prm.nx, ny = prm.ny, nz = prm.nz;
    const i64 n = prm.n;
    const Off cell = Off(x) + Off(nx) * (Off(y) + Off(ny) * Off(z));
    const Off cell_b = cell * SB;
    // the AD half of the split step takes the region and frozen bits from the row tables instead
    const std::uint32_t lk = ADONLY ? 0u : ld_b(p.links, cell * 4u);
    const bool solid = (OP == FMX_NS_BGK) && (lk & FMX_LINK_SOLID);
    // a frozen solid cell alternates between the two states held in f and f_new and has zero
    // post-collision velocity in both: no NS loads, collision or stores
    bool frozen = FZ && (OP == FMX_NS_BGK) && (lk & FMX_LINK_FROZEN);
    bool in_region = (lk & FMX_LINK_REGION) != 0u;

    // Source of the pull along (cx, cy, cz): the cell - c, or the cell itself when that lies outside
    // the box (x is never periodic; y / z wrap when periodic).  The row part is wave-uniform: a scalar
    // byte delta per direction, applied to this cell's offset; only x = 0 / nx - 1 lanes keep local
    // for the x-directions.
    const bool xm_ok = x > 0, xp_ok = x < nx - 1;
    // rows y - cy and z - cz (index cy + 1, cz + 1) relative to this one, and whether they exist
    int dy[3] = {1, 0, -1}, dz[3] = {1, 0, -1};
    bool yin[3] = {y + 1 < ny, true, y > 0}, zin[3] = {z + 1 < nz, true, z > 0};
    if (prm.py) {
        if (!yin[0]) dy[0] = 1 - ny;
        if (!yin[2]) dy[2] = ny - 1;
        yin[0] = yin[2] = true;
    }
    if (prm.pz) {
        if (!zin[0]) dz[0] = 1 - nz;
        if (!zin[2]) dz[2] = nz - 1;
        zin[0] = zin[2] = true;
    }
    // byte delta of row (y - cy, z - cz) for cy, cz in {-1, 0, 1}, computed once (wave-uniform) so the
    // pulls below form one run of loads with no branches between them
    Off rowd[3][3];
    bool rowok[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            rowd[a][b] = (Off(dy[a]) * Off(nx) + Off(dz[b]) * (Off(nx) * Off(ny))) * Off(SB);
            rowok[a][b] = yin[a] && zin[b];
        }
    auto src_b = [&](int cx, int cy, int cz) -> Off {  // byte offset of the source cell within a plane
        const bool ok = rowok[cy + 1][cz + 1];
        const Off d = ok ? rowd[cy + 1][cz + 1] - Off(cx) * Off(SB) : Off(0);
        if (cx == 0) return cell_b + d;
        return cell_b + (((cx == 1) ? xm_ok : xp_ok) ? d : Off(0));
    };
    // Row-compacted NS planes: the compact index of x in row (cy, cz) counts that row's active cells
    // before x (mbcnt of the segment's mask, which this wave's 64 lanes cover exactly, plus the segment
    // base); x - 1 is one less and x + 1 adds the row's bit at x, also across the segment's ends.  A
    // pull whose source is solid never uses its index (the bounce-back offset replaces it), so only
    // the indices of active sources matter.  Keep-local pulls use this cell's own index.
    Off rsl[3][3] = {};
    unsigned rbit[3][3] = {};
    Off uidx = 0;  // ADONLY: compact index of this cell (of the next active one in its row if frozen)
    if constexpr (ADONLY) {
        const int seg = __builtin_amdgcn_readfirstlane(x) >> 6;
        const i64 ent = (i64(y) + i64(ny) * z) * rc.nseg + seg;
        const RowSeg &e = rc.rs[ent];
        const unsigned lo = __builtin_amdgcn_readfirstlane(e.lo), hi = __builtin_amdgcn_readfirstlane(e.hi);
        const unsigned base = __builtin_amdgcn_readfirstlane(e.base);
        const unsigned long long rg = rc.region[ent];
        const unsigned rlo = __builtin_amdgcn_readfirstlane(unsigned(rg)), rhi = __builtin_amdgcn_readfirstlane(unsigned(rg >> 32));
        const unsigned lane = unsigned(x) & 63u;
        uidx = Off(__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, base)));
        frozen = (((lane < 32u ? (lo >> lane) : (hi >> (lane - 32u))) & 1u) == 0u);
        in_region = (((lane < 32u ? (rlo >> lane) : (rhi >> (lane - 32u))) & 1u) != 0u);
    }
    if constexpr (RC) {
        const int seg = __builtin_amdgcn_readfirstlane(x) >> 6;
        const unsigned lane = unsigned(x) & 63u;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int yy = rowok[a][b] ? y + dy[a] : y, zz = rowok[a][b] ? z + dz[b] : z;
                const RowSeg &e = rc.rs[(i64(yy) + i64(ny) * zz) * rc.nseg + seg];
                const unsigned lo = __builtin_amdgcn_readfirstlane(e.lo), hi = __builtin_amdgcn_readfirstlane(e.hi);
                const unsigned base = __builtin_amdgcn_readfirstlane(e.base);
                rsl[a][b] = Off(__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, base)));
                rbit[a][b] = ((lane < 32u ? (lo >> lane) : (hi >> (lane - 32u))) & 1u);
            }
    }
    const Off ncell = RC ? rsl[1][1] : cell;  // this cell's index in the NS planes
    const Off ncell_b = ncell * SB;
    const i64 nn = RC ? rc.nc : n;  // NS plane stride
    const Off nplane_b = Off(nn) * SB;
    auto src_n = [&](int cx, int cy, int cz) -> Off {  // byte offset of the NS source cell within a plane
        if constexpr (!RC) {
            return src_b(cx, cy, cz);
        } else {
            bool ok = rowok[cy + 1][cz + 1];
            Off si = rsl[cy + 1][cz + 1];
            if (cx == 1) {
                ok = ok && xm_ok;
                si = si - Off(1);
            } else if (cx == -1) {
                ok = ok && xp_ok;
                si = si + Off(rbit[cy + 1][cz + 1]);
            }
            return (ok ? si : ncell) * Off(SB);
        }
    };

    // ---------------------------------------------------------------- NS pull
    // With bounce-back links, direction q and its opposite are addressed from the lower of the two
    // planes, so the bounce-back pull (this cell along opp(q)) and the streaming pull differ only in
    // the offset, chosen by the link bit.
    // RC: a frozen lane loads too (no branch around the run of loads): every link of a solid cell
    // bounces back, so its offsets are its compact index -- the next active cell of its row, whose
    // lines the wave reads anyway -- and its results are discarded (zero velocity, no stores).
    C fo[19];
    if (!ADONLY && (RC || !frozen)) static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
        constexpr int o = Q19::opp[q];
        constexpr bool BB = OP == FMX_NS_BGK && q != 0;
        constexpr int bq = BB ? (q < o ? q : o) : q;
        Off off = src_n(cx, cy, cz) + Off(q - bq) * nplane_b;
        if constexpr (BB) {
            const Off boff = ncell_b + Off(o - bq) * nplane_b;
            const Off m = Off(int(lk << (31 - q)) >> 31);  // all ones where the q link bounces back
            off = (boff & m) | (off & ~m);
        }
        fo[q] = decode<P>(ld_b(plane(p.f, bq, nn), off), w19<C>(q), prm.f_ref);
    });

    // ------------------------------------------- AD pull and sorption inputs
    // (issued with the NS pulls; nothing below writes what they read).  Every lane loads from an
    // address in the box and the per-lane choices (series substitution, sorbing cell) are selects,
    // not branches: a lane-dependent branch around a load closes a basic block, and the loads and
    // their uses inside it then wait for the memory round trip one direction at a time.
    const C phi_pre = (prm.ad_inlet == 2) ? C(p.inlet_series[step]) : C(0);
    C gi[7];
    static_for<0, 7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
        const Off s = src_b(cx, cy, cz);
        // x of the source (or keep-local) cell: x - cx when in the domain, else x
        const int sx = (cx == 1) ? (x > 0 ? x - 1 : x) : ((cx == -1) ? (x < nx - 1 ? x + 1 : x) : x);
        const S raw = ld_b(plane(p.g, q, n), s);
        C v;
        if constexpr (IsHalf<P>::value) {
            // deviation about the source cell's lagged concentration; a Neumann outlet cell holds the
            // populations of its x - 1 neighbour, stored about that neighbour's concentration
            Off sref = s * (CB / SB);
            if (prm.ad_outlet && sx == nx - 1) sref -= CB;
            v = decode_g<P>(raw, w7<C>(q), ld_b(p.phi, sref));
        } else {
            v = raw;
        }
        gi[q] = (prm.ad_inlet == 2 && sx == 0) ? phi_pre * w7<C>(q) : v;
    });
    const bool sorbs = prm.isotherm && in_region;
    // read by every lane (the sorption region is most of a pack's waves); used only where sorbs
    const C phi_lag = ld_b(p.phi, cell * CB);
    const C q_old = ld_b(p.q, cell * CB);

    // ------------------------------------------ NS collide (in place) and moments
    C r2 = C(0), ux2 = C(0), uy2 = C(0), uz2 = C(0);
    if constexpr (ADONLY) {
        if (prm.coupling) {  // every lane loads (its index is in the box of active cells); frozen: zero
            // (ux, uy, uz) of a cell side by side: one wide load
            C uv[3];
            __builtin_memcpy(uv, static_cast<const C *>(rc.u) + std::size_t(uidx) * 3u, sizeof(uv));
            ux2 = frozen ? C(0) : uv[0];
            uy2 = frozen ? C(0) : uv[1];
            uz2 = frozen ? C(0) : uv[2];
        }
    } else if (RC || !frozen) {
        const C fld = p.field ? ld_b(p.field, cell * CB) : ((OP == FMX_NS_POROUS) ? C(1) : C(0));
        ns_collide_cell<OP, IsHalf<P>::value>(fo, solid, fld, prm);
        post_moments<IsHalf<P>::value>(fo, r2, ux2, uy2, uz2);
        if constexpr (RC) {
            r2 = frozen ? C(0) : r2;
            ux2 = frozen ? C(0) : ux2;
            uy2 = frozen ? C(0) : uy2;
            uz2 = frozen ? C(0) : uz2;
        }
    }

    // ------------------------------------------------------------- NS stores
    const bool inlet_cell = prm.ns_inlet && x == 0;
    const bool outlet_cell = prm.ns_outlet && x == nx - 1;
    if constexpr (!ADONLY) {
    if (inlet_cell && p.inlet_raw) {  // the x = 0 populations as the reference holds them (last step of a run)
        const unsigned np = unsigned(ny) * unsigned(nz), pc = unsigned(y) + unsigned(ny) * unsigned(z);
#pragma unroll
        for (int q = 0; q < 19; ++q) p.inlet_raw[i64(q) * np + pc] = fo[q];
    }
    if (prm.lane_bc) {
        // Boundary values computed by the boundary lane itself, from its x +- 1 neighbour's values
        // handed over by lane shuffles (the neighbour is the adjacent lane of the same wave), so every
        // population plane takes one full-wave store:
        //   x = 0:      next step's inlet f: feq(rho0, u_BC), or NEQ from the x = 1 populations
        //   x = nx-1:   PressureDirichlet feq(rho_out, u) with u of the x = nx-2 cell
        C um_x = C(0), um_y = C(0), um_z = C(0);
        if (prm.ns_outlet) {
            um_x = __shfl_up(ux2, 1, 64);
            um_y = __shfl_up(uy2, 1, 64);
            um_z = __shfl_up(uz2, 1, 64);
        }
        C fp[19], rp = C(0), up_x = C(0), up_y = C(0), up_z = C(0);
        if (prm.ns_inlet == 1) {
#pragma unroll
            for (int q = 0; q < 19; ++q) fp[q] = __shfl_down(fo[q], 1, 64);
            rp = __shfl_down(r2, 1, 64);
            up_x = __shfl_down(ux2, 1, 64);
            up_y = __shfl_down(uy2, 1, 64);
            up_z = __shfl_down(uz2, 1, 64);
        }
        if (!frozen) static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            C v = fo[q];
            if (inlet_cell) v = (prm.ns_inlet == 2) ? inlet_eq<q>(prm) : inlet_neq<q>(fp[q], rp, up_x, up_y, up_z, prm);
            else if (outlet_cell) v = outlet_pressure<q>(um_x, um_y, um_z, prm);
            st_plane_nt(plane(p.fn, q, nn), ncell_b, encode<P>(v, C(Q19::inv_w[q]), prm.f_ref));
        });
    } else {
        // the x = 1 thread writes the NEQ inlet of x = 0, the x = nx-2 thread the pressure outlet
        if (inlet_cell) {
            if (prm.ns_inlet == 2) {  // next step's VelocityDirichletEq: f(0,y,z) = feq(rho0, u_BC)
                static_for<0, 19>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    st_plane_nt(plane(p.fn, q, nn), ncell_b, encode<P>(inlet_eq<q>(prm), C(Q19::inv_w[q]), prm.f_ref));
                });
            }
        } else if (!outlet_cell && !frozen) {
#pragma unroll
            for (int q = 0; q < 19; ++q)
                st_plane_nt(plane(p.fn, q, nn), ncell_b, encode<P>(fo[q], C(Q19::inv_w[q]), prm.f_ref));
        }
        if (prm.ns_inlet == 1 && x == 1) {  // next step's NEQ update of (0, y, z)
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                st_b(plane(p.fn, q, nn), ncell_b - SB,  // x = 0 and 1 are never frozen: adjacent in compact rows too
                     encode<P>(inlet_neq<q>(fo[q], r2, ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
            });
        }
        if (prm.ns_outlet && x == nx - 2) {  // PressureDirichlet of (nx-1, y, z)
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                st_b(plane(p.fn, q, nn), ncell_b + SB,  // nor x = nx-2 and nx-1
                     encode<P>(outlet_pressure<q>(ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
            });
        }
    }
    }  // !ADONLY

    // ------------------------------------------------------ sorption source and AD collision
    C src_term = C(0);
    if (sorbs) {
        C q_new;
        src_term = ldf_cell<IsHalf<P>::value>(phi_lag, q_old, prm, q_new);
        st_b(p.q_new, cell * CB, q_new);
    }
    const C om = p.tau_field ? recip<IsHalf<P>::value>(ld_b(p.tau_field, cell * CB)) : prm.omega_ad;
    const C ax = prm.coupling ? ux2 : C(0), ay = prm.coupling ? uy2 : C(0), az = prm.coupling ? uz2 : C(0);
    C go[7];
    const C phi = ad_collide_cell(gi, src_term, om, ax, ay, az, prm.min_conc, go);
    st_b(p.phi_new, cell * CB, phi);

    // ------------------------------------------------------------- AD stores
    // (half storage: deviation about this cell's new concentration, which the readers load as phi; the
    // Neumann outlet cell holds its x - 1 neighbour's populations, about that neighbour's concentration)
    const bool ad_in_cell = prm.ad_inlet == 1 && x == 0;
    const bool ad_out_cell = prm.ad_outlet && x == nx - 1;
    constexpr bool FAST = IsHalf<P>::value;
    if (prm.lane_bc) {
        C gm[7], phim = C(0);
        if (prm.ad_outlet) {
#pragma unroll
            for (int q = 0; q < 7; ++q) gm[q] = __shfl_up(go[q], 1, 64);
            phim = __shfl_up(phi, 1, 64);
        }
        const C ref = ad_out_cell ? phim : phi;
        const C inv_s = FAST ? recip<FAST>(g_scale(ref)) : C(0);
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            C v = go[q];
            if (ad_in_cell) v = prm.phi_in * w7<C>(q);
            else if (ad_out_cell) v = gm[q];
            st_plane_nt(plane(p.gn, q, n), cell_b, encode_g<P>(v, C(Q7::inv_w[q]), ref, inv_s));
        });
    } else {
        const C inv_s = FAST ? recip<FAST>(g_scale(phi)) : C(0);
        if (ad_in_cell) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                st_plane_nt(plane(p.gn, q, n), cell_b, encode_g<P>(prm.phi_in * w7<C>(q), C(Q7::inv_w[q]), phi, inv_s));
        } else if (!ad_out_cell) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                st_plane_nt(plane(p.gn, q, n), cell_b, encode_g<P>(go[q], C(Q7::inv_w[q]), phi, inv_s));
        }
        if (prm.ad_outlet && x == nx - 2) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                st_b(plane(p.gn, q, n), cell_b + SB, encode_g<P>(go[q], C(Q7::inv_w[q]), phi, inv_s));
        }
    }

    // ------------------------------------------------- outlet breakthrough sum
    if (p.outlet_acc && x == nx - 1) {
        if (acc_row < 0) {  // a row of the outlet ring, summed later in a fixed order
            p.outlet_acc[unsigned(y) + unsigned(ny) * unsigned(z)] = double(phi);
        } else {
            const int slot = int((unsigned(y) + unsigned(ny) * unsigned(z)) % FMX_OUTLET_SLOTS);
            unsafeAtomicAdd(p.outlet_acc + acc_row * (FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE) + slot * FMX_OUTLET_STRIDE,
                            double(phi));
        }
    }
}

constexpr int coupled_waves(int P) { return (P == FMX_F16 || P == FMX_BF16) ? 7 : (P == FMX_F64 ? 4 : 1); }

template <int P, int OP, bool WIDE, bool FZ>
__global__ void __launch_bounds__(256, coupled_waves(P))
    k_coupled(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, i64 step, i64 acc_row) {
    coupled_body<P, OP, WIDE, FZ, false>(p, prm, step, acc_row, RcArgs{});
}

template <int P, bool WIDE>
__global__ void __launch_bounds__(256, coupled_waves(P))
    k_coupled_rc(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, i64 step, i64 acc_row, RcArgs rc) {
    coupled_body<P, FMX_NS_BGK, WIDE, true, true>(p, prm, step, acc_row, rc);
}

// The NS work of active cell k (compact index) at (x, y, z) with link word lk, given the compact index
// of x in each neighbouring row (rsl; this row's is k), that row's active bit at x and whether the row
// exists (keep-local otherwise): the NS pulls, collision and stores of k_coupled_rc, with the inlet and
// outlet values computed by the boundary lanes from lane shuffles (the x = 0 / 1 and nx-2 / nx-1 cells
// are never frozen, so they are adjacent in the compact order too), and the cell's velocity stored for
// the AD half.
template <int P, bool WIDE>
__device__ __forceinline__ void ns_split_cell(const CoupledPtrs<P> &p, const CoupledParams<typename Storage<P>::C> &prm,
                                              const SplitArgs &sa, i64 k, int x, int y, int z, std::uint32_t lk,
                                              const std::conditional_t<WIDE, i64, unsigned> (&rsl)[3][3],
                                              const unsigned (&rbit)[3][3], const bool (&rowok)[3][3]) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    using Off = std::conditional_t<WIDE, i64, unsigned>;
    constexpr unsigned SB = sizeof(S);
    const int nx = prm.nx, ny = prm.ny, nz = prm.nz;
    const bool solid = lk & FMX_LINK_SOLID;
    const bool xm_ok = x > 0, xp_ok = x < nx - 1;
    const Off ncell_b = Off(k) * SB;
    const Off nplane_b = Off(sa.nc) * SB;
    auto src_n = [&](int cx, int cy, int cz) -> Off {
        bool ok = rowok[cy + 1][cz + 1];
        Off si = rsl[cy + 1][cz + 1];
        if (cx == 1) {
            ok = ok && xm_ok;
            si = si - Off(1);
        } else if (cx == -1) {
            ok = ok && xp_ok;
            si = si + Off(rbit[cy + 1][cz + 1]);
        }
        return (ok ? si : Off(k)) * Off(SB);
    };
    C fo[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
        constexpr int o = Q19::opp[q];
        constexpr bool BB = q != 0;
        constexpr int bq = BB ? (q < o ? q : o) : q;
        Off off = src_n(cx, cy, cz) + Off(q - bq) * nplane_b;
        if constexpr (BB) {
            const Off boff = ncell_b + Off(o - bq) * nplane_b;
            const Off m = Off(int(lk << (31 - q)) >> 31);
            off = (boff & m) | (off & ~m);
        }
        fo[q] = decode<P>(ld_b(plane(p.f, bq, sa.nc), off), w19<C>(q), prm.f_ref);
    });
    C r2 = C(0), ux2 = C(0), uy2 = C(0), uz2 = C(0);
    ns_collide_cell<FMX_NS_BGK, IsHalf<P>::value>(fo, solid, C(0), prm);
    post_moments<IsHalf<P>::value>(fo, r2, ux2, uy2, uz2);
    C *u = static_cast<C *>(sa.u) + 3 * k;  // one wide store per lane
    u[0] = ux2;
    u[1] = uy2;
    u[2] = uz2;
    const bool inlet_cell = prm.ns_inlet && x == 0;
    const bool outlet_cell = prm.ns_outlet && x == nx - 1;
    if (inlet_cell && p.inlet_raw) {
        const unsigned np = unsigned(ny) * unsigned(nz), pc = unsigned(y) + unsigned(ny) * unsigned(z);
#pragma unroll
        for (int q = 0; q < 19; ++q) p.inlet_raw[i64(q) * np + pc] = fo[q];
    }
    // Boundary lanes compute their own values from their neighbour's, handed over by lane shuffles:
    // x = 1 is the next compact cell after x = 0, x = nx-2 the one before x = nx-1, so the neighbour is
    // the adjacent lane unless the pair straddles a wave boundary (lane 63 / lane 0 of the next wave).
    // Those pairs keep the neighbouring-thread write below; every other wave stores each population
    // plane with one full-wave store.
    const unsigned lane = unsigned(k) & 63u;
    const bool in_edge = prm.ns_inlet == 1 && inlet_cell && lane == 63u;
    const bool out_edge = outlet_cell && lane == 0u;
    if (__any(inlet_cell || outlet_cell)) {
        C fp[19], rp = C(0), up_x = C(0), up_y = C(0), up_z = C(0), um_x = C(0), um_y = C(0), um_z = C(0);
        if (prm.ns_inlet == 1) {
#pragma unroll
            for (int q = 0; q < 19; ++q) fp[q] = __shfl_down(fo[q], 1, 64);
            rp = __shfl_down(r2, 1, 64);
            up_x = __shfl_down(ux2, 1, 64);
            up_y = __shfl_down(uy2, 1, 64);
            up_z = __shfl_down(uz2, 1, 64);
        }
        if (prm.ns_outlet) {
            um_x = __shfl_up(ux2, 1, 64);
            um_y = __shfl_up(uy2, 1, 64);
            um_z = __shfl_up(uz2, 1, 64);
        }
        if (!(in_edge || out_edge)) static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            C v = fo[q];
            if (inlet_cell) v = (prm.ns_inlet == 2) ? inlet_eq<q>(prm) : inlet_neq<q>(fp[q], rp, up_x, up_y, up_z, prm);
            else if (outlet_cell) v = outlet_pressure<q>(um_x, um_y, um_z, prm);
            st_plane_nt(plane(p.fn, q, sa.nc), ncell_b, encode<P>(v, C(Q19::inv_w[q]), prm.f_ref));
        });
    } else {
#pragma unroll
        for (int q = 0; q < 19; ++q)
            st_plane_nt(plane(p.fn, q, sa.nc), ncell_b, encode<P>(fo[q], C(Q19::inv_w[q]), prm.f_ref));
    }
    if (prm.ns_inlet == 1 && x == 1 && lane == 0u) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            st_b(plane(p.fn, q, sa.nc), ncell_b - SB,
                 encode<P>(inlet_neq<q>(fo[q], r2, ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
        });
    }
    if (prm.ns_outlet && x == nx - 2 && lane == 63u) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            st_b(plane(p.fn, q, sa.nc), ncell_b + SB,
                 encode<P>(outlet_pressure<q>(ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
        });
    }
}

// Rows (y - cy, z - cz) around (y, z): the row deltas with periodic wrap, and whether each
// row exists (a = cy + 1, b = cz + 1).
template <class C>
__device__ __forceinline__ void neighbour_rows(const CoupledParams<C> &prm, int y, int z, int (&dy)[3], int (&dz)[3],
                                               bool (&rowok)[3][3]) {
    const int ny = prm.ny, nz = prm.nz;
    bool yin[3] = {y + 1 < ny, true, y > 0}, zin[3] = {z + 1 < nz, true, z > 0};
    dy[0] = 1;
    dy[1] = 0;
    dy[2] = -1;
    dz[0] = 1;
    dz[1] = 0;
    dz[2] = -1;
    if (prm.py) {
        dy[0] = yin[0] ? 1 : 1 - ny;
        dy[2] = yin[2] ? -1 : ny - 1;
        yin[0] = yin[2] = true;
    }
    if (prm.pz) {
        dz[0] = zin[0] ? 1 : 1 - nz;
        dz[2] = zin[2] ? -1 : nz - 1;
        zin[0] = zin[2] = true;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) rowok[a][b] = yin[a] && zin[b];
}

// The NS half of the split step, one thread per active cell in compact order (box index and link word
// in coords[k]; the neighbouring rows' table entries gathered per lane, from L1 / L2).
template <int P, bool WIDE>
__global__ void __launch_bounds__(256, coupled_waves(P))
    k_ns_split(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, SplitArgs sa) {
    using Off = std::conditional_t<WIDE, i64, unsigned>;
    const unsigned total = gridDim.x;
    unsigned b = blockIdx.x;
    if (prm.xcd_remap && (total & 7u) == 0u) b = (b & 7u) * (total >> 3) + (b >> 3);
    const i64 k = i64(b) * 256 + threadIdx.x;
    if (k >= sa.nact) return;
    const int nx = prm.nx, ny = prm.ny;
    const uint2 cl = sa.coords[k];
    const unsigned c = cl.x;
    const int x = int(c % unsigned(nx));
    const unsigned t = c / unsigned(nx);
    const int y = int(t % unsigned(ny)), z = int(t / unsigned(ny));
    int dy[3], dz[3];
    bool rowok[3][3];
    neighbour_rows(prm, y, z, dy, dz, rowok);
    const int seg = x >> 6;
    const unsigned jl = unsigned(x) & 63u;
    const unsigned long long below = (1ull << jl) - 1ull;
    Off rsl[3][3];
    unsigned rbit[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            if (a == 1 && bb == 1) {
                rsl[a][bb] = Off(k);
                rbit[a][bb] = 1u;
                continue;
            }
            const int yy = rowok[a][bb] ? y + dy[a] : y, zz = rowok[a][bb] ? z + dz[bb] : z;
            const RowSeg e = sa.rs[(i64(yy) + i64(ny) * zz) * sa.nseg + seg];
            const unsigned long long m = (static_cast<unsigned long long>(e.hi) << 32) | e.lo;
            rsl[a][bb] = Off(e.base + unsigned(__builtin_popcountll(m & below)));
            rbit[a][bb] = unsigned(m >> jl) & 1u;
        }
    ns_split_cell<P, WIDE>(p, prm, sa, k, x, y, z, cl.y, rsl, rbit, rowok);
}

template <int P, bool WIDE>
__global__ void __launch_bounds__(256, coupled_waves(P))
    k_ad_split(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, i64 step, i64 acc_row, SplitArgs sa) {
    coupled_body<P, FMX_NS_BGK, WIDE, true, false, true, SplitArgs>(p, prm, step, acc_row, sa);
}

template <int P>
int launch_coupled_split(const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                         i64 acc_row, const SplitArgs &sa, hipStream_t s) {
    const char *fw = std::getenv("FMX_WIDE");
    const bool wide = (fw && std::atoi(fw) != 0) || 4 * prm.n * i64(sizeof(typename Storage<P>::S)) > (i64(1) << 32);
    if (sa.nact > 0) {
        // a multiple of 8 workgroups, so that each XCD takes a contiguous range (the remap in the kernel)
        const dim3 g1(unsigned(((sa.nact + 255) / 256 + 7) / 8 * 8)), b1(256);
        if (wide) hipLaunchKernelGGL((k_ns_split<P, true>), g1, b1, 0, s, ptr, prm, sa);
        else hipLaunchKernelGGL((k_ns_split<P, false>), g1, b1, 0, s, ptr, prm, sa);
        if (int rc = after_launch("k_ns_split")) return rc;
    }
    const dim3 grid(coupled_blocks(prm.gx, prm.gy, prm.gz)), block(prm.block_x, prm.block_y, 1);
    if (wide) hipLaunchKernelGGL((k_ad_split<P, true>), grid, block, 0, s, ptr, prm, step, acc_row, sa);
    else hipLaunchKernelGGL((k_ad_split<P, false>), grid, block, 0, s, ptr, prm, step, acc_row, sa);
    return after_launch("k_ad_split");
}

template <int P>
int launch_coupled_rc(const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                      i64 acc_row, const RcArgs &rc, hipStream_t s) {
    const dim3 grid(coupled_blocks(prm.gx, prm.gy, prm.gz)), block(prm.block_x, prm.block_y, 1);
    // compact planes are no larger than the box's, and the box's AD planes decide as in k_coupled
    const char *fw = std::getenv("FMX_WIDE");
    const bool wide = (fw && std::atoi(fw) != 0) || 4 * prm.n * i64(sizeof(typename Storage<P>::S)) > (i64(1) << 32);
    if (wide) hipLaunchKernelGGL((k_coupled_rc<P, true>), grid, block, 0, s, ptr, prm, step, acc_row, rc);
    else hipLaunchKernelGGL((k_coupled_rc<P, false>), grid, block, 0, s, ptr, prm, step, acc_row, rc);
    return after_launch("k_coupled_rc");
}

template <int P>
int launch_coupled_step(int ns_op, const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm,
                        i64 step, i64 acc_row, hipStream_t s) {
    if constexpr (P != FMX_F64) {
        if (prm.pair) return launch_coupled_pair<P>(ns_op, ptr, prm, step, acc_row, s);
    }
    const dim3 grid(coupled_blocks(prm.gx, prm.gy, prm.gz)), block(prm.block_x, prm.block_y, 1);
    // 32-bit offsets reach four population planes from a plane base (a bounce-back pair is at most
    // three planes apart)
    // (FMX_WIDE=1 forces the 64-bit form, to check it against the 32-bit one on small boxes)
    const char *fw = std::getenv("FMX_WIDE");
    const bool wide = (fw && std::atoi(fw) != 0) || 4 * prm.n * i64(sizeof(typename Storage<P>::S)) > (i64(1) << 32);
#define FMX_LAUNCH(OP, W, FZ) hipLaunchKernelGGL((k_coupled<P, OP, W, FZ>), grid, block, 0, s, ptr, prm, step, acc_row)
    switch (ns_op) {
        case FMX_NS_BGK:
            if (prm.freeze) {
                if (wide) FMX_LAUNCH(FMX_NS_BGK, true, true);
                else FMX_LAUNCH(FMX_NS_BGK, false, true);
            } else {
                if (wide) FMX_LAUNCH(FMX_NS_BGK, true, false);
                else FMX_LAUNCH(FMX_NS_BGK, false, false);
            }
            break;
        case FMX_NS_POROUS:
            if (wide) FMX_LAUNCH(FMX_NS_POROUS, true, false);
            else FMX_LAUNCH(FMX_NS_POROUS, false, false);
            break;
        case FMX_NS_PBB:
            if (wide) FMX_LAUNCH(FMX_NS_PBB, true, false);
            else FMX_LAUNCH(FMX_NS_PBB, false, false);
            break;
        default: return set_error(FMX_ERR_ARG, "fmx_coupled: unknown ns_op %d", ns_op);
    }
#undef FMX_LAUNCH
    return after_launch("k_coupled");
}

}  // namespace fmx
