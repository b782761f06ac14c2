// coupled_kernel.hpp -- the fused coupled step kernel (one cell per thread) and its launcher.
//
// Included by coupled.hip (FP32 / FP64 storage: IEEE division, no FMA contraction, bit-identical
// to the serial restatement of the reference) and coupled_half.hip (FP16 / BF16 storage, compiled
// with FMA contraction and the hardware reciprocal: its accuracy is set by the 16-bit storage).
// Each precision is instantiated in exactly one of the two files.
#pragma once

#include "coupled_cell.hpp"

namespace fmx {

template <int P, int OP, bool NTL>
__global__ void __launch_bounds__(256) k_coupled(CoupledPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, i64 step,
                                                 i64 acc_row) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    constexpr unsigned SB = sizeof(S);
    constexpr unsigned CB = sizeof(C);
    int bx, by, bz;
    block_coords(prm, bx, by, bz);
    const int x = bx * prm.block_x + int(threadIdx.x);
    const int y = by * prm.block_y + int(threadIdx.y);
    const int z = bz;
    if (x >= prm.nx || y >= prm.ny) return;
    const int nx = prm.nx, ny = prm.ny, nz = prm.nz;
    const i64 n = prm.n;
    const unsigned cell = unsigned(x) + unsigned(nx) * (unsigned(y) + unsigned(ny) * unsigned(z));
    const unsigned cell_b = cell * SB;
    const std::uint32_t lk = ld_b(p.links, cell * 4u);
    const bool solid = (OP == FMX_NS_BGK) && (lk & FMX_LINK_SOLID);
    // a frozen solid cell alternates between the two states held in f and f_new and has zero
    // post-collision velocity in both: no NS loads, collision or stores
    const bool frozen = (OP == FMX_NS_BGK) && (lk & FMX_LINK_FROZEN);

    // neighbour rows: x keeps local out of domain, y/z wrap when periodic
    const bool xm_ok = x > 0, xp_ok = x < nx - 1;
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
    const bool xok[3] = {xm_ok, true, xp_ok};
    // source cell of the pull along (cx, cy, cz) = cell - c, or the cell itself when out of domain
    auto src_of = [&](int cx, int cy, int cz) -> unsigned {
        const bool ok = xok[1 - cx] && yok[1 - cy] && zok[1 - cz];
        return ok ? unsigned(x - cx) + unsigned(nx) * (ys[1 - cy] + unsigned(ny) * zs[1 - cz]) : cell;
    };

    // ---------------------------------------------------------------- NS pull
    C fo[19];
    if (!frozen) static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
        const unsigned src = src_of(cx, cy, cz);
        bool bb = false;
        if constexpr (OP == FMX_NS_BGK && q != 0) bb = (lk >> q) & 1u;
        S raw;
        if constexpr (NTL) {
            if (bb) raw = ld_b_nt(plane(p.f, Q19::opp[q], n), cell_b);
            else raw = ld_b_nt(plane(p.f, q, n), src * SB);
        } else {
            if (bb) raw = ld_b(plane(p.f, Q19::opp[q], n), cell_b);
            else raw = ld_b(plane(p.f, q, n), src * SB);
        }
        fo[q] = decode<P>(raw, w19<C>(q), prm.f_ref);
    });

    // ------------------------------------------- AD pull and sorption inputs
    // (issued with the NS pulls; nothing below writes what they read)
    const C phi_pre = (prm.ad_inlet == 2) ? C(p.inlet_series[step]) : C(0);
    C gi[7];
    static_for<0, 7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
        const unsigned s = src_of(cx, cy, cz);
        // x of the source (or keep-local) cell: x - cx when in the domain, else x
        const int sx = (cx == 1) ? (x > 0 ? x - 1 : x) : ((cx == -1) ? (x < nx - 1 ? x + 1 : x) : x);
        if (prm.ad_inlet == 2 && sx == 0) {
            gi[q] = phi_pre * w7<C>(q);
        } else {
            const S raw = NTL ? ld_b_nt(plane(p.g, q, n), s * SB) : ld_b(plane(p.g, q, n), s * SB);
            if constexpr (IsHalf<P>::value) {
                // deviation about the source cell's lagged concentration; a Neumann outlet cell holds the
                // populations of its x - 1 neighbour, stored about that neighbour's concentration
                const unsigned sref = (prm.ad_outlet && sx == nx - 1) ? s - 1u : s;
                gi[q] = w7<C>(q) * (ld_b(p.phi, sref * CB) + widen(raw));
            } else {
                gi[q] = raw;
            }
        }
    });
    const bool sorbs = prm.isotherm && (lk & FMX_LINK_REGION);
    C phi_lag = C(0), q_old = C(0);
    if (sorbs) {
        phi_lag = ld_b(p.phi, cell * CB);
        q_old = ld_b(p.q, cell * CB);
    }

    // ------------------------------------------ NS collide (in place) and moments
    C r2 = C(0), ux2 = C(0), uy2 = C(0), uz2 = C(0);
    if (!frozen) {
        const C fld = p.field ? ld_b(p.field, cell * CB) : ((OP == FMX_NS_POROUS) ? C(1) : C(0));
        ns_collide_cell<OP, IsHalf<P>::value>(fo, solid, fld, prm);
        post_moments<IsHalf<P>::value>(fo, r2, ux2, uy2, uz2);
    }

    // ------------------------------------------------------------- NS stores
    const bool inlet_cell = prm.ns_inlet && x == 0;
    const bool outlet_cell = prm.ns_outlet && x == nx - 1;
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
            st_b_nt(plane(p.fn, q, n), cell_b, encode<P>(v, C(Q19::inv_w[q]), prm.f_ref));
        });
    } else {
        // the x = 1 thread writes the NEQ inlet of x = 0, the x = nx-2 thread the pressure outlet
        if (inlet_cell) {
            if (prm.ns_inlet == 2) {  // next step's VelocityDirichletEq: f(0,y,z) = feq(rho0, u_BC)
                static_for<0, 19>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    st_b_nt(plane(p.fn, q, n), cell_b, encode<P>(inlet_eq<q>(prm), C(Q19::inv_w[q]), prm.f_ref));
                });
            }
        } else if (!outlet_cell && !frozen) {
#pragma unroll
            for (int q = 0; q < 19; ++q)
                st_b_nt(plane(p.fn, q, n), cell_b, encode<P>(fo[q], C(Q19::inv_w[q]), prm.f_ref));
        }
        if (prm.ns_inlet == 1 && x == 1) {  // next step's NEQ update of (0, y, z)
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                st_b(plane(p.fn, q, n), cell_b - SB,
                     encode<P>(inlet_neq<q>(fo[q], r2, ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
            });
        }
        if (prm.ns_outlet && x == nx - 2) {  // PressureDirichlet of (nx-1, y, z)
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                st_b(plane(p.fn, q, n), cell_b + SB,
                     encode<P>(outlet_pressure<q>(ux2, uy2, uz2, prm), C(Q19::inv_w[q]), prm.f_ref));
            });
        }
    }

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
    if (prm.lane_bc) {
        C gm[7], phim = C(0);
        if (prm.ad_outlet) {
#pragma unroll
            for (int q = 0; q < 7; ++q) gm[q] = __shfl_up(go[q], 1, 64);
            phim = __shfl_up(phi, 1, 64);
        }
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            C v = go[q], ref = phi;
            if (ad_in_cell) {
                v = prm.phi_in * w7<C>(q);
            } else if (ad_out_cell) {
                v = gm[q];
                ref = phim;
            }
            st_b_nt(plane(p.gn, q, n), cell_b, encode<P>(v, C(Q7::inv_w[q]), ref));
        });
    } else {
        if (ad_in_cell) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                st_b_nt(plane(p.gn, q, n), cell_b, encode<P>(prm.phi_in * w7<C>(q), C(Q7::inv_w[q]), phi));
        } else if (!ad_out_cell) {
#pragma unroll
            for (int q = 0; q < 7; ++q) st_b_nt(plane(p.gn, q, n), cell_b, encode<P>(go[q], C(Q7::inv_w[q]), phi));
        }
        if (prm.ad_outlet && x == nx - 2) {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                st_b(plane(p.gn, q, n), cell_b + SB, encode<P>(go[q], C(Q7::inv_w[q]), phi));
        }
    }

    // ------------------------------------------------- outlet breakthrough sum
    if (p.outlet_acc && x == nx - 1) {
        const int slot = int((unsigned(y) + unsigned(ny) * unsigned(z)) % FMX_OUTLET_SLOTS);
        unsafeAtomicAdd(p.outlet_acc + acc_row * (FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE) + slot * FMX_OUTLET_STRIDE,
                        double(phi));
    }
}

template <int P>
int launch_coupled_step(int ns_op, const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm,
                        i64 step, i64 acc_row, hipStream_t s) {
    if constexpr (P != FMX_F64) {
        if (prm.pair) return launch_coupled_pair<P>(ns_op, ptr, prm, step, acc_row, s);
    }
    const dim3 grid(coupled_blocks(prm.gx, prm.gy, prm.gz)), block(prm.block_x, prm.block_y, 1);
    // FMX_NT_LOADS=1 streams the population pulls with non-temporal loads (measured slower; kept for study)
    static const bool nt_loads = [] {
        const char *e = std::getenv("FMX_NT_LOADS");
        return e != nullptr && std::atoi(e) != 0;
    }();
#define FMX_LAUNCH(OP, NT) hipLaunchKernelGGL((k_coupled<P, OP, NT>), grid, block, 0, s, ptr, prm, step, acc_row)
    switch (ns_op) {
        case FMX_NS_BGK:
            if (nt_loads) FMX_LAUNCH(FMX_NS_BGK, true);
            else FMX_LAUNCH(FMX_NS_BGK, false);
            break;
        case FMX_NS_POROUS:
            if (nt_loads) FMX_LAUNCH(FMX_NS_POROUS, true);
            else FMX_LAUNCH(FMX_NS_POROUS, false);
            break;
        case FMX_NS_PBB:
            if (nt_loads) FMX_LAUNCH(FMX_NS_PBB, true);
            else FMX_LAUNCH(FMX_NS_PBB, false);
            break;
        default: return set_error(FMX_ERR_ARG, "fmx_coupled: unknown ns_op %d", ns_op);
    }
#undef FMX_LAUNCH
    return after_launch("k_coupled");
}

}  // namespace fmx
