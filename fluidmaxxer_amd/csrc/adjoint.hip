// adjoint.hip -- discrete adjoint of the fused coupled step (k_coupled in coupled.hip).
//
// The reference has no adjoint; the map differentiated here is exactly the fused forward step
// (same pulls, boundary handling, clamps and branches), so the gradients are those of the
// reference's update sequence (main.cpp:319-336).
//
// Representation.  The adjoint state after the reverse of step s is held READER-indexed:
//     lam_f[q][c] = dL / d(the value cell c pulled along q at the start of step s)
//     lam_g[q][c] = the same for the D3Q7 pulls,
// plus lam_q[c] = dL/dq(c) and lam_phi[c] = dL/d(lagged concentration at c).  Every stored
// population slot of the forward state is read by at most two pulls: its streaming neighbour,
// or the owner itself through bounce-back, plus the owner's keep-local pull where the opposite
// source lies outside the box.  The reverse step of cell c first GATHERS the adjoint of the
// slots it wrote in step s from those readers (its own link word says which reader applies),
// then back-propagates through its local update and writes the adjoint of what it pulled.
// No atomics are needed for the populations and every adjoint slot is written once per step.
//
// Boundary rules (the boundary conditions as the fused step folds them in):
//   * x = 0 with an equilibrium NS inlet, or an ADDirichlet AD inlet: the stored values are
//     constants, so the owner's adjoint of them is zero.
//   * x = 0 with the NEQ velocity inlet: the stored values are the next step's
//     feq(r2, u_BC) + f - feq(r2, u2) of the (1, y, z) cell's post-collision populations f and their
//     moments r2, u2 (VelocityDirichletNEQ.cu:81-95), so that cell's reverse step gathers their
//     adjoint and back-propagates it; the x = 0 cell's own collision result is replaced and gets none.
//   * x = nx-1 with a pressure outlet / Neumann outlet: the outlet slots hold functions of the
//     (nx-2, y, z) cell's post-collision state (the pressure equilibrium of its velocity, a copy
//     of its g), so that cell's reverse step gathers their adjoint and back-propagates it; the
//     outlet cell's own collision result is discarded and gets none.
//   * per-step inlet series: pulls whose source lies on x = 0 read the series value; their
//     adjoint accumulates into grad_series[step].
//   * clamps (phi >= min_conc, field in [0, 1], tau >= 0.5001) differentiate the active branch;
//     on the closed interval [0, 1] the field passes its gradient through.
#include "coupled_common.hpp"

namespace fmx {

template <int P> struct AdjointPtrs {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    // forward state at the start of the step
    const S *f;
    const S *g;
    const C *q;
    const C *phi;
    const std::uint32_t *links;
    const C *field;
    const C *tau_field;
    const double *inlet_series;
    // adjoint: reader-indexed adjoint of the next step's pulls (in) and of this step's pulls (out)
    const C *lf_in;
    const C *lg_in;
    C *lf_out;
    C *lg_out;
    C *lam_q;    // in place: dL/dq after the step -> before the step
    C *lam_phi;  // in place: dL/d(concentration written by the step) -> dL/d(lagged concentration)
    double *grad_field;
    double *grad_series;
    const double *outlet_weight;
    double *iso_part;            // 2 per block: the block's running dL/d iso_a, dL/d iso_b
    double *ser_row;             // 1 per block: the block's inlet-series sum of this step (or null)
};

// Block sum of three values; lane 0 of the block receives the totals.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c) {
    __shared__ double red[3][16];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    const int tid = int(threadIdx.x + threadIdx.y * blockDim.x);
    const int wave = tid >> 6, lane = tid & 63;
    const int nwaves = int(blockDim.x * blockDim.y) >> 6;
    if (lane == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
        red[2][wave] = c;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < nwaves; ++w) {
            a += red[0][w];
            b += red[1][w];
            c += red[2][w];
        }
    }
}

// Addressing as in k_coupled: byte offsets are 32-bit (a scalar plane base plus a vector offset)
// unless WIDE; rows are whole waves, so the y / z neighbour deltas are wave-uniform.  The collided
// populations are recomputed where they are needed rather than held (fewer live registers, more
// waves in flight); every sum runs in the same order as before, so the results do not depend on it.
template <int P, int OP, bool WIDE>
__global__ void __launch_bounds__(256) k_adjoint(AdjointPtrs<P> p, CoupledParams<typename Storage<P>::C> prm, i64 step,
                                                 i64 widx, double inv_count) {
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
    const int nx = prm.nx, ny = prm.ny, nz = prm.nz;
    const i64 n = prm.n;
    double acc_a = 0.0, acc_b = 0.0, acc_s = 0.0;

    if (x < nx && y < ny) {
        const Off cell = Off(x) + Off(nx) * (Off(y) + Off(ny) * Off(z));
        const Off cell_c = cell * CB;
        const Off plane_s = Off(n) * SB, plane_c = Off(n) * CB;
        const std::uint32_t lk = ld_b(p.links, cell * 4u);
        const bool solid = (OP == FMX_NS_BGK) && (lk & FMX_LINK_SOLID);

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
        // does the cell (ox - cx, y - cy, z - cz) exist, and its index relative to the cell ox
        auto rel_ok = [&](int ox, int cx, int cy, int cz) -> bool {
            return ox - cx >= 0 && ox - cx < nx && yin[cy + 1] && zin[cz + 1];
        };
        auto rel = [&](int cx, int cy, int cz) -> Off {
            return Off(dy[cy + 1]) * Off(nx) + Off(dz[cz + 1]) * (Off(nx) * Off(ny)) - Off(cx);
        };
        // source of this cell's pull along c: the cell - c, or the cell itself outside the box
        auto src = [&](int cx, int cy, int cz) -> Off { return rel_ok(x, cx, cy, cz) ? cell + rel(cx, cy, cz) : cell; };

        // ------------------------------------------------ forward recompute: pulls
        C fo[19];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
            constexpr int o = Q19::opp[q];
            constexpr bool BB = OP == FMX_NS_BGK && q != 0;
            constexpr int bq = BB ? (q < o ? q : o) : q;
            Off off = src(cx, cy, cz) * SB + Off(q - bq) * plane_s;
            if constexpr (BB) {  // the bounce-back pull: this cell along opp(q)
                const Off boff = cell * SB + Off(o - bq) * plane_s;
                const Off m = Off(int(lk << (31 - q)) >> 31);
                off = (boff & m) | (off & ~m);
            }
            fo[q] = decode<P>(ld_b(plane(p.f, bq, n), off), w19<C>(q), prm.f_ref);
        });
        // (every load below is unconditional, from an address in the box, and selected afterwards:
        // loads inside branches would each wait for their data before the next one is issued)
        const C phi_pre = (prm.ad_inlet == 2) ? C(p.inlet_series[step]) : C(0);
        C gi[7];
        unsigned subst = 0;  // bit q: the pull along q reads the inlet series value
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
            const bool in = rel_ok(x, cx, cy, cz);
            const int sx = in ? x - cx : x;
            const bool sub = prm.ad_inlet == 2 && sx == 0;
            subst |= unsigned(sub) << q;
            const Off s = (in ? cell + rel(cx, cy, cz) : cell) * SB;
            const S raw = ld_b(plane(p.g, q, n), s);
            C gq;
            if constexpr (IsHalf<P>::value) {
                // 16-bit states hold g about the source cell's lagged concentration (the x - 1 cell's
                // on a Neumann outlet plane), as the forward step stores them
                const Off sref = s * (CB / SB) - ((prm.ad_outlet && sx == nx - 1) ? Off(CB) : Off(0));
                gq = decode_g<P>(raw, w7<C>(q), ld_b(p.phi, sref));
            } else {
                gq = raw;
            }
            gi[q] = sub ? phi_pre * w7<C>(q) : gq;
        });
        C phis = C(0);
#pragma unroll
        for (int q = 0; q < 7; ++q) phis += gi[q];
        const bool sorbs = prm.isotherm && (lk & FMX_LINK_REGION);
        const C phi_lag = sorbs ? ld_b(p.phi, cell_c) : C(0);
        const C fv = p.field ? ld_b(p.field, cell_c) : ((OP == FMX_NS_POROUS) ? C(1) : C(0));
        const C fcl = fv < C(0) ? C(0) : (fv > C(1) ? C(1) : fv);
        const bool f_pass = fv >= C(0) && fv <= C(1);

        // ------------------------------------------------ forward recompute: NS collision
        C rho = C(0), jx = C(0), jy = C(0), jz = C(0);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            rho += fo[q];
            jx = axpy1<Q19::cx[q]>(jx, fo[q]);
            jy = axpy1<Q19::cy[q]>(jy, fo[q]);
            jz = axpy1<Q19::cz[q]>(jz, fo[q]);
        });
        const bool rho_ok = rho > C(0);
        C ux = C(0), uy = C(0), uz = C(0), invR = C(0);
        if (rho_ok) {
            invR = C(1) / rho;
            ux = jx * invR;
            uy = jy * invR;
            uz = jz * invR;
            if (OP == FMX_NS_POROUS) {
                ux *= fcl;
                uy *= fcl;
                uz *= fcl;
            }
        } else {
            rho = prm.rho0;
        }
        const C u2x15 = C(1.5) * (ux * ux + uy * uy + uz * uz);
        const C ns = (OP == FMX_NS_PBB) ? fcl : C(0);
        // collided population q (before the partial bounce-back blend)
        auto fc = [&](auto qc) -> C {
            constexpr int q = decltype(qc)::value;
            return fo[q] + prm.omega * (feq_bgk<q>(rho, ux, uy, uz, u2x15) - fo[q]);
        };
        // post-collision moments of the stored result
        C r2 = C(0), vx = C(0), vy = C(0), vz = C(0);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            C fr = fc(qc);
            if (OP == FMX_NS_PBB) fr = (C(1) - ns) * fr + ns * fo[Q19::opp[q]];
            if (solid) fr = fo[q];
            r2 += fr;
            vx = axpy1<Q19::cx[q]>(vx, fr);
            vy = axpy1<Q19::cy[q]>(vy, fr);
            vz = axpy1<Q19::cz[q]>(vz, fr);
        });
        const C inv2 = (r2 > C(0)) ? (C(1) / r2) : C(0);
        const C ux2 = vx * inv2, uy2 = vy * inv2, uz2 = vz * inv2;

        // ------------------------------------------------ forward recompute: sorption and AD
        C ce = C(0), bc = C(0);
        if (sorbs) {
            const C c = phi_lag < prm.min_conc ? prm.min_conc : phi_lag;
            ce = c - prm.min_conc;
            bc = prm.iso_b * ce;
        }
        const bool clamped = phis < prm.min_conc;
        const C phi = clamped ? prm.min_conc : phis;
        const C om = p.tau_field ? C(1) / ld_b(p.tau_field, cell_c) : prm.omega_ad;
        const C ax = prm.coupling ? ux2 : C(0), ay = prm.coupling ? uy2 : C(0), az = prm.coupling ? uz2 : C(0);

        // ------------------------------------------------ gather the adjoint of what the step stored
        // adjoint of slot (q, owner) of f, the owner at x = ox, index oc, link word lo: its readers'
        // adjoint (the owner + c, or the owner itself through bounce-back; plus the owner's
        // keep-local pull where the owner - c lies outside the box)
        // The streaming / bounce-back reader's part (gf_main, gg_main; `on` false: none) is loaded
        // unconditionally, the keep-local part only where the owner - c lies outside the box.
        auto gf_main = [&](auto qc, int ox, Off oc, std::uint32_t lo, bool on) -> C {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q], oq = Q19::opp[q];
            constexpr bool BB = OP == FMX_NS_BGK && q != 0;
            constexpr int bq = BB ? (q < oq ? q : oq) : q;
            const bool nb_ok = rel_ok(ox, -cx, -cy, -cz);
            Off off = (nb_ok ? oc + rel(-cx, -cy, -cz) : oc) * CB + Off(q - bq) * plane_c;
            bool take = nb_ok;
            if constexpr (OP == FMX_NS_BGK) {
                // a solid owner's slot q is read back by itself along opp(q); so is a bounce-back link's
                const bool self = (lo & FMX_LINK_SOLID) || (BB && ((lo >> oq) & 1u));
                off = self ? oc * CB + Off(oq - bq) * plane_c : off;
                take = take || self;
            }
            const C v = ld_b(plane(p.lf_in, bq, n), off);
            return (on && take) ? v : C(0);
        };
        auto gf_local = [&](auto qc, int ox, Off oc, std::uint32_t lo, bool on, C v) -> C {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q19::cx[q], cy = Q19::cy[q], cz = Q19::cz[q];
            const bool o_solid = OP == FMX_NS_BGK && (lo & FMX_LINK_SOLID);
            if (on && !o_solid && !rel_ok(ox, cx, cy, cz)) v += ld_b(plane(p.lf_in, q, n), oc * CB);
            return v;
        };
        auto gg_main = [&](auto qc, int ox, Off oc, bool on) -> C {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
            const bool nb_ok = rel_ok(ox, -cx, -cy, -cz);
            const C v = ld_b(plane(p.lg_in, q, n), (nb_ok ? oc + rel(-cx, -cy, -cz) : oc) * CB);
            return (on && nb_ok) ? v : C(0);
        };
        auto gg_local = [&](auto qc, int ox, Off oc, bool on, C v) -> C {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = Q7::cx[q], cy = Q7::cy[q], cz = Q7::cz[q];
            if (on && !rel_ok(ox, cx, cy, cz)) v += ld_b(plane(p.lg_in, q, n), oc * CB);
            return v;
        };

        const bool f_const = (prm.ns_inlet && x == 0) || (prm.ns_outlet && x == nx - 1);
        const bool g_const = (prm.ad_inlet == 1 && x == 0) || (prm.ad_outlet && x == nx - 1);
        const bool writes_outlet_f = prm.ns_outlet && x == nx - 2;
        const bool writes_outlet_g = prm.ad_outlet && x == nx - 2;
        // slots of g on x = 0 under a series inlet: every reader takes the series value instead
        const bool g_gather = !g_const && !(prm.ad_inlet == 2 && x == 0);
        C lfr[19], lgo[7];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lfr[q] = gf_main(qc, x, cell, lk, !f_const);
        });
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lgo[q] = gg_main(qc, x, cell, g_gather);
        });
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lfr[q] = gf_local(qc, x, cell, lk, !f_const, lfr[q]);
        });
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lgo[q] = gg_local(qc, x, cell, g_gather, lgo[q]);
        });
        if (writes_outlet_g) {  // the Neumann outlet's copy of this cell's g (x + 1 = nx - 1 is never 0)
            static_for<0, 7>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                lgo[q] += gg_local(qc, x + 1, cell + 1u, true, gg_main(qc, x + 1, cell + 1u, true));
            });
        }
        C lux = C(0), luy = C(0), luz = C(0);  // adjoint of the post-collision velocity
        if (prm.ns_inlet == 1 && x == 1) {
            // the NEQ inlet values of (0, y, z): fB = feq(r2, u_BC) + f - feq(r2, u2) with f, r2, u2 this
            // cell's post-collision populations and moments.  dfB/df is the identity; through r2 (u2
            // held) every fB gains w (P(u_BC) - P(u2)), P(u) = 1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2; through
            // u2, -r2 w (3 e + 9 (e.u2) e - 3 u2); u2 = v / r2 goes back with the coupling below.
            const std::uint32_t lo = ld_b(p.links, (cell - 1u) * 4u);
            const C u2N = C(1.5) * (ux2 * ux2 + uy2 * uy2 + uz2 * uz2), u2B = C(1.5) * prm.uin2;
            C lr2 = C(0), sw = C(0), tx = C(0), ty = C(0), tz = C(0);
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const C lb = gf_local(qc, 0, cell - 1u, lo, true, gf_main(qc, 0, cell - 1u, lo, true));
                lfr[q] += lb;
                const C w = w19<C>(q);
                const C euN = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(ux2, uy2, uz2);
                const C euB = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(prm.uin_x, prm.uin_y, prm.uin_z);
                const C cuN = C(3) * euN, cuB = C(3) * euB;
                lr2 += lb * w * ((cuB + C(0.5) * cuB * cuB - u2B) - (cuN + C(0.5) * cuN * cuN - u2N));
                const C t = lb * w * C(3) * (C(1) + cuN);
                tx = axpy1<Q19::cx[q]>(tx, t);
                ty = axpy1<Q19::cy[q]>(ty, t);
                tz = axpy1<Q19::cz[q]>(tz, t);
                sw += lb * w;
            });
            lux -= r2 * (tx - C(3) * ux2 * sw);
            luy -= r2 * (ty - C(3) * uy2 * sw);
            luz -= r2 * (tz - C(3) * uz2 * sw);
#pragma unroll
            for (int q = 0; q < 19; ++q) lfr[q] += lr2;
        }
        if (writes_outlet_f) {  // the outlet cell's populations feq(rho_out, u2) (PressureDirichlet.cu:81-95)
            const std::uint32_t lo = ld_b(p.links, (cell + 1u) * 4u);
            C sw = C(0);
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const C lp = gf_local(qc, x + 1, cell + 1u, lo, true, gf_main(qc, x + 1, cell + 1u, lo, true)) *
                             w19<C>(q) * prm.rho_out;
                const C eu = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(ux2, uy2, uz2);
                const C t = lp * C(3) * (C(1) + C(3) * eu);
                lux = axpy1<Q19::cx[q]>(lux, t);
                luy = axpy1<Q19::cy[q]>(luy, t);
                luz = axpy1<Q19::cz[q]>(luz, t);
                sw += lp;
            });
            lux -= C(3) * ux2 * sw;
            luy -= C(3) * uy2 * sw;
            luz -= C(3) * uz2 * sw;
        }
        const C lq_next = ld_b(p.lam_q, cell_c);
        const C seed = (p.outlet_weight && x == nx - 1) ? C(p.outlet_weight[widx] * inv_count) : C(0);
        C lphi = ld_b(p.lam_phi, cell_c) + seed;

        // ------------------------------------------------ AD collision: go = gi + om (geq - gi) + w S
        C lS = C(0), lax = C(0), lay = C(0), laz = C(0);
        C lgi[7];
        static_for<0, 7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const C w = w7<C>(q);
            const C eu = dot3<Q7::cx[q], Q7::cy[q], Q7::cz[q]>(ax, ay, az);
            const C lgeq = om * lgo[q];
            lS += w * lgo[q];
            lphi += lgeq * w * (C(1) + C(4) * eu);
            const C t = lgeq * w * phi * C(4);
            lax = axpy1<Q7::cx[q]>(lax, t);
            lay = axpy1<Q7::cy[q]>(lay, t);
            laz = axpy1<Q7::cz[q]>(laz, t);
            lgi[q] = (C(1) - om) * lgo[q];
        });
        if (!clamped) {
#pragma unroll
            for (int q = 0; q < 7; ++q) lgi[q] += lphi;
        }
        C ser = C(0);
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if ((subst >> q) & 1u) ser += w7<C>(q) * lgi[q];
            st_plane_nt(plane(p.lg_out, q, n), cell_c, lgi[q]);
        }
        acc_s = double(ser);

        // ------------------------------------------------ sorption: q' = q + dq, S = -dq
        if (sorbs) {
            const C ldq = lq_next - lS;
            const C lqeq = prm.k_ldf * ldq;
            C lce;
            if (prm.isotherm == FMX_ISO_LANGMUIR) {
                const C den = C(1) + bc;
                acc_a = double(lqeq * (bc / den));
                const C lbc = lqeq * prm.iso_a / (den * den);
                acc_b = double(lbc * ce);
                lce = lbc * prm.iso_b;
            } else {
                acc_a = double(lqeq * ce);
                lce = lqeq * prm.iso_a;
            }
            st_b(p.lam_q, cell_c, C(lq_next - prm.k_ldf * ldq));
            st_b(p.lam_phi, cell_c, phi_lag < prm.min_conc ? C(0) : lce);
        } else {
            st_b(p.lam_q, cell_c, C(0));
            st_b(p.lam_phi, cell_c, C(0));
        }

        // ------------------------------------------------ NS: post-collision velocity u2 = v / r2
        if (prm.coupling) {
            lux += lax;
            luy += lay;
            luz += laz;
        }
        if (r2 > C(0)) {
            const C lu_u = lux * ux2 + luy * uy2 + luz * uz2;
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const C lc = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(lux, luy, luz);
                lfr[q] += (lc - lu_u) * inv2;
            });
        }

        // ------------------------------------------------ NS collision and pull
        if (solid) {
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                st_plane_nt(plane(p.lf_out, q, n), cell_c, lfr[q]);
            });
        } else {
            // fr = (1 - ns) fc + ns fo[opp] (partial bounce-back), fc = fo + omega (feq(rho, u) - fo):
            // first the sums over q, then each pulled population's adjoint
            C lfield = C(0);
            C lrho = C(0), lvx = C(0), lvy = C(0), lvz = C(0), sw = C(0);
            // The per-direction terms below are recomputed from copies of rho and u that the compiler
            // cannot see through: otherwise it keeps the equilibria of the forward recompute above
            // live (19 or more registers) all the way here instead of recomputing them.
            C rr = rho, rux = ux, ruy = uy, ruz = uz, ru2 = u2x15;
            asm volatile("" : "+v"(rr), "+v"(rux), "+v"(ruy), "+v"(ruz), "+v"(ru2));
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                C lfc = lfr[q];
                if (OP == FMX_NS_PBB) {
                    const C fo_q = fo[q];
                    lfield += lfr[q] * (fo[Q19::opp[q]] - (fo_q + prm.omega * (feq_bgk<q>(rr, rux, ruy, ruz, ru2) - fo_q)));
                    lfc = (C(1) - ns) * lfr[q];
                }
                const C lfeq = prm.omega * lfc;
                const C w = w19<C>(q);
                const C eu = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(rux, ruy, ruz);
                const C cu = C(3) * eu;
                lrho += lfeq * w * (C(1) + cu + C(0.5) * cu * cu - ru2);
                const C t = lfeq * w * rr * C(3) * (C(1) + cu);
                lvx = axpy1<Q19::cx[q]>(lvx, t);
                lvy = axpy1<Q19::cy[q]>(lvy, t);
                lvz = axpy1<Q19::cz[q]>(lvz, t);
                sw += lfeq * w;
            });
            C ljx = C(0), ljy = C(0), ljz = C(0);
            if (rho_ok) {  // the rho <= 0 fallback uses constants
                lvx -= C(3) * rho * ux * sw;
                lvy -= C(3) * rho * uy * sw;
                lvz -= C(3) * rho * uz * sw;
                // u = s j / rho (s = clamped porosity for the porous operator, else 1)
                const C s = (OP == FMX_NS_POROUS) ? fcl : C(1);
                ljx = lvx * s * invR;
                ljy = lvy * s * invR;
                ljz = lvz * s * invR;
                lrho -= (lvx * ux + lvy * uy + lvz * uz) * invR;
                if (OP == FMX_NS_POROUS) lfield = (lvx * jx + lvy * jy + lvz * jz) * invR;
            } else {
                lrho = C(0);
            }
            C rns = ns;  // (recomputed, not carried over from the sums above; see rr)
            asm volatile("" : "+v"(rns));
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                C lfo = C(0);
                if (OP == FMX_NS_PBB) lfo += rns * lfr[Q19::opp[q]];
                const C lfc = (OP == FMX_NS_PBB) ? (C(1) - rns) * lfr[q] : lfr[q];
                lfo += (C(1) - prm.omega) * lfc;
                lfo += lrho + dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(ljx, ljy, ljz);
                st_plane_nt(plane(p.lf_out, q, n), cell_c, lfo);
            });
            if ((OP == FMX_NS_PBB || OP == FMX_NS_POROUS) && p.grad_field && p.field && f_pass) {
                const Off gb = cell * 8u;
                st_b(p.grad_field, gb, ld_b(p.grad_field, gb) + double(lfield));
            }
        }
    }

    // ---------------------------------------------------- parameter gradients
    // Block b covers the same cells in every step: it keeps its own running isotherm sums and stores
    // its inlet-series sum of this step; k_series_rows and k_iso_final add them in block order.  No
    // atomics, so the gradients repeat to the bit from run to run.
    const bool series = prm.ad_inlet == 2 && p.ser_row;
    if (!prm.isotherm && !series) return;  // grid-uniform
    block_sum3(acc_a, acc_b, acc_s);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        if (prm.isotherm) {
            double *ip = p.iso_part + 2 * i64(blockIdx.x);
            ip[0] += acc_a;
            ip[1] += acc_b;
        }
        if (series) p.ser_row[blockIdx.x] = bx == 0 ? acc_s : 0.0;  // series pulls lie in the x = 0 / 1 cells
    }
}

// Row r of the series buffer holds the per-block sums of reverse step top - r; one workgroup per
// row adds them in block order.
__global__ void __launch_bounds__(256) k_series_rows(const double *__restrict__ rows, int nblocks,
                                                     double *__restrict__ series, i64 top) {
    const double *row = rows + i64(blockIdx.x) * nblocks;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = int(threadIdx.x); i < nblocks; i += int(blockDim.x)) c += row[i];
    block_sum3(a, b, c);
    if (threadIdx.x == 0) series[top - i64(blockIdx.x)] += c;
}

// The per-block isotherm sums added in block order.
__global__ void __launch_bounds__(1024) k_iso_final(const double *__restrict__ part, int nblocks,
                                                    double *__restrict__ grad) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = int(threadIdx.x); i < nblocks; i += int(blockDim.x)) {
        a += part[2 * i64(i)];
        b += part[2 * i64(i) + 1];
    }
    block_sum3(a, b, c);
    if (threadIdx.x == 0) {
        grad[0] = a;
        grad[1] = b;
    }
}

// 16-bit checkpoints of an FP32 run: f about f_ref and g about the source cell's concentration
// (this cell's, or the x - 1 cell's on a Neumann outlet plane, whose populations it holds) relative to
// it (encode_g), the layout the 16-bit forward step itself stores; q and phi are copied in FP32.
template <int CP>
__global__ void __launch_bounds__(256) k_ckpt_encode(const float *__restrict__ f, const float *__restrict__ g,
                                                     const float *__restrict__ phi,
                                                     typename Storage<CP>::S *__restrict__ hf,
                                                     typename Storage<CP>::S *__restrict__ hg, i64 n, int nx,
                                                     int ad_outlet, float f_ref) {
    using S = typename Storage<CP>::S;
    const i64 c = gtid();
    if (c >= n) return;
    const i64 sref = (ad_outlet && c % nx == nx - 1) ? c - 1 : c;
#pragma unroll
    for (int q = 0; q < 19; ++q) hf[i64(q) * n + c] = S(f[i64(q) * n + c] * float(Q19::inv_w[q]) - f_ref);
    const float ph = phi[sref];
    const float inv_s = 1.0f / g_scale(ph);
#pragma unroll
    for (int q = 0; q < 7; ++q) hg[i64(q) * n + c] = encode_g<CP>(g[i64(q) * n + c], float(Q7::inv_w[q]), ph, inv_s);
}

template <int CP>
__global__ void __launch_bounds__(256) k_ckpt_decode(const typename Storage<CP>::S *__restrict__ hf,
                                                     const typename Storage<CP>::S *__restrict__ hg,
                                                     const float *__restrict__ phi, float *__restrict__ f,
                                                     float *__restrict__ g, i64 n, int nx, int ad_outlet,
                                                     float f_ref) {
    const i64 c = gtid();
    if (c >= n) return;
    const i64 sref = (ad_outlet && c % nx == nx - 1) ? c - 1 : c;
#pragma unroll
    for (int q = 0; q < 19; ++q) f[i64(q) * n + c] = w19<float>(q) * (f_ref + widen(hf[i64(q) * n + c]));
    const float ph = phi[sref];
#pragma unroll
    for (int q = 0; q < 7; ++q) g[i64(q) * n + c] = decode_g<CP>(hg[i64(q) * n + c], w7<float>(q), ph);
}

// Element-wise precision change of one state section (FP32 checkpoints of an FP64 run).
template <class From, class To>
__global__ void __launch_bounds__(256) k_cast(const From *__restrict__ src, To *__restrict__ dst, i64 count) {
    for (i64 i = gtid(); i < count; i += i64(gridDim.x) * blockDim.x) dst[i] = To(src[i]);
}

// ------------------------------------------------------------- host side
inline i64 align256(i64 b) { return (b + 255) & ~i64(255); }

template <class From, class To> static int cast_section(const void *src, void *dst, i64 count, hipStream_t s) {
    const i64 want = (count + 255) / 256;
    const unsigned grid = unsigned(want < 1 ? 1 : (want > 65536 ? 65536 : want));
    hipLaunchKernelGGL((k_cast<From, To>), dim3(grid), dim3(256), 0, s, static_cast<const From *>(src),
                       static_cast<To *>(dst), count);
    return after_launch("k_cast");
}

// Reverse steps whose per-block inlet-series sums the workspace holds before they are added up.
constexpr int kSeriesRows = 64;

// Workspace of a reverse sweep: 2 running isotherm sums per launch block, and with a per-step inlet
// series kSeriesRows rows of one sum per block.
inline i64 workspace_bytes(const fmx_coupled_desc &d) {
    const CoupledParams<float> prm = make_params<float>(d);
    const i64 nblocks = coupled_blocks(prm.gx, prm.gy, prm.gz);
    return (2 + (d.ad_inlet == 2 ? kSeriesRows : 0)) * nblocks * i64(sizeof(double));
}

// One stored state: f (19 planes), g (7 planes), q, phi, each section 256-byte aligned.
template <int P> struct StateView {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    S *f;
    S *g;
    C *q;
    C *phi;
};

template <int P> i64 state_bytes(i64 n) {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    return align256(19 * n * i64(sizeof(S))) + align256(7 * n * i64(sizeof(S))) + 2 * align256(n * i64(sizeof(C)));
}

template <int P> StateView<P> state_at(void *base, i64 slot, i64 n) {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    char *b = static_cast<char *>(base) + slot * state_bytes<P>(n);
    StateView<P> v;
    v.f = reinterpret_cast<S *>(b);
    b += align256(19 * n * i64(sizeof(S)));
    v.g = reinterpret_cast<S *>(b);
    b += align256(7 * n * i64(sizeof(S)));
    v.q = reinterpret_cast<C *>(b);
    b += align256(n * i64(sizeof(C)));
    v.phi = reinterpret_cast<C *>(b);
    return v;
}

template <int P> static int save_state(const fmx_coupled_bufs &b, int cur, const StateView<P> &dst, i64 n, hipStream_t s) {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    if (int rc = check_hip(hipMemcpyAsync(dst.f, b.f[cur], 19 * n * sizeof(S), hipMemcpyDeviceToDevice, s), "save f"))
        return rc;
    if (int rc = check_hip(hipMemcpyAsync(dst.g, b.g[cur], 7 * n * sizeof(S), hipMemcpyDeviceToDevice, s), "save g"))
        return rc;
    if (int rc = check_hip(hipMemcpyAsync(dst.q, b.q, n * sizeof(C), hipMemcpyDeviceToDevice, s), "save q")) return rc;
    return check_hip(hipMemcpyAsync(dst.phi, b.phi[cur], n * sizeof(C), hipMemcpyDeviceToDevice, s), "save phi");
}

template <int P> static int restore_state(const fmx_coupled_bufs &b, int cur, const StateView<P> &src, i64 n,
                                          hipStream_t s) {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    if (int rc = check_hip(hipMemcpyAsync(b.f[cur], src.f, 19 * n * sizeof(S), hipMemcpyDeviceToDevice, s), "load f"))
        return rc;
    if (int rc = check_hip(hipMemcpyAsync(b.g[cur], src.g, 7 * n * sizeof(S), hipMemcpyDeviceToDevice, s), "load g"))
        return rc;
    if (int rc = check_hip(hipMemcpyAsync(b.q, src.q, n * sizeof(C), hipMemcpyDeviceToDevice, s), "load q")) return rc;
    return check_hip(hipMemcpyAsync(b.phi[cur], src.phi, n * sizeof(C), hipMemcpyDeviceToDevice, s), "load phi");
}

// The current FP32 state of b into a 16-bit checkpoint slot.
template <int CP>
static int save_half_ckpt(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, const StateView<CP> &dst,
                          i64 n, hipStream_t s) {
    hipLaunchKernelGGL(k_ckpt_encode<CP>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const float *>(b.f[cur]),
                       static_cast<const float *>(b.g[cur]), static_cast<const float *>(b.phi[cur]), dst.f, dst.g, n,
                       d.nx, d.ad_outlet, float(d.f_ref));
    if (int rc = after_launch("k_ckpt_encode")) return rc;
    if (int rc = check_hip(hipMemcpyAsync(dst.q, b.q, n * sizeof(float), hipMemcpyDeviceToDevice, s), "save q")) return rc;
    return check_hip(hipMemcpyAsync(dst.phi, b.phi[cur], n * sizeof(float), hipMemcpyDeviceToDevice, s), "save phi");
}

// A 16-bit checkpoint slot decoded into an FP32 state slot.
template <int CP>
static int load_half_ckpt(const fmx_coupled_desc &d, const StateView<CP> &src, const StateView<FMX_F32> &dst, i64 n,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_ckpt_decode<CP>, dim3(blocks_for(n)), dim3(256), 0, s, src.f, src.g, src.phi, dst.f, dst.g,
                       n, d.nx, d.ad_outlet, float(d.f_ref));
    if (int rc = after_launch("k_ckpt_decode")) return rc;
    if (int rc = check_hip(hipMemcpyAsync(dst.q, src.q, n * sizeof(float), hipMemcpyDeviceToDevice, s), "load q"))
        return rc;
    return check_hip(hipMemcpyAsync(dst.phi, src.phi, n * sizeof(float), hipMemcpyDeviceToDevice, s), "load phi");
}

// The current FP64 state of b rounded into an FP32 checkpoint slot, and such a slot widened back.
static int save_narrow_ckpt(const fmx_coupled_bufs &b, int cur, const StateView<FMX_F32> &dst, i64 n, hipStream_t s) {
    if (int rc = cast_section<double, float>(b.f[cur], dst.f, 19 * n, s)) return rc;
    if (int rc = cast_section<double, float>(b.g[cur], dst.g, 7 * n, s)) return rc;
    if (int rc = cast_section<double, float>(b.q, dst.q, n, s)) return rc;
    return cast_section<double, float>(b.phi[cur], dst.phi, n, s);
}

static int load_narrow_ckpt(const fmx_coupled_bufs &b, int cur, const StateView<FMX_F32> &src, i64 n, hipStream_t s) {
    if (int rc = cast_section<float, double>(src.f, b.f[cur], 19 * n, s)) return rc;
    if (int rc = cast_section<float, double>(src.g, b.g[cur], 7 * n, s)) return rc;
    if (int rc = cast_section<float, double>(src.q, b.q, n, s)) return rc;
    return cast_section<float, double>(src.phi, b.phi[cur], n, s);
}

// The state of b into checkpoint slot j: as stored, (FP32 run, 16-bit checkpoints) encoded, or (FP64
// run, FP32 checkpoints) rounded.
template <int P>
static int save_ckpt(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, void *ckpt, int cp, i64 j, i64 n,
                     hipStream_t s) {
    if constexpr (P == FMX_F32) {
        if (cp == FMX_F16) return save_half_ckpt<FMX_F16>(d, b, cur, state_at<FMX_F16>(ckpt, j, n), n, s);
        if (cp == FMX_BF16) return save_half_ckpt<FMX_BF16>(d, b, cur, state_at<FMX_BF16>(ckpt, j, n), n, s);
    }
    if constexpr (P == FMX_F64) {
        if (cp == FMX_F32) return save_narrow_ckpt(b, cur, state_at<FMX_F32>(ckpt, j, n), n, s);
    }
    return save_state<P>(b, cur, state_at<P>(ckpt, j, n), n, s);
}

// One forward step from stored state a into stored state b (no outlet sums, no inlet_raw).
template <int P>
static int step_between(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, const StateView<P> &a,
                        const StateView<P> &o, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                        hipStream_t s) {
    using C = typename Storage<P>::C;
    CoupledPtrs<P> ptr{};
    ptr.f = a.f;
    ptr.fn = o.f;
    ptr.g = a.g;
    ptr.gn = o.g;
    ptr.q = a.q;
    ptr.q_new = o.q;
    ptr.phi = a.phi;
    ptr.phi_new = o.phi;
    ptr.links = b.links;
    ptr.field = static_cast<const C *>(b.field);
    ptr.tau_field = static_cast<const C *>(b.tau_field);
    ptr.inlet_raw = nullptr;
    ptr.inlet_series = b.inlet_series;
    ptr.outlet_acc = nullptr;
    return launch_coupled_step<P>(d.ns_op, ptr, prm, step, 0, s);
}

template <int P>
static int adjoint_step(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, const StateView<P> &st,
                        const fmx_adjoint_bufs &adj, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                        i64 widx, int lam_cur, double *iso_part, double *ser_row, hipStream_t s) {
    using C = typename Storage<P>::C;
    AdjointPtrs<P> ap{};
    ap.iso_part = iso_part;
    ap.ser_row = ser_row;
    ap.f = st.f;
    ap.g = st.g;
    ap.q = st.q;
    ap.phi = st.phi;
    ap.links = b.links;
    ap.field = static_cast<const C *>(b.field);
    ap.tau_field = static_cast<const C *>(b.tau_field);
    ap.inlet_series = b.inlet_series;
    ap.lf_in = static_cast<const C *>(adj.lam_f[lam_cur]);
    ap.lg_in = static_cast<const C *>(adj.lam_g[lam_cur]);
    ap.lf_out = static_cast<C *>(adj.lam_f[1 - lam_cur]);
    ap.lg_out = static_cast<C *>(adj.lam_g[1 - lam_cur]);
    ap.lam_q = static_cast<C *>(adj.lam_q);
    ap.lam_phi = static_cast<C *>(adj.lam_phi);
    ap.grad_field = adj.grad_field;
    ap.grad_series = adj.grad_series;
    ap.outlet_weight = adj.outlet_weight;
    const dim3 grid(coupled_blocks(prm.gx, prm.gy, prm.gz)), block(prm.block_x, prm.block_y, 1);
    const double inv_count = 1.0 / double(i64(d.ny) * d.nz);
    // 64-bit offsets when a pair of FP64 planes could pass 4 GiB (FMX_WIDE=1 forces them)
    const char *fw = std::getenv("FMX_WIDE");
    const bool wide = (fw && std::atoi(fw) != 0) || 4 * prm.n * i64(sizeof(double)) > (i64(1) << 32);
#define FMX_ADJ_LAUNCH(OPV, W) \
    hipLaunchKernelGGL((k_adjoint<P, OPV, W>), grid, block, 0, s, ap, prm, step, widx, inv_count)
    switch (d.ns_op) {
        case FMX_NS_BGK:
            if (wide) FMX_ADJ_LAUNCH(FMX_NS_BGK, true);
            else FMX_ADJ_LAUNCH(FMX_NS_BGK, false);
            break;
        case FMX_NS_POROUS:
            if (wide) FMX_ADJ_LAUNCH(FMX_NS_POROUS, true);
            else FMX_ADJ_LAUNCH(FMX_NS_POROUS, false);
            break;
        default:
            if (wide) FMX_ADJ_LAUNCH(FMX_NS_PBB, true);
            else FMX_ADJ_LAUNCH(FMX_NS_PBB, false);
            break;
    }
#undef FMX_ADJ_LAUNCH
    return after_launch("k_adjoint");
}

// Checkpoints at the storage precision, 16-bit checkpoints of an FP32 run, or FP32 checkpoints of an
// FP64 run (whose reverse sweep then runs in FP32).
static bool ckpt_prec_ok(int prec, int cp) {
    return cp == prec || (prec == FMX_F32 && (cp == FMX_F16 || cp == FMX_BF16)) || (prec == FMX_F64 && cp == FMX_F32);
}
// 16-bit checkpoints of an FP32 run: decoded into a segment slot before the segment is recomputed
static bool decoded_ckpt(int prec, int cp) { return prec == FMX_F32 && cp != FMX_F32; }

static int adjoint_check(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int64_t step0, int32_t nsteps,
                         int32_t k, int cp, const char *who, bool writes_outlet = false) {
    if (int rc = fmx_coupled_validate(d, b)) return rc;
    if (!ckpt_prec_ok(d->prec, cp))
        return set_error(FMX_ERR_UNSUPPORTED,
                         "%s: checkpoint precision %d with storage precision %d (checkpoints keep the storage "
                         "precision, or FP16 / BF16 for an FP32 run, or FP32 for an FP64 run)",
                         who, cp, d->prec);

    if (step0 < 0 || nsteps < 1 || k < 1) return set_error(FMX_ERR_ARG, "%s: bad step range or interval", who);
    return check_run_buffers(*d, *b, step0, nsteps, who, writes_outlet);
}

// The loss weights a reverse sweep reads: outlet_weight[0 .. need-1].
static int check_weights(const fmx_adjoint_bufs *adj, i64 need, const char *who) {
    if (adj && adj->outlet_weight && adj->outlet_weight_len < need)
        return set_error(FMX_ERR_ARG, "%s: outlet_weight has %lld entries, %lld needed", who,
                         (long long)adj->outlet_weight_len, (long long)need);
    return FMX_OK;
}

template <int P>
static int forward_ckpt(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, int nsteps, int k, int cur,
                        void *ckpt, int cp, hipStream_t s) {
    const i64 n = i64(d.nx) * d.ny * d.nz;
    if (int rc = clear_frozen(d, b.links, s)) return rc;
    fmx_coupled_bufs bf = b;
    bf.compact_work = nullptr;  // every cell is on the full path while it is recorded for the sweep
    const int nseg = (nsteps + k - 1) / k;
    for (int j = 0; j < nseg; ++j) {
        const int len = (j + 1) * k <= nsteps ? k : nsteps - j * k;
        if (int rc = save_ckpt<P>(d, b, cur, ckpt, cp, j, n, s)) return rc;
        if (int rc = run_coupled_steps<P>(d, bf, step0 + i64(j) * k, len, cur, s, i64(j) * k)) return rc;
        cur = (cur + len) & 1;
    }
    return FMX_OK;
}

// Segment-store slots a reverse sweep over segments of `k` steps needs: one state per step of the
// segment after its first (plus the decoded checkpoint when checkpoints are 16-bit), or, two-level,
// sub-checkpoints every m steps and one sub-segment of recomputed states: (nsub - 1) + (m - 1).
inline i64 seg_slots_single(int k, bool recast) { return recast ? k : (k > 1 ? k - 1 : 0); }
inline int sub_length(int len, bool recast, i64 slots) {
    for (int m = 1; m <= len; ++m) {
        const int nsub = (len + m - 1) / m;
        if ((recast ? 1 : 0) + (nsub - 1) + (m - 1) <= slots) return m;
    }
    return 0;
}
inline i64 seg_slots_min(int k, bool recast) {
    for (i64 slots = 0;; ++slots)
        if (sub_length(k, recast, slots) > 0) return slots;
}

template <int P>
static int backward(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, int nsteps, int k,
                    void *ckpt, int cp, void *seg, i64 seg_slots, const fmx_adjoint_bufs &adj, hipStream_t s) {
    using C = typename Storage<P>::C;
    const i64 n = i64(d.nx) * d.ny * d.nz;
    const CoupledParams<C> prm = make_params<C>(d);
    if (int rc = clear_frozen(d, b.links, s)) return rc;
    // terminal condition: nothing after the last step depends on its state
    struct Z {
        void *p;
        size_t bytes;
    } zeros[] = {{adj.lam_f[0], size_t(19 * n) * sizeof(C)},
                 {adj.lam_g[0], size_t(7 * n) * sizeof(C)},
                 {adj.lam_q, size_t(n) * sizeof(C)},
                 {adj.lam_phi, size_t(n) * sizeof(C)},
                 {adj.grad_field, adj.grad_field ? size_t(n) * sizeof(double) : 0},
                 {adj.grad_series ? adj.grad_series + step0 : nullptr, size_t(nsteps) * sizeof(double)}};
    for (const Z &zz : zeros)
        if (zz.p && zz.bytes)
            if (int rc = fmx_fill_zero(zz.p, zz.bytes, s)) return rc;  // a kernel: safe under graph capture
    // the workspace: per-block isotherm sums, then the per-block series sums of up to kSeriesRows
    // reverse steps
    const int nblocks = int(coupled_blocks(prm.gx, prm.gy, prm.gz));
    const bool want_series = d.ad_inlet == 2 && adj.grad_series;
    double *iso_part = static_cast<double *>(adj.workspace);
    double *ser_rows = iso_part + 2 * i64(nblocks);
    if (int rc = fmx_fill_zero(iso_part, 2 * size_t(nblocks) * sizeof(double), s))
        return rc;
    int rows = 0;  // rows filled; row r holds reverse step top - r
    i64 top = 0;
    auto flush_rows = [&]() -> int {
        if (rows == 0) return FMX_OK;
        hipLaunchKernelGGL(k_series_rows, dim3(rows), dim3(256), 0, s, ser_rows, nblocks, adj.grad_series, top);
        rows = 0;
        return after_launch("k_series_rows");
    };
    const bool recast = cp != P;  // 16-bit checkpoints of an FP32 run: decoded into seg slot 0
    const int nseg = (nsteps + k - 1) / k;
    int lam = 0;
    for (int j = nseg - 1; j >= 0; --j) {
        const int len = (j + 1) * k <= nsteps ? k : nsteps - j * k;
        const i64 s0 = step0 + i64(j) * k;
        if constexpr (P == FMX_F32) {
            if (cp == FMX_F16)
                if (int rc = load_half_ckpt<FMX_F16>(d, state_at<FMX_F16>(ckpt, j, n), state_at<P>(seg, 0, n), n, s))
                    return rc;
            if (cp == FMX_BF16)
                if (int rc = load_half_ckpt<FMX_BF16>(d, state_at<FMX_BF16>(ckpt, j, n), state_at<P>(seg, 0, n), n, s))
                    return rc;
        }
        // Sub-segments of m steps (m = len: one level).  Slots: [decoded checkpoint], the states at
        // the starts of sub-segments 1 .. nsub-1, then the recomputed states 1 .. m-1 of one sub-segment.
        const int m = seg_slots >= seg_slots_single(len, recast) ? len : sub_length(len, recast, seg_slots);
        const int nsub = (len + m - 1) / m;
        const i64 base = recast ? 1 : 0;
        // state at the start of step s0 + i*m + r of segment j (r < m)
        auto state = [&](int i, int r) {
            if (r > 0) return state_at<P>(seg, base + (nsub - 1) + (r - 1), n);
            if (i > 0) return state_at<P>(seg, base + (i - 1), n);
            return recast ? state_at<P>(seg, 0, n) : state_at<P>(ckpt, j, n);
        };
        // forward through sub-segments 0 .. nsub-2 to their successors' starting states
        for (int i = 0; i + 1 < nsub; ++i)
            for (int r = 1; r <= m; ++r) {
                const auto dst = r < m ? state(i, r) : state(i + 1, 0);
                if (int rc = step_between<P>(d, b, state(i, r - 1), dst, prm, s0 + i64(i) * m + r - 1, s)) return rc;
            }
        // each sub-segment from the last: recompute its states, then its adjoint steps in reverse
        for (int i = nsub - 1; i >= 0; --i) {
            const int mi = (i + 1) * m <= len ? m : len - i * m;
            const i64 t0 = s0 + i64(i) * m;
            for (int r = 1; r < mi; ++r)
                if (int rc = step_between<P>(d, b, state(i, r - 1), state(i, r), prm, t0 + r - 1, s)) return rc;
            for (int r = mi - 1; r >= 0; --r) {
                const i64 t = t0 + r;
                double *row = nullptr;
                if (want_series) {
                    if (rows > 0 && t != top - rows)
                        if (int rc = flush_rows()) return rc;
                    if (rows == 0) top = t;
                    row = ser_rows + i64(rows) * nblocks;
                }
                if (int rc = adjoint_step<P>(d, b, state(i, r), adj, prm, t, t - step0, lam, iso_part, row, s)) return rc;
                lam ^= 1;
                if (want_series && ++rows == kSeriesRows)
                    if (int rc = flush_rows()) return rc;
            }
        }
    }
    if (int rc = flush_rows()) return rc;
    hipLaunchKernelGGL(k_iso_final, dim3(1), dim3(1024), 0, s, iso_part, nblocks, adj.grad_iso);
    return after_launch("k_iso_final");
}

// One reverse step on the state held in b (f[cur], g[cur], q, phi[cur]: the state at the start of
// `step`), for callers that keep their own checkpoints: the kernel and sums of the sweep in
// backward(), with outlet_weight indexed by the step itself; grad_iso is rewritten from the running
// per-block sums after every step.
template <int P>
static int single_step(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, i64 step, int lam_cur,
                       bool first, const fmx_adjoint_bufs &adj, hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const i64 n = i64(d.nx) * d.ny * d.nz;
    const CoupledParams<C> prm = make_params<C>(d);
    const int nblocks = int(coupled_blocks(prm.gx, prm.gy, prm.gz));
    double *iso_part = static_cast<double *>(adj.workspace);
    if (first) {  // the terminal condition, and the accumulators cleared
        if (int rc = clear_frozen(d, b.links, s)) return rc;
        struct Z {
            void *p;
            size_t bytes;
        } zeros[] = {{adj.lam_f[lam_cur], size_t(19 * n) * sizeof(C)},
                     {adj.lam_g[lam_cur], size_t(7 * n) * sizeof(C)},
                     {adj.lam_q, size_t(n) * sizeof(C)},
                     {adj.lam_phi, size_t(n) * sizeof(C)},
                     {adj.grad_field, adj.grad_field ? size_t(n) * sizeof(double) : 0},
                     {iso_part, 2 * size_t(nblocks) * sizeof(double)}};
        for (const Z &zz : zeros)
            if (zz.p && zz.bytes)
                if (int rc = fmx_fill_zero(zz.p, zz.bytes, s)) return rc;
    }
    const StateView<P> st{static_cast<S *>(b.f[cur]), static_cast<S *>(b.g[cur]), static_cast<C *>(b.q),
                          static_cast<C *>(b.phi[cur])};
    const bool want_series = d.ad_inlet == 2 && adj.grad_series;
    double *row = want_series ? iso_part + 2 * i64(nblocks) : nullptr;
    if (int rc = adjoint_step<P>(d, b, st, adj, prm, step, step, lam_cur, iso_part, row, s)) return rc;
    if (want_series) {
        hipLaunchKernelGGL(k_series_rows, dim3(1), dim3(256), 0, s, row, nblocks, adj.grad_series, step);
        if (int rc = after_launch("k_series_rows")) return rc;
    }
    hipLaunchKernelGGL(k_iso_final, dim3(1), dim3(1024), 0, s, iso_part, nblocks, adj.grad_iso);
    return after_launch("k_iso_final");
}

}  // namespace fmx

using namespace fmx;

extern "C" {

int64_t fmx_adjoint_state_bytes(const fmx_coupled_desc *d) {
    if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1) return -1;
    const i64 n = i64(d->nx) * d->ny * d->nz;
    switch (d->prec) {
        case FMX_F32: return state_bytes<FMX_F32>(n);
        case FMX_F64: return state_bytes<FMX_F64>(n);
        case FMX_F16: return state_bytes<FMX_F16>(n);
        case FMX_BF16: return state_bytes<FMX_BF16>(n);
        default: return -1;
    }
}

int64_t fmx_adjoint_ckpt_bytes(const fmx_coupled_desc *d, int ckpt_prec) {
    if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1 || !ckpt_prec_ok(d->prec, ckpt_prec)) return -1;
    const i64 n = i64(d->nx) * d->ny * d->nz;
    switch (ckpt_prec) {
        case FMX_F32: return state_bytes<FMX_F32>(n);
        case FMX_F64: return state_bytes<FMX_F64>(n);
        case FMX_F16: return state_bytes<FMX_F16>(n);
        default: return state_bytes<FMX_BF16>(n);
    }
}

int fmx_adjoint_forward(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int64_t step0, int32_t nsteps,
                        int32_t k, int cur, void *ckpt, int ckpt_prec, void *stream) {
    FMX_TRY_BEGIN
    if (int rc = adjoint_check(d, b, step0, nsteps, k, ckpt_prec, "fmx_adjoint_forward", true)) return rc;
    if ((cur != 0 && cur != 1) || !ckpt) return set_error(FMX_ERR_ARG, "fmx_adjoint_forward: bad cur or checkpoint store");
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return forward_ckpt<FMX_F32>(*d, *b, step0, nsteps, k, cur, ckpt, ckpt_prec, s);
        case FMX_F64: return forward_ckpt<FMX_F64>(*d, *b, step0, nsteps, k, cur, ckpt, ckpt_prec, s);
        case FMX_F16: return forward_ckpt<FMX_F16>(*d, *b, step0, nsteps, k, cur, ckpt, ckpt_prec, s);
        default: return forward_ckpt<FMX_BF16>(*d, *b, step0, nsteps, k, cur, ckpt, ckpt_prec, s);
    }
    FMX_TRY_END
}

int64_t fmx_adjoint_workspace_bytes(const fmx_coupled_desc *d) {
    if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1) return -1;
    return workspace_bytes(*d);
}

int64_t fmx_adjoint_seg_slots(int32_t k, int recast, int two_level) {
    if (k < 1) return -1;
    return two_level ? seg_slots_min(k, recast != 0) : seg_slots_single(k, recast != 0);
}

int fmx_checkpoint_save(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, void *ckpt, int ckpt_prec,
                        int64_t slot, void *stream) {
    FMX_TRY_BEGIN
    if (int rc = fmx_coupled_validate(d, b)) return rc;
    if ((cur != 0 && cur != 1) || !ckpt || slot < 0) return set_error(FMX_ERR_ARG, "fmx_checkpoint_save: bad arguments");
    if (!ckpt_prec_ok(d->prec, ckpt_prec))
        return set_error(FMX_ERR_UNSUPPORTED, "fmx_checkpoint_save: checkpoint precision %d with storage precision %d",
                         ckpt_prec, d->prec);
    const i64 n = i64(d->nx) * d->ny * d->nz;
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return save_ckpt<FMX_F32>(*d, *b, cur, ckpt, ckpt_prec, slot, n, s);
        case FMX_F64: return save_ckpt<FMX_F64>(*d, *b, cur, ckpt, ckpt_prec, slot, n, s);
        case FMX_F16: return save_ckpt<FMX_F16>(*d, *b, cur, ckpt, ckpt_prec, slot, n, s);
        default: return save_ckpt<FMX_BF16>(*d, *b, cur, ckpt, ckpt_prec, slot, n, s);
    }
    FMX_TRY_END
}

int fmx_checkpoint_load(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, const void *ckpt, int ckpt_prec,
                        int64_t slot, void *stream) {
    FMX_TRY_BEGIN
    if (int rc = fmx_coupled_validate(d, b)) return rc;
    if ((cur != 0 && cur != 1) || !ckpt || slot < 0) return set_error(FMX_ERR_ARG, "fmx_checkpoint_load: bad arguments");
    if (!ckpt_prec_ok(d->prec, ckpt_prec))
        return set_error(FMX_ERR_UNSUPPORTED, "fmx_checkpoint_load: checkpoint precision %d with storage precision %d",
                         ckpt_prec, d->prec);
    const i64 n = i64(d->nx) * d->ny * d->nz;
    hipStream_t s = as_stream(stream);
    void *ck = const_cast<void *>(ckpt);
    int rc = FMX_OK;
    if (d->prec == FMX_F32 && ckpt_prec != FMX_F32) {
        const StateView<FMX_F32> dst{static_cast<float *>(b->f[cur]), static_cast<float *>(b->g[cur]),
                                     static_cast<float *>(b->q), static_cast<float *>(b->phi[cur])};
        rc = ckpt_prec == FMX_F16 ? load_half_ckpt<FMX_F16>(*d, state_at<FMX_F16>(ck, slot, n), dst, n, s)
                                  : load_half_ckpt<FMX_BF16>(*d, state_at<FMX_BF16>(ck, slot, n), dst, n, s);
    } else if (d->prec == FMX_F64 && ckpt_prec == FMX_F32) {
        rc = load_narrow_ckpt(*b, cur, state_at<FMX_F32>(ck, slot, n), n, s);
    } else {
        switch (d->prec) {
            case FMX_F32: rc = restore_state<FMX_F32>(*b, cur, state_at<FMX_F32>(ck, slot, n), n, s); break;
            case FMX_F64: rc = restore_state<FMX_F64>(*b, cur, state_at<FMX_F64>(ck, slot, n), n, s); break;
            case FMX_F16: rc = restore_state<FMX_F16>(*b, cur, state_at<FMX_F16>(ck, slot, n), n, s); break;
            default: rc = restore_state<FMX_BF16>(*b, cur, state_at<FMX_BF16>(ck, slot, n), n, s);
        }
    }
    if (rc) return rc;
    // frozen solid cells keep their other phase in f[1 - cur], which a checkpoint does not hold: they
    // are marked again from the restored state (desc.freeze_solids), or every cell goes on with the
    // full update (the same values)
    return mark_frozen(*d, *b, cur, s);
    FMX_TRY_END
}

int fmx_adjoint_step(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, int64_t step, int lam_cur,
                     int first, const fmx_adjoint_bufs *adj, void *stream) {
    FMX_TRY_BEGIN
    if (int rc = adjoint_check(d, b, step, 1, 1, d ? d->prec : 0, "fmx_adjoint_step")) return rc;
    if ((cur != 0 && cur != 1) || (lam_cur != 0 && lam_cur != 1) || !adj)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_step: bad arguments");
    if (!adj->lam_f[0] || !adj->lam_f[1] || !adj->lam_g[0] || !adj->lam_g[1] || !adj->lam_q || !adj->lam_phi ||
        !adj->grad_iso)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_step: missing adjoint buffer");
    const i64 wb = workspace_bytes(*d);
    if (!adj->workspace || adj->workspace_bytes < wb)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_step: workspace of %lld bytes given, %lld needed",
                         (long long)(adj->workspace ? adj->workspace_bytes : 0), (long long)wb);
    if (d->ad_inlet != 2 && adj->grad_series)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_step: grad_series needs the per-step inlet series (ad_inlet 2)");
    if (int rc = check_weights(adj, step + 1, "fmx_adjoint_step")) return rc;
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return single_step<FMX_F32>(*d, *b, cur, step, lam_cur, first != 0, *adj, s);
        case FMX_F64: return single_step<FMX_F64>(*d, *b, cur, step, lam_cur, first != 0, *adj, s);
        case FMX_F16: return single_step<FMX_F16>(*d, *b, cur, step, lam_cur, first != 0, *adj, s);
        default: return single_step<FMX_BF16>(*d, *b, cur, step, lam_cur, first != 0, *adj, s);
    }
    FMX_TRY_END
}

int fmx_adjoint_backward(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int64_t step0, int32_t nsteps,
                         int32_t k, const void *ckpt, int ckpt_prec, void *seg, int64_t seg_slots,
                         const fmx_adjoint_bufs *adj, void *stream) {
    FMX_TRY_BEGIN
    if (int rc = adjoint_check(d, b, step0, nsteps, k, ckpt_prec, "fmx_adjoint_backward")) return rc;
    const bool recast = decoded_ckpt(d->prec, ckpt_prec);
    const int kk = nsteps < k ? nsteps : k;
    const i64 need = seg_slots_min(kk, recast);
    if (!ckpt || !adj || (need > 0 && (!seg || seg_slots < need)))
        return set_error(FMX_ERR_ARG,
                         "fmx_adjoint_backward: missing checkpoint or segment store (segments of %d steps need at "
                         "least %lld segment slots, %lld for one level; %lld given)",
                         kk, (long long)need, (long long)seg_slots_single(kk, recast), (long long)seg_slots);
    if (!adj->lam_f[0] || !adj->lam_f[1] || !adj->lam_g[0] || !adj->lam_g[1] || !adj->lam_q || !adj->lam_phi ||
        !adj->grad_iso)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_backward: missing adjoint buffer");
    const i64 wb = workspace_bytes(*d);
    if (!adj->workspace || adj->workspace_bytes < wb)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_backward: workspace of %lld bytes given, %lld needed "
                                      "(fmx_adjoint_workspace_bytes)",
                         (long long)(adj->workspace ? adj->workspace_bytes : 0), (long long)wb);
    if (d->ad_inlet != 2 && adj->grad_series)
        return set_error(FMX_ERR_ARG, "fmx_adjoint_backward: grad_series needs the per-step inlet series (ad_inlet 2)");
    if (int rc = check_weights(adj, nsteps, "fmx_adjoint_backward")) return rc;
    hipStream_t s = as_stream(stream);
    void *ck = const_cast<void *>(ckpt);
    if (d->prec == FMX_F64 && ckpt_prec == FMX_F32) {
        // The FP32 sweep over the FP64 run's rounded checkpoints: the FP32 step and adjoint with the
        // caller's FP32 copies of the per-cell fields (the FP64 run's are not readable as FP32).
        if ((b->field && !adj->field_rev) || (b->tau_field && !adj->tau_field_rev))
            return set_error(FMX_ERR_ARG, "fmx_adjoint_backward: FP32 checkpoints of an FP64 run need field_rev / "
                                          "tau_field_rev, the FP32 copies of field / tau_field");
        fmx_coupled_desc d32 = *d;
        d32.prec = FMX_F32;
        fmx_coupled_bufs b32 = *b;
        b32.field = b->field ? adj->field_rev : nullptr;
        b32.tau_field = b->tau_field ? adj->tau_field_rev : nullptr;
        return backward<FMX_F32>(d32, b32, step0, nsteps, k, ck, FMX_F32, seg, seg_slots, *adj, s);
    }
    switch (d->prec) {
        case FMX_F32: return backward<FMX_F32>(*d, *b, step0, nsteps, k, ck, ckpt_prec, seg, seg_slots, *adj, s);
        case FMX_F64: return backward<FMX_F64>(*d, *b, step0, nsteps, k, ck, ckpt_prec, seg, seg_slots, *adj, s);
        case FMX_F16: return backward<FMX_F16>(*d, *b, step0, nsteps, k, ck, ckpt_prec, seg, seg_slots, *adj, s);
        default: return backward<FMX_BF16>(*d, *b, step0, nsteps, k, ck, ckpt_prec, seg, seg_slots, *adj, s);
    }
    FMX_TRY_END
}

}  // extern "C"
