// ref_cpu.hpp -- CPU oracle for FluidMaxxer's coupled D3Q19 + D3Q7 step.
//
// TEST INFRASTRUCTURE ONLY. This file restates, kernel by kernel, the
// reference's CUDA update in plain C++ (array-of-structures layout
// f[cell*Q + q], x-fastest cells, the reference's expression order), templated
// on T in {float, double}; serial, or OpenMP-parallel over independent cells with
// identical results. It is the checker for the HIP product path and the CPU
// baseline; it is never linked into, loaded by or called from the
// product library.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use it.
//
// Each function names the reference file:line it follows (paths relative to the
// reference repository root).  Departures from the reference are deliberate and
// listed where they occur:
//   * 64-bit cell indices (the reference overflows `cell*19` above 113M cells,
//     BGKCollision.cu:88);
//   * k_ad_neumann gets the missing `c < N` guard (ADNeumannOutlet.cu:14);
//   * BounceBack is a true pairwise swap; the reference's in-place loop
//     (BounceBack.cu:29-30) leaves every opposite pair holding the same value;
//   * no --use_fast_math: IEEE division and no flush-to-zero;
//   * PartialBounceBack (Walsh et al. 2009) and the piecewise-linear inlet
//     series have no reference counterpart; they follow their published / stated
//     formulas and are pinned by limits (see the test suite).
#pragma once

// Built with OpenMP, the per-cell loops of the update run in parallel: each cell of such a loop
// writes only its own outputs and reads nothing another cell of the same loop writes, so the
// results do not depend on the thread count (the update has no reductions).  The boundary loops
// that read a neighbour in the array they write (pressure outlet, Neumann copy) and the outlet
// statistics stay serial.
#ifdef _OPENMP
#define ORC_PAR_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PAR_FOR
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <vector>

namespace orc {

using i64 = std::int64_t;

// ---------------------------------------------------------------- descriptors
// D3Q19.hpp:5-50 (velocity order: rest, 6 axis, 12 edge; weights T(1)/T(n)).
struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    template <class T> static T w(int q) {
        return q == 0 ? T(1) / T(3) : (q < 7 ? T(1) / T(18) : T(1) / T(36));
    }
};

// D3Q7.hpp:5-28 (w = 1/4 rest, 1/8 axis; cs^2 = 1/4).
struct D3Q7 {
    static constexpr int Q = 7;
    static constexpr int cx[7] = {0, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[7] = {0, 0, 0, 1, -1, 0, 0};
    static constexpr int cz[7] = {0, 0, 0, 0, 0, 1, -1};
    static constexpr int opp[7] = {0, 2, 1, 4, 3, 6, 5};
    template <class T> static T w(int q) { return q == 0 ? T(1) / T(4) : T(1) / T(8); }
};

struct Grid {
    int Nx, Ny, Nz;
    bool px, py, pz;
    i64 n() const { return i64(Nx) * Ny * Nz; }
    i64 idx(int x, int y, int z) const { return x + i64(Nx) * (y + i64(Ny) * z); }
    void coords(i64 c, int &x, int &y, int &z) const {
        const i64 plane = i64(Nx) * Ny;
        z = int(c / plane);
        const i64 rem = c - i64(z) * plane;
        y = int(rem / Nx);
        x = int(rem - i64(y) * Nx);
    }
    // mapCoord: periodic wrap or reject (BGKCollision.cu:26-41, ADBGKCollision.cu:16-26).
    bool map(int &x, int &y, int &z) const {
        if (px) x = (x % Nx + Nx) % Nx;
        else if (x < 0 || x >= Nx) return false;
        if (py) y = (y % Ny + Ny) % Ny;
        else if (y < 0 || y >= Ny) return false;
        if (pz) z = (z % Nz + Nz) % Nz;
        else if (z < 0 || z >= Nz) return false;
        return true;
    }
};

// ------------------------------------------------------------------ NS init
// k_init_eq_d3q19_1d (BGKCollision.cu:44-71): feq with u = 0.
template <class T> void ns_init_eq_bgk(T *f, const Grid &g, T rho0) {
    const T ux = T(0), uy = T(0), uz = T(0), u2 = T(0);
    ORC_PAR_FOR
    for (i64 c = 0; c < g.n(); ++c)
        for (int q = 0; q < 19; ++q) {
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            const T cu = T(3) * eu;
            f[c * 19 + q] = D3Q19::w<T>(q) * rho0 * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
        }
}

// k_init_eq_d3q19 (PorousBGKCollision.cu:43-58): f = w * rho0.
template <class T> void ns_init_eq_porous(T *f, const Grid &g, T rho0) {
    ORC_PAR_FOR
    for (i64 c = 0; c < g.n(); ++c)
        for (int q = 0; q < 19; ++q) f[c * 19 + q] = D3Q19::w<T>(q) * rho0;
}

// NS tau clamp + omega (BGKCollision.cu:202-205, PorousBGKCollision.cu:180-181).
template <class T> T ns_omega(T tau) {
    if (tau < T(0.5001)) tau = T(0.5001);
    return T(1) / tau;
}

// --------------------------------------------------------- NS collide+stream
// k_bgk_stream_pull_d3q19 (BGKCollision.cu:74-177).
template <class T>
void ns_collide_bgk(const T *f, T *f_new, const std::uint8_t *obst, const Grid &g, T omega, T rho0) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        const i64 base = cell * 19;
        if (obst && obst[cell]) {  // solid: on-site reverse (:90-97)
            for (int q = 0; q < 19; ++q) f_new[base + q] = f[base + D3Q19::opp[q]];
            continue;
        }
        int x, y, z;
        g.coords(cell, x, y, z);
        T fi[19];
        for (int q = 0; q < 19; ++q) {  // pull (:108-130)
            int xs = x - D3Q19::cx[q], ys = y - D3Q19::cy[q], zs = z - D3Q19::cz[q];
            if (!g.map(xs, ys, zs)) {
                fi[q] = f[base + q];
            } else {
                const i64 src = g.idx(xs, ys, zs);
                if (obst && obst[src]) fi[q] = f[base + D3Q19::opp[q]];
                else fi[q] = f[src * 19 + q];
            }
        }
        T rho = T(0), ux = T(0), uy = T(0), uz = T(0);  // (:133-146)
        for (int q = 0; q < 19; ++q) {
            const T v = fi[q];
            rho += v;
            ux += v * T(D3Q19::cx[q]);
            uy += v * T(D3Q19::cy[q]);
            uz += v * T(D3Q19::cz[q]);
        }
        if (!(rho > T(0))) {  // (:149-158)
            rho = rho0;
            ux = uy = uz = T(0);
        } else {
            const T invR = T(1) / rho;
            ux *= invR;
            uy *= invR;
            uz *= invR;
        }
        const T u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; ++q) {  // (:163-176)
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            const T cu = T(3) * eu;
            const T feq = D3Q19::w<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
            f_new[base + q] = fi[q] + omega * (feq - fi[q]);
        }
    }
}

// k_porous_bgk_stream_pull_d3q19 (PorousBGKCollision.cu:64-155).
template <class T>
void ns_collide_porous(const T *f, T *f_new, const T *porosity, const Grid &g, T omega, T rho0) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        const i64 base = cell * 19;
        int x, y, z;
        g.coords(cell, x, y, z);
        T fi[19];
        for (int q = 0; q < 19; ++q) {  // pull, no obstacle handling (:87-101)
            int xs = x - D3Q19::cx[q], ys = y - D3Q19::cy[q], zs = z - D3Q19::cz[q];
            if (!g.map(xs, ys, zs)) fi[q] = f[base + q];
            else fi[q] = f[g.idx(xs, ys, zs) * 19 + q];
        }
        T rho = T(0), jx = T(0), jy = T(0), jz = T(0);
        for (int q = 0; q < 19; ++q) {
            const T v = fi[q];
            rho += v;
            jx += v * T(D3Q19::cx[q]);
            jy += v * T(D3Q19::cy[q]);
            jz += v * T(D3Q19::cz[q]);
        }
        if (!(rho > T(0))) {  // (:118-121)
            rho = rho0;
            jx = jy = jz = T(0);
        }
        const T invR = T(1) / rho;  // (:124-127)
        T ux = jx * invR, uy = jy * invR, uz = jz * invR;
        T phi = porosity ? porosity[cell] : T(1);  // (:130-136)
        if (phi < T(0)) phi = T(0);
        if (phi > T(1)) phi = T(1);
        ux *= phi;
        uy *= phi;
        uz *= phi;
        const T u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; ++q) {
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            const T cu = T(3) * eu;
            const T feq = D3Q19::w<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
            f_new[base + q] = fi[q] + omega * (feq - fi[q]);
        }
    }
}

// Partial bounce-back BGK (no reference counterpart).  Walsh, Burwinkle & Saar,
// "A new partial-bounceback lattice-Boltzmann method for fluid flow through
// heterogeneous media", Computers & Geosciences 35 (2009) 1186-1193, eq. (5):
//   f_i(x + c_i, t+1) = (1 - n_s) f_i^c(x, t) + n_s f_{-i}(x, t)
// with f^c the BGK post-collision population and f_{-i} the pre-collision
// (post-streaming) opposite population at the same node.  In pull form the
// blend is applied where the collision happens, so the buffer keeps holding
// post-collision values.  Streaming is the porous kernel's (keep-local out of
// domain, no obstacle test); the density fallback is BGKCollision's.
template <class T>
void ns_collide_pbb(const T *f, T *f_new, const T *ns_field, const Grid &g, T omega, T rho0) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        const i64 base = cell * 19;
        int x, y, z;
        g.coords(cell, x, y, z);
        T fi[19];
        for (int q = 0; q < 19; ++q) {
            int xs = x - D3Q19::cx[q], ys = y - D3Q19::cy[q], zs = z - D3Q19::cz[q];
            if (!g.map(xs, ys, zs)) fi[q] = f[base + q];
            else fi[q] = f[g.idx(xs, ys, zs) * 19 + q];
        }
        T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
        for (int q = 0; q < 19; ++q) {
            const T v = fi[q];
            rho += v;
            ux += v * T(D3Q19::cx[q]);
            uy += v * T(D3Q19::cy[q]);
            uz += v * T(D3Q19::cz[q]);
        }
        if (!(rho > T(0))) {
            rho = rho0;
            ux = uy = uz = T(0);
        } else {
            const T invR = T(1) / rho;
            ux *= invR;
            uy *= invR;
            uz *= invR;
        }
        T ns = ns_field ? ns_field[cell] : T(0);
        if (ns < T(0)) ns = T(0);
        if (ns > T(1)) ns = T(1);
        const T u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; ++q) {
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            const T cu = T(3) * eu;
            const T feq = D3Q19::w<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
            const T fc = fi[q] + omega * (feq - fi[q]);
            f_new[base + q] = (T(1) - ns) * fc + ns * fi[D3Q19::opp[q]];
        }
    }
}

// ------------------------------------------------------------- NS moments
// k_compute_macros_d3q19 (Macros.cu:20-61).
template <class T> void ns_moments(const T *f, T *rho, T *ux, T *uy, T *uz, const Grid &g) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        const i64 base = cell * 19;
        T r = T(0), vx = T(0), vy = T(0), vz = T(0);
        for (int q = 0; q < 19; ++q) {
            const T fi = f[base + q];
            r += fi;
            vx += fi * T(D3Q19::cx[q]);
            vy += fi * T(D3Q19::cy[q]);
            vz += fi * T(D3Q19::cz[q]);
        }
        const T invR = (r > T(0)) ? (T(1) / r) : T(0);
        rho[cell] = r;
        ux[cell] = vx * invR;
        uy[cell] = vy * invR;
        uz[cell] = vz * invR;
    }
}

// ---------------------------------------------------------------- NS BCs
// VelocityDirichletNEQ host set-up (VelocityDirichletNEQ.hpp:26-107): compact the
// mask, require exactly one axis-aligned boundary plane, neighbour one cell inward.
inline void neq_build_lists(const std::uint8_t *mask, const Grid &g, std::vector<i64> &bnd,
                            std::vector<i64> &nei) {
    bnd.clear();
    nei.clear();
    for (i64 i = 0; i < g.n(); ++i)
        if (mask[i]) bnd.push_back(i);
    if (bnd.empty()) return;
    const int nx = g.Nx, ny = g.Ny, nz = g.Nz;
    int minx = nx - 1, maxx = 0, miny = ny - 1, maxy = 0, minz = nz - 1, maxz = 0;
    for (i64 c : bnd) {
        int x, y, z;
        g.coords(c, x, y, z);
        minx = std::min(minx, x); maxx = std::max(maxx, x);
        miny = std::min(miny, y); maxy = std::max(maxy, y);
        minz = std::min(minz, z); maxz = std::max(maxz, z);
    }
    const bool isX = (minx == maxx), isY = (miny == maxy), isZ = (minz == maxz);
    if (int(isX) + int(isY) + int(isZ) != 1)
        throw std::runtime_error(
            "VelocityDirichletNEQ: mask must describe exactly one axis-aligned plane (x=const OR y=const OR z=const).");
    int dx = 0, dy = 0, dz = 0;
    if (isX) {
        if (minx == 0) dx = 1;
        else if (minx == nx - 1) dx = -1;
        else throw std::runtime_error("VelocityDirichletNEQ: x-plane not at boundary");
    }
    if (isY) {
        if (miny == 0) dy = 1;
        else if (miny == ny - 1) dy = -1;
        else throw std::runtime_error("VelocityDirichletNEQ: y-plane not at boundary");
    }
    if (isZ) {
        if (minz == 0) dz = 1;
        else if (minz == nz - 1) dz = -1;
        else throw std::runtime_error("VelocityDirichletNEQ: z-plane not at boundary");
    }
    for (i64 c : bnd) {
        int x, y, z;
        g.coords(c, x, y, z);
        const int xn = x + dx, yn = y + dy, zn = z + dz;
        if (xn < 0 || xn >= nx || yn < 0 || yn >= ny || zn < 0 || zn >= nz)
            throw std::runtime_error("VelocityDirichletNEQ: inferred neighbor out of domain.");
        nei.push_back(g.idx(xn, yn, zn));
    }
}

// feq_d3q19 / compute_rho_u_d3q19 (VelocityDirichletNEQ.cu:13-46).
template <class T> T feq19(int q, T rho, T ux, T uy, T uz) {
    const T u2 = ux * ux + uy * uy + uz * uz;
    const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
    const T cu = T(3) * eu;
    return D3Q19::w<T>(q) * rho * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
}

// k_velocity_dirichlet_neq_pre_d3q19 (VelocityDirichletNEQ.cu:49-96), on f, PreCollision.
template <class T>
void ns_bc_velocity_neq(T *f, const std::vector<i64> &bnd, const std::vector<i64> &nei, T ux, T uy, T uz) {
    for (std::size_t k = 0; k < bnd.size(); ++k) {
        const i64 baseB = bnd[k] * 19, baseN = nei[k] * 19;
        T r = T(0), jx = T(0), jy = T(0), jz = T(0);
        for (int q = 0; q < 19; ++q) {
            const T fq = f[baseN + q];
            r += fq;
            jx += fq * T(D3Q19::cx[q]);
            jy += fq * T(D3Q19::cy[q]);
            jz += fq * T(D3Q19::cz[q]);
        }
        const T inv = (r > T(0)) ? (T(1) / r) : T(0);
        const T uxN = jx * inv, uyN = jy * inv, uzN = jz * inv;
        for (int q = 0; q < 19; ++q) {
            const T fN = f[baseN + q];
            const T feqN = feq19(q, r, uxN, uyN, uzN);
            const T feqB = feq19(q, r, ux, uy, uz);
            f[baseB + q] = feqB + (fN - feqN);
        }
    }
}

// k_velocity_dirichlet_eq_pre_d3q19 (VelocityDirichlet.cu:54-84), on f, PreCollision.
template <class T>
void ns_bc_velocity_eq(T *f, const std::uint8_t *mask, const Grid &g, T rho0, T ux, T uy, T uz) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        if (!mask[cell]) continue;
        const i64 base = cell * 19;
        const T u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; ++q) {
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            const T cu = T(3) * eu;
            f[base + q] = D3Q19::w<T>(q) * rho0 * (T(1) + cu + T(0.5) * cu * cu - T(1.5) * u2);
        }
    }
}

// k_pressure_dirichlet_outlet_d3q19 (PressureDirichlet.cu:26-96), on f_new, PostStreaming.
template <class T> void ns_bc_pressure(T *f, const std::uint8_t *mask, const Grid &g, T rho_out) {
    for (i64 cell = 0; cell < g.n(); ++cell) {
        if (!mask[cell]) continue;
        int x, y, z;
        g.coords(cell, x, y, z);
        if (x <= 0) continue;
        const i64 baseN = g.idx(x - 1, y, z) * 19;
        T rhoN = T(0), ux = T(0), uy = T(0), uz = T(0);
        for (int q = 0; q < 19; ++q) {
            const T fi = f[baseN + q];
            rhoN += fi;
            ux += fi * T(D3Q19::cx[q]);
            uy += fi * T(D3Q19::cy[q]);
            uz += fi * T(D3Q19::cz[q]);
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
        const i64 base = cell * 19;
        for (int q = 0; q < 19; ++q) {
            const T eu = T(D3Q19::cx[q]) * ux + T(D3Q19::cy[q]) * uy + T(D3Q19::cz[q]) * uz;
            f[base + q] = D3Q19::w<T>(q) * rho_out * (T(1) + T(3) * eu + T(4.5) * eu * eu - T(1.5) * u2);
        }
    }
}

// BounceBack (BounceBack.cu:12-31) as a true swap of each opposite pair.
template <class T> void ns_bc_bounceback(T *f, const std::uint8_t *mask, const Grid &g) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < g.n(); ++cell) {
        if (!mask[cell]) continue;
        const i64 base = cell * 19;
        T tmp[19];
        for (int q = 0; q < 19; ++q) tmp[q] = f[base + D3Q19::opp[q]];
        for (int q = 0; q < 19; ++q) f[base + q] = tmp[q];
    }
}

// ------------------------------------------------------------------ AD
// k_ad_init_eq_d3q7 (ADBGKCollision.cu:29-38).
template <class T> void ad_init_eq(T *g, const Grid &gr, T phi0) {
    ORC_PAR_FOR
    for (i64 c = 0; c < gr.n(); ++c)
        for (int q = 0; q < 7; ++q) g[c * 7 + q] = phi0 * D3Q7::w<T>(q);
}

// k_ad_bgk_stream_pull_d3q7_source (ADBGKCollision.cu:41-128).  `inlet_pre`
// (nullable) carries a PreCollision Dirichlet value already written into g by the
// caller; this function only streams and collides.
template <class T>
void ad_collide(const T *g, T *g_new, T *phi_out, const T *tau, const T *ux, const T *uy, const T *uz,
                const T *S, const Grid &gr, T phiMin) {
    ORC_PAR_FOR
    for (i64 cell = 0; cell < gr.n(); ++cell) {
        int x, y, z;
        gr.coords(cell, x, y, z);
        const T t = tau[cell];
        const T omega = T(1) / t;
        const T vx = ux[cell], vy = uy[cell], vz = uz[cell];
        const i64 base = cell * 7;
        T gi[7];
        for (int q = 0; q < 7; ++q) {
            int xs = x - D3Q7::cx[q], ys = y - D3Q7::cy[q], zs = z - D3Q7::cz[q];
            if (!gr.map(xs, ys, zs)) gi[q] = g[base + q];
            else gi[q] = g[(xs + i64(gr.Nx) * (ys + i64(gr.Ny) * zs)) * 7 + q];
        }
        T phi = T(0);
        for (int q = 0; q < 7; ++q) phi += gi[q];
        if (phi < phiMin) phi = phiMin;
        const T invCs2 = T(4);
        const T src = S ? S[cell] : T(0);
        for (int q = 0; q < 7; ++q) {
            const T eu = T(D3Q7::cx[q]) * vx + T(D3Q7::cy[q]) * vy + T(D3Q7::cz[q]) * vz;
            const T geq = D3Q7::w<T>(q) * phi * (T(1) + invCs2 * eu);
            g_new[base + q] = gi[q] + omega * (geq - gi[q]) + D3Q7::w<T>(q) * src;
        }
        if (phi_out) phi_out[cell] = phi;
    }
}

// k_ad_dirichlet (ADDirichlet.cu:11-30) and k_ad_dirichlet_tanh_pulse_d3q7
// (ADDirichletTanhPulse.cu:11-31): g <- phi * w on the mask.
template <class T> void ad_bc_dirichlet(T *g, const std::uint8_t *mask, const Grid &gr, T phi) {
    ORC_PAR_FOR
    for (i64 c = 0; c < gr.n(); ++c) {
        if (!mask[c]) continue;
        for (int q = 0; q < 7; ++q) g[c * 7 + q] = phi * D3Q7::w<T>(q);
    }
}

// Tanh pulse inlet value (ADDirichletTanhPulse.cu:72-83), host double arithmetic.
template <class T> T tanh_pulse_value(int step, T base, T amp, int t_on, int t_off, T k) {
    const double tt = double(step), ton = double(t_on), toff = double(t_off), kk = double(k);
    const double pulse = 0.5 * (std::tanh((tt - ton) / kk) - std::tanh((tt - toff) / kk));
    return base + static_cast<T>(pulse) * amp;
}

// Piecewise-linear inlet series (no reference counterpart; generalises the tanh
// pulse to a knot vector): knot k sits at step k * dt_knot; the value at `step`
// interpolates linearly in double and is rounded to T once.  Beyond the last knot
// the last value holds.
template <class T> T series_value(int step, const double *knots, int n_knots, double dt_knot) {
    if (n_knots <= 0) return T(0);
    const double s = double(step) / dt_knot;
    if (s <= 0.0) return T(knots[0]);
    const int k = int(std::floor(s));
    if (k >= n_knots - 1) return T(knots[n_knots - 1]);
    const double a = s - double(k);
    return static_cast<T>(knots[k] * (1.0 - a) + knots[k + 1] * a);
}

// k_ad_neumann (ADNeumannOutlet.cu:5-31) with the missing bounds guard.
template <class T> void ad_bc_neumann(T *g, const std::uint8_t *mask, const Grid &gr) {
    for (i64 c = 0; c < gr.n(); ++c) {
        if (!mask[c]) continue;
        int x, y, z;
        gr.coords(c, x, y, z);
        if (x != gr.Nx - 1) continue;
        const i64 src = gr.idx(x - 1, y, z);
        for (int q = 0; q < 7; ++q) g[c * 7 + q] = g[src * 7 + q];
    }
}

// ------------------------------------------------------------------ LDF
// k_henry_ldf_source (LinearLDF.cu:13-48).
template <class T>
void ldf_linear(i64 n, const std::uint8_t *region, const T *phi, T *S, T *q, T K, T kLDF, T phiMin) {
    ORC_PAR_FOR
    for (i64 i = 0; i < n; ++i) {
        if (!region[i]) {
            S[i] = T(0);
            continue;
        }
        T c = phi[i];
        if (c < phiMin) c = phiMin;
        const T cEff = c - phiMin;
        const T qeq = K * cEff;
        const T qi = q[i];
        const T dq = kLDF * (qeq - qi);
        q[i] = qi + dq;
        S[i] = -dq;
    }
}

// k_langmuir_ldf_source (LangmuirLDF.cu:7-41).
template <class T>
void ldf_langmuir(i64 n, const std::uint8_t *region, const T *phi, T *S, T *q, T qMax, T b, T kLDF, T phiMin) {
    ORC_PAR_FOR
    for (i64 i = 0; i < n; ++i) {
        if (!region[i]) {
            S[i] = T(0);
            continue;
        }
        T c = phi[i];
        if (c < phiMin) c = phiMin;
        const T cc = c - phiMin;
        const T bc = b * cc;
        const T qeq = qMax * (bc / (T(1) + bc));
        const T qi = q[i];
        const T dq = kLDF * (qeq - qi);
        q[i] = qi + dq;
        S[i] = -dq;
    }
}

// OutletCSV::log statistics (OutletCSV.hpp:51-80): avg/min/max of phi over the mask in double.
template <class T>
void outlet_stats(const T *phi, const std::uint8_t *mask, i64 n, double &avg, double &mn, double &mx) {
    double sum = 0.0;
    mn = std::numeric_limits<double>::infinity();
    mx = -std::numeric_limits<double>::infinity();
    std::size_t cnt = 0;
    for (i64 i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        const double v = static_cast<double>(phi[i]);
        sum += v;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
        ++cnt;
    }
    avg = (cnt > 0) ? sum / double(cnt) : 0.0;
    if (cnt == 0) {
        mn = 0.0;
        mx = 0.0;
    }
}

}  // namespace orc
