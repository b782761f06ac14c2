// coupled_common.hpp -- parameters, pointers, addressing and launch shape shared by the fused
// coupled step (coupled.hip) and its discrete adjoint (adjoint.hip).
#pragma once

#include <cstdio>
#include <cstdlib>

#include "fmx_exact.hpp"
#include "fmx_half.hpp"
#include "fmx_internal.hpp"

namespace fmx {

template <class C> struct CoupledParams {
    int nx, ny, nz, py, pz;
    int ns_inlet, ns_outlet, ad_inlet, ad_outlet, isotherm, coupling;
    int gx, gy, gz;          // logical block grid
    int block_x, block_y;
    int xcd_remap;
    int tile_y, tile_z;      // row tiles (0: plain y-then-z order); they divide gy and gz
    int pair;                // two cells per thread (k_coupled_pair) where the storage and nx allow it
    int freeze;              // frozen solid cells may be present (FMX_FREEZE=1): the step checks their link bit
    int lane_bc;             // boundary lanes take their x +- 1 neighbour's values by lane shuffles
    i64 n;
    C omega, rho0, uin_x, uin_y, uin_z, uin2, rho_out, phi_in, omega_ad, min_conc, iso_a, iso_b, k_ldf, f_ref;
};

template <int P> struct CoupledPtrs {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    const S *f;
    S *fn;
    const S *g;
    S *gn;
    const C *q;     // adsorbed load at the start of the step
    C *q_new;       // adsorbed load after the step (may alias q: each cell touches only its own)
    const C *phi;   // previous step's concentration (sorption input, half-storage reference of g)
    C *phi_new;     // this step's concentration
    std::uint32_t *links;
    const C *field;
    const C *tau_field;
    C *inlet_raw;   // nullable: x = 0 populations as the reference holds them between steps
    const double *inlet_series;
    double *outlet_acc;       // accumulator rows (acc_row >= 0), or this step's outlet-ring row (acc_row < 0)
};

// Logical block id -> (bx, by, bz).  With xcd_remap the hardware's round-robin
// assignment of consecutive workgroups to the 8 XCDs is inverted, so each XCD
// sweeps one contiguous range of the domain and its L2 sees the neighbouring rows
// its pulls touch.  With tiles (tile_y x tile_z rows, y fastest inside a tile and
// tiles y-fastest), the rows at z +- 1 are tile_y blocks away instead of a whole
// y-plane of blocks, so they are still in L2 when their neighbours pull from them.
template <class C> __device__ __forceinline__ void block_coords(const CoupledParams<C> &prm, int &bx, int &by, int &bz) {
    const unsigned total = unsigned(prm.gx) * unsigned(prm.gy) * unsigned(prm.gz);
    unsigned b = blockIdx.x;
    if (prm.xcd_remap && (total & 7u) == 0u) b = (b & 7u) * (total >> 3) + (b >> 3);
    bx = int(b % unsigned(prm.gx));
    const unsigned r = b / unsigned(prm.gx);
    if (prm.tile_y) {
        const unsigned tsz = unsigned(prm.tile_y) * unsigned(prm.tile_z);
        const unsigned t = r / tsz, w = r % tsz;
        const unsigned nty = unsigned(prm.gy) / unsigned(prm.tile_y);
        by = int((t % nty) * unsigned(prm.tile_y) + w % unsigned(prm.tile_y));
        bz = int((t / nty) * unsigned(prm.tile_z) + w / unsigned(prm.tile_y));
    } else {
        by = int(r % unsigned(prm.gy));
        bz = int(r / unsigned(prm.gy));
    }
}

// Population plane q of a structure-of-arrays buffer (wave-uniform pointer) and a load /
// store at a byte offset from it: with a 32-bit unsigned offset the address is a scalar base plus
// a vector offset (the host guarantees every plane is below 4 GiB).
template <class T> __device__ __forceinline__ T *plane(T *base, int q, i64 n) { return base + i64(q) * n; }
template <class T, class O> __device__ __forceinline__ T ld_b(const T *base, O boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <class T, class O> __device__ __forceinline__ void st_b(T *base, O boff, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff) = v;
}
template <class T, class O> __device__ __forceinline__ void st_b_nt(T *base, O boff, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff));
}
// Non-temporal store into a population plane whose base is wave-uniform.  With a 32-bit offset the
// base and the offset go through empty asm statements (base in scalar registers, one pair per
// plane; offset a 32-bit value of the store's own block): otherwise the compiler derives every
// plane's address from the first one's in vector registers, a 64-bit vector add per store, instead
// of using the scalar-base store.
template <class T, class O> __device__ __forceinline__ void st_plane_nt(T *base, O boff, T v) {
    if constexpr (std::is_same_v<O, unsigned>) {
        asm("" : "+s"(base));
        asm("" : "+v"(boff));
        using G = __attribute__((address_space(1))) T;
        using GC = __attribute__((address_space(1))) char;
        __builtin_nontemporal_store(v, (G *)((GC *)(G *)base + boff));
    } else {
        st_b_nt(base, boff, v);
    }
}

// A stored 16-bit value as float: FP16 through half_bits_to_float (fmx_half.hpp), BF16 is the top half
// of an FP32.
template <class S> __device__ __forceinline__ float widen(S v) {
    if constexpr (std::is_same_v<S, _Float16>) {
        unsigned short h;
        __builtin_memcpy(&h, &v, 2);
        return half_bits_to_float(h);
    } else {
        return float(v);
    }
}

// w (ref + h) for a stored deviation h
template <int P>
__device__ __forceinline__ typename Storage<P>::C decode(typename Storage<P>::S v, typename Storage<P>::C w,
                                                          typename Storage<P>::C ref) {
    if constexpr (IsHalf<P>::value) return w * (ref + widen(v));
    else return v;
}
template <int P>
__device__ __forceinline__ typename Storage<P>::S encode(typename Storage<P>::C v, typename Storage<P>::C inv_w,
                                                          typename Storage<P>::C ref) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    if constexpr (IsHalf<P>::value) {
        C t = v * inv_w - ref;
        // keep the 16-bit rounding a separate conversion: fused into a mixed-precision FMA it writes
        // half a register, which serialises on the register's previous value
        asm volatile("" : "+v"(t));
        return S(t);
    } else {
        return v;
    }
}

// 16-bit storage of an AD population: its deviation from the reference concentration phi_ref (the
// source cell's lagged concentration), relative to it, g = w (phi_ref + s h) with s = max(phi_ref, 2^-14),
// so the 16-bit rounding is relative to the local concentration level down to 2^-14, and below it
// absolute at s 2^-24 = 3.6e-12 (FP16) instead of 2^-24 = 6e-8: a front's tail near a 1e-8 clamp floor
// keeps its significant bits.  The floor keeps s positive where phi_ref is 0 or negative, and h within
// FP16's range (|h| < 65504) wherever |g / w| < 4 -- concentrations of order one, as the lattice units
// here have them (inlet value 1); a larger h would store as Inf, which the mass diagnostic counts.
template <class C> __device__ __forceinline__ C g_scale(C phi_ref) {
    return phi_ref > C(0x1p-14) ? phi_ref : C(0x1p-14);
}
template <int P>
__device__ __forceinline__ typename Storage<P>::C decode_g(typename Storage<P>::S v, typename Storage<P>::C w,
                                                            typename Storage<P>::C phi_ref) {
    if constexpr (IsHalf<P>::value) return w * (phi_ref + g_scale(phi_ref) * widen(v));
    else return v;
}
// inv_s: 1 / g_scale(phi_ref), computed once per reference by the caller
template <int P>
__device__ __forceinline__ typename Storage<P>::S encode_g(typename Storage<P>::C v, typename Storage<P>::C inv_w,
                                                            typename Storage<P>::C phi_ref,
                                                            typename Storage<P>::C inv_s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    if constexpr (IsHalf<P>::value) {
        C t = (v * inv_w - phi_ref) * inv_s;
        asm volatile("" : "+v"(t));  // a separate conversion (see encode)
        return S(t);
    } else {
        return v;
    }
}

// feq in the BGK form w rho (1 + cu + 0.5 cu^2 - 1.5 u^2), cu = 3 e.u (BGKCollision.cu:166-173)
template <int q, class C> __device__ __forceinline__ C feq_bgk(C rho, C ux, C uy, C uz, C u2x15) {
    const C eu = dot3<Q19::cx[q], Q19::cy[q], Q19::cz[q]>(ux, uy, uz);
    const C cu = C(3) * eu;
    return w19<C>(q) * rho * (C(1) + cu + C(0.5) * cu * cu - u2x15);
}

// ------------------------------------------------------------- host side
// One wave-multiple row of x per block: 256 x 1 for long rows, 128 x 2 / 64 x 4 for short ones,
// so no lane idles when nx is 64 or 128.  FMX_BLOCK=AxB and FMX_XCD_REMAP=0|1 override.
struct LaunchShape {
    int bx = 256, by = 1, remap = 1, ty = 0, tz = 0;
    bool tiled = false;  // FMX_TILE given: no default slab order
};

inline LaunchShape launch_shape(int nx) {
    LaunchShape s;
    const int rx = ((nx + 63) / 64) * 64;
    s.bx = rx < 256 ? rx : 256;
    s.by = 256 / s.bx;
    if (const char *e = std::getenv("FMX_BLOCK")) {
        int a = 0, b = 0;
        // rows of whole waves: the kernel takes y as wave-uniform
        if (std::sscanf(e, "%dx%d", &a, &b) == 2 && a > 0 && b > 0 && a * b <= 256 && a % 64 == 0) {
            s.bx = a;
            s.by = b;
        }
    }
    if (const char *e = std::getenv("FMX_XCD_REMAP")) s.remap = std::atoi(e) != 0;
    if (const char *e = std::getenv("FMX_TILE")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%dx%d", &a, &b) == 2 && a > 0 && b > 0) {
            s.ty = a;
            s.tz = b;
        }
        s.tiled = true;  // FMX_TILE=none (or any non-tile value): the plain y-then-z order
    }
    return s;
}

template <class C> CoupledParams<C> make_params(const fmx_coupled_desc &d) {
    CoupledParams<C> p{};
    const LaunchShape ls = launch_shape(d.nx);
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = d.nz;
    p.py = d.periodic_y != 0;
    p.pz = d.periodic_z != 0;
    p.ns_inlet = d.ns_inlet;
    p.ns_outlet = d.ns_outlet;
    p.ad_inlet = d.ad_inlet;
    p.ad_outlet = d.ad_outlet;
    p.isotherm = d.isotherm;
    p.coupling = d.coupling;
    p.block_x = ls.bx;
    p.block_y = ls.by;
    p.gx = (d.nx + ls.bx - 1) / ls.bx;
    p.gy = (d.ny + ls.by - 1) / ls.by;
    p.gz = d.nz;
    p.xcd_remap = ls.remap;
    // tiles only where they divide the block grid exactly
    if (ls.ty > 0 && ls.tz > 0 && p.gy % ls.ty == 0 && p.gz % ls.tz == 0) {
        p.tile_y = ls.ty;
        p.tile_z = ls.tz;
    } else if (!ls.tiled && p.gx * p.gy > 256) {
        // Large z-planes: slabs of block rows in y running through the whole z extent, so the rows a
        // block shares bounce-back lines with (z +- 1) come about 128 blocks later in the launch
        // order instead of a whole y-plane of blocks later, while each z-plane of a slab stays a run
        // of contiguous rows (512^3 FP32: +12.6%; no gain where a z-plane is 128 blocks)
        int ty = 1;
        for (int t = 1; t <= p.gy; ++t)
            if (p.gy % t == 0 && t * p.gx <= 128) ty = t;
        if (ty < p.gy) {
            p.tile_y = ty;
            p.tile_z = p.gz;
        }
    }
    // FMX_PAIR=1 selects the two-cells-per-thread kernel where nx % 128 == 0 (measured no faster for
    // FP32 and slower for FP16/BF16 storage, so off by default); the frozen-solid path (FMX_FREEZE=1)
    // runs only in the one-cell-per-thread kernel
    const char *pe = std::getenv("FMX_PAIR");
    const char *fz = std::getenv("FMX_FREEZE");
    p.freeze = (fz && std::atoi(fz) != 0) ? 1 : 0;
    p.pair = (pe && std::atoi(pe) != 0) && !p.freeze && d.nx >= 128 && d.nx % 128 == 0;
    // the x = nx-2 / nx-1 and x = 0 / 1 cells share a wave when rows start at wave boundaries and
    // nx - 1 is not a multiple of 64
    p.lane_bc = (ls.bx % 64 == 0) && ((d.nx - 1) % 64 != 0);
    p.n = i64(d.nx) * d.ny * d.nz;
    C tau = C(d.tau);
    if (tau < C(0.5001)) tau = C(0.5001);  // BGKCollision.cu:203-204
    p.omega = C(1) / tau;
    p.rho0 = C(d.rho0);
    p.uin_x = C(d.u_in[0]);
    p.uin_y = C(d.u_in[1]);
    p.uin_z = C(d.u_in[2]);
    p.uin2 = p.uin_x * p.uin_x + p.uin_y * p.uin_y + p.uin_z * p.uin_z;
    p.rho_out = C(d.rho_out);
    p.phi_in = C(d.phi_in);
    p.omega_ad = C(1) / C(d.tau_ad);
    p.min_conc = C(d.min_conc);
    p.iso_a = C(d.iso_a);
    p.iso_b = C(d.iso_b);
    p.k_ldf = C(d.k_ldf);
    p.f_ref = C(d.f_ref);
    return p;
}

inline unsigned coupled_blocks(int gx, int gy, int gz) { return unsigned(gx) * unsigned(gy) * unsigned(gz); }

// The fused forward step (coupled.hip): one launch over the whole box.
template <int P>
int launch_coupled_step(int ns_op, const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm,
                        i64 step, i64 acc_row, hipStream_t s);
// The same step with two cells per thread (coupled_pair.hip); FP32 / FP16 / BF16 storage, nx % 128 == 0.
template <int P>
int launch_coupled_pair(int ns_op, const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm,
                        i64 step, i64 acc_row, hipStream_t s);
// Row-compacted NS storage.  The NS planes hold only the cells that are not frozen, row (y, z) after
// row in x order, so a run over a sphere pack moves no population of its frozen solid cells and every
// cache line it touches is full.  One entry per (row, 64-cell segment): the segment's active-cell mask
// and the compact index of its first active cell; the compact index of active cell x is base +
// popcount(mask below x).
struct alignas(16) RowSeg {
    unsigned lo, hi;  // active mask of x = 64 seg .. 64 seg + 63 (bit x % 64)
    unsigned base;    // compact index of the segment's first active cell
    unsigned count;   // active cells in the segment
};
struct RcArgs {
    const RowSeg *rs;  // (y + ny z) * nseg + seg
    i64 nc;            // compact plane stride (>= active cells)
    int nseg;          // segments per row, ceil(nx / 64)
};
// One step of the fused kernel on row-compacted NS planes (ptr.f / ptr.fn compact, everything else
// in the box layout); BGK with frozen solid cells.
template <int P>
int launch_coupled_rc(const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                      i64 acc_row, const RcArgs &rc, hipStream_t s);
// The split step (16-bit storage on row-compacted NS planes): an NS kernel whose waves hold only active
// cells (their box index and link word in `coords`) stores each active cell's velocity in u, then an AD kernel over
// the box reads it.  The 16-bit step is bound by its memory instructions, and the NS ones are then
// issued for the active cells only.
struct SplitArgs {
    const RowSeg *rs;
    const unsigned long long *region;  // sorption-region mask of each (row, segment), bit x % 64
    i64 nc;                  // compact plane stride
    int nseg;
    const uint2 *coords;     // box index and link word of active cell k
    i64 nact;
    void *u;                 // ux, uy, uz of active cell k at 3k, 3k + 1, 3k + 2 (compute precision)
};
template <int P>
int launch_coupled_split(const CoupledPtrs<P> &ptr, const CoupledParams<typename Storage<P>::C> &prm, i64 step,
                         i64 acc_row, const SplitArgs &sa, hipStream_t s);
// nsteps fused steps on the ping-pong buffers of b, outlet sums of step step0+k into row acc_row0+k.
template <int P>
int run_coupled_steps(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, int nsteps, int cur,
                      hipStream_t s, i64 acc_row0 = 0);
// FMX_ERR_ARG unless the inlet series (and, when the run writes it, the outlet accumulator) of b cover
// steps step0 .. step0+nsteps-1.
int check_run_buffers(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, i64 nsteps, const char *who,
                      bool writes_outlet = true);
// Clears the frozen-solid bits of the link words (every cell back on the full update path).
int clear_frozen(const fmx_coupled_desc &d, std::uint32_t *links, hipStream_t s);
// Marks the frozen solid cells of state f[cur] when desc.freeze_solids asks for them (BGK), writing
// their next phase into f[1 - cur]; otherwise clears the frozen bits.
int mark_frozen(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, hipStream_t s);

}  // namespace fmx
