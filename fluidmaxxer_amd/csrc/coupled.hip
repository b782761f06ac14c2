// coupled.hip -- the fused coupled step: D3Q19 NS stream-collide, velocity hand-off,
// LDF sorption source and D3Q7 AD stream-collide in ONE kernel launch per step.
//
// Reference step (main.cpp:319-336, NSLattice.hpp:188-209, ADLattice.hpp:144-162):
//   NS:  PreCollision BCs (VelocityDirichletNEQ on f) -> collide+stream f -> f_new
//        -> PostStreaming BCs (PressureDirichlet on f_new) -> swap
//   coupling: u = moments of the post-BC populations (Macros.cu:20-61) copied to the AD lattice
//   sorption: S = -dq from the previous step's phi (LinearLDF.cu / LangmuirLDF.cu)
//   AD:  PreCollision BCs (tanh / series inlet on g) -> collide+stream g -> g_new, phi
//        -> PostStreaming BCs (ADDirichlet inlet, ADNeumannOutlet) -> swap
//
// How the boundary conditions fold into one launch without changing results:
//   * Pressure outlet (x = nx-1): its populations are feq(rho_out, u) with u the moments of the
//     post-collision cell at x = nx-2 (PressureDirichlet.cu:46-92).  The outlet cell's lane takes
//     that velocity from the neighbouring lane (a lane shuffle; both cells lie in one wave) and
//     stores the outlet populations itself, in the same full-wave store as every other cell.
//   * AD Neumann outlet: g_new(nx-1) = g_new(nx-2) (ADNeumannOutlet.cu:21-30): the outlet lane
//     stores its neighbour's g.  The outlet cell's own AD collision result (and so its velocity)
//     is never observable, which is why a pressure outlet requires the Neumann outlet.
//   * NEQ velocity inlet (x = 0): step s+1 starts by replacing f(0,y,z) with
//     feq(rho_N, u_BC) + f_N - feq(rho_N, u_N), f_N the populations of (1,y,z) after step s
//     (VelocityDirichletNEQ.cu:49-96).  The x = 0 lane takes f_N and its moments from the x = 1
//     lane at the end of step s and stores the result as the next step's f(0,y,z); on the last
//     step of a run it also stores its own post-collision populations into the `inlet_raw` plane,
//     which is what the reference holds in f(0,y,z) between steps (state read-back).
//     fmx_coupled_prologue applies the first step's update.
//   * Equilibrium velocity inlet (VelocityDirichletEq, VelocityDirichlet.cu:54-84): the x = 0
//     lane stores feq(rho0, u_BC) as the next step's f(0,y,z) (and `inlet_raw` as above).
//   * ADDirichlet inlet (post-stream): the x = 0 lane stores phi_in * w.
//   * Pre-collision inlet series (ADDirichletTanhPulse pattern): every pull whose source cell
//     (or keep-local cell) lies on x = 0 uses phi_in(step) * w_q.
//   * The LDF of step s reads phi written by step s-1's AD collision (phi ping-pong).
//   * Outlet breakthrough (OutletCSV.hpp:51-80): the outlet lanes add their phi into
//     FMX_OUTLET_SLOTS FP64 accumulators per step.
// When x = nx-2 and nx-1 fall in different waves ((nx - 1) % 64 == 0) the x = 1 and x = nx-2
// threads write the neighbouring boundary cell's populations instead.
#include <cmath>

#include "coupled_kernel.hpp"
#include "diagnostics.hpp"

namespace fmx {

// Link words: bounce-back bits of the D3Q19 pulls (BGKCollision.cu:90-97, 124-126), region, solid.
__global__ void __launch_bounds__(256) k_build_links(const std::uint8_t *__restrict__ flags,
                                                     std::uint32_t *__restrict__ links, int nx, int ny, int nz,
                                                     int py, int pz, int bgk) {
    const i64 c = gtid();
    const i64 n = i64(nx) * ny * nz;
    if (c >= n) return;
    const int x = int(c % nx);
    const int y = int((c / nx) % ny);
    const int z = int(c / (i64(nx) * ny));
    const std::uint8_t f = flags[c];
    std::uint32_t w = 0;
    if (f & FMX_FLAG_REGION) w |= FMX_LINK_REGION;
    if (f & FMX_FLAG_SOLID) w |= FMX_LINK_SOLID;
    if (bgk) {
        for (int q = 1; q < 19; ++q) {
            if (f & FMX_FLAG_SOLID) {
                w |= 1u << q;
                continue;
            }
            int xs = x - Q19::cx[q], ys = y - Q19::cy[q], zs = z - Q19::cz[q];
            if (xs < 0 || xs >= nx) continue;  // keep local
            if (ys < 0 || ys >= ny) {
                if (!py) continue;
                ys = (ys + ny) % ny;
            }
            if (zs < 0 || zs >= nz) {
                if (!pz) continue;
                zs = (zs + nz) % nz;
            }
            if (flags[cell_index(xs, ys, zs, nx, ny)] & FMX_FLAG_SOLID) w |= 1u << q;
        }
    }
    links[c] = w;
}

// First step's inlet update (NEQ on f(0,y,z) from f(1,y,z)) and the raw x = 0 plane.
template <int P>
__global__ void __launch_bounds__(256) k_coupled_prologue(typename Storage<P>::S *f,
                                                          typename Storage<P>::C *inlet_raw,
                                                          CoupledParams<typename Storage<P>::C> prm) {
    using C = typename Storage<P>::C;
    const i64 np = i64(prm.ny) * prm.nz;
    const i64 k = gtid();
    if (k >= np) return;
    const i64 n = prm.n;
    const i64 c0 = k * prm.nx;  // (0, y, z) with k = y + ny*z
    const i64 c1 = c0 + 1;
    C f1[19];
    C r = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        inlet_raw[q * np + k] = ld_pop<P>(f, i64(q) * n + c0, w19<C>(q), prm.f_ref);
        f1[q] = ld_pop<P>(f, i64(q) * n + c1, w19<C>(q), prm.f_ref);
        r += f1[q];
        jx += f1[q] * C(Q19::cx[q]);
        jy += f1[q] * C(Q19::cy[q]);
        jz += f1[q] * C(Q19::cz[q]);
    }
    if (prm.ns_inlet == 2) {  // VelocityDirichletEq (VelocityDirichlet.cu:54-84)
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const C euB = C(Q19::cx[q]) * prm.uin_x + C(Q19::cy[q]) * prm.uin_y + C(Q19::cz[q]) * prm.uin_z;
            const C cuB = C(3) * euB;
            const C v = w19<C>(q) * prm.rho0 * (C(1) + cuB + C(0.5) * cuB * cuB - C(1.5) * prm.uin2);
            st_pop<P>(f, i64(q) * n + c0, v, C(Q19::inv_w[q]), prm.f_ref);
        }
        return;
    }
    if (prm.ns_inlet != 1) return;
    const C inv = (r > C(0)) ? (C(1) / r) : C(0);
    const C ux = jx * inv, uy = jy * inv, uz = jz * inv;
    const C u2N = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const C euN = C(Q19::cx[q]) * ux + C(Q19::cy[q]) * uy + C(Q19::cz[q]) * uz;
        const C cuN = C(3) * euN;
        const C feqN = w19<C>(q) * r * (C(1) + cuN + C(0.5) * cuN * cuN - C(1.5) * u2N);
        const C euB = C(Q19::cx[q]) * prm.uin_x + C(Q19::cy[q]) * prm.uin_y + C(Q19::cz[q]) * prm.uin_z;
        const C cuB = C(3) * euB;
        const C feqB = w19<C>(q) * r * (C(1) + cuB + C(0.5) * cuB * cuB - C(1.5) * prm.uin2);
        st_pop<P>(f, i64(q) * n + c0, feqB + (f1[q] - feqN), C(Q19::inv_w[q]), prm.f_ref);
    }
}

// Frozen-solid set-up: a BGK solid cell away from the inlet/outlet update planes only reverses its
// own populations (BGKCollision.cu:90-97), so it alternates between its state A and A[opp].  When
// both phases give exactly zero post-collision velocity in the kernel's arithmetic, the cell is
// marked frozen and its next phase is written to the other buffer once.
template <int P>
__global__ void __launch_bounds__(256) k_freeze(const typename Storage<P>::S *__restrict__ fcur,
                                                typename Storage<P>::S *__restrict__ fnext,
                                                std::uint32_t *__restrict__ links,
                                                CoupledParams<typename Storage<P>::C> prm, int allow) {
    using C = typename Storage<P>::C;
    const i64 c = gtid();
    const i64 n = prm.n;
    if (c >= n) return;
    const int x = int(c % prm.nx);
    std::uint32_t lk = links[c] & ~FMX_LINK_FROZEN;
    if (allow && (lk & FMX_LINK_SOLID) && x >= 2 && x <= prm.nx - 3) {
        C a[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) a[q] = ld_pop<P>(fcur, i64(q) * n + c, w19<C>(q), prm.f_ref);
        bool still = true;
        for (int phase = 0; phase < 2; ++phase) {  // phase 0: post-collision = a[opp]; phase 1: = a
            C r = C(0), vx = C(0), vy = C(0), vz = C(0);
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const C v = phase == 0 ? a[Q19::opp[q]] : a[q];
                r += v;
                vx = axpy1<Q19::cx[q]>(vx, v);
                vy = axpy1<Q19::cy[q]>(vy, v);
                vz = axpy1<Q19::cz[q]>(vz, v);
            });
            const C inv = (r > C(0)) ? (C(1) / r) : C(0);
            still = still && vx == C(0) && vy == C(0) && vz == C(0) && inv - inv == C(0);
        }
        if (still) {
            lk |= FMX_LINK_FROZEN;
#pragma unroll
            for (int q = 0; q < 19; ++q) st_pop<P>(fnext, i64(q) * n + c, a[Q19::opp[q]], C(Q19::inv_w[q]), prm.f_ref);
        }
    }
    links[c] = lk;
}

// Reference-view read-back: decode populations; the x = 0 plane comes from inlet_raw.  Half
// storage of g is relative to the lagged concentration (the x - 1 cell's on a Neumann outlet).
template <int P, int Q>
__global__ void __launch_bounds__(256) k_coupled_view(const typename Storage<P>::S *pop,
                                                      const typename Storage<P>::C *inlet_raw,
                                                      const typename Storage<P>::C *phi, typename Storage<P>::C *out,
                                                      int nx, i64 np, i64 n, typename Storage<P>::C ref, int ad_outlet) {
    using C = typename Storage<P>::C;
    const i64 c = gtid();
    if (c >= n) return;
    const i64 xx = c % nx;
    const i64 sref = (ad_outlet && xx == nx - 1) ? c - 1 : c;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        C v;
        if constexpr (Q == 19) {
            v = ld_pop<P>(pop, i64(q) * n + c, w19<C>(q), ref);
            if (inlet_raw && xx == 0) v = inlet_raw[q * np + c / nx];
        } else if constexpr (IsHalf<P>::value) {
            v = decode_g<P>(pop[i64(q) * n + c], w7<C>(q), phi[sref]);
        } else {
            v = pop[i64(q) * n + c];
        }
        out[q * n + c] = v;
    }
}

// Moves the density reference of 16-bit NS storage from ref_old to ref_new: h' = h + (ref_old - ref_new)
// (delta, exact for references within a factor of two) with one rounding to the storage type, so
// w (ref_new + h') is w (ref_old + h) up to that rounding.  Grid-stride over the 19 n stored values.
template <int P>
__global__ void __launch_bounds__(256) k_recenter(typename Storage<P>::S *f, i64 count, float delta) {
    using S = typename Storage<P>::S;
    for (i64 i = gtid(); i < count; i += i64(gridDim.x) * blockDim.x) f[i] = S(widen(f[i]) + delta);
}

// Encode populations into storage: dst[q*n + c] = src ? src[q*n + c] : w_q * value.
template <int P, int Q>
__global__ void __launch_bounds__(256) k_encode_pop(typename Storage<P>::S *dst, const typename Storage<P>::C *src,
                                                    i64 n, typename Storage<P>::C value, typename Storage<P>::C ref,
                                                    const typename Storage<P>::C *phi, int nx, int ad_outlet) {
    using C = typename Storage<P>::C;
    const i64 c = gtid();
    if (c >= n) return;
    const i64 xx = c % nx;
    const i64 sref = (ad_outlet && xx == nx - 1) ? c - 1 : c;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C w = (Q == 19) ? w19<C>(q) : w7<C>(q);
        const int iw = (Q == 19) ? Q19::inv_w[q] : Q7::inv_w[q];
        const C v = src ? src[q * n + c] : w * value;
        if constexpr (Q == 7 && IsHalf<P>::value)
            dst[i64(q) * n + c] = encode_g<P>(v, C(iw), phi[sref], C(1) / g_scale(phi[sref]));
        else
            st_pop<P>(dst, i64(q) * n + c, v, C(iw), ref);
    }
}

template <class C>
__global__ void __launch_bounds__(256) k_copy_or_zero(C *dst, const C *src, i64 n) {
    const i64 c = gtid();
    if (c >= n) return;
    dst[c] = src ? src[c] : C(0);
}

__global__ void __launch_bounds__(256) k_outlet_mean(const double *__restrict__ acc, i64 nsteps, double inv_count,
                                                     double *__restrict__ mean) {
    const i64 k = gtid();
    if (k >= nsteps) return;
    double s = 0.0;
    for (int j = 0; j < FMX_OUTLET_SLOTS; ++j) s += acc[k * FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE + j * FMX_OUTLET_STRIDE];
    mean[k] = s * inv_count;
}

// ------------------------------------------------------------- host side
template <int P> static CoupledPtrs<P> make_ptrs(const fmx_coupled_bufs &b, int cur) {
    using S = typename Storage<P>::S;
    using C = typename Storage<P>::C;
    CoupledPtrs<P> p{};
    p.f = static_cast<const S *>(b.f[cur]);
    p.fn = static_cast<S *>(b.f[1 - cur]);
    p.g = static_cast<const S *>(b.g[cur]);
    p.gn = static_cast<S *>(b.g[1 - cur]);
    p.q = static_cast<const C *>(b.q);
    p.q_new = static_cast<C *>(b.q);
    p.phi = static_cast<const C *>(b.phi[cur]);
    p.phi_new = static_cast<C *>(b.phi[1 - cur]);
    p.links = b.links;
    p.field = static_cast<const C *>(b.field);
    p.tau_field = static_cast<const C *>(b.tau_field);
    p.inlet_raw = static_cast<C *>(b.inlet_raw);
    p.inlet_series = b.inlet_series;
    p.outlet_acc = b.outlet_acc;
    return p;
}

// Row r of the outlet ring holds the outlet-plane concentrations of the step whose accumulator row is
// acc_row0 + r; one workgroup per row adds them in a fixed order (plane index order per thread, then
// a fixed tree) into slot 0 of that accumulator row.
__global__ void __launch_bounds__(256) k_outlet_ring_rows(const double *__restrict__ ring, i64 plane,
                                                          double *__restrict__ acc, i64 acc_row0) {
    __shared__ double red[256];
    const double *row = ring + i64(blockIdx.x) * plane;
    double s = 0.0;
    for (i64 i = threadIdx.x; i < plane; i += blockDim.x) s += row[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[(acc_row0 + i64(blockIdx.x)) * (FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE)] += red[0];
}

template int launch_coupled_step<FMX_F32>(int, const CoupledPtrs<FMX_F32> &, const CoupledParams<float> &, i64, i64,
                                          hipStream_t);
template int launch_coupled_step<FMX_F64>(int, const CoupledPtrs<FMX_F64> &, const CoupledParams<double> &, i64, i64,
                                          hipStream_t);

// FP16 / BF16 storage: instantiated in coupled_half.hip only
extern template int launch_coupled_step<FMX_F16>(int, const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64,
                                                 i64, hipStream_t);
extern template int launch_coupled_step<FMX_BF16>(int, const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64,
                                                  i64, hipStream_t);

template int launch_coupled_rc<FMX_F32>(const CoupledPtrs<FMX_F32> &, const CoupledParams<float> &, i64, i64,
                                        const RcArgs &, hipStream_t);
template int launch_coupled_rc<FMX_F64>(const CoupledPtrs<FMX_F64> &, const CoupledParams<double> &, i64, i64,
                                        const RcArgs &, hipStream_t);
extern template int launch_coupled_rc<FMX_F16>(const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                               const RcArgs &, hipStream_t);
extern template int launch_coupled_rc<FMX_BF16>(const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64, i64,
                                                const RcArgs &, hipStream_t);
template int launch_coupled_split<FMX_F32>(const CoupledPtrs<FMX_F32> &, const CoupledParams<float> &, i64, i64,
                                           const SplitArgs &, hipStream_t);
template int launch_coupled_split<FMX_F64>(const CoupledPtrs<FMX_F64> &, const CoupledParams<double> &, i64, i64,
                                           const SplitArgs &, hipStream_t);
extern template int launch_coupled_split<FMX_F16>(const CoupledPtrs<FMX_F16> &, const CoupledParams<float> &, i64, i64,
                                                  const SplitArgs &, hipStream_t);
extern template int launch_coupled_split<FMX_BF16>(const CoupledPtrs<FMX_BF16> &, const CoupledParams<float> &, i64,
                                                   i64, const SplitArgs &, hipStream_t);

// ------------------------------------------------------------ row-compacted NS storage
// Workspace (fmx_coupled_bufs::compact_work): the RowSeg table, the scan's chunk sums, a header whose
// first word is the number of active cells, the sorption-region masks, then the two compact population
// buffers of 19 planes
// each, whose plane stride nc is what the remaining bytes hold (a multiple of 128 cells); with 16-bit
// storage (the split step) also the box index and link word (8 bytes) and velocity (three values of
// the compute precision) of each active cell.
constexpr int kRcChunk = 4096;  // table entries per block of the scan (256 threads x 16)
// Shorter runs stay in the box layout: building the table, the host synchronisation and the two
// copies cost about two compact steps per run.
constexpr int kRcMinSteps = 16;

struct RcLayout {
    i64 ents, chunks, part_off, hdr_off, region_off, planes_off, nc;
};
// Compact runs take the split step (FMX_SPLIT=0: the fused compact kernel instead); the workspace then
// holds the per-active-cell record and velocity as well
static bool split_step(int) {
    const char *sp = std::getenv("FMX_SPLIT");
    return !(sp && std::atoi(sp) == 0);
}
// the split step's bytes per active cell beyond the NS planes: box index and link word (uint2), then
// the velocity, three values of the compute precision (FP64 for FP64 storage, FP32 otherwise)
static i64 split_cell_bytes(int prec) { return 8 + 3 * (prec == FMX_F64 ? 8 : 4); }
static i64 align256(i64 v) { return (v + 255) & ~i64(255); }
static RcLayout rc_layout(const fmx_coupled_desc &d, i64 bytes) {
    RcLayout l{};
    const i64 nseg = (d.nx + 63) / 64;
    l.ents = i64(d.ny) * d.nz * nseg;
    l.chunks = (l.ents + kRcChunk - 1) / kRcChunk;
    l.part_off = align256(l.ents * i64(sizeof(RowSeg)));
    l.hdr_off = l.part_off + align256(l.chunks * 4);
    l.region_off = l.hdr_off + 256;  // sorption-region mask per (row, segment), for the split step's AD half
    l.planes_off = l.region_off + align256(l.ents * 8);
    const i64 sb = d.prec == FMX_F64 ? 8 : (d.prec == FMX_F32 ? 4 : 2);
    const i64 per_cell = 2 * 19 * sb + (split_step(d.prec) ? split_cell_bytes(d.prec) : 0);
    l.nc = bytes > l.planes_off ? ((bytes - l.planes_off) / per_cell) / 128 * 128 : 0;
    return l;
}

// box index and link word of every active cell at its compact index (the split step's NS threads)
__global__ void __launch_bounds__(256) k_rc_coords(const RowSeg *__restrict__ rs, const std::uint32_t *__restrict__ links,
                                                   uint2 *__restrict__ coords, int nx, int nseg, i64 n) {
    const i64 c = gtid();
    if (c >= n) return;
    const i64 row = c / nx;
    const int x = int(c - row * nx);
    const RowSeg e = rs[row * nseg + (x >> 6)];
    const unsigned l = unsigned(x) & 63u;
    const unsigned long long m = (static_cast<unsigned long long>(e.hi) << 32) | e.lo;
    if (!((m >> l) & 1ull)) return;
    coords[i64(e.base) + __popcll(m & ((1ull << l) - 1ull))] = make_uint2(unsigned(c), links[c]);
}

// one wave per (row, segment): the mask of cells without the frozen bit and its popcount
__global__ void __launch_bounds__(256) k_rc_masks(const std::uint32_t *__restrict__ links, RowSeg *__restrict__ rs,
                                                  unsigned long long *__restrict__ region, int nx, int nseg, i64 ents) {
    const i64 w = i64(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (w >= ents) return;  // wave-uniform
    const unsigned lane = threadIdx.x & 63u;
    const i64 row = w / nseg;
    const int x = int(w - row * nseg) * 64 + int(lane);
    const std::uint32_t lk = x < nx ? links[row * nx + x] : FMX_LINK_FROZEN;
    const unsigned long long m = __ballot(!(lk & FMX_LINK_FROZEN));
    const unsigned long long r = __ballot((lk & FMX_LINK_REGION) != 0u);
    if (lane == 0) {
        rs[w].lo = unsigned(m);
        rs[w].hi = unsigned(m >> 32);
        rs[w].count = unsigned(__popcll(m));
        region[w] = r;
    }
}

// exclusive prefix sum of one value per thread over a block of 256; *total gets the block's sum
__device__ unsigned rc_block_scan(unsigned v, unsigned *total) {
    __shared__ unsigned wsum[4];
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    unsigned inc = v;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const unsigned t = __shfl_up(inc, dd, 64);
        if (lane >= unsigned(dd)) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned off = 0, tot = 0;
#pragma unroll
    for (unsigned i = 0; i < 4; ++i) {
        off += i < w ? wsum[i] : 0u;
        tot += wsum[i];
    }
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(256) k_rc_chunk_sums(const RowSeg *__restrict__ rs, i64 ents,
                                                       unsigned *__restrict__ part) {
    const i64 e0 = i64(blockIdx.x) * kRcChunk + i64(threadIdx.x) * 16;
    unsigned v = 0;
    for (int i = 0; i < 16; ++i)
        if (e0 + i < ents) v += rs[e0 + i].count;
    unsigned tot;
    rc_block_scan(v, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// one block: the chunk sums become their exclusive prefix; hdr[0] = all active cells
__global__ void __launch_bounds__(256) k_rc_scan_chunks(unsigned *__restrict__ part, i64 chunks,
                                                        unsigned long long *__restrict__ hdr) {
    const i64 per = (chunks + 255) / 256;
    const i64 c0 = i64(threadIdx.x) * per;
    unsigned v = 0;
    for (i64 i = 0; i < per; ++i)
        if (c0 + i < chunks) v += part[c0 + i];
    unsigned tot;
    unsigned run = rc_block_scan(v, &tot);
    for (i64 i = 0; i < per; ++i)
        if (c0 + i < chunks) {
            const unsigned t = part[c0 + i];
            part[c0 + i] = run;
            run += t;
        }
    if (threadIdx.x == 0) hdr[0] = tot;
}

__global__ void __launch_bounds__(256) k_rc_bases(RowSeg *__restrict__ rs, i64 ents, const unsigned *__restrict__ part) {
    const i64 e0 = i64(blockIdx.x) * kRcChunk + i64(threadIdx.x) * 16;
    unsigned v = 0;
    for (int i = 0; i < 16; ++i)
        if (e0 + i < ents) v += rs[e0 + i].count;
    unsigned tot;
    unsigned run = part[blockIdx.x] + rc_block_scan(v, &tot);
    for (int i = 0; i < 16; ++i)
        if (e0 + i < ents) {
            rs[e0 + i].base = run;
            run += rs[e0 + i].count;
        }
}

// Copies the 19 stored populations of every active cell between the box layout (plane stride n) and
// the compact one (stride nc), without conversion.
template <class S>
__global__ void __launch_bounds__(256) k_rc_move(const S *__restrict__ src, S *__restrict__ dst,
                                                 const RowSeg *__restrict__ rs, int nx, int nseg, i64 n, i64 nc,
                                                 int to_compact) {
    const i64 c = gtid();
    if (c >= n) return;
    const i64 row = c / nx;
    const int x = int(c - row * nx);
    const RowSeg e = rs[row * nseg + (x >> 6)];
    const unsigned l = unsigned(x) & 63u;
    const unsigned long long m = (static_cast<unsigned long long>(e.hi) << 32) | e.lo;
    if (!((m >> l) & 1ull)) return;
    const i64 k = i64(e.base) + __popcll(m & ((1ull << l) - 1ull));
    const i64 si = to_compact ? c : k, sn = to_compact ? n : nc;
    const i64 di = to_compact ? k : c, dn = to_compact ? nc : n;
#pragma unroll
    for (int q = 0; q < 19; ++q) dst[q * dn + di] = src[q * sn + si];
}

// Builds the table from the current link words, checks that the workspace holds the active cells
// (one host synchronisation) and copies f[cur] into compact buffer cur; *fits is false (and nothing
// copied) when they do not fit, e.g. when fewer cells are frozen than when the workspace was sized.
template <int P>
static int rc_begin(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, hipStream_t s, RcArgs &rc,
                    void *cbuf[2], bool *fits, SplitArgs *sa) {
    *fits = false;
    using S = typename Storage<P>::S;
    const RcLayout l = rc_layout(d, b.compact_work_bytes);
    if (l.nc <= 0) return set_error(FMX_ERR_ARG, "fmx_coupled_run: compact_work_bytes %lld is too small",
                                    (long long)b.compact_work_bytes);
    char *w = static_cast<char *>(b.compact_work);
    RowSeg *rs = reinterpret_cast<RowSeg *>(w);
    unsigned *part = reinterpret_cast<unsigned *>(w + l.part_off);
    unsigned long long *hdr = reinterpret_cast<unsigned long long *>(w + l.hdr_off);
    const int nseg = (d.nx + 63) / 64;
    const i64 n = i64(d.nx) * d.ny * d.nz;
    unsigned long long *region = reinterpret_cast<unsigned long long *>(w + l.region_off);
    hipLaunchKernelGGL(k_rc_masks, dim3(unsigned((l.ents + 3) / 4)), dim3(256), 0, s, b.links, rs, region, d.nx, nseg,
                       l.ents);
    if (int rc2 = after_launch("k_rc_masks")) return rc2;
    hipLaunchKernelGGL(k_rc_chunk_sums, dim3(unsigned(l.chunks)), dim3(256), 0, s, rs, l.ents, part);
    if (int rc2 = after_launch("k_rc_chunk_sums")) return rc2;
    hipLaunchKernelGGL(k_rc_scan_chunks, dim3(1), dim3(256), 0, s, part, l.chunks, hdr);
    if (int rc2 = after_launch("k_rc_scan_chunks")) return rc2;
    hipLaunchKernelGGL(k_rc_bases, dim3(unsigned(l.chunks)), dim3(256), 0, s, rs, l.ents, part);
    if (int rc2 = after_launch("k_rc_bases")) return rc2;
    unsigned long long nact = 0;
    if (int rc2 = check_hip(hipMemcpyAsync(&nact, hdr, sizeof(nact), hipMemcpyDeviceToHost, s),
                            "fmx_coupled_run: hipMemcpyAsync"))
        return rc2;
    if (int rc2 = check_hip(hipStreamSynchronize(s), "fmx_coupled_run: hipStreamSynchronize")) return rc2;
    if (i64(nact) > l.nc) return FMX_OK;
    *fits = true;
    cbuf[0] = w + l.planes_off;
    cbuf[1] = w + l.planes_off + l.nc * 19 * i64(sizeof(S));
    rc.rs = rs;
    rc.nc = l.nc;
    rc.nseg = nseg;
    if (sa) {  // the split step: box index and velocity of every active cell after the two buffers
        uint2 *coords = reinterpret_cast<uint2 *>(w + l.planes_off + 2 * l.nc * 19 * i64(sizeof(S)));
        sa->rs = rs;
        sa->region = region;
        sa->nc = l.nc;
        sa->nseg = nseg;
        sa->coords = coords;
        sa->nact = i64(nact);
        sa->u = coords + l.nc;
        hipLaunchKernelGGL(k_rc_coords, dim3(blocks_for(n)), dim3(256), 0, s, rs, b.links, coords, d.nx, nseg, n);
        if (int rc2 = after_launch("k_rc_coords")) return rc2;
    }
    hipLaunchKernelGGL(k_rc_move<S>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const S *>(b.f[cur]),
                       static_cast<S *>(cbuf[cur]), rs, d.nx, nseg, n, l.nc, 1);
    return after_launch("k_rc_move");
}

// Steps step0 .. step0+nsteps-1 on the ping-pong buffers; the outlet sums of step step0+k go to row
// acc_row0 + k of outlet_acc.  With a compact workspace (BGK, no mass log, at least kRcMinSteps
// steps) the NS populations of the cells that are not frozen are stepped in row-compacted buffers and
// copied back at the end; frozen cells keep their two alternating states in f[0] / f[1] as always.
template <int P>
int run_coupled_steps(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, int nsteps, int cur,
                      hipStream_t s, i64 acc_row0) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const CoupledParams<C> prm = make_params<C>(d);
    const bool ring = b.outlet_acc && b.outlet_ring;
    const i64 plane = i64(d.ny) * d.nz;
    bool compact = b.compact_work && d.ns_op == FMX_NS_BGK && !b.mass_log && nsteps >= kRcMinSteps;
    const bool split = split_step(d.prec);
    RcArgs rca{};
    SplitArgs sa{};
    void *cbuf[2] = {nullptr, nullptr};
    if (compact)
        if (int rc = rc_begin<P>(d, b, cur, s, rca, cbuf, &compact, split ? &sa : nullptr)) return rc;
    int rows = 0;  // ring rows filled since the last fold
    for (int k = 0; k < nsteps; ++k) {
        CoupledPtrs<P> ptr = make_ptrs<P>(b, (cur + k) & 1);
        if (k != nsteps - 1) ptr.inlet_raw = nullptr;  // the raw x = 0 plane is state read-back only
        if (ring) ptr.outlet_acc = b.outlet_ring + i64(rows) * plane;  // acc_row -1: store into this ring row
        if (compact) {
            ptr.f = static_cast<const S *>(cbuf[(cur + k) & 1]);
            ptr.fn = static_cast<S *>(cbuf[(cur + k + 1) & 1]);
            const i64 row = ring ? -1 : acc_row0 + k;
            if (int rc = split ? launch_coupled_split<P>(ptr, prm, step0 + k, row, sa, s)
                               : launch_coupled_rc<P>(ptr, prm, step0 + k, row, rca, s))
                return rc;
        } else if (int rc = launch_coupled_step<P>(d.ns_op, ptr, prm, step0 + k, ring ? -1 : acc_row0 + k, s)) {
            return rc;
        }
        if (ring && (++rows == FMX_OUTLET_RING_ROWS || k == nsteps - 1)) {
            hipLaunchKernelGGL(k_outlet_ring_rows, dim3(rows), dim3(256), 0, s, b.outlet_ring, plane, b.outlet_acc,
                               acc_row0 + k + 1 - rows);
            if (int rc = after_launch("k_outlet_ring_rows")) return rc;
            rows = 0;
        }
        const i64 done = acc_row0 + k + 1;  // steps of the caller's run completed
        if (b.mass_log && b.mass_every > 0 && done % b.mass_every == 0) {
            const int nxt = (cur + k + 1) & 1;  // the buffers this step wrote
            if (int rc = launch_mass_any(d.prec, b.phi[nxt], b.q, b.f[nxt], prm.n, d.f_ref,
                                         static_cast<double *>(b.mass_workspace),
                                         b.mass_log + (done / b.mass_every - 1) * kMassCols, s))
                return rc;
        }
    }
    if (compact) {
        const int last = (cur + nsteps) & 1;
        const i64 n = prm.n;
        hipLaunchKernelGGL(k_rc_move<S>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const S *>(cbuf[last]),
                           static_cast<S *>(b.f[last]), rca.rs, d.nx, rca.nseg, n, rca.nc, 0);
        if (int rc = after_launch("k_rc_move")) return rc;
    }
    return FMX_OK;
}

template int run_coupled_steps<FMX_F32>(const fmx_coupled_desc &, const fmx_coupled_bufs &, i64, int, int, hipStream_t,
                                        i64);
template int run_coupled_steps<FMX_F64>(const fmx_coupled_desc &, const fmx_coupled_bufs &, i64, int, int, hipStream_t,
                                        i64);
template int run_coupled_steps<FMX_F16>(const fmx_coupled_desc &, const fmx_coupled_bufs &, i64, int, int, hipStream_t,
                                        i64);
template int run_coupled_steps<FMX_BF16>(const fmx_coupled_desc &, const fmx_coupled_bufs &, i64, int, int, hipStream_t,
                                         i64);

template <int P>
static int freeze_pass(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, hipStream_t s, int allow) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const CoupledParams<C> prm = make_params<C>(d);
    hipLaunchKernelGGL(k_freeze<P>, dim3(blocks_for(prm.n)), dim3(256), 0, s, static_cast<const S *>(b.f[cur]),
                       static_cast<S *>(b.f[1 - cur]), b.links, prm, allow);
    return after_launch("k_freeze");
}

template <int P>
static int prologue(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const CoupledParams<C> prm = make_params<C>(d);
    if (b.inlet_raw) {
        const i64 np = i64(d.ny) * d.nz;
        hipLaunchKernelGGL(k_coupled_prologue<P>, dim3(blocks_for(np)), dim3(256), 0, s, static_cast<S *>(b.f[cur]),
                           static_cast<C *>(b.inlet_raw), prm);
        if (int rc = after_launch("k_coupled_prologue")) return rc;
    }
    if (d.ns_op != FMX_NS_BGK) return FMX_OK;
    // Opt-in (desc.freeze_solids, or FMX_FREEZE=1): frozen cells pay off in runs with a compact
    // workspace, whose NS planes leave them out.  In the box layout a frozen cell only saves traffic
    // when whole cache lines are frozen, which dense sphere packs rarely give (measured slower on the
    // 256x128x128 pack).
    const char *fz = std::getenv("FMX_FREEZE");
    return freeze_pass<P>(d, b, cur, s, (fz != nullptr && std::atoi(fz) != 0) || d.freeze_solids != 0);
}

int mark_frozen(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, hipStream_t s) {
    if (d.ns_op != FMX_NS_BGK || !d.freeze_solids) return clear_frozen(d, b.links, s);
    switch (d.prec) {
        case FMX_F32: return freeze_pass<FMX_F32>(d, b, cur, s, 1);
        case FMX_F64: return freeze_pass<FMX_F64>(d, b, cur, s, 1);
        case FMX_F16: return freeze_pass<FMX_F16>(d, b, cur, s, 1);
        default: return freeze_pass<FMX_BF16>(d, b, cur, s, 1);
    }
}

int clear_frozen(const fmx_coupled_desc &d, std::uint32_t *links, hipStream_t s) {
    const CoupledParams<float> prm = make_params<float>(d);
    hipLaunchKernelGGL(k_freeze<FMX_F32>, dim3(blocks_for(prm.n)), dim3(256), 0, s, nullptr, nullptr, links, prm, 0);
    return after_launch("k_freeze");
}

template <int P>
static int load_state(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, int cur, const void *f_c, const void *g_c,
                      const void *q_c, const void *phi_c, bool init, double phi0, hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const i64 n = i64(d.nx) * d.ny * d.nz;
    const unsigned nb = blocks_for(n);
    C *phi = static_cast<C *>(b.phi[cur]);
    if (init || phi_c) {  // first: g is stored relative to phi in half precision
        hipLaunchKernelGGL(k_copy_or_zero<C>, dim3(nb), dim3(256), 0, s, phi, static_cast<const C *>(phi_c), n);
        if (int rc = after_launch("k_copy_or_zero")) return rc;
    }
    if (init || f_c) {
        hipLaunchKernelGGL((k_encode_pop<P, 19>), dim3(nb), dim3(256), 0, s, static_cast<S *>(b.f[cur]),
                           static_cast<const C *>(f_c), n, C(d.rho0), C(d.f_ref), phi, d.nx, d.ad_outlet);
        if (int rc = after_launch("k_encode_pop")) return rc;
    }
    if (init || g_c) {
        hipLaunchKernelGGL((k_encode_pop<P, 7>), dim3(nb), dim3(256), 0, s, static_cast<S *>(b.g[cur]),
                           static_cast<const C *>(g_c), n, C(phi0), C(0), phi, d.nx, d.ad_outlet);
        if (int rc = after_launch("k_encode_pop")) return rc;
    }
    if (init || q_c) {
        hipLaunchKernelGGL(k_copy_or_zero<C>, dim3(nb), dim3(256), 0, s, static_cast<C *>(b.q), static_cast<const C *>(q_c), n);
        if (int rc = after_launch("k_copy_or_zero")) return rc;
    }
    return prologue<P>(d, b, cur, s);
}

template <int P, int Q>
static int view(const fmx_coupled_desc &d, const void *pop, const void *inlet_raw, const void *phi, void *out,
                double ref, hipStream_t s) {
    using C = typename Storage<P>::C;
    using S = typename Storage<P>::S;
    const i64 n = i64(d.nx) * d.ny * d.nz;
    hipLaunchKernelGGL((k_coupled_view<P, Q>), dim3(blocks_for(n)), dim3(256), 0, s, static_cast<const S *>(pop),
                       static_cast<const C *>(inlet_raw), static_cast<const C *>(phi), static_cast<C *>(out), d.nx,
                       i64(d.ny) * d.nz, n, C(ref), d.ad_outlet);
    return after_launch("k_coupled_view");
}


// The lengths a run over steps step0 .. step0+nsteps-1 reads and writes: the inlet series (entries step0 ..
// step0+nsteps-1) and the outlet accumulator (rows 0 .. nsteps-1).
int check_run_buffers(const fmx_coupled_desc &d, const fmx_coupled_bufs &b, i64 step0, i64 nsteps, const char *who,
                      bool writes_outlet) {
    if (d.ad_inlet == 2 && b.inlet_series_len < step0 + nsteps)
        return set_error(FMX_ERR_ARG, "%s: steps %lld .. %lld read inlet_series[%lld], but inlet_series_len is %lld", who,
                         (long long)step0, (long long)(step0 + nsteps - 1), (long long)(step0 + nsteps - 1),
                         (long long)b.inlet_series_len);
    if (writes_outlet && b.outlet_acc && b.outlet_acc_rows < nsteps)
        return set_error(FMX_ERR_ARG, "%s: %lld steps write %lld outlet_acc rows, but outlet_acc_rows is %lld", who,
                         (long long)nsteps, (long long)nsteps, (long long)b.outlet_acc_rows);
    if (writes_outlet && b.mass_log) {
        if (b.mass_every < 1) return set_error(FMX_ERR_ARG, "%s: mass_log needs mass_every >= 1", who);
        const i64 n = i64(d.nx) * d.ny * d.nz;
        const i64 need = i64(mass_blocks(n)) * kMassCols * i64(sizeof(double));
        if (!b.mass_workspace || b.mass_workspace_bytes < need)
            return set_error(FMX_ERR_ARG, "%s: mass_log needs a mass_workspace of %lld bytes (%lld given)", who,
                             (long long)need, (long long)(b.mass_workspace ? b.mass_workspace_bytes : 0));
        if (b.mass_log_rows < nsteps / b.mass_every)
            return set_error(FMX_ERR_ARG, "%s: %lld steps log %lld mass rows, but mass_log_rows is %lld", who,
                             (long long)nsteps, (long long)(nsteps / b.mass_every), (long long)b.mass_log_rows);
    }
    return FMX_OK;
}

}  // namespace fmx

using namespace fmx;

extern "C" {

int fmx_coupled_validate(const fmx_coupled_desc *d, const fmx_coupled_bufs *b) {
    if (!d || !b) return set_error(FMX_ERR_ARG, "fmx_coupled: null descriptor");
    if (d->nx < 3 || d->ny < 1 || d->nz < 1 || d->nz >= 65536)
        return set_error(FMX_ERR_ARG, "fmx_coupled: need nx >= 3, ny >= 1, 1 <= nz < 65536 (got %d x %d x %d)", d->nx,
                         d->ny, d->nz);
    const i64 n = i64(d->nx) * d->ny * d->nz;
    const i64 sb = d->prec == FMX_F64 ? 8 : 4;  // widest per-cell element the kernel addresses
    if (n * sb >= (i64(1) << 32))
        return set_error(FMX_ERR_ARG, "fmx_coupled: %lld cells exceed the 4 GiB population-plane limit", (long long)n);
    if (d->prec < FMX_F32 || d->prec > FMX_BF16) return set_error(FMX_ERR_ARG, "fmx_coupled: bad precision %d", d->prec);
    if (d->ns_op < FMX_NS_BGK || d->ns_op > FMX_NS_PBB) return set_error(FMX_ERR_ARG, "fmx_coupled: bad ns_op");
    if (d->ns_inlet < 0 || d->ns_inlet > 2 || d->ns_outlet < 0 || d->ns_outlet > 1 || d->ad_inlet < 0 ||
        d->ad_inlet > 2 || d->ad_outlet < 0 || d->ad_outlet > 1 || d->isotherm < 0 || d->isotherm > 2)
        return set_error(FMX_ERR_ARG, "fmx_coupled: bad boundary/isotherm selector");
    if (d->ns_outlet && !d->ad_outlet)
        return set_error(FMX_ERR_UNSUPPORTED,
                         "fmx_coupled: a pressure outlet needs the AD Neumann outlet in the fused step (use the "
                         "operator API for other combinations)");
    if (!(d->tau_ad > 0.0) && !b->tau_field) return set_error(FMX_ERR_ARG, "fmx_coupled: tau_ad must be > 0");
    if (!b->f[0] || !b->f[1] || !b->g[0] || !b->g[1] || !b->q || !b->phi[0] || !b->phi[1] || !b->links)
        return set_error(FMX_ERR_ARG, "fmx_coupled: missing state buffer");
    if (d->ns_inlet && !b->inlet_raw) return set_error(FMX_ERR_ARG, "fmx_coupled: ns_inlet needs inlet_raw");
    if (d->freeze_solids < 0 || d->freeze_solids > 1) return set_error(FMX_ERR_ARG, "fmx_coupled: freeze_solids is 0 or 1");
    if (b->compact_work && b->compact_work_bytes < fmx_coupled_compact_bytes(d, 1))
        return set_error(FMX_ERR_ARG, "fmx_coupled: compact_work_bytes %lld is below the row table's %lld",
                         (long long)b->compact_work_bytes, (long long)fmx_coupled_compact_bytes(d, 1));
    if (d->ad_inlet == 2 && !b->inlet_series) return set_error(FMX_ERR_ARG, "fmx_coupled: ad_inlet 2 needs inlet_series");
    if (d->ns_op == FMX_NS_POROUS && !b->field)
        return FMX_OK;  // porosity defaults to 1 (PorousBGKCollision.cu:130)
    return FMX_OK;
}

int fmx_coupled_build_links(const fmx_coupled_desc *d, const std::uint8_t *flags, std::uint32_t *links,
                            void *stream) {
    FMX_TRY_BEGIN
    if (!d || !flags || !links || d->nx < 1 || d->ny < 1 || d->nz < 1)
        return set_error(FMX_ERR_ARG, "fmx_coupled_build_links: bad arguments");
    const i64 n = i64(d->nx) * d->ny * d->nz;
    hipLaunchKernelGGL(k_build_links, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), flags, links, d->nx, d->ny,
                       d->nz, d->periodic_y != 0, d->periodic_z != 0, d->ns_op == FMX_NS_BGK);
    return after_launch("k_build_links");
    FMX_TRY_END
}

int fmx_coupled_prologue(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return set_error(FMX_ERR_ARG, "fmx_coupled_prologue: cur must be 0 or 1");
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return prologue<FMX_F32>(*d, *b, cur, s);
        case FMX_F64: return prologue<FMX_F64>(*d, *b, cur, s);
        case FMX_F16: return prologue<FMX_F16>(*d, *b, cur, s);
        default: return prologue<FMX_BF16>(*d, *b, cur, s);
    }
    FMX_TRY_END
}

int fmx_coupled_init(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, double phi0, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return set_error(FMX_ERR_ARG, "fmx_coupled_init: cur must be 0 or 1");
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return load_state<FMX_F32>(*d, *b, cur, nullptr, nullptr, nullptr, nullptr, true, phi0, s);
        case FMX_F64: return load_state<FMX_F64>(*d, *b, cur, nullptr, nullptr, nullptr, nullptr, true, phi0, s);
        case FMX_F16: return load_state<FMX_F16>(*d, *b, cur, nullptr, nullptr, nullptr, nullptr, true, phi0, s);
        default: return load_state<FMX_BF16>(*d, *b, cur, nullptr, nullptr, nullptr, nullptr, true, phi0, s);
    }
    FMX_TRY_END
}

int fmx_coupled_load(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, const void *f_c, const void *g_c,
                     const void *q_c, const void *phi_c, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return set_error(FMX_ERR_ARG, "fmx_coupled_load: cur must be 0 or 1");
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return load_state<FMX_F32>(*d, *b, cur, f_c, g_c, q_c, phi_c, false, 0.0, s);
        case FMX_F64: return load_state<FMX_F64>(*d, *b, cur, f_c, g_c, q_c, phi_c, false, 0.0, s);
        case FMX_F16: return load_state<FMX_F16>(*d, *b, cur, f_c, g_c, q_c, phi_c, false, 0.0, s);
        default: return load_state<FMX_BF16>(*d, *b, cur, f_c, g_c, q_c, phi_c, false, 0.0, s);
    }
    FMX_TRY_END
}

int fmx_coupled_run(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int64_t step0, int32_t nsteps, int cur,
                    void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if (cur != 0 && cur != 1) return set_error(FMX_ERR_ARG, "fmx_coupled_run: cur must be 0 or 1");
    if (nsteps < 0 || step0 < 0) return set_error(FMX_ERR_ARG, "fmx_coupled_run: bad step range");
    if (int rc2 = check_run_buffers(*d, *b, step0, nsteps, "fmx_coupled_run")) return rc2;
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return run_coupled_steps<FMX_F32>(*d, *b, step0, nsteps, cur, s);
        case FMX_F64: return run_coupled_steps<FMX_F64>(*d, *b, step0, nsteps, cur, s);
        case FMX_F16: return run_coupled_steps<FMX_F16>(*d, *b, step0, nsteps, cur, s);
        default: return run_coupled_steps<FMX_BF16>(*d, *b, step0, nsteps, cur, s);
    }
    FMX_TRY_END
}

int fmx_coupled_view_f(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, void *f_out, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if ((cur != 0 && cur != 1) || !f_out) return set_error(FMX_ERR_ARG, "fmx_coupled_view_f: bad arguments");
    const void *raw = d->ns_inlet ? b->inlet_raw : nullptr;
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return view<FMX_F32, 19>(*d, b->f[cur], raw, nullptr, f_out, d->f_ref, s);
        case FMX_F64: return view<FMX_F64, 19>(*d, b->f[cur], raw, nullptr, f_out, d->f_ref, s);
        case FMX_F16: return view<FMX_F16, 19>(*d, b->f[cur], raw, nullptr, f_out, d->f_ref, s);
        default: return view<FMX_BF16, 19>(*d, b->f[cur], raw, nullptr, f_out, d->f_ref, s);
    }
    FMX_TRY_END
}

int fmx_coupled_view_g(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, void *g_out, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if ((cur != 0 && cur != 1) || !g_out) return set_error(FMX_ERR_ARG, "fmx_coupled_view_g: bad arguments");
    hipStream_t s = as_stream(stream);
    switch (d->prec) {
        case FMX_F32: return view<FMX_F32, 7>(*d, b->g[cur], nullptr, b->phi[cur], g_out, 0.0, s);
        case FMX_F64: return view<FMX_F64, 7>(*d, b->g[cur], nullptr, b->phi[cur], g_out, 0.0, s);
        case FMX_F16: return view<FMX_F16, 7>(*d, b->g[cur], nullptr, b->phi[cur], g_out, 0.0, s);
        default: return view<FMX_BF16, 7>(*d, b->g[cur], nullptr, b->phi[cur], g_out, 0.0, s);
    }
    FMX_TRY_END
}

int64_t fmx_coupled_recenter_workspace_bytes(const fmx_coupled_desc *d) {
    if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return -1;
    const i64 n = i64(d->nx) * d->ny * d->nz;
    return int64_t(mass_blocks(n) + 1) * kMassCols * int64_t(sizeof(double));
}

int64_t fmx_coupled_compact_bytes(const fmx_coupled_desc *d, int64_t active) {
    if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || active < 0 || active > i64(d->nx) * d->ny * d->nz) return -1;
    if (d->prec < FMX_F32 || d->prec > FMX_BF16) return -1;
    const RcLayout l = rc_layout(*d, 0);
    const i64 sb = d->prec == FMX_F64 ? 8 : (d->prec == FMX_F32 ? 4 : 2);
    const i64 per_cell = 2 * 19 * sb + (split_step(d->prec) ? split_cell_bytes(d->prec) : 0);
    return l.planes_off + per_cell * ((active + 127) / 128 * 128);
}

int fmx_coupled_recenter(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int cur, double target,
                         double *f_ref_out, void *workspace, int64_t workspace_bytes, void *stream) {
    FMX_TRY_BEGIN
    int rc = fmx_coupled_validate(d, b);
    if (rc) return rc;
    if ((cur != 0 && cur != 1) || !f_ref_out) return set_error(FMX_ERR_ARG, "fmx_coupled_recenter: bad arguments");
    hipStream_t s = as_stream(stream);
    const i64 n = i64(d->nx) * d->ny * d->nz;
    double ref = target;
    if (std::isnan(target)) {  // the mean density of the cells without the solid link bit
        const int64_t need = fmx_coupled_recenter_workspace_bytes(d);
        if (!workspace || workspace_bytes < need)
            return set_error(FMX_ERR_ARG, "fmx_coupled_recenter: workspace of %lld bytes given, %lld needed "
                                          "(fmx_coupled_recenter_workspace_bytes)",
                             (long long)(workspace ? workspace_bytes : 0), (long long)need);
        double *partial = static_cast<double *>(workspace);
        double *out = partial + i64(mass_blocks(n)) * kMassCols;
        if (int rc2 = launch_fluid_rho_any(d->prec, b->f[cur], b->links, n, d->f_ref, partial, out, s)) return rc2;
        double host[2] = {0.0, 0.0};
        if (int rc2 = check_hip(hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, s),
                                "fmx_coupled_recenter: hipMemcpyAsync"))
            return rc2;
        if (int rc2 = check_hip(hipStreamSynchronize(s), "fmx_coupled_recenter: hipStreamSynchronize")) return rc2;
        ref = host[1] > 0.0 ? host[0] / host[1] : d->f_ref;
    }
    if (!std::isfinite(ref)) return set_error(FMX_ERR_ARG, "fmx_coupled_recenter: non-finite reference %g", ref);
    if (d->prec == FMX_F16 || d->prec == FMX_BF16) {
        if (!(ref > 0.5 * d->f_ref) || !(ref < 2.0 * d->f_ref))
            return set_error(FMX_ERR_ARG, "fmx_coupled_recenter: reference %g is not within a factor of two of f_ref %g",
                             ref, d->f_ref);
        // the kernels decode with the reference rounded to FP32; so does the shift
        const float delta = float(d->f_ref) - float(ref);
        const i64 count = 19 * n;
        const i64 want = (count + 255) / 256;
        const unsigned grid = unsigned(want < 65536 ? want : 65536);
        // both buffers: the other one holds the second phase of any frozen solid cell (FMX_FREEZE), which
        // no step rewrites; the rest of it is overwritten by the next step anyway
        for (int k = 0; k < 2; ++k) {
            if (d->prec == FMX_F16)
                hipLaunchKernelGGL(k_recenter<FMX_F16>, dim3(grid), dim3(256), 0, s, static_cast<_Float16 *>(b->f[k]),
                                   count, delta);
            else
                hipLaunchKernelGGL(k_recenter<FMX_BF16>, dim3(grid), dim3(256), 0, s, static_cast<__bf16 *>(b->f[k]),
                                   count, delta);
            if (int rc2 = after_launch("k_recenter")) return rc2;
        }
    }
    *f_ref_out = ref;
    return FMX_OK;
    FMX_TRY_END
}

int fmx_coupled_step(const fmx_coupled_desc *d, const fmx_coupled_bufs *b, int64_t step, int cur, void *stream) {
    return fmx_coupled_run(d, b, step, 1, cur, stream);
}

int fmx_outlet_mean(const double *acc, int64_t nsteps, int64_t count, double *mean, void *stream) {
    FMX_TRY_BEGIN
    if (nsteps == 0) return FMX_OK;
    if (!acc || !mean || nsteps < 0 || count <= 0) return set_error(FMX_ERR_ARG, "fmx_outlet_mean: bad arguments");
    hipLaunchKernelGGL(k_outlet_mean, dim3(blocks_for(nsteps)), dim3(256), 0, as_stream(stream), acc, nsteps,
                       1.0 / double(count), mean);
    return after_launch("k_outlet_mean");
    FMX_TRY_END
}

}  // extern "C"
