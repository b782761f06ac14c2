/* fmx_abi.h -- C ABI of libfmx_hip.so, the gfx950 implementation of
 * FluidMaxxer's coupled D3Q19 Navier-Stokes + D3Q7 advection-diffusion step.
 *
 * Conventions
 *   - Every entry point returns an int status (FMX_OK = 0) and never throws;
 *     fmx_last_error() returns the thread's last message.
 *   - Buffers are caller-owned device pointers (fmx_alloc or any HIP
 *     allocation).  Scalar fields (phi, q, S, tau, porosity, rho, u) are float
 *     for FMX_F32/FMX_F16/FMX_BF16 and double for FMX_F64.
 *   - Populations are stored structure-of-arrays: population q of cell c at
 *     pop[q*N + c], N = nx*ny*nz, c = x + nx*(y + ny*z).  The reference stores
 *     array-of-structures f[c*Q + q]; fmx_pack_aos / fmx_unpack_aos convert.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *     kernels are asynchronous on it.  fmx_set_sync_check(1) restores the
 *     reference's synchronise-and-check after every launch
 *     (CudaHelpers.cu:16-20).
 *   - Reference citations are paths relative to the reference repository.
 */
#ifndef FMX_ABI_H
#define FMX_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: buffer lengths in fmx_coupled_bufs (inlet_series_len, outlet_acc_rows) and fmx_adjoint_bufs
 * (outlet_weight_len), checked before any launch; fmx_reduce_mass with a caller workspace and a
 * non-finite count, and its periodic log inside fmx_coupled_run (mass_log).
 * 3: FP32 checkpoints of an FP64 run, whose reverse sweep runs in FP32 (fmx_adjoint_bufs field_rev,
 * tau_field_rev)
 * 4: fmx_coupled_recenter (the density reference of 16-bit NS storage moved to the flow's mean)
 * 5: fmx_coupled_desc::freeze_solids, fmx_coupled_bufs::compact_work, fmx_coupled_compact_bytes
 *    (row-compacted NS storage that leaves frozen solid cells out of a run) */
#define FMX_ABI_VERSION 5

/* storage precision (DATA_TYPE, Types.hpp:6, plus half-precision deviation storage) */
enum { FMX_F32 = 0, FMX_F64 = 1, FMX_F16 = 2, FMX_BF16 = 3 };
/* NS collision operators */
enum { FMX_NS_BGK = 0, FMX_NS_POROUS = 1, FMX_NS_PBB = 2 };
/* isotherms */
enum { FMX_ISO_NONE = 0, FMX_ISO_LINEAR = 1, FMX_ISO_LANGMUIR = 2 };
/* status codes */
enum { FMX_OK = 0, FMX_ERR_ARG = 1, FMX_ERR_HIP = 2, FMX_ERR_UNSUPPORTED = 3 };
/* flag byte bits of the coupled step */
enum { FMX_FLAG_SOLID = 1, FMX_FLAG_REGION = 2 };
/* per-cell link word of the coupled step (fmx_coupled_build_links): bit q (1..18) set when the
 * D3Q19 pull along direction q bounces back (the cell is solid, or the in-domain source cell is
 * solid; FMX_NS_BGK only), plus the region and solid bits */
#define FMX_LINK_REGION (1u << 19)
#define FMX_LINK_SOLID (1u << 20)
/* set by fmx_coupled_prologue on FMX_NS_BGK solid cells (2 <= x <= nx-3) whose populations and their
 * reversal both carry exactly zero momentum: such a cell alternates between two fixed states, both are
 * kept in the two population buffers, and the step skips its NS loads and stores */
#define FMX_LINK_FROZEN (1u << 21)

typedef struct fmx_grid {
    int32_t nx, ny, nz;
    int32_t periodic_x, periodic_y, periodic_z; /* LBMParams::periodic{X,Y,Z}, Params.hpp:21-23 */
} fmx_grid;

/* ------------------------------------------------------------- runtime */
int fmx_abi_version(void);
const char *fmx_last_error(void);
int fmx_device_count(int *count);
/* name (NUL-terminated, truncated to name_len), total HBM bytes, CU count, gcnArchName */
int fmx_device_info(int device, char *name, int name_len, int64_t *total_mem, int *cu_count, char *arch,
                    int arch_len);
int fmx_set_device(int device);
/* replaces cudaMallocBytes / cudaFreeBytes (CudaHelpers.hpp:8-11, CudaHelpers.cu:23-35) */
int fmx_alloc(void **ptr, size_t bytes);
int fmx_free(void *ptr);
/* replace cudaMemcpyHtoD / DtoH / D2D / cudaMemsetBytes (CudaHelpers.hpp:14-23, CudaHelpers.cu:37-59);
 * asynchronous on `stream`, host buffers must stay valid until the stream is synchronised */
int fmx_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int fmx_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int fmx_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* cudaMemset semantics (every byte set to value & 0xff), done by a fill kernel rather than a memset
 * node, so it is safe inside captured graphs; fmx_fill_zero is fmx_memset(dst, 0, ...) */
int fmx_memset(void *dst, int value, size_t bytes, void *stream);
int fmx_fill_zero(void *dst, size_t bytes, void *stream);
int fmx_stream_create(void **stream);
int fmx_stream_destroy(void *stream);
int fmx_stream_sync(void *stream);
/* cudaCheckThrow (CudaHelpers.cu:16-20): last launch error + device synchronise */
int fmx_device_sync(void);
int fmx_set_sync_check(int on);
int fmx_event_create(void **event);
int fmx_event_destroy(void *event);
int fmx_event_record(void *event, void *stream);
int fmx_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---------------------------------------------------------- layout I/O */
/* soa[q*n + c] <-> aos[c*Q + q] for Q in {19, 7} (reference layout, parity I/O) */
int fmx_pack_aos(int prec, const void *soa, void *aos, int Q, int64_t n, void *stream);
int fmx_unpack_aos(int prec, const void *aos, void *soa, int Q, int64_t n, void *stream);

/* ----------------------------------------------------- NS operators (F32/F64) */
/* BGKCollision::initEq -> k_init_eq_d3q19_1d (BGKCollision.cu:44-71, 179-195);
 * PorousBGKCollision::initEq -> k_init_eq_d3q19 (PorousBGKCollision.cu:43-58): f = w * rho0 */
int fmx_ns_init_eq(int prec, void *f, const fmx_grid *grid, double rho0, void *stream);
/* NSCollisionOperator::collideStream (NSInterfaces.hpp:9-15), f -> f_new:
 *   FMX_NS_BGK    k_bgk_stream_pull_d3q19 (BGKCollision.cu:74-177, 197-223); obstacle nullable
 *   FMX_NS_POROUS k_porous_bgk_stream_pull_d3q19 (PorousBGKCollision.cu:64-155, 175-197); field = porosity
 *   FMX_NS_PBB    partial bounce-back (Walsh et al. 2009; no reference counterpart); field = solid fraction
 * tau is clamped to >= 0.5001 as in BGKCollision.cu:203-204 */
int fmx_ns_collide_stream(int prec, int op, const void *f, void *f_new, const uint8_t *obstacle, const void *field,
                          const fmx_grid *grid, double tau, double rho0, void *stream);
/* Lattice::computeMacroscopicDevice -> k_compute_macros_d3q19 (Macros.cu:20-61, 95-119) */
int fmx_ns_moments(int prec, const void *f, void *rho, void *ux, void *uy, void *uz, const fmx_grid *grid,
                   void *stream);
/* VelocityDirichletNEQ::apply, PreCollision on f (VelocityDirichletNEQ.cu:49-115); bnd/nei are the
 * compacted boundary and inward-neighbour cell lists built on the host (VelocityDirichletNEQ.hpp:26-107) */
int fmx_ns_bc_velocity_neq(int prec, void *f, const int64_t *bnd, const int64_t *nei, int64_t count, double ux,
                           double uy, double uz, int64_t ncells, void *stream);
/* VelocityDirichletEq::apply, PreCollision on f (VelocityDirichlet.cu:54-84, 111-130) */
int fmx_ns_bc_velocity_eq(int prec, void *f, const int64_t *cells, int64_t count, double rho0, double ux, double uy,
                          double uz, int64_t ncells, void *stream);
/* PressureDirichlet::apply, PostStreaming on f_new (PressureDirichlet.cu:26-96, 123-145); nei[k] is the
 * x-1 neighbour of cells[k] (mask cells with x = 0 are dropped on the host, PressureDirichlet.cu:50) */
int fmx_ns_bc_pressure(int prec, void *f_new, const int64_t *cells, const int64_t *nei, int64_t count, double rho_out,
                       int64_t ncells, void *stream);
/* BounceBack::apply, PostStreaming on f_new (BounceBack.cu:12-53), as a true swap of opposite pairs */
int fmx_ns_bc_bounceback(int prec, void *f_new, const int64_t *cells, int64_t count, int64_t ncells, void *stream);

/* ----------------------------------------------------- AD operators (F32/F64) */
/* ADBGKCollision::initEq -> k_ad_init_eq_d3q7 (ADBGKCollision.cu:29-38, 130-144): g = phi0 * w */
int fmx_ad_init_eq(int prec, void *g, const fmx_grid *grid, double phi0, void *stream);
/* ADBGKCollision::collideStream -> k_ad_bgk_stream_pull_d3q7_source (ADBGKCollision.cu:41-128, 146-166);
 * S and phi_out nullable */
int fmx_ad_collide_stream(int prec, const void *g, void *g_new, void *phi_out, const void *tau, const void *ux,
                          const void *uy, const void *uz, const void *S, const fmx_grid *grid, double min_conc,
                          void *stream);
/* ADDirichlet (post, on g_new: ADDirichlet.cu:11-53) and ADDirichletTanhPulse (pre, on g:
 * ADDirichletTanhPulse.cu:66-100): g <- phi * w on the listed cells */
int fmx_ad_bc_dirichlet(int prec, void *g, const int64_t *cells, int64_t count, double phi, int64_t ncells,
                        void *stream);
/* ADDirichletTanhPulse (pre, on g: ADDirichletTanhPulse.cu:66-100) with the per-step inlet value on
 * the device: g <- T(series[step]) * w on the listed cells (the pulse, or any inlet concentration
 * curve, tabulated per step; the host computes nothing per step) */
int fmx_ad_bc_dirichlet_series(int prec, void *g, const int64_t *cells, int64_t count, const double *series,
                               int64_t step, int64_t ncells, void *stream);
/* ADNeumannOutlet::apply (ADNeumannOutlet.cu:5-53): g_new[cells[k]] <- g_new[src[k]], all 7 directions */
int fmx_ad_bc_neumann(int prec, void *g_new, const int64_t *cells, const int64_t *src, int64_t count, int64_t ncells,
                      void *stream);

/* ------------------------------------------------------------- sorption */
/* launch_henry_ldf_source_d3q7 (LinearLDF.hpp:8-17, LinearLDF.cu:13-75) */
int fmx_ldf_linear(int prec, int64_t n, const uint8_t *region, const void *phi, void *S, void *q, double K,
                   double k_ldf, double min_conc, void *stream);
/* launch_langmuir_ldf_source_d3q7 (LangmuirLDF.hpp:7-14, LangmuirLDF.cu:7-59) */
int fmx_ldf_langmuir(int prec, int64_t n, const uint8_t *region, const void *phi, void *S, void *q, double q_max,
                     double b, double k_ldf, double min_conc, void *stream);

/* ------------------------------------------------ fused coupled step */
/* One coupled step of main.cpp:319-336 (NS.step -> moments/velocity hand-off -> sorption source ->
 * AD.step) as a single kernel launch per step, for the configuration the reference's main builds:
 * x is the flow axis with the inlet plane x = 0 and the outlet plane x = nx-1 (buildXPlaneMask,
 * helpers.hpp:7-14).  Requires nx >= 3. */
typedef struct fmx_coupled_desc {
    int32_t nx, ny, nz;
    int32_t periodic_y, periodic_z; /* x is never periodic */
    int32_t prec;                   /* population storage, FMX_F32/F64/F16/BF16 */
    int32_t ns_op;                  /* FMX_NS_* */
    int32_t ns_inlet;               /* 0 none, 1 VelocityDirichletNEQ, 2 VelocityDirichletEq, at x = 0 */
    int32_t ns_outlet;              /* 0 none, 1 PressureDirichlet at x = nx-1 (requires ad_outlet = 1) */
    int32_t ad_inlet;               /* 0 none, 1 ADDirichlet (post-stream) at x = 0,
                                       2 pre-collision Dirichlet at x = 0 with a per-step value */
    int32_t ad_outlet;              /* 0 none, 1 ADNeumannOutlet at x = nx-1 */
    int32_t isotherm;               /* FMX_ISO_* */
    int32_t coupling;               /* 1: AD advected by the NS velocity (main.cpp:322-326) */
    double rho0, tau;               /* LBMParams::rho0, tau */
    double u_in[3];                 /* inlet velocity */
    double rho_out;                 /* outlet density */
    double phi_in;                  /* ADDirichlet value */
    double tau_ad;                  /* AD relaxation time when tau_field is NULL */
    double min_conc;                /* LBMParams::minConcentration */
    double iso_a, iso_b, k_ldf;     /* linear: a = KHenry, k = kA; Langmuir: a = qMax, b = bLangmuir, k = kLDF */
    double f_ref;                   /* half storage of f: f = w * (f_ref + h) (fmx_coupled_recenter moves
                                       it); half storage of g holds g = w * (phi_src + s * h),
                                       s = max(phi_src, 2^-14), phi_src the source cell's lagged
                                       concentration (the x = nx-2 value on a Neumann outlet plane) */
    int32_t freeze_solids;          /* BGK: 1 = the prologue and fmx_checkpoint_load mark frozen solid cells
                                       (FMX_LINK_FROZEN: a solid cell at rest away from the x boundary
                                       planes, whose state only alternates), so that runs given a compact
                                       workspace skip them */
    int32_t reserved0;
} fmx_coupled_desc;

#define FMX_OUTLET_SLOTS 64
#define FMX_OUTLET_STRIDE 16 /* doubles between slots (one 128-byte line each) */

typedef struct fmx_coupled_bufs {
    void *f[2];               /* 19*N populations, ping-pong */
    void *g[2];               /* 7*N populations, ping-pong */
    void *q;                  /* N adsorbed load (AdsorbedField) */
    void *phi[2];             /* N concentration, ping-pong with f/g: phi[cur] is the previous step's clamped
                                 pre-collision sum (sorption input), the step writes phi[1-cur] */
    uint32_t *links;          /* N link words from fmx_coupled_build_links (the prologue updates the
                                 FMX_LINK_FROZEN bits) */
    const void *field;        /* N porosity (POROUS) or solid fraction (PBB); NULL = 1 / 0 */
    const void *tau_field;    /* N AD relaxation times; NULL = tau_ad */
    void *inlet_raw;          /* 19*ny*nz, required with ns_inlet != 0: the x = 0 plane's populations as the
                                 reference holds them between steps (before the next inlet update) */
    const double *inlet_series; /* ad_inlet = 2: inlet value of step s at inlet_series[s] */
    double *outlet_acc;       /* nullable: FMX_OUTLET_SLOTS*FMX_OUTLET_STRIDE doubles per step, accumulating
                                 phi over x = nx-1; step s (of a run starting at step0) at
                                 outlet_acc[(s - step0) * SLOTS * STRIDE]; caller zeroes it */
    double *outlet_ring;      /* nullable scratch of FMX_OUTLET_RING_ROWS*ny*nz doubles: with it, the outlet
                                 cells store phi there and every step's sum is added in a fixed order (every
                                 FMX_OUTLET_RING_ROWS steps and at the end of a run), so the breakthrough
                                 curve repeats to the bit; without it the sums are FP64 atomic adds */
    int64_t inlet_series_len; /* entries of inlet_series: a run over steps step0 .. step0+nsteps-1 reads
                                 entries step0 .. step0+nsteps-1, so it needs step0 + nsteps of them */
    int64_t outlet_acc_rows;  /* rows of outlet_acc: a run of nsteps steps writes rows 0 .. nsteps-1 */
    double *mass_log;         /* nullable: every mass_every-th step of a run appends one row of FMX_MASS_COLS
                                 doubles, fmx_reduce_mass of the state after that step (sum phi, sum q, sum
                                 of the stored f, non-finite count); row r after step step0 + (r+1)*mass_every - 1 */
    void *mass_workspace;     /* with mass_log: fmx_reduce_mass_workspace_bytes(nx*ny*nz) bytes of scratch */
    int64_t mass_log_rows;    /* rows of mass_log (>= nsteps / mass_every) */
    int64_t mass_workspace_bytes;
    int64_t mass_every;       /* steps between mass_log rows (>= 1 with mass_log) */
    void *compact_work;       /* nullable, BGK: workspace of fmx_coupled_compact_bytes(desc, active) bytes; a
                                 run then steps the NS populations of the cells without FMX_LINK_FROZEN in
                                 row-compacted buffers there (the frozen ones are never read or written), each
                                 step an NS launch over those cells and an AD launch over the box (FMX_SPLIT=0:
                                 one fused launch), and leaves f[] as without it, to the bit.  Ignored with mass_log, by runs of
                                 fewer than 16 steps and when the active cells do not fit (those runs use
                                 the box layout).  A run given it synchronises the stream once, after
                                 building its row table */
    int64_t compact_work_bytes;
} fmx_coupled_bufs;
#define FMX_OUTLET_RING_ROWS 64

int fmx_coupled_validate(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs);
/* Builds the per-cell link words from the flag bytes (FMX_FLAG_SOLID | FMX_FLAG_REGION per cell)
 * for the geometry and periodicity of `desc`; run once per geometry */
int fmx_coupled_build_links(const fmx_coupled_desc *desc, const uint8_t *flags, uint32_t *links, void *stream);
/* Lattice::initEquilibrium + ADLattice::initEquilibrium (NSLattice.hpp:177-185, ADLattice.hpp:134-141,
 * BGKCollision.cu:44-71, ADBGKCollision.cu:29-38) for the fused state: f[cur] = w*rho0, g[cur] = w*phi0,
 * q = 0, phi[cur] = 0 (AdsorbedField.hpp:19-20, ADLattice.hpp:49), followed by fmx_coupled_prologue */
int fmx_coupled_init(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, double phi0, void *stream);
/* Loads a state given in compute precision SoA (f_c 19*N, g_c 7*N, q_c N, phi_c N; each nullable = keep)
 * into f[cur], g[cur], q, phi[cur], followed by fmx_coupled_prologue */
int fmx_coupled_load(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, const void *f_c,
                     const void *g_c, const void *q_c, const void *phi_c, void *stream);
/* Copies the x = 0 plane of f[cur] to inlet_raw, applies the first step's inlet update to f[cur] and
 * marks frozen solid cells in bufs->links, writing their next phase into f[1-cur] (call after any
 * change of state made outside fmx_coupled_run, before the next fmx_coupled_run) */
int fmx_coupled_prologue(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, void *stream);
/* Steps step0 .. step0+nsteps-1.  State is read from f[cur], g[cur], phi[cur]; step k writes
 * buffer (cur + k + 1) % 2, so the result is in (cur + nsteps) % 2. */
int fmx_coupled_run(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int64_t step0, int32_t nsteps,
                    int cur, void *stream);
/* One step of fmx_coupled_run (step0 = step, nsteps = 1): the loop body of main.cpp:319-357 */
int fmx_coupled_step(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int64_t step, int cur, void *stream);
/* The NS populations of f[cur] as the reference holds them (x = 0 plane from inlet_raw), converted to
 * compute precision (float or double) SoA in f_out (19*N). */
int fmx_coupled_view_f(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, void *f_out,
                       void *stream);
/* g[cur] converted to compute precision SoA in g_out (7*N). */
int fmx_coupled_view_g(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, void *g_out,
                       void *stream);
/* No counterpart in the reference (its populations are FP32 or FP64).  Moves the density reference
 * f_ref of FP16 / BF16 NS storage (f = w * (f_ref + h)) to `target`, or, when target is NaN, to the
 * FP64 mean density of the cells of f[cur] without the FMX_LINK_SOLID bit (a reduction into
 * `workspace`, fmx_coupled_recenter_workspace_bytes(desc) bytes, then one synchronisation of the
 * stream); f[0] and f[1] are re-encoded about it (h += f_ref - target, one rounding) and *f_ref_out
 * receives it.  The caller sets desc->f_ref = *f_ref_out before the next call; a 16-bit checkpoint
 * (fmx_checkpoint_save) keeps the f_ref it was saved with.  Keeping f_ref at the flow's mean
 * density keeps the stored deviations small where a pressure drop builds up.  FP32 / FP64 storage:
 * the state is unchanged and *f_ref_out is still the mean (or target).  The reference must stay within
 * a factor of two of desc->f_ref (FMX_ERR_ARG otherwise). */
int64_t fmx_coupled_recenter_workspace_bytes(const fmx_coupled_desc *desc);
/* No counterpart in the reference (its BGK kernel reads and writes every solid cell).  Bytes of
 * fmx_coupled_bufs::compact_work for a run in which `active` cells lack FMX_LINK_FROZEN (count them in
 * the link words after the prologue): a row table of 16 bytes per 64 cells of a row, two compact
 * copies of the 19 population planes of the active cells, and per active cell an 8-byte record and a
 * velocity of three compute-precision values (none of the last two with FMX_SPLIT=0). */
int64_t fmx_coupled_compact_bytes(const fmx_coupled_desc *desc, int64_t active);
int fmx_coupled_recenter(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, double target,
                         double *f_ref_out, void *workspace, int64_t workspace_bytes, void *stream);
/* OutletCSV::log average (OutletCSV.hpp:51-80) for nsteps steps from outlet_acc: mean[k] = sum / count */
int fmx_outlet_mean(const double *outlet_acc, int64_t nsteps, int64_t count, double *mean, void *stream);

/* ------------------------------------------------------ discrete adjoint */
/* No counterpart in the reference (it has no adjoint): the exact reverse-mode derivative of the
 * fused step above (fmx_coupled_run), with any of its NS inlets (none, NEQ, equilibrium).  With FP16 / BF16 storage the
 * checkpoints and segment states keep the 16-bit deviation encoding and the reverse steps run in FP32.
 *
 * Loss: any L whose dependence on the run is through the outlet means C_out(s) (fmx_outlet_mean);
 * the caller supplies outlet_weight[s - step0] = dL/dC_out(s).  Gradients returned:
 *   grad_iso[0..1]  dL/d iso_a, dL/d iso_b (Langmuir q_max and K_L; linear: K_Henry, -)
 *   grad_field[c]   dL/d field[c] (partial-bounce-back fraction or porosity; clamps pass the gradient
 *                   on [0, 1])
 *   grad_series[s]  dL/d inlet_series[s] (ad_inlet = 2)
 * Checkpointing: fmx_adjoint_forward keeps the state before steps step0, step0+k, ... in `ckpt`
 * (ceil(nsteps/k) slots of fmx_adjoint_ckpt_bytes); fmx_adjoint_backward recomputes each segment
 * into `seg` (slots of fmx_adjoint_state_bytes) and runs the adjoint steps over it in reverse.
 * Checkpoint precision `ckpt_prec`: the storage precision (desc->prec: states kept as they are), or,
 * for an FP32 run, FMX_F16 / FMX_BF16: f and g are checkpointed as 16-bit deviations (f/w - f_ref,
 * g/w - phi_src, the 16-bit storage layout of the forward step), q and phi in FP32, and each segment
 * is recomputed in FP32 from its decoded checkpoint (seg then holds k slots instead of k-1).
 * For an FP64 run, FMX_F32: the FP64 state rounded to FP32 at each checkpoint, and the whole reverse
 * sweep -- segment recompute and adjoint steps -- in FP32 (seg slots, lam_* and field_rev /
 * tau_field_rev in FP32; seg slots of fmx_adjoint_ckpt_bytes(desc, FMX_F32) bytes, k-1 of them).  The
 * forward curve is the FP64 run's; the gradients are those of the FP32 update along the FP64 run's
 * trajectory, restarted from it at every checkpoint, so the FP32 run's own drift is confined to
 * one segment.
 * Two levels: given fewer segment slots than that (at least fmx_adjoint_seg_slots(k, recast, 1),
 * about 2 sqrt(k)), the sweep keeps sub-checkpoints every m steps within a segment and recomputes
 * one sub-segment at a time (one more forward pass over each segment; the same gradients, bit for
 * bit).  All sums over the grid run in a fixed order (no atomics): a call repeats its gradients to
 * the bit.  The sweep allocates nothing; its scratch is the caller's `workspace`. */

typedef struct fmx_adjoint_bufs {
    void *lam_f[2];              /* 19*N each, compute precision: adjoint populations (scratch) */
    void *lam_g[2];              /* 7*N each (scratch) */
    void *lam_q, *lam_phi;       /* N each (scratch) */
    void *workspace;             /* fmx_adjoint_workspace_bytes(desc) bytes (scratch) */
    int64_t workspace_bytes;     /* size of workspace */
    const double *outlet_weight; /* nsteps: dL/dC_out(step0 + i); NULL = no loss (all gradients 0) */
    int64_t outlet_weight_len;   /* entries of outlet_weight (>= nsteps; fmx_adjoint_step: > step) */
    double *grad_iso;            /* 2 doubles (written) */
    double *grad_field;          /* N doubles (written) or NULL */
    double *grad_series;         /* indexed like inlet_series; entries step0 .. step0+nsteps-1 written, or NULL */
    const void *field_rev;       /* FP64 run with FP32 checkpoints: bufs->field in FP32 (N; NULL where field is) */
    const void *tau_field_rev;   /* likewise bufs->tau_field in FP32; both unused otherwise */
} fmx_adjoint_bufs;

/* Bytes of one stored state (f, g, q, phi in the storage precision; each section 256-byte aligned) */
int64_t fmx_adjoint_state_bytes(const fmx_coupled_desc *desc);
/* Bytes of one checkpoint slot at precision ckpt_prec; -1 if the run's storage cannot be
 * checkpointed at that precision (ckpt_prec is the storage precision, FP16 / BF16 for an FP32 run,
 * or FP32 for an FP64 run) */
int64_t fmx_adjoint_ckpt_bytes(const fmx_coupled_desc *desc, int ckpt_prec);
/* fmx_coupled_run with checkpoints: stores the state before steps step0 + j*k into slot j of ckpt
 * (at ckpt_prec), outlet sums as fmx_coupled_run (row s - step0); result in buffer
 * (cur + nsteps) % 2.  Clears the frozen-solid link bits (the adjoint needs every cell on the full
 * update path). */
int fmx_adjoint_forward(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int64_t step0, int32_t nsteps,
                        int32_t k, int cur, void *ckpt, int ckpt_prec, void *stream);
/* Device scratch bytes fmx_adjoint_backward needs (per-launch-block running sums; with a per-step
 * inlet series also the per-block sums of 64 reverse steps): 16 bytes per 256-cell launch block,
 * 528 with ad_inlet 2.  It follows the launch shape (FMX_BLOCK / FMX_TILE, as the kernels do).
 * -1 for a bad descriptor. */
int64_t fmx_adjoint_workspace_bytes(const fmx_coupled_desc *desc);
/* Segment slots (of fmx_adjoint_state_bytes; of fmx_adjoint_ckpt_bytes(desc, FMX_F32) for an FP64 run
 * with FP32 checkpoints) a reverse sweep over segments of k steps needs: one level (two_level = 0)
 * or the two-level minimum (two_level = 1); recast = 1 for 16-bit checkpoints of an FP32 run (0 for
 * every other pair).  -1 for k < 1. */
int64_t fmx_adjoint_seg_slots(int32_t k, int recast, int two_level);
/* Reverse sweep over the steps of fmx_adjoint_forward (same desc, links, field, tau_field,
 * inlet_series and ckpt_prec; f/g/q/phi of bufs are not used).  Segment j is recomputed from ckpt
 * slot j into seg (seg_slots state slots), then its adjoint steps run last to first. */
int fmx_adjoint_backward(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int64_t step0, int32_t nsteps,
                         int32_t k, const void *ckpt, int ckpt_prec, void *seg, int64_t seg_slots,
                         const fmx_adjoint_bufs *adj, void *stream);
/* Checkpoints of a run (no counterpart in the reference, which keeps no run state): slot `slot` of
 * ckpt (slots of fmx_adjoint_ckpt_bytes(desc, ckpt_prec)) <- the fused state f[cur], g[cur], q,
 * phi[cur], as stored, (FP32 run) as FP16 / BF16 deviations or (FP64 run) rounded to FP32.
 * fmx_checkpoint_load puts it back (a 16-bit slot decoded to FP32, an FP32 slot widened to FP64) and
 * marks the frozen solid cells again from it (desc->freeze_solids; else it clears their link bits), so
 * fmx_coupled_run goes on from it as from the step it was saved before (exactly, for a checkpoint at
 * the storage precision). */
int fmx_checkpoint_save(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, void *ckpt, int ckpt_prec,
                        int64_t slot, void *stream);
int fmx_checkpoint_load(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, const void *ckpt,
                        int ckpt_prec, int64_t slot, void *stream);
/* One reverse step (no counterpart in the reference), for callers that keep their own checkpoints:
 * bufs holds the state at the start of `step` (f[cur], g[cur], q, phi[cur]; fmx_checkpoint_load)
 * and adj->lam_f / lam_g[lam_cur] the adjoint of step + 1's pulls; writes the adjoint of this step's
 * pulls into lam_f / lam_g[1 - lam_cur], updates lam_q / lam_phi, adds into grad_field and (ad_inlet
 * 2) grad_series[step], and rewrites grad_iso with the running total.  The loss weight of this step
 * is adj->outlet_weight[step] (indexed by the step itself).  first = 1 for the last step of the run:
 * sets the terminal condition (lam_*[lam_cur], lam_q, lam_phi = 0) and clears grad_field and the
 * running sums; grad_series is not cleared.  Steps taken last to first, lam_cur alternating, give
 * fmx_adjoint_backward's gradients to the bit. */
int fmx_adjoint_step(const fmx_coupled_desc *desc, const fmx_coupled_bufs *bufs, int cur, int64_t step, int lam_cur,
                     int first, const fmx_adjoint_bufs *adj, void *stream);

/* ---------------------------------------------------------- diagnostics */
/* No counterpart in the reference (its only reduction is OutletCSV::log's host loop, OutletCSV.hpp:29-58).
 * FP64 sums over all cells as wave64 reductions in a fixed order (a call repeats to the bit):
 * out[0] = sum phi, out[1] = sum q, out[2] = sum rho (= sum of all stored f), out[3] = number of
 * non-finite values seen (phi, q, or a cell whose populations sum to Inf / NaN).  f is SoA in `prec`
 * storage (half storage needs f_ref); phi / q in its compute precision; any input may be NULL (sum 0).
 * `workspace`: fmx_reduce_mass_workspace_bytes(n) bytes of device scratch (the per-block partials). */
#define FMX_MASS_COLS 4
int64_t fmx_reduce_mass_workspace_bytes(int64_t n);
int fmx_reduce_mass(int prec, const void *phi, const void *q, const void *f, int64_t n, double f_ref, double *out,
                    void *workspace, int64_t workspace_bytes, void *stream);
/* No counterpart in the reference.  fmx_reduce_mass for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc): the same kernels and the same bits,
 * with the partials in a 64 KiB array the library holds per GPU until the process ends instead of a
 * workspace argument.  Allocates nothing; at most 2^35 cells (FMX_ERR_ARG beyond).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (different streams, threads) race. */
int fmx_reduce_mass_static(int prec, const void *phi, const void *q, const void *f, int64_t n, double f_ref,
                           double *out, void *stream);
/* OutletCSV::log (OutletCSV.hpp:29-58) on the device: out[0] = mean, out[1] = min, out[2] = max of phi
 * (compute precision of `prec`) over the listed cells, in FP64 and a fixed summation order */
int fmx_reduce_outlet_mean(int prec, const void *phi, const int64_t *cells, int64_t count, double *out, void *stream);

/* Streaming-bandwidth probes: mode 0 copies `bytes` from src to dst (16 B per lane, non-temporal),
 * mode 1 only reads src (dst: a 4-byte sink); modes 2 / 3 / 4 copy 26 planes of n = bytes / (26 * e)
 * elements of e = 2 / 4 / 8 bytes, one element per plane per thread (the population access
 * pattern of the fused step without its arithmetic).  bytes must be a multiple of 16. */
int fmx_stream_probe(int mode, void *dst, const void *src, size_t bytes, void *stream);

/* ------------------------------------------------------------ geometry */
/* Seeded sphere pack (SplitMix64 -> xoshiro256**): solid walls on y in {0, ny-1} and z in {0, nz-1},
 * random spheres with integer-voxelised radius r ~ U[r_min, r_max], centres kept `clearance` cells away
 * from x = 0 and x = nx-1, added until the fluid fraction reaches target_porosity or max_tries
 * consecutive rejections (overlap = 0 forbids sphere overlap).  solid[c] = 1 for walls and spheres,
 * sphere[c] = 1 for sphere voxels only (nullable).  *porosity = fluid cells / all cells. */
int fmx_geom_sphere_pack(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double r_min, double r_max,
                         double target_porosity, int32_t overlap, int32_t clearance, int64_t max_tries,
                         uint8_t *solid, uint8_t *sphere, double *porosity);
/* buildXPlaneMask (helpers.hpp:7-14) */
int fmx_build_x_plane_mask(int32_t nx, int32_t ny, int32_t nz, int32_t x_plane, uint8_t *mask);
/* ASCII VTK ImageData with one UInt8 DataArray, x-fastest (the layout readVTI expects) */
int fmx_vti_write_mask(const char *path, int32_t nx, int32_t ny, int32_t nz, const uint8_t *mask, const char *name);
/* readVTI (VTIReader.hpp:35-160): pass mask = NULL to query dims; mask[c] = value > threshold
 * (3-component arrays use the magnitude; a short array fills the rest with 0) */
int fmx_vti_read_mask(const char *path, double threshold, int32_t *dims, uint8_t *mask, int64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* FMX_ABI_H */
