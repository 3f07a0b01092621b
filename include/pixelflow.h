/*
 * pixelflow.h -- C-ABI of the MI355X (gfx950) solver for PixelFlow's time step.
 *
 * The reference has no FFI: its time step is a set of Fortran subroutines called from
 * each executable's `program main` with explicit-shape arrays dimension(0:md,0:nd[,0:ld])
 * (src/omp_parallel/ibm_3d_uniform_omp_cpu.f90:81-132, ibm_2d_uniform_omp_cpu.f90:80-125,
 * ibm_3d_air_condition_omp_cpu.f90:95-146).  Each entry point below names the reference
 * code it replaces.  Plain C types only; every call returns 0 on success or a negative
 * PF_E* code, with a message in pf_last_error().  A context owns all device memory; host
 * pointers are never kept after a call returns.  One host thread at a time per context.
 *
 * Host arrays use the reference's storage: Fortran column-major with ghost layers,
 * element (i,j,k), 0<=i<=m+1, 0<=j<=n+1, 0<=k<=l+1, at offset i + ldi*(j + ldj*k), where
 * ldi >= m+2 and ldj >= n+2 are the leading dimensions (md+1, nd+1 for the reference's
 * static arrays; m+2, n+2 for exact-size allocatables).  2D fields ignore k.
 */
#ifndef PIXELFLOW_H
#define PIXELFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 3

enum pf_status {
    PF_OK = 0,
    PF_EINVAL = -1,      /* bad argument or descriptor */
    PF_EHIP = -2,        /* HIP runtime error */
    PF_ENOMEM = -3,      /* device allocation failed */
    PF_ESTATE = -4,      /* call out of order (e.g. step before porosity) */
    PF_EDIVERGED = -5    /* a non-finite value appeared (reported, never aborts) */
};

enum pf_bc { PF_BC_UNIFORM = 0, PF_BC_AIR = 1 };
enum pf_precision { PF_PREC_FP64 = 0, PF_PREC_MIXED = 1 };
/* Pressure solve of tolerance mode (residual_tol > 0).  No reference counterpart: the
 * reference runs iter_max SOR iterations with no convergence test
 * (ibm_3d_uniform_omp_cpu.f90:460-611). */
enum pf_solver {
    PF_SOLVER_SOR = 0,   /* cycles of inner_iters red/black SOR iterations */
    PF_SOLVER_MG = 1     /* BiCGStab preconditioned by a multigrid V-cycle, to inner_tol per cycle */
};
enum pf_field { PF_FIELD_U = 0, PF_FIELD_V = 1, PF_FIELD_W = 2, PF_FIELD_P = 3, PF_FIELD_POROSITY = 4 };

/* Everything one case needs: the controlDict variables (src/omp_parallel/lib/global.f90:
 * 17-26) after the derivations of grid_conditions* (dx = width/(m-1), dt = time/istep_max,
 * src/omp_parallel/lib/grid.f90:50-52, 163-166, 297-300). */
typedef struct pf_desc {
    int32_t dim;              /* 2 or 3 */
    int32_t m, n, l;          /* interior extents (l = 1 in 2D) */
    int32_t bc;               /* pf_bc */
    int32_t nonslip;          /* wall-force term on/off */
    int32_t iter_max;         /* FP64: fixed SOR iterations per step (reference semantics) */
    int32_t face[6];          /* air: top,bottom,east,west,north,south; 0 wall,1 inlet,2 outlet */
    int32_t precision;        /* pf_precision */
    int32_t inner_iters;      /* mixed: FP32 SOR iterations per defect-correction cycle */
    int32_t max_cycles;       /* mixed, tolerance mode: cap on cycles per step */
    int32_t report_residual;  /* mixed, fixed-count mode: also evaluate the final FP64 residual */
    int32_t inlet_porosity;   /* 2D uniform: inlet and initial velocity scaled by porosity, the
                                 backward-facing-step executable (ibm_2d_backstep_omp_cpu.f90:
                                 533-534, 615-618); 0 for the others */
    int32_t solver;           /* pf_solver (tolerance mode; fixed-count mode always runs SOR) */
    double dx, dy, dz, dt;
    double xnue, xlambda, density, thickness;
    double inlet_velocity, outlet_pressure, aoa;
    double omega;             /* relux_factor */
    double residual_tol;      /* mixed: 0 = run iter_max iterations in cycles of inner_iters (the
                                 reference's fixed count); > 0 = cycles of inner_iters until the
                                 FP64 residual max|r/ap| <= residual_tol (or max_cycles) */
    double inner_tol;         /* PF_SOLVER_MG: relative residual of each inner solve (0 = 1e-3);
                                 inner_iters then caps its BiCGStab iterations (0 = 30) */
} pf_desc;

/* Per-step report. sor_error = the reference's "SOR iteration no. .. -- p error:" value
 * (running max of |p - p_old|, ibm_3d_uniform_omp_cpu.f90:575-583). */
typedef struct pf_stats {
    double sor_error;         /* PF_SOLVER_MG: max |p_new - p_old| of the step's updates */
    double residual;          /* mixed: FP64 residual max|r/ap| (tolerance mode: the last one
                                 evaluated; fixed-count mode: after the step if report_residual,
                                 else -1); FP64 path: -1 */
    int32_t sor_iters;        /* SOR iterations executed in the step (PF_SOLVER_MG: V-cycles) */
    int32_t cycles;           /* mixed: defect-correction cycles (FP64 path: 0) */
} pf_stats;

typedef struct pf_ctx pf_ctx;

/* Replaces the static arrays and set-up of program main (ibm_3d_uniform_omp_cpu.f90:13-19,
 * 42-59): validates the descriptor and allocates every device field. */
int pf_create(const pf_desc *desc, pf_ctx **out);
int pf_destroy(pf_ctx *ctx);
const char *pf_last_error(const pf_ctx *ctx);   /* NULL ctx: the last pf_create failure */
int pf_abi_version(void);

/* grid_conditions / grid_conditions_wall / grid_conditions_yz_periodic
 * (src/omp_parallel/lib/grid.f90:36-47 + 92-106, 151-158 + 215-243, 285-292 + 349-377):
 * raw interior porosity (i fastest, m*n[*l] values) clamped to threshold, ghosts filled;
 * then the time-invariant Poisson face coefficients are built once on the device. */
int pf_set_porosity(pf_ctx *ctx, const double *raw_interior, double threshold);

/* initial_conditions followed by boundary (ibm_3d_uniform_omp_cpu.f90:68-71, :756-793;
 * 2D :67-70, :542-568; air ibm_3d_air_condition_omp_cpu.f90:1174-1211). */
int pf_initial(pf_ctx *ctx);

/* Host <-> device copies of one field in the reference's host storage (see top). */
int pf_upload(pf_ctx *ctx, int field, const double *host, int64_t ldi, int64_t ldj);
int pf_download(pf_ctx *ctx, int field, double *host, int64_t ldi, int64_t ldj);

/* nsteps iterations of the time loop body (ibm_3d_uniform_omp_cpu.f90:85-128: copy to
 * *_old, solve_p, projection, boundary; 2D :84-118; air :99-142), device resident.
 * stats (may be NULL) receives nsteps entries.  Returns after the work is queued and the
 * statistics are read back. */
int pf_step(pf_ctx *ctx, int nsteps, pf_stats *stats);

/* output_divergent_2d/3d (src/omp_parallel/lib/output.f90:217-224, 929-940): the
 * porosity-weighted flux divergence of the current velocity, interior only, i fastest. */
int pf_divergence(pf_ctx *ctx, double *out_interior);

/* output_force_log_2d/3d (src/omp_parallel/lib/output.f90:244-305, 1090-1165), the drag
 * executable's per-step log (ibm_2d_drag_omp_cpu.f90:121): out[0..5] = pressure force x,
 * y, z and wall-force (viscous) force x, y, z on the immersed body for the current fields
 * (2D: z entries 0).  Each cell's terms are the reference's; the sum is a fixed-order
 * device tree (deterministic, equal to the reference's serial sum up to rounding). */
int pf_force(pf_ctx *ctx, double *out);

/* No reference counterpart.  pf_force without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same kernels and the same result bit for bit, with the reduction's
 * partial sums in a fixed device array of the library (1024 partial blocks, about 48 KB
 * per GPU, held until the process ends) instead of a buffer pf_force allocates on its first
 * call.  Sizes past that array are rejected with PF_EINVAL (the force reduction never
 * needs more).  All calls on one GPU share the array: calls that overlap on one GPU race. */
int pf_force_noalloc(pf_ctx *ctx, double *out);

/* Device bytes the context holds (admission control sizes cases with this). */
int64_t pf_footprint_bytes(const pf_desc *desc);

/* Wait for all queued work of the context. */
int pf_sync(pf_ctx *ctx);

/* Measurement hooks (no reference counterpart: the reference prints wall-clock stamps
 * only, src/omp_parallel/lib/utils.f90:7-16).  With timing on, pf_step brackets each
 * step's SOR sweeps with HIP events on the context's stream; pf_last_sweep_time_us then
 * gives the average time per colour sweep (half an SOR iteration; a fused launch does
 * two) of the last pf_step, gaps between the back-to-back launches included, and the
 * number of colour sweeps it covers. */
int pf_set_timing(pf_ctx *ctx, int on);
/* Which colour-sweep kernel the context uses (0 stored face coefficients, 1 faces rebuilt
 * from porosity with IEEE division, 2 the same with the constant-divisor division that
 * pf_set_porosity verified face by face).  All three give identical results. */
int pf_sor_kernel(const pf_ctx *ctx);
/* Bandwidth calibration (no reference counterpart): a plain device-to-device copy of
 * `bytes`, `iters` times; *gbps = (read + written bytes) / second. */
int pf_probe_copy(int64_t bytes, int iters, double *gbps);
/* 1 if the context runs each SOR iteration as one fused launch (both colours, ping-pong
 * buffers; even periodic extents; opt-in with PF_FUSED=1 in the environment), 0 if as two
 * colour-sweep launches (the default, faster on MI355X).  Same results. */
int pf_sor_fused(const pf_ctx *ctx);
/* 1 if the FP64 colour sweeps read the porosity coded in 4 bytes per value (values FP32
 * holds exactly as floats, the others through a table of at most 2^20 distinct doubles:
 * the same doubles, 4 fewer bytes per cell and sweep; the default past 512 MB of sweep
 * arrays, PF_PHI_CODE=0/1 in the environment decides), 0 if they read it as stored FP64
 * (smaller grids, too many distinct values, or a precision other than FP64).  Same
 * results.  No reference counterpart. */
int pf_sor_coded(const pf_ctx *ctx);
/* Cells of both colours that the colour sweeps update in waves skipping the porosity
 * loads: waves whose cells and their six neighbours all have porosity exactly 1 take 1 for
 * those operands (the same values, 8 bytes fewer per updated cell and sweep in FP64; one
 * flag byte per 32 colour slots of a row marks them, built by pf_set_porosity).  FP64
 * contexts count the FP64 sweeps, mixed-precision ones the FP32 sweeps (whose unit values
 * are those that round to 1).  PF_UNIT_PHI=0 in the environment turns it off (*cells is
 * then 0).  Same results.  No reference counterpart. */
int pf_sor_unit_cells(const pf_ctx *ctx, int64_t *cells);
int pf_last_sweep_time_us(const pf_ctx *ctx, double *avg_us, int64_t *launches);

/* ------------------------------------------------------------------------------------
 * Batches: many independent cases of uneven size advanced together, one launch per stage
 * of the step for all of them, the step captured in a hipGraph.  No reference
 * counterpart (the reference runs one case per process); every case of a batch computes
 * exactly what pf_step computes for it alone.  Cases of one batch share dim, bc, precision
 * and the iteration settings (iter_max, inner_iters, max_cycles, residual_tol); they may be
 * added and removed between steps (continuous batching).
 * ------------------------------------------------------------------------------------ */
typedef struct pf_batch pf_batch;

typedef struct pf_batch_config {
    int32_t persistent_blocks;  /* 0: one block per tile; > 0: this many blocks walk the tiles */
    int32_t use_graph;          /* capture the fixed-count step in a hipGraph and replay it */
    int32_t inner_iters;        /* mixed precision: FP32 iterations per defect-correction cycle */
    int32_t tile_x;             /* colour-sweep tile width in slots: 32, 64 or 128; 0 = the
                                   default (32 in 3D, 64 in 2D) */
} pf_batch_config;

/* kind: the shared settings; arena_bytes: device memory for the cases' fields. */
int pf_batch_create(const pf_desc *kind, int max_cases, int64_t arena_bytes, pf_batch **out);
int pf_batch_destroy(pf_batch *b);
const char *pf_batch_last_error(const pf_batch *b);
/* Device bytes one case takes in a batch arena (-1 for an invalid descriptor). */
int64_t pf_batch_case_bytes(const pf_desc *desc);
/* Largest free contiguous range of the arena. */
int64_t pf_batch_free_bytes(const pf_batch *b);
/* Add a case (grid_conditions + initial_conditions + boundary, as pf_set_porosity and
 * pf_initial); *slot receives its slot number in [0, max_cases). */
int pf_batch_add(pf_batch *b, const pf_desc *desc, const double *raw_interior, double threshold, int *slot);
int pf_batch_remove(pf_batch *b, int slot);
int pf_batch_count(const pf_batch *b);
/* nsteps steps of every case; stats[s * max_cases + slot] for the occupied slots. */
int pf_batch_step(pf_batch *b, int nsteps, pf_stats *stats);
int pf_batch_download(pf_batch *b, int slot, int field, double *host, int64_t ldi, int64_t ldj);
int pf_batch_divergence(pf_batch *b, int slot, double *out_interior);
int pf_batch_set_config(pf_batch *b, const pf_batch_config *cfg);
int pf_batch_get_config(const pf_batch *b, pf_batch_config *cfg);
int pf_batch_sync(pf_batch *b);
/* Steps the case in `slot` has taken since it was added (-1 if the slot is empty). */
int64_t pf_batch_case_steps(const pf_batch *b, int slot);
/* Free and total device memory (hipMemGetInfo), for admission control. */
int pf_device_memory(int64_t *free_bytes, int64_t *total_bytes);

/* ------------------------------------------------------------------------------------
 * Serving: continuous batching over one resident batch, driven by a worker thread of the
 * library (no reference counterpart).  Requests submitted from any host thread are
 * admitted in arrival order as slots and arena memory free up, advanced in shared step
 * chunks with the cases already resident, and retired when their step count is reached
 * (the fields asked for are kept in host memory) or when their step became non-finite.
 * Every result equals the case run alone.  Times are seconds since pf_serve_create.
 * ------------------------------------------------------------------------------------ */
enum pf_serve_state { PF_SERVE_QUEUED = 0, PF_SERVE_RUNNING = 1, PF_SERVE_DONE = 2, PF_SERVE_FAILED = 3 };
typedef struct pf_server pf_server;

/* kind and max_cases as for pf_batch_create; arena_bytes <= 0: half the free device memory;
 * chunk: steps advanced between admissions. */
int pf_serve_create(const pf_desc *kind, int max_cases, int64_t arena_bytes, int chunk, pf_server **out);
int pf_serve_destroy(pf_server *s);
/* The server's batch (read-only use: pf_batch_free_bytes, pf_batch_get_config; a
 * pf_batch_set_config is safe only while no request is queued or running). */
pf_batch *pf_serve_batch(pf_server *s);
/* fields: bit PF_FIELD_* set = keep that field for pf_serve_result.  The porosity is copied. */
int pf_serve_submit(pf_server *s, const pf_desc *desc, const double *raw_interior, double threshold, int64_t nsteps,
                    int32_t fields, int64_t *request_id);
/* state (pf_serve_state), steps taken so far, SOR iterations over those steps, the
 * statistics of its latest step (pointers other than state may be NULL). */
int pf_serve_poll(pf_server *s, int64_t id, int *state, int64_t *steps_done, int64_t *sor_iters, pf_stats *last);
/* Per field, the maximum of the request's per-step statistics over the steps run so far
 * (the worst residual shows whether every step met a tolerance; a NaN is kept). */
int pf_serve_worst(pf_server *s, int64_t id, pf_stats *worst);
/* Block until the request is done or failed, or timeout_s (< 0: no limit) passes. */
int pf_serve_wait(pf_server *s, int64_t id, double timeout_s, int *state);
int pf_serve_times(pf_server *s, int64_t id, double *t_submit, double *t_admit, double *t_finish);
/* A kept field of a done request, in the reference's host storage (see top). */
int pf_serve_result(pf_server *s, int64_t id, int field, double *host, int64_t ldi, int64_t ldj);
/* Forget a finished request and its fields. */
int pf_serve_release(pf_server *s, int64_t id);
/* Steps run, seconds spent inside batch steps, requests queued and running. */
int pf_serve_counters(pf_server *s, int64_t *steps_run, double *busy_s, int *queued, int *running);
/* Why a request failed; for an unknown id, the server's last fatal error.  The text is a
 * copy owned by the calling thread, valid until that thread's next pf_serve_error call. */
const char *pf_serve_error(pf_server *s, int64_t id);

#ifdef __cplusplus
}
#endif
#endif
