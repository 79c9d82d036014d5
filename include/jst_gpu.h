/*
 * jst_gpu.h — C ABI of the MI355X (gfx950) JST + Spalart-Allmaras solver library
 * (libjstgpu.so).  Plain pointers and sizes only; no torch or HIP types cross it.
 *
 * The reference (jst-rs) has no FFI: its seam is the crate's Rust API.  Each entry point
 * below names the reference item it replaces.  Handles are opaque and not thread-safe:
 * drive one context from one host thread; distinct contexts may run concurrently (each
 * owns its own HIP stream).  The caller owns every host buffer; nothing is retained after
 * a call returns.
 */
#ifndef JST_GPU_H
#define JST_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..6 map 1:1 onto jst-rs MeshError (mesh.rs:18-25); CONFIG is the
 * InvalidData io::Error of Config::from_str (config.rs:165-170); NONPHYSICAL is
 * NonPhysical (state.rs:241-259, solver.rs:151-185). */
typedef enum {
    JST_OK = 0,
    JST_ERR_IO = 1,
    JST_ERR_BAD_HEADER = 2,
    JST_ERR_EMPTY = 3,
    JST_ERR_COUNT_MISMATCH = 4,
    JST_ERR_BAD_POINT = 5,
    JST_ERR_TOO_SMALL = 6,
    JST_ERR_NONPHYSICAL = 7,
    JST_ERR_CONFIG = 8,
    JST_ERR_HIP = 9,
    JST_ERR_ARG = 10,
    JST_ERR_NOSPACE = 11 /* batch: the HBM arena or the case table cannot take the case now */
} jst_status;

typedef enum {
    JST_FP64 = 0,  /* plain path: every operation in FP64, reference operation order */
    JST_MIXED = 1  /* FP32 stencil arithmetic, FP64 geometry/accumulation/reductions */
} jst_precision;

typedef struct jst_ctx jst_ctx;

/* jst-rs state.rs:241-259 NonPhysical {i, j, rho, p} of the first row-major offender. */
typedef struct {
    long i, j;
    double rho, p;
} jst_nonphysical;

/* jst-rs solver.rs:38-43 RunReport. */
typedef struct {
    long steps;
    double residual;
    double totaltime;
    int converged;
} jst_report;

/* Per-step callback of Solver::run (solver.rs:203-226, `on_step(step, residual)`). */
typedef void (*jst_step_cb)(long step, double residual, void *user);

/* Thread-local text of the last error, worded like the reference's Display impls
 * (mesh.rs:27-44, state.rs:249-257).  Never NULL. */
const char *jst_last_error(void);

/* config.json follows the crate's schema (config.rs:10-72).  Default-off extension keys,
 * absent from the reference (left out, they reproduce it exactly):
 *   simulation.Re, simulation.ref_length   Reynolds number (by a mu0 rescale) and the
 *                                          reference length of Re and CL/CD (default 1)
 *   simulation.local_time_stepping (bool)  each cell advances with its own dt
 *   simulation.viscous_time_step (bool)    the viscous spectral radius joins the local dt
 *   spalart_allmaras.nut_ratio             nu~ / nu of the free stream (reference: 0.1)
 *   spalart_allmaras.negative (bool)       the negative-nu~ branch of the SA model
 *                                          (Allmaras, Johnson & Spalart, ICCFD7-1902):
 *                                          no eddy viscosity, fn-scaled nu~ diffusion and
 *                                          a vorticity-based source where nu~ < 0, wall
 *                                          ghosts included
 *   solver.residual_drop                   stop once residual < residual(step 1) 10^-drop
 *
 * Config::from_str (config.rs:165-170) + Config::finish (173-206) on the host only.
 * derived_out (may be NULL) receives 20 doubles: cv cp cw1 cv1_cubed c_inf u_inf v_inf
 * rho_inf mu_inf nut_inf supersonic suth_num inv_t_ref inv_pr inv_prt inv_kappa2 cw3_6
 * one_plus_cw3_6 mu0_eff halo.  No GPU is touched. */
jst_status jst_config_check(const char *config_json, double *derived_out);

/* Mesh::parse (mesh.rs:109-166) incl. trim_closing_duplicate (169-189) on the host only.
 * Fills n_rings / n_theta; xy (may be NULL) receives 2*n_rings*n_theta doubles. */
jst_status jst_mesh_parse(const char *mesh_text, int *n_rings, int *n_theta, double *xy,
                          size_t xy_len);

/* Solver::new(cfg, &Mesh) (solver.rs:47-62): geometry (FP64, on the GPU), freestream
 * initialisation and boundary state.  Validates exactly like Config::from_str and
 * Mesh::parse and returns their status. */
jst_status jst_create(const char *config_json, const char *mesh_text, jst_precision prec,
                      jst_ctx **out);
/* Solver::from_paths (solver.rs:64-69). */
jst_status jst_create_from_paths(const char *config_path, const char *mesh_path,
                                 jst_precision prec, jst_ctx **out);
void jst_destroy(jst_ctx *ctx);

/* Solver::{ni, nj} (solver.rs:71-78). */
jst_status jst_dims(const jst_ctx *ctx, int *ni, int *nj);

/* nsteps x Solver::advance (solver.rs:97-120) followed by Solver::residual (190-200):
 * res_out (may be NULL) receives one residual per completed step.  On a non-physical state
 * returns JST_ERR_NONPHYSICAL and fills err (may be NULL). */
jst_status jst_advance(jst_ctx *ctx, int nsteps, double *res_out, jst_nonphysical *err);

/* Solver::run (solver.rs:203-226).  max_steps <= 0 uses solver.iteration.  cb (may be
 * NULL) fires once per step in order; residuals are read back from the device in batches,
 * the step count is exact (the device stops on convergence). */
jst_status jst_run(jst_ctx *ctx, long max_steps, jst_step_cb cb, void *user, jst_report *out,
                   jst_nonphysical *err);

/* Solver::residual (solver.rs:190-200) of the last completed step (NaN before any). */
jst_status jst_residual(const jst_ctx *ctx, double *out);

/* timestep::compute (timestep.rs:24-47) + Solver::residual_terms (solver.rs:86-94): one
 * stage's Fc, Fv, Fd, S, gradients and SA auxiliaries on the current state, without
 * touching U.  dt_out receives the global minimum dt.  Read the terms with jst_get_field. */
jst_status jst_residual_terms(jst_ctx *ctx, double *dt_out);

/* Field<T>::to_interior_vec (field.rs:343-345): interior values row-major NI x NJ
 * ((NI+1) x NJ for tau_* faces).  Names:
 *   state:     rho p T u v E H c nut U0..U4 localdt
 *   terms:     Fc0..4 Fv0..4 Fd0..4 S grad0..7 (dudx dudy dvdx dvdy dTdx dTdy dnudx dnudy)
 *              chi fv1  (valid after jst_residual_terms)
 *   geometry:  vol cx cy wall_dist inv_vol inv_wd2 tau_nx tau_ny tau_mx tau_my n_nx n_ny
 *              n_mx n_my
 * n is the capacity of dst in doubles. */
jst_status jst_get_field(jst_ctx *ctx, const char *name, double *dst, size_t n);

/* Overwrite the interior conserved state (5 planes of NI x NJ, row-major) — test
 * injection such as rho = -1 (solver.rs:358-365).  Runs the unpack check and returns
 * JST_ERR_NONPHYSICAL with err filled when the state is not physical. */
jst_status jst_set_state(jst_ctx *ctx, const double *u_planes, jst_nonphysical *err);

/* No counterpart in the reference (it computes no forces): wall force from the wall
 * tau-face momentum fluxes, F = -sum_j (Fc - Fv)_tau(0,j)[mx,my], and CL/CD by the angle
 * of attack over q_inf * L.  Summed in FP64 in j order. */
jst_status jst_forces(jst_ctx *ctx, double *fx, double *fy, double *cl, double *cd);

/* No counterpart in the reference (wall diagnostics): per wall face j = 0..NJ-1 (the tau
 * faces (0, j), normal n into the fluid with |n| = edge length, unit tangent
 * s = (-n_y, n_x)/|n| in the direction of increasing j), from the same FP64 face fluxes as
 * jst_forces: x, y = face midpoint; Cp = (p_w - p_inf)/q_inf with
 * p_w = (Fc_mx n_x + Fc_my n_y)/|n|^2 (the mirrored wall ghost cancels the face momentum,
 * so the convective momentum flux is p_w n); Cf = (Fv[mx,my] . s)/(|n| q_inf), the
 * tangential traction the fluid exerts on the body.  Arrays of NJ doubles; any may be
 * NULL. */
jst_status jst_wall(jst_ctx *ctx, double *x, double *y, double *cp, double *cf);

/* No counterpart in the reference: jst_wall as text, columns j,x,y,cp,cf (j 1-based,
 * {:.8e} like result.csv). */
jst_status jst_write_wall_csv(jst_ctx *ctx, const char *path);

/* Solver::write_csv (solver.rs:229-255): columns i,j,x,y,rho,p,T,u,v,miubl, 1-based,
 * {:.8e}. */
jst_status jst_write_csv(jst_ctx *ctx, const char *path);

/* Step counter and accumulated physical time (Solver::{step, totaltime}). */
jst_status jst_progress(const jst_ctx *ctx, long *step, double *totaltime);

/* Timing hook for benchmarks (no reference counterpart): runs nsteps steps bracketed by HIP
 * events on the context's stream.  total_ms = whole region; when stage_ms is not NULL a
 * second pass of nsteps steps records events around every stage-kernel launch and
 * stage_ms receives the mean stage-kernel duration.  The device does not stop on
 * convergence during timing. */
jst_status jst_time_steps(jst_ctx *ctx, int nsteps, double *total_ms, double *stage_ms);
/* Timing hooks that let the caller time steps with its own clock (no reference counterpart):
 * jst_time_begin switches off the convergence and step-limit stops and does every set-up the
 * steps need (on the FP32-mixed path: launch lists and HIP graphs built and launched once),
 * then waits; jst_time_run queues exactly nsteps steps and waits for them; jst_time_end
 * restores the stops and reports a non-physical state reached meanwhile. */
jst_status jst_time_begin(jst_ctx *ctx, int nsteps);
jst_status jst_time_run(jst_ctx *ctx, int nsteps);
jst_status jst_time_end(jst_ctx *ctx);


/* ======================================================================================
 * Batched solver and serving loop.  No counterpart in the reference (it solves one case per
 * process); these are the callers of the same residual-and-update loop that keep many
 * cases resident in HBM and advance all of them with one kernel launch per RK stage.
 * ====================================================================================== */

typedef struct jst_batch jst_batch;

/* Per-case run state (device-resident, read back after each synchronised step batch). */
typedef enum {
    JST_CASE_RUNNING = 0,
    JST_CASE_CONVERGED = 1,   /* residual < targetres, or the relative residual_drop */
    JST_CASE_STOPPED = 2,     /* reached its step limit */
    JST_CASE_NONPHYSICAL = 3  /* diverged; the other cases continue */
} jst_case_status;

typedef struct {
    long step;
    int status;
    double residual, res_first, totaltime;
    long cells;
} jst_case_info;

/* A batch owns one HIP stream and an HBM arena of hbm_bytes for case state; max_cases is the
 * size of its case table.  prec must be JST_MIXED. */
jst_status jst_batch_create(jst_precision prec, size_t hbm_bytes, int max_cases, jst_batch **out);
void jst_batch_destroy(jst_batch *b);
/* Device bytes a case occupies (state, geometry, scratch): its admission footprint. */
jst_status jst_case_footprint(const char *config_json, const char *mesh_text, jst_precision prec,
                              size_t *bytes);
/* Admit a case (geometry precompute + freestream start on the device).  max_steps <= 0 uses
 * the config's solver.iteration.  Returns JST_ERR_NOSPACE when it does not fit now. */
jst_status jst_batch_add(jst_batch *b, const char *config_json, const char *mesh_text,
                         long max_steps, int *case_id);
/* Retire a case and return its memory to the arena. */
jst_status jst_batch_remove(jst_batch *b, int case_id);
/* Advance every running case by up to nsteps steps (5 stage launches + 1 step-end launch per
 * step for the whole batch, replayed from a HIP graph).  Each case stops by itself on
 * convergence, its step limit or a non-physical state.  sync != 0 waits and refreshes the
 * per-case info. */
jst_status jst_batch_step(jst_batch *b, int nsteps, int sync);
jst_status jst_batch_info(const jst_batch *b, int case_id, jst_case_info *out);
jst_status jst_batch_forces(jst_batch *b, int case_id, double *cl, double *cd);
/* State injection into a resident case (no reference counterpart; the batch form of
 * jst_set_state, and its FP32-mixed implementation): u_planes = 5 x NI x NJ conserved
 * variables, row-major per component, become U^n; the FP32 base and the frozen time step
 * are re-derived from it.  The Solver::unpack check then runs (solver.rs:151-171): a
 * non-physical cell fills err and returns JST_ERR_NONPHYSICAL.  With jst_batch_get_field's
 * U0..U4 this is checkpoint / resume of a case. */
jst_status jst_batch_set_state(jst_batch *b, int case_id, const double *u_planes,
                               jst_nonphysical *err);
/* jst_wall for one case of a batch (FP64 wall terms from the case's FP64 state). */
jst_status jst_batch_wall(jst_batch *b, int case_id, double *x, double *y, double *cp, double *cf);
/* Field export for one case (names as jst_get_field: state, geometry, and after
 * jst_batch_residual_terms the FP32 residual R0..R4 = Fc-Fv-Fd per component and S). */
jst_status jst_batch_get_field(jst_batch *b, int case_id, const char *name, double *dst,
                               size_t n);
jst_status jst_batch_residual_terms(jst_batch *b, int case_id, double *dt_out);
/* No counterpart in the reference.  For callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc): jst_batch_residual_terms without any
 * allocation.  The case's work-item list goes into a fixed device array of the library
 * (65536 items, 1 MiB per device, held until the process ends) instead of a fresh buffer; a
 * case needing more items fails with JST_ERR_ARG.  Call it with the batch's device current.
 * Results are bitwise those of jst_batch_residual_terms.  The array is shared by every call
 * on one device: calls that overlap on one device (two host threads, two batches) race. */
jst_status jst_batch_residual_terms_noalloc(jst_batch *b, int case_id, double *dt_out);
/* Residual history of a case: res[k] for steps first..first+n-1 (1-based), within the last
 * 4096 steps. */
jst_status jst_batch_history(jst_batch *b, int case_id, long first, long n, double *res);
/* Work decomposition of the stage kernel: variant 0..9 = row-segment length (16, 24, 32, 40,
 * 48, 64, 8, 4, 2 or 1 rows per wave work item of 58 columns) for every case, variant 10 =
 * each case its own length from its size alone (the shortest whose work items number at
 * most 2048, the waves the GPU keeps resident), variant 11 = one length for every case,
 * fitted to the whole batch whenever its cases change (the one minimising
 * ceil(items / 2048) x (L + 6), the row steps of the busiest resident wave), variant
 * 100 + L (1 <= L <= 256) = L rows for every case; grid = workgroups of 4 waves (0 = one
 * per 4 work items).  The grid never changes results; the segment length changes them only
 * at FP32 rounding level (where a segment's edge rows are evaluated), so under variant 11 a
 * case's result depends on which other cases share the batch at that level. */
jst_status jst_batch_set_tiling(jst_batch *b, int variant, int grid);
/* Row-segment length the stage kernel runs case case_id with under the current tiling
 * (under variant 11, the batch's fitted length for its current cases). */
jst_status jst_batch_segment_rows(const jst_batch *b, int case_id, int *rows);
/* Autotuner: times every variant with three grid sizes (0, 512, 1024) on the current batch
 * for trial_steps steps each, in two rounds (each candidate keeps its faster time; the steps
 * are real work and advance the cases), and keeps the fastest; returns the choice. */
jst_status jst_batch_tune(jst_batch *b, int trial_steps, int *variant, int *grid);
/* Benchmark hooks that let the caller time steps with its own clock: jst_batch_time_begin
 * switches off the convergence and step-limit stops of every running case, builds the
 * launch lists and the HIP graphs that queueing nsteps steps uses and launches each graph
 * once (up to GRAPH_STEPS + 1 untimed warm steps), then waits; jst_batch_step(b, nsteps, 1)
 * is then exactly nsteps steps of every running case plus one wait; jst_batch_time_end
 * restores the stops and fails with JST_ERR_NONPHYSICAL if a case diverged meanwhile. */
jst_status jst_batch_time_begin(jst_batch *b, int nsteps);
jst_status jst_batch_time_end(jst_batch *b);
/* Benchmark hook: nsteps steps of every running case with convergence stops disabled;
 * total_ms from HIP events around the region, stage_ms (may be NULL) the mean duration of
 * one stage launch from a second pass with events around every stage launch. */
jst_status jst_batch_time_steps(jst_batch *b, int nsteps, double *total_ms, double *stage_ms);
/* Benchmark hook: nsteps steps with HIP events around every launch; mean durations of the
 * RK stage 1-4 launches (inner_ms), the stage-5 launch (last_ms) and the step end (end_ms). */
jst_status jst_batch_time_kernels(jst_batch *b, int nsteps, double *inner_ms, double *last_ms,
                                  double *end_ms);
/* Cases of the batch still running (active, no non-physical state) as of the last
 * synchronisation with the device, and their cells in total. */
jst_status jst_batch_counts(const jst_batch *b, int *active_cases, long *active_cells);
/* Profiling calibration (no reference counterpart): one device-to-device copy of `bytes`
 * bytes, so that profiler byte counters can be scaled against a transfer of known size. */
jst_status jst_copy_probe(size_t bytes, double *ms);

/* One solve request of the serving loop. */
typedef struct {
    double arrival_s;         /* seconds after the loop starts */
    const char *config_json;
    const char *mesh_text;
    long max_steps;           /* <= 0: config solver.iteration */
} jst_request;

typedef struct {
    double admit_s, done_s;   /* seconds after the loop starts */
    long steps;
    int status;               /* jst_case_status at retirement, or -1 if never admitted */
    double residual, cl, cd;  /* cl, cd: NaN for a request that ends non-physical or is
                                 never admitted */
} jst_request_result;

typedef struct {
    size_t hbm_bytes;         /* arena for resident case state */
    int max_cases;            /* case-table slots (concurrent cases) */
    int steps_per_round;      /* steps launched between admission/retirement checks */
    int tune;                 /* retune the launch grid whenever the batch changes size class */
    int segment_rows;         /* row-segment length of the stage kernel.  0 = adaptive: each
                                 case its own length from its size alone (variant 10 of
                                 jst_batch_set_tiling: short segments, for latency) while
                                 the GPU has waves to spare, one length fitted to the batch
                                 while it is saturated (variant 11: fewer halo rows, every
                                 item in one round where possible, for throughput); a switch
                                 or a refit moves results at FP32 rounding level.  -1 = each case its
                                 own length always (results bitwise independent of the
                                 load).  1 .. 256 = that length for every case */
} jst_serve_opts;

typedef struct {
    double wall_s;
    double cell_stages;       /* sum over cases of NI*NJ*5*steps */
    double p50_s, p99_s;      /* time to completion (done - arrival) over every retired
                                 request: converged, stopped at its step limit or
                                 non-physical */
    int completed, nonphysical, max_concurrent;
    int tiling_switches;      /* adaptive segment_rows: times the batch changed tiling */
    int converged;            /* requests retired as JST_CASE_CONVERGED */
    int rejected;             /* requests never admitted (invalid, or larger than the arena) */
    double p50_converged_s, p99_converged_s; /* time to convergence (done - arrival) over the
                                                converged requests only; NaN if none */
} jst_serve_summary;

/* Serving loop: requests are admitted in arrival order as slots and arena memory free up,
 * run concurrently in one batch, and retired on convergence / step limit / divergence.  A
 * request that is invalid, or that does not fit the arena even when no other case is
 * resident, is rejected (its result status -1) and the loop goes on with the rest. */
jst_status jst_serve(const jst_request *reqs, int n, const jst_serve_opts *opts,
                     jst_request_result *results, jst_serve_summary *summary);

#ifdef __cplusplus
}
#endif
#endif
