/*
 * iblbm_hip.h -- C ABI of libiblbm_hip.so, the gfx950 IB-LBM per-step hot path.
 *
 * The reference binds its hot path as Python-level operators over CuPy arrays
 * (iblbm/gpu_kernels.py:760-1223, selected by iblbm/backend.py:14-36).  This library
 * replaces that surface with two layers behind plain C types (no torch or CuPy types):
 *
 *  1. A state-owning, fused step interface (iblbm_ctx_*, iblbm_instance_*, iblbm_step):
 *     one launch per time step runs pull-streaming, the boundary closure, the macroscopic
 *     moments, the equilibrium and the collision for every node, and one workgroup per
 *     instance runs the immersed-boundary stage (DF, MDF or DFC) over the marker window in
 *     the same launch.  It replaces the step sequence of iblbm/runtime/step.py:21-78 and
 *     iblbm/runtime/loop.py:59-87.
 *  2. Operator-level entry points with the argument meaning of the reference's
 *     gpu_kernels.py wrappers (iblbm_op_*): host arrays in, host arrays out.
 *
 * Conventions (identical to the reference): populations are (N, 9) row-major AoS on the
 * host with node n = j*nx + i; U is (N, 2); Eulerian fields are (ny, nx); marker
 * coordinates Lx, Ly are in domain units (x = i*dx); the lattice is D2Q9 with
 * e = [(0,0),(1,0),(0,1),(-1,0),(0,-1),(1,1),(-1,1),(-1,-1),(1,-1)].
 *
 * Every function returns 0 on success or a negative IBLBM_E* code; iblbm_last_error()
 * returns the message of the most recent failure on the calling thread.  A context is not
 * thread-safe; iblbm_step is asynchronous on the context's stream and the download and
 * reduction calls synchronise it.
 */
#ifndef IBLBM_HIP_H
#define IBLBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBLBM_OK 0
#define IBLBM_EINVAL (-1)   /* invalid argument: maps to ValueError */
#define IBLBM_EKEY (-2)     /* unknown model / method / kernel name: maps to KeyError */
#define IBLBM_EDEVICE (-3)  /* HIP runtime or device failure: maps to RuntimeError */
#define IBLBM_ENOMEM (-4)   /* device allocation failed: maps to MemoryError */

/* Collision models: iblbm/lbm/collision/{bgk,trt,cm_mrt}.py */
#define IBLBM_COLL_BGK 0
#define IBLBM_COLL_TRT 1
#define IBLBM_COLL_CM_MRT 2
/* Boundary closures: iblbm/boundary/{dirichlet_neumann,settling_channel}.py */
#define IBLBM_BC_DIRICHLET_NEUMANN 0
#define IBLBM_BC_SETTLING_CHANNEL 1
/* Closures that read the previous step's velocity or an analytic velocity (run by a small
 * boundary kernel beside the bulk): iblbm/boundary/{velocity_inlet,open_boundary,
 * tg_analytical}.py.  Marker windows must stay three nodes away from every wall. */
#define IBLBM_BC_VELOCITY_INLET 2
#define IBLBM_BC_OPEN 3
#define IBLBM_BC_TG_ANALYTICAL 4
/* IBM methods: iblbm/ibm/strategies/{df,mdf,dfc}.py */
#define IBLBM_IBM_NONE 0
#define IBLBM_IBM_DF 1
#define IBLBM_IBM_MDF 2
#define IBLBM_IBM_DFC 3
/* Delta kernels: iblbm/gpu_kernels.py:952-953 ids (peskin4pt = 0, hat = 1) */
#define IBLBM_DELTA_PESKIN4PT 0
#define IBLBM_DELTA_HAT 1
/* Storage precision of the populations */
#define IBLBM_FP64 0
#define IBLBM_FP32_DEV 1 /* FP32 deviations f_i - w_i * rho0, FP64 IBM reductions */

/* One lattice instance (a SimConfig after iblbm/init.py:116-265 has run). */
typedef struct iblbm_case {
    int32_t nx, ny;
    int32_t collision;       /* IBLBM_COLL_* */
    int32_t bc;              /* IBLBM_BC_* */
    int32_t ibm;             /* IBLBM_IBM_* */
    int32_t delta;           /* IBLBM_DELTA_* */
    int32_t incompressible;  /* SimConfig.incompressible_lbgk */
    int32_t precision;       /* IBLBM_FP64 | IBLBM_FP32_DEV */
    double tau;              /* init.py:165-168 */
    double tau_minus;        /* TRT: 0.5 + Lambda/(tau - 0.5), trt.py:75 */
    double s_diag[9];        /* CM_MRT relaxation rates, cm_mrt.py:171-188 */
    double dt, dx, dy;       /* dt = 1 (lattice), dx = dy = 1/(NN-1) (domain units) */
    double in_u, dens;       /* Dirichlet inflow velocity and reference density */
    int32_t mdf_iterations, mdf_min_iterations;
    double mdf_tolerance, mdf_omega; /* omega from iblbm_omega_estimate */
    int32_t lb_capacity;     /* maximum markers this instance will hold */
    int32_t box_capacity;    /* maximum window nodes (bw*bh) the marker set will cover */
    int32_t fixed_body;      /* 1: markers never move (window tables and DFC lambda cached) */
    int32_t reserved;
} iblbm_case;

/* Per-step IBM statistics (iblbm/ibm/mdf.py:224-229; coupled iteration of several particles,
 * iblbm/ibm/multi_particle.py:214-221, adds guard_trips and omega_used). */
typedef struct iblbm_mdf_stats {
    int32_t iter_count;
    int32_t diverged;
    int32_t converged;
    int32_t guard_trips;  /* coupled: rejected iterations (omega halved); else 0 */
    double last_residual;
    double omega_used;    /* coupled: omega at the end of the step; else the instance's omega */
} iblbm_mdf_stats;

/* Diagnostic reductions of the most recent step that requested them (FP64). */
typedef struct iblbm_diag {
    double rho_sum;     /* sum of rho over all N nodes (mean = rho_sum / N) */
    double force[2];    /* DF/MDF: sum of fib over the grid; DFC: sum of marker forces q_k */
    double torque_fib;  /* DF/MDF: sum over the grid of (i - cx)*fib_y - (j - cy)*fib_x */
    double torque_q;    /* DFC: sum_k (Lx_k - cx)/dx * q_y - (Ly_k - cy)/dy * q_x */
    int64_t step;       /* step counter at which these were taken */
} iblbm_diag;

/* --- context and instances (replaces iblbm/runtime/step.py + loop.py:59-87) --- */
int iblbm_ctx_create(int device, void **ctx);
int iblbm_ctx_destroy(void *ctx);
const char *iblbm_last_error(void);
int iblbm_device_name(int device, char *buf, int buflen);

/* Registers an instance; the device buffers are allocated here.  *id receives its index. */
int iblbm_instance_add(void *ctx, const iblbm_case *c, int *id);

/* Post-collision populations f (the array the next step streams), AoS (N,9) FP64. */
int iblbm_upload_f(void *ctx, int id, const double *f_aos);
int iblbm_download_f(void *ctx, int id, double *f_aos);

/* Markers in domain units, desired velocity (Lb,2) in lattice units and arc length Larea.
 * The window box (every clamped stencil node, see iblbm_marker_box) must fit box_capacity.
 * Replaces the per-step H2D of iblbm/runtime/scenarios/sedimentation_single.py:156-160. */
int iblbm_set_markers(void *ctx, int id, const double *Lx, const double *Ly, const double *ud,
                      int Lb, double Larea);
/* Several particles in one DF / MDF window (replaces the per-particle loops of
 * iblbm/ibm/multi_particle.py:62-455): the markers passed to iblbm_set_markers are the
 * particles' rings concatenated in order, lb[p] markers of arc length larea[p] each.  DF spreads
 * each particle's closure on its own and adds them in particle order (:347-421); MDF runs the
 * coupled iteration (:62-316) from the instance's mdf_omega, which the caller sets to
 * min(0.8 omega_est, 0.8).  DFC windows ignore this (their rings are spread as one).  Call it
 * before iblbm_set_markers; np <= 8. */
int iblbm_set_particles(void *ctx, int id, int np, const int *lb, const double *larea);
/* Coupled MDF: particle `part`'s accumulated window fib of the last step, box as in
 * iblbm_download_window (its force and torque are grid sums of it, multi_particle.py:199-211). */
int iblbm_download_particle_fib(void *ctx, int id, int part, double *fibx, double *fiby,
                                int box[4]);
/* DF of several particles: the marker forces (Lb, 2) of the last step (a particle's force is
 * -sum(Lf * Larea) over its markers, iblbm/ibm/forcing.py:263-275). */
int iblbm_download_marker_force(void *ctx, int id, double *F);
/* Centre (domain units) about which the diagnostic torque sums are taken. */
int iblbm_set_torque_center(void *ctx, int id, double cx, double cy);
/* iblbm_set_markers (and, with center != NULL, iblbm_set_torque_center) for n instances of
 * one context at once: one host-to-device copy, one scatter launch and one table launch for
 * all, one synchronisation.  Instance ids[b] gets lb[b] markers taken in order from the
 * concatenated Lx, Ly (sum lb) and ud (sum lb, 2), arc length larea[b] and torque centre
 * center[2b], center[2b+1].  Every instance is checked before any changes.  Replaces the
 * per-particle marker uploads of a batched settling ensemble, each member as
 * iblbm/runtime/scenarios/sedimentation_multi.py:49-94 (pre_step) and
 * iblbm/runtime/scenarios/sedimentation_single.py:156-160 set them. */
int iblbm_set_markers_batch(void *ctx, int n, const int *ids, const int *lb, const double *Lx,
                            const double *Ly, const double *ud, const double *larea,
                            const double *center);
/* The marker forces of n instances in one read, concatenated in ids order, (Lb, 2) each: the
 * DFC force q of a DFC instance (iblbm/ibm/dfc.py:232-253, as iblbm_download_dfc_force) or the
 * marker forces of a several-particle DF window (as iblbm_download_marker_force).  Replaces the
 * per-member force reads of iblbm/ibm/strategies/dfc.py:129-150 and
 * iblbm/ibm/forcing.py:263-275 in a batched settling ensemble. */
int iblbm_download_marker_forces_batch(void *ctx, int n, const int *ids, double *F,
                                       int64_t capacity);
/* capacity: the marker rows (pairs of doubles) F holds; a call whose instances carry more
 * markers than that fails with IBLBM_EINVAL and writes nothing. */
/* Window box [box[0], box[0]+box[2]) x [box[1], box[1]+box[3]) covering every clamped
 * stencil node of the markers (stencil radius 2 for peskin4pt, 1 for hat). */
int iblbm_marker_box(const double *Lx, const double *Ly, int Lb, double dx, double dy, int nx,
                     int ny, int delta, int box[4]);

/* Closures that read the previous step's velocity (IBLBM_BC_VELOCITY_INLET, IBLBM_BC_OPEN):
 * the velocity U (N,2) the first step's closure reads, i.e. the initial state's velocity
 * (iblbm/init.py:175-195); the library keeps the three outermost rings of it. */
int iblbm_set_boundary_velocity(void *ctx, int id, const double *U);
/* IBLBM_BC_TG_ANALYTICAL (iblbm/boundary/tg_analytical.py:29-168): the boundary velocity at
 * step t is (A_k d(t), B_k d(t)) for boundary node k, coef = (A_k, B_k) in the order left (ny
 * nodes, j = 0..ny-1), right (ny), top (i = 1..nx-2), bottom (i = 1..nx-2), corners (0,0),
 * (nx-1,0), (0,ny-1), (nx-1,ny-1); ncoef = 2 ny + 2 (nx - 2) + 4.  decay[t] for t < nt (the
 * last entry holds beyond).  The host forms both with the reference's NumPy expressions
 * (iblbm/diagnostics/core.py:216-239), so the products reproduce its velocities. */
int iblbm_set_analytic_boundary(void *ctx, int id, const double *coef, int ncoef,
                                const double *decay, int nt);
/* Replaces the markers' desired velocity (Lb,2) only, keeping their positions and window
 * tables (iblbm/runtime/scenarios/taylor_green.py:21-32 updates it every step). */
int iblbm_set_desired_velocity(void *ctx, int id, const double *ud);

/* Runs nsteps steps of every instance (asynchronous).  diag_every > 0 takes the diagnostic
 * reductions on steps whose counter is a multiple of it. */
int iblbm_step(void *ctx, int nsteps, int diag_every);
int iblbm_sync(void *ctx);

/* Captures nsteps steps into a hipGraph and replays it (the step counter advances). */
int iblbm_graph_run(void *ctx, int nsteps, int diag_every);
/* Runs nsteps steps as ONE persistent cooperative launch of the fused step kernel, with a
 * grid-wide barrier between steps (replaces the per-step loop body of
 * iblbm/runtime/loop.py:59-87 like iblbm_step; no reference counterpart for the form).  Needs
 * every window on the fused path, one collision model, the tile row layout and no boundary
 * kernel (IBLBM_EINVAL otherwise); returns after the launch has finished (IBLBM_EDEVICE if a
 * grid barrier timed out).  Results equal iblbm_step's bit for bit. */
int iblbm_step_persistent(void *ctx, int nsteps, int diag_every);
/* on = 1 (default): a step runs as one fused launch when every marker window of the context
 * qualifies for it, otherwise the marker-window workgroups run on a second stream concurrently
 * with the bulk tiles of the same step; on = 0: window and bulk kernels run in order on one
 * stream. */
int iblbm_set_concurrent(void *ctx, int on);
/* How the next steps will be launched: IBLBM_MODE_FUSED (one launch per step),
 * IBLBM_MODE_TWO_STREAMS or IBLBM_MODE_SERIAL. */
#define IBLBM_MODE_AUTO (-1)
#define IBLBM_MODE_FUSED 0
#define IBLBM_MODE_TWO_STREAMS 1
#define IBLBM_MODE_SERIAL 2
int iblbm_step_mode(void *ctx, int *mode);

/* --- launch schedule and autotuner ---
 * No counterpart in the reference (its kernels use fixed 256-thread blocks,
 * iblbm/gpu_kernels.py:733-736).  A schedule only changes how the same per-node arithmetic
 * is distributed over workgroups, so every schedule gives bit-identical results. */
typedef struct {
    int32_t mode;         /* IBLBM_MODE_AUTO (fused where eligible) or an explicit mode */
    int32_t bulk_threads; /* bulk workgroup size of the two-kernel modes: 64..1024, x64;
                           * a workgroup covers a 64 x (bulk_threads / 64) tile of rows */
    int32_t bulk_rows;    /* rows a bulk wavefront keeps in flight: 1, 2 or 4; 3: one row
                           * at a time, software-pipelined (the next row's loads issued
                           * before the current row's collision).  3 and 4 apply to
                           * FP32-deviation storage (FP64 contexts take 2 for them; in the
                           * fused step 3 also needs the BGK or TRT collision) */
    int32_t bulk_waves;   /* 0: one row per wavefront (grid covers every row); > 0: that
                           * many wavefronts walk the rows with a grid stride */
    int32_t step_rows;    /* rows in flight of a bulk wavefront in the fused launch: 1 to
                           * 4, as bulk_rows */
    int32_t step_order;   /* fused launch: 0 rows by grid stride; 1 one block of adjacent rows
                           * per bulk wavefront, all wavefronts; k >= 2 blocks of k rows (the
                           * wavefronts past the last block idle, leaving memory bandwidth to
                           * the windows); at most 64; -1 one block per bulk workgroup, its
                           * wavefronts taking the next 1 or 2 rows (step_rows) from a counter
                           * in LDS; with step_rows = 3 one row at a time, the next row's
                           * populations copied into the workgroup's LDS while the current
                           * row computes (BGK and TRT contexts without the predictor; others
                           * take 2 rows) */
    int32_t row_layout;   /* 0: a bulk row is a 64-column tile of one lattice row; 1: a bulk
                           * row is 64 consecutive nodes in row-major node order (no part-
                           * empty tile at the end of a lattice row).  1 applies to contexts
                           * of FP32-deviation storage whose instances are all at least 64
                           * nodes wide; others keep tiles */
    int32_t store_mode;   /* two-kernel modes: 0: the bulk stores the new populations
                           * normally; 1: with non-temporal stores, which leave the last-level
                           * cache to the loads (pays once the populations outgrow it: about
                           * 256 MB on MI355X).  The fused step always stores normally */
    int32_t bulk_delay;   /* fused launch: the bulk wavefronts start this many 10 ns ticks
                           * (0..1000) after their workgroup, so that the windows' first
                           * loads meet an idle memory system; 0 starts them at once */
} iblbm_schedule;
int iblbm_get_schedule(void *ctx, iblbm_schedule *out);
int iblbm_set_schedule(void *ctx, const iblbm_schedule *in);
/* The row layout the context's instances use now (0 tiles, 1 flat): the schedule's, where
 * the context qualifies for it. */
int iblbm_row_layout_in_use(void *ctx, int *layout);
/* Times every candidate schedule for `steps` steps (after 2 warm-up steps) on the context's
 * current state, then the four fastest again for 4 x `steps` steps, twice in turn; restores
 * the populations and the step counter before each trial, and keeps the fastest of the
 * re-timed ones.  best_ms (may be NULL) receives its mean time per step over those two
 * trials; n_tried the number of candidates.  Refused while a learned-predictor mode is on (trial steps would record
 * samples).  The last step's diagnostics and MDF statistics afterwards are a trial's. */
int iblbm_autotune(void *ctx, int steps, double *best_ms, int *n_tried);

/* Device-side initial state for large ensembles: f = f_eq(rho, u) with
 * u = (ux (1 + amp xi0), uy (1 + amp xi1)), xi uniform in (-1, 1) per node and component from
 * a counter-based hash of (seed, node, component) -- a seeded synthetic field, not numpy's
 * stream (iblbm/init.py:116-265 forms the same equilibrium on the host).  Both population
 * buffers are written. */
int iblbm_init_flow(void *ctx, int id, double rho, double ux, double uy, double amp,
                    uint64_t seed);

int iblbm_get_mdf_stats(void *ctx, int id, iblbm_mdf_stats *out);
int iblbm_get_diag(void *ctx, int id, iblbm_diag *out);
/* Asynchronous diagnostics (no counterpart in the reference, whose loop.py:113-114 reads
 * them synchronously): iblbm_diag_fetch queues copies of the last diagnostic step's sums into
 * pinned host slot 0 or 1 behind the steps already queued; iblbm_get_diag_slot waits for that
 * slot only and returns what iblbm_get_diag would have returned then.  A host loop can queue
 * the next steps before reading a slot, so the device never waits for the host. */
int iblbm_diag_fetch(void *ctx, int slot);
int iblbm_get_diag_slot(void *ctx, int id, int slot, iblbm_diag *out);
/* Timing marks k = 0..3 on the context stream; elapsed(a, b) waits for b. */
int iblbm_event_record(void *ctx, int k);
int iblbm_event_elapsed(void *ctx, int a, int b, double *ms);
/* Window fib of the last step scattered into a zero (N,2) array. */
int iblbm_download_fib(void *ctx, int id, double *fib_aos);
/* Post-stream/boundary/IBM populations fstar, rho and U (velocity-corrected for DF/MDF) of
 * the last step, i.e. state.fstar, state.ro, state.U of the reference after that step. */
int iblbm_download_state(void *ctx, int id, double *fstar_aos, double *ro, double *U);
/* rho and U (velocity-corrected for DF/MDF) of the last step on the box [x0, x0 + w) x
 * [y0, y0 + h) inside the grid, box-major (ro: w*h, U: (w*h, 2)): the sub-arrays of state.ro
 * and state.U that the explicit Euler integrator's internal angular momentum reads around a
 * disc (iblbm/diagnostics/internal_flow.py:220-258, called from
 * iblbm/runtime/scenarios/sedimentation_single.py:198-206 and sedimentation_multi.py:148-156)
 * without downloading the whole grid. */
int iblbm_download_state_box(void *ctx, int id, int x0, int y0, int w, int h, double *ro,
                             double *U);
/* DFC marker forces q_k (Lb,2) of the last step (iblbm/ibm/dfc.py:232-253). */
int iblbm_download_dfc_force(void *ctx, int id, double *q);
/* fib of the last step on the window box only (box-major, bw*bh each) and the box
 * [x0, y0, bw, bh]; fib is zero outside it.  Replaces the per-step full-grid to_cpu(fib) of
 * iblbm/ibm/strategies/mdf.py:172-185 (force and torque extraction). */
int iblbm_download_window(void *ctx, int id, double *fibx, double *fiby, int box[4]);
int64_t iblbm_step_counter(void *ctx);

/* Time of the last iblbm_step/iblbm_graph_run measured with HIP events on the context
 * stream, in milliseconds, and the number of steps it covered. */
int iblbm_last_timing(void *ctx, double *ms, int *steps);

/* --- host-side helper (iblbm/ibm/a_norm.py:204-266) --- */
/* omega = 1/||A||_inf for markers in domain units; returns NaN on bad input. */
double iblbm_omega_estimate(const double *Lx, const double *Ly, int Lb, int delta,
                            double Larea, double dx);

/* --- operator-level entry points (iblbm/gpu_kernels.py wrappers; host arrays) --- */
/* feq_gpu (gpu_kernels.py:906): ro (N), U (N,2) -> feq (N,9) */
int iblbm_op_feq(const double *ro, const double *U, int64_t N, int incompressible, double *feq);
/* macroscopic_gpu (:878): fstar (N,9) -> ro (N), U (N,2) */
int iblbm_op_macroscopic(const double *fstar, int64_t N, int incompressible, double *ro,
                         double *U);
/* streaming_gpu (:846) with the CPU push semantics of streaming.py:38-77 */
int iblbm_op_streaming(const double *fstar_old, const double *f, int nx, int ny, double *out);
/* collision_gpu / collision_trt_gpu / collision_cm_mrt_gpu (:760, :787, :815).
 * p0 = tau (BGK, TRT) or unused (CM_MRT); p1 = tau_minus (TRT); s_diag for CM_MRT. */
int iblbm_op_collision(int model, const double *fstar, const double *feq, const double *U,
                       const double *fib, int64_t N, double tau, double tau_minus,
                       const double *s_diag, double dt, double *out);
/* apply_bc_dirichlet_neumann / apply_bc_settling_channel on a post-stream fstar, in place */
int iblbm_op_boundary(int bc, double *fstar, int nx, int ny, double in_u, double dens,
                      int incompressible);
/* ibm_direct_forcing (df.py:37-154): fib (N,2), Lux, Luy (Lb) */
int iblbm_op_direct_forcing(const double *Lx, const double *Ly, const double *ud, int Lb,
                            const double *Eux, const double *Euy, const double *Ro, double dx,
                            double dy, double dt, double Larea, int ny, int nx, int delta,
                            int incompressible, double *fib, double *Lux, double *Luy);
/* ibm_multi_direct_forcing (mdf.py:61-229) without rotation coupling */
int iblbm_op_multi_direct_forcing(const double *Lx, const double *Ly, const double *ud, int Lb,
                                  const double *U, const double *ro, double dx, double dy,
                                  double dt, double Larea, int ny, int nx, int n_iter,
                                  int min_iter, int delta, double omega, double tol,
                                  double *fib_total, iblbm_mdf_stats *stats);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the MDF
 * window of iblbm_op_multi_direct_forcing on caller-owned DEVICE arrays -- Lx, Ly (Lb),
 * ud (Lb, 2), U (N, 2) interleaved, ro (N) and the output fib_total (N, 2) -- with every
 * working array (descriptor, marker copies, window tables, touched-node fields, statistics)
 * carved from one fixed 6 MiB device array of the library per GPU, held until the process
 * ends; the call allocates nothing.  Limits: 1 <= Lb <= 512 and a window box (every clamped
 * stencil node) of at most 16384 nodes; a larger box, or a non-finite marker, fails with
 * IBLBM_EINVAL after the launch and leaves fib_total zero.  Work is ordered on `stream`
 * (a hipStream_t; NULL for the default stream) and the call waits for it before it returns
 * the statistics.  Calls that overlap on one GPU (several host threads or streams) share
 * that array and race. */
int iblbm_op_multi_direct_forcing_noalloc(const double *Lx, const double *Ly, const double *ud,
                                          int Lb, const double *U, const double *ro, double dx,
                                          double dy, double dt, double Larea, int ny, int nx,
                                          int n_iter, int min_iter, int delta, double omega,
                                          double tol, double *fib_total, iblbm_mdf_stats *stats,
                                          void *stream);
/* compute_lambda / compute_lambda_gpu (dfc.py:126-186, gpu_kernels.py:1091) */
int iblbm_op_dfc_lambda(const double *Lx, const double *Ly, int Lb, double rho_f, double Larea,
                        double dx, double dy, int ny, int nx, int delta, double *lambda_k);
/* apply_dfc_correction with a given lambda (dfc.py:256-342): fstar corrected in place */
int iblbm_op_dfc_correction(const double *Lx, const double *Ly, const double *ud, int Lb,
                            double *fstar, const double *lambda_k, double dx, double dy,
                            double Larea, int ny, int nx, int delta, double correction_scale,
                            double *force);

/* dfc_correction_gpu (gpu_kernels.py:1147-1223): the Eulerian correction field alone,
 * ef (N,9) = sum_k dF_k W(x - X_k) ds gathered from zero in np.add.at order (dfc.py:189-229),
 * zero outside the window, and the marker forces q (Lb,2) (:232-253); the caller adds it to
 * fstar (dfc.py:307-316).  lambda_k NULL: recomputed as compute_lambda does. */
int iblbm_op_dfc_correction_field(const double *Lx, const double *Ly, const double *ud, int Lb,
                                  const double *fstar, const double *lambda_k, double dx,
                                  double dy, double Larea, int ny, int nx, int delta, double *ef,
                                  double *force);
/* velocity_correction_gpu (gpu_kernels.py:930-953, dispatch.py:18-33): U (N,2) in place,
 * u + fib dt / (2 rho), or u + 0.5 fib dt when incompressible. */
int iblbm_op_velocity_correction(double *U, const double *fib, const double *ro, int64_t N,
                                 double dt, int incompressible);
/* ibm_interp_force_kernel (gpu_kernels.py:236-286, bound from ibm/forcing.py:57-63): per
 * marker the interpolated velocity Lux, Luy and the force Lf = 2 rho~ (ud - U~) / dt
 * (incompressible: 2 (ud - U~) / dt), in the CPU summation order of df.py:94-128. */
int iblbm_op_interp_force(const double *Lx, const double *Ly, const double *ud, int Lb,
                          const double *Eux, const double *Euy, const double *Ro, double dx,
                          double dy, double dt, int ny, int nx, int delta, int incompressible,
                          double *Lfx, double *Lfy, double *Lux, double *Luy);
/* ibm_spread_kernel (gpu_kernels.py:294-330, bound from ibm/forcing.py:157-164): adds the
 * spread of the marker forces, sum_k Lf_k W ds in np.add.at order (df.py:131-153), to the
 * (ny, nx) fields Efx, Efy in place (the reference passes zeroed fields). */
int iblbm_op_spread(const double *Lfx, const double *Lfy, const double *Lx, const double *Ly,
                    int Lb, double dx, double dy, double Larea, int ny, int nx, int delta,
                    double *Efx, double *Efy);

/* --- learned MDF force predictor ---
 * No counterpart in the reference: the MDF sub-step surrogate (an 8 -> 64 -> 64 -> 2 tanh MLP
 * per marker from (change of q / sig, direction of q, marker direction, sub-cell offset) to
 * the change of plain MDF's force before its last sub-step (since the previous converged
 * force) / sig, where q = R (ud - U) is the forcing slip before
 * forcing and sig the largest change of q over the markers of the marker's wavefront (64
 * markers); trained on the device on the plain MDF loop's own converged forces, BF16 MFMA
 * inference in the window workgroup).  After a prediction the step always applies at least
 * one true MDF correction and restarts plain MDF once if the divergence guard trips. */
#define IBLBM_PRED_OFF 0     /* plain MDF */
#define IBLBM_PRED_COLLECT 1 /* plain MDF, recording (features, scaled force change) */
#define IBLBM_PRED_WARM 2    /* previous step's force before its last sub-step, no network */
#define IBLBM_PRED_LEARNED 3 /* previous force + the network's predicted change */
#define IBLBM_MLP_IN 8
#define IBLBM_MLP_OUT 2
#define IBLBM_MLP_NP 4866   /* W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2] */
#define IBLBM_MLP_NORM 18   /* feature mean[8], feature scale[8], target scale[2] */
/* Sets the mode of a one-particle MDF instance of at most 512 markers (lb_capacity), whose
 * window runs in the fused step or in the window kernel of the two-kernel step (the same
 * arithmetic in both); the first non-zero mode allocates the predictor with room for
 * `capacity` training samples.  IBLBM_EINVAL for several particles or a larger capacity. */
int iblbm_predictor_mode(void *ctx, int id, int mode, int64_t capacity);
/* Training samples recorded so far (capped at the capacity). */
int iblbm_predictor_samples(void *ctx, int id, int64_t *n);
/* All samples written since the predictor was enabled (sample r sits in row r % capacity). */
int iblbm_predictor_samples_written(void *ctx, int id, int64_t *n);
/* Copies n samples (features[8], target[2] as float) starting at sample number `first`;
 * the target is (change of plain MDF's force before its last sub-step) / sig. */
int iblbm_predictor_data(void *ctx, int id, int64_t first, int64_t n, float *rows);
/* Normalisation from the samples, then `steps` Adam steps (minibatch 128, seed) on the device;
 * losses (may be NULL) receives the minibatch loss of every step. */
int iblbm_predictor_train(void *ctx, int id, int steps, double lr, uint32_t seed,
                          float *losses);
int iblbm_predictor_params(void *ctx, int id, float *params, float *norm);
int iblbm_predictor_set_params(void *ctx, int id, const float *params, const float *norm);
/* The previous step's marker forces before its last MDF sub-step (Lb, 2) followed by its
 * forcing slips q before forcing (Lb, 2), 4 Lb values: read (set = 0), or write and mark valid
 * (set = 1). */
int iblbm_predictor_previous_force(void *ctx, int id, double *F, int set);
/* Sums since the last reset: [true MDF sub-steps applied, steps, first-sub-step residual,
 * plain-MDF restarts]. */
int iblbm_predictor_stats(void *ctx, int id, double out[4], int reset);
/* Operator level: FP32 loss mean((y_hat - y)^2) and its gradient for nb <= 128 rows of
 * normalised features x (nb, 8) and targets y (nb, 2); BF16 MFMA inference of n rows. */
int iblbm_mlp_loss_grad(const float *params, const float *x, const float *y, int nb,
                        float *grad, float *loss);
int iblbm_mlp_infer_bf16(const float *params, const float *x, int n, float *out);

#ifdef __cplusplus
}
#endif
#endif /* IBLBM_HIP_H */
