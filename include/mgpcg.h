/*
 * mgpcg.h -- C-ABI of the MI355X-native MGPCG Poisson solver (libmgpcg.so).
 *
 * The boundary replaces the reference's in-process Python interface
 *   MGPCGPoissonSolver(dim, res, n_mg_levels, pre_and_post_smoothing,
 *                      bottom_smoothing, real)              mgpcg.py:6
 *   .reinitialize(cell_type, strategy)                      mgpcg.py:79-87
 *   .full_reinitialize(strategy)                            mgpcg.py:89-94
 *   .solve(max_iters, verbose, rel_tol, abs_tol, eps)       mgpcg.py:150-209
 *   result in .x                                            mgpcg.py:39
 * The strategies fill the solver's public fields b, grid_type[l], Adiag[l],
 * Ax[l] (pressure_project.py:41-70); here they fill caller-owned device
 * arrays in the same canonical layout (C-order over res >> l, Ax as
 * [*res_l, dim] array-of-structs, cell types int8 FLUID=0 AIR=1 SOLID=2) and
 * hand them to mgpcg_load_level / mgpcg_load_rhs.
 *
 * Conventions: every function returns 0 on success, -1 on an invalid
 * argument, -2 on a HIP runtime error; mgpcg_last_error() returns a
 * thread-local message.  All pointers to field data are device pointers.
 * A handle is not re-entrant; distinct handles may be used from distinct
 * threads.  Work is ordered after everything previously enqueued on the
 * stream given to mgpcg_set_stream (NULL = the legacy default stream) and that stream is
 * made to wait for the handle's work before the call returns.
 */
#ifndef MGPCG_H
#define MGPCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPCG_ABI_VERSION 1

/* precision modes */
#define MGPCG_PLAIN 0 /* FP32 everywhere, FP32 scalars: the reference's arithmetic (mgpcg.py:42-44) */
#define MGPCG_MIXED 1 /* FP32 V-cycle/Laplacian, FP64 x, dots, scalars, true-residual refinement */
#define MGPCG_F64 2   /* FP64 everywhere */

/* solve status */
#define MGPCG_RUNNING 0
#define MGPCG_CONVERGED 1
#define MGPCG_MAX_ITERS 2
#define MGPCG_BREAKDOWN 3 /* alpha or beta not finite */

typedef struct mgpcg_solver mgpcg_solver_t;

typedef struct {
  int32_t iters;   /* completed non-final loop bodies (mgpcg.py:176-209) */
  int32_t status;  /* MGPCG_CONVERGED / MGPCG_MAX_ITERS / MGPCG_BREAKDOWN */
  double r0r0;     /* initial r.r over FLUID cells (mgpcg.py:158-159) */
  double rr;       /* last r.r tested against tol (mgpcg.py:188-194) */
  double true_rr;  /* last recomputed |b - A x|^2 (MIXED), -1 if none */
  double tol;      /* max(abs_tol, rel_tol * r0r0) (mgpcg.py:164) */
  double alpha, beta;
} mgpcg_result_t;

/* Replaces MGPCGPoissonSolver.__init__ (mgpcg.py:6-44): allocates the whole
 * hierarchy (levels res >> l, l < n_levels) in device memory.
 * n_mg_levels, pre_and_post_smoothing, bottom_smoothing as in mgpcg.py:6;
 * every res[a] must be divisible by 2^(n_levels-1) (mgpcg.py:23). */
int mgpcg_create(int dim, const int32_t* res, int n_levels, int pre_and_post_smoothing,
                 int bottom_smoothing, int precision, mgpcg_solver_t** out);
int mgpcg_destroy(mgpcg_solver_t* solver);

/* Orders the handle's work after / before the caller's stream (hipStream_t). */
int mgpcg_set_stream(mgpcg_solver_t* solver, void* stream);

/* Process-wide tuning knobs (no reference counterpart; `solver` is unused). */
#define MGPCG_OPT_COARSE_MAX_SLOTS 3 /* levels with at most this many slots per colour are
                                        solved by one workgroup per system (default 4096);
                                        applies to handles created afterwards */
#define MGPCG_OPT_MAX_GRID 4         /* workgroups of a persistent stage launch (default 2048) */
#define MGPCG_OPT_FUSE_FIRST 5    /* fill(0)+smooth(0)+smooth(1) in one kernel: 1 on every
                                     level, 2 on the levels below level 0 (default 0) */
#define MGPCG_OPT_FUSE_PROLONG 6  /* 1: prolongate inside the next smooth(1) (default 1) */
#define MGPCG_OPT_FUSE_P 7        /* 1: update_p inside compute_Ap (default 0) */
#define MGPCG_OPT_COMPACT 8       /* 1: levels whose nonzero off-diagonals are all one value
                                     read per-cell coupling flags instead of Ax (default 1);
                                     applies to systems solved or submitted afterwards */
#define MGPCG_OPT_COARSE_LDS 9    /* 1: the coarse-level kernel runs on a copy of its levels'
                                     cell types, flags, Adiag, r and z in LDS (160 KB) when
                                     they fit (FP32 hierarchies; default 1) */
#define MGPCG_OPT_RESTRICT_ROWS 10 /* restriction kernels: 2 a residual pass into scratch
                                      then a child gather (default), 1 tiled by coarse
                                      rows with the residuals in LDS (fine rows up to 512
                                      cells), 0 one lane per child */
#define MGPCG_OPT_PROBE_EVERY 11  /* the launch probe times one eligible launch in this
                                     many (default 4) */
#define MGPCG_OPT_SERVER_DEPTH 12 /* iterations a server keeps queued on the device ahead of
                                     the host, 1..3 (default 2) */
#define MGPCG_OPT_SERVER_STREAMS 13 /* engines (each with its own stream) a server created
                                       afterwards spreads its members over, 1..4 (default 2) */
#define MGPCG_OPT_GRAPH 14        /* 1: an engine captures one iteration of its current member
                                     set into a hipGraph on the set's second iteration and
                                     replays it until the set or an option changes (default
                                     0); the launch probe and mgpcg_solve_hostsync launch
                                     kernel by kernel */
#define MGPCG_OPT_RESTRICT_COARSE 15 /* restriction kernels on the levels below level 0, the
                                        modes of MGPCG_OPT_RESTRICT_ROWS; -1 (default): the
                                        MGPCG_OPT_RESTRICT_ROWS mode */
int mgpcg_set_option(mgpcg_solver_t* solver, int option, int value);
/* The current value of a knob (no reference counterpart; `solver` is unused). */
int mgpcg_get_option(const mgpcg_solver_t* solver, int option, int* value);

/* Padded layout of a level: out6 = {P0, P1, W, slots per colour, cells,
 * first coarse level}; bytes = device bytes held by the handle. */
int mgpcg_info(const mgpcg_solver_t* solver, int level, int32_t* out6, int64_t* bytes);
/* Device bytes mgpcg_create would allocate for this shape and precision (the
 * handle's pool), without allocating: what an HBM-budget admission charges.
 * No reference counterpart (the reference allocates Taichi fields). */
int mgpcg_pool_bytes(int dim, const int32_t* res, int n_levels, int precision, int64_t* bytes);

/* Loads one level of the operator from the canonical fields a strategy's
 * build_A wrote (pressure_project.py:45-64 and siblings) and zeroes that
 * level's r and z.  Called for l = 0..n_levels-1 after the strategy ran;
 * together with mgpcg_load_rhs this is the tail of reinitialize /
 * full_reinitialize (mgpcg.py:79-94). */
int mgpcg_load_level(mgpcg_solver_t* solver, int level, const int8_t* grid_type,
                     const float* Adiag, const float* Ax);
/* PressureProjectStrategy's reinitialize (pressure_project.py:15-64 via
 * mgpcg.py:79-87) straight into the device hierarchy: b and every level's
 * operator are computed in the packed layout, bit for bit what
 * mgpcg_pressure_build_b / _A, mgpcg_coarsen, mgpcg_load_level and
 * mgpcg_load_rhs produce, without the canonical b, Adiag and Ax.
 * grid_type[l]: canonical int8 cell types per level; level 0 is read,
 * levels >= 1 are written (coarsened, mgpcg.py:46-60).  velocity[k]: MAC
 * component k (res + e_k); phi: the level set (ghost-fluid weights at level 0
 * when ghost_fluid_method).  add_c != 0 adds f32(c) to b on FLUID cells
 * (PressureProjectWithVolumeControlStrategy, volume_control.py:33-37).
 * Ordered on the handle's stream. */
int mgpcg_assemble_pressure(mgpcg_solver_t* solver, int8_t* const* grid_type,
                            const float* const* velocity, const float* phi, double scale_A,
                            double scale_b, double p0, int ghost_fluid_method, double c_max,
                            int add_c, double c);
/* Whether every nonzero off-diagonal coefficient of a loaded level has one
 * value (then *coef is it, or 0 when the level has no coupling); waits for
 * the handle's stream.  No reference counterpart: inspection only. */
int mgpcg_level_coupling(mgpcg_solver_t* solver, int level, int* uniform, float* coef);
/* Copies a loaded level's per-slot flag words to flags (device memory, 2 *
 * slots-per-colour uint16: colour 0 then colour 1) and its per-block FLUID
 * flags to wet (device memory, ceil(slots / 256) bytes), in the padded
 * colour-split layout (mgpcg_info).  A flag word holds the nonzero-coupling
 * bits (upper neighbour on axis a: bit a, lower: bit 3 + a) and in bits 6..8
 * the diagonal code: n < 7 when Adiag equals n couplings of the slot's
 * magnitude added in order from 0, 7 when it is read from Adiag.  Waits for
 * the handle's stream.  No reference counterpart: inspection only. */
int mgpcg_level_flags(mgpcg_solver_t* solver, int level, uint16_t* flags, uint8_t* wet);
/* Copies a loaded level's work orders (device memory, int32): ord1 gets one
 * entry per block of 256 slots (natural order), ord2 one per half-sweep
 * virtual block of 512 slots (a permutation with the FLUID-holding ones
 * spread evenly over the positions; only level 0, when the batched kernels
 * solve it, has one: ord2 is left alone for the others and may be NULL).  Entry k
 * names the block (low 31 bits) that the k-th virtual block of a system's
 * stage processes; bit 31 is set when that block holds a FLUID slot.  Waits
 * for the handle's stream.  No reference counterpart: inspection only. */
int mgpcg_level_order(mgpcg_solver_t* solver, int level, int32_t* ord1, int32_t* ord2);
/* Loads b (the strategy's build_b output, mgpcg.py:20). */
int mgpcg_load_rhs(mgpcg_solver_t* solver, const float* b);

/* init_gridtype (mgpcg.py:46-60) on canonical int8 arrays, asynchronous on
 * `stream`: coarse = AIR if any child is AIR, else FLUID if any child is
 * FLUID, else SOLID.  fine_res has dim entries. */
int mgpcg_coarsen(void* stream, int dim, const int32_t* fine_res, const int8_t* fine,
                  int8_t* coarse);

/* solve (mgpcg.py:150-209) with the loop, the reductions and every scalar
 * resident on the device.  max_iters = -1 means no cap.  refine_every is the
 * true-residual period of MGPCG_MIXED (ignored otherwise).  Blocks until the
 * solve finished; the result stays in the handle (mgpcg_store_x). */
int mgpcg_solve(mgpcg_solver_t* solver, int max_iters, double rel_tol, double abs_tol,
                double eps, int refine_every, mgpcg_result_t* out);
/* The same solve driven like the reference's host loop: one launch per
 * kernel and a device->host read after every reduction (mgpcg.py:159, 172,
 * 180, 189, 202).  The plain-GPU baseline. */
int mgpcg_solve_hostsync(mgpcg_solver_t* solver, int max_iters, double rel_tol, double abs_tol,
                         double eps, int refine_every, mgpcg_result_t* out);

/* The plain FP32 GPU port of mgpcg.py structured like the reference: one
 * kernel launch per reference kernel (fills, smooth, restrict, prolongate,
 * compute_Ap, reduce, update_xr, update_p), alpha / beta / the residual test
 * on the host, and a device->host read of every reduced scalar.  MGPCG_PLAIN
 * handles only.  It is the baseline the batched solver is measured against;
 * its x is bit-identical to mgpcg_solve's in MGPCG_PLAIN. */
int mgpcg_solve_reference_structured(mgpcg_solver_t* solver, int max_iters, double rel_tol,
                                     double abs_tol, double eps, mgpcg_result_t* out);

/* Copies x (mgpcg.py:39) into a canonical array; exactly one of the two
 * pointers is non-NULL. */
int mgpcg_store_x(mgpcg_solver_t* solver, float* x_f32, double* x_f64);

/* ---------------------------------------------------------------------------
 * Batched solves (no reference counterpart: the reference solves one system
 * at a time).  Every stage of an iteration is one persistent kernel launch
 * over all systems of the batch, which may have different resolutions but
 * share dim, precision, n_levels and the smoothing counts.  results[i]
 * belongs to systems[i]; each system's x is as after mgpcg_solve.
 * ------------------------------------------------------------------------- */
typedef struct mgpcg_batch mgpcg_batch_t;
int mgpcg_batch_create(mgpcg_batch_t** out);
int mgpcg_batch_destroy(mgpcg_batch_t* batch);
int mgpcg_batch_set_stream(mgpcg_batch_t* batch, void* stream);
int mgpcg_batch_solve(mgpcg_batch_t* batch, mgpcg_solver_t* const* systems, int n,
                      int max_iters, double rel_tol, double abs_tol, double eps,
                      int refine_every, mgpcg_result_t* results);

/* ---------------------------------------------------------------------------
 * Server: continuous batching.  mgpcg_server_submit queues an assembled
 * system: its assembly is the work enqueued on the system's stream
 * (mgpcg_set_stream) before the call, and systems may be assembled on
 * different streams.  Each mgpcg_server_step admits queued systems at the
 * iteration boundary -- while others are running, those whose assembly has
 * finished on the device; when none is running, all of them -- enqueues one
 * iteration for every member and retires the members whose state (read one
 * iteration behind) says they are done; their results are returned by
 * mgpcg_server_poll with the submit tag.  A system must not be modified
 * between submit and its result.  Admission policy (HBM budget) is the
 * caller's: it decides when to create and submit systems.
 * ------------------------------------------------------------------------- */
typedef struct mgpcg_server mgpcg_server_t;
int mgpcg_server_create(int max_iters, double rel_tol, double abs_tol, double eps,
                        int refine_every, mgpcg_server_t** out);
int mgpcg_server_destroy(mgpcg_server_t* server);
int mgpcg_server_set_stream(mgpcg_server_t* server, void* stream);
int mgpcg_server_submit(mgpcg_server_t* server, mgpcg_solver_t* system, int64_t tag);
int mgpcg_server_step(mgpcg_server_t* server, int* n_active);
int mgpcg_server_poll(mgpcg_server_t* server, int64_t* tags, mgpcg_result_t* results, int cap,
                      int* n_out);
int mgpcg_server_pending(const mgpcg_server_t* server, int* n_pending, int* n_active,
                         int* inflight);

/* Measurement (no reference counterpart): enable = 1 starts timing one in
 * MGPCG_OPT_PROBE_EVERY launches of the level-0 red-black half-sweep
 * smooth(0, 0) on the device; a timed launch runs as its own instantiation,
 * mg::k_stage<..., 0, 2> (the others are mg::k_stage<..., 0, 1>), so a
 * profiler's statistics of that name cover exactly the timed launches, and
 * records the device clock when its first workgroup starts and when its last
 * one ends.  enable = 0 stops and returns the number of timed launches, their
 * summed device time, and the level-0 cells and FLUID cells those launches
 * swept (counted on the device: systems already finished are not swept).  A
 * server sums over its engine streams.  Output pointers may be NULL. */
int mgpcg_server_probe(mgpcg_server_t* server, int enable, int* n_launches, double* total_ms,
                       int64_t* total_cells, int64_t* total_fluid);
int mgpcg_batch_probe(mgpcg_batch_t* batch, int enable, int* n_launches, double* total_ms,
                      int64_t* total_cells, int64_t* total_fluid);
/* Measurement (no reference counterpart): iteration graphs captured and
 * iterations launched from a graph so far by a solver's own engine, a batch
 * or a server (summed over its engine streams).  Exactly one handle is given;
 * the other two are NULL. */
int mgpcg_graph_counts(const mgpcg_solver_t* solver, const mgpcg_batch_t* batch,
                       const mgpcg_server_t* server, int64_t* captures, int64_t* launches);

/* Kernel-level access for parity tests (no reference counterpart): move a
 * level vector between canonical layout (float, or double for MGPCG_F64 and
 * for x outside MGPCG_PLAIN) and the handle, and run one stage. */
#define MGPCG_FIELD_R 0
#define MGPCG_FIELD_Z 1
#define MGPCG_FIELD_P 2      /* p buffer of even iterations */
#define MGPCG_FIELD_AP 3
#define MGPCG_FIELD_X 4
#define MGPCG_FIELD_P_NEXT 5 /* p buffer of odd iterations */
#define MGPCG_OP_SMOOTH 0                /* smooth(level, arg) (mgpcg.py:104-109) */
#define MGPCG_OP_SMOOTH_PAIR_FROM_ZERO 1 /* z.fill(0); smooth(level,0); smooth(level,1) */
#define MGPCG_OP_RESTRICT 2              /* restrict(level) into r[level+1] (mgpcg.py:111-118) */
#define MGPCG_OP_PROLONGATE 3            /* prolongate(level) (mgpcg.py:120-123) */
#define MGPCG_OP_VCYCLE 4                /* v_cycle() (mgpcg.py:125-148) */
#define MGPCG_OP_COMPUTE_AP 5 /* p_next = z + 0 * p; Ap = A p_next (mgpcg.py:234-238, 218-224) */
#define MGPCG_OP_PROLONG_SMOOTH1 6 /* prolongate(level); smooth(level, 1), FLUID cells */
int mgpcg_debug_set(mgpcg_solver_t* solver, int field, int level, const void* canon);
int mgpcg_debug_get(mgpcg_solver_t* solver, int field, int level, void* canon);
int mgpcg_debug_op(mgpcg_solver_t* solver, int op, int level, int arg);
/* Average device time of one launch of a stage (SMOOTH, RESTRICT, PROLONGATE
 * or VCYCLE), measured with HIP events around `reps` back-to-back launches
 * on the handle's stream (no reference counterpart: measurement only). */
int mgpcg_time_op(mgpcg_solver_t* solver, int op, int level, int arg, int reps,
                  double* ms_per_launch);

/* ---------------------------------------------------------------------------
 * Operator assembly on the device (the strategies' Taichi kernels).  All
 * fields are canonical C-order float32 device arrays; MAC components
 * velocity[k], H[k], H_ext[k] have shape res + e_k; n is [*res, dim].
 * Every kernel writes every cell of its outputs (the reference relies on
 * initialize() zeroing them, mgpcg.py:63-77).  Asynchronous on `stream`.
 * ------------------------------------------------------------------------- */
/* PressureProjectStrategy.build_b_kernel (pressure_project.py:15-39);
 * c_max is the ghost-fluid clamp (1e3 in the reference, :37). */
int mgpcg_pressure_build_b(void* stream, int dim, const int32_t* res, const int8_t* cell_type,
                           const float* const* velocity, const float* phi, double scale_A,
                           double scale_b, double p0, int ghost_fluid_method, double c_max,
                           float* b);
/* PressureProjectStrategy.build_A_kernel (pressure_project.py:45-64) on level
 * `level` (grid_type, Adiag, Ax have shape res >> level; phi is level 0). */
int mgpcg_pressure_build_A(void* stream, int dim, const int32_t* res, int level,
                           const int8_t* grid_type, const float* phi, double scale_A,
                           int ghost_fluid_method, double c_max, float* Adiag, float* Ax);
/* MagneticForceStrategy.build_b_kernel (magnetic_force.py:17-39). */
int mgpcg_magnetic_build_b(void* stream, int dim, const int32_t* res, const int8_t* cell_type,
                           const float* const* velocity, const float* const* H, double scale_A,
                           double scale_b, double p0, double mu0, double k, float* b);
/* calc_volume (volume_control.py:22-31): sum over cells, FP64, written to
 * *out_device; partials needs ceil(cells / 256) doubles of device memory. */
int mgpcg_calc_volume(void* stream, int dim, const int32_t* res, const float* phi, double dx,
                      double* partials, int partials_cap, double* out_device);
/* No counterpart in the reference: mgpcg_calc_volume without a partials
 * buffer and without any allocation, for callers that cannot allocate
 * (inside HIP graph capture, tight serving loops).  The per-block partial
 * sums live in a file-scope device array of 2^18 doubles (2 MiB per GPU, held
 * until the process ends), so grids of at most 2^26 cells; larger grids are
 * rejected.  Calls that overlap on one GPU share that array and race. */
int mgpcg_calc_volume_static(void* stream, int dim, const int32_t* res, const float* phi,
                             double dx, double* out_device);
/* add_c (volume_control.py:33-37): b[FLUID] += f32(c). */
int mgpcg_add_c(void* stream, int dim, const int32_t* res, const int8_t* cell_type, double c,
                float* b);
/* PotentialFunctionStrategy (potential_function.py:19-34 / 40-55). */
int mgpcg_potential_build_b(void* stream, int dim, const int32_t* res, const float* chi,
                            const float* const* H_ext, float* b);
int mgpcg_potential_build_A(void* stream, int dim, const int32_t* res, int level, double dx,
                            const float* chi, float* Adiag, float* Ax);
/* SurfaceTension.calc_n / calc_kappa (surface_tension.py:115-136) and
 * SurfaceTensionStrategy (surface_tension.py:23-66 / 71-89) for axis d. */
int mgpcg_calc_n(void* stream, int dim, const int32_t* res, int d, const float* phi, double dx,
                 float* n);
int mgpcg_calc_kappa(void* stream, int dim, const int32_t* res, const float* n, double dx,
                     float* kappa);
int mgpcg_surface_build_b(void* stream, int dim, const int32_t* res, int d,
                          const float* velocity_d, const float* phi, const float* kappa,
                          const float* n, double dt, double dx, double sigma, float* b);
int mgpcg_surface_build_A(void* stream, int dim, const int32_t* res, int level, int d,
                          const float* phi, double dt, double dx, double sigma, float* Adiag,
                          float* Ax);

/* ---------------------------------------------------------------------------
 * The step after the solve: the pressure gradient applied to the MAC velocity
 * (in place; velocity[k] has shape res + e_k, pressure is the solve's x).
 * ------------------------------------------------------------------------- */
/* FluidSimulator.apply_pressure (fluid_simulator.py:192-214); c_max is the
 * ghost-fluid clamp (1e3 in the reference, :205/:211). */
int mgpcg_apply_pressure(void* stream, int dim, const int32_t* res, const int8_t* cell_type,
                         const float* phi, const float* pressure, float* const* velocity,
                         double dt, double rho, double dx, double p0, int ghost_fluid_method,
                         double c_max);
/* FerrofluidSimulator.apply_magnetic_force (ferrofluid_simulator.py:114-137). */
int mgpcg_apply_magnetic_force(void* stream, int dim, const int32_t* res,
                               const int8_t* cell_type, const float* pressure,
                               const float* const* H, float* const* velocity, double dt,
                               double rho, double dx, double p0, double mu0, double k);

/* ---------------------------------------------------------------------------
 * The step before the solve: cell classification and redistancing.
 * ------------------------------------------------------------------------- */
/* FluidSimulator.apply_markers (fluid_simulator.py:151-159): non-SOLID cells
 * become FLUID where phi <= 0, AIR elsewhere (cell_type in place). */
int mgpcg_apply_markers(void* stream, int dim, const int32_t* res, const float* phi,
                        int8_t* cell_type);
/* FluidSimulator.init_boundary (fluid_simulator.py:337-341): SOLID border. */
int mgpcg_init_boundary(void* stream, int dim, const int32_t* res, int8_t* cell_type);
/* FastSweepingLevelSet.redistance (level_set.py:260-333), phi in place;
 * phi_temp, snap (floats) and fixed (bytes) are caller-provided work arrays
 * of the grid's size.  Neighbouring lines are read as they were before each
 * sweep (deterministic; the reference's parallel lines race). */
int mgpcg_redistance(void* stream, int dim, const int32_t* res, double dx, int repeat_times,
                     float* phi, float* phi_temp, float* snap, int8_t* fixed);

/* FluidSimulator.extrap_velocity (fluid_simulator.py:240-270), velocity in
 * place: per component, faces next to a FLUID cell are valid; `iterations`
 * (10 in the reference) passes average the valid faces of each invalid face's
 * 3^dim neighbourhood into it.  backup (floats), valid and valid_temp (bytes)
 * are work arrays of at least the largest component's size.  Faces outside a
 * component's shape count as invalid. */
int mgpcg_extrap_velocity(void* stream, int dim, const int32_t* res, const int8_t* cell_type,
                          float* const* velocity, int iterations, float* backup, int8_t* valid,
                          int8_t* valid_temp);

const char* mgpcg_last_error(void);
int mgpcg_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MGPCG_H */
