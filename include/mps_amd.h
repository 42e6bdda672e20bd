/*
 * mps_amd.h — C ABI of the MI355X MPS time-step library (libmps_amd.so).
 *
 * The library owns all device memory, one HIP stream per handle and the
 * captured graphs; the caller owns every host buffer and the library copies in
 * and out.  Every entry point returns 0 on success and a non-zero status
 * (MPS_E_*) on failure; mps_last_error() then returns a message.  No C++
 * exception crosses this boundary.  One handle is driven by one host thread.
 *
 * Particle arrays are indexed by particle id and ids must be 0..N-1 in order,
 * exactly as the reference requires (src/mps.cpp:192 indexes particles[id],
 * src/mps.cpp:290 indexes sourceTerm[pi.id]).  Vectors are 3 doubles per
 * particle (x,y,z), as the reference's Eigen::Vector3d (src/particle.hpp:27).
 *
 * Enumerations keep the reference's integer values (src/particle.hpp:5,19):
 *   type: Ghost=-1, Fluid=0, Wall=1, DummyWall=2
 *   boundary condition: Ignored=0, Surface=1, Inner=2
 */
#ifndef MPS_AMD_H
#define MPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPS_OK 0
#define MPS_E_INVALID 1   /* bad argument or input state */
#define MPS_E_HIP 2       /* HIP runtime error */
#define MPS_E_NOMEM 3     /* device allocation failed */
#define MPS_E_RANGE 4     /* a particle fell outside the bucket grid */
#define MPS_E_SOLVER 5    /* the step completed, but a PPE solve stopped above its
                             tolerance (the true FP64 relative residual); the
                             reference ignores its solver's status
                             (src/mps.cpp:457-460) and so may the caller */
/* The reference's asymmetric-matrix exit (src/mps.cpp:418-443) has no status:
 * the matrix is assembled symmetric by construction (one density per scene,
 * the same distance both ways). */

#define MPS_GHOST (-1)
#define MPS_FLUID 0
#define MPS_WALL 1
#define MPS_DUMMY_WALL 2

#define MPS_BC_IGNORED 0
#define MPS_BC_SURFACE 1
#define MPS_BC_INNER 2

/* POD mirror of Settings (src/settings.hpp:35-86) after Settings::load
 * (src/settings.cpp:21-84): radii and the collision distance are already
 * multiplied by the particle distance, re_max = max(re_pressure, re_viscosity,
 * re_surface) (src/settings.cpp:56-61). */
typedef struct mps_settings {
    int32_t dim;               /* 2 or 3 */
    int32_t qc_on;             /* pressure stabilization: quasi compressibility on */
    int32_t hos_on;            /* pressure stabilization: higher order source term on */
    int32_t reserved0;
    double particle_distance;  /* l0 */
    double dt;
    double finish_time;
    double output_interval;
    double cfl_condition;
    double domain_min[3];      /* domain range X/Y/Z [min, max] */
    double domain_max[3];
    double density;
    double kinematic_viscosity;
    double gravity[3];
    double re_pressure;
    double re_viscosity;
    double re_surface;
    double re_max;
    double surface_threshold;  /* threshold for surface detection (beta) */
    double compressibility;    /* quasi compressibility alpha */
    double hos_gamma;
    double relaxation;         /* relaxation coefficient for pressure */
    double collision_distance; /* collision distance ratio * l0 */
    double restitution;        /* coefficient of restitution */
} mps_settings;

/* Pressure-solver controls.  No counterpart in the reference, whose solver is
 * Eigen's ConjugateGradient with fixed defaults (src/mps.cpp:452-460). */
#define MPS_PREC_FP64 0        /* FP64 Jacobi-PCG */
#define MPS_PREC_MIXED 1       /* FP32 inner Jacobi-PCG + FP64 iterative refinement */
typedef struct mps_solver_options {
    int32_t precision;         /* MPS_PREC_* */
    int32_t warm_start;        /* 1: start from the previous step's unclamped solution;
                                  2: from the line through the last two (where both are nonzero) */
    int32_t max_iterations;    /* cap on total CG iterations per solve; <=0: 2*rows */
    int32_t max_refinements;   /* mixed: cap on outer refinement steps; <=0: 20 */
    double tolerance;          /* stop when ||b-Ax||_2 <= tolerance*||b||_2 */
    double inner_tolerance;    /* mixed: relative reduction asked of each inner solve */
    int32_t check_interval;    /* CG iterations between host convergence checks (>=1) */
    int32_t use_graph;         /* 1: replay captured HIP graphs for the CG chunks */
    int32_t preconditioner;    /* MPS_PRECOND_*: JACOBI is Eigen's default; MULTIGRID
                                  changes the iterations, not the tolerance met */
    int32_t reserved;          /* 0 */
} mps_solver_options;

/* Preconditioners of the pressure CG. */
#define MPS_PRECOND_JACOBI 0    /* diagonal (Eigen's DiagonalPreconditioner) */
#define MPS_PRECOND_MULTIGRID 1 /* aggregation multigrid V-cycle on the bucket
                                   cells, one hierarchy per scene; systems
                                   under 8192 rows use JACOBI */

/* Statistics of the last pressure solve. */
typedef struct mps_solver_stats {
    int32_t rows;              /* Inner rows in the system */
    int32_t iterations;        /* CG iterations (mixed: inner iterations summed) */
    int32_t refinements;       /* mixed: outer refinement steps (FP64 residuals);
                                  FP64: restarts from the iterate after its true
                                  residual missed the tolerance */
    int32_t converged;         /* 1 if the true relative residual met the tolerance */
    double rhs_norm;           /* ||b||_2 */
    double rel_residual;       /* ||b-Ax||_2 / ||b||_2, true FP64 residual */
    double initial_residual;   /* ||b-Ax0||_2 / ||b||_2 of the starting guess */
} mps_solver_stats;

typedef struct mps_handle mps_handle;

/* Human-readable message of the last failure on h (or of the last failed
 * mps_create when h is NULL). Never NULL. */
const char* mps_last_error(const mps_handle* h);

/* Replaces MPS::MPS(path) (src/mps.cpp:18-36) plus the particle import of
 * MPS::loadInitialState (src/mps.cpp:96-102): takes settings and the initial
 * particle table, computes the reference values n0/lambda (src/mps.cpp:132-163),
 * sizes the bucket grid (src/bucket.cpp:10-22) and uploads the state.
 * pos/vel are N x 3 doubles; type holds MPS_* type codes. Pressure starts at 0
 * and the boundary condition at Ignored (the reference leaves them
 * uninitialised, src/particle.hpp:28,34). */
int mps_create(const mps_settings* s, int32_t n, const int32_t* id,
               const int8_t* type, const double* pos, const double* vel,
               mps_handle** out);
int mps_destroy(mps_handle* h);

/* Ensembles (no counterpart in the reference, which runs one scene per
 * process): nscene independent scenes with their own settings (dt, l0,
 * viscosity, domain, ...) in one handle, stepped together.  Particles are
 * packed scene after scene, counts[k] per scene, with local ids 0..n_k-1 in
 * order; every other entry point addresses them by global index (scene
 * offset + local id).  Each scene keeps its own bucket grid (a key range of
 * the shared cell table) and its own CG scalars and convergence test, so a
 * scene's results are bit-identical to the same scene run alone.  All scenes
 * share the solver options and must share dim. */
int mps_create_ensemble(int32_t nscene, const mps_settings* s, const int32_t* counts,
                        const int32_t* id, const int8_t* type, const double* pos,
                        const double* vel, mps_handle** out);
int mps_scene_count(const mps_handle* h);
/* Per-scene statistics of the last step: solver stats and Courant number
 * (either pointer may be NULL). mps_step's Courant is their maximum, and
 * mps_get_solver_stats sums iterations and takes the worst residual. */
int mps_ensemble_stats(const mps_handle* h, mps_solver_stats* stats, double* courant);

int mps_set_solver(mps_handle* h, const mps_solver_options* o);
int mps_get_solver_stats(const mps_handle* h, mps_solver_stats* out);

/* Replaces MPS::loadInitialState (src/mps.cpp:96-109): the initial
 * setNumberDensity (src/mps.cpp:627-641). */
int mps_load_initial_state(mps_handle* h);

/* Replaces MPS::stepForward (src/mps.cpp:111-130). Writes the Courant number
 * (src/mps.cpp:547-571) to *courant when courant is not NULL.  Returns
 * MPS_E_SOLVER after a completed step whose pressure solve stopped above its
 * tolerance (iteration or refinement cap, or a tolerance below what FP64
 * rounding of A x allows); the step used the solution as it stood. */
int mps_step(mps_handle* h, int32_t is_export, double* courant);

/* Runs n_steps steps with no export and no host synchronisation between
 * them; writes each step's Courant number to courants[n_steps] if not NULL.
 * Every step runs; MPS_E_SOLVER names the first whose solve missed. */
int mps_run_steps(mps_handle* h, int32_t n_steps, double* courants);

/* Copies the particle state out in id order (the fields Exporter reads,
 * src/exporter.cpp:35-51,96-136). Any pointer may be NULL. */
int mps_download(const mps_handle* h, int8_t* type, double* pos, double* vel,
                 double* pressure, double* ndr, int8_t* bc);

/* Minimum pressure per particle (src/mps.cpp:481-503), id order. */
int mps_download_min_pressure(const mps_handle* h, double* pmin);

/* Overwrites the particle state (id order); used for restarts from a
 * lossless checkpoint. Any pointer may be NULL (field unchanged). */
int mps_upload(mps_handle* h, const int8_t* type, const double* pos,
               const double* vel, const double* pressure, const double* ndr,
               const int8_t* bc);

/* Debug-level events the reference logs with spdlog::debug and counts in
 * debugLogCount (src/mps.cpp:405-410, 559-564, 655-663), accumulated since
 * mps_create or the last mps_clear_debug_events:
 *   counts[0] = particles ghosted (out of domain),
 *   counts[1] = Inner particles with no Dirichlet connection,
 *   counts[2] = steps whose Courant number exceeded the CFL condition.
 * ghost_ids / unconnected_ids receive up to cap ids each, ascending. */
int mps_debug_events(const mps_handle* h, int64_t counts[3],
                     int32_t* ghost_ids, int32_t* unconnected_ids, int32_t cap);
int mps_clear_debug_events(mps_handle* h);

/* ---- test and inspection entry points ---- */

/* Internal counters (no counterpart in the reference; for tests):
 *   out[0] = times a solver's captured CG graphs were dropped because a
 *            pointer or launch shape they bake in changed,
 *   out[1] = CG vector reallocations,
 *   out[2] = pressure-matrix storage reallocations (entries or 3D anchors),
 *   out[3..6] = blocks of the LDS-staged neighbour walks on this device since
 *            the process started, counted only for handles created with
 *            MPS_AMD_STAGE_STATS=1: all of them, those that staged every
 *            stencil row, those that staged some rows but not all, and of
 *            those the ones that left a row out before a later staged row,
 *   out[7] = CG iterations enqueued by the handle's solves, counting those
 *            of a chunk that return at once after convergence,
 *   out[8] = CG chunk graphs captured (one per solver and chunk length while
 *            the pointers they bake in stay; out[0] counts the drops).
 * Writes min(cap, 9) values. */
int mps_debug_counters(const mps_handle* h, int64_t* out, int32_t cap);

/* Bucket grid geometry (src/bucket.hpp:11-14). */
int mps_bucket_shape(const mps_handle* h, int32_t dims[4] /* num,numX,numY,numZ */,
                     double* length);
/* Rebuilds the bucket grid from the current positions
 * (Bucket::storeParticles, src/bucket.cpp:24-51) and writes the reference's
 * linked-list view of it: first/last per cell (num entries), next per
 * particle (N entries), -1 for none. */
int mps_dump_bucket(mps_handle* h, int32_t* first, int32_t* last, int32_t* next);
/* Neighbour lists as MPS::setNeighbors (src/mps.cpp:585-619) builds them from
 * the current positions, in the reference's order. offsets has N+1 entries;
 * call with ids == NULL to get offsets only. Ghost particles get an empty
 * list here (the reference keeps a stale one). */
int mps_dump_neighbors(mps_handle* h, int32_t* offsets, int32_t* ids,
                       double* dists);
/* mps_dump_neighbors without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The lists
 * are staged in a static device array of the library (48 MiB per GPU, held
 * until the process ends) instead of buffers allocated per call, so a
 * system with more than 4,194,304 neighbour entries is refused (MPS_E_INVALID,
 * the message says so; mps_dump_neighbors takes any size).  Every handle on
 * one GPU shares that array: calls that overlap on one GPU (other host
 * threads, other handles) race.  Same arguments and output as
 * mps_dump_neighbors. */
int mps_dump_neighbors_noalloc(mps_handle* h, int32_t* offsets, int32_t* ids, double* dists);
/* Pressure system of the last step in id order: source term b (N), matrix
 * diagonal after the Dirichlet-connection doubling (N, 0 for non-Inner rows),
 * and the off-diagonal entries in CSR by id (row_ptr N+1, col, val: rows
 * sorted by column). Call with col == NULL to get row_ptr only. Also the
 * solution before clamping (N). */
int mps_get_pressure_system(const mps_handle* h, double* b, double* diag,
                            int32_t* row_ptr, int32_t* col, double* val,
                            double* x);
/* Positions and velocities as they were just before the pressure stage of the
 * last step (after collision), id order; needs the snapshot flag. */
int mps_set_snapshot(mps_handle* h, int32_t on);
int mps_get_snapshot(const mps_handle* h, double* pos, double* vel);

/* Launch block sizes of the per-particle kernel families (no counterpart in
 * the reference; the GPU path's tuning), in this order: bucket build,
 * viscosity+move, collision, BC+source, diagonal+connectivity, matrix fill,
 * pressure out+minimum pressure, gradient+move, number density.  Multiples of
 * 64 in [64, 512]; defaults 256.  Results do not depend on them.
 * mps_autotune runs 4 x steps_per_candidate real steps (the state advances
 * exactly as it would untuned), times each family at 64/128/256/512 threads
 * with HIP events and keeps the fastest per family. */
int mps_set_block_sizes(mps_handle* h, const int32_t block_sizes[9]);
int mps_autotune(mps_handle* h, int32_t steps_per_candidate, int32_t block_sizes_out[9]);

/* Times the pressure-solver kernels on the system of the last step, with HIP
 * events on the handle's stream, averaged over reps launches each (no
 * counterpart in the reference; a measurement hook).  The variants timed are
 * the ones the last solve ran (with the multigrid preconditioner: the apply
 * that gathers u = M r and the step that reads it).  out_ms / out_bytes:
 *   [0] FP64 SpMV, [1] FP64 vector update, [2] FP32 SpMV, [3] FP32 update,
 * bytes = algorithmic bytes of one launch (matrix entries, column indices and
 * vectors each read or written once).  Leaves the solver scratch undefined. */
int mps_time_pressure_kernels(mps_handle* h, int32_t reps, double out_ms[4], int64_t out_bytes[4]);

/* Reverse-mode gradients of the step (no counterpart in the reference, which
 * has no derivatives; an extension for inverse problems).
 *
 * mps_adjoint_record runs `steps` steps from the current state, exactly as
 * mps_run_steps does (courants may be NULL), and keeps a device checkpoint of
 * the state before every `checkpoint_interval`-th step.  The handle ends at
 * the final state.
 *
 * mps_adjoint_backward then takes dL/dx and dL/dv at that final state
 * (pos_bar, vel_bar: N x 3, id order; either may be NULL for zero) and returns
 * dL/dx and dL/dv at the recorded run's start state (pos_bar0, vel_bar0:
 * N x 3) and param_bar = {dL/dnu (kinematic viscosity), dL/de (coefficient of
 * restitution), dL/drelaxation (relaxation coefficient for pressure)}.  Each
 * segment between checkpoints is recomputed once to recover its step starts;
 * each step is then run again with its hooks on and reversed.  Every discrete
 * decision of the recorded run is held fixed (neighbour sets, boundary
 * conditions, Dirichlet connectivity, collision pairs that fired, the
 * negative-pressure clamp, the minimum-pressure argument, ghosting), so the
 * result is the derivative of the piecewise-smooth map on the piece the run
 * lies in.  The adjoint pressure equation is solved with the recorded run's
 * solver options (zero start); adjoint_stats (may be NULL) gets the total
 * iterations and the worst final relative residual.  Afterwards the handle is
 * back at the final state with its events and statistics unchanged.
 * Single-scene handles without the higher-order source term only. */
int mps_adjoint_record(mps_handle* h, int32_t steps, int32_t checkpoint_interval, double* courants);
int mps_adjoint_backward(mps_handle* h, const double* pos_bar, const double* vel_bar, double* pos_bar0,
                         double* vel_bar0, double* param_bar, mps_solver_stats* adjoint_stats);

/* Asynchronous export snapshot (no counterpart in the reference, whose
 * exporter reads the particle vector synchronously, src/simulation.cpp:231-245).
 * mps_snapshot_begin queues, on the handle's stream, the gather of type,
 * position, velocity, pressure, number density ratio and boundary condition
 * into id order and their copy into library-owned pinned host memory of slot
 * 0 or 1, and returns without waiting: later steps queue behind the copy, so
 * the state copied is the one of the call.  mps_snapshot_end waits for that
 * copy and returns pointers into the slot (N int8 types, N x 3 positions and
 * velocities, N pressures and number density ratios, N int8 boundary
 * conditions; any pointer argument may be NULL).  They stay valid until the
 * slot is begun again or the handle is destroyed.  mps_snapshot_end touches
 * only the slot, so another host thread may call it while the handle's owner
 * keeps stepping; it reports failure by status only (mps_last_error is not
 * updated). */
int mps_snapshot_begin(mps_handle* h, int32_t slot);
int mps_snapshot_end(mps_handle* h, int32_t slot, const int8_t** type, const double** pos, const double** vel,
                     const double** p, const double** ndr, const int8_t** bc);

/* Lossless checkpoint (the reference restarts from its 6-decimal CSV,
 * src/mps.cpp:38-94): the full particle state as a byte image, every field in
 * FP64 plus types, boundary conditions and the solver's warm start, so a run
 * resumed from it continues bit for bit as if uninterrupted.  The image is
 * mps_checkpoint_size() bytes (a 64-byte header, then 12 doubles and 2 bytes
 * per particle); mps_checkpoint_load accepts only an image of a handle of the
 * same shape (particle count, dim, scenes, each scene's particle count and
 * settings) and validates it first.  Debug
 * events and solver statistics are not part of the state. */
int64_t mps_checkpoint_size(const mps_handle* h);
int mps_checkpoint_save(mps_handle* h, void* buf, int64_t cap);
int mps_checkpoint_load(mps_handle* h, const void* buf, int64_t size);

#ifdef __cplusplus
}
#endif
#endif /* MPS_AMD_H */
