/*
 * pcc.h — C ABI of the MI355X collision-pipeline library (libpcc.so).
 *
 * One call of pcc_step() is one timestep of the proliferating-cell-colony
 * collision pipeline on one GPU: spatial-grid contact detection, constraint
 * assembly, then either the hard-model ReLCP/BBPGD complementarity solve or the
 * soft-model Hertz force assembly, then FP64 state integration and growth.
 *
 * It replaces, on one GPU and without MPI/PETSc:
 *   ParticleManager::step(int, std::function<void()>)
 *       code/cpp/src/simulation/ParticleManager.h:50, ParticleManager.cpp:150-166
 *   solveHardModel(...)      code/cpp/src/dynamics/models/Hard.h:9, Hard.cpp:91-193
 *   solveSoftPotential(...)  code/cpp/src/dynamics/models/Soft.h:8, Soft.cpp:21-114
 * and returns the same SolverSolution summary
 *   (ParticleManager.h:32-38; built at Hard.cpp:192 and Soft.cpp:113).
 *
 * Conventions
 *  - Plain C: no exceptions cross this boundary; every entry point returns a
 *    pcc_status and pcc_last_error() gives the message.  The three solver
 *    failures of the reference map 1:1 to PCC_E_BBPGD_NOT_CONVERGED
 *    ("BBPGD did not converge", BBPGD.cpp:118-121), PCC_E_RELCP_MAX_ITERS
 *    ("Maximum number of constraint iterations reached", Hard.cpp:187-190) and
 *    PCC_E_INVALID_MODE ("Invalid mode", ParticleManager.cpp:165).
 *  - Body index == global id (gid).  The reference numbers bodies 0..N-1 by
 *    ascending x on one rank (Domain::assignGlobalIDs, Domain.cpp:492-508);
 *    the caller passes bodies in gid order.
 *  - Host buffers are owned by the caller; device buffers by the context.  A
 *    context owns one HIP stream on one device; every call is stream-ordered
 *    and synchronous at return.  One context per host thread.
 *  - State stays resident on the device between steps.
 */
#ifndef PCC_H_
#define PCC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pcc_status {
  PCC_OK = 0,
  PCC_E_BBPGD_NOT_CONVERGED = 1, /* BBPGD.cpp:118-121 */
  PCC_E_RELCP_MAX_ITERS = 2,     /* Hard.cpp:187-190 */
  PCC_E_INVALID_MODE = 3,        /* ParticleManager.cpp:165 */
  PCC_E_INVALID_ARGUMENT = 4,
  PCC_E_DEVICE = 5,              /* a HIP runtime call failed */
  PCC_E_OUT_OF_MEMORY = 6,
  PCC_E_UNSUPPORTED = 7,         /* a request the library does not serve */
  PCC_E_NO_STATE = 8             /* pcc_step before pcc_set_state */
} pcc_status;

enum { PCC_MODE_HARD = 0, PCC_MODE_SOFT = 1 };

/* FP64: every kernel computes in double; the whole step is bit-identical to the
 *       FP64 restatement of the reference algorithm.
 * MIXED: constraint Jacobian/stress blocks, impulses and gradients are stored
 *       and applied in FP32 (D, D^T, L, L^T, Hertz sums); the BB step-size dot
 *       products, the projected-gradient residual, contact geometry and the
 *       position/orientation/length integration stay FP64. */
enum { PCC_PREC_FP64 = 0, PCC_PREC_MIXED = 1 };

/* Mirrors SimulationParameters (code/cpp/src/util/Config.h:9-56) for the fields
 * the step reads, with the reference's single `tolerance` (Q8: it drives the
 * BBPGD residual, the ReLCP overlap test and the phase>=1 contact cutoff,
 * Hard.cpp:121,132,153) split into three.  pcc_default_params() ties them as
 * the reference does. */
typedef struct pcc_params {
  int32_t mode;      /* PCC_MODE_HARD | PCC_MODE_SOFT  (-mode) */
  int32_t precision; /* PCC_PREC_FP64 | PCC_PREC_MIXED */
  double dt;         /* sim_config.dt_s (-dt) */
  double xi;         /* drag coefficient (-xi) */
  double tau;        /* growth time constant (-tau) */
  double lambda;     /* stress sensitivity (-lambda) */
  double temperature;/* Brownian temperature (-temperature): see brownian_seed */
  double kcc;        /* soft-model collision constant (-kcc) */
  double bbpgd_tol;      /* BBPGD projected-gradient residual bound */
  double relcp_tol;      /* ReLCP stops once max overlap <= relcp_tol (k > 0) */
  double cutoff_phase0;  /* contact cutoff of ReLCP phase 0 (reference: 0.25) */
  double cutoff_phaseK;  /* contact cutoff of phases k >= 1 (reference: tolerance) */
  int64_t max_bbpgd_iterations;     /* -max_bbpgd_iterations */
  int64_t max_recursive_iterations; /* -max_recursive_iterations */
  double linked_cell_size;          /* -linked_cell_size (default 2.2*l0) */
  double grid_origin;               /* -1 on every axis (CollisionDetector.cpp:24) */
  double warm_start_factor;         /* gamma0 = -factor*phi for phi < 0 (500, CollisionDetector.cpp:308) */
  /* Brownian input (calculate_external_velocities, Physics.cpp:59-81, and
   * Particle::calculateBrownianVelocity, Particle.cpp:174-190).  With
   * temperature T > 0, ReLCP phase 0 of a hard step and every soft step add
   * to each body's angular velocity
   *   omega_z = ((T u1) u_z) iota iota iota,
   *   omega_x = ((T u1) u_x) iota^3, omega_y = ((T u1) u_y) iota^3 unless
   *   monolayer,
   * where iota is the body's impedance from its last step (0 for a body that
   * has not been stepped, as Particle.cpp:25,34 initialise it) and the u are
   * uniform on [-1, 1), drawn in the reference's order (u1, then u_x, u_y
   * unless monolayer, then u_z).  The reference draws them from one unseeded
   * global std::mt19937 (Physics.cpp:19); here they come from a counter-based
   * generator keyed by (brownian_seed, the colony's step count, the body's
   * index within its colony, draw number), so a run is reproducible and a
   * colony's draws are the same alone and in any batch.  The step count is
   * the number of step calls since the colony was set (pcc_set_state,
   * pcc_batch_set_state) or added (pcc_batch_add); pcc_snapshot and
   * pcc_restore keep it with the state, pcc_batch_set_step_count sets it
   * (to resume a checkpoint). */
  int32_t monolayer;                /* -monolayer (default true, CLIParser.h:183) */
  int32_t reserved0;
  uint64_t brownian_seed;
} pcc_params;

/* Host-side body state, gid order.  Vectors are interleaved per body
 * (pos[3*i+k], quat[4*i+k] with quat = (w, x, y, z)).  Fields marked (out)
 * are written by pcc_get_state and ignored by pcc_set_state (may be NULL). */
typedef struct pcc_bodies {
  int64_t n;
  double* pos;        /* 3n */
  double* quat;       /* 4n, unit (w, x, y, z) */
  double* length;     /* n */
  double* l0;         /* n */
  double* diameter;   /* n */
  double* ldot;       /* n */
  double* force;      /* 3n (out): last ReLCP phase's force increment (Q9) */
  double* torque;     /* 3n (out) */
  double* velocity;   /* 3n (out) */
  double* omega;      /* 3n (out) */
  double* impedance;  /* n: exp(-lambda stress) of the last step.  Read by
                       * pcc_set_state as the impedance the Brownian input of
                       * the next step uses (NULL: 0, a new body's value) and
                       * written by pcc_get_state. */
  double* stress;     /* n (out) */
  int32_t* num_constraints; /* n (out) */
  int32_t* age;       /* n: generation (ParticleData::age, ParticleData.h:16): 1 for a new cell,
                       * +1 for both halves at each division (Particle.cpp:26,144,148).  May be
                       * NULL: pcc_set_state then starts every cell at 1. */
} pcc_bodies;

/* SoA image of std::vector<Constraint> (code/cpp/src/dynamics/Constraint.h:6-49).
 * `capacity` is the length of every array on input (vectors hold 3*capacity). */
typedef struct pcc_constraints {
  int64_t capacity;
  int32_t* gidI;            /* lower body index of the pair */
  int32_t* gidJ;
  double* signed_distance;  /* phi */
  double* normal;           /* 3c, normI = normalize(closest_I - closest_J) */
  double* rPosI;            /* 3c */
  double* rPosJ;            /* 3c */
  double* contact_point;    /* 3c */
  double* stressI;
  double* stressJ;
  double* gamma;
  int32_t* gid;             /* index within its detection batch */
  int32_t* iteration;       /* ReLCP phase that created it */
} pcc_constraints;

/* SolverSolution without the constraint vector (fetch it with
 * pcc_get_constraints). */
typedef struct pcc_step_result {
  int64_t n_constraints;
  int32_t constraint_iterations;
  int32_t reserved;
  int64_t bbpgd_iterations;
  double residual;
  double max_overlap;
} pcc_step_result;

/* One BBPGD iteration of a traced step: BBPGDStep (logger/BBPGDLogger.h:7-17)
 * as collected at BBPGD.cpp:57-80, with the energy terms of
 * HardModelGradient::energy (HardModelGradient.cpp:160-203) at the iterate:
 *   linear    = gamma . Phi
 *   quadratic = dt/2 (D gamma) . M (D gamma)
 *   growth    = dt/(lambda tau) sum_i l_i exp(-lambda (L gamma)_i)
 *   total     = linear + quadratic + growth. */
typedef struct pcc_bbpgd_record {
  int64_t step;                /* BBPGD iteration (0 = the initial gradient) */
  int64_t recursive_iteration; /* ReLCP phase of the solve */
  int64_t total_constraints;   /* constraints of the phase */
  double residual;             /* projected-gradient residual at the iterate */
  double step_size;            /* BB step that produced the iterate (1/res at step 0) */
  double linear;
  double quadratic;
  double growth;
  double total;
} pcc_bbpgd_record;

typedef struct pcc_ctx pcc_ctx;

/* Reference defaults (code/cpp/src/util/CLIParser.h:167-194,232) with the
 * tolerance coupling of Q8, for mode hard, FP64, l0 = 1: temperature 1e-28,
 * monolayer on, brownian_seed 0. */
void pcc_default_params(pcc_params* p);

pcc_status pcc_create(const pcc_params* params, int device, pcc_ctx** out);
void pcc_destroy(pcc_ctx* ctx);
pcc_status pcc_set_params(pcc_ctx* ctx, const pcc_params* params);
/* The message of the context's last failed call; with ctx == NULL, that of
 * the calling thread's last failed call of an entry point without a context
 * (pcc_detect_noalloc). */
const char* pcc_last_error(const pcc_ctx* ctx);
const char* pcc_status_string(pcc_status s);

/* Upload/download the body state (host buffers, gid order). */
pcc_status pcc_set_state(pcc_ctx* ctx, const pcc_bodies* bodies);
pcc_status pcc_get_state(pcc_ctx* ctx, pcc_bodies* bodies);

/* Device-side copy of the current body state into a second resident buffer,
 * and back: lets a caller replay the same step from identical inputs without
 * host traffic. */
pcc_status pcc_snapshot(pcc_ctx* ctx);
pcc_status pcc_restore(pcc_ctx* ctx);

/* One timestep (ParticleManager::step).  `iter` is the step index, used by the
 * reference only to gate tracing (Hard.cpp:109).  On PCC_E_RELCP_MAX_ITERS the
 * state has been advanced and grown, as the reference does before throwing. */
pcc_status pcc_step(pcc_ctx* ctx, int32_t iter, pcc_step_result* out);

/* Constraint set of the last step (all ReLCP phases, in solver order).
 * *count receives the number of constraints; arrays are filled up to
 * out->capacity entries. */
pcc_status pcc_get_constraints(pcc_ctx* ctx, pcc_constraints* out, int64_t* count);

/* Contact detection only, at the current state and a given cutoff: fills the
 * constraints of one detection batch (CollisionDetector::detectCollisions,
 * CollisionDetector.cpp:84-111) and, if non-NULL, the grid ordering
 * (bucket of every body and the stable body permutation by bucket). */
pcc_status pcc_detect(pcc_ctx* ctx, double cutoff, pcc_constraints* out, int64_t* count,
                      uint32_t* bucket_out, int32_t* perm_out);

/* ---------------------------------------------------------------------------
 * Contact detection that allocates nothing.  No reference counterpart: it is
 * for callers that cannot allocate device memory (set-up code around HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * The same detection batch as pcc_detect (CollisionDetector::detectCollisions,
 * CollisionDetector.cpp:84-111; records bit-identical, in (contact x, I, J)
 * order, gid = position, iteration = 0) of ONE colony whose state is already
 * in device memory, written into device arrays the caller owns.  There is no
 * context: the library's scratch (grid, per-slot counts, raw records, sort
 * storage) is one fixed array in device memory per GPU, reserved when the
 * library's code for this entry point is first loaded on that GPU and held
 * until the process ends (about 80 MiB).  It serves at most
 * PCC_NOALLOC_MAX_BODIES bodies and PCC_NOALLOC_MAX_CONTACTS contacts, and
 * at most PCC_NOALLOC_MAX_BLOCK_CANDIDATES candidate pairs (pairs whose
 * centres pass the reference's bounding-sphere filter) among the neighbour
 * cells of each group of 256 (body, neighbour cell) slots in grid order
 * (dense 2D packings with the default grid give ~100); larger inputs are
 * rejected with PCC_E_INVALID_ARGUMENT (a batch found to exceed the contact
 * limit returns that status with *count set and nothing written).  Every call on one GPU shares that GPU's scratch and leaves its
 * contents behind: calls that overlap on one GPU (from several threads, or on
 * several streams) race and must be serialised by the caller.
 *
 *   device  the HIP device; it is made current for the call and the
 *           previous current device restored on return
 *   stream  a hipStream_t on that device (NULL: the null stream); the call
 *           is ordered on it and synchronous at return (it reads the record
 *           count back to the host once)
 *   params  linked_cell_size, grid_origin and warm_start_factor are used
 *   bodies  device pointers, gid order: pos (3n, interleaved), quat (4n),
 *           length (n), diameter (n)
 *   out     device pointers (pcc_constraints layout); up to out->capacity
 *           records are written; NULL arrays are skipped
 *   count   receives the number of contacts found
 * Errors that have no context are reported by pcc_last_error(NULL) on the
 * calling thread. */
#define PCC_NOALLOC_MAX_BODIES 32768
#define PCC_NOALLOC_MAX_CONTACTS 262144
#define PCC_NOALLOC_MAX_BLOCK_CANDIDATES 512
typedef struct pcc_device_bodies {
  int64_t n;
  const double* pos;
  const double* quat;
  const double* length;
  const double* diameter;
} pcc_device_bodies;
pcc_status pcc_detect_noalloc(int device, void* stream, const pcc_params* params, double cutoff,
                              const pcc_device_bodies* bodies, const pcc_constraints* out,
                              int64_t* count);

/* Per-ReLCP-phase diagnostics of the last step (the data the reference's
 * SimulationStep/BBPGD loggers print, SimulationLogger.h:7-30): BBPGD
 * iterations and constraint-set size of each solved phase.  *count receives
 * the number of solved phases; arrays are filled up to `max` entries. */
pcc_status pcc_get_phase_stats(pcc_ctx* ctx, int64_t* bbpgd_iterations, int64_t* n_constraints,
                               int64_t max, int64_t* count);

/* ---------------------------------------------------------------------------
 * Batched ensembles: one context holds many independent colonies (same mode
 * and precision; every other parameter per colony).  A batched step runs all
 * colonies' ReLCP loops in lockstep with shared launches: every BBPGD
 * iteration is one launch sequence over all colonies still solving, with a
 * per-colony step size, iteration count and convergence flag on the device
 * (a converged colony's blocks exit at once).  Each colony's results are the
 * ones a context holding it alone produces, bit for bit.  Colonies are
 * indexed 0..n-1 in the order they were set or added; removing one shifts the
 * later ones down.  The single-colony entry points above require a context
 * with exactly one colony.  No reference counterpart: the reference runs one
 * colony per MPI job (Domain::run, Domain.cpp:121-143).
 * ------------------------------------------------------------------------- */

/* Replaces the context's colonies: bodies[q] and, if params is not NULL,
 * params[q] (else the context's parameters) for q < n_colonies. */
pcc_status pcc_batch_set_state(pcc_ctx* ctx, int64_t n_colonies, const pcc_bodies* bodies,
                               const pcc_params* params);
/* Appends one colony (state stays resident; *colony receives its index). */
pcc_status pcc_batch_add(pcc_ctx* ctx, const pcc_bodies* bodies, const pcc_params* params,
                         int64_t* colony);
/* Removes colony `colony`; later colonies move down by one.  Removing the
 * last colony leaves the context without state (pcc_batch_add starts a new
 * batch). */
pcc_status pcc_batch_remove(pcc_ctx* ctx, int64_t colony);
pcc_status pcc_batch_set_params(pcc_ctx* ctx, int64_t colony, const pcc_params* params);
pcc_status pcc_batch_size(pcc_ctx* ctx, int64_t* n_colonies, int64_t* n_bodies);
/* One timestep of every colony.  results[q] and statuses[q] (may be NULL)
 * receive each colony's SolverSolution summary and solver status (the three
 * solver failures of the reference are per colony; the call itself returns
 * PCC_OK unless the device or the arguments fail). */
pcc_status pcc_batch_step(pcc_ctx* ctx, int32_t iter, pcc_step_result* results,
                          pcc_status* statuses);
pcc_status pcc_batch_get_state(pcc_ctx* ctx, int64_t colony, pcc_bodies* bodies);
/* Colony `colony`'s constraints; gidI / gidJ are indices within the colony. */
pcc_status pcc_batch_get_constraints(pcc_ctx* ctx, int64_t colony, pcc_constraints* out,
                                     int64_t* count);
pcc_status pcc_batch_get_phase_stats(pcc_ctx* ctx, int64_t colony, int64_t* bbpgd_iterations,
                                     int64_t* n_constraints, int64_t max, int64_t* count);
/* Colony `colony`'s step count (0 for a single-colony context), which keys
 * its Brownian draws (see pcc_params.brownian_seed): get it to checkpoint a
 * run, set it to resume one.  No reference counterpart (the reference's
 * generator is one global stream). */
pcc_status pcc_batch_get_step_count(pcc_ctx* ctx, int64_t colony, int64_t* count);
pcc_status pcc_batch_set_step_count(pcc_ctx* ctx, int64_t colony, int64_t count);

/* Cell division and gid reassignment between steps, on the device
 * (ParticleManager::divideParticles / Particle::divide, ParticleManager.cpp:139-148,
 * Particle.cpp:122-150, then Domain::assignGlobalIDs, Domain.cpp:492-508):
 * every cell with length >= 2 l0 splits into l1 = U[1-length_noise,
 * 1+length_noise] L/2 and l2 = L - l1 at x -/+ (l/2) t (the reference draws
 * length_noise 0.02 from an unseeded std::random_device); each child is then
 * turned in-plane about z by an angle drawn from N(0, angle_noise) (a seeded
 * Irwin-Hall approximation; angle_noise <= 0.2).  Draws depend on (seed, step,
 * index of the parent within its colony) only.  Cells are then renumbered by
 * ascending x (ties keep parents first, right children after, in parent
 * order), also when no cell divides, as the reference renumbers before every
 * step.  The per-body outputs of the last step are reset to zero.
 * pcc_divide: single-colony context, *n_bodies receives the new count;
 * pcc_batch_divide: seeds[q] per colony, sizes[q] receives the new counts. */
pcc_status pcc_divide(pcc_ctx* ctx, uint64_t seed, int32_t step, double length_noise,
                      double angle_noise, int64_t* n_bodies);
pcc_status pcc_batch_divide(pcc_ctx* ctx, const uint64_t* seeds, int32_t step,
                            double length_noise, double angle_noise, int64_t* sizes);

/* Autotuner state.  The first BBPGD solve with at least 1e5 constraints
 * times every variant of the body kernel on its own operators and keeps the
 * fastest for the context (PCC_AUTOTUNE=0 in the environment keeps the
 * default, variant 1).  The choice is also kept for the process, per (device
 * architecture, planar, precision, power-of-two bucket of the bodies, bucket
 * of 4 C / N): a later context whose first such solve falls in a known class
 * takes it without timing.  Variants v = 0..5: staging size (incidences per LDS
 * pass) 512/1024/2048 planar or 256/512/1024 3D for v % 3, threads per block
 * 256 (v < 3) or 128 (two bodies per thread); v = 6..8: 256 threads per
 * block, staging 256/512/1024, the next pass's incidence records loaded while
 * the current pass is summed.  The sums depend on none of these.  *variant,
 * *body_chunk, *body_threads: the variant in use; *tuned: 1 once timed in
 * this context, 2 when taken from the process's memory, 0 before;
 * variant_us[PCC_BODY_VARIANTS] (may be NULL): the measured launch times.
 * Any output pointer may be NULL.  No reference counterpart. */
#define PCC_BODY_VARIANTS 9
pcc_status pcc_get_tuning(pcc_ctx* ctx, int32_t* variant, int32_t* body_chunk,
                          int32_t* body_threads, int32_t* tuned, double* variant_us);

/* Counters of the context's multi-kernel BBPGD solves since its creation:
 * *compactions = times a solve's grids were rebuilt over the colonies still
 * iterating (once those hold at most half of the body blocks).  Any output
 * pointer may be NULL.  No reference counterpart. */
pcc_status pcc_get_solve_stats(pcc_ctx* ctx, int64_t* compactions);

/* CFL time-step adaptation between steps (Domain::adaptDt, Domain.cpp:69-118),
 * on the device: the median over the colony's cells of |v| + ldot after the
 * last step (mean of the two middle values for an even count), then
 * dt' = clamp(dt (1 - alpha) + alpha cfl_factor tolerance / median,
 * 0.9 dt, 1.1 dt); dt is kept when the median is 0.  The reference uses
 * cfl_factor 0.5 (-cfl_factor), alpha 0.01 and its single `tolerance`
 * (Domain.cpp:106), which is the colony's relcp_tol here in both entry
 * points.  pcc_adapt_dt updates the context's dt; pcc_batch_adapt_dt every
 * colony's.  *dt / dts[q] receive the new step, *median / medians[q] the
 * median speed (either may be NULL). */
pcc_status pcc_adapt_dt(pcc_ctx* ctx, double cfl_factor, double alpha, double* dt,
                        double* median);
pcc_status pcc_batch_adapt_dt(pcc_ctx* ctx, double cfl_factor, double alpha, double* dts,
                              double* medians);

/* BBPGD trace (the reference's bbpgd_logger, Hard.cpp:107-116,181-183, and
 * BBPGD.cpp:57-80).  While enabled, every hard step records, per colony, one
 * pcc_bbpgd_record per BBPGD iteration of every ReLCP phase (up to
 * `capacity` records per colony and step; later ones are counted, not
 * stored), and, as in the reference with a logger attached, checks the
 * residual at every iteration instead of every tenth, so a traced solve may
 * stop earlier than an untraced one.  The energy terms are summed on the
 * device in a fixed order.  enable = 0 turns it off (capacity ignored). */
pcc_status pcc_set_trace(pcc_ctx* ctx, int enable, int64_t capacity);
/* Records of the last step for colony `colony` (0 for a single-colony
 * context): copies min(count, capacity, trace capacity) records to `out` (may
 * be NULL with capacity 0); *count = the iterations traced in that step. */
pcc_status pcc_get_trace(pcc_ctx* ctx, int64_t colony, pcc_bbpgd_record* out, int64_t capacity,
                         int64_t* count);

/* Device memory held by the context (every device buffer it has allocated),
 * for admission against an HBM budget.  No reference counterpart. */
pcc_status pcc_device_bytes(pcc_ctx* ctx, int64_t* bytes);

/* Achievable device copy bandwidth (bytes read + written per second / 1e9)
 * of a grid-stride copy of `bytes` bytes, averaged over `reps` launches: the
 * practical ceiling beside the published HBM peak.  No reference
 * counterpart. */
pcc_status pcc_measure_copy_bandwidth(int device, int64_t bytes, int reps, double* gbs);

/* Achievable device read bandwidth (bytes read per second / 1e9): a
 * grid-stride sum over `bytes` bytes, four independent 32-byte loads in
 * flight per thread, averaged over `reps` launches.  The BBPGD applies are
 * read-dominated, so this is their practical ceiling.  No reference
 * counterpart. */
pcc_status pcc_measure_read_bandwidth(int device, int64_t bytes, int reps, double* gbs);

/* Per-kernel GPU time accounting with HIP events on the context stream.
 * Kernel ids: see pcc_kernel_name(). */
enum {
  PCC_K_BODY_APPLY = 0,   /* D^T/L^T segmented reduction + mobility + growth law */
  PCC_K_CONSTR_APPLY = 1, /* D/L gather + projection + BB dots + residual */
  PCC_K_DETECT = 2,
  PCC_K_ASSEMBLE = 3,
  PCC_K_INTEGRATE = 4,
  PCC_K_HERTZ = 5,
  PCC_K_COUNT = 6
};
pcc_status pcc_set_timing(pcc_ctx* ctx, int enable);
pcc_status pcc_get_timing(pcc_ctx* ctx, int kernel, int64_t* launches, double* total_ms,
                          double* bytes_per_launch);
const char* pcc_kernel_name(int kernel);

#ifdef __cplusplus
}
#endif

#endif /* PCC_H_ */
