/*
 * pcc_oracle.h — FP64 CPU restatement of the reference's per-timestep collision
 * pipeline.  TEST INFRASTRUCTURE ONLY: it is the checker for libpcc (the HIP
 * library) and the CPU baseline of bench.py.  Nothing in the product links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout, code/cpp/src/...).  Deviations from the reference,
 * each deliberate and shared with the GPU path:
 *   - pair orientation is canonical (I = lower gid); the reference's depends on
 *     std::unordered_map iteration order (SpatialGrid.cpp:141);
 *   - ties of the contact-x sort (CollisionDetector.cpp:94-96, an unstable
 *     std::sort) are broken by (gidI, gidJ);
 *   - the broad-phase grid is a hashed cell list with a stable counting sort
 *     (same cell geometry as SpatialGrid.cpp:29-56, same candidate set);
 *   - BB dot products (BBPGD.cpp:3-29) are summed in a fixed blocked tree
 *     (256-wide blocks, 64-lane butterflies) instead of one sequential loop;
 *   - exp() (HardModelGradient.cpp:40, VecExp) is a fixed Cody-Waite +
 *     degree-13 Taylor evaluation so CPU and GPU round identically;
 *   - the Brownian draws (temperature > 0) come from a counter-based
 *     generator keyed by (brownian_seed, step index, body index, draw) in
 *     place of the reference's unseeded global std::mt19937 (Physics.cpp:19).
 */
#ifndef PCC_ORACLE_H_
#define PCC_ORACLE_H_

#include <stdint.h>

#include "pcc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Segment-segment closest points, SpherocylinderCell.h:64-261.
 * out = {s, t, c0[3], c1[3], sqrDistance, distance}. */
void oracle_dcp(const double P0[3], const double P1[3], const double Q0[3], const double Q1[3],
                double out[10]);
/* Quaternion.cpp:7-38 (rotate e_x by q = (w,x,y,z)). */
void oracle_direction(const double q[4], double t[3]);
/* ArrayMath.h:219-231. */
void oracle_normalize(const double a[3], double out[3]);
double oracle_exp(double x);

/* Division of every cell with length >= 2 l0 and gid reassignment by x
 * (Particle.cpp:122-150, Domain.cpp:492-508), seeded; arrays hold `cap`
 * bodies; age (generation, may be NULL) is carried as a double; impedance
 * (may be NULL): the parent keeps its value, the new cell starts at 0
 * (Particle.cpp:25).  Returns the new body count, -1 if it exceeds cap. */
int64_t oracle_divide(int64_t n, int64_t cap, double* pos, double* quat, double* length,
                      double* l0, double* diameter, double* ldot, double* age,
                      double* impedance, uint64_t seed, int32_t step, double length_noise,
                      double angle_noise);

/* Broad phase.  cell = floor((x - origin)/h) per axis (SpatialGrid.cpp:29-36);
 * bucket = slot(cell) & (table_size-1), the slot row-major (1024 cells to a
 * row) when every body lies in z = 0 with its axis in the plane (quat NULL:
 * positions decide alone), else a hash; stable counting sort of bodies by
 * bucket.  Returns table_size; fills cell (3n), bucket (n), perm (n) and
 * bucket_start (table_size+1) when non-NULL and large enough
 * (bucket_start_cap >= table_size+1). */
int64_t oracle_grid(int64_t n, const double* pos, const double* quat, double h, double origin,
                    int32_t* cell, uint32_t* bucket, int32_t* perm, int32_t* bucket_start,
                    int64_t bucket_start_cap);
/* 1 if every body lies in z = 0 with its axis in the plane (the grid's
 * row-major slots). */
int oracle_grid_flat(int64_t n, const double* pos, const double* quat);
int64_t oracle_table_size(int64_t n);

/* One detection batch (CollisionDetector::detectCollisions,
 * CollisionDetector.cpp:84-111): candidates from the grid (brute == 0) or all
 * pairs (brute != 0), filtered and built by tryCreateConstraint
 * (CollisionDetector.cpp:201-311), sorted by (contact x, gidI, gidJ).
 * Returns the count; fills `out` up to its capacity.  iteration is stored in
 * every record; num_constraints (if non-NULL) is incremented per contact. */
int64_t oracle_detect(const pcc_bodies* b, double cutoff, double h, double origin,
                      double warm_start_factor, int32_t iteration, int brute,
                      pcc_constraints* out, int32_t* num_constraints);

/* One timestep: ParticleManager::step (ParticleManager.cpp:150-166).
 * Bodies are updated in place (all (out) fields written); b->impedance (may
 * be NULL: 0) is also read, as the impedance of the Brownian input, and iter
 * keys the Brownian draws: it is the colony's step count (step calls since
 * its state was set), which is what the library uses.  The constraint set
 * of the step is written to `cons` (may be NULL) up to its capacity;
 * result->n_constraints is the full count.  Optional diagnostics: if non-NULL,
 * phase_iters[k] receives the BBPGD iterations of phase k (up to
 * max_phase_entries). */
pcc_status oracle_step(const pcc_params* p, pcc_bodies* b, int32_t iter, pcc_step_result* result,
                       pcc_constraints* cons, int64_t* phase_iters, int64_t max_phase_entries,
                       char* err, int64_t err_len);

/* BBPGD iteration budget for timing: run the hard step but stop every BBPGD
 * solve after `iteration_cap` iterations (no error), so a bounded CPU sample of
 * a large step can be timed.  0 = no cap. */
void oracle_set_iteration_cap(int64_t iteration_cap);

/* Wall time (s) of each completed ReLCP phase of the last hard step;
 * returns the number of phases. */
int64_t oracle_phase_times(double* out, int64_t max);

/* BBPGD trace of the following hard steps (BBPGDLogger; BBPGD.cpp:57-80):
 * one record per BBPGD iteration of every phase, up to `capacity`; the
 * residual is checked at every iteration while a buffer is set.  NULL turns
 * it off.  oracle_trace_count(): iterations traced since the last set. */
void oracle_set_trace(pcc_bbpgd_record* buf, int64_t capacity);
int64_t oracle_trace_count(void);

/* Phase 0 of a hard step at the given state, for tests of the energy: detects
 * with cutoff_phase0, builds the operators (gamma_old = the warm start,
 * ldot_prev = b->ldot) and, if gamma (one value per constraint) is non-NULL,
 * writes E = HardModelGradient::energy (HardModelGradient.cpp:160-203) and
 * g = HardModelGradient::gradient (:89-121) at gamma (either may be NULL).
 * Returns the number of constraints. */
int64_t oracle_phase0_energy(const pcc_params* p, const pcc_bodies* b, const double* gamma,
                             double E[4], double* g);

/* Deterministic blocked sum used for the BB dot products (exposed for tests). */
double oracle_block_sum(const double* v, int64_t n);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int oracle_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* PCC_ORACLE_H_ */
