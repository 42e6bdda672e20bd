/*
 * mps_oracle.h — FP64 CPU restatement of the reference MPS time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker and the timed CPU path; the
 * product library (libmps_amd.so) never links or calls it.
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#include <stdint.h>
#include "../include/mps_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_state oracle_state;

/* RefValues (src/mps.cpp:132-163). */
void oracle_ref_values(int32_t dim, double l0, double re, double* n0, double* lambda);
/* weight (src/mps.cpp:9-16). */
double oracle_weight(double dist, double re);

oracle_state* oracle_create(const mps_settings* s, int32_t n, const int32_t* type,
                            const double* pos, const double* vel);
void oracle_destroy(oracle_state* o);

/* CG controls: tol (relative, Eigen default = DBL_EPSILON), max_iter (<=0:
 * 2N, Eigen default), duplicate (1: solve twice as src/mps.cpp:457,460). */
void oracle_set_solver(oracle_state* o, double tol, int32_t max_iter, int32_t duplicate);

void oracle_load_initial_state(oracle_state* o);
double oracle_step(oracle_state* o, int32_t is_export);

void oracle_get_state(const oracle_state* o, int32_t* type, double* pos, double* vel,
                      double* p, double* pmin, double* ndr, int32_t* bc);
void oracle_set_state(oracle_state* o, const int32_t* type, const double* pos,
                      const double* vel, const double* p, const double* ndr,
                      const int32_t* bc);

/* debugLogCount split by cause: [0] out of domain, [1] unconnected Inner,
 * [2] CFL exceeded. Ghosted / unconnected ids in the order logged. */
void oracle_debug_counts(const oracle_state* o, int64_t counts[3]);
int32_t oracle_event_ids(const oracle_state* o, int32_t kind, int32_t* ids, int32_t cap);

/* Stage-level access for tests. */
void oracle_set_neighbors(oracle_state* o);
void oracle_bucket_shape(const oracle_state* o, int32_t dims[4], double* length);
void oracle_dump_bucket(const oracle_state* o, int32_t* first, int32_t* last, int32_t* next);
int64_t oracle_neighbor_total(const oracle_state* o);
void oracle_dump_neighbors(const oracle_state* o, int32_t* offsets, int32_t* ids, double* dists);

/* Last step's pressure system: b (N), diag (N, after doubling; 0 for
 * non-Inner), off-diagonal CSR by id (rows sorted by column), x (N, CG
 * solution before clamping). */
int64_t oracle_offdiag_nnz(const oracle_state* o);
void oracle_get_pressure_system(const oracle_state* o, double* b, double* diag,
                                int32_t* row_ptr, int32_t* col, double* val, double* x);
/* Positions / velocities just before calcPressure of the last step. */
void oracle_get_snapshot(const oracle_state* o, double* pos, double* vel);
/* Iterations and final relative (recursive) residual of the last CG solve. */
void oracle_cg_stats(const oracle_state* o, int32_t* iterations, double* rel_residual);
/* Accumulated wall seconds per stage: [0] neighbour builds, [1] explicit stages,
 * [2] collision, [3] BC+source+matrix+BFS, [4] CG, [5] pressure post + gradient. */
void oracle_stage_seconds(const oracle_state* o, double out[6]);

/* Stand-alone Jacobi-PCG with Eigen's ConjugateGradient semantics on a CSR
 * matrix (full rows, sorted columns). Returns iterations. */
int32_t oracle_pcg(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                   const double* b, double* x, double tol, int32_t max_iter, double* rel_res);

#ifdef __cplusplus
}
#endif
#endif
