/*
 * mps_oracle.c — FP64 CPU restatement of the reference MPS time step
 * (src/mps.cpp:9-667, src/bucket.cpp:10-51, src/particle.cpp:19-29).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Nothing in the product library links or calls this file.
 *
 * Fidelity rules followed throughout:
 *  - same loop orders and the same FP64 expression order as the reference;
 *    compiled with -ffp-contract=off so no multiply-add is fused;
 *  - Eigen::Vector3d reductions (norm, squaredNorm, dot) evaluate as
 *    (a0+a1)+a2: Vector3d is packet-accessible under Eigen's default unaligned
 *    vectorisation, so redux sums the first two lanes as one packet and adds
 *    the third (Eigen Core/Redux.h, LinearVectorizedTraversal);
 *  - neighbour lists are per-particle arrays kept across builds, so a Ghost
 *    keeps its stale list exactly as src/mps.cpp:589-593 does;
 *  - OpenMP only on the loops the reference marks `omp parallel for` in
 *    mps.cpp; bucket insertion, collision, assembly, BFS and CG are serial.
 *
 * The CG follows Eigen's ConjugateGradient<SparseMatrix<double>> algorithm
 * (Jacobi preconditioner with 1 for a missing or zero diagonal, zero initial
 * guess, stop when ||r||^2 < max(tol^2 ||b||^2, DBL_MIN), at most 2N
 * iterations).  Its SpMV sums each row in ascending column order; Eigen's
 * lower-triangle self-adjoint product orders those sums differently, so the
 * CG iterates are pinned by the solution (direct solves), not by bits.
 */
#include "mps_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { T_GHOST = -1, T_FLUID = 0, T_WALL = 1, T_DUMMY = 2 };
enum { BC_IGNORED = 0, BC_SURFACE = 1, BC_INNER = 2 };

typedef struct {
    int32_t id;
    double dist;
} onbr;

typedef struct {
    double n0, lambda;
} refv;

struct oracle_state {
    int32_t n;
    mps_settings s;
    refv ref_p, ref_v, ref_s;
    /* particles (src/particle.hpp:21-46), id order */
    int32_t* type;
    double* pos;  /* 3n */
    double* vel;  /* 3n */
    double* acc;  /* 3n */
    double* p;
    double* pmin;
    double* rho;
    double* ndr;
    int32_t* bc;
    unsigned char* dconn;
    /* neighbour lists */
    int32_t* nb_cnt;
    int32_t* nb_cap;
    onbr** nb;
    /* bucket (src/bucket.hpp:11-14) */
    int32_t num, numX, numY, numZ;
    double len;
    int32_t *first, *last, *next;
    /* pressure system of the last step */
    double* b;
    double* diag;
    int32_t* rp;  /* n+1 */
    int32_t* cl;
    double* vl;
    int64_t nnz_cap;
    double* x;
    /* solver */
    double tol;
    int32_t max_iter;
    int32_t duplicate;
    int32_t cg_iters;
    double cg_relres;
    /* events */
    int64_t counts[3];
    int32_t* ev_ids[2];
    int32_t ev_n[2], ev_cap[2];
    /* snapshot before calcPressure */
    double* snap_pos;
    double* snap_vel;
    double secs[6];
};

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

/* Eigen Vector3d helpers: (a0+a1)+a2 reductions. */
static inline double sqnorm3(double x, double y, double z) { return (x * x + y * y) + z * z; }
static inline double dot3(const double* a, const double* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* src/mps.cpp:9-16 */
double oracle_weight(double dist, double re) {
    double w = 0.0;
    if (dist < re) w = (re / dist) - 1.0;
    return w;
}

/* src/mps.cpp:132-163 */
void oracle_ref_values(int32_t dim, double l0, double re, double* n0_out, double* lambda_out) {
    int iz0, iz1;
    if (dim == 2) {
        iz0 = 0;
        iz1 = 1;
    } else {
        iz0 = -4;
        iz1 = 5;
    }
    double n0 = 0.0, lambda = 0.0;
    for (int ix = -4; ix < 5; ix++)
        for (int iy = -4; iy < 5; iy++)
            for (int iz = iz0; iz < iz1; iz++) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                double rx = l0 * (double) ix, ry = l0 * (double) iy, rz = l0 * (double) iz;
                double dist = sqrt(sqnorm3(rx, ry, rz));
                double dist2 = dist * dist;
                n0 += oracle_weight(dist, re);
                lambda += dist2 * oracle_weight(dist, re);
            }
    lambda /= n0;
    *n0_out = n0;
    *lambda_out = lambda;
}

/* src/particle.cpp:19-29 */
static inline double inverse_density(const oracle_state* o, int32_t i) {
    return o->type[i] == T_FLUID ? 1 / o->rho[i] : 0;
}

static void push_event(oracle_state* o, int kind, int32_t id) {
    if (o->ev_n[kind] == o->ev_cap[kind]) {
        o->ev_cap[kind] = o->ev_cap[kind] ? 2 * o->ev_cap[kind] : 64;
        o->ev_ids[kind] = (int32_t*) realloc(o->ev_ids[kind], sizeof(int32_t) * o->ev_cap[kind]);
    }
    o->ev_ids[kind][o->ev_n[kind]++] = id;
}

oracle_state* oracle_create(const mps_settings* s, int32_t n, const int32_t* type,
                            const double* pos, const double* vel) {
    oracle_state* o = (oracle_state*) calloc(1, sizeof(oracle_state));
    o->n = n;
    o->s = *s;
    oracle_ref_values(s->dim, s->particle_distance, s->re_pressure, &o->ref_p.n0, &o->ref_p.lambda);
    oracle_ref_values(s->dim, s->particle_distance, s->re_viscosity, &o->ref_v.n0, &o->ref_v.lambda);
    oracle_ref_values(s->dim, s->particle_distance, s->re_surface, &o->ref_s.n0, &o->ref_s.lambda);
    size_t N = (size_t) n;
    o->type = (int32_t*) malloc(sizeof(int32_t) * N);
    o->pos = (double*) malloc(sizeof(double) * 3 * N);
    o->vel = (double*) malloc(sizeof(double) * 3 * N);
    o->acc = (double*) calloc(3 * N, sizeof(double));
    o->p = (double*) calloc(N, sizeof(double));
    o->pmin = (double*) calloc(N, sizeof(double));
    o->rho = (double*) malloc(sizeof(double) * N);
    o->ndr = (double*) calloc(N, sizeof(double));
    o->bc = (int32_t*) calloc(N, sizeof(int32_t)); /* Ignored: defined start value */
    o->dconn = (unsigned char*) calloc(N, 1);
    o->nb_cnt = (int32_t*) calloc(N, sizeof(int32_t));
    o->nb_cap = (int32_t*) calloc(N, sizeof(int32_t));
    o->nb = (onbr**) calloc(N, sizeof(onbr*));
    memcpy(o->type, type, sizeof(int32_t) * N);
    memcpy(o->pos, pos, sizeof(double) * 3 * N);
    memcpy(o->vel, vel, sizeof(double) * 3 * N);
    for (size_t i = 0; i < N; i++) o->rho[i] = s->density; /* src/mps.cpp:84-90 */
    /* Bucket::Bucket (src/bucket.cpp:10-22) */
    o->len = s->re_max;
    o->numX = (int) ((s->domain_max[0] - s->domain_min[0]) / o->len) + 3;
    o->numY = (int) ((s->domain_max[1] - s->domain_min[1]) / o->len) + 3;
    o->numZ = (int) ((s->domain_max[2] - s->domain_min[2]) / o->len) + 3;
    o->num = o->numX * o->numY * o->numZ;
    o->first = (int32_t*) malloc(sizeof(int32_t) * (size_t) o->num);
    o->last = (int32_t*) malloc(sizeof(int32_t) * (size_t) o->num);
    o->next = (int32_t*) malloc(sizeof(int32_t) * N);
    o->b = (double*) calloc(N, sizeof(double));
    o->diag = (double*) calloc(N, sizeof(double));
    o->rp = (int32_t*) calloc(N + 1, sizeof(int32_t));
    o->x = (double*) calloc(N, sizeof(double));
    o->snap_pos = (double*) calloc(3 * N, sizeof(double));
    o->snap_vel = (double*) calloc(3 * N, sizeof(double));
    o->tol = DBL_EPSILON;
    o->max_iter = 0;
    o->duplicate = 1;
    return o;
}

void oracle_destroy(oracle_state* o) {
    if (!o) return;
    for (int32_t i = 0; i < o->n; i++) free(o->nb[i]);
    free(o->nb);
    free(o->nb_cnt);
    free(o->nb_cap);
    free(o->type);
    free(o->pos);
    free(o->vel);
    free(o->acc);
    free(o->p);
    free(o->pmin);
    free(o->rho);
    free(o->ndr);
    free(o->bc);
    free(o->dconn);
    free(o->first);
    free(o->last);
    free(o->next);
    free(o->b);
    free(o->diag);
    free(o->rp);
    free(o->cl);
    free(o->vl);
    free(o->x);
    free(o->ev_ids[0]);
    free(o->ev_ids[1]);
    free(o->snap_pos);
    free(o->snap_vel);
    free(o);
}

void oracle_set_solver(oracle_state* o, double tol, int32_t max_iter, int32_t duplicate) {
    o->tol = tol;
    o->max_iter = max_iter;
    o->duplicate = duplicate;
}

/* ---------------- bucket + neighbours ---------------- */

static inline int cell_of(const oracle_state* o, const double* x, int* ix, int* iy, int* iz) {
    *ix = (int) ((x[0] - o->s.domain_min[0]) / o->len) + 1;
    *iy = (int) ((x[1] - o->s.domain_min[1]) / o->len) + 1;
    *iz = (int) ((x[2] - o->s.domain_min[2]) / o->len) + 1;
    return *ix + *iy * o->numX + *iz * o->numX * o->numY;
}

/* src/bucket.cpp:24-51 */
static void store_particles(oracle_state* o) {
    for (int i = 0; i < o->num; i++) {
        o->first[i] = -1;
        o->last[i] = -1;
    }
    for (int i = 0; i < o->n; i++) o->next[i] = -1;
    for (int i = 0; i < o->n; i++) {
        if (o->type[i] == T_GHOST) continue;
        int ix, iy, iz;
        int k = cell_of(o, o->pos + 3 * i, &ix, &iy, &iz);
        if (ix < 1 || iy < 1 || iz < 1 || ix > o->numX - 2 || iy > o->numY - 2 || iz > o->numZ - 2) {
            fprintf(stderr, "oracle: particle %d outside the bucket grid\n", i);
            abort();
        }
        if (o->last[k] == -1)
            o->first[k] = i;
        else
            o->next[o->last[k]] = i;
        o->last[k] = i;
    }
}

static inline void nb_push(oracle_state* o, int32_t i, int32_t j, double dist) {
    if (o->nb_cnt[i] == o->nb_cap[i]) {
        o->nb_cap[i] = o->nb_cap[i] ? 2 * o->nb_cap[i] : 32;
        o->nb[i] = (onbr*) realloc(o->nb[i], sizeof(onbr) * (size_t) o->nb_cap[i]);
    }
    o->nb[i][o->nb_cnt[i]].id = j;
    o->nb[i][o->nb_cnt[i]].dist = dist;
    o->nb_cnt[i]++;
}

/* src/mps.cpp:585-619 */
void oracle_set_neighbors(oracle_state* o) {
    double t0 = now_s();
    store_particles(o);
    const double remax = o->s.re_max;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] == T_GHOST) continue;
        o->nb_cnt[i] = 0;
        const double* xi = o->pos + 3 * i;
        int ix, iy, iz;
        cell_of(o, xi, &ix, &iy, &iz);
        for (int jx = ix - 1; jx <= ix + 1; jx++)
            for (int jy = iy - 1; jy <= iy + 1; jy++)
                for (int jz = iz - 1; jz <= iz + 1; jz++) {
                    int k = jx + jy * o->numX + jz * o->numX * o->numY;
                    int j = o->first[k];
                    while (j != -1) {
                        const double* xj = o->pos + 3 * j;
                        double dist = sqrt(sqnorm3(xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]));
                        if (j != i && dist < remax) nb_push(o, i, j, dist);
                        j = o->next[j];
                    }
                }
    }
    o->secs[0] += now_s() - t0;
}

/* src/mps.cpp:573-583 */
static double number_density(const oracle_state* o, int32_t i, double re) {
    double nd = 0.0;
    for (int k = 0; k < o->nb_cnt[i]; k++) {
        double dist = o->nb[i][k].dist;
        if (dist < re) nd += oracle_weight(dist, re);
    }
    return nd;
}

/* src/mps.cpp:643-667 */
static void check_boundary_violation(oracle_state* o, int32_t i) {
    const double* x = o->pos + 3 * i;
    int in = 1;
    for (int d = 0; d < 3; d++)
        if (x[d] < o->s.domain_min[d] || o->s.domain_max[d] < x[d]) in = 0;
    if (!in) {
#pragma omp critical(oracle_events)
        {
            o->counts[0]++;
            push_event(o, 0, i);
        }
        o->type[i] = T_GHOST;
    }
}

/* src/mps.cpp:627-641 */
static void set_number_density(oracle_state* o) {
    oracle_set_neighbors(o);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->bc[i] == BC_SURFACE)
            o->ndr[i] = 0.0;
        else
            o->ndr[i] = number_density(o, i, o->s.re_pressure) / o->ref_p.n0;
    }
}

void oracle_load_initial_state(oracle_state* o) { set_number_density(o); }

/* ---------------- explicit stages ---------------- */

/* src/mps.cpp:165-175 */
static void calc_gravity(oracle_state* o) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        for (int d = 0; d < 3; d++) o->acc[3 * i + d] = o->type[i] == T_FLUID ? o->s.gravity[d] : 0.0;
    }
}

/* src/mps.cpp:177-203 */
static void calc_viscosity(oracle_state* o) {
    const double re = o->s.re_viscosity, n0 = o->ref_v.n0, lambda = o->ref_v.lambda;
    const double a = (o->s.kinematic_viscosity) * (2.0 * (double) o->s.dim) / (n0 * lambda);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] != T_FLUID) continue;
        double vt[3] = {0.0, 0.0, 0.0};
        const double* vi = o->vel + 3 * i;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            double dist = o->nb[i][k].dist;
            if (dist < re) {
                double w = oracle_weight(dist, re);
                const double* vj = o->vel + 3 * j;
                for (int d = 0; d < 3; d++) vt[d] += (vj[d] - vi[d]) * w;
            }
        }
        for (int d = 0; d < 3; d++) vt[d] *= a;
        for (int d = 0; d < 3; d++) o->acc[3 * i + d] += vt[d];
    }
}

/* src/mps.cpp:205-215 */
static void move_particles(oracle_state* o) {
    const double dt = o->s.dt;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] == T_FLUID) {
            for (int d = 0; d < 3; d++) o->vel[3 * i + d] += o->acc[3 * i + d] * dt;
            for (int d = 0; d < 3; d++) o->pos[3 * i + d] += o->vel[3 * i + d] * dt;
            check_boundary_violation(o, i);
        }
        for (int d = 0; d < 3; d++) o->acc[3 * i + d] = 0.0;
    }
}

/* src/mps.cpp:217-250 (serial, order-dependent) */
static void collision(oracle_state* o) {
    double t0 = now_s();
    const double dc = o->s.collision_distance, e = o->s.restitution;
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] != T_FLUID) continue;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            double dist = o->nb[i][k].dist;
            if (o->type[j] == T_FLUID && j >= i) continue;
            if (dist < dc) {
                double invMi = inverse_density(o, i);
                double invMj = inverse_density(o, j);
                double reduced = 1.0 / (invMi + invMj);
                double* xi = o->pos + 3 * i;
                double* xj = o->pos + 3 * j;
                double* vi = o->vel + 3 * i;
                double* vj = o->vel + 3 * j;
                double nrm[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
                double z = sqnorm3(nrm[0], nrm[1], nrm[2]);
                if (z > 0.0) {
                    double sq = sqrt(z);
                    for (int d = 0; d < 3; d++) nrm[d] = nrm[d] / sq;
                }
                double dv[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
                double rv = dot3(dv, nrm);
                if (rv < 0.0) {
                    double impulse = -(1.0 + e) * rv * reduced;
                    double si = impulse * invMi, sj = impulse * invMj;
                    for (int d = 0; d < 3; d++) vi[d] -= si * nrm[d];
                    for (int d = 0; d < 3; d++) vj[d] += sj * nrm[d];
                    double depth = dc - dist;
                    double pimp = depth * reduced;
                    double pi_ = pimp * invMi, pj_ = pimp * invMj;
                    for (int d = 0; d < 3; d++) xi[d] -= pi_ * nrm[d];
                    for (int d = 0; d < 3; d++) xj[d] += pj_ * nrm[d];
                }
            }
        }
    }
    o->secs[2] += now_s() - t0;
}

/* ---------------- pressure ---------------- */

/* src/mps.cpp:262-278 */
static void set_boundary_condition(oracle_state* o) {
    const double n0 = o->ref_s.n0, beta = o->s.surface_threshold;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] == T_GHOST || o->type[i] == T_DUMMY)
            o->bc[i] = BC_IGNORED;
        else if (number_density(o, i, o->s.re_surface) < beta * n0)
            o->bc[i] = BC_SURFACE;
        else
            o->bc[i] = BC_INNER;
    }
}

/* src/mps.cpp:280-320 */
static void set_source_term(oracle_state* o) {
    const double re = o->s.re_pressure, n0 = o->ref_p.n0, dt = o->s.dt;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->bc[i] == BC_INNER) {
            if (o->s.hos_on) {
                const double gamma = o->s.hos_gamma;
                double bi = gamma * (1.0 / (dt * dt)) * (o->ndr[i] - 1.0);
                const double* xi = o->pos + 3 * i;
                const double* vi = o->vel + 3 * i;
                for (int k = 0; k < o->nb_cnt[i]; k++) {
                    int32_t j = o->nb[i][k].id;
                    if (o->bc[j] == BC_IGNORED) continue;
                    const double* xj = o->pos + 3 * j;
                    const double* vj = o->vel + 3 * j;
                    double rij[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
                    double uij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
                    double dist = o->nb[i][k].dist;
                    if (dist < re) bi -= (1.0 / (n0 * dt)) * re * dot3(rij, uij) / (dist * dist * dist);
                }
                o->b[i] = bi;
            } else {
                const double gamma = o->s.relaxation;
                const double nd = number_density(o, i, re);
                o->b[i] = gamma * (1.0 / (dt * dt)) * (nd - n0) / n0;
            }
        } else {
            o->b[i] = 0.0;
        }
    }
}

static int cmp_col(const void* a, const void* b) {
    const int64_t* x = (const int64_t*) a;
    const int64_t* y = (const int64_t*) b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

/* src/mps.cpp:322-360: off-diagonals in CSR by id, each row sorted by column
 * as setFromTriplets leaves a RowMajor matrix; the diagonal kept apart. */
static void set_matrix(oracle_state* o) {
    const double re = o->s.re_pressure, n0 = o->ref_p.n0, lambda = o->ref_p.lambda;
    const double D = (double) o->s.dim;
    int64_t nnz = 0;
    for (int32_t i = 0; i < o->n; i++) {
        o->rp[i] = (int32_t) nnz;
        o->diag[i] = 0.0;
        if (o->bc[i] != BC_INNER) continue;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            if (o->bc[j] == BC_INNER && o->nb[i][k].dist < re) nnz++;
        }
    }
    o->rp[o->n] = (int32_t) nnz;
    if (nnz > o->nnz_cap) {
        o->nnz_cap = nnz + nnz / 4 + 16;
        o->cl = (int32_t*) realloc(o->cl, sizeof(int32_t) * (size_t) o->nnz_cap);
        o->vl = (double*) realloc(o->vl, sizeof(double) * (size_t) o->nnz_cap);
    }
    int64_t* tmp = (int64_t*) malloc(sizeof(int64_t) * 2 * 256);
    int tmp_cap = 256;
    for (int32_t i = 0; i < o->n; i++) {
        if (o->bc[i] != BC_INNER) continue;
        double cii = 0.0;
        int m = 0;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            double dist = o->nb[i][k].dist;
            if (o->bc[j] == BC_IGNORED) continue;
            if (dist < re) {
                double a = (2.0 * D / (lambda * n0)) * oracle_weight(dist, re) / o->rho[i];
                cii += a;
                if (o->bc[j] == BC_INNER) {
                    if (m == tmp_cap) {
                        tmp_cap *= 2;
                        tmp = (int64_t*) realloc(tmp, sizeof(int64_t) * 2 * (size_t) tmp_cap);
                    }
                    double na = -a;
                    tmp[2 * m] = j;
                    memcpy(&tmp[2 * m + 1], &na, sizeof(double));
                    m++;
                }
            }
        }
        if (o->s.qc_on) {
            const double alpha = o->s.compressibility;
            cii += alpha / (o->s.dt * o->s.dt);
        }
        o->diag[i] = cii;
        qsort(tmp, (size_t) m, 2 * sizeof(int64_t), cmp_col);
        for (int q = 0; q < m; q++) {
            o->cl[o->rp[i] + q] = (int32_t) tmp[2 * q];
            memcpy(&o->vl[o->rp[i] + q], &tmp[2 * q + 1], sizeof(double));
        }
    }
    free(tmp);
}

/* src/mps.cpp:362-415 (serial multi-source BFS) */
static void ensure_dirichlet_connection(oracle_state* o) {
    const double re = o->s.re_pressure;
    for (int32_t i = 0; i < o->n; i++) o->dconn[i] = o->bc[i] == BC_SURFACE;
    int32_t* queue = (int32_t*) malloc(sizeof(int32_t) * (size_t) (o->n + 1));
    for (int32_t s = 0; s < o->n; s++) {
        if (o->bc[s] != BC_SURFACE) continue;
        int32_t head = 0, tail = 0;
        queue[tail++] = s;
        while (head < tail) {
            int32_t i = queue[head++];
            for (int k = 0; k < o->nb_cnt[i]; k++) {
                int32_t j = o->nb[i][k].id;
                if (o->nb[i][k].dist < re && !o->dconn[j]) {
                    o->dconn[j] = 1;
                    queue[tail++] = j;
                }
            }
        }
    }
    free(queue);
    for (int32_t i = 0; i < o->n; i++) {
        if (!o->dconn[i] && o->bc[i] == BC_INNER) {
            o->counts[1]++;
            push_event(o, 1, i);
            o->diag[i] *= 2.0;
        }
    }
}

/* y = A x with A = diag + off-diagonal CSR, each row summed in ascending
 * column order (the diagonal at its sorted place). */
static void spmv(const oracle_state* o, const double* xin, double* y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        double acc = 0.0;
        int placed = 0;
        if (o->bc[i] == BC_INNER) {
            for (int32_t q = o->rp[i]; q < o->rp[i + 1]; q++) {
                if (!placed && o->cl[q] > i) {
                    acc += o->diag[i] * xin[i];
                    placed = 1;
                }
                acc += o->vl[q] * xin[o->cl[q]];
            }
            if (!placed) acc += o->diag[i] * xin[i];
        }
        y[i] = acc;
    }
}

typedef void (*spmv_fn)(const void* ctx, const double* x, double* y);

static int32_t pcg_core(int32_t n, spmv_fn mv, const void* ctx, const double* invdiag,
                        const double* b, double* x, double tol, int32_t max_iter,
                        double* rel_res) {
    /* Eigen ConjugateGradient: internal::conjugate_gradient */
    size_t N = (size_t) n;
    double* r = (double*) malloc(sizeof(double) * N);
    double* p = (double*) malloc(sizeof(double) * N);
    double* z = (double*) malloc(sizeof(double) * N);
    double* t = (double*) malloc(sizeof(double) * N);
    for (size_t i = 0; i < N; i++) x[i] = 0.0;
    for (size_t i = 0; i < N; i++) r[i] = b[i]; /* b - A*0 */
    double rhs2 = 0.0;
    for (size_t i = 0; i < N; i++) rhs2 += b[i] * b[i];
    int32_t it = 0;
    if (rhs2 == 0.0) {
        *rel_res = 0.0;
        goto done;
    }
    {
        const double thr = fmax(tol * tol * rhs2, DBL_MIN);
        double res2 = rhs2;
        if (res2 < thr) {
            *rel_res = sqrt(res2 / rhs2);
            goto done;
        }
        for (size_t i = 0; i < N; i++) p[i] = invdiag[i] * r[i];
        double absNew = 0.0;
        for (size_t i = 0; i < N; i++) absNew += r[i] * p[i];
        while (it < max_iter) {
            mv(ctx, p, t);
            double pt = 0.0;
            for (size_t i = 0; i < N; i++) pt += p[i] * t[i];
            double alpha = absNew / pt;
            for (size_t i = 0; i < N; i++) x[i] += alpha * p[i];
            for (size_t i = 0; i < N; i++) r[i] -= alpha * t[i];
            res2 = 0.0;
            for (size_t i = 0; i < N; i++) res2 += r[i] * r[i];
            if (res2 < thr) break;
            for (size_t i = 0; i < N; i++) z[i] = invdiag[i] * r[i];
            double absOld = absNew;
            absNew = 0.0;
            for (size_t i = 0; i < N; i++) absNew += r[i] * z[i];
            double beta = absNew / absOld;
            for (size_t i = 0; i < N; i++) p[i] = z[i] + beta * p[i];
            it++;
        }
        *rel_res = sqrt(res2 / rhs2);
    }
done:
    free(r);
    free(p);
    free(z);
    free(t);
    return it;
}

static void state_mv(const void* ctx, const double* x, double* y) { spmv((const oracle_state*) ctx, x, y); }

typedef struct {
    int32_t n;
    const int32_t *rp, *cl;
    const double* vl;
} csr_ctx;

static void csr_mv(const void* ctx, const double* x, double* y) {
    const csr_ctx* c = (const csr_ctx*) ctx;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; i++) {
        double acc = 0.0;
        for (int32_t q = c->rp[i]; q < c->rp[i + 1]; q++) acc += c->vl[q] * x[c->cl[q]];
        y[i] = acc;
    }
}

int32_t oracle_pcg(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                   const double* b, double* x, double tol, int32_t max_iter, double* rel_res) {
    double* invdiag = (double*) malloc(sizeof(double) * (size_t) n);
    for (int32_t i = 0; i < n; i++) {
        invdiag[i] = 1.0;
        for (int32_t q = row_ptr[i]; q < row_ptr[i + 1]; q++)
            if (col[q] == i && val[q] != 0.0) invdiag[i] = 1.0 / val[q];
    }
    csr_ctx c = {n, row_ptr, col, val};
    int32_t it = pcg_core(n, csr_mv, &c, invdiag, b, x, tol, max_iter > 0 ? max_iter : 2 * n, rel_res);
    free(invdiag);
    return it;
}

/* src/mps.cpp:417-470 */
static void solve_poisson(oracle_state* o) {
    double t0 = now_s();
    size_t N = (size_t) o->n;
    double* invdiag = (double*) malloc(sizeof(double) * N);
    for (size_t i = 0; i < N; i++) invdiag[i] = (o->bc[i] == BC_INNER && o->diag[i] != 0.0) ? 1.0 / o->diag[i] : 1.0;
    int32_t maxit = o->max_iter > 0 ? o->max_iter : 2 * o->n;
    o->cg_iters = pcg_core(o->n, state_mv, o, invdiag, o->b, o->x, o->tol, maxit, &o->cg_relres);
    if (o->duplicate) o->cg_iters = pcg_core(o->n, state_mv, o, invdiag, o->b, o->x, o->tol, maxit, &o->cg_relres);
    free(invdiag);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) o->p[i] = o->x[i];
    o->secs[4] += now_s() - t0;
}

/* src/mps.cpp:472-479 */
static void remove_negative_pressure(oracle_state* o) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++)
        if (o->p[i] < 0.0) o->p[i] = 0.0;
}

/* src/mps.cpp:481-503 */
static void set_minimum_pressure(oracle_state* o) {
    const double re = o->s.re_pressure;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->bc[i] == BC_IGNORED) continue;
        o->pmin[i] = o->p[i];
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            if (o->type[j] == T_DUMMY) continue;
            if (o->nb[i][k].dist < re)
                if (o->pmin[i] > o->p[j]) o->pmin[i] = o->p[j];
        }
    }
}

/* src/mps.cpp:252-260 */
static void calc_pressure(oracle_state* o) {
    double t0 = now_s();
    set_boundary_condition(o);
    set_source_term(o);
    set_matrix(o);
    ensure_dirichlet_connection(o);
    o->secs[3] += now_s() - t0;
    solve_poisson(o);
    t0 = now_s();
    remove_negative_pressure(o);
    set_minimum_pressure(o);
    o->secs[5] += now_s() - t0;
}

/* src/mps.cpp:505-533 */
static void calc_pressure_gradient(oracle_state* o) {
    const double re = o->s.re_pressure, n0 = o->ref_p.n0;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] != T_FLUID) continue;
        double g[3] = {0.0, 0.0, 0.0};
        const double* xi = o->pos + 3 * i;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            int32_t j = o->nb[i][k].id;
            double dist = o->nb[i][k].dist;
            if (o->type[j] == T_DUMMY) continue;
            if (dist < re) {
                const double* xj = o->pos + 3 * j;
                double dp = o->p[j] - o->pmin[i];
                double w = oracle_weight(dist, re);
                double d2 = dist * dist;
                for (int d = 0; d < 3; d++) g[d] += (xj[d] - xi[d]) * dp * w / d2;
            }
        }
        const double c = (double) o->s.dim / n0;
        for (int d = 0; d < 3; d++) g[d] *= c;
        for (int d = 0; d < 3; d++) o->acc[3 * i + d] = -1.0 * g[d] / o->rho[i];
    }
}

/* src/mps.cpp:535-545 */
static void move_with_pressure_gradient(oracle_state* o) {
    const double dt = o->s.dt;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] == T_FLUID) {
            for (int d = 0; d < 3; d++) o->vel[3 * i + d] += o->acc[3 * i + d] * dt;
            for (int d = 0; d < 3; d++) o->pos[3 * i + d] += o->acc[3 * i + d] * dt * dt;
            check_boundary_violation(o, i);
        }
    }
}

/* src/mps.cpp:547-571 */
static double courant_number(oracle_state* o) {
    double cmax = 0.0;
    for (int32_t i = 0; i < o->n; i++) {
        if (o->type[i] != T_FLUID) continue;
        const double* v = o->vel + 3 * i;
        double c = (sqrt(sqnorm3(v[0], v[1], v[2])) * o->s.dt) / o->s.particle_distance;
        cmax = (cmax < c) ? c : cmax; /* std::max */
    }
    if (cmax > o->s.cfl_condition) o->counts[2]++;
    return cmax;
}

/* src/mps.cpp:111-130 */
double oracle_step(oracle_state* o, int32_t is_export) {
    oracle_set_neighbors(o);
    double t0 = now_s();
    calc_gravity(o);
    calc_viscosity(o);
    move_particles(o);
    o->secs[1] += now_s() - t0;

    oracle_set_neighbors(o);
    collision(o);

    oracle_set_neighbors(o);
    memcpy(o->snap_pos, o->pos, sizeof(double) * 3 * (size_t) o->n);
    memcpy(o->snap_vel, o->vel, sizeof(double) * 3 * (size_t) o->n);
    calc_pressure(o);
    t0 = now_s();
    calc_pressure_gradient(o);
    move_with_pressure_gradient(o);
    o->secs[5] += now_s() - t0;

    if (is_export || o->s.hos_on) set_number_density(o);
    return courant_number(o);
}

/* ---------------- accessors ---------------- */

void oracle_get_state(const oracle_state* o, int32_t* type, double* pos, double* vel, double* p,
                      double* pmin, double* ndr, int32_t* bc) {
    size_t N = (size_t) o->n;
    if (type) memcpy(type, o->type, sizeof(int32_t) * N);
    if (pos) memcpy(pos, o->pos, sizeof(double) * 3 * N);
    if (vel) memcpy(vel, o->vel, sizeof(double) * 3 * N);
    if (p) memcpy(p, o->p, sizeof(double) * N);
    if (pmin) memcpy(pmin, o->pmin, sizeof(double) * N);
    if (ndr) memcpy(ndr, o->ndr, sizeof(double) * N);
    if (bc) memcpy(bc, o->bc, sizeof(int32_t) * N);
}

void oracle_set_state(oracle_state* o, const int32_t* type, const double* pos, const double* vel,
                      const double* p, const double* ndr, const int32_t* bc) {
    size_t N = (size_t) o->n;
    if (type) memcpy(o->type, type, sizeof(int32_t) * N);
    if (pos) memcpy(o->pos, pos, sizeof(double) * 3 * N);
    if (vel) memcpy(o->vel, vel, sizeof(double) * 3 * N);
    if (p) memcpy(o->p, p, sizeof(double) * N);
    if (ndr) memcpy(o->ndr, ndr, sizeof(double) * N);
    if (bc) memcpy(o->bc, bc, sizeof(int32_t) * N);
}

void oracle_debug_counts(const oracle_state* o, int64_t counts[3]) {
    for (int k = 0; k < 3; k++) counts[k] = o->counts[k];
}

int32_t oracle_event_ids(const oracle_state* o, int32_t kind, int32_t* ids, int32_t cap) {
    if (kind < 0 || kind > 1) return 0;
    int32_t m = o->ev_n[kind] < cap ? o->ev_n[kind] : cap;
    if (ids) memcpy(ids, o->ev_ids[kind], sizeof(int32_t) * (size_t) m);
    return o->ev_n[kind];
}

void oracle_bucket_shape(const oracle_state* o, int32_t dims[4], double* length) {
    dims[0] = o->num;
    dims[1] = o->numX;
    dims[2] = o->numY;
    dims[3] = o->numZ;
    *length = o->len;
}

void oracle_dump_bucket(const oracle_state* o, int32_t* first, int32_t* last, int32_t* next) {
    memcpy(first, o->first, sizeof(int32_t) * (size_t) o->num);
    memcpy(last, o->last, sizeof(int32_t) * (size_t) o->num);
    memcpy(next, o->next, sizeof(int32_t) * (size_t) o->n);
}

int64_t oracle_neighbor_total(const oracle_state* o) {
    int64_t t = 0;
    for (int32_t i = 0; i < o->n; i++) t += o->nb_cnt[i];
    return t;
}

void oracle_dump_neighbors(const oracle_state* o, int32_t* offsets, int32_t* ids, double* dists) {
    int64_t t = 0;
    for (int32_t i = 0; i < o->n; i++) {
        offsets[i] = (int32_t) t;
        for (int k = 0; k < o->nb_cnt[i]; k++) {
            if (ids) ids[t] = o->nb[i][k].id;
            if (dists) dists[t] = o->nb[i][k].dist;
            t++;
        }
    }
    offsets[o->n] = (int32_t) t;
}

int64_t oracle_offdiag_nnz(const oracle_state* o) { return o->rp[o->n]; }

void oracle_get_pressure_system(const oracle_state* o, double* b, double* diag, int32_t* row_ptr,
                                int32_t* col, double* val, double* x) {
    size_t N = (size_t) o->n;
    if (b) memcpy(b, o->b, sizeof(double) * N);
    if (diag) memcpy(diag, o->diag, sizeof(double) * N);
    if (row_ptr) memcpy(row_ptr, o->rp, sizeof(int32_t) * (N + 1));
    if (col) memcpy(col, o->cl, sizeof(int32_t) * (size_t) o->rp[o->n]);
    if (val) memcpy(val, o->vl, sizeof(double) * (size_t) o->rp[o->n]);
    if (x) memcpy(x, o->x, sizeof(double) * N);
}

void oracle_get_snapshot(const oracle_state* o, double* pos, double* vel) {
    size_t N = (size_t) o->n;
    if (pos) memcpy(pos, o->snap_pos, sizeof(double) * 3 * N);
    if (vel) memcpy(vel, o->snap_vel, sizeof(double) * 3 * N);
}

void oracle_cg_stats(const oracle_state* o, int32_t* iterations, double* rel_residual) {
    *iterations = o->cg_iters;
    *rel_residual = o->cg_relres;
}

void oracle_stage_seconds(const oracle_state* o, double out[6]) {
    for (int k = 0; k < 6; k++) out[k] = o->secs[k];
}
