// adjoint.hip — reverse-mode kernels of the MPS step (MPS::stepForward,
// src/mps.cpp:111-130): for every stage, the transpose of its Jacobian with
// the forward step's discrete decisions held fixed (neighbour sets, boundary
// conditions, Dirichlet connectivity, which collision pairs fired, the
// negative-pressure clamp, the minimum-pressure argument, ghosting).  The
// adjoint pressure Poisson equation A^T lambda = dL/dp is solved by the same
// symmetric CG as the forward one (A is symmetric).
//
// Adjoint accumulators (xb, vb: dL/dx, dL/dv) are in particle id order.  The
// pressure-stage kernels run on the slots of the forward step's third bucket
// build; the viscosity-stage kernels on a rebuild of the step's start state.
// Every kernel gathers: a thread writes only its own particle, so sums are
// deterministic and no atomics are used.
//
// Derivatives used (r = |x_j - x_i|, d = x_j - x_i, w(r) = re/r - 1):
//   w'(r) = -re / r^2
//   phi(r) = w(r) / r^2,  phi'(r) / r = 2 / r^4 - 3 re / r^5
//   dr/dx_j = d / r = -dr/dx_i
//
// No multiply-add is fused here, as in the forward stages: the neighbour
// walks recompute distances, and a pair within an ulp of a radius must take
// the same side of dist < re as the recorded forward run did.
#pragma clang fp contract(off)

#include "mps_internal.hpp"

namespace mps {

static inline dim3 grid_for(int n, int bs) { return dim3((unsigned) ((n + bs - 1) / bs)); }

#define LAUNCH_D(kern, grid, block, st, ...)                                  \
    do {                                                                      \
        if (dim == 2)                                                         \
            hipLaunchKernelGGL((kern<2>), grid, block, 0, st, __VA_ARGS__);   \
        else                                                                  \
            hipLaunchKernelGGL((kern<3>), grid, block, 0, st, __VA_ARGS__);   \
    } while (0)

__device__ __forceinline__ double wprime(double dist, double re) { return -re / (dist * dist); }

// ---------------- device state save / restore (id order) ----------------
//
// A saved state holds, per particle id: x y z vx vy vz (z and vz are the
// frozen plane values in 2D), the warm-start solution, pressure, minimum
// pressure, number density ratio, a ghost's stale number density and the
// solution before the warm-start one (12 n doubles), then type and boundary
// condition (2 n bytes).  Restoring writes
// slot = id, the layout of a fresh handle; the next bucket build sorts it, and
// the build's result does not depend on the slot order it starts from.

__global__ void __launch_bounds__(kBlock) k_save_state(int n, int dim, Hot h, Cold cold, double* __restrict__ f,
                                                       int8_t* __restrict__ t8) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t N = (size_t) n;
    const int id = h.id[s];
    f[id] = h.px[s];
    f[N + id] = h.py[s];
    f[2 * N + id] = dim == 3 ? h.pz[s] : cold.z[id];
    f[3 * N + id] = h.vx[s];
    f[4 * N + id] = h.vy[s];
    f[5 * N + id] = dim == 3 ? h.vz[s] : cold.vz[id];
    f[6 * N + id] = cold.pprev[id];
    f[7 * N + id] = cold.p[id];
    f[8 * N + id] = cold.pmin[id];
    f[9 * N + id] = cold.ndr[id];
    f[10 * N + id] = cold.ghost_n[id];
    f[11 * N + id] = cold.pprev2[id];
    t8[id] = h.type[s];
    t8[N + id] = cold.bc[id];
}

__global__ void __launch_bounds__(kBlock) k_restore_state(int n, int dim, Hot h, Cold cold, int* __restrict__ slot_of,
                                                          const double* __restrict__ f,
                                                          const int8_t* __restrict__ t8) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t N = (size_t) n;
    h.px[i] = f[i];
    h.py[i] = f[N + i];
    h.vx[i] = f[3 * N + i];
    h.vy[i] = f[4 * N + i];
    if (dim == 3) {
        h.pz[i] = f[2 * N + i];
        h.vz[i] = f[5 * N + i];
    } else {
        cold.z[i] = f[2 * N + i];
        cold.vz[i] = f[5 * N + i];
    }
    cold.pprev[i] = f[6 * N + i];
    cold.p[i] = f[7 * N + i];
    cold.pmin[i] = f[8 * N + i];
    cold.ndr[i] = f[9 * N + i];
    cold.ghost_n[i] = f[10 * N + i];
    cold.pprev2[i] = f[11 * N + i];
    h.id[i] = i;
    h.type[i] = t8[i];
    cold.bc[i] = t8[N + i];
    slot_of[i] = i;
}

hipError_t save_state(int dim, int n, const Hot& h, const Cold& cold, double* f, int8_t* t8, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_save_state, grid_for(n, kBlock), dim3(kBlock), 0, st, n, dim, h, cold, f, t8);
    return hipGetLastError();
}

hipError_t restore_state(int dim, int n, Hot& h, Cold& cold, int* slot_of, const double* f, const int8_t* t8,
                         hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_restore_state, grid_for(n, kBlock), dim3(kBlock), 0, st, n, dim, h, cold, slot_of, f, t8);
    return hipGetLastError();
}

// ---------------- deterministic sums ----------------

constexpr int kSumChunk = 8 * kBlock;

__device__ __forceinline__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

__global__ void __launch_bounds__(kBlock) k_sum_partials(const double* __restrict__ v, int n,
                                                         double* __restrict__ part) {
    __shared__ double red[kBlock];
    const size_t base = (size_t) blockIdx.x * kSumChunk;
    double t = 0.0;
    for (int k = 0; k < kSumChunk / kBlock; k++) {
        const size_t i = base + (size_t) k * kBlock + threadIdx.x;
        if (i < (size_t) n) t += v[i];
    }
    const double s = block_sum(t, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_sum_final(const double* __restrict__ part, int nb,
                                                      double* __restrict__ acc) {
    __shared__ double red[kBlock];
    double t = 0.0;
    for (int i = threadIdx.x; i < nb; i += kBlock) t += part[i];
    const double s = block_sum(t, red);
    if (threadIdx.x == 0) acc[0] += s;
}

// acc[0] += sum(v[0..n)), in a fixed order.
static hipError_t sum_into(const double* v, int n, double* part, double* acc, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int nb = (n + kSumChunk - 1) / kSumChunk;
    hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(kBlock), 0, st, v, n, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, part, nb, acc);
    return hipGetLastError();
}

size_t adj_partials(int n) { return (size_t) (n + kSumChunk - 1) / kSumChunk + 1; }

// ---------------- pressure gradient + implicit move (src/mps.cpp:505-545) ----------------
//
// Forward, fluid i: g_i = sum_j d_ij (p_j - phat_i) phi(r_ij) over j not dummy
// wall with r < re_p; a_i = -(D/n0) g_i / rho; v += a dt; x += a dt^2.
// Adjoint: abar_i = vbar_i dt + xbar_i dt^2, gbar_i = -(D/n0)/rho abar_i.

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_grad_pre(GridView g, Consts c, int n, Hot h,
                                                         const int8_t* __restrict__ tg,
                                                         const int8_t* __restrict__ bc_s,
                                                         const double* __restrict__ p_s, AdjBufs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    const double k = -(c.gscale / c.rho);
    for (int d = 0; d < DIM; d++) {
        double v = 0.0;
        if (tg[s] == MPS_FLUID) v = k * (a.vb[d][id] * c.dt + a.xb[d][id] * (c.dt * c.dt));
        a.gb[d][s] = v;
    }
    // argument of the minimum pressure (src/mps.cpp:487-500): the first
    // strict improvement in list order, starting from p_i itself
    int m = -1;
    if (bc_s[s] != MPS_BC_IGNORED) {
        double pm = p_s[s];
        m = s;
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, c.re_p,
                           [&](int q, double, double, double, double) {
                               if (tg[q] == MPS_DUMMY_WALL) return;
                               if (pm > p_s[q]) {
                                   pm = p_s[q];
                                   m = q;
                               }
                           });
    }
    a.amin[s] = m;
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_grad(GridView g, Consts c, int n, Hot h, const int8_t* __restrict__ tg,
                                                     const double* __restrict__ p_s,
                                                     const double* __restrict__ pmin_s, AdjBufs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    const bool fl = tg[s] == MPS_FLUID;
    const bool notdummy = tg[s] != MPS_DUMMY_WALL;
    double gs[3] = {0.0, 0.0, 0.0};
    if (fl)
        for (int d = 0; d < DIM; d++) gs[d] = a.gb[d][s];
    const double ps = p_s[s], pmi = pmin_s[s];
    const double re = c.re_p;
    double xacc[3] = {0.0, 0.0, 0.0};
    double pbar = 0.0, phbar = 0.0;
    const double zi = DIM == 3 ? h.pz[s] : 0.0;
    if (tg[s] != MPS_GHOST)  // ghosts at the build sit outside every cell
    for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                       [&](int q, double r, double dx, double dy, double dz) {
                           const int tq = tg[q];
                           const double dd[3] = {dx, dy, dz};
                           const double r2 = r * r;
                           const double phi = (re / r - 1.0) / r2;
                           const double dphi_r = 2.0 / (r2 * r2) - 3.0 * re / (r2 * r2 * r);
                           if (fl && tq != MPS_DUMMY_WALL) {  // own row
                               double gd = 0.0;
                               for (int d = 0; d < DIM; d++) gd += gs[d] * dd[d];
                               const double dp = p_s[q] - pmi;
                               for (int d = 0; d < DIM; d++) xacc[d] -= dp * (phi * gs[d] + gd * dphi_r * dd[d]);
                               phbar -= gd * phi;
                           }
                           if (notdummy && tq == MPS_FLUID) {  // as a neighbour of fluid q
                               double gq[3], gd = 0.0;
                               for (int d = 0; d < DIM; d++) {
                                   gq[d] = a.gb[d][q];
                                   gd += gq[d] * dd[d];
                               }
                               const double dp = ps - pmin_s[q];
                               for (int d = 0; d < DIM; d++) xacc[d] += dp * (phi * gq[d] + gd * dphi_r * dd[d]);
                               pbar -= gd * phi;
                           }
                       });
    for (int d = 0; d < DIM; d++) a.xb[d][id] += xacc[d];
    a.pb[s] = pbar;
    a.phb[s] = phbar;
}

// minimum pressure + clamp (src/mps.cpp:472-503): pbar_{argmin_i} += phatbar_i,
// then the adjoint right-hand side on the Inner rows where the solution was
// not clamped.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_minp(GridView g, Consts c, int n, Hot h,
                                                     const int8_t* __restrict__ tg,
                                                     const int8_t* __restrict__ bc_s, const double* __restrict__ x_s,
                                                     AdjBufs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double v = a.pb[s];
    if (a.amin[s] == s) v += a.phb[s];
    if (tg[s] != MPS_GHOST) {
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, c.re_p,
                           [&](int q, double, double, double, double) {
                               if (a.amin[q] == s) v += a.phb[q];
                           });
    }
    a.rhs[s] = (bc_s[s] == MPS_BC_INNER && !(x_s[s] < 0.0)) ? v : 0.0;
}

// ---------------- source term + matrix (src/mps.cpp:280-360, 412) ----------------
//
// Inner row i: b_i = relax/dt^2 (n_i - n0)/n0 with n_i = sum_j w(r_ij) over the
// list within re_p; A_ii = delta_i (sum_j a_ij + alpha/dt^2) over j not Ignored,
// A_ij = -a_ij for Inner j, a_ij = coefA w(r_ij) / rho, delta_i = 2 for an
// Inner particle without a path to a Surface particle.  With lambda = A^-1 pbar
// and p the unclamped solution (0 off the Inner rows):
//   bbar_i = lambda_i,  abar_ij = lambda_i (p_j - delta_i p_i).

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_ppe(GridView g, Consts c, int n, Hot h,
                                                    const int8_t* __restrict__ tg,
                                                    const int8_t* __restrict__ bc_s, const double* __restrict__ x_s,
                                                    const int* __restrict__ parent, const int8_t* __restrict__ surf,
                                                    AdjBufs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    const int8_t bs = bc_s[s];
    const bool inner = bs == MPS_BC_INNER;
    const double ls = inner ? a.lam[s] : 0.0;
    const double prs = inner ? x_s[s] : 0.0;
    const double ds = (inner && !surf[parent[s]]) ? 2.0 : 1.0;
    const double nbs = ls * c.src_scale / c.n0_p;
    const double cA = c.coefA / c.rho;
    const double re = c.re_p;
    double R[3] = {0.0, 0.0, 0.0};
    double nd = 0.0;
    if (tg[s] != MPS_GHOST) {
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                           [&](int q, double r, double dx, double dy, double dz) {
                               const double wp = wprime(r, re);
                               const int8_t bq = bc_s[q];
                               double rb = 0.0;
                               if (inner) {  // own row
                                   nd += re / r - 1.0;
                                   rb += nbs * wp;
                                   if (bq != MPS_BC_IGNORED) {
                                       const double prq = bq == MPS_BC_INNER ? x_s[q] : 0.0;
                                       rb += ls * (prq - ds * prs) * cA * wp;
                                   }
                               }
                               if (bq == MPS_BC_INNER) {  // as a neighbour in row q
                                   const double lq = a.lam[q];
                                   const double prq = x_s[q];
                                   const double dq = surf[parent[q]] ? 1.0 : 2.0;
                                   rb += lq * c.src_scale / c.n0_p * wp;
                                   if (bs != MPS_BC_IGNORED) rb += lq * (prs - dq * prq) * cA * wp;
                               }
                               const double k = rb / r;
                               R[0] += k * dx;
                               R[1] += k * dy;
                               if (DIM == 3) R[2] += k * dz;
                           });
    }
    for (int d = 0; d < DIM; d++) a.xb[d][id] -= R[d];
    a.crelax[s] = inner ? ls * (1.0 / (c.dt * c.dt)) * (nd - c.n0_p) / c.n0_p : 0.0;
}

// ---------------- collision (src/mps.cpp:217-250) ----------------
//
// Pairs of one level touch disjoint fluid particles, so levels are reversed
// one at a time and a level's pairs run in parallel.  A pair's values before
// its update come from the forward step's record.  A wall partner never
// changes: its adjoint is a plain sum of the pairs' direct terms, added after
// the sweep together with the terms through the build-time distance (which
// depends on the positions before any pair acted).

constexpr int kAdjCollThreads = 1024;
constexpr int kPairScratch = 8;  // per pair: dist-bar, wall x-bar (3), wall v-bar (3), e-bar

__global__ void __launch_bounds__(kAdjCollThreads) k_adj_coll(int dim, Consts c, const int* __restrict__ np_dev,
                                                              int cap, const int* __restrict__ level,
                                                              const double* __restrict__ pdist, CollTape tp,
                                                              AdjBufs a) {
    int np = *np_dev;
    if (np > cap) np = cap;
    if (np == 0) return;
    const int nlev = *tp.nlev;
    const int D = dim;
    volatile double* xb[3] = {a.xb[0], a.xb[1], a.xb[2]};
    volatile double* vb[3] = {a.vb[0], a.vb[1], a.vb[2]};
    const double e = c.restitution, dc = c.dc;
    for (int L = nlev; L >= 1; L--) {
        for (int p = threadIdx.x; p < np; p += blockDim.x) {
            if (level[p] != L) continue;
            const int fl = tp.flags[p];
            double* cp = a.cpair + (size_t) kPairScratch * p;
            if (!(fl & 1)) {
                for (int k = 0; k < kPairScratch; k++) cp[k] = 0.0;
                continue;
            }
            const bool jf = (fl & 2) != 0;
            const int i = tp.pid[2 * p], j = tp.pid[2 * p + 1];
            const double* pre = tp.pre + 12 * (size_t) p;
            const double invMi = 1 / c.rho;
            const double invMj = jf ? 1 / c.rho : 0;
            const double m = 1.0 / (invMi + invMj);
            double nrm[3], dv[3];
            double zz = 0.0;
            for (int d = 0; d < 3; d++) {
                nrm[d] = pre[3 + d] - pre[d];
                dv[d] = pre[9 + d] - pre[6 + d];
                zz += nrm[d] * nrm[d];
            }
            const double sq = sqrt(zz);
            for (int d = 0; d < 3; d++) nrm[d] /= sq;
            double rv = 0.0;
            for (int d = 0; d < 3; d++) rv += dv[d] * nrm[d];
            const double J = -(1.0 + e) * rv * m;
            const double si = J * invMi, sj = J * invMj;
            const double P = (dc - pdist[p]) * m;
            const double pi_ = P * invMi, pj_ = P * invMj;
            double Xi[3] = {0, 0, 0}, Xj[3] = {0, 0, 0}, Vi[3] = {0, 0, 0}, Vj[3] = {0, 0, 0};
            for (int d = 0; d < D; d++) {
                Xi[d] = xb[d][i];
                Vi[d] = vb[d][i];
                if (jf) {
                    Xj[d] = xb[d][j];
                    Vj[d] = vb[d][j];
                }
            }
            double vin = 0.0, vjn = 0.0, xin = 0.0, xjn = 0.0;
            for (int d = 0; d < 3; d++) {
                vin += Vi[d] * nrm[d];
                vjn += Vj[d] * nrm[d];
                xin += Xi[d] * nrm[d];
                xjn += Xj[d] * nrm[d];
            }
            const double Jb = -vin * invMi + vjn * invMj;
            const double Pb = -xin * invMi + xjn * invMj;
            const double rvb = -(1.0 + e) * m * Jb;
            double nb[3];
            for (int d = 0; d < 3; d++) nb[d] = -si * Vi[d] + sj * Vj[d] - pi_ * Xi[d] + pj_ * Xj[d] + rvb * dv[d];
            double nn = 0.0;
            for (int d = 0; d < 3; d++) nn += nb[d] * nrm[d];
            double db[3];
            for (int d = 0; d < 3; d++) db[d] = (nb[d] - nn * nrm[d]) / sq;
            for (int d = 0; d < D; d++) {
                xb[d][i] = Xi[d] - db[d];
                vb[d][i] = Vi[d] - rvb * nrm[d];
                if (jf) {
                    xb[d][j] = Xj[d] + db[d];
                    vb[d][j] = Vj[d] + rvb * nrm[d];
                }
            }
            cp[0] = -m * Pb;  // dL/d(build-time distance)
            for (int d = 0; d < 3; d++) {
                cp[1 + d] = jf ? 0.0 : db[d];
                cp[4 + d] = jf ? 0.0 : rvb * nrm[d];
            }
            cp[7] = -rv * m * Jb;  // dL/de
        }
        __threadfence();
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double esum = 0.0;
        for (int p = 0; p < np; p++) {
            if (!(tp.flags[p] & 1)) continue;
            const double* cp = a.cpair + (size_t) kPairScratch * p;
            const int i = tp.pid[2 * p], j = tp.pid[2 * p + 1];
            const double* u = tp.dir + 3 * (size_t) p;
            for (int d = 0; d < D; d++) {
                xb[d][i] = xb[d][i] - cp[0] * u[d];
                xb[d][j] = xb[d][j] + cp[0] * u[d] + cp[1 + d];
                vb[d][j] = vb[d][j] + cp[4 + d];
            }
            esum += cp[7];
        }
        a.acc[1] += esum;
    }
}

// ---------------- gravity + viscosity + explicit move (src/mps.cpp:165-215) ----------------

// v1 = v0 + a dt, x1 = x0 + v1 dt (fluid): vbar0 = vbar1 + xbar1 dt, abar = vbar0 dt.
__global__ void __launch_bounds__(kBlock) k_adj_move1(int n, int dim, double dt, const int8_t* __restrict__ type0,
                                                      AdjBufs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool fl = type0[i] == MPS_FLUID;
    for (int d = 0; d < dim; d++) {
        if (fl) {
            const double vt = a.vb[d][i] + a.xb[d][i] * dt;
            a.vb[d][i] = vt;
            a.ab[d][i] = vt * dt;
        } else {
            a.ab[d][i] = 0.0;
        }
    }
}

// Fluid i: a_i = g + avisc sum_j (v_j - v_i) w(r_ij) over r < re_v, avisc =
// nu 2D / (n0_v lambda_v).  h is the start state rebuilt into slots.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_visc(GridView g, Consts c, int n, Hot h, AdjBufs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    const bool fl = h.type[s] == MPS_FLUID;
    double as_[3] = {0.0, 0.0, 0.0};
    if (fl)
        for (int d = 0; d < DIM; d++) as_[d] = a.ab[d][id];
    const double vs[3] = {h.vx[s], h.vy[s], DIM == 3 ? h.vz[s] : 0.0};
    const double re = c.re_v, av = c.avisc;
    double xacc[3] = {0.0, 0.0, 0.0}, vacc[3] = {0.0, 0.0, 0.0};
    double num = 0.0;
    if (h.type[s] != MPS_GHOST) {
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                           [&](int q, double r, double dx, double dy, double dz) {
                               const double w = re / r - 1.0;
                               const double wp = wprime(r, re);
                               const double dd[3] = {dx, dy, dz};
                               const double dv[3] = {h.vx[q] - vs[0], h.vy[q] - vs[1],
                                                     DIM == 3 ? h.vz[q] - vs[2] : 0.0};
                               double rb = 0.0;
                               if (fl) {  // own row
                                   double ad = 0.0;
                                   for (int d = 0; d < DIM; d++) {
                                       vacc[d] -= av * w * as_[d];
                                       ad += as_[d] * dv[d];
                                   }
                                   rb += av * ad * wp;
                                   num += ad * w;
                               }
                               if (h.type[q] == MPS_FLUID) {  // as a neighbour of fluid q
                                   const int iq = h.id[q];
                                   double ad = 0.0;
                                   for (int d = 0; d < DIM; d++) {
                                       const double aq = a.ab[d][iq];
                                       vacc[d] += av * w * aq;
                                       ad -= aq * dv[d];
                                   }
                                   rb += av * ad * wp;
                               }
                               const double k = rb / r;
                               for (int d = 0; d < DIM; d++) xacc[d] -= k * dd[d];
                           });
    }
    for (int d = 0; d < DIM; d++) {
        a.xb[d][id] += xacc[d];
        a.vb[d][id] += vacc[d];
    }
    a.cnu[s] = fl ? num * ((2.0 * (double) DIM) / (c.n0_v * c.lam_v)) : 0.0;
}

// ---------------- launchers ----------------

hipError_t adj_pressure_stage(int dim, const GridView& g, int n, const Consts& c, const Hot& h,
                              const int8_t* type_g, const StepScratch& w, const double* x_s, AdjBufs& a,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 gr = grid_for(n, kBlock), bl(kBlock);
    LAUNCH_D(k_adj_grad_pre, gr, bl, st, g, c, n, h, type_g, w.bc_s, w.p_s, a);
    LAUNCH_D(k_adj_grad, gr, bl, st, g, c, n, h, type_g, w.p_s, w.pmin_s, a);
    LAUNCH_D(k_adj_minp, gr, bl, st, g, c, n, h, type_g, w.bc_s, x_s, a);
    return hipGetLastError();
}

hipError_t adj_ppe_stage(int dim, const GridView& g, int n, const Consts& c, const Hot& h, const int8_t* type_g,
                         const StepScratch& w, const double* x_s, AdjBufs& a, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 gr = grid_for(n, kBlock), bl(kBlock);
    LAUNCH_D(k_adj_ppe, gr, bl, st, g, c, n, h, type_g, w.bc_s, x_s, w.parent, w.surf, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sum_into(a.crelax, n, a.part, a.acc + 2, st);
}

hipError_t adj_collision_stage(int dim, int n, const Consts& c, const StepScratch& w, const CollTape& tp, AdjBufs& a,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_coll, dim3(1), dim3(kAdjCollThreads), 0, st, dim, c, w.coll_off + n, w.coll_cap,
                       w.coll_level, w.coll_dist, tp, a);
    return hipGetLastError();
}

hipError_t adj_explicit_stage(int dim, const GridView& g, int n, const Consts& c, const int8_t* type0_id,
                              const Hot& hs, AdjBufs& a, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 gr = grid_for(n, kBlock), bl(kBlock);
    hipLaunchKernelGGL(k_adj_move1, gr, bl, 0, st, n, dim, c.dt, type0_id, a);
    LAUNCH_D(k_adj_visc, gr, bl, st, g, c, n, hs, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sum_into(a.cnu, n, a.part, a.acc + 0, st);
}

size_t adj_pair_scratch(int cap) { return (size_t) kPairScratch * (size_t) cap; }

}  // namespace mps
