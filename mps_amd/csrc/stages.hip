// stages.hip — the explicit and pressure-assembly stages of the MPS step on
// the cell-sorted layout, FP64, in the reference's expression order.
//
// Floating-point contraction is off in this file: every sum, product and
// quotient rounds as the reference's scalar C++ does (x86-64 SSE2, no FMA), so
// the stages up to the pressure solve reproduce the reference bit for bit.
#pragma clang fp contract(off)

#include <algorithm>
#include <type_traits>

#include "mps_internal.hpp"

namespace mps {

static inline dim3 grid_for(int n, int bs) { return dim3((unsigned) ((n + bs - 1) / bs)); }

// Launches kern<DIM, MULTI> for the handle's dimension and scene layout.
// The single-scene 2D walks that stage candidates in LDS run at six waves
// per SIMD (at most 80 VGPRs) instead of the four or five their free
// register use gives: measured at 2M, k_visc_move 292 -> 272 us,
// k_bc_source 190 -> 185 us, k_coll_count 72 -> 67 us.  k_grad_move spills
// at that bound and runs slower (210 -> 225 us), and so would the ensemble
// variants (their per-scene constants take registers): they keep their own.
#define MPS_WALK_WAVES __attribute__((amdgpu_waves_per_eu(DIM == 2 && !MULTI ? 6 : 1)))

#define LAUNCH_DM(kern, grid, block, st, ...)                                                   \
    do {                                                                                        \
        if (dim == 2) {                                                                         \
            if (tab.g)                                                                          \
                hipLaunchKernelGGL((kern<2, true>), grid, block, 0, st, __VA_ARGS__);           \
            else                                                                                \
                hipLaunchKernelGGL((kern<2, false>), grid, block, 0, st, __VA_ARGS__);          \
        } else {                                                                                \
            if (tab.g)                                                                          \
                hipLaunchKernelGGL((kern<3, true>), grid, block, 0, st, __VA_ARGS__);           \
            else                                                                                \
                hipLaunchKernelGGL((kern<3, false>), grid, block, 0, st, __VA_ARGS__);          \
        }                                                                                       \
    } while (0)

__device__ __forceinline__ void push_event(int* counter, int* ids, int cap, int id) {
    int k = atomicAdd(counter, 1);
    if (k < cap) ids[k] = id;
}

// src/mps.cpp:643-667 (strict comparisons on every axis)
__device__ __forceinline__ bool in_domain(const Consts& c, double x, double y, double z) {
    return !(x < c.dmin[0] || c.dmax[0] < x || y < c.dmin[1] || c.dmax[1] < y || z < c.dmin[2] ||
             c.dmax[2] < z);
}

// Number density over the current list within re_p, kept for a particle that
// turns Ghost: the reference leaves a Ghost's list stale (src/mps.cpp:589-593)
// and setNumberDensity later reads it (src/mps.cpp:630-640).
template <int DIM>
__device__ __forceinline__ double ndens(const GridView& g, const Hot& h, int s, double xi, double yi, double zi,
                                        double re) {
    double nd = 0.0;
    for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, xi, yi, zi, re,
                       [&](int, double dist, double, double, double) { nd += wfun(dist, re); });
    return nd;
}

// LDS staging of a block's candidate positions (see nb_stage, nb_stage_f):
// capacity in slots per block; every thread of the block declares and fills
// it first.  2D: FP64, 3 ranges of ~280 slots for 256-thread blocks (24 KB).
// 3D: FP32 offsets, 9 ranges of ~330 slots: 3392 slots (40.7 KB with the
// metadata) is the most that still lets four 256-thread blocks share a CU's
// 160 KB; on the 3.6M-particle dam break at t = 0 one block in ten needs
// more and walks global memory (at 3072 slots, four in ten did).  In FP64
// (72 KB) two blocks fit a CU and the walks were slower than from global
// memory.
template <int DIM>
struct NbStageOf {
    using type = NbStage<2>;
};
template <>
struct NbStageOf<3> {
    using type = NbStageF;
};
template <int DIM>
constexpr int kStageCap = DIM == 2 ? 1536 : 3392;
template <int DIM>
using StageT = typename std::conditional<DIM == 2, double, float>::type;

// Staging counts for tests (see stage_stats_read); only touched when the
// grid's stage_stats flag is set, one atomic per block.
static __device__ unsigned long long g_stage_stats[4];

__device__ __forceinline__ void stage_stats_record(int mask, int full) {
    atomicAdd(&g_stage_stats[0], 1ull);
    if (mask == full) {
        atomicAdd(&g_stage_stats[1], 1ull);
    } else if (mask != 0) {
        atomicAdd(&g_stage_stats[2], 1ull);
        const int top = 31 - __clz(mask);           // highest staged row
        if (mask != (1 << (top + 1)) - 1) atomicAdd(&g_stage_stats[3], 1ull);  // a row left out below it
    }
}

hipError_t stage_stats_read(long long out[4]) {
    unsigned long long v[4];
    const hipError_t e = hipMemcpyFromSymbol(v, HIP_SYMBOL(g_stage_stats), sizeof v, 0, hipMemcpyDeviceToHost);
    for (int k = 0; k < 4; k++) out[k] = e == hipSuccess ? (long long) v[k] : 0;
    return e;
}

template <int DIM>
__device__ __forceinline__ typename NbStageOf<DIM>::type nb_stage_any(const GridView& g, const SceneTab* tab,
                                                                      const Hot& h, int n, int* meta,
                                                                      StageT<DIM>* sx, StageT<DIM>* sy,
                                                                      StageT<DIM>* sz) {
    const int cap = kStageCap<DIM> < g.stage_cap ? kStageCap<DIM> : g.stage_cap;
    if constexpr (DIM == 2) {
        const auto st = nb_stage<2>(g, tab, h, n, cap, meta, sx, sy, sz);
        if (g.stage_stats && threadIdx.x == 0) stage_stats_record(st.ok ? 1 : 0, 1);
        return st;
    } else {
        const auto st = nb_stage_f(g, tab, h, n, cap, meta, sx, sy, sz);
        if (g.stage_stats && threadIdx.x == 0) stage_stats_record(st.ok, 0x1ff);
        return st;
    }
}

#define NB_STAGE(DIM, GRID, HOT, N)                                                                        \
    __shared__ int s_nb_meta[nb_stage_meta_ints<DIM>()];                                                 \
    __shared__ StageT<DIM> s_nb_x[kStageCap<DIM>];                                                       \
    __shared__ StageT<DIM> s_nb_y[kStageCap<DIM>];                                                       \
    __shared__ StageT<DIM> s_nb_z[DIM == 3 ? kStageCap<DIM> : 1];                                        \
    const auto nbs = nb_stage_any<DIM>(GRID, MULTI ? &tab : nullptr, HOT, N, s_nb_meta, s_nb_x, s_nb_y, s_nb_z)

// ---------------- gravity + viscosity + explicit move (src/mps.cpp:165-215) ----------------

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) MPS_WALK_WAVES k_visc_move(GridView g0, Consts c0, SceneTab tab, int n, Hot cur, Hot out, Cold cold,
                                                      Events ev) {
    NB_STAGE(DIM, g0, cur, n);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, cur.id[s], g, c);
    double x = cur.px[s], y = cur.py[s], z = DIM == 3 ? cur.pz[s] : 0.0;
    double vx = cur.vx[s], vy = cur.vy[s], vz = DIM == 3 ? cur.vz[s] : 0.0;
    if (cur.type[s] == MPS_FLUID) {
        double ax = c.g[0], ay = c.g[1], az = c.g[2];
        double tx = 0.0, ty = 0.0, tz = 0.0;
        const double re = c.re_v;
        // 3D: the grouped loop in batches of 3 (of 4 the kernel needs more
        // than 128 registers, a wave fewer per SIMD, and runs 6 % slower;
        // of 3 it runs 9 % faster than cell by cell)
        for_neighbors<DIM, 3>(nbs, g, cur.px, cur.py, cur.pz, cur.key[s], s, x, y, z, re,
                           [&](int q, double dist, double, double, double) {
                               const double w = wfun(dist, re);
                               tx += (cur.vx[q] - vx) * w;
                               ty += (cur.vy[q] - vy) * w;
                               if (DIM == 3) tz += (cur.vz[q] - vz) * w;
                           });
        tx *= c.avisc;
        ty *= c.avisc;
        tz *= c.avisc;
        ax += tx;
        ay += ty;
        az += tz;
        vx += ax * c.dt;
        vy += ay * c.dt;
        if (DIM == 3) vz += az * c.dt;
        const double x0 = x, y0 = y, z0 = z;
        x += vx * c.dt;
        y += vy * c.dt;
        if (DIM == 3) z += vz * c.dt;
        if (!in_domain(c, x, y, DIM == 3 ? z : c.dmin[2])) {
            cur.type[s] = MPS_GHOST;
            const int id = cur.id[s];
            push_event(ev.n_ghost, ev.ghost_ids, ev.cap, id);
            cold.ghost_n[id] = ndens<DIM>(g, cur, s, x0, y0, z0, c.re_p);
        }
    }
    out.px[s] = x;
    out.py[s] = y;
    out.vx[s] = vx;
    out.vy[s] = vy;
    if (DIM == 3) {
        out.pz[s] = z;
        out.vz[s] = vz;
    }
}

hipError_t launch_visc_move(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& cur,
                            Hot& out, Cold& cold, Events& ev, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    LAUNCH_DM(k_visc_move, grid_for(n, bs), dim3(bs), st, g, c, tab, n, cur, out, cold, ev);
    return hipGetLastError();
}

// ---------------- collision (src/mps.cpp:217-250) ----------------
//
// The reference visits fluid particles in ascending id and, for each, its list
// in order; a pair is skipped when j is fluid with j.id >= i.id, and acts when
// the list distance (fixed at the build) is below the collision distance.
// Candidate pairs are therefore fixed before any update.  They are listed in
// the reference's order (per-id counts, scan, emit), then each pair gets a
// level = 1 + the highest level of earlier pairs sharing a fluid particle;
// walls never change, so they carry no dependency.  Pairs of one level touch
// disjoint fluid particles and are resolved together; levels run in order.

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) MPS_WALK_WAVES k_coll_count(GridView g0, Consts c0, SceneTab tab, int n, Hot h, int* __restrict__ cnt) {
    NB_STAGE(DIM, g0, h, n);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    const int id = h.id[s];
    int k = 0;
    if (h.type[s] == MPS_FLUID) {
        const double re = c.dc < c.re_max ? c.dc : c.re_max;
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(nbs, g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                           [&](int q, double, double, double, double) {
                               if (!(h.type[q] == MPS_FLUID && h.id[q] >= id)) k++;
                           });
    }
    cnt[id] = k;
}

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_coll_emit(GridView g0, Consts c0, SceneTab tab, int n, Hot h, const int* __restrict__ cnt,
                                                      const int* __restrict__ off, int* __restrict__ pairs,
                                                      double* __restrict__ pdist, int cap, int* __restrict__ err,
                                                      double* __restrict__ pdir) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    const int id = h.id[s];
    if (cnt[id] == 0) return;
    int o = off[id];
    const double re = c.dc < c.re_max ? c.dc : c.re_max;
    const double zi = DIM == 3 ? h.pz[s] : 0.0;
    for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                       [&](int q, double dist, double dx, double dy, double dz) {
                           if (h.type[q] == MPS_FLUID && h.id[q] >= id) return;
                           if (o < cap) {
                               pairs[2 * o] = s;
                               pairs[2 * o + 1] = q;
                               pdist[o] = dist;
                               if (pdir) {  // d dist / d x_j, kept for the reverse sweep
                                   pdir[3 * (size_t) o] = dx / dist;
                                   pdir[3 * (size_t) o + 1] = dy / dist;
                                   pdir[3 * (size_t) o + 2] = dz / dist;
                               }
                           } else {
                               atomicOr(err, kErrCollCap);
                           }
                           o++;
                       });
}

constexpr int kResolveThreads = 1024;
// Pair lists up to this long get their levels out of LDS; longer ones out of
// global scratch, by the same steps.
constexpr int kCollLds = 4096;

// Levels of np pairs, in parallel (one workgroup).  Pairs come grouped by the
// id of their first (fluid) particle i, ascending, and a fluid partner j has
// a smaller id than i, so it appears only in its own, earlier group and as
// the partner of later pairs.  The latest earlier pair touching i is
// therefore the pair before p when that has the same i; the latest earlier
// pair touching j is the previous pair with partner j (found by sorting the
// (j, p) keys), or else the last pair of j's own group.  level[p] = 1 + the
// higher level of those two (0 for none); rounds of that update over every
// pair reach the same values as the sequential assignment once no level
// changes (each round settles at least one more link of every chain).
// s_key (a power of two >= np entries), s_prev and s_lev (np) are LDS for
// short lists, global scratch for long ones.
template <typename LV>
__device__ void coll_levels_core(int np, const Hot& h, const int* __restrict__ off, const int* __restrict__ pairs,
                                 unsigned long long* s_key, int* s_prev, LV* s_lev, int* level, int& lmax) {
    __shared__ int s_changed, s_max;
    int n2 = 1;
    while (n2 < np) n2 <<= 1;
    for (int p = threadIdx.x; p < n2; p += blockDim.x) {
        unsigned long long k = ~0ull;
        if (p < np) {
            const int j = pairs[2 * p + 1];
            if (h.type[j] == MPS_FLUID) k = ((unsigned long long) (unsigned) j << 32) | (unsigned) p;
            s_prev[p] = -1;
            s_lev[p] = (LV) 1;
        }
        s_key[p] = k;
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {  // bitonic sort, ascending
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int ixj = i ^ jj;
                if (ixj > i) {
                    const unsigned long long a = s_key[i], b = s_key[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        s_key[i] = b;
                        s_key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < n2; k += blockDim.x) {
        const unsigned long long key = s_key[k];
        if (key == ~0ull) continue;
        const int j = (int) (key >> 32), p = (int) (key & 0xffffffffu);
        int prev;
        if (k > 0 && (int) (s_key[k - 1] >> 32) == j) {
            prev = (int) (s_key[k - 1] & 0xffffffffu);
        } else {
            const int gid = h.id[j];
            const int lo = off[gid], hi = off[gid + 1];
            prev = hi > lo ? hi - 1 : -1;
        }
        s_prev[p] = prev;
    }
    __syncthreads();
    volatile LV* lev = s_lev;
    for (;;) {
        if (threadIdx.x == 0) s_changed = 0;
        __syncthreads();
        int ch = 0;
        for (int p = threadIdx.x; p < np; p += blockDim.x) {
            const int pi = p > 0 && pairs[2 * (p - 1)] == pairs[2 * p] ? p - 1 : -1;
            const int pj = s_prev[p];
            int L = 1;
            if (pi >= 0 && lev[pi] + 1 > L) L = lev[pi] + 1;
            if (pj >= 0 && lev[pj] + 1 > L) L = lev[pj] + 1;
            if (L != lev[p]) {
                lev[p] = (LV) L;
                ch = 1;
            }
        }
        if (ch) s_changed = 1;
        __syncthreads();
        if (!s_changed) break;
        __syncthreads();
    }
    if (threadIdx.x == 0) s_max = 0;
    __syncthreads();
    int mx = 0;
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        const int L = lev[p];
        level[p] = L;
        mx = L > mx ? L : mx;
    }
    atomicMax(&s_max, mx);
    __syncthreads();
    lmax = s_max;
    __syncthreads();
}

__device__ void coll_levels_lds(int np, const Hot& h, const int* __restrict__ off, const int* __restrict__ pairs,
                                int* __restrict__ level, int& lmax) {
    __shared__ unsigned long long s_key[kCollLds];
    __shared__ int s_prev[kCollLds];
    __shared__ short s_lev[kCollLds];
    coll_levels_core<short>(np, h, off, pairs, s_key, s_prev, s_lev, level, lmax);
}

template <bool MULTI>
__global__ void __launch_bounds__(kResolveThreads) k_coll_resolve(int n, Consts c0, SceneTab tab, int dim, Hot h,
                                                                  const int* __restrict__ off,
                                                                  const int* __restrict__ pairs,
                                                                  const double* __restrict__ pdist,
                                                                  int* level, unsigned long long* gkey,
                                                                  int* gprev, int cap, CollTape tp) {
    __shared__ int s_max;
    int np = off[n];
    if (np > cap) np = cap;
    if (np == 0) return;
    int mx = 0;
    if (np <= kCollLds)
        coll_levels_lds(np, h, off, pairs, level, mx);
    else  // the levels themselves are the rounds' working copy
        coll_levels_core<int>(np, h, off, pairs, gkey, gprev, level, level, mx);
    if (threadIdx.x == 0) {
        s_max = mx;
        if (tp.nlev) *tp.nlev = mx;
    }
    __syncthreads();
    volatile double* px = h.px;
    volatile double* py = h.py;
    volatile double* pz = h.pz;
    volatile double* vx = h.vx;
    volatile double* vy = h.vy;
    volatile double* vz = h.vz;
    const bool d3 = dim == 3;
    for (int L = 1; L <= s_max; L++) {
        for (int p = threadIdx.x; p < np; p += blockDim.x) {
            if (level[p] != L) continue;
            const int i = pairs[2 * p], j = pairs[2 * p + 1];
            Consts c = c0;
            GridView gdummy;
            scene_of<MULTI>(tab, h.id[i], gdummy, c);
            const double e = c.restitution, dc = c.dc;
            const double dist = pdist[p];
            const bool jf = h.type[j] == MPS_FLUID;
            // src/particle.cpp:19-29: 1/density for fluid, 0 otherwise
            const double invMi = 1 / c.rho;
            const double invMj = jf ? 1 / c.rho : 0;
            const double reduced = 1.0 / (invMi + invMj);
            double nx = px[j] - px[i], ny = py[j] - py[i], nz = d3 ? pz[j] - pz[i] : 0.0;
            const double zz = (nx * nx + ny * ny) + nz * nz;
            if (zz > 0.0) {  // Eigen normalized()
                const double sq = sqrt(zz);
                nx = nx / sq;
                ny = ny / sq;
                nz = nz / sq;
            }
            const double dvx = vx[j] - vx[i], dvy = vy[j] - vy[i], dvz = d3 ? vz[j] - vz[i] : 0.0;
            const double rv = (dvx * nx + dvy * ny) + dvz * nz;
            if (tp.flags) {  // record for the reverse sweep
                double* pre = tp.pre + 12 * (size_t) p;
                pre[0] = px[i];
                pre[1] = py[i];
                pre[2] = d3 ? pz[i] : 0.0;
                pre[3] = px[j];
                pre[4] = py[j];
                pre[5] = d3 ? pz[j] : 0.0;
                pre[6] = vx[i];
                pre[7] = vy[i];
                pre[8] = d3 ? vz[i] : 0.0;
                pre[9] = vx[j];
                pre[10] = vy[j];
                pre[11] = d3 ? vz[j] : 0.0;
                tp.pid[2 * p] = h.id[i];
                tp.pid[2 * p + 1] = h.id[j];
                tp.flags[p] = (rv < 0.0 ? 1 : 0) | (jf ? 2 : 0);
            }
            if (rv < 0.0) {
                const double impulse = -(1.0 + e) * rv * reduced;
                const double si = impulse * invMi, sj = impulse * invMj;
                vx[i] = vx[i] - si * nx;
                vy[i] = vy[i] - si * ny;
                if (d3) vz[i] = vz[i] - si * nz;
                const double depth = dc - dist;
                const double pimp = depth * reduced;
                const double pi_ = pimp * invMi, pj_ = pimp * invMj;
                px[i] = px[i] - pi_ * nx;
                py[i] = py[i] - pi_ * ny;
                if (d3) pz[i] = pz[i] - pi_ * nz;
                if (jf) {  // a wall's update adds zero: skipped
                    vx[j] = vx[j] + sj * nx;
                    vy[j] = vy[j] + sj * ny;
                    if (d3) vz[j] = vz[j] + sj * nz;
                    px[j] = px[j] + pj_ * nx;
                    py[j] = py[j] + pj_ * ny;
                    if (d3) pz[j] = pz[j] + pj_ * nz;
                }
            }
        }
        __threadfence();
        __syncthreads();
    }
}

hipError_t launch_collision(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, Hot& h,
                            StepScratch& w, int* err, hipStream_t st, int bs, const CollTape* tape) {
    if (n == 0) return hipSuccess;
    const CollTape tp = tape ? *tape : CollTape{nullptr, nullptr, nullptr, nullptr, nullptr};
    LAUNCH_DM(k_coll_count, grid_for(n, bs), dim3(bs), st, g, c, tab, n, h, w.coll_cnt);
    hipError_t e = scan_exclusive(w.coll_cnt, w.coll_off, n, w.scan_tmp, st);
    if (e != hipSuccess) return e;
    LAUNCH_DM(k_coll_emit, grid_for(n, bs), dim3(bs), st, g, c, tab, n, h, w.coll_cnt, w.coll_off, w.coll_pairs,
              w.coll_dist, w.coll_cap, err, tp.dir);
    if (tab.g)
        hipLaunchKernelGGL(k_coll_resolve<true>, dim3(1), dim3(kResolveThreads), 0, st, n, c, tab, dim, h, w.coll_off,
                           w.coll_pairs, w.coll_dist, w.coll_level, w.coll_key, w.coll_prev, w.coll_cap, tp);
    else
        hipLaunchKernelGGL(k_coll_resolve<false>, dim3(1), dim3(kResolveThreads), 0, st, n, c, tab, dim, h, w.coll_off,
                           w.coll_pairs, w.coll_dist, w.coll_level, w.coll_key, w.coll_prev, w.coll_cap, tp);
    return hipGetLastError();
}

// ---------------- boundary condition + source term (src/mps.cpp:262-320) ----------------

// BC and source term, and the start of the Dirichlet connectivity forest:
// parent = the smallest slot among the particle and its neighbours within
// re_p (every non-ghost particle, dummy walls included, src/mps.cpp:387-397).
template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) MPS_WALK_WAVES k_bc_source(GridView g0, Consts c0, SceneTab tab, int n, Hot h, Cold cold,
                                                      int8_t* __restrict__ bc_s, double* __restrict__ b_s,
                                                      int* __restrict__ parent, NbList nb) {
    NB_STAGE(DIM, g0, h, n);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    const int t = h.type[s];
    int8_t bc = MPS_BC_IGNORED;
    double b = 0.0;
    int m = s;
    int cnt = 0;
    if (t != MPS_GHOST) {
        const double rs = c.re_s, rp = c.re_p;
        const double re = rs > rp ? rs : rp;
        const bool same = rs == rp;  // one weight serves both sums (the usual settings)
        double ns = 0.0, np = 0.0;
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(nbs, g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                           [&](int q, double dist, double, double, double) {
                               if (cnt < nb.cap) {
                                   const size_t e = (size_t) cnt * nb.stride + s;
                                   nb.q[e] = q;
                                   nb.d[e] = dist;
                               }
                               cnt++;
                               double ws = 0.0;
                               if (dist < rs) {
                                   ws = wfun(dist, rs);
                                   ns += ws;
                               }
                               if (dist < rp) {
                                   np += same ? ws : wfun(dist, rp);
                                   if (q < m) m = q;
                               }
                           });
        if (t != MPS_DUMMY_WALL) {
            bc = ns < c.beta * c.n0_s ? MPS_BC_SURFACE : MPS_BC_INNER;
            if (bc == MPS_BC_INNER && !c.hos_on) b = c.src_scale * (np - c.n0_p) / c.n0_p;
        }
    }
    bc_s[s] = bc;
    b_s[s] = b;
    parent[s] = m;
    nb.cnt[s] = cnt;
    cold.bc[h.id[s]] = bc;
}

// Higher-order source term (src/mps.cpp:288-308); needs every slot's BC.
template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_source_hos(GridView g0, Consts c0, SceneTab tab, int n, Hot h, Cold cold,
                                                       const int8_t* __restrict__ bc_s, double* __restrict__ b_s,
                                                       NbList nb) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    if (!c.hos_on || bc_s[s] != MPS_BC_INNER) return;
    double b = c.hos_scale * (cold.ndr[h.id[s]] - 1.0);
    const double re = c.re_p;
    const double xi = h.px[s], yi = h.py[s], zi = DIM == 3 ? h.pz[s] : 0.0;
    const double vxi = h.vx[s], vyi = h.vy[s], vzi = DIM == 3 ? h.vz[s] : 0.0;
    // neighbours beyond re_p add nothing, so the list (radius >= re_p) covers the walk
    struct Nbr {
        int bc;
        double vx, vy, vz;
    };
    for_listed_g<DIM>(nb, g, h.px, h.py, h.pz, h.key[s], s, xi, yi, zi, re,
                      [&](int q) { return Nbr{(int) bc_s[q], h.vx[q], h.vy[q], DIM == 3 ? h.vz[q] : 0.0}; },
                      [&](int q, double dist, double rx, double ry, double rz, const Nbr& nq) {
                           if (nq.bc == MPS_BC_IGNORED) return;
                           const double ux = nq.vx - vxi, uy = nq.vy - vyi;
                           const double uz = DIM == 3 ? nq.vz - vzi : 0.0;
                           if (dist < re) b -= c.hos_coef * re * ((rx * ux + ry * uy) + rz * uz) / (dist * dist * dist);
                       });
    b_s[s] = b;
}

hipError_t launch_bc_source(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                            Cold& cold, StepScratch& w, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    LAUNCH_DM(k_bc_source, grid_for(n, bs), dim3(bs), st, g, c, tab, n, h, cold, w.bc_s, w.b_s, w.parent, w.nb);
    if (c.hos_on) LAUNCH_DM(k_source_hos, grid_for(n, bs), dim3(bs), st, g, c, tab, n, h, cold, w.bc_s, w.b_s, w.nb);
    return hipGetLastError();
}

// ---------------- matrix diagonal + Dirichlet connectivity (src/mps.cpp:322-415) ----------------

// Connectivity: the reference runs a multi-source BFS from every Surface
// particle over edges dist < re_p through any particle (src/mps.cpp:373-400).
// The visited set is the union of connected components that hold a Surface
// particle, found here with lock-free union-find (hook the larger root under
// the smaller with atomicCAS, path halving on find), started from the
// parents k_bc_source wrote.
__device__ __forceinline__ int cc_find(volatile int* parent, int v) {
    int curr = parent[v];
    if (curr != v) {
        int prev = v, next;
        while (curr > (next = parent[curr])) {
            parent[prev] = next;
            prev = curr;
            curr = next;
        }
    }
    return curr;
}

// Matrix diagonal (src/mps.cpp:336-356) and off-diagonal count of Inner
// rows, and the union-find hooking over the same re_p neighbours.
template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_diag_join(GridView g0, Consts c0, SceneTab tab, int n, Hot h,
                                                             const int8_t* __restrict__ bc_s,
                                                             double* __restrict__ diag, int* __restrict__ nnz,
                                                             int* parent, NbList nb) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    double d = 0.0;
    int k = 0;
    if (h.type[s] != MPS_GHOST) {
        const bool inner = bc_s[s] == MPS_BC_INNER;
        const double re = c.re_p;
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        volatile int* vp = parent;
        for_listed_g<DIM>(nb, g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                          [&](int q) { return (int) bc_s[q]; },
                          [&](int q, double dist, double, double, double, int bq) {
                               if (inner) {
                                   if (bq != MPS_BC_IGNORED) {
                                       const double a = c.coefA * wfun(dist, re) / c.rho;
                                       d += a;
                                       if (bq == MPS_BC_INNER) k++;
                                   }
                               }
                               if (q <= s) return;
                               // parents only ever move to ancestors, so equal
                               // parents (even read stale) share a component
                               if (parent[q] == parent[s]) return;
                               int a = cc_find(vp, s);
                               int b = cc_find(vp, q);
                               while (a != b) {
                                   if (a < b) {
                                       const int r = atomicCAS(&parent[b], b, a);
                                       if (r == b) break;
                                       b = r;
                                   } else {
                                       const int r = atomicCAS(&parent[a], a, b);
                                       if (r == a) break;
                                       a = r;
                                   }
                               }
                           });
        if (inner && c.qc_on) d += c.qc_add;
    }
    diag[s] = d;
    nnz[s] = k;
}

// One round of pointer jumping on the connectivity forest before the
// hooking: parent[s] = parent[parent[s]].  k_bc_source starts each particle
// at its smallest neighbour, so chains run across the fluid (hundreds of
// links at 2M particles) and the hooking pass's finds would chase them with
// uncached loads; each round halves them.  Parents only move to ancestors,
// so a jump that reads a parent another thread has already moved is still
// an ancestor: the races change how far a round gets, never the components.
__global__ void __launch_bounds__(kMaxTuneBlock) k_cc_jump(int n, int hops, int* __restrict__ parent) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p0 = parent[s];
    int p = p0;
    for (int k = 0; k < hops; k++) {
        const int q = parent[p];
        if (q == p) break;
        p = q;
    }
    if (p != p0) parent[s] = p;
}

struct CcJumps {
    int rounds, hops;
};
// Read at every step (a getenv is nothing beside a step), so one process
// can compare settings.
static CcJumps cc_jumps() {
    CcJumps j{4, 7};
    if (const char* e = getenv("MPS_AMD_CC_JUMPS")) {
        j.hops = 1;
        sscanf(e, "%d,%d", &j.rounds, &j.hops);
        j.rounds = std::max(0, std::min(j.rounds, 64));
        j.hops = std::max(1, std::min(j.hops, 64));
    }
    return j;
}

__global__ void __launch_bounds__(kMaxTuneBlock) k_cc_finalize(int n, int* parent, const int8_t* __restrict__ bc_s,
                                                        int8_t* __restrict__ surf) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    volatile int* vp = parent;
    int r = vp[s];
    while (r != vp[r]) r = vp[r];
    parent[s] = r;
    if (bc_s[s] == MPS_BC_SURFACE) surf[r] = 1;
}

// Dirichlet doubling (src/mps.cpp:403-414) and, for the Jacobi scaling of
// the solve, 1/sqrt of each Inner row's final diagonal.
__global__ void __launch_bounds__(kMaxTuneBlock) k_cc_apply(int n, Hot h, const int* __restrict__ parent,
                                                     const int8_t* __restrict__ bc_s, const int8_t* __restrict__ surf,
                                                     double* __restrict__ diag, double* __restrict__ isq,
                                                     int* __restrict__ inner_flag, Events ev) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const bool inner = bc_s[s] == MPS_BC_INNER;
    inner_flag[s] = inner ? 1 : 0;
    double d = diag[s];
    if (inner && !surf[parent[s]]) {
        d *= 2.0;
        diag[s] = d;
        push_event(ev.n_unconn, ev.unconn_ids, ev.cap, h.id[s]);
    }
    isq[s] = inner ? 1.0 / sqrt(d) : 0.0;
}

hipError_t launch_diag_cc(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                          StepScratch& w, Events& ev, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    const dim3 gr = grid_for(n, bs), bl(bs);
    const CcJumps jp = cc_jumps();
    for (int r = jp.rounds; r > 0; r--) hipLaunchKernelGGL(k_cc_jump, gr, bl, 0, st, n, jp.hops, w.parent);
    LAUNCH_DM(k_diag_join, gr, bl, st, g, c, tab, n, h, w.bc_s, w.diag_s, w.nnz_s, w.parent, w.nb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // surf[] is a byte array: clear it with the int fill over its rounded-up size
    e = fill_int(reinterpret_cast<int*>(w.surf), 0, ((size_t) n + 3) / 4, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cc_finalize, gr, bl, 0, st, n, w.parent, w.bc_s, w.surf);
    hipLaunchKernelGGL(k_cc_apply, gr, bl, 0, st, n, h, w.parent, w.bc_s, w.surf, w.diag_s, w.isq_s, w.inner_flag, ev);
    return hipGetLastError();
}

// ---------------- pressure out, clamp (src/mps.cpp:466-479) ----------------

__global__ void __launch_bounds__(kMaxTuneBlock) k_pressure_out(int n, Hot h, const int8_t* __restrict__ bc_s,
                                                         const double* __restrict__ x_s, Cold cold,
                                                         double* __restrict__ p_s) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    const double x = bc_s[s] == MPS_BC_INNER ? x_s[s] : 0.0;  // empty rows solve to 0
    cold.pprev2[id] = cold.pprev[id];
    cold.pprev[id] = x;
    const double p = x < 0.0 ? 0.0 : x;  // removeNegativePressure
    p_s[s] = p;
    cold.p[id] = p;
}

hipError_t launch_pressure_out(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                               const double* x_s, Cold& cold, StepScratch& w, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pressure_out, grid_for(n, bs), dim3(bs), 0, st, n, h, w.bc_s, x_s, cold, w.p_s);
    return hipGetLastError();
}

// ---------------- minimum pressure, pressure gradient, implicit move (src/mps.cpp:481-545) ----------------
//
// One thread per slot walks its re_p neighbours twice: first for the minimum
// pressure (src/mps.cpp:487-500: p_i itself and every neighbour that is not a
// dummy wall; Ignored particles keep their old value), which needs only the
// clamped pressures of the pass before and the list's indices, then for the
// gradient, which needs only its own minimum.  The second walk finds the
// neighbours in cache.

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_grad_move(GridView g0, Consts c0, SceneTab tab, int n, Hot cur, Hot out, Cold cold,
                                                      const int8_t* __restrict__ bc_s,
                                                      const double* __restrict__ p_s,
                                                      double* __restrict__ pmin_s, Events ev, NbList nb) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, cur.id[s], g, c);
    double x = cur.px[s], y = cur.py[s], z = DIM == 3 ? cur.pz[s] : 0.0;
    double vx = cur.vx[s], vy = cur.vy[s], vz = DIM == 3 ? cur.vz[s] : 0.0;
    const int t = cur.type[s];
    const int id = cur.id[s];
    struct NbrTP {
        int type;
        double p;
    };
    auto nbr_tp = [&](int q) { return NbrTP{(int) cur.type[q], p_s[q]}; };
    double pmi;
    if (bc_s[s] == MPS_BC_IGNORED) {
        pmi = cold.pmin[id];
    } else {
        pmi = p_s[s];
        for_listed_ids<DIM>(nb, g, cur.px, cur.py, cur.pz, cur.key[s], s, x, y, z, c.re_p, nbr_tp,
                            [&](int, const NbrTP& nq) {
                                if (nq.type == MPS_DUMMY_WALL) return;
                                if (pmi > nq.p) pmi = nq.p;
                            });
        cold.pmin[id] = pmi;
    }
    pmin_s[s] = pmi;
    if (t == MPS_FLUID) {
        const double re = c.re_p;
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for_listed_g<DIM>(nb, g, cur.px, cur.py, cur.pz, cur.key[s], s, x, y, z, re, nbr_tp,
                          [&](int, double dist, double dx, double dy, double dz, const NbrTP& nq) {
                               if (nq.type == MPS_DUMMY_WALL) return;
                               const double dp = nq.p - pmi;
                               const double w = wfun(dist, re);
                               const double d2 = dist * dist;
                               gx += dx * dp * w / d2;
                               gy += dy * dp * w / d2;
                               if (DIM == 3) gz += dz * dp * w / d2;
                           });
        gx *= c.gscale;
        gy *= c.gscale;
        gz *= c.gscale;
        const double ax = -1.0 * gx / c.rho, ay = -1.0 * gy / c.rho, az = -1.0 * gz / c.rho;
        vx += ax * c.dt;
        vy += ay * c.dt;
        if (DIM == 3) vz += az * c.dt;
        const double x0 = x, y0 = y, z0 = z;
        x += ax * c.dt * c.dt;
        y += ay * c.dt * c.dt;
        if (DIM == 3) z += az * c.dt * c.dt;
        if (!in_domain(c, x, y, DIM == 3 ? z : c.dmin[2])) {
            cur.type[s] = MPS_GHOST;
            push_event(ev.n_ghost, ev.ghost_ids, ev.cap, id);
            cold.ghost_n[id] = ndens<DIM>(g, cur, s, x0, y0, z0, c.re_p);
        }
    }
    out.px[s] = x;
    out.py[s] = y;
    out.vx[s] = vx;
    out.vy[s] = vy;
    if (DIM == 3) {
        out.pz[s] = z;
        out.vz[s] = vz;
    }
}

hipError_t launch_grad_move(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& cur,
                            Hot& out, Cold& cold, StepScratch& w, Events& ev, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    LAUNCH_DM(k_grad_move, grid_for(n, bs), dim3(bs), st, g, c, tab, n, cur, out, cold, w.bc_s, w.p_s, w.pmin_s, ev,
              w.nb);
    return hipGetLastError();
}

// ---------------- number density ratio (src/mps.cpp:627-641) ----------------

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_ndr(GridView g0, Consts c0, SceneTab tab, int n, Hot h, Cold cold) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts c = c0;
    scene_of<MULTI>(tab, h.id[s], g, c);
    const int id = h.id[s];
    double nd;
    if (h.type[s] != MPS_GHOST) {
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        nd = ndens<DIM>(g, h, s, h.px[s], h.py[s], zi, c.re_p);
    } else {
        nd = cold.ghost_n[id];
    }
    cold.ndr[id] = cold.bc[id] == MPS_BC_SURFACE ? 0.0 : nd / c.n0_p;
}

hipError_t launch_ndr(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                      Cold& cold, hipStream_t st, int bs) {
    if (n == 0) return hipSuccess;
    LAUNCH_DM(k_ndr, grid_for(n, bs), dim3(bs), st, g, c, tab, n, h, cold);
    return hipGetLastError();
}

// ---------------- Courant number (src/mps.cpp:547-571) ----------------
// max |v| over fluid; (|v| dt)/l0 is monotone in |v|, so the host applies it
// to the maximum and gets the reference's value bit for bit.

template <bool MULTI>
__global__ void __launch_bounds__(kBlock) k_vmax(int n, Hot h, int dim3d, const int16_t* __restrict__ sid,
                                                 unsigned long long* __restrict__ out) {
    __shared__ double red[kBlock / 64];
    if constexpr (MULTI) {
        // One slot per thread.  A scene's slots are contiguous (slots are
        // sorted by key and every scene owns one key range), so a wavefront
        // nearly always holds one scene: then one atomic for the wavefront,
        // else one per lane that carries a value.
        const int s = blockIdx.x * blockDim.x + threadIdx.x;
        double v = 0.0;
        int si = -1;
        if (s < n && h.type[s] == MPS_FLUID) {
            const double a = h.vx[s], b = h.vy[s], cz = dim3d ? h.vz[s] : 0.0;
            v = sqrt((a * a + b * b) + cz * cz);
            si = sid[h.id[s]];
        }
        int lo = si < 0 ? 0x7fffffff : si, hi = si;
        double m = v;
        for (int o = 32; o > 0; o >>= 1) {
            const int a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
            const double u = __shfl_xor(m, o, 64);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
            m = u > m ? u : m;
        }
        // the block's wavefronts: one atomic if they all hold the same scene
        __shared__ int s_sc[kBlock / 64];
        __shared__ double s_m[kBlock / 64];
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            s_sc[wv] = hi < 0 ? -1 : (lo == hi ? hi : -2);  // -1 none, -2 mixed
            s_m[wv] = m;
        }
        __syncthreads();
        int bs = -1;
        double bm = 0.0;
        for (int k = 0; k < (int) (blockDim.x >> 6); k++) {
            const int c = s_sc[k];
            if (c == -1) continue;
            bs = (bs == -1 || bs == c) ? c : -2;
            bm = s_m[k] > bm ? s_m[k] : bm;
        }
        if (bs >= 0) {
            if (threadIdx.x == 0 && bm > 0.0) atomicMax(&out[bs], (unsigned long long) __double_as_longlong(bm));
        } else if (bs == -2) {
            if (hi < 0) return;
            if (lo == hi) {
                if ((threadIdx.x & 63) == 0 && m > 0.0)
                    atomicMax(&out[hi], (unsigned long long) __double_as_longlong(m));
            } else if (v > 0.0) {
                atomicMax(&out[si], (unsigned long long) __double_as_longlong(v));
            }
        }
        return;
    }
    // grid-stride: a few hundred blocks, so a few hundred atomics on the result
    double v = 0.0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        if (h.type[s] != MPS_FLUID) continue;
        const double a = h.vx[s], b = h.vy[s], cz = dim3d ? h.vz[s] : 0.0;
        const double u = sqrt((a * a + b * b) + cz * cz);
        v = u > v ? u : v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_down(v, o, 64);
        v = u > v ? u : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < kBlock / 64; k++) m = red[k] > m ? red[k] : m;
        atomicMax(out, (unsigned long long) __double_as_longlong(m));
    }
}

hipError_t launch_courant(int dim, int n, const Hot& h, const SceneTab& tab, int nscene, double* vmax,
                          hipStream_t st) {
    hipError_t e = fill_double(vmax, 0.0, (size_t) nscene, st);
    if (e != hipSuccess || n == 0) return e;
    const dim3 grid(std::min(512u, grid_for(n, kBlock).x));
    if (tab.g)
        hipLaunchKernelGGL(k_vmax<true>, grid_for(n, kBlock), dim3(kBlock), 0, st, n, h, dim == 3 ? 1 : 0, tab.sid,
                           reinterpret_cast<unsigned long long*>(vmax));
    else
        hipLaunchKernelGGL(k_vmax<false>, grid, dim3(kBlock), 0, st, n, h, dim == 3 ? 1 : 0, tab.sid,
                           reinterpret_cast<unsigned long long*>(vmax));
    return hipGetLastError();
}

// ---------------- id-order snapshots ----------------

__global__ void __launch_bounds__(kMaxTuneBlock) k_gather_id(int n, int dim, Hot h, Cold cold, double* __restrict__ pos,
                                                      double* __restrict__ vel) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = h.id[s];
    pos[3 * (size_t) id] = h.px[s];
    pos[3 * (size_t) id + 1] = h.py[s];
    pos[3 * (size_t) id + 2] = dim == 3 ? h.pz[s] : cold.z[id];
    vel[3 * (size_t) id] = h.vx[s];
    vel[3 * (size_t) id + 1] = h.vy[s];
    vel[3 * (size_t) id + 2] = dim == 3 ? h.vz[s] : cold.vz[id];
}

__global__ void __launch_bounds__(kBlock) k_gather_type(int n, Hot h, int8_t* __restrict__ type_id) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    type_id[h.id[s]] = h.type[s];
}

hipError_t launch_gather_type(int n, const Hot& h, int8_t* type_id, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_type, grid_for(n, kBlock), dim3(kBlock), 0, st, n, h, type_id);
    return hipGetLastError();
}

hipError_t launch_snapshot(int dim, int n, const Hot& h, const Cold& cold, double* pos_id, double* vel_id,
                           hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_id, grid_for(n, kBlock), dim3(kBlock), 0, st, n, dim, h, cold, pos_id, vel_id);
    return hipGetLastError();
}

}  // namespace mps
