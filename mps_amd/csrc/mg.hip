// mg.hip — aggregation multigrid preconditioner for the pressure CG.
//
// The Jacobi-scaled system Ahat (unit diagonal, SELL storage) is coarsened
// along the bucket grid the step already has:
//   level 1: one aggregate per bucket cell.  Rows are compacted in slot order
//            and slots are sorted by cell, so a cell's Inner rows are one
//            contiguous row range;
//   level l+1 (l >= 1): 2 x 2 (x 2 in 3D) blocks of level-l cells.
// Interpolation is piecewise constant (P_{i,K} = 1 when row i lies in
// aggregate K); coarse operators are Galerkin, A_{l+1} = P^T A_l P, built by
// gathers (one thread per coarse row sums its children's entries in a fixed
// order), so they are deterministic.  Particles interact only within
// re_p <= re_max = cell side, so a coarse row couples at most to its 8 (26)
// neighbouring aggregates: every coarse operator is an ELL matrix of width
// 9 (27) with the diagonal kept apart.
//
// The preconditioner is one V(1,1) cycle: damped Jacobi pre-smoothing from a
// zero guess, residual, restriction, recursion, prolongation with a factor
// (over-correction for piecewise-constant interpolation), damped Jacobi
// post-smoothing; the coarsest level (<= 64 rows) gets a fixed number of
// Jacobi sweeps in one workgroup.  Pre- and post-smoother are the same
// symmetric operator and the coarsest solve is a fixed polynomial, so the
// cycle is a symmetric positive definite linear operator and plain
// (preconditioned) CG applies.  Every kernel returns at once when the CG's
// done flag is up, like the CG kernels around it.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mps_internal.hpp"

namespace mps {

constexpr int kMgMaxLevels = 16;
constexpr int kMgCoarse = 64;  // rows of the coarsest level at most

static inline dim3 mg_grid(int n) { return dim3((unsigned) std::max(1, (n + kBlock - 1) / kBlock)); }
static inline dim3 fill_grid_mg(int n) { return dim3((unsigned) std::min(4096, std::max(1, (n + kBlock - 1) / kBlock))); }

// The cycle is skipped once every segment's CG has converged (a segment
// still iterating needs it; converged segments ignore its output).
__device__ __forceinline__ bool mg_done(MgDone d) {
    __shared__ int s_done;
    if (threadIdx.x == 0) {
        int all = d.flag ? 1 : 0;
        for (int i = 0; i < d.count && all; i++) all = *(volatile const int*) (d.flag + (size_t) i * d.stride);
        s_done = all;
    }
    __syncthreads();
    return s_done != 0;
}

// The same test made by each wave on its own (lane i reads segment i's
// flag), for kernels without a workgroup barrier: every wave reads the same
// flags, written by earlier launches, so all waves agree.
__device__ __forceinline__ bool mg_done_wave(MgDone d) {
    if (!d.flag) return false;
    const int lane = threadIdx.x & 63;
    bool all = true;
    for (int i0 = 0; i0 < d.count; i0 += 64) {
        const int i = i0 + lane;
        const int f = i < d.count ? *(volatile const int*) (d.flag + (size_t) i * d.stride) : 1;
        all = all && __all(f != 0);
    }
    return all;
}

// Scene of a key / a dense index: the last i with first[i] <= v (first has
// count + 1 ascending entries).
__device__ __forceinline__ int mg_bucket(const int* __restrict__ first, int count, int v) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}


// ---------------- setup ----------------

// Cell key of each compact row and the first row of each cell.
__global__ void __launch_bounds__(kBlock) k_mg_heads(int R, const int* __restrict__ row_slot,
                                                     const int32_t* __restrict__ key_s, int* __restrict__ head,
                                                     int* __restrict__ rkey) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int k = key_s[row_slot[r]];
    rkey[r] = k;
    head[r] = (r == 0 || key_s[row_slot[r - 1]] != k) ? 1 : 0;
}

// Level 1: fine row -> aggregate, aggregate -> first fine row, its scene and
// its cell on the scene's grid.
__global__ void __launch_bounds__(kBlock) k_mg_level1(int R, const int* __restrict__ head,
                                                      const int* __restrict__ hscan, const int* __restrict__ rkey,
                                                      const int* __restrict__ key0, int nscene,
                                                      int* __restrict__ agg0, unsigned char* __restrict__ agg0b,
                                                      int* __restrict__ cst1, int* __restrict__ coord1,
                                                      int* __restrict__ scene1) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int K = hscan[r] + head[r] - 1;
    agg0[r] = K;
    agg0b[r] = (unsigned char) K;  // the cell's place in its up-pass block of 256 cells
    if (head[r]) {
        const int sc = mg_bucket(key0, nscene, rkey[r]);
        cst1[K] = r;
        coord1[K] = rkey[r] - key0[sc];  // the cell's key is its place on the scene's grid
        scene1[K] = sc;
    }
    if (r == R - 1) cst1[K + 1] = R;
}

struct MgDims {
    int x, y, z;
};

__device__ __forceinline__ void mg_decode(int key, MgDims d, int& cx, int& cy, int& cz) {
    cx = key % d.x;
    const int t = key / d.x;
    cy = t % d.y;
    cz = t / d.y;
}

// Stencil position of a neighbouring cell: dx + 3 dy (+ 9 dz in 3D), each
// offset shifted to 0..2, from the key difference on a grid of d.x * d.y (*
// d.z) cells (keys x fastest).  Needs d.x, d.y >= 3 (then the offsets are the
// nearest multiples of the strides); -1 if dkey is no neighbour offset.
template <int W>
__device__ __forceinline__ int mg_stencil_of_dkey(int dkey, MgDims d) {
    int dz = 0;
    if (W == 27 && d.z > 1) {
        const int nxy = d.x * d.y;
        dz = dkey > nxy / 2 ? 1 : (dkey < -(nxy / 2) ? -1 : 0);
        dkey -= dz * nxy;
    }
    const int dy = dkey > d.x / 2 ? 1 : (dkey < -(d.x / 2) ? -1 : 0);
    const int dx = dkey - dy * d.x;
    if (dx < -1 || dx > 1) return -1;
    int j = (dx + 1) + 3 * (dy + 1);
    if (W == 27) j += 9 * (dz + 1);
    return j;
}

// The same from decoded cell coordinates (any grid size).
template <int W>
__device__ __forceinline__ int mg_stencil_of_cells(int k0, int k1, MgDims d) {
    int x0, y0, z0, x1, y1, z1;
    mg_decode(k0, d, x0, y0, z0);
    mg_decode(k1, d, x1, y1, z1);
    const int dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1 || (W == 9 && dz != 0)) return -1;
    int j = (dx + 1) + 3 * (dy + 1);
    if (W == 27) j += 9 * (dz + 1);
    return j;
}

// Galerkin level 1 from the fine SELL matrix (unit diagonal not stored;
// padding entries point at their own row with value 0).  Aggregate K is a
// bucket cell of scene scene1[K]; rkey holds each fine row's cell key.  A block takes A consecutive
// aggregates, whose rows are one contiguous range, in chunks of kAsm1Rows
// rows: one thread per row sums its entries into the row's stencil bins
// (LDS, entry order), then one thread per (aggregate, bin) adds the rows'
// bins in row order.  Row reads are coalesced; the sums are in a fixed order.
constexpr int kAsm1Rows = 256;  // threads per block, rows per chunk
constexpr int kMgBatch = 9;     // ELL entries per round of loads in the coarse assembly (W is a multiple)

template <typename T, typename IDX, int W, int A>
__global__ void __launch_bounds__(kAsm1Rows) k_mg_assemble1(int n1, const int* __restrict__ cst1,
                                                            const int* __restrict__ slice_pos,
                                                            const IDX* __restrict__ col, const int* __restrict__ anc,
                                                            const T* __restrict__ val,
                                                            const int* __restrict__ agg0,
                                                            const int* __restrict__ rkey,
                                                            const int* __restrict__ scene1,
                                                            const MgDims* __restrict__ dims, int cap1,
                                                            int* __restrict__ col1, T* __restrict__ val1,
                                                            T* __restrict__ diag1, T* __restrict__ dinv1,
                                                            int* __restrict__ err) {
    __shared__ double s_bin[kAsm1Rows][W + 1];  // [W]: the row's entries inside its own aggregate
    __shared__ double s_acc[A][W + 1];
    __shared__ int s_nbr[A][W];                 // the aggregate at each stencil place (-1: none)
    __shared__ int s_cst[A + 1];
    __shared__ MgDims s_dim[A];
    const int tid = threadIdx.x;
    const int K0 = blockIdx.x * A;
    const int na = n1 - K0 < A ? n1 - K0 : A;
    if (na <= 0) return;  // block-uniform
    if (tid <= na) s_cst[tid] = cst1[K0 + tid];
    if (tid < na) {
        s_dim[tid] = dims[scene1[K0 + tid]];
    }
    for (int i = tid; i < A * W; i += kAsm1Rows) s_nbr[i / W][i % W] = -1;
    for (int i = tid; i < A * (W + 1); i += kAsm1Rows) s_acc[i / (W + 1)][i % (W + 1)] = 0.0;
    __syncthreads();
    const int r0 = s_cst[0], r1 = s_cst[na];
    for (int c0 = r0; c0 < r1; c0 += kAsm1Rows) {
        const int r = c0 + tid;
        for (int j = 0; j <= W; j++) s_bin[tid][j] = 0.0;
        if (r < r1) {
            int a = 0, hi = na - 1;  // the last aggregate whose first row is <= r
            while (a < hi) {
                const int mid = (a + hi + 1) >> 1;
                if (s_cst[mid] <= r)
                    a = mid;
                else
                    hi = mid - 1;
            }
            const int K = K0 + a, rk = rkey[r];
            const MgDims d = s_dim[a];
            const int sl = r >> 6, lane = r & 63;
            const int sp = slice_pos[sl], w = slice_pos[sl + 1] - sp;
            const size_t base = (size_t) sp * 64 + lane;
            int alo, ahi;
            sell_anchors<IDX>(anc, r, alo, ahi);
            // four entries per round: columns and values, then the columns'
            // aggregates and cell keys (a row's cell is its aggregate's, so
            // the key difference to the row's own gives the stencil place),
            // each round's loads issued together (entries past w re-read the
            // last and are dropped)
            for (int k0 = 0; k0 < w; k0 += 4) {
                int K2[4], cd[4];
                double v[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const size_t idx = base + (size_t) (k0 + q < w ? k0 + q : w - 1) * 64;
                    K2[q] = sell_col(col, idx, r, alo, ahi);
                    v[q] = (double) val[idx];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    cd[q] = rkey[K2[q]];
                    K2[q] = agg0[K2[q]];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (k0 + q >= w) break;
                    if (cd[q] == rk) {  // the row's own cell: its own aggregate
                        s_bin[tid][W] += v[q];
                        continue;
                    }
                    const int j = mg_stencil_of_dkey<W>(cd[q] - rk, d);
                    if (j < 0) {
                        atomicOr(err, 1);
                        continue;
                    }
                    s_bin[tid][j] += v[q];
                    s_nbr[a][j] = K2[q];  // every writer of this place writes the same aggregate
                }
            }
        }
        __syncthreads();
        for (int t = tid; t < na * (W + 1); t += kAsm1Rows) {
            const int a = t / (W + 1), j = t % (W + 1);
            const int ra = s_cst[a] > c0 ? s_cst[a] : c0;
            const int rb = s_cst[a + 1] < c0 + kAsm1Rows ? s_cst[a + 1] : c0 + kAsm1Rows;
            double acc = s_acc[a][j];
            for (int q = ra; q < rb; q++) acc += j == W ? 1.0 + s_bin[q - c0][W] : s_bin[q - c0][j];
            s_acc[a][j] = acc;
        }
        __syncthreads();
    }
    // neighbours in stencil order, then padding
    for (int t = tid; t < na * W; t += kAsm1Rows) {
        const int a = t / W, j = t % W;
        const int K = K0 + a;
        int m = 0, slot = 0;
        for (int q = 0; q < W; q++) {
            const bool h = s_nbr[a][q] >= 0;
            m += h ? 1 : 0;
            slot += (h && q < j) ? 1 : 0;
        }
        if (s_nbr[a][j] >= 0) {
            col1[(size_t) slot * cap1 + K] = s_nbr[a][j];
            val1[(size_t) slot * cap1 + K] = (T) s_acc[a][j];
        }
        if (j >= m) {
            col1[(size_t) j * cap1 + K] = -1;
            val1[(size_t) j * cap1 + K] = (T) 0;
        }
    }
    if (tid < na) {
        const double dg = s_acc[tid][W];
        diag1[K0 + tid] = (T) dg;
        dinv1[K0 + tid] = (T) (1.0 / dg);
    }
}

// Tables for building the structure of all coarse levels in one pass each
// (levels lo..hi; per level its scenes' grids and first dense indices, the
// first cell of its grid in the joint flag array of levels 2.., the joint
// scan's value there, the first row of its rows in the joint row list of
// levels 1.., and its row arrays).
struct MgTabs {
    const MgDims* sdims[kMgMaxLevels + 1];
    const int* goff[kMgMaxLevels + 1];
    int foff[kMgMaxLevels + 2];
    int pbase[kMgMaxLevels + 2];
    int rbase[kMgMaxLevels + 2];
    int cbase[kMgMaxLevels + 2];
    int* coord[kMgMaxLevels + 1];
    int* scene[kMgMaxLevels + 1];
    int* agg[kMgMaxLevels + 1];
    int* child[kMgMaxLevels + 1];
    int cap[kMgMaxLevels + 1];
    int lo, hi, nscene, C;
};

// The level (lo..hi) whose range [X[l], X[l + 1]) holds t.
__device__ __forceinline__ int mg_level_of(const int* X, int lo, int hi, int t) {
    int l = lo;
    while (l < hi && t >= X[l + 1]) l++;
    return l;
}

// Every level-1 cell flags the cell above it on each level 2..hi (level l
// merges 2^(l-1) level-1 cells per axis: its coordinates are c >> (l - 1)).
__global__ void __launch_bounds__(kBlock) k_mg_mark_all(int n1, const int* __restrict__ coord1,
                                                        const int* __restrict__ scene1,
                                                        const MgDims* __restrict__ d1, MgTabs T,
                                                        int* __restrict__ flag) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= n1) return;
    const int sc = scene1[K];
    int cx, cy, cz;
    mg_decode(coord1[K], d1[sc], cx, cy, cz);
    for (int l = 2; l <= T.hi; l++) {
        const int sh = l - 1;
        const MgDims q = T.sdims[l][sc];
        flag[T.foff[l] + T.goff[l][sc] + (cx >> sh) + q.x * ((cy >> sh) + q.y * (cz >> sh))] = 1;
    }
}

// Rows of each level 2..hi: differences of the joint scan at the grids' starts.
__global__ void k_mg_sizes(const int* __restrict__ pscan, MgTabs T, int* __restrict__ out) {
    const int l = 2 + (int) threadIdx.x;
    if (l <= T.hi) out[l] = pscan[T.foff[l + 1]] - pscan[T.foff[l]];
}

// Each flagged grid cell of levels 2..hi becomes a row: its cell and scene.
__global__ void __launch_bounds__(kBlock) k_mg_parents_all(const int* __restrict__ flag,
                                                           const int* __restrict__ pscan, MgTabs T) {
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= T.foff[T.hi + 1] || !flag[P]) return;
    const int l = mg_level_of(T.foff, 2, T.hi, P);
    const int loc = P - T.foff[l];
    const int sc = mg_bucket(T.goff[l], T.nscene, loc);
    const int r = pscan[P] - T.pbase[l];
    T.coord[l][r] = loc - T.goff[l][sc];
    T.scene[l][r] = sc;
}

// Child tables of levels 2..hi to -1 (empty slots).
__global__ void __launch_bounds__(kBlock) k_mg_clear_children(MgTabs T) {
    const int tot = T.cbase[T.hi + 1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        const int l = mg_level_of(T.cbase, 2, T.hi, i);
        T.child[l][i - T.cbase[l]] = -1;
    }
}

// Rows of levels lo..hi - 1: agg = parent row on the level above, and the
// parent's child slot (cx&1) + 2(cy&1) + 4(cz&1) = the row.
__global__ void __launch_bounds__(kBlock) k_mg_link_all(const int* __restrict__ pscan, MgTabs T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.rbase[T.hi]) return;
    const int l = mg_level_of(T.rbase, T.lo, T.hi - 1, t);
    const int K = t - T.rbase[l];
    const int sc = T.scene[l][K];
    int cx, cy, cz;
    mg_decode(T.coord[l][K], T.sdims[l][sc], cx, cy, cz);
    const int slot = (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1);
    const MgDims q = T.sdims[l + 1][sc];
    const int key = T.goff[l + 1][sc] + (cx >> 1) + q.x * ((cy >> 1) + q.y * (cz >> 1));
    const int pr = pscan[T.foff[l + 1] + key] - T.pbase[l + 1];
    T.agg[l][K] = pr;
    T.child[l + 1][(size_t) slot * T.cap[l + 1] + pr] = K;
}

// Galerkin level l + 1 from level l: parent P sums its children's rows.  A
// block takes kAsm1Rows / C parents: one thread per (parent, child slot) sums
// the child's entries into its stencil bins (LDS, entry order), then one
// thread per (parent, bin) adds the children's bins in slot order.
template <typename T, int W, int C>
__global__ void __launch_bounds__(kAsm1Rows) k_mg_assemble(int np, const int* __restrict__ child, int capp,
                                                           const int* __restrict__ col, const T* __restrict__ val,
                                                           const T* __restrict__ diag, int cap,
                                                           const int* __restrict__ agg,
                                                           const int* __restrict__ coordp,
                                                           const int* __restrict__ scenep,
                                                           const MgDims* __restrict__ dp, int* __restrict__ colp,
                                                           T* __restrict__ valp, T* __restrict__ diagp,
                                                           T* __restrict__ dinvp, int* __restrict__ err) {
    static_assert(W % kMgBatch == 0, "stencil width is a multiple of the batch");
    constexpr int A = kAsm1Rows / C;
    __shared__ double s_bin[kAsm1Rows][W + 1];  // [W]: the child's diagonal and entries inside P
    __shared__ int s_nbr[A][W];
    const int tid = threadIdx.x;
    const int P0 = blockIdx.x * A;
    const int na = np - P0 < A ? np - P0 : A;
    if (na <= 0) return;  // block-uniform
    for (int i = tid; i < A * W; i += kAsm1Rows) s_nbr[i / W][i % W] = -1;
    for (int j = 0; j <= W; j++) s_bin[tid][j] = 0.0;
    __syncthreads();
    {
        const int a = tid / C, q = tid % C;
        const int P = P0 + a;
        const int K = a < na ? child[(size_t) q * capp + P] : -1;
        if (K >= 0) {
            const MgDims d = dp[scenep[P]];
            const int key = coordp[P];
            s_bin[tid][W] = (double) diag[K];
            // the row's entries in rounds of kMgBatch: columns and values,
            // then their parents, then the parents' cells, each round's
            // loads issued together (padding, column -1, is read at K and
            // ends the row)
            for (int j0 = 0; j0 < W; j0 += kMgBatch) {
                int cj[kMgBatch], pj[kMgBatch], kj[kMgBatch];
                double vj[kMgBatch];
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) {
                    cj[q] = col[(size_t) (j0 + q) * cap + K];
                    vj[q] = (double) val[(size_t) (j0 + q) * cap + K];
                }
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) pj[q] = agg[cj[q] >= 0 ? cj[q] : K];
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) kj[q] = coordp[pj[q]];
                bool end = false;
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) {
                    if (cj[q] < 0) {
                        end = true;
                        break;
                    }
                    const int P2 = pj[q];
                    if (P2 == P) {
                        s_bin[tid][W] += vj[q];
                        continue;
                    }
                    const int sj = mg_stencil_of_cells<W>(key, kj[q], d);
                    if (sj < 0) {
                        atomicOr(err, 1);
                        continue;
                    }
                    s_bin[tid][sj] += vj[q];
                    s_nbr[a][sj] = P2;  // every writer of this place writes the same parent
                }
                if (end) break;
            }
        }
    }
    __syncthreads();
    // reduce the children's bins into the first child's row of LDS
    for (int t = tid; t < na * (W + 1); t += kAsm1Rows) {
        const int a = t / (W + 1), j = t % (W + 1);
        double acc = 0.0;
        for (int q = 0; q < C; q++) acc += s_bin[a * C + q][j];
        s_bin[a * C][j] = acc;  // only this thread reads or writes [a * C][j] now
    }
    __syncthreads();
    for (int t = tid; t < na * W; t += kAsm1Rows) {
        const int a = t / W, j = t % W;
        const int P = P0 + a;
        int m = 0, slot = 0;
        for (int q = 0; q < W; q++) {
            const bool h = s_nbr[a][q] >= 0;
            m += h ? 1 : 0;
            slot += (h && q < j) ? 1 : 0;
        }
        if (s_nbr[a][j] >= 0) {
            colp[(size_t) slot * capp + P] = s_nbr[a][j];
            valp[(size_t) slot * capp + P] = (T) s_bin[a * C][j];
        }
        if (j >= m) {
            colp[(size_t) j * capp + P] = -1;
            valp[(size_t) j * capp + P] = (T) 0;
        }
    }
    if (tid < na) {
        const double dg = s_bin[tid * C][W];
        diagp[P0 + tid] = (T) dg;
        dinvp[P0 + tid] = (T) (1.0 / dg);
    }
}

// ---------------- the V-cycle ----------------

// acc + sum over fine row r's entries of Ahat_rc x_c, in entry order (SELL-64
// slices); entries are read four at a time, the loads of a group issued
// before any is used.
template <typename T, typename IDX>
__device__ __forceinline__ T mg_fine_row(const int* __restrict__ slice_pos, const IDX* __restrict__ col,
                                         const int* __restrict__ anc, const T* __restrict__ val,
                                         const T* __restrict__ x, int r, T acc) {
#pragma clang fp contract(fast)
    const int sl = r >> 6, lane = r & 63;
    const int s0 = slice_pos[sl], w = slice_pos[sl + 1] - s0;
    const size_t base = (size_t) s0 * 64 + lane;
    int lo, hi;
    sell_anchors<IDX>(anc, r, lo, hi);
    constexpr bool NT = sizeof(T) == 8;  // the FP64 matrix: streamed
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const size_t i0 = base + (size_t) k * 64;
        const int c0 = sell_col<IDX, NT>(col, i0, r, lo, hi), c1 = sell_col<IDX, NT>(col, i0 + 64, r, lo, hi);
        const int c2 = sell_col<IDX, NT>(col, i0 + 128, r, lo, hi), c3 = sell_col<IDX, NT>(col, i0 + 192, r, lo, hi);
        const T v0 = ld_stream<NT>(val + i0), v1 = ld_stream<NT>(val + i0 + 64);
        const T v2 = ld_stream<NT>(val + i0 + 128), v3 = ld_stream<NT>(val + i0 + 192);
        const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc += v0 * x0;
        acc += v1 * x1;
        acc += v2 * x2;
        acc += v3 * x3;
    }
    for (; k < w; k++) {
        const size_t idx = base + (size_t) k * 64;
        acc += ld_stream<NT>(val + idx) * x[sell_col<IDX, NT>(col, idx, r, lo, hi)];
    }
    return acc;
}

// Fine level, down: x0 = omega b (unit diagonal); res = b - Ahat x0.
template <typename T, typename IDX>
__global__ void __launch_bounds__(kBlock) k_mg_down0(const int* __restrict__ nrow, const int* __restrict__ slice_pos,
                                                     const IDX* __restrict__ col, const int* __restrict__ anc,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ b, T* __restrict__ res, T omega,
                                                     MgDone done) {
#pragma clang fp contract(fast)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = *nrow;
    if (mg_done_wave(done) || r >= nr) return;
    const T br = b[r];
    const T acc = mg_fine_row<T, IDX>(slice_pos, col, anc, val, b, r, br);
    res[r] = br - omega * acc;
}

// Restriction to level 1: sum of the cell's fine residuals.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_restrict1(const int* __restrict__ nrow, const int* __restrict__ cst1,
                                                         const T* __restrict__ res, T* __restrict__ b1,
                                                         MgDone done) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = *nrow;
    if (mg_done_wave(done) || K >= nr) return;
    b1[K] = cell_sum(res, cst1[K], cst1[K + 1]);
}

// Coarse levels are small: their kernels are bound by the round trips of
// dependent loads, not by bytes.  A row's W ELL entries are read without a
// branch, every load of a round issued before any is used: padding (column
// -1, value 0, after the row's columns) is read at the row's own index and
// left out of the sum, which keeps the entries' order of summation.

// Residual of level-l row K at x = omega D^-1 b (x recomputed where read).
template <typename T, int W>
__device__ __forceinline__ T mg_child_residual(int K, const int* __restrict__ col, const T* __restrict__ val,
                                               const T* __restrict__ diag, const T* __restrict__ dinv, int cap,
                                               const T* __restrict__ b, T omega) {
#pragma clang fp contract(fast)
    int c[W];
    T v[W], bb[W], dd[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
        c[j] = col[(size_t) j * cap + K];
        v[j] = val[(size_t) j * cap + K];
    }
    const T bK = b[K], gK = diag[K], dK = dinv[K];
#pragma unroll
    for (int j = 0; j < W; j++) {
        const int cc = c[j] >= 0 ? c[j] : K;
        bb[j] = b[cc];
        dd[j] = dinv[cc];
    }
    T acc = gK * (omega * bK * dK);
#pragma unroll
    for (int j = 0; j < W; j++)
        if (c[j] >= 0) acc += v[j] * (omega * bb[j] * dd[j]);
    return bK - acc;
}

// Coarse level l, down: b_{l+1}[P] = sum over P's children K of (b - A x)_K.
// One lane per child slot, the C lanes of a parent summed in a fixed tree.
// The level size, the done flags and the child are read in one round (the
// child is in range for every P < capp and used only for P < np).
template <typename T, int W, int C>
__global__ void __launch_bounds__(kBlock) k_mg_down(const int* __restrict__ nparent, const int* __restrict__ child,
                                                    int capp, const int* __restrict__ col, const T* __restrict__ val,
                                                    const T* __restrict__ diag, const T* __restrict__ dinv, int cap,
                                                    const T* __restrict__ b, T* __restrict__ bp, T omega,
                                                    MgDone done) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = t / C, q = t % C;
    const int np = *nparent;
    const int Kc = P < capp ? child[(size_t) q * capp + P] : -1;
    if (mg_done_wave(done)) return;  // wave-uniform
    const bool live = P < np;
    T v = (T) 0;
    if (live && Kc >= 0) v = mg_child_residual<T, W>(Kc, col, val, diag, dinv, cap, b, omega);
    for (int o = C / 2; o > 0; o >>= 1) v += __shfl_down(v, o, C);
    if (live && q == 0) bp[P] = v;
}

// Coarse level l, up at row K: xx = omega D^-1 b + cf x_{l+1}[agg] (recomputed
// where read), then x = xx + omega D^-1 (b - A xx).  The row's own values
// and entries are read in the round that reads the level size and the done
// flags (in range for every K < cap; rows past cap read row 0 and store
// nothing).  PRO (level 1): also the fine level's first step up for the
// cell's rows cst1[K] .. cst1[K + 1] - 1: xxf = omega bf + cf x[K].
template <typename T, int W, bool PRO>
__global__ void __launch_bounds__(kBlock) k_mg_up(const int* __restrict__ nrow, const int* __restrict__ col,
                                                  const T* __restrict__ val, const T* __restrict__ diag,
                                                  const T* __restrict__ dinv, int cap, const int* __restrict__ agg,
                                                  const T* __restrict__ b, const T* __restrict__ xc,
                                                  T* __restrict__ x, T omega, T cf, MgDone done,
                                                  const int* __restrict__ cst1, const T* __restrict__ bf,
                                                  T* __restrict__ xxf, const unsigned char* __restrict__ aggf) {
#pragma clang fp contract(fast)
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    const int Kr = K < cap ? K : 0;
    const int nr = *nrow;
    int c[W], ag[W];
    T v[W], bb[W], dd[W], xv[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
        c[j] = col[(size_t) j * cap + Kr];
        v[j] = val[(size_t) j * cap + Kr];
    }
    const T bK = b[Kr], dK = dinv[Kr], gK = diag[Kr];
    const int aK = agg[Kr];
    // the done flags are the whole launch's, so every block takes the same branch
    if (mg_done_wave(done)) return;
    T xK = (T) 0;
    if (K < nr) {
#pragma unroll
        for (int j = 0; j < W; j++) {
            const int cc = c[j] >= 0 ? c[j] : K;
            bb[j] = b[cc];
            dd[j] = dinv[cc];
            ag[j] = agg[cc];
        }
        const T xcK = xc[aK];
#pragma unroll
        for (int j = 0; j < W; j++) xv[j] = xc[ag[j]];
        const T xx = omega * bK * dK + cf * xcK;
        T acc = gK * xx;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (c[j] >= 0) acc += v[j] * (omega * bb[j] * dd[j] + cf * xv[j]);
        xK = xx + omega * dK * (bK - acc);
        x[K] = xK;
    }
    if (PRO) {
        // the fine level's first step up, xxf = omega bf + cf x[cell], over
        // the block's cells' fine rows (one contiguous range) with
        // consecutive threads on consecutive rows
        __shared__ T s_x[kBlock];
        s_x[threadIdx.x] = xK;
        __syncthreads();
        const int K0 = blockIdx.x * blockDim.x;
        const int K1 = K0 + (int) blockDim.x < nr ? K0 + (int) blockDim.x : nr;
        if (K0 >= K1) return;
        const int f0 = cst1[K0], f1 = cst1[K1];
        // the low byte of a row's cell is its place among the block's cells
        static_assert(kBlock == 256, "one byte holds a cell's place in the block");
        for (int r = f0 + (int) threadIdx.x; r < f1; r += blockDim.x) xxf[r] = omega * bf[r] + cf * s_x[aggf[r]];
    }
}

// Device table of one level's arrays, for the tail kernel.
struct MgLevelDesc {
    int cap;
    int pad;
    const int* col;
    const void* val;
    const void* diag;
    const void* dinv;
    const int* agg;
    const int* child;
    void* b;
    void* x;
};

constexpr int kMgTailRows = 1024;       // levels this small may run inside the tail kernel
constexpr size_t kMgAdditiveEntries = 22u << 20;  // fine entries from which single scenes take the additive cycle
constexpr int kMgTailThreads = 1024;
constexpr int kMgTailLds = 156 * 1024;  // dynamic LDS of the tail kernel

// LDS bytes one row of a tail level takes: ELL columns (16-bit) and values,
// diagonal, inverse diagonal, b, x, a second x (coarsest sweeps), the parent
// index and the children indices (16-bit).
__host__ __device__ constexpr int mg_tail_row_bytes(int W, int C, int tsize) {
    return W * (2 + tsize) + 5 * tsize + 2 + 2 * C;
}

// One tail level's arrays in LDS for n rows (16-byte aligned pieces).
template <typename T, int W, int C>
struct MgLdsLevel {
    short* col;  // W x n
    T* val;      // W x n
    T* diag;
    T* dinv;
    T* b;
    T* x;
    T* x2;
    short* agg;
    short* child;  // C x n
    __device__ static int align16(int v) { return (v + 15) & ~15; }
    __device__ static int bytes(int n) {
        return align16(W * n * 2) + align16(W * n * (int) sizeof(T)) + 5 * align16(n * (int) sizeof(T)) +
               align16(n * 2) + align16(C * n * 2);
    }
    __device__ void carve(unsigned char* base, int n) {
        int o = 0;
        col = (short*) (base + o);
        o += align16(W * n * 2);
        val = (T*) (base + o);
        o += align16(W * n * (int) sizeof(T));
        diag = (T*) (base + o);
        o += align16(n * (int) sizeof(T));
        dinv = (T*) (base + o);
        o += align16(n * (int) sizeof(T));
        b = (T*) (base + o);
        o += align16(n * (int) sizeof(T));
        x = (T*) (base + o);
        o += align16(n * (int) sizeof(T));
        x2 = (T*) (base + o);
        o += align16(n * (int) sizeof(T));
        agg = (short*) (base + o);
        o += align16(n * 2);
        child = (short*) (base + o);
    }
};

// Levels l0..L in one workgroup, out of LDS: every level's operator, transfer
// maps and b[l0] are loaded once, then the restrictions down to the
// coarsest, `sweeps` damped Jacobi sweeps there and the prolongation +
// smoothing phases back to level l0 run on LDS, separated by workgroup
// barriers; x[l0] goes back to global memory for the next kernel.  The load
// takes the rows of all levels as one list, each thread reading all of a
// row's values into registers before it writes any to LDS, so the whole
// load costs about one round trip per 1024 rows.
template <typename T, int W, int C>
__global__ void __launch_bounds__(kMgTailThreads) k_mg_tail(const MgLevelDesc* __restrict__ lv,
                                                            const int* __restrict__ nd, int l0, int L, T omega, T cf,
                                                            int sweeps, MgDone done) {
#pragma clang fp contract(fast)
    static_assert(W >= C, "child reads of the first level fall inside its column array");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_off[kMgMaxLevels + 2];
    __shared__ int s_row[kMgMaxLevels + 2];  // first row of each level in the joint list
    __shared__ int s_n[kMgMaxLevels + 1];
    __shared__ MgLevelDesc s_lv[kMgMaxLevels + 1];
    const int tid = threadIdx.x;
    if (tid <= L - l0) {
        s_n[l0 + tid] = nd[l0 + tid];
        s_lv[l0 + tid] = lv[l0 + tid];
    }
    if (mg_done(done)) return;  // its barrier also publishes the writes above
    if (tid == 0) {
        int o = 0, r = 0;
        for (int l = l0; l <= L; l++) {
            s_off[l] = o;
            s_row[l] = r;
            o += MgLdsLevel<T, W, C>::bytes(s_n[l]);
            r += s_n[l];
        }
        s_row[L + 1] = r;
    }
    __syncthreads();
    auto level = [&](int l) {
        MgLdsLevel<T, W, C> q;
        q.carve(smem + s_off[l], s_n[l]);
        return q;
    };
    // load
    for (int t = tid; t < s_row[L + 1]; t += kMgTailThreads) {
        int l = l0;
        while (t >= s_row[l + 1]) l++;
        const int K = t - s_row[l];
        const MgLevelDesc& d = s_lv[l];
        // arrays a level does not have are read at in-range places of its
        // column array and the values dropped
        const int* agp = l < L ? d.agg : d.col;
        const int* chp = l > l0 ? d.child : d.col;
        const T* bp = l == l0 ? (const T*) d.b : (const T*) d.val;
        int cc[W], ch[C];
        T vv[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            cc[j] = d.col[(size_t) j * d.cap + K];
            vv[j] = ((const T*) d.val)[(size_t) j * d.cap + K];
        }
#pragma unroll
        for (int c = 0; c < C; c++) ch[c] = chp[(size_t) c * d.cap + K];
        const T dg = ((const T*) d.diag)[K], di = ((const T*) d.dinv)[K], bb = bp[K];
        const int ag = agp[K];
        const int n = s_n[l];
        const MgLdsLevel<T, W, C> q = level(l);
#pragma unroll
        for (int j = 0; j < W; j++) {
            q.col[j * n + K] = (short) cc[j];
            q.val[j * n + K] = vv[j];
        }
        q.diag[K] = dg;
        q.dinv[K] = di;
        if (l == l0) q.b[K] = bb;
        if (l < L) q.agg[K] = (short) ag;
        if (l > l0)
#pragma unroll
            for (int c = 0; c < C; c++) q.child[c * n + K] = (short) ch[c];
    }
    __syncthreads();
    // down
    for (int l = l0; l < L; l++) {
        const MgLdsLevel<T, W, C> a = level(l), p = level(l + 1);
        const int n = s_n[l], np = s_n[l + 1];
        const int items = np * C;
        for (int base = 0; base < items; base += kMgTailThreads) {
            const int t = base + tid;
            const int P = t / C, qs = t % C;
            T v = (T) 0;
            if (P < np) {
                const int K = p.child[qs * np + P];
                if (K >= 0) {
                    const T bK = a.b[K];
                    T acc = a.diag[K] * (omega * bK * a.dinv[K]);
                    for (int j0 = 0; j0 < W; j0 += kMgBatch) {
                        int c[kMgBatch];
                        T vv[kMgBatch], xb[kMgBatch];
#pragma unroll
                        for (int q = 0; q < kMgBatch; q++) {
                            c[q] = a.col[(j0 + q) * n + K];
                            vv[q] = a.val[(j0 + q) * n + K];
                        }
                        if (W > kMgBatch && c[0] < 0) break;
#pragma unroll
                        for (int q = 0; q < kMgBatch; q++) {
                            const int cc = c[q] >= 0 ? c[q] : K;
                            xb[q] = omega * a.b[cc] * a.dinv[cc];
                        }
#pragma unroll
                        for (int q = 0; q < kMgBatch; q++)
                            if (c[q] >= 0) acc += vv[q] * xb[q];
                    }
                    v = bK - acc;
                }
            }
            for (int o = C / 2; o > 0; o >>= 1) v += __shfl_down(v, o, C);
            if (P < np && qs == 0) p.b[P] = v;
        }
        __syncthreads();
    }
    // coarsest
    {
        const MgLdsLevel<T, W, C> c = level(L);
        const int n = s_n[L];
        T* xa = c.x;
        T* xb = c.x2;
        for (int K = tid; K < n; K += kMgTailThreads) xa[K] = (T) 0;
        __syncthreads();
        for (int it = 0; it < sweeps; it++) {
            for (int K = tid; K < n; K += kMgTailThreads) {
                const T xt = xa[K];
                T acc = c.diag[K] * xt;
                for (int j0 = 0; j0 < W; j0 += kMgBatch) {
                    int cj[kMgBatch];
                    T vv[kMgBatch], xv[kMgBatch];
#pragma unroll
                    for (int q = 0; q < kMgBatch; q++) {
                        cj[q] = c.col[(j0 + q) * n + K];
                        vv[q] = c.val[(j0 + q) * n + K];
                    }
                    if (W > kMgBatch && cj[0] < 0) break;
#pragma unroll
                    for (int q = 0; q < kMgBatch; q++) xv[q] = xa[cj[q] >= 0 ? cj[q] : K];
#pragma unroll
                    for (int q = 0; q < kMgBatch; q++)
                        if (cj[q] >= 0) acc += vv[q] * xv[q];
                }
                xb[K] = xt + omega * c.dinv[K] * (c.b[K] - acc);
            }
            __syncthreads();
            T* t = xa;
            xa = xb;
            xb = t;
        }
        if (xa != c.x)
            for (int K = tid; K < n; K += kMgTailThreads) c.x[K] = xa[K];
        __syncthreads();
    }
    // up (a level's own x is written, never read, in its phase)
    for (int l = L - 1; l >= l0; l--) {
        const MgLdsLevel<T, W, C> a = level(l);
        const T* xc = level(l + 1).x;
        const int n = s_n[l];
        for (int K = tid; K < n; K += kMgTailThreads) {
            const T bK = a.b[K], dK = a.dinv[K];
            const T xx = omega * bK * dK + cf * xc[a.agg[K]];
            T acc = a.diag[K] * xx;
            for (int j0 = 0; j0 < W; j0 += kMgBatch) {
                int c[kMgBatch], ag[kMgBatch];
                T vv[kMgBatch], bb[kMgBatch], dd[kMgBatch], xv[kMgBatch];
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) {
                    c[q] = a.col[(j0 + q) * n + K];
                    vv[q] = a.val[(j0 + q) * n + K];
                }
                if (W > kMgBatch && c[0] < 0) break;
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) {
                    const int cc = c[q] >= 0 ? c[q] : K;
                    bb[q] = a.b[cc];
                    dd[q] = a.dinv[cc];
                    ag[q] = a.agg[cc];
                }
#pragma unroll
                for (int q = 0; q < kMgBatch; q++) xv[q] = xc[ag[q]];
#pragma unroll
                for (int q = 0; q < kMgBatch; q++)
                    if (c[q] >= 0) acc += vv[q] * (omega * bb[q] * dd[q] + cf * xv[q]);
            }
            a.x[K] = xx + omega * dK * (bK - acc);
        }
        __syncthreads();
    }
    T* xo = (T*) lv[l0].x;
    const MgLdsLevel<T, W, C> top = level(l0);
    for (int K = tid; K < s_n[l0]; K += kMgTailThreads) xo[K] = top.x[K];
}

// Fine level, up: xx = omega b + cf x1[agg0] ...
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_prolong0(const int* __restrict__ nrow, const T* __restrict__ b,
                                                        const int* __restrict__ agg0, const T* __restrict__ x1,
                                                        T* __restrict__ xx, T omega, T cf,
                                                        MgDone done) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = *nrow;
    if (mg_done_wave(done) || r >= nr) return;
    xx[r] = omega * b[r] + cf * x1[agg0[r]];
}

// ... then u = xx + omega (b - Ahat xx).
template <typename T, typename IDX>
__global__ void __launch_bounds__(kBlock) k_mg_up0(const int* __restrict__ nrow, const int* __restrict__ slice_pos,
                                                   const IDX* __restrict__ col, const int* __restrict__ anc,
                                                   const T* __restrict__ val,
                                                   const T* __restrict__ b, const T* __restrict__ xx,
                                                   T* __restrict__ u, T omega, MgDone done) {
#pragma clang fp contract(fast)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = *nrow;
    if (mg_done_wave(done) || r >= nr) return;
    const T xr = xx[r], br = b[r];
    const T acc = mg_fine_row<T, IDX>(slice_pos, col, anc, val, xx, r, xr);
    u[r] = xr + omega * (br - acc);
}

// ---------------- host side ----------------

struct MgWork {
    int L = 0;                       // coarse levels 1..L (L = coarsest)
    int W = 9, C = 4;
    int esize = 8;                   // bytes per value the buffers are sized for
    int n[kMgMaxLevels + 1] = {};    // rows per level (host copy of the last setup)
    int cap[kMgMaxLevels + 1] = {};
    int nscene = 1;
    std::vector<MgDims> sdims[kMgMaxLevels + 1];  // per level: each scene's grid
    std::vector<int> goff[kMgMaxLevels + 1];      // per level: each scene's first dense index (+ total)
    MgDims* d_sdims[kMgMaxLevels + 1] = {};
    int* d_goff[kMgMaxLevels + 1] = {};
    int scap = 0;                    // scenes the device tables hold
    int* scene[kMgMaxLevels + 1] = {};  // per level: scene of each row
    int gen = 0;                     // bumped whenever a launch argument changes
    int* nd = nullptr;               // device row counts, kMgMaxLevels + 1
    int* err = nullptr;
    // per level (index l): agg[l]: level l rows -> level l+1 rows
    int* agg[kMgMaxLevels + 1] = {};
    int* coord[kMgMaxLevels + 1] = {};
    int* child[kMgMaxLevels + 1] = {};  // level l >= 2: C x cap children
    int* col[kMgMaxLevels + 1] = {};
    void* val[kMgMaxLevels + 1] = {};
    void* diag[kMgMaxLevels + 1] = {};
    void* dinv[kMgMaxLevels + 1] = {};
    void* b[kMgMaxLevels + 1] = {};
    void* x[kMgMaxLevels + 1] = {};
    int* cst1 = nullptr;
    int fcap = 0;                    // fine rows the fine scratch holds
    int* head = nullptr;
    int* hscan = nullptr;
    int* rkey = nullptr;
    unsigned char* agg0b = nullptr;  // low byte of each fine row's cell
    void* res0 = nullptr;
    void* xx0 = nullptr;
    int gcap = 0;                    // dense grid scratch
    int* lsz = nullptr;              // device: rows per level from the joint scan
    int* flag = nullptr;
    int* pscan = nullptr;
    int* scan_tmp = nullptr;
    size_t scan_cap = 0;
    MgLevelDesc* desc = nullptr;     // device table of the levels
    MgLevelDesc hdesc[kMgMaxLevels + 1] = {};
    // what the device copies of the per-level tables last received: the grid
    // tables depend only on the scenes, so most setups upload nothing (an
    // upload from pageable memory waits for the stream)
    std::vector<MgDims> up_sdims[kMgMaxLevels + 1];
    std::vector<int> up_goff[kMgMaxLevels + 1];
    MgLevelDesc up_desc[kMgMaxLevels + 1] = {};
    bool up_desc_valid = false;
    Mailbox mb;  // level-size reads
    int tail = 1;                    // first level handled by the tail kernel
    // smoother weight, coarse-correction factor and coarsest sweeps, chosen by
    // measurement: 2D (2M dam break, 64-scene ensemble) 1.0, 1.4, 1; 3D
    // (3.6M dam break) 1.0, 1.3, 8.  Jacobi with weight 1 is a convergent
    // smoother here: the scaled matrix has unit diagonal and off-diagonal row
    // sums below 1, so its eigenvalues lie in (0, 2).
    double omega = 1.0, cf = 1.4;
    int sweeps = 1;
    bool user_params = false;        // set by mg_set_params / MPS_AMD_MG
    // Fine level of the cycle.  Multiplicative (the symmetric V(1,1)): Jacobi
    // from zero, restriction of the residual, coarse cycle, prolongation,
    // Jacobi: u = x + omega (r - Ahat x), x = omega r + cf P c(R (r - omega Ahat r)),
    // two fine SpMVs.  Additive: u = omega r + cf P c(R r), no fine SpMV;
    // still symmetric positive definite (the sum of omega I and cf P C P^T
    // with C the symmetric coarse cycle), so the CG stays a CG.  The additive
    // cycle needs 1.5-1.9x the iterations and saves two fine SpMVs each.
    // Measured (single dam breaks, steps 5-35): it is 2-6 % faster at the 2D
    // 2M (25M matrix entries) and 7-9 % at the 3D 3.6M (85M), even at 3D
    // 1.1-1.5M (25-35M), 4-20 % slower at 2D 0.8-1.6M (10-21M) and 3D 0.6M
    // (15M), and 7-11 % slower on the 64-scene ensemble.  The default takes it for
    // single scenes whose fine matrix holds kMgAdditiveEntries or more;
    // MPS_AMD_MG_FINE=additive|multiplicative overrides.
    int fine_mode = 0;               // 0: by size, 1: additive, 2: multiplicative
    bool fine_additive = false;      // the current hierarchy's choice
    int tail_rows = kMgTailRows;
};

void mg_set_params(MgWork* m, double omega, double cf, int sweeps) {
    if (!m) return;
    if (omega > 0) m->omega = omega;
    if (cf > 0) m->cf = cf;
    if (sweeps > 0) m->sweeps = sweeps;
    m->user_params = true;
    m->gen++;
}

// MPS_AMD_MG="omega,factor,sweeps[,tail_rows]" overrides the smoother
// weight, the coarse-correction factor, the coarsest sweeps and the size up
// to which levels run in the tail kernel (tuning aid).
MgWork* mg_create() {
    MgWork* m = new MgWork();
    if (const char* v = getenv("MPS_AMD_MG")) {
        double om = 0, cf = 0;
        int sw = 0, tr = 0;
        if (sscanf(v, "%lf,%lf,%d,%d", &om, &cf, &sw, &tr) >= 1) mg_set_params(m, om, cf, sw);
        if (tr > 0) m->tail_rows = tr;
    }
    if (const char* v = getenv("MPS_AMD_MG_FINE"))
        m->fine_mode = std::strcmp(v, "additive") == 0 ? 1 : (std::strcmp(v, "multiplicative") == 0 ? 2 : 0);
    return m;
}

static void mg_free_level(MgWork* m, int l) {
    void* ps[] = {m->agg[l],  m->coord[l], m->child[l], m->col[l], m->val[l],
                  m->diag[l], m->dinv[l],  m->b[l],     m->x[l],   m->scene[l]};
    for (void* p : ps) hipFree(p);
    m->scene[l] = nullptr;
    m->agg[l] = m->coord[l] = m->child[l] = m->col[l] = nullptr;
    m->val[l] = m->diag[l] = m->dinv[l] = m->b[l] = m->x[l] = nullptr;
    m->cap[l] = 0;
}

void mg_destroy(MgWork* m) {
    if (!m) return;
    for (int l = 0; l <= kMgMaxLevels; l++) {
        mg_free_level(m, l);
        hipFree(m->d_sdims[l]);
        hipFree(m->d_goff[l]);
    }
    void* ps[] = {m->nd,  m->err,   m->cst1,  m->head,     m->hscan, m->rkey,     m->agg0b,
                  m->res0, m->xx0, m->flag, m->pscan, m->scan_tmp, m->desc, m->lsz};
    for (void* p : ps) hipFree(p);
    mailbox_free(m->mb);
    delete m;
}

int mg_generation(const MgWork* m) { return m ? m->gen : -1; }
int mg_levels(const MgWork* m) { return m ? m->L : 0; }

template <typename P>
static hipError_t mg_alloc(P** p, size_t bytes) {
    return hipMalloc((void**) p, bytes < 8 ? 8 : bytes);
}

// Level l holds rows for up to `need`; level 0 holds only agg.
static hipError_t mg_reserve_level(MgWork* m, int l, int need) {
    if (need <= m->cap[l] && m->agg[l]) return hipSuccess;
    mg_free_level(m, l);
    const int cap = need + need / 4 + 64;
    const size_t W = (size_t) m->W, C = (size_t) m->C, e = 8;
    hipError_t er;
    if ((er = mg_alloc(&m->agg[l], sizeof(int) * (size_t) cap))) return er;
    if (l >= 1) {
        if ((er = mg_alloc(&m->coord[l], sizeof(int) * (size_t) cap))) return er;
        if ((er = mg_alloc(&m->scene[l], sizeof(int) * (size_t) cap))) return er;
        if ((er = mg_alloc(&m->child[l], sizeof(int) * C * cap))) return er;
        if ((er = mg_alloc(&m->col[l], sizeof(int) * W * cap))) return er;
        if ((er = mg_alloc(&m->val[l], e * W * cap))) return er;
        if ((er = mg_alloc(&m->diag[l], e * cap))) return er;
        if ((er = mg_alloc(&m->dinv[l], e * cap))) return er;
        if ((er = mg_alloc(&m->b[l], e * cap))) return er;
        if ((er = mg_alloc(&m->x[l], e * cap))) return er;
    }
    m->cap[l] = cap;
    m->gen++;
    return hipSuccess;
}

static hipError_t mg_scan(MgWork* m, const int* in, int* out, int len, hipStream_t st) {
    const size_t need = scan_tmp_ints(len) + 16;
    if (need > m->scan_cap) {
        hipFree(m->scan_tmp);
        m->scan_tmp = nullptr;
        hipError_t e = mg_alloc(&m->scan_tmp, sizeof(int) * need);
        if (e) return e;
        m->scan_cap = need;
    }
    return scan_exclusive(in, out, len, m->scan_tmp, st);
}

static hipError_t mg_read_int(MgWork* m, const int* dev, int* host, hipStream_t st) {
    const MailSrc src{dev, (int) sizeof(int)};
    return mailbox_read(m->mb, &src, 1, host, st);
}

// Host-to-device copy of a small table unless the device already holds it.
template <typename V>
static hipError_t mg_upload(V* dev, const std::vector<V>& host, std::vector<V>& last, hipStream_t st) {
    if (last.size() == host.size() && std::equal(host.begin(), host.end(), last.begin(), [](const V& a, const V& b) {
            return std::memcmp(&a, &b, sizeof(V)) == 0;
        }))
        return hipSuccess;
    last = host;
    return hipMemcpyAsync(dev, host.data(), sizeof(V) * host.size(), hipMemcpyHostToDevice, st);
}

template <typename T, int W, int C>
static hipError_t mg_allow_tail_lds() {
    static hipError_t once = hipFuncSetAttribute((const void*) k_mg_tail<T, W, C>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, kMgTailLds);
    return once;
}

// The fine level's form for this system (see MgWork::fine_mode).  A system
// that took the additive cycle keeps it down to 7/8 of the threshold, so a
// size near it does not switch the cycle, and with it the captured CG
// graphs, from one step to the next.
static bool mg_fine_additive(const MgWork* m, const Sell& A) {
    if (m->fine_mode) return m->fine_mode == 1;
    if (A.nseg != 1) return false;
    const size_t lim = m->fine_additive ? kMgAdditiveEntries - kMgAdditiveEntries / 8 : kMgAdditiveEntries;
    return A.entries >= lim;
}

template <typename T, typename IDX>
static hipError_t mg_setup_t(MgWork* m, const Sell& A, const Hot& h, int dim, const T* fval, const IDX* fcol,
                             hipStream_t st) {
    hipError_t e;
    // without the large LDS allowance the cycle cannot run: solve without it
    if ((dim == 2 ? mg_allow_tail_lds<T, 9, 4>() : mg_allow_tail_lds<T, 27, 8>()) != hipSuccess)
        return hipErrorInvalidValue;
    const int R = A.rows;
    m->W = dim == 2 ? 9 : 27;
    m->C = dim == 2 ? 4 : 8;
    const bool add = mg_fine_additive(m, A);
    if (add != m->fine_additive) m->gen++;
    m->fine_additive = add;
    if (!m->user_params) {
        // the additive cycle prefers 1.2 in 3D (3.6M dam break: 1.2, 1.3, 1.5
        // gave 20.9, 21.5, 25.7 iterations)
        const double cf = dim == 2 ? 1.4 : (add ? 1.2 : 1.3);
        const int sw = dim == 2 ? 1 : 8;
        if (cf != m->cf || sw != m->sweeps || m->omega != 1.0) m->gen++;
        m->omega = 1.0;
        m->cf = cf;
        m->sweeps = sw;
    }
    if (!m->nd) {
        if ((e = mg_alloc(&m->nd, sizeof(int) * (kMgMaxLevels + 1)))) return e;
        if ((e = mg_alloc(&m->lsz, sizeof(int) * (kMgMaxLevels + 1)))) return e;
        if ((e = fill_int(m->lsz, 0, kMgMaxLevels + 1, st))) return e;
        if ((e = mg_alloc(&m->err, sizeof(int)))) return e;
        if ((e = mg_alloc(&m->desc, sizeof(MgLevelDesc) * (kMgMaxLevels + 1)))) return e;
    }
    if (R > m->fcap) {
        void* ps[] = {m->head, m->hscan, m->rkey, m->res0, m->xx0, m->cst1, m->agg0b};
        for (void* p : ps) hipFree(p);
        const int cap = R + R / 4 + 1024;
        if ((e = mg_alloc(&m->head, sizeof(int) * (size_t) cap)) || (e = mg_alloc(&m->hscan, sizeof(int) * (cap + 1))) ||
            (e = mg_alloc(&m->agg0b, (size_t) cap)) ||
            (e = mg_alloc(&m->rkey, sizeof(int) * (size_t) cap)) || (e = mg_alloc(&m->res0, 8 * (size_t) cap)) ||
            (e = mg_alloc(&m->xx0, 8 * (size_t) cap)) || (e = mg_alloc(&m->cst1, sizeof(int) * (cap + 1))))
            return e;
        m->fcap = cap;
        m->gen++;
    }
    if ((e = mg_reserve_level(m, 0, R))) return e;
    if ((e = fill_int(m->err, 0, 1, st))) return e;
    const int ns = A.nseg;
    if (ns > m->scap) {
        for (int k = 0; k <= kMgMaxLevels; k++) {
            hipFree(m->d_sdims[k]);
            hipFree(m->d_goff[k]);
            m->d_sdims[k] = nullptr;
            m->d_goff[k] = nullptr;
        }
        for (int k = 0; k <= kMgMaxLevels; k++) {
            m->up_sdims[k].clear();  // the new buffers hold nothing yet
            m->up_goff[k].clear();
            if ((e = mg_alloc(&m->d_sdims[k], sizeof(MgDims) * (size_t) ns)) ||
                (e = mg_alloc(&m->d_goff[k], sizeof(int) * (size_t) (ns + 1))))
                return e;
        }
        m->scap = ns;
        m->gen++;
    }
    if (m->nscene != ns) m->gen++;
    m->nscene = ns;
    // level 1: cells
    hipLaunchKernelGGL(k_mg_heads, mg_grid(R), dim3(kBlock), 0, st, R, A.row_slot, h.key, m->head, m->rkey);
    if ((e = mg_scan(m, m->head, m->hscan, R, st))) return e;
    int n1 = 0;
    if ((e = mg_read_int(m, m->hscan + R, &n1, st))) return e;
    if ((e = mg_reserve_level(m, 1, n1))) return e;
    hipLaunchKernelGGL(k_mg_level1, mg_grid(R), dim3(kBlock), 0, st, R, m->head, m->hscan, m->rkey, A.scene_key0, ns,
                       m->agg[0], m->agg0b, m->cst1, m->coord[1], m->scene[1]);
    m->n[0] = R;
    m->n[1] = n1;
    m->sdims[1].resize(ns);
    for (int k = 0; k < ns; k++) {
        const GridView& gk = A.scene_grids[k];
        m->sdims[1][k] = MgDims{gk.nx, gk.ny, gk.flat ? 1 : gk.nz};
        if (gk.nx < 3 || gk.ny < 3 || (!gk.flat && dim == 3 && gk.nz < 3)) return hipErrorInvalidValue;
    }
    if ((e = mg_upload(m->d_sdims[1], m->sdims[1], m->up_sdims[1], st))) return e;
    const dim3 ab(kAsm1Rows);
    if (dim == 2)
        hipLaunchKernelGGL((k_mg_assemble1<T, IDX, 9, 48>), dim3((n1 + 47) / 48), dim3(kAsm1Rows), 0, st, n1, m->cst1,
                           A.slice_off, fcol, A.anc, fval, m->agg[0], m->rkey, m->scene[1], m->d_sdims[1],
                           m->cap[1], m->col[1], (T*) m->val[1], (T*) m->diag[1], (T*) m->dinv[1], m->err);
    else
        hipLaunchKernelGGL((k_mg_assemble1<T, IDX, 27, 16>), dim3((n1 + 15) / 16), dim3(kAsm1Rows), 0, st, n1,
                           m->cst1, A.slice_off, fcol, A.anc, fval, m->agg[0], m->rkey, m->scene[1],
                           m->d_sdims[1], m->cap[1], m->col[1], (T*) m->val[1], (T*) m->diag[1], (T*) m->dinv[1],
                           m->err);
    // Coarse levels 2..: level l + 1 merges 2x2 (2x2x2) cells of level l.
    // Their structure is built for all levels at once: the grid dims of
    // every level that still shrinks, one flag per grid cell set by the
    // level-1 cells under it, one scan over all the grids and one read of
    // the level sizes; then rows, child tables and parent links of the kept
    // levels in one launch each, and the Galerkin products level by level.
    // Levels stop at <= 64 rows, or once every scene's grid is a single cell.
    const int coarse = std::max(kMgCoarse, ns);
    MgTabs tb{};
    tb.nscene = ns;
    tb.C = m->C;
    int lmax = 1;
    {
        int f = 0;
        for (int k = 1; k < kMgMaxLevels; k++) {
            std::vector<MgDims>& dp = m->sdims[k + 1];
            std::vector<int>& go = m->goff[k + 1];
            dp.resize(ns);
            go.assign(ns + 1, 0);
            bool shrinks = false;
            for (int q = 0; q < ns; q++) {
                const MgDims d = m->sdims[k][q];
                dp[q] = MgDims{(d.x + 1) / 2, (d.y + 1) / 2, (d.z + 1) / 2};
                shrinks = shrinks || d.x > 1 || d.y > 1 || d.z > 1;
                go[q + 1] = go[q] + dp[q].x * dp[q].y * dp[q].z;
            }
            if (!shrinks) break;
            if ((e = mg_upload(m->d_sdims[k + 1], dp, m->up_sdims[k + 1], st)) ||
                (e = mg_upload(m->d_goff[k + 1], go, m->up_goff[k + 1], st)))
                return e;
            tb.foff[k + 1] = f;
            f += go[ns];
            lmax = k + 1;
        }
        tb.foff[lmax + 1] = f;
    }
    for (int k = 1; k <= lmax; k++) {
        tb.sdims[k] = m->d_sdims[k];
        tb.goff[k] = m->d_goff[k];
    }
    tb.lo = 2;
    tb.hi = lmax;
    int l = 1;
    if (lmax >= 2 && n1 > coarse) {
        const int G = tb.foff[lmax + 1];
        if (G + 1 > m->gcap) {
            hipFree(m->flag);
            hipFree(m->pscan);
            m->flag = m->pscan = nullptr;
            const int cap = G + 1024;
            if ((e = mg_alloc(&m->flag, sizeof(int) * (size_t) cap)) ||
                (e = mg_alloc(&m->pscan, sizeof(int) * (size_t) (cap + 1))))
                return e;
            m->gcap = cap;
        }
        if ((e = fill_int(m->flag, 0, (size_t) G, st))) return e;
        hipLaunchKernelGGL(k_mg_mark_all, mg_grid(n1), dim3(kBlock), 0, st, n1, m->coord[1], m->scene[1],
                           m->d_sdims[1], tb, m->flag);
        if ((e = mg_scan(m, m->flag, m->pscan, G, st))) return e;
        hipLaunchKernelGGL(k_mg_sizes, dim3(1), dim3(64), 0, st, m->pscan, tb, m->lsz);
        int sz[kMgMaxLevels + 1] = {};
        {
            const MailSrc src{m->lsz, (int) sizeof(int) * (kMgMaxLevels + 1)};
            if ((e = mailbox_read(m->mb, &src, 1, sz, st))) return e;
        }
        for (int k = 2; k <= lmax; k++) m->n[k] = sz[k];
        while (l < lmax && m->n[l] > coarse) l++;
        const int L = l;
        // the kept levels' rows, then their tables
        for (int k = 2; k <= L; k++)
            if ((e = mg_reserve_level(m, k, m->n[k]))) return e;
        tb.hi = L;
        tb.pbase[2] = 0;
        tb.cbase[2] = 0;
        for (int k = 2; k <= L; k++) {
            tb.pbase[k + 1] = tb.pbase[k] + m->n[k];
            tb.cbase[k + 1] = tb.cbase[k] + m->C * m->cap[k];
        }
        tb.lo = 1;
        tb.rbase[1] = 0;
        for (int k = 1; k <= L; k++) {
            tb.rbase[k + 1] = tb.rbase[k] + m->n[k];
            tb.coord[k] = m->coord[k];
            tb.scene[k] = m->scene[k];
            tb.agg[k] = m->agg[k];
            tb.child[k] = m->child[k];
            tb.cap[k] = m->cap[k];
        }
        hipLaunchKernelGGL(k_mg_parents_all, mg_grid(tb.foff[L + 1]), dim3(kBlock), 0, st, m->flag, m->pscan, tb);
        hipLaunchKernelGGL(k_mg_clear_children, fill_grid_mg(tb.cbase[L + 1]), dim3(kBlock), 0, st, tb);
        hipLaunchKernelGGL(k_mg_link_all, mg_grid(tb.rbase[L]), dim3(kBlock), 0, st, m->pscan, tb);
        for (int k = 1; k < L; k++) {
            const int np = m->n[k + 1];
            const dim3 pg((np + kAsm1Rows / m->C - 1) / (kAsm1Rows / m->C));
            if (dim == 2)
                hipLaunchKernelGGL((k_mg_assemble<T, 9, 4>), pg, ab, 0, st, np, m->child[k + 1], m->cap[k + 1],
                                   m->col[k], (const T*) m->val[k], (const T*) m->diag[k], m->cap[k], m->agg[k],
                                   m->coord[k + 1], m->scene[k + 1], m->d_sdims[k + 1], m->col[k + 1],
                                   (T*) m->val[k + 1], (T*) m->diag[k + 1], (T*) m->dinv[k + 1], m->err);
            else
                hipLaunchKernelGGL((k_mg_assemble<T, 27, 8>), pg, ab, 0, st, np, m->child[k + 1], m->cap[k + 1],
                                   m->col[k], (const T*) m->val[k], (const T*) m->diag[k], m->cap[k], m->agg[k],
                                   m->coord[k + 1], m->scene[k + 1], m->d_sdims[k + 1], m->col[k + 1],
                                   (T*) m->val[k + 1], (T*) m->diag[k + 1], (T*) m->dinv[k + 1], m->err);
        }
    }
    if (m->L != l) m->gen++;
    m->L = l;
    if (m->n[l] > coarse) return hipErrorInvalidValue;  // the level cap was hit before the grids got small
    int tail = l;
    {
        const int rb = mg_tail_row_bytes(m->W, m->C, (int) sizeof(T)) + 16;
        long long used = (long long) m->n[l] * rb + 256;
        while (tail > 1 && m->n[tail - 1] <= m->tail_rows &&
               used + (long long) m->n[tail - 1] * rb + 256 <= kMgTailLds) {
            tail--;
            used += (long long) m->n[tail] * rb + 256;
        }
        if (used > kMgTailLds) return hipErrorInvalidValue;
    }
    if (tail != m->tail) m->gen++;
    m->tail = tail;
    for (int k = 1; k <= l; k++)
        m->hdesc[k] = MgLevelDesc{m->cap[k], 0,         m->col[k], m->val[k],   m->diag[k],
                                  m->dinv[k], m->agg[k], m->child[k], m->b[k], m->x[k]};
    if (!m->up_desc_valid || std::memcmp(m->up_desc, m->hdesc, sizeof(m->hdesc)) != 0) {
        std::memcpy(m->up_desc, m->hdesc, sizeof(m->hdesc));
        m->up_desc_valid = true;
        if ((e = hipMemcpyAsync(m->desc, m->up_desc, sizeof(MgLevelDesc) * (kMgMaxLevels + 1), hipMemcpyHostToDevice,
                                st)))
            return e;
    }
    if ((e = hipMemcpyAsync(m->nd, m->n, sizeof(int) * (kMgMaxLevels + 1), hipMemcpyHostToDevice, st))) return e;
    return hipGetLastError();  // m->err is read by the caller with its next read
}

const int* mg_error_flag(const MgWork* m) { return m->err; }

hipError_t mg_setup(MgWork* m, const Sell& A, const Hot& h, int dim, int fp32, hipStream_t st) {
    const unsigned short* acol = (const unsigned short*) A.dcol;  // 3D codes (see sell_col)
    if (fp32) {
        if (A.use16 && A.dim == 3) return mg_setup_t<float, unsigned short>(m, A, h, dim, A.val32, acol, st);
        return A.use16 ? mg_setup_t<float, short>(m, A, h, dim, A.val32, A.dcol, st)
                       : mg_setup_t<float, int>(m, A, h, dim, A.val32, A.col, st);
    }
    if (A.use16 && A.dim == 3) return mg_setup_t<double, unsigned short>(m, A, h, dim, A.val64, acol, st);
    return A.use16 ? mg_setup_t<double, short>(m, A, h, dim, A.val64, A.dcol, st)
                   : mg_setup_t<double, int>(m, A, h, dim, A.val64, A.col, st);
}

template <typename T, typename IDX, int W, int C>
static void mg_cycle_t(MgWork* m, const Sell& A, const IDX* fcol, const T* fval, const T* r, T* u, MgDone done,
                       bool b1_given, hipStream_t st) {
    const T om = (T) m->omega, cf = (T) m->cf;
    const int L = m->L, lt = m->tail;
    int* nd = m->nd;
    T* res0 = (T*) m->res0;
    T* xx0 = (T*) m->xx0;
    const bool add = m->fine_additive;
    if (!add)
        hipLaunchKernelGGL((k_mg_down0<T, IDX>), mg_grid(m->cap[0]), dim3(kBlock), 0, st, nd + 0, A.slice_off, fcol,
                           A.anc, fval, r, res0, om, done);
    if (!(add && b1_given))
        hipLaunchKernelGGL((k_mg_restrict1<T>), mg_grid(m->cap[1]), dim3(kBlock), 0, st, nd + 1, m->cst1,
                           add ? r : (const T*) res0, (T*) m->b[1], done);
    T* const xf = add ? u : xx0;  // the fine level's first step up
    for (int l = 1; l < lt; l++)
        hipLaunchKernelGGL((k_mg_down<T, W, C>), mg_grid(C * m->cap[l + 1]), dim3(kBlock), 0, st, nd + l + 1,
                           m->child[l + 1], m->cap[l + 1], m->col[l], (const T*) m->val[l], (const T*) m->diag[l],
                           (const T*) m->dinv[l], m->cap[l], (const T*) m->b[l], (T*) m->b[l + 1], om, done);
    hipLaunchKernelGGL((k_mg_tail<T, W, C>), dim3(1), dim3(kMgTailThreads), kMgTailLds, st, m->desc, nd, lt, L, om,
                       cf, m->sweeps, done);
    for (int l = lt - 1; l >= 2; l--)
        hipLaunchKernelGGL((k_mg_up<T, W, false>), mg_grid(m->cap[l]), dim3(kBlock), 0, st, nd + l, m->col[l],
                           (const T*) m->val[l], (const T*) m->diag[l], (const T*) m->dinv[l], m->cap[l], m->agg[l],
                           (const T*) m->b[l], (const T*) m->x[l + 1], (T*) m->x[l], om, cf, done, nullptr, nullptr,
                           nullptr, nullptr);
    if (lt >= 2)  // level 1's up pass also takes the fine level's first step up
        hipLaunchKernelGGL((k_mg_up<T, W, true>), mg_grid(m->cap[1]), dim3(kBlock), 0, st, nd + 1, m->col[1],
                           (const T*) m->val[1], (const T*) m->diag[1], (const T*) m->dinv[1], m->cap[1], m->agg[1],
                           (const T*) m->b[1], (const T*) m->x[2], (T*) m->x[1], om, cf, done, m->cst1, r, xf,
                           m->agg0b);
    else
        hipLaunchKernelGGL((k_mg_prolong0<T>), mg_grid(m->cap[0]), dim3(kBlock), 0, st, nd + 0, r, m->agg[0],
                           (const T*) m->x[1], xf, om, cf, done);
    if (!add)
        hipLaunchKernelGGL((k_mg_up0<T, IDX>), mg_grid(m->cap[0]), dim3(kBlock), 0, st, nd + 0, A.slice_off, fcol,
                           A.anc, fval, r, xx0, u, om, done);
}

template <typename T, typename IDX>
void mg_cycle(MgWork* m, const Sell& A, const IDX* fcol, const T* fval, const T* r, T* u, MgDone done, int dim,
              bool b1_given, hipStream_t st) {
    if (dim == 2)
        mg_cycle_t<T, IDX, 9, 4>(m, A, fcol, fval, r, u, done, b1_given, st);
    else
        mg_cycle_t<T, IDX, 27, 8>(m, A, fcol, fval, r, u, done, b1_given, st);
}

bool mg_fine_restriction(const MgWork* m, const int** cst1, int* n1, void** b1) {
    if (!m || !m->fine_additive || m->L < 1) return false;
    *cst1 = m->cst1;
    *n1 = m->n[1];
    *b1 = m->b[1];
    return true;
}

template void mg_cycle<float, short>(MgWork*, const Sell&, const short*, const float*, const float*, float*, MgDone,
                                     int, bool, hipStream_t);
template void mg_cycle<float, int>(MgWork*, const Sell&, const int*, const float*, const float*, float*, MgDone, int,
                                   bool, hipStream_t);
template void mg_cycle<double, short>(MgWork*, const Sell&, const short*, const double*, const double*, double*,
                                      MgDone, int, bool, hipStream_t);
template void mg_cycle<double, int>(MgWork*, const Sell&, const int*, const double*, const double*, double*, MgDone,
                                    int, bool, hipStream_t);
template void mg_cycle<float, unsigned short>(MgWork*, const Sell&, const unsigned short*, const float*, const float*,
                                              float*, MgDone, int, bool, hipStream_t);
template void mg_cycle<double, unsigned short>(MgWork*, const Sell&, const unsigned short*, const double*,
                                               const double*, double*, MgDone, int, bool, hipStream_t);

}  // namespace mps
