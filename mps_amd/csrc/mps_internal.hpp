// mps_internal.hpp — device state, grid view and shared device helpers of the
// MI355X MPS step.  Internal to libmps_amd.so.
//
// Layout in HBM.  After every bucket build the hot particle arrays are
// physically reordered by cell key ("slots"): positions, velocities, id and
// type.  Cells are contiguous slot ranges [cs[k], cs[k+1]) and, within a
// cell, slots ascend by particle id, which is the reference's in-cell order
// (src/bucket.cpp:36-50 appends to the tail in id order).  Ghost particles sit
// after the last cell.  Cold per-particle fields (pressure, minimum pressure,
// number density ratio, boundary condition, warm-start solution, the stale
// number density of ghosts) stay in id order and are reached through id[].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/mps_amd.h"

namespace mps {

constexpr int kBlock = 256;
constexpr int kMaxTuneBlock = 512;  // largest block the autotuner may pick for per-particle kernels

struct GridView {
    double xmin, ymin, zmin;
    double len;         // bucket side = re_max (src/bucket.cpp:11)
    int nx, ny, nz;     // numX, numY, numZ (src/bucket.cpp:13-16)
    int flat;           // 1: every in-domain particle has iz = 1 (numZ == 3), keys drop z
    int ncell;          // keys of this scene: flat ? nx*ny : nx*ny*nz
    int key0;           // first key of this scene in the handle's key space
    int ghost_key;      // key of the ghost segment (after every scene's cells)
    const int* cs;      // cell starts over the whole key space (+ ghost segment end)
    int stage_cap;      // most slots a block copies into LDS (the kernels' LDS size caps it further)
    int stage_stats;    // MPS_AMD_STAGE_STATS=1: count staged blocks by how much they staged (tests)
};

// Hot particle arrays in slot order (one of two ping-pong sets).
struct Hot {
    double* px;
    double* py;
    double* pz;
    double* vx;
    double* vy;
    double* vz;
    int32_t* id;
    int8_t* type;
    int32_t* key;   // cell key of the slot at the last build (ncell for ghosts)
};

struct Consts {
    int dim;
    double dt;
    double l0;
    double rho;
    double nu;
    double g[3];
    double re_p, re_v, re_s, re_max;
    double n0_p, lam_p, n0_v, lam_v, n0_s, lam_s;
    double beta;
    double relax;
    int qc_on;
    double alpha;
    int hos_on;
    double hos_gamma;
    double dc;
    double restitution;
    double dmin[3], dmax[3];
    double cfl;
    // coefficients, each computed once in the reference's expression order
    double avisc;      // nu * (2 D) / (n0_v lambda_v)          src/mps.cpp:182
    double src_scale;  // gamma * (1 / (dt dt))                  src/mps.cpp:313
    double hos_scale;  // gamma_hos * (1 / (dt dt))              src/mps.cpp:291
    double hos_coef;   // 1 / (n0 dt)                            src/mps.cpp:304
    double coefA;      // 2 D / (lambda n0)                      src/mps.cpp:343
    double qc_add;     // alpha / (dt dt)                        src/mps.cpp:354
    double gscale;     // D / n0                                 src/mps.cpp:530
};

// Per-scene lookup for handles holding several independent scenes packed in
// one particle array (ensembles).  Single-scene kernels ignore it.
struct SceneTab {
    const GridView* g;      // per scene
    const Consts* c;        // per scene
    const int16_t* sid;     // scene of each particle id
};

template <bool MULTI>
__device__ __forceinline__ void scene_of(const SceneTab& t, int id, GridView& g, Consts& c) {
    if constexpr (MULTI) {
        const int si = t.sid[id];
        g = t.g[si];
        c = t.c[si];
    }
}

// Device error flags (bit mask written with atomicOr).
constexpr int kErrRange = 1;     // particle outside the bucket grid
constexpr int kErrCollCap = 2;   // collision pair buffer overflow
constexpr int kErrSellCap = 4;   // matrix storage overflow

// Event buffers and counters on the device.
struct Events {
    int* n_ghost;        // appended count (may exceed cap)
    int* ghost_ids;
    int* n_unconn;
    int* unconn_ids;
    int cap;
    int* err;
};

// ---------------- small device helpers ----------------

// blockIdx -> logical block so that each of the 8 XCDs (which take the blocks
// of a dispatch round robin) works on one contiguous range of the index
// space and keeps that range's neighbourhood in its own L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b & 7, idx = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Sum of v[r0 .. r1) in row order, four loads in flight (the level-1
// restriction of the multigrid cycle; the CG step forms the same sums where
// it writes r, so both give the same bits).
template <typename T>
__device__ __forceinline__ T cell_sum(const T* __restrict__ v, int r0, int r1) {
    T s = (T) 0;
    for (int r = r0; r < r1; r += 4) {
        T q[4];
#pragma unroll
        for (int j = 0; j < 4; j++) q[j] = v[r + j < r1 ? r + j : r1 - 1];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (r + j < r1) s += q[j];
    }
    return s;
}

// Cell coordinate as src/bucket.cpp:40-42 computes it: (int)((x-min)/len)+1.
// Returns false if the value cannot index the grid.
__device__ __forceinline__ bool cell_coord(double x, double xmin, double len, int n, int& ic) {
    double q = (x - xmin) / len;
    if (!(q > -1.0 && q < (double) (n - 2))) return false;  // NaN fails too
    ic = (int) q + 1;
    return ic >= 1 && ic <= n - 2;
}

__device__ __forceinline__ double wfun(double dist, double re) {
    // src/mps.cpp:9-16
    double w = 0.0;
    if (dist < re) w = (re / dist) - 1.0;
    return w;
}

// Candidates q in [b, e) of one cell, in ascending order.  Positions are
// loaded kNbBatch at a time (indices clamped to e - 1, so every load stays in
// the cell), so one loop trip has kNbBatch independent loads in flight
// instead of one; the visit order and the arithmetic of what f receives are
// those of a plain q = b .. e-1 loop.  d2 > re2 skips the square root (see
// for_neighbors); re2 = re^2 (1 + 1e-12) leaves room for the fused
// prefilter, whose d2 differs from the reference's rounding by a few ulp.
constexpr int kNbBatch = 4;

template <int DIM, typename F>
__device__ __forceinline__ void scan_cell(const double* __restrict__ px, const double* __restrict__ py,
                                          const double* __restrict__ pz, int b, int e, int s, double xi,
                                          double yi, double zi, double re, double re2, F& f) {
    // 32-bit byte offsets (handles hold fewer than 2^29 particles, so 8 q
    // fits): one offset register
    // serves the three arrays, with their bases in scalar registers
    const char* __restrict__ bx = reinterpret_cast<const char*>(px);
    const char* __restrict__ by = reinterpret_cast<const char*>(py);
    const char* __restrict__ bz = reinterpret_cast<const char*>(pz);
    const unsigned last = (unsigned) (e - 1);
    for (int q0 = b; q0 < e; q0 += kNbBatch) {
        double x[kNbBatch], y[kNbBatch], z[kNbBatch];
#pragma unroll
        for (int i = 0; i < kNbBatch; i++) {
            const unsigned q = (unsigned) (q0 + i) < last ? (unsigned) (q0 + i) : last;
            const unsigned off = q * 8u;
            x[i] = *reinterpret_cast<const double*>(bx + off);
            y[i] = *reinterpret_cast<const double*>(by + off);
            if constexpr (DIM == 3) z[i] = *reinterpret_cast<const double*>(bz + off);
        }
#pragma unroll
        for (int i = 0; i < kNbBatch; i++) {
            const int q = q0 + i;
            if (q >= e) break;
            const double dx = x[i] - xi;
            const double dy = y[i] - yi;
            const double dz = DIM == 3 ? z[i] - zi : 0.0;
            // cheap test first: d2 with fused multiply-adds is within a few
            // ulp of the reference's expression, far inside re2's margin
            const double d2f = DIM == 3 ? fma(dz, dz, fma(dy, dy, dx * dx)) : fma(dy, dy, dx * dx);
            if (d2f > re2) continue;
            const double d2 = DIM == 3 ? (dx * dx + dy * dy) + dz * dz : dx * dx + dy * dy;
            if (d2 > re2) continue;
            const double dist = sqrt(d2);
            if (q != s && dist < re) f(q, dist, dx, dy, dz);
        }
    }
}

// Distances from a particle at p to the two neighbouring cells' boxes along
// one axis (cell c spans [pmin + (c - 1) len, pmin + c len) on the axis; the
// particle's cell is ic), less a margin far above the rounding of the cell
// index and of the distance.  Every slot in a cell lies in the cell's box
// when the walk reads the positions its bucket build keyed (which every walk
// here does), so a cell whose box is farther than re from the particle holds
// no neighbour and is skipped whole; the visit order of the other cells and
// what f receives are unchanged.
__device__ __forceinline__ void cell_gaps(double p, double pmin, double len, int ic, double& lo, double& hi) {
    const double m = 1e-9 * len;
    const double l = p - (pmin + (double) (ic - 1) * len) - m;
    const double h = (pmin + (double) ic * len) - p - m;
    lo = l > 0.0 ? l : 0.0;
    hi = h > 0.0 ? h : 0.0;
}
__device__ __forceinline__ double cell_gap(int d, double lo, double hi) {  // d = offset + 1 in 0..2
    return d == 0 ? lo : (d == 2 ? hi : 0.0);
}

// Visits the neighbours of slot s in the reference's order (src/mps.cpp:599-613):
// jx outer, jy, jz inner, ascending id within a cell.  Calls f(q, dist, dx, dy, dz)
// for every slot q != s with dist < re, dist = sqrt((dx*dx+dy*dy)+dz*dz) and
// dx = x_q - x_s (Eigen's (pj.position - pi.position).norm()).
// DIM == 2 runs on the flat key space with z identically zero.
template <int DIM, typename F>
__device__ __forceinline__ void for_neighbors(const GridView& g, const double* __restrict__ px,
                                              const double* __restrict__ py,
                                              const double* __restrict__ pz, int key, int s,
                                              double xi, double yi, double zi, double re, F&& f) {
    key -= g.key0;  // key inside this scene's grid
    const int* __restrict__ cs = g.cs + g.key0;
    if constexpr (DIM == 2) {
        const int ix = key % g.nx;
        const int iy = key / g.nx;
        // the 9 cells' slot ranges first (independent loads, one round trip),
        // then the walk in the reference's order.  Candidates with
        // d2 > re^2 (1 + 1e-12) are skipped before the square root: sqrt is
        // correctly rounded and monotone, so for them sqrt(d2) >= re and the
        // reference's test dist < re fails too.
        const double re2 = re * re * (1.0 + 1e-12);
        double gxl, gxh, gyl, gyh;
        cell_gaps(xi, g.xmin, g.len, ix, gxl, gxh);
        cell_gaps(yi, g.ymin, g.len, iy, gyl, gyh);
        const double rr = re * re;
        int cb[9], ce[9];
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int jx = ix - 1 + a, jy = iy - 1 + b;
                const double gx = cell_gap(a, gxl, gxh), gy = cell_gap(b, gyl, gyh);
                const bool ok = jx >= 0 && jx < g.nx && jy >= 0 && jy < g.ny && gx * gx + gy * gy <= rr;
                const int k = jx + jy * g.nx;
                cb[3 * a + b] = ok ? cs[k] : 0;
                ce[3 * a + b] = ok ? cs[k + 1] : 0;
            }
        }
#pragma unroll
        for (int c = 0; c < 9; c++) scan_cell<2>(px, py, pz, cb[c], ce[c], s, xi, yi, 0.0, re, re2, f);
    } else {
        const double re2 = re * re * (1.0 + 1e-12);  // see the 2D branch
        int ix, iy, iz;
        if (g.flat) {
            ix = key % g.nx;
            iy = key / g.nx;
            iz = 1;
        } else {
            ix = key % g.nx;
            iy = (key / g.nx) % g.ny;
            iz = key / (g.nx * g.ny);
        }
        double gxl, gxh, gyl, gyh, gzl, gzh;
        cell_gaps(xi, g.xmin, g.len, ix, gxl, gxh);
        cell_gaps(yi, g.ymin, g.len, iy, gyl, gyh);
        cell_gaps(zi, g.zmin, g.len, iz, gzl, gzh);
        const double rr = re * re;
        for (int jx = ix - 1; jx <= ix + 1; jx++) {
            if (jx < 0 || jx >= g.nx) continue;
            const double gx = cell_gap(jx - ix + 1, gxl, gxh);
            for (int jy = iy - 1; jy <= iy + 1; jy++) {
                if (jy < 0 || jy >= g.ny) continue;
                const double gy = cell_gap(jy - iy + 1, gyl, gyh);
                for (int jz = iz - 1; jz <= iz + 1; jz++) {
                    if (g.flat) {
                        if (jz != 1) continue;  // layers 0 and 2 are empty
                    } else if (jz < 0 || jz >= g.nz) {
                        continue;
                    }
                    const double gz = g.flat ? 0.0 : cell_gap(jz - iz + 1, gzl, gzh);
                    if (gx * gx + gy * gy + gz * gz > rr) continue;
                    const int k = g.flat ? (jx + jy * g.nx) : (jx + jy * g.nx + jz * g.nx * g.ny);
                    scan_cell<3>(px, py, pz, cs[k], cs[k + 1], s, xi, yi, zi, re, re2, f);
                }
            }
        }
    }
}

// ---------------- walks from LDS ----------------
//
// A walk's candidate loads are gathers: the 64 lanes of a wavefront sit in
// several cells, so one load instruction touches many cache lines, and the
// L1 tag lookups (one line per cycle) bound the walk.  A block can instead
// copy, coalesced, the slot ranges its walks can reach into LDS and read the
// candidates there.  The block's slots are consecutive, so their cell keys
// form one range [k_lo, k_hi]; every neighbour cell of a key k in it lies in
// [k_lo - 1 + o, k_hi + 1 + o] for one stencil-row offset o = dy nx (+ dz nx
// ny), and cells of consecutive keys are consecutive slot ranges: one slot
// range per stencil row (3 in 2D, 9 in 3D).  Blocks whose ranges exceed the
// LDS capacity, that hold ghosts or another scene's cells, or that are on a
// flat 3D grid read global memory as before.  Either way the walk visits the
// same candidates in the same order with the same arithmetic.
template <int DIM>
struct NbStage {
    static constexpr int NR = DIM == 2 ? 3 : 9;  // stencil rows (dy, then dz)
    int ok;           // this block's walks read LDS
    const int* meta;  // LDS: [0] staged, [1 + r] LDS index minus slot for row r
    const double* x;  // LDS copies of the staged positions
    const double* y;
    const double* z;
};

template <int DIM>
constexpr int nb_stage_meta_ints() {
    return 1 + 3 * NbStage<DIM>::NR + 2;
}

// The grid a block's staging uses: the handle's, or for several scenes the
// scene of the block's first and last slot (no staging if they differ).
__device__ __forceinline__ bool nb_block_grid(const GridView& g0, const SceneTab* tab, const Hot& h, int s0, int s1,
                                              GridView& g) {
    g = g0;
    if (!tab) return true;
    const int a = tab->sid[h.id[s0]], b = tab->sid[h.id[s1]];
    if (a != b) return false;
    g = tab->g[a];
    return true;
}

// Every thread of the block must call it (it holds two block barriers when
// cap > 0).  meta: nb_stage_meta_ints<DIM>() ints of LDS; sx, sy, sz: cap
// doubles each (sz unused in 2D).  cap == 0: no staging, no barrier.
template <int DIM>
__device__ __forceinline__ NbStage<DIM> nb_stage(const GridView& g0, const SceneTab* tab, const Hot& h, int n,
                                                 int cap, int* meta, double* sx, double* sy, double* sz) {
    constexpr int NR = NbStage<DIM>::NR;
    if (cap <= 0) return NbStage<DIM>{0, meta, sx, sy, sz};
    int* first = meta + 1 + NR;
    int* base = meta + 1 + 2 * NR;  // NR + 1 entries
    if (threadIdx.x == 0) {
        int ok = 0;
        const int s0 = blockIdx.x * blockDim.x;
        const int s1 = (s0 + (int) blockDim.x < n ? s0 + (int) blockDim.x : n) - 1;
        GridView g;
        if (s0 < n && nb_block_grid(g0, tab, h, s0, s1, g) && !(DIM == 3 && g.flat)) {
            const int k_lo = h.key[s0] - g.key0, k_hi = h.key[s1] - g.key0;
            if (k_lo >= 0 && k_hi < g.ncell) {
                const int* cs = g.cs + g.key0;
                const int nx = g.nx, nxy = g.nx * g.ny;
                int tot = 0;
                for (int r = 0; r < NR; r++) {
                    const int dy = r % 3 - 1, dz = DIM == 3 ? r / 3 - 1 : 0;
                    int lo = k_lo - 1 + dy * nx + dz * nxy, hi = k_hi + 1 + dy * nx + dz * nxy;
                    lo = lo < 0 ? 0 : lo;
                    hi = hi > g.ncell - 1 ? g.ncell - 1 : hi;
                    const int a = lo <= hi ? cs[lo] : 0, b = lo <= hi ? cs[hi + 1] : 0;
                    first[r] = a;
                    base[r] = tot;
                    meta[1 + r] = tot - a;
                    tot += b - a;
                }
                base[NR] = tot;
                ok = tot <= cap ? 1 : 0;
            }
        }
        meta[0] = ok;
    }
    __syncthreads();
    if (meta[0]) {
        const int tot = base[NR];
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            int r = 0;
            while (r + 1 < NR && base[r + 1] <= i) r++;
            const int q = first[r] + (i - base[r]);
            sx[i] = h.px[q];
            sy[i] = h.py[q];
            if constexpr (DIM == 3) sz[i] = h.pz[q];
        }
    }
    __syncthreads();
    return NbStage<DIM>{meta[0], meta, sx, sy, sz};
}

// Pass mask of cnt <= 64 candidates: bit i is set when pass(i).  The
// candidate index runs the same in every lane of a wavefront (only the
// count differs), so it stays scalar and the mask is built in two 32-bit
// halves: a select and an OR per candidate instead of 64-bit shifts.
template <typename P>
__device__ __forceinline__ uint64_t pass_mask(int cnt, P&& pass) {
    uint32_t lo = 0, hi = 0;
    const int n0 = cnt < 32 ? cnt : 32;
    for (int i = 0; i < n0; i++) lo |= (pass(i) ? 1u : 0u) << i;
    for (int i = 32; i < cnt; i++) hi |= (pass(i) ? 1u : 0u) << (i - 32);
    return ((uint64_t) hi << 32) | lo;
}

// scan_cell on the LDS copy (candidate slot q sits at index q + off), in two
// passes: the candidates' prefilter results go into a 64-bit mask first (64
// candidates at a time), then the exact test, the square root and f run once
// per set bit, in slot order.  A wavefront then
// runs them as often as its busiest lane has passes, not at every candidate
// that any lane passes.
template <int DIM, typename F>
__device__ __forceinline__ void scan_cell_lds_mask(const double* __restrict__ sx, const double* __restrict__ sy,
                                                   const double* __restrict__ sz, int off, int b, int e, int s,
                                                   double xi, double yi, double zi, double re, double re2, F& f) {
    for (int c0 = b; c0 < e; c0 += 64) {
        const int c1 = e - c0 < 64 ? e : c0 + 64;
        uint64_t m = pass_mask(c1 - c0, [&](int i) {
            const int q = c0 + i;
            const double dx = sx[q + off] - xi;
            const double dy = sy[q + off] - yi;
            const double dz = DIM == 3 ? sz[q + off] - zi : 0.0;
            const double d2f = DIM == 3 ? fma(dz, dz, fma(dy, dy, dx * dx)) : fma(dy, dy, dx * dx);
            return d2f <= re2;
        });
        while (m) {
            const int q = c0 + __builtin_ctzll(m);
            m &= m - 1;
            const double dx = sx[q + off] - xi;
            const double dy = sy[q + off] - yi;
            const double dz = DIM == 3 ? sz[q + off] - zi : 0.0;
            const double d2 = DIM == 3 ? (dx * dx + dy * dy) + dz * dz : dx * dx + dy * dy;
            if (d2 > re2) continue;
            const double dist = sqrt(d2);
            if (q != s && dist < re) f(q, dist, dx, dy, dz);
        }
    }
}

// for_neighbors reading the block's LDS copy when it has one.
template <int DIM, int GROUP = kNbBatch, typename F>  // GROUP: see the 3D overload
__device__ __forceinline__ void for_neighbors(const NbStage<DIM>& st, const GridView& g,
                                              const double* __restrict__ px, const double* __restrict__ py,
                                              const double* __restrict__ pz, int key, int s, double xi, double yi,
                                              double zi, double re, F&& f) {
    if (!st.ok) {
        for_neighbors<DIM>(g, px, py, pz, key, s, xi, yi, zi, re, f);
        return;
    }
    key -= g.key0;
    const int* __restrict__ cs = g.cs + g.key0;
    const double re2 = re * re * (1.0 + 1e-12);
    if constexpr (DIM == 2) {
        const int ix = key % g.nx;
        const int iy = key / g.nx;
        double gxl, gxh, gyl, gyh;  // cells out of reach are skipped (see cell_gaps)
        cell_gaps(xi, g.xmin, g.len, ix, gxl, gxh);
        cell_gaps(yi, g.ymin, g.len, iy, gyl, gyh);
        const double rr = re * re;
        int cb[9], ce[9];
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int jx = ix - 1 + a, jy = iy - 1 + b;
                const double gx = cell_gap(a, gxl, gxh), gy = cell_gap(b, gyl, gyh);
                const bool ok = jx >= 0 && jx < g.nx && jy >= 0 && jy < g.ny && gx * gx + gy * gy <= rr;
                const int k = jx + jy * g.nx;
                cb[3 * a + b] = ok ? cs[k] : 0;
                ce[3 * a + b] = ok ? cs[k + 1] : 0;
            }
        }
        const int o0 = st.meta[1], o1 = st.meta[2], o2 = st.meta[3];
        // The passes of all nine cells in one loop.  Each stencil column's
        // three cells (jy ascending) share a 64-bit word, kF bits a cell, so
        // a lane's passes come out of the three words in the reference's
        // order (columns, cells, slots), and a wavefront runs the exact test
        // and f as often as its busiest lane has passes in all, instead of
        // the sum over cells of the busiest lane's passes in each (about
        // half as often: 28 against 58 trips for the viscosity walk of a 2D
        // lattice).  A lane with a cell of more than kF candidates takes the
        // cells one at a time.
        constexpr int kF = 21;
        bool fits = true;
#pragma unroll
        for (int c = 0; c < 9; c++) fits = fits && ce[c] - cb[c] <= kF;
        if (!fits) {
#pragma unroll
            for (int c = 0; c < 9; c++) {
                const int b = c % 3;
                scan_cell_lds_mask<2>(st.x, st.y, st.z, b == 0 ? o0 : (b == 1 ? o1 : o2), cb[c], ce[c], s, xi, yi,
                                      0.0, re, re2, f);
            }
            return;
        }
        uint64_t wd[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            uint64_t w = 0;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int c0 = cb[3 * a + b];
                const int off = b == 0 ? o0 : (b == 1 ? o1 : o2);
                const uint64_t m = pass_mask(ce[3 * a + b] - c0, [&](int i) {
                    const double dx = st.x[c0 + i + off] - xi;
                    const double dy = st.y[c0 + i + off] - yi;
                    return fma(dy, dy, dx * dx) <= re2;
                });
                w |= m << (kF * b);
            }
            wd[a] = w;
        }
        for (;;) {
            const int a = wd[0] ? 0 : (wd[1] ? 1 : (wd[2] ? 2 : 3));
            if (a == 3) break;
            uint64_t w = a == 0 ? wd[0] : (a == 1 ? wd[1] : wd[2]);
            const int p = __builtin_ctzll(w);
            w &= w - 1;
            if (a == 0)
                wd[0] = w;
            else if (a == 1)
                wd[1] = w;
            else
                wd[2] = w;
            const int b = p < kF ? 0 : (p < 2 * kF ? 1 : 2);
            const int c = 3 * a + b;
            int base = cb[0];  // cb[c] by selects: constant indices keep cb in registers
#pragma unroll
            for (int k = 1; k < 9; k++) base = c == k ? cb[k] : base;
            const int q = base + (p - kF * b);
            const int off = b == 0 ? o0 : (b == 1 ? o1 : o2);
            const double dx = st.x[q + off] - xi;
            const double dy = st.y[q + off] - yi;
            const double d2 = dx * dx + dy * dy;
            if (d2 > re2) continue;
            const double dist = sqrt(d2);
            if (q != s && dist < re) f(q, dist, dx, dy, 0.0);
        }
    } else {
        const int ix = key % g.nx;
        const int iy = (key / g.nx) % g.ny;
        const int iz = key / (g.nx * g.ny);
        double gxl, gxh, gyl, gyh, gzl, gzh;  // cells out of reach are skipped (see cell_gaps)
        cell_gaps(xi, g.xmin, g.len, ix, gxl, gxh);
        cell_gaps(yi, g.ymin, g.len, iy, gyl, gyh);
        cell_gaps(zi, g.zmin, g.len, iz, gzl, gzh);
        const double rr = re * re;
        for (int jx = ix - 1; jx <= ix + 1; jx++) {
            if (jx < 0 || jx >= g.nx) continue;
            const double gx = cell_gap(jx - ix + 1, gxl, gxh);
            for (int jy = iy - 1; jy <= iy + 1; jy++) {
                if (jy < 0 || jy >= g.ny) continue;
                const double gy = cell_gap(jy - iy + 1, gyl, gyh);
                for (int jz = iz - 1; jz <= iz + 1; jz++) {
                    if (jz < 0 || jz >= g.nz) continue;
                    const double gz = cell_gap(jz - iz + 1, gzl, gzh);
                    if (gx * gx + gy * gy + gz * gz > rr) continue;
                    const int k = jx + jy * g.nx + jz * g.nx * g.ny;
                    const int r = (jy - iy + 1) + 3 * (jz - iz + 1);
                    scan_cell_lds_mask<3>(st.x, st.y, st.z, st.meta[1 + r], cs[k], cs[k + 1], s, xi, yi, zi, re, re2,
                                          f);
                }
            }
        }
    }
}

// 3D variant: the LDS copy holds FP32 offsets from the block's first slot
// (12 bytes a slot instead of 24, so 9 ranges fit with room for occupancy).
// They only skip candidates that are certainly beyond re: slack covers the
// rounding of the offsets (half an FP32 ulp of the largest offset, bounded by
// the grid's diagonal) and of the FP32 distance; every candidate not skipped
// is tested in FP64 from global memory exactly as before.  The walk tests a
// cell's candidates (64 at a time) in FP32 first and keeps the passes as a
// bit mask, then loads the passes' FP64 positions kNbBatch at a time: the
// FP64 test and f run in rounds of kNbBatch passes of the wavefront's
// busiest lane instead of at every candidate that any lane passes, and
// candidates still reach them in slot order.
struct NbStageF {
    int ok;           // bit r: stencil row r is in LDS
    const int* meta;  // as NbStage<3>
    const float* x;
    const float* y;
    const float* z;
    double ox, oy, oz;  // the block's origin (its first slot)
    float slack;        // absolute, added to re
};

__device__ __forceinline__ NbStageF nb_stage_f(const GridView& g0, const SceneTab* tab, const Hot& h, int n, int cap,
                                              int* meta, float* sx, float* sy, float* sz) {
    constexpr int NR = 9;
    int* first = meta + 1 + NR;
    int* base = meta + 1 + 2 * NR;
    const int s0 = blockIdx.x * blockDim.x;
    if (threadIdx.x == 0) {
        int ok = 0;
        float slack = 0.0f;
        const int s1 = (s0 + (int) blockDim.x < n ? s0 + (int) blockDim.x : n) - 1;
        GridView g;
        if (s0 < n && nb_block_grid(g0, tab, h, s0, s1, g) && !g.flat) {
            const int k_lo = h.key[s0] - g.key0, k_hi = h.key[s1] - g.key0;
            if (k_lo >= 0 && k_hi < g.ncell) {
                const int* cs = g.cs + g.key0;
                const int nx = g.nx, nxy = g.nx * g.ny;
                // rows are staged in order while they fit; a row that does
                // not is left in global memory (the walk reads its cells
                // there) and later rows may still fit
                int tot = 0;
                for (int r = 0; r < NR; r++) {
                    const int dy = r % 3 - 1, dz = r / 3 - 1;
                    int lo = k_lo - 1 + dy * nx + dz * nxy, hi = k_hi + 1 + dy * nx + dz * nxy;
                    lo = lo < 0 ? 0 : lo;
                    hi = hi > g.ncell - 1 ? g.ncell - 1 : hi;
                    const int a = lo <= hi ? cs[lo] : 0, b = lo <= hi ? cs[hi + 1] : 0;
                    first[r] = a;
                    base[r] = tot;
                    meta[1 + r] = tot - a;
                    if (tot + (b - a) <= cap) {
                        tot += b - a;
                        ok |= 1 << r;
                    }
                }
                base[NR] = tot;
                // every offset is at most the grid's diagonal; an FP32 offset
                // is off by half an ulp of that, a difference of two by one
                // ulp per axis
                const double diag =
                    g.len * sqrt((double) g.nx * g.nx + (double) g.ny * g.ny + (double) g.nz * g.nz);
                slack = (float) (4.0 * diag * 1.2e-7);
            }
        }
        meta[0] = ok;
        meta[1 + 3 * NR + 1] = __float_as_int(slack);
    }
    __syncthreads();
    const int ok = meta[0];
    const int o = s0 < n ? s0 : 0;
    const double ox = h.px[o], oy = h.py[o], oz = h.pz[o];
    if (ok) {
        const int tot = base[NR];
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            int r = 0;
            while (r + 1 < NR && base[r + 1] <= i) r++;
            const int q = first[r] + (i - base[r]);
            sx[i] = (float) (h.px[q] - ox);
            sy[i] = (float) (h.py[q] - oy);
            sz[i] = (float) (h.pz[q] - oz);
        }
    }
    __syncthreads();
    return NbStageF{ok, meta, sx, sy, sz, ox, oy, oz, __int_as_float(meta[1 + 3 * NR + 1])};
}

template <int DIM, int GROUP = kNbBatch, typename F>
__device__ __forceinline__ void for_neighbors(const NbStageF& st, const GridView& g, const double* __restrict__ px,
                                              const double* __restrict__ py, const double* __restrict__ pz, int key,
                                              int s, double xi, double yi, double zi, double re, F&& f) {
    static_assert(DIM == 3, "the FP32 staging is the 3D walk's");
    if (!st.ok) {
        for_neighbors<3>(g, px, py, pz, key, s, xi, yi, zi, re, f);
        return;
    }
    key -= g.key0;
    const int* __restrict__ cs = g.cs + g.key0;
    const double re2 = re * re * (1.0 + 1e-12);
    const float xf = (float) (xi - st.ox), yf = (float) (yi - st.oy), zf = (float) (zi - st.oz);
    // the FP32 distance itself is within a few ulp (relative 1e-6 covers it)
    const float rf = ((float) re + st.slack) * 1.000001f;
    const float rf2 = rf * rf;
    const char* __restrict__ bx = reinterpret_cast<const char*>(px);
    const char* __restrict__ by = reinterpret_cast<const char*>(py);
    const char* __restrict__ bz = reinterpret_cast<const char*>(pz);
    const int ix = key % g.nx;
    const int iy = (key / g.nx) % g.ny;
    const int iz = key / (g.nx * g.ny);
    double gxl, gxh, gyl, gyh, gzl, gzh;  // cells out of reach are skipped (see cell_gaps)
    cell_gaps(xi, g.xmin, g.len, ix, gxl, gxh);
    cell_gaps(yi, g.ymin, g.len, iy, gyl, gyh);
    cell_gaps(zi, g.zmin, g.len, iz, gzl, gzh);
    const double rr = re * re;
    // One pass loop per (jx, jy) over its three jz cells: the cells' FP32
    // pass masks (64 candidates a cell at most) are taken first, then the
    // passes come out of them in the reference's order (jz, then slot) in
    // batches of kNbBatch, so a wavefront runs the FP64 tests as often as
    // its busiest lane has passes in the three cells together, not the sum
    // of the busiest lane's in each.  A (jx, jy) with a cell of more than 64
    // candidates, or one left in global memory, goes cell by cell.  GROUP
    // is the batch size (a kernel near the 128 registers of four waves per
    // SIMD takes a smaller batch to make room for the three words).
    auto batch_loop = [&](uint64_t& w0, uint64_t& w1, uint64_t& w2, int b0, int b1, int b2) {
        while (w0 | w1 | w2) {
            int qq[GROUP];
            int np = 0;
#pragma unroll
            for (int i = 0; i < GROUP; i++) {  // past the last pass: the first again
                const int d = w0 ? 0 : (w1 ? 1 : (w2 ? 2 : 3));
                if (d == 3) {
                    qq[i] = qq[0];
                    continue;
                }
                uint64_t w = d == 0 ? w0 : (d == 1 ? w1 : w2);
                const int q = (d == 0 ? b0 : (d == 1 ? b1 : b2)) + __builtin_ctzll(w);
                w &= w - 1;
                if (d == 0)
                    w0 = w;
                else if (d == 1)
                    w1 = w;
                else
                    w2 = w;
                qq[i] = q;
                np = i + 1;
            }
            double x[GROUP], y[GROUP], z[GROUP];
#pragma unroll
            for (int i = 0; i < GROUP; i++) {
                const unsigned boff = (unsigned) qq[i] * 8u;
                x[i] = *reinterpret_cast<const double*>(bx + boff);
                y[i] = *reinterpret_cast<const double*>(by + boff);
                z[i] = *reinterpret_cast<const double*>(bz + boff);
            }
#pragma unroll
            for (int i = 0; i < GROUP; i++) {
                if (i >= np) break;
                const int q = qq[i];
                const double dx = x[i] - xi;
                const double dy = y[i] - yi;
                const double dz = z[i] - zi;
                const double d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 > re2) continue;
                const double dist = sqrt(d2);
                if (q != s && dist < re) f(q, dist, dx, dy, dz);
            }
        }
    };
    // FP32 pass mask of candidates [c0, c0 + cnt), cnt <= 64, staged at c0 + off
    auto fmask = [&](int c0, int cnt, int off) {
        // the FP32 squared distance by fused multiply-adds: within the same
        // few ulp the slack of rf covers
        return pass_mask(cnt, [&](int i) {
            const int q = c0 + i + off;
            const float dxf = st.x[q] - xf, dyf = st.y[q] - yf, dzf = st.z[q] - zf;
            return fmaf(dzf, dzf, fmaf(dyf, dyf, dxf * dxf)) <= rf2;
        });
    };
    for (int jx = ix - 1; jx <= ix + 1; jx++) {
        if (jx < 0 || jx >= g.nx) continue;
        const double gx = cell_gap(jx - ix + 1, gxl, gxh);
        for (int jy = iy - 1; jy <= iy + 1; jy++) {
            if (jy < 0 || jy >= g.ny) continue;
            const double gy = cell_gap(jy - iy + 1, gyl, gyh);
            int cb[3], ce[3], co[3];
            bool grouped = true;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int jz = iz - 1 + d;
                const double gz = cell_gap(d, gzl, gzh);
                const bool in = jz >= 0 && jz < g.nz && gx * gx + gy * gy + gz * gz <= rr;
                const int k = jx + jy * g.nx + jz * g.nx * g.ny;
                const int r = (jy - iy + 1) + 3 * d;
                cb[d] = in ? cs[k] : 0;
                ce[d] = in ? cs[k + 1] : 0;
                co[d] = st.meta[1 + r];
                grouped = grouped && (!in || (((st.ok >> r) & 1) && ce[d] - cb[d] <= 64));
            }
            if (grouped) {
                uint64_t w0 = fmask(cb[0], ce[0] - cb[0], co[0]);
                uint64_t w1 = fmask(cb[1], ce[1] - cb[1], co[1]);
                uint64_t w2 = fmask(cb[2], ce[2] - cb[2], co[2]);
                batch_loop(w0, w1, w2, cb[0], cb[1], cb[2]);
                continue;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int r = (jy - iy + 1) + 3 * d;
                const int b = cb[d], e = ce[d];
                if (b >= e) continue;
                if (!((st.ok >> r) & 1)) {  // row left in global memory
                    scan_cell<3>(px, py, pz, b, e, s, xi, yi, zi, re, re2, f);
                    continue;
                }
                for (int c0 = b; c0 < e; c0 += 64) {
                    const int c1 = e - c0 < 64 ? e : c0 + 64;
                    uint64_t w0 = fmask(c0, c1 - c0, co[d]), w1 = 0, w2 = 0;
                    batch_loop(w0, w1, w2, c0, 0, 0);
                }
            }
        }
    }
}

// Neighbour lists of the pressure stages.  The boundary-condition pass walks
// the cells once at radius max(re_s, re_p) and records each slot's neighbours
// in the walk's (the reference's) order with their distances; the later walks
// over the same bucket build (diagonal and connectivity, matrix fill, minimum
// pressure, gradient, higher-order source) read the list instead of scanning
// the cells again.  Entry k of slot s sits at [k * stride + s], so the lanes
// of a wavefront read entry k together.  cnt[s] counts every neighbour; a
// slot with more than cap neighbours is walked again by its readers.
struct NbList {
    int* q;
    double* d;
    int* cnt;
    int cap;
    size_t stride;
};

// Calls f(q, dist, dx, dy, dz, gather(q)) for slot s's neighbours with
// dist < re (re at most the list's radius), in for_neighbors' order and with
// its values.  Entries go four at a time: the next four entries' indices are
// loaded while this four's positions and gather(q) values are read, all of
// them before f runs, so a group costs one round trip of dependent loads.
// gather(q) must only read (it runs for entries f then skips).
template <int DIM, typename G, typename F>
__device__ __forceinline__ void for_listed_g(const NbList& L, const GridView& g, const double* __restrict__ px,
                                             const double* __restrict__ py, const double* __restrict__ pz, int key,
                                             int s, double xi, double yi, double zi, double re, G&& gather, F&& f) {
    const int m = L.cnt[s];
    if (m > L.cap) {
        for_neighbors<DIM>(g, px, py, pz, key, s, xi, yi, zi, re,
                           [&](int q, double dist, double dx, double dy, double dz) {
                               f(q, dist, dx, dy, dz, gather(q));
                           });
        return;
    }
    if (m <= 0) return;
    int qn[4];
    double dn[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {  // entries past m re-read the last one
        const size_t e = (size_t) (i < m ? i : m - 1) * L.stride + s;
        qn[i] = L.q[e];
        dn[i] = L.d[e];
    }
    for (int k0 = 0; k0 < m; k0 += 4) {
        int qq[4];
        double dd[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            qq[i] = qn[i];
            dd[i] = dn[i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + 4 + i < m ? k0 + 4 + i : m - 1;
            const size_t e = (size_t) k * L.stride + s;
            qn[i] = L.q[e];
            dn[i] = L.d[e];
        }
        decltype(gather(0)) gv[4];
        double xq[4], yq[4], zq[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            gv[i] = gather(qq[i]);
            xq[i] = px[qq[i]];
            yq[i] = py[qq[i]];
            zq[i] = DIM == 3 ? pz[qq[i]] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (k0 + i >= m) break;
            const double dist = dd[i];
            if (!(dist < re)) continue;
            f(qq[i], dist, xq[i] - xi, yq[i] - yi, DIM == 3 ? zq[i] - zi : 0.0, gv[i]);
        }
    }
}

// Calls f(q, gather(q)) for slot s's listed neighbours, in list order, when
// re is the list's own radius (every entry passes dist < re, so neither the
// distances nor the positions are read: 4 bytes an entry instead of 12).
// Entries go four at a time as in for_listed_g; a slot whose list overflowed
// is walked with for_neighbors.
template <int DIM, typename G, typename F>
__device__ __forceinline__ void for_listed_ids(const NbList& L, const GridView& g, const double* __restrict__ px,
                                               const double* __restrict__ py, const double* __restrict__ pz,
                                               int key, int s, double xi, double yi, double zi, double re,
                                               G&& gather, F&& f) {
    const int m = L.cnt[s];
    if (m > L.cap) {
        for_neighbors<DIM>(g, px, py, pz, key, s, xi, yi, zi, re,
                           [&](int q, double, double, double, double) { f(q, gather(q)); });
        return;
    }
    if (m <= 0) return;
    int qn[4];
#pragma unroll
    for (int i = 0; i < 4; i++) qn[i] = L.q[(size_t) (i < m ? i : m - 1) * L.stride + s];
    for (int k0 = 0; k0 < m; k0 += 4) {
        int qq[4];
#pragma unroll
        for (int i = 0; i < 4; i++) qq[i] = qn[i];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + 4 + i < m ? k0 + 4 + i : m - 1;
            qn[i] = L.q[(size_t) k * L.stride + s];
        }
        decltype(gather(0)) gv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) gv[i] = gather(qq[i]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (k0 + i >= m) break;
            f(qq[i], gv[i]);
        }
    }
}

// for_listed_g without per-neighbour gathers: f(q, dist, dx, dy, dz).
template <int DIM, typename F>
__device__ __forceinline__ void for_listed(const NbList& L, const GridView& g, const double* __restrict__ px,
                                           const double* __restrict__ py, const double* __restrict__ pz, int key,
                                           int s, double xi, double yi, double zi, double re, F&& f) {
    for_listed_g<DIM>(L, g, px, py, pz, key, s, xi, yi, zi, re, [](int) { return 0; },
                      [&](int q, double dist, double dx, double dy, double dz, int) { f(q, dist, dx, dy, dz); });
}

// ---------------- launchers (defined in the .hip translation units) ----------------

// grid.hip
hipError_t scan_exclusive(const int* in, int* out, int m, int* tmp, hipStream_t st);
size_t scan_tmp_ints(int m);
hipError_t fill_int(int* p, int v, size_t n, hipStream_t st);
hipError_t fill_double(double* p, double v, size_t n, hipStream_t st);

// Small device-to-host reads that skip the copy engine and the stream wait:
// a one-block kernel stores the words into pinned host memory, then a
// sequence number with a system-scope release; the host spins on the
// number (after 2 ms it waits on the stream instead, which also reports a
// failed launch).  Reads are ordered with the stream's other work.
struct Mailbox {
    uint32_t* host = nullptr;  // [0] sequence, then the words (pinned, mapped)
    uint32_t* dev = nullptr;   // device address of host
    int cap = 0;               // words after the sequence
    uint32_t seq = 0;
};
struct MailSrc {
    const void* p;  // device address, 4-byte aligned
    int bytes;      // multiple of 4
};
// Reads the sources, back to back, into out (host memory).
hipError_t mailbox_read(Mailbox& mb, const MailSrc* src, int nsrc, void* out, hipStream_t st);
void mailbox_free(Mailbox& mb);

struct BuildBufs {
    int* cnt;        // ncell + 1 counters (kept at zero between builds)
    int* cs;         // ncell + 2 cell starts
    int* bin_id;     // n: ids binned per cell (arbitrary order within a cell)
    int* dst;        // n: binned position of each slot
    int* scan_tmp;
    int* slot_of;    // n: id -> slot, filled on demand by slot_of_ids
};
hipError_t slot_of_ids(int n, const int* id, int* slot_of, hipStream_t st);
hipError_t build_grid(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& cur, Hot& nxt,
                      const BuildBufs& b, int* err, hipStream_t st, int bs);
hipError_t dump_neighbor_counts(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& h,
                                int* cnt_by_id, double re, hipStream_t st);
hipError_t dump_neighbor_lists(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& h,
                               const int* off_by_id, int* ids, double* dists, double re, hipStream_t st);

// stages.hip
struct Cold {
    double* p;        // pressure (id order)
    double* pmin;     // minimum pressure (id order)
    double* ndr;      // number density ratio (id order)
    int8_t* bc;       // boundary condition (id order)
    double* pprev;    // last unclamped solution (id order), warm start
    double* pprev2;   // the one before it (extrapolated warm start)
    double* ghost_n;  // number density (re_p) of the last list a ghost had
    double* z;        // 2D: frozen z of every particle (id order)
    double* vz;       // 2D: frozen vz
};

struct StepScratch {
    int8_t* bc_s;       // slot order
    double* b_s;        // source term
    double* diag_s;     // matrix diagonal
    double* isq_s;      // 1 / sqrt(diagonal) of Inner slots (the Jacobi scaling)
    int* nnz_s;         // off-diagonal count per slot (Inner rows)
    int* parent;        // connectivity forest
    int8_t* surf;       // component holds a Surface particle
    int* inner_flag;    // 1 for Inner slots
    int* row_of;        // slot -> compact row (exclusive scan of inner_flag)
    double* p_s;        // pressure, slot order
    double* pmin_s;
    double* vmax;       // Courant reduction target (one double)
    int* coll_cnt;      // per id
    int* coll_off;      // per id + 1
    int* coll_pairs;    // 2 * cap
    double* coll_dist;  // cap
    int* coll_level;    // cap
    unsigned long long* coll_key;  // coll_n2: sort keys of pair lists too long for LDS
    int* coll_prev;     // cap
    int coll_cap;
    int coll_n2;        // power of two >= coll_cap
    int* scan_tmp;
    NbList nb;          // pressure-stage neighbour lists (written by the BC pass)
};

// Record of one step's collision pass for the reverse sweep (every pointer
// null: recording off).  Indexed like the candidate pair list.
struct CollTape {
    int* pid;     // 2 per pair: particle ids (i, j)
    int* flags;   // per pair: bit 0 the impulse was applied, bit 1 j is fluid
    double* pre;  // 12 per pair: x_i, x_j, v_i, v_j (3 each) just before the pair acted
    double* dir;  // 3 per pair: (x_j - x_i) / dist at the bucket build
    int* nlev;    // number of levels
};

hipError_t launch_visc_move(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& cur,
                            Hot& out, Cold& cold, Events& ev, hipStream_t st, int bs);
hipError_t launch_collision(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, Hot& h,
                            StepScratch& w, int* err, hipStream_t st, int bs, const CollTape* tape = nullptr);
hipError_t launch_bc_source(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                            Cold& cold, StepScratch& w, hipStream_t st, int bs);
hipError_t launch_diag_cc(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                          StepScratch& w, Events& ev, hipStream_t st, int bs);
hipError_t launch_pressure_out(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                               const double* x_s, Cold& cold, StepScratch& w, hipStream_t st, int bs);
hipError_t launch_grad_move(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& cur,
                            Hot& out, Cold& cold, StepScratch& w, Events& ev, hipStream_t st, int bs);
hipError_t launch_ndr(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                      Cold& cold, hipStream_t st, int bs);
hipError_t launch_courant(int dim, int n, const Hot& h, const SceneTab& tab, int nscene, double* vmax,
                          hipStream_t st);
hipError_t launch_snapshot(int dim, int n, const Hot& h, const Cold& cold, double* pos_id, double* vel_id,
                           hipStream_t st);
hipError_t launch_gather_type(int n, const Hot& h, int8_t* type_id, hipStream_t st);

// ppe.hip — matrix storage and CG
struct Sell {
    int dim;             // 2 or 3
    const GridView* scene_grids;  // host: each segment's bucket grid (multigrid aggregation)
    int rows;            // Inner rows
    const int* rows_dev; // Inner rows, on the device
    int nseg;            // independent segments (scenes)
    const int* scene_key0;  // device, nseg + 1: first key of each scene (last = ghost key)
    const int* cs;       // device: cell starts
    const int* row_of;   // device: slot -> compact row
    int nslices;
    int* slice_off;      // nslices + 1 (in entries)
    int* slice_w;        // nslices (max nnz of the slice)
    int* col;            // entries: column (compact row index)
    short* dcol;         // entries: 16-bit column codes, when use16 (see sell_col)
    int use16;           // the 16-bit codes hold every column
    int* anc;            // 3D, 2 per row: row offsets of the cells below and above the row's cell
    size_t anc_cap;      // rows anc holds
    int* maxdelta;       // device: max |column - row| of the last assembly
    int wide_hint;       // host: the last assembly needed 32-bit columns (write them in the first pass)
    double* val64;       // scaled off-diagonal values (FP64)
    float* val32;        // scaled off-diagonal values (FP32), mixed path
    int* row_slot;       // rows -> slot
    double* sdiag;       // 1/sqrt(diag) per row
    double* dgl;         // diag per row
    size_t cap;          // entry capacity
    size_t entries = 0;  // host: entries of the last assembly (slice widths x 64)
    // debug CSR (raw values by slot), filled when capture is on
    int* raw_off;        // n + 1 (by slot)
    int* raw_col;        // slots
    double* raw_val;
    size_t raw_cap;
    long long grows = 0;  // times the entry or anchor storage was reallocated
};

// MPS_AMD_TIGHT_CAPS=1 (tests): capacities grow to exactly what a system
// needs, so every larger system reallocates its storage.
bool tight_caps();

// *p, as a nontemporal load when NT: for data a kernel streams once per
// pass, so that it does not evict what the next kernels read again.
template <bool NT, typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Column of SELL entry idx of row `row`.  int: the column itself.  short
// (2D): column - row.  unsigned short (3D): the row's neighbours lie in
// three z layers of cells, each a compact run of rows, so the code holds
// the layer (0 the row's own, 1 below, 2 above; top 2 bits) and the offset
// + 8192 (14 bits) from the row itself in its own layer, else from the row's
// anchor in that layer (lo, hi: the first row of the cell directly below
// or above the row's cell, minus the row).
// NT: the load is marked nontemporal (the FP64 matrix, which a pass
// streams once and which does not fit the Infinity Cache beside the
// solver's vectors, so it should not push them out).
template <typename IDX, bool NT = false>
__device__ __forceinline__ int sell_col(const IDX* __restrict__ c, size_t idx, int row, int lo, int hi) {
    const IDX raw = ld_stream<NT>(c + idx);
    if constexpr (sizeof(IDX) == 4) {
        return (int) raw;
    } else if constexpr (static_cast<IDX>(-1) < 0) {
        return row + (int) raw;
    } else {
        const unsigned v = raw;
        const unsigned l = v >> 14;
        return row + ((int) (v & 16383u) - 8192) + (l == 1u ? lo : (l == 2u ? hi : 0));
    }
}
// The row's anchors (unsigned short codes only).
template <typename IDX>
__device__ __forceinline__ void sell_anchors(const int* __restrict__ anc, int row, int& lo, int& hi) {
    if constexpr (sizeof(IDX) == 2 && !(static_cast<IDX>(-1) < 0)) {
        lo = anc[2 * (size_t) row];
        hi = anc[2 * (size_t) row + 1];
    } else {
        lo = hi = 0;
    }
}
constexpr int kAnchMax = 8191;  // largest |offset| an unsigned short code holds

hipError_t assemble_sell(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                         const StepScratch& w, Sell& A, int want_val32, int want_raw, int* err,
                         int* host_rows, size_t* host_entries, Mailbox& mb, hipStream_t st, int bs);

// mg.hip — aggregation multigrid preconditioner (single segment)
struct MgWork;
MgWork* mg_create();
void mg_destroy(MgWork* m);
int mg_generation(const MgWork* m);
int mg_levels(const MgWork* m);
void mg_set_params(MgWork* m, double omega, double cf, int sweeps);
// Builds the hierarchy for the current system (every segment its own
// grids); hipErrorInvalidValue when it cannot (the caller then solves
// without it).  One check is left on the device: the int at mg_error_flag
// is nonzero after the stream has run the setup if a coarse row had more
// neighbours than the stencil allows, and the hierarchy must then not be
// used.
hipError_t mg_setup(MgWork* m, const Sell& A, const Hot& h, int dim, int fp32, hipStream_t st);
const int* mg_error_flag(const MgWork* m);
// count flags at stride ints from flag: the cycle is skipped once all are set
struct MgDone {
    const int* flag;
    int stride;
    int count;
};
// Enqueues u = M r (one V-cycle).  b1_given: the level-1 right-hand side
// (the cells' sums of r, see mg_fine_restriction) is already in place and
// the cycle does not form it; honoured only by the additive fine level.
template <typename T, typename IDX>
void mg_cycle(MgWork* m, const Sell& A, const IDX* fcol, const T* fval, const T* r, T* u, MgDone done, int dim,
              bool b1_given, hipStream_t st);
// When the current hierarchy's cycle restricts the fine residual r itself
// (the additive fine level): the first fine row of each level-1 cell
// (n1 + 1 entries, the last = rows) and the level-1 right-hand side b1[K] =
// sum of r over cell K's rows in row order, for a caller that forms b1
// where it writes r.  False otherwise.
bool mg_fine_restriction(const MgWork* m, const int** cst1, int* n1, void** b1);

struct CgWork;  // defined in ppe.hip
CgWork* cg_create();
void cg_destroy(CgWork* cw);
// Selects which record of last iteration counts sizes the CG chunks: 0 for
// the step's own solves, 1 for the adjoint's.
void cg_prediction_bank(CgWork* cw, int bank);
// stages.hip: counts of the LDS staging of the walks since the process
// started, kept when GridView::stage_stats is set: [0] blocks, [1] blocks
// that staged every stencil row, [2] some but not all rows, [3] of those,
// blocks that left a row out before a later staged one.
hipError_t stage_stats_read(long long out[4]);
hipError_t cg_reserve(CgWork* cw, int rows);
// [0] times a solver's cached graphs were dropped, [1] CG vector reallocations
void cg_debug_counts(const CgWork* cw, long long out[4]);
// Solves A x = b for the Inner rows.  b_s / x_s in slot order; x0_id (id order,
// may be NULL) gives the warm start.  Fills stats.
// stats: one entry per segment.
hipError_t cg_solve(CgWork* cw, const Sell& A, const Hot& h, int n, const double* b_s, const double* pprev_id,
                    const double* pprev2_id,
                    double* x_s, const mps_solver_options& o, mps_solver_stats* stats, hipStream_t st,
                    std::string& msg);
hipError_t cg_time_kernels(CgWork* cw, const Sell& A, int reps, long long nnz, hipStream_t st, double ms[4],
                           long long bytes[4]);

// adjoint.hip — reverse-mode step
struct AdjBufs {
    double* xb[3];   // dL/dx, id order
    double* vb[3];   // dL/dv, id order
    double* gb[3];   // adjoint of the pressure-gradient sum, slots of the pressure build
    double* pb;      // adjoint of the clamped pressure (slots)
    double* phb;     // adjoint of the minimum pressure (slots)
    int* amin;       // slot holding each slot's minimum pressure (-1: Ignored)
    double* rhs;     // adjoint Poisson right-hand side (slots)
    double* lam;     // adjoint Poisson solution (slots)
    double* ab[3];   // adjoint of the explicit acceleration, id order
    double* cnu;     // per-slot terms of dL/dnu
    double* crelax;  // per-slot terms of dL/drelaxation
    double* cpair;   // per-pair scratch of the collision sweep
    double* part;    // block partials of the parameter sums
    double* acc;     // dL/dnu, dL/drestitution, dL/drelaxation
};
constexpr int kStateDoubles = 12;  // doubles per particle of a saved state
constexpr int kStateBytes = 2;     // bytes per particle of a saved state
hipError_t save_state(int dim, int n, const Hot& h, const Cold& cold, double* f, int8_t* t8, hipStream_t st);
hipError_t restore_state(int dim, int n, Hot& h, Cold& cold, int* slot_of, const double* f, const int8_t* t8,
                         hipStream_t st);
size_t adj_partials(int n);
size_t adj_pair_scratch(int cap);
hipError_t adj_pressure_stage(int dim, const GridView& g, int n, const Consts& c, const Hot& h,
                              const int8_t* type_g, const StepScratch& w, const double* x_s, AdjBufs& a,
                              hipStream_t st);
hipError_t adj_ppe_stage(int dim, const GridView& g, int n, const Consts& c, const Hot& h, const int8_t* type_g,
                         const StepScratch& w, const double* x_s, AdjBufs& a, hipStream_t st);
hipError_t adj_collision_stage(int dim, int n, const Consts& c, const StepScratch& w, const CollTape& tp, AdjBufs& a,
                               hipStream_t st);
hipError_t adj_explicit_stage(int dim, const GridView& g, int n, const Consts& c, const int8_t* type0_id,
                              const Hot& hs, AdjBufs& a, hipStream_t st);

}  // namespace mps
