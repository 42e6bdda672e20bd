// ppe.hip — pressure Poisson system storage and the GPU conjugate-gradient
// solver that replaces Eigen's ConjugateGradient (src/mps.cpp:452-460).
//
// Storage.  Only Inner rows carry equations (src/mps.cpp:330-332); they are
// compacted in slot (cell) order, so a row's neighbours are a few hundred to a
// few thousand rows away and stay in L2.  Off-diagonals are kept in SELL-64:
// slices of 64 consecutive rows (one wavefront), each padded to its widest
// row and stored entry-major, so every load in the SpMV is one coalesced
// 64-lane access.  The system is Jacobi-scaled symmetrically,
//   Ahat = S A S, S = diag(1/sqrt(a_ii)),  Ahat yhat = S b,  x = S yhat,
// which is Jacobi-preconditioned CG in exact arithmetic (Eigen's default
// DiagonalPreconditioner) and leaves a unit diagonal that needs no storage.
//
// Solver.  One CG iteration is two kernels:
//   spmv  : p = r + beta p (own row, and each neighbour's on the fly), q = Ahat p,
//           partial sums of p.q -> last block forms alpha
//   update: y += alpha p, r -= alpha q, partial sums of r.r and of the
//           unscaled residual sum a_ii r_i^2 -> last block forms beta, tests
//           convergence and raises the device-side done flag.
// Reductions are deterministic: blocks write partials, the last block to
// finish (atomic ticket) sums them in block order.  Kernels return at once
// once the done flag is up, so the host enqueues fixed chunks of iterations
// (or replays a captured graph of them) and only reads the flag between
// chunks.
//
// Mixed precision: FP32 inner CG on the FP32 copy of Ahat, with an FP64
// residual b - A x recomputed from the FP64 values after each inner solve
// (iterative refinement) until the true relative residual meets the tolerance.
// Contraction off: distances and weights must round as the reference's
// scalar FP64 code does (no fused multiply-add).
#pragma clang fp contract(off)

#include <math.h>
#include <stdlib.h>
#include <cstring>

#include <algorithm>
#include <vector>

#include "mps_internal.hpp"

namespace mps {

static inline dim3 grid_for(int n) { return dim3((unsigned) ((n + kBlock - 1) / kBlock)); }


__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int) (blockDim.x >> 6); k++) t += red[k];
    __syncthreads();
    return t;  // valid in thread 0
}

// ---------------- assembly ----------------

__global__ void __launch_bounds__(kMaxTuneBlock) k_rows(int n, const int* __restrict__ inner_flag,
                                                 const int* __restrict__ row_of, const int* __restrict__ nnz_s,
                                                 int* __restrict__ row_slot, int* __restrict__ slice_w) {
    // The rows of one wavefront's slots are consecutive, so they lie in at
    // most two slices: the widths are reduced across the wavefront and set
    // with at most two atomics.  Every lane reaches the shuffles.
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < n && inner_flag[s];
    int r = 0, v = 0, sl = 0x7fffffff;
    if (act) {
        r = row_of[s];
        row_slot[r] = s;
        sl = r >> 6;
        v = nnz_s[s];
    }
    int smin = sl;
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(smin, o, 64);
        smin = u < smin ? u : smin;
    }
    if (smin == 0x7fffffff) return;  // no Inner slot in this wavefront (uniform)
    int v0 = act && sl == smin ? v : -1, v1 = act && sl != smin ? v : -1;
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(v0, o, 64), b = __shfl_xor(v1, o, 64);
        v0 = a > v0 ? a : v0;
        v1 = b > v1 ? b : v1;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&slice_w[smin], v0);
        if (v1 >= 0) atomicMax(&slice_w[smin + 1], v1);
    }
}

// SELL assembly: one wavefront per block and per slice.  A row's first
// kFillStage entries are staged in LDS while its neighbours are walked (each
// lane finds them in a different loop trip), then all 64 lanes write entry k
// of their rows together, so every store is one coalesced 64-lane line.
// Entries past kFillStage (dense 3D rows) are written directly.  The 32-bit
// column array is written only when asked (some offset does not fit 16 bits).
constexpr int kFillStage = 12;

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(64) k_fill_sell(GridView g0, Consts c0, SceneTab tab, Hot h,
                                                  const int* __restrict__ rows_p, const int* __restrict__ row_slot,
                                                  const int* __restrict__ row_of, const int8_t* __restrict__ bc_s,
                                                  const double* __restrict__ diag_s,
                                                  const double* __restrict__ isq_s,
                                                  const int* __restrict__ slice_pos, int* __restrict__ col,
                                                  short* __restrict__ dcol, int* __restrict__ maxdelta,
                                                  double* __restrict__ val64, float* __restrict__ val32,
                                                  double* __restrict__ sdiag, double* __restrict__ dgl,
                                                  const int* __restrict__ raw_off, int* __restrict__ raw_col,
                                                  double* __restrict__ raw_val, int write_col, NbList nb,
                                                  int* __restrict__ anc) {
    // 3D: 16-bit codes relative to the row's anchors in the cell layers below
    // and above (see sell_col); staged entries keep (column - row) * 4 + layer
    constexpr bool kAnch = DIM == 3;
    __shared__ double s_val[kFillStage][64];
    __shared__ int s_dl[kFillStage][64];
    int lo = 0, hi = 0;
    auto code16 = [&](int dl, int layer) {
        const int off = dl - (layer == 1 ? lo : (layer == 2 ? hi : 0));
        return kAnch ? (short) ((layer << 14) | (off + 8192)) : (short) dl;
    };
    const int lane = threadIdx.x;
    const int sl = blockIdx.x;
    const int r = sl * 64 + lane;
    const int rows = *rows_p;
    if (sl * 64 >= rows) return;  // block-uniform
    const size_t base = (size_t) slice_pos[sl] * 64 + lane;
    const int width = slice_pos[sl + 1] - slice_pos[sl];
    int k = 0;
    int md = 0;
    if (r < rows) {
        const int s = row_slot[r];
        GridView g = g0;
        Consts c = c0;
        scene_of<MULTI>(tab, h.id[s], g, c);
        const double di = diag_s[s];
        const double si = isq_s[s];  // 1 / sqrt(di), from the connectivity pass
        sdiag[r] = si;
        dgl[r] = di;
        int ro = raw_off ? raw_off[s] : 0;
        const double re = c.re_p;
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        // the row's layer and anchors: first rows of the cells directly below
        // and above its own (a flat grid has one layer)
        const int nxy = g.nx * g.ny;
        int iz = 0;
        if (kAnch && !g.flat) {
            const int* __restrict__ csl = g.cs + g.key0;
            const int kl = h.key[s] - g.key0;
            iz = kl / nxy;
            if (iz > 0) lo = row_of[csl[kl - nxy]] - r;
            if (iz < g.nz - 1) hi = row_of[csl[kl + nxy]] - r;
        }
        if (kAnch) {
            anc[2 * (size_t) r] = lo;
            anc[2 * (size_t) r + 1] = hi;
        }
        struct Nbr {
            int bc;      // boundary class
            double isq;  // 1 / sqrt(diag_s[q])
            int row;     // compacted row
            int key;     // cell key
        };
        for_listed_g<DIM>(nb, g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                          [&](int q) { return Nbr{(int) bc_s[q], isq_s[q], row_of[q], kAnch ? h.key[q] : 0}; },
                          [&](int q, double dist, double, double, double, const Nbr& nq) {
                               if (nq.bc != MPS_BC_INNER) return;
                               // the off-diagonal value of src/mps.cpp:343-348
                               const double a = c.coefA * wfun(dist, re) / c.rho;
                               const double v = -a;
                               const double sj = nq.isq;
                               const double sv = v * (si * sj);
                               const int cq = nq.row;
                               const int dl = cq - r;
                               int layer = 0;
                               if (kAnch && !g.flat) {
                                   const int zq = (nq.key - g.key0) / nxy;
                                   layer = zq < iz ? 1 : (zq > iz ? 2 : 0);
                               }
                               const int off = dl - (layer == 1 ? lo : (layer == 2 ? hi : 0));
                               const int ad = off < 0 ? -off : off;
                               md = ad > md ? ad : md;
                               if (k < kFillStage) {
                                   s_val[k][lane] = sv;
                                   s_dl[k][lane] = dl * 4 + layer;
                               } else {
                                   const size_t idx = base + (size_t) k * 64;
                                   dcol[idx] = code16(dl, layer);
                                   if (write_col) col[idx] = cq;
                                   val64[idx] = sv;
                                   if (val32) val32[idx] = (float) sv;
                               }
                               if (raw_col) {
                                   raw_col[ro] = q;
                                   raw_val[ro] = v;
                                   ro++;
                               }
                               k++;
                           });
    }
    // coalesced write-out of the staged entries; padding (value 0, own row) to the slice width
    const int ws = width < kFillStage ? width : kFillStage;
    for (int e = 0; e < ws; e++) {
        const size_t idx = base + (size_t) e * 64;
        const bool have = e < k;
        const int dq = have ? s_dl[e][lane] : 0;
        const int dl = dq >> 2, layer = dq & 3;  // padding: own row, own layer
        const double sv = have ? s_val[e][lane] : 0.0;
        dcol[idx] = code16(dl, layer);
        if (write_col) col[idx] = r + dl;
        val64[idx] = sv;
        if (val32) val32[idx] = (float) sv;
    }
    for (int e = k > kFillStage ? k : kFillStage; e < width; e++) {
        const size_t idx = base + (size_t) e * 64;
        dcol[idx] = code16(0, 0);
        if (write_col) col[idx] = r;
        val64[idx] = 0.0;
        if (val32) val32[idx] = 0.0f;
    }
    // at most one atomic per wavefront (the block is one wavefront; every
    // lane is here).  Every wavefront aims at the same word, so the atomic is
    // skipped when the word already holds md: it only grows during the
    // launch, so a stale read is too low, never too high, and costs at most
    // an atomic that was not needed.
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(md, o, 64);
        md = u > md ? u : md;
    }
    if (lane == 0 && md > 0 && __hip_atomic_load(maxdelta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < md)
        atomicMax(maxdelta, md);
}

hipError_t assemble_sell(int dim, const GridView& g, const SceneTab& tab, int n, const Consts& c, const Hot& h,
                         const StepScratch& w,
                         Sell& A, int want_val32, int want_raw, int* err, int* host_rows, size_t* host_entries,
                         Mailbox& mb, hipStream_t st, int bs) {
    (void) err;
    hipError_t e;
    const int nsl_max = (n + 63) / 64;
    // compact Inner rows in slot order
    e = scan_exclusive(w.inner_flag, w.row_of, n, w.scan_tmp, st);
    if (e != hipSuccess) return e;
    e = fill_int(A.slice_w, 0, (size_t) nsl_max + 1, st);
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(k_rows, dim3((n + bs - 1) / bs), dim3(bs), 0, st, n, w.inner_flag, w.row_of, w.nnz_s, A.row_slot,
                           A.slice_w);
    e = scan_exclusive(A.slice_w, A.slice_off, nsl_max, w.scan_tmp, st);
    if (e != hipSuccess) return e;
    int hv[2];
    {
        const MailSrc src[2] = {{w.row_of + n, (int) sizeof(int)}, {A.slice_off + nsl_max, (int) sizeof(int)}};
        if ((e = mailbox_read(mb, src, 2, hv, st)) != hipSuccess) return e;
    }
    const int rows = hv[0];
    const size_t entries = (size_t) hv[1] * 64;
    *host_rows = rows;
    *host_entries = entries;
    A.entries = entries;
    A.rows = rows;
    A.rows_dev = w.row_of + n;
    A.nslices = (rows + 63) / 64;
    if (entries > A.cap) {
        size_t cap = tight_caps() ? entries : entries + entries / 4 + 64 * 64;
        A.grows++;
        hipFree(A.col);
        hipFree(A.dcol);
        hipFree(A.val64);
        hipFree(A.val32);
        A.col = nullptr;
        A.dcol = nullptr;
        A.val64 = nullptr;
        A.val32 = nullptr;
        A.cap = 0;
        if ((e = hipMalloc(&A.col, cap * sizeof(int))) != hipSuccess) return e;
        if ((e = hipMalloc(&A.dcol, cap * sizeof(short))) != hipSuccess) return e;
        if ((e = hipMalloc(&A.val64, cap * sizeof(double))) != hipSuccess) return e;
        if ((e = hipMalloc(&A.val32, cap * sizeof(float))) != hipSuccess) return e;
        A.cap = cap;
    }
    if (want_raw) {
        e = scan_exclusive(w.nnz_s, A.raw_off, n, w.scan_tmp, st);
        if (e != hipSuccess) return e;
        int tot = 0;
        e = hipMemcpyAsync(&tot, A.raw_off + n, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        if ((size_t) tot > A.raw_cap) {
            hipFree(A.raw_col);
            hipFree(A.raw_val);
            A.raw_col = nullptr;
            A.raw_val = nullptr;
            A.raw_cap = 0;
            size_t cap = (size_t) tot + 1024;
            if ((e = hipMalloc(&A.raw_col, cap * sizeof(int))) != hipSuccess) return e;
            if ((e = hipMalloc(&A.raw_val, cap * sizeof(double))) != hipSuccess) return e;
            A.raw_cap = cap;
        }
    }
    A.use16 = 0;
    if (rows == 0) return hipSuccess;
    const int* rows_p = w.row_of + n;
    if ((e = fill_int(A.maxdelta, 0, 1, st)) != hipSuccess) return e;
    // one wavefront per slice; the 32-bit columns only if some offset needs
    // them.  The first pass writes them when the previous assembly needed
    // them (the particle order changes little between steps); a second pass
    // walks the neighbours again only when that guess was wrong.
    const dim3 fg((rows + 63) / 64), fb(64);
    const int lim16 = dim == 3 ? kAnchMax : 32767;
    if (dim == 3 && (size_t) rows > A.anc_cap) {
        hipFree(A.anc);
        A.anc = nullptr;
        A.anc_cap = 0;
        A.grows++;
        const size_t cap = tight_caps() ? (size_t) rows : (size_t) rows + rows / 4 + 64;
        if ((e = hipMalloc(&A.anc, 2 * cap * sizeof(int))) != hipSuccess) return e;
        A.anc_cap = cap;
    }
    int md = 0;
    for (int pass = 0; pass < 2; pass++) {
        const int write_col = pass || A.wide_hint;
#define FILL_ARGS                                                                                                  \
    g, c, tab, h, rows_p, A.row_slot, w.row_of, w.bc_s, w.diag_s, w.isq_s, A.slice_off, A.col, A.dcol, A.maxdelta, A.val64, \
        want_val32 ? A.val32 : nullptr, A.sdiag, A.dgl, (want_raw && !pass) ? A.raw_off : nullptr,                \
        (want_raw && !pass) ? A.raw_col : nullptr, (want_raw && !pass) ? A.raw_val : nullptr, write_col, w.nb, A.anc
        if (dim == 2 && tab.g)
            hipLaunchKernelGGL((k_fill_sell<2, true>), fg, fb, 0, st, FILL_ARGS);
        else if (dim == 2)
            hipLaunchKernelGGL((k_fill_sell<2, false>), fg, fb, 0, st, FILL_ARGS);
        else if (tab.g)
            hipLaunchKernelGGL((k_fill_sell<3, true>), fg, fb, 0, st, FILL_ARGS);
        else
            hipLaunchKernelGGL((k_fill_sell<3, false>), fg, fb, 0, st, FILL_ARGS);
#undef FILL_ARGS
        // 16-bit column codes when every offset fits: within +-32767 rows of
        // the row (2D), within +-8191 of the row's anchor in the entry's
        // layer (3D)
        const MailSrc src{A.maxdelta, (int) sizeof(int)};
        if ((e = mailbox_read(mb, &src, 1, &md, st)) != hipSuccess) return e;
        if (md <= lim16 || write_col) break;
    }
    A.use16 = md <= lim16 ? 1 : 0;
    A.wide_hint = A.use16 ? 0 : 1;
    return hipGetLastError();
}

// ---------------- CG ----------------
//
// The system holds one or more independent segments (one per scene of the
// handle); a segment is a contiguous range of rows with its own scalars,
// stopping test and done flag.  Segment s is solved by nb_s = cg_blocks(rows_s)
// consecutive blocks of every CG launch; block v of the segment owns a
// contiguous range of its rows (a multiple of 64) and writes its partial sums
// to part[(first block of s + v) * NV + k].  A kernel that needs the sums
// starts by reducing the segment's partials in one fixed order, redundantly
// in every block of the segment: the kernel boundary makes them visible on
// every XCD, no fence or atomic is needed, and every block derives the same
// scalars bit for bit.  A segment's arithmetic is the same whether it is
// solved alone or next to others, so an ensemble member reproduces its solo
// run exactly.
//
// Iteration (Chronopoulos-Gear form of CG, one reduction point):
//   apply : w = Ahat r; partials of r.r, w.r and the stopping norm
//   step  : beta = g/g_old, alpha = g/(d - beta g/alpha_old);
//           p = r + beta p; s = w + beta s; y += alpha p; r -= alpha s
// apply gathers only r, once per stored entry; step is elementwise.

constexpr int kMaxBlocks = 1024;  // per segment
// Threads per block of the per-iteration kernels (apply, step): 1024 blocks
// of 512 fill every SIMD with 8 waves, twice the loads in flight of 256.
constexpr int kCgBlock = 512;
constexpr int kStepPre = 4;  // rows a thread of k_cg_step loads ahead of the partials' reduction
// Below this many rows the V-cycle's setup and launches cost more than the
// iterations it saves (measured: 584-particle dam break), so small systems
// keep the Jacobi scaling alone.
constexpr int kMgMinRows = 8192;
// FP64 CG restarts from the current iterate when the true residual of the
// recurrence's answer misses the tolerance.
constexpr int kMaxRestarts = 3;

// Rows per block of a segment below kMaxBlocks blocks: one per thread.
// Segments large enough to fill the chip (the 2M and 3.6M dam breaks) take
// kMaxBlocks blocks whatever this is.  It depends on the segment's rows
// alone, so an ensemble member keeps its solo run's block layout and bits.
// Measured (ms/step FP64 / mixed at 256, 512 and 1024 rows a block): 64-scene
// ensemble 16.67 / 15.13, 15.60 / 13.62, 15.71 / 13.25 (at 256 half of each
// 512-thread block idled, and each of a segment's many blocks reduced all
// of its partials); 2D 134k 2.76 / 2.87, 2.70 / 2.81, 2.85 / 2.91 (1024
// leaves CUs idle); 2D 35k 1.82 / 1.98, 1.85 / 2.00, 1.93 / 2.09.
// MPS_AMD_CG_ROWS overrides it (a tuning aid, read once).
static int cg_rows_per_block() {
    static const int v = [] {
        const char* e = getenv("MPS_AMD_CG_ROWS");
        const int r = e ? atoi(e) : 0;
        return r >= 64 ? r : kCgBlock;
    }();
    return v;
}

static int cg_blocks(int rows) {
    const int rpb = cg_rows_per_block();
    int nb = (rows + rpb - 1) / rpb;
    if (nb > kMaxBlocks) nb = kMaxBlocks;
    return nb < 1 ? 1 : nb;
}

struct CgScalars {
    double gam[2];     // r.r by iteration parity
    double alp[2];     // alpha by iteration parity
    double thr;        // stop when rho2 <= thr
    double rho2;       // latest residual^2 (unscaled sum a_ii r_i^2, or r.r for inner solves)
    double red[3];     // results of k_reduce_parts
    int iter;          // apply launches run (read by every block of k_cg_step)
    int max_iter;
    int done;
    int converged;
    int frozen;        // timing mode: kernels leave every scalar as it is
    int fin;           // completed updates, written once when the segment stops
};

struct Seg {
    int nseg;
    const int* blk_off;  // nseg + 1: first block of each segment
    const int* row_off;  // nseg + 1: first row of each segment
};

__device__ __forceinline__ int seg_of_block(const Seg& sg, int b) {
    int lo = 0, hi = sg.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sg.blk_off[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct BlockRows {
    int seg, v, nb, b0;  // segment, block within it, its block count, its first block
    int lb;              // the block's place among the segment's contiguous row ranges
    int r0, r1;          // rows of this block
};

// Row range of this block: contiguous within the segment, whole multiples of
// 64 rows from the segment start, XCD-grouped.
__device__ __forceinline__ BlockRows block_rows(const Seg& sg) {
    BlockRows br;
    br.seg = seg_of_block(sg, blockIdx.x);
    br.b0 = sg.blk_off[br.seg];
    br.v = blockIdx.x - br.b0;
    br.nb = sg.blk_off[br.seg + 1] - br.b0;
    const int base = sg.row_off[br.seg];
    const int rows = sg.row_off[br.seg + 1] - base;
    const int per = ((rows + br.nb - 1) / br.nb + 63) & ~63;
    const int lb = xcd_block(br.v, br.nb);
    br.lb = lb;
    br.r0 = base + lb * per;
    const int end = base + rows;
    br.r1 = br.r0 + per < end ? br.r0 + per : end;
    if (br.r0 > end) br.r0 = end;
    return br;
}

// Fixed-order sum of NV interleaved partials (part[b*NV + k], b < nb) in every
// block; result broadcast to all threads.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nb, double (&out)[NV],
                                                double* red) {
    double acc[NV];
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < NV; k++) acc[k] += part[(size_t) NV * b + k];
    for (int k = 0; k < NV; k++) {
        double a = acc[k];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * NV + k] = a;
    }
    __syncthreads();
    for (int k = 0; k < NV; k++) {
        double t = 0.0;
        for (int w = 0; w < (int) (blockDim.x >> 6); w++) t += red[w * NV + k];
        out[k] = t;
    }
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void write_partials(const double (&v)[NV], double* __restrict__ part, double* red) {
    for (int k = 0; k < NV; k++) {
        double a = v[k];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * NV + k] = a;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = 0.0;
        for (int w = 0; w < (int) (blockDim.x >> 6); w++) t += red[w * NV + threadIdx.x];
        part[(size_t) NV * blockIdx.x + threadIdx.x] = t;
    }
}

__device__ __forceinline__ int block_uniform_done(CgScalars* sc) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = *(volatile int*) &sc->done;
    __syncthreads();
    return s_done;
}

// acc + sum_k val[k] x[col[k]] over the row's SELL entries, gathers issued in
// batches of 4 for memory-level parallelism.  NT: the matrix is read with
// nontemporal loads (the FP64 solve's apply: its 10 bytes an entry do not
// fit the Infinity Cache beside the solver's vectors, and streaming them
// keeps the vectors and coarse levels cached: 2M FP64 step -6 %; the FP32
// matrix does fit and is read normally, as is the FP64 residual's, where
// nontemporal loads made the mixed step 2 % slower).
template <typename T, typename IDX, bool NT = false>
__device__ __forceinline__ T sell_row(const int* __restrict__ slice_pos, const IDX* __restrict__ col,
                                      const int* __restrict__ anc, const T* __restrict__ val,
                                      const T* __restrict__ x, int row, T acc) {
    const int sl = row >> 6, lane = row & 63;
    const int s0 = slice_pos[sl], w = slice_pos[sl + 1] - s0;
    const size_t base = (size_t) s0 * 64 + lane;
    int lo, hi;
    sell_anchors<IDX>(anc, row, lo, hi);
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const size_t i0 = base + (size_t) k * 64;
        const int c0 = sell_col<IDX, NT>(col, i0, row, lo, hi), c1 = sell_col<IDX, NT>(col, i0 + 64, row, lo, hi);
        const int c2 = sell_col<IDX, NT>(col, i0 + 128, row, lo, hi), c3 = sell_col<IDX, NT>(col, i0 + 192, row, lo, hi);
        const T v0 = ld_stream<NT>(val + i0), v1 = ld_stream<NT>(val + i0 + 64);
        const T v2 = ld_stream<NT>(val + i0 + 128), v3 = ld_stream<NT>(val + i0 + 192);
        const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc += v0 * x0;
        acc += v1 * x1;
        acc += v2 * x2;
        acc += v3 * x3;
    }
    for (; k < w; k++) {
        const size_t i0 = base + (size_t) k * 64;
        acc += ld_stream<NT>(val + i0) * x[sell_col<IDX, NT>(col, i0, row, lo, hi)];
    }
    return acc;
}

// apply: w = Ahat u with u = M r (u = r without a preconditioner; unit
// diagonal); partials r.u, w.u and the stopping norm of r.
template <typename T, typename IDX, bool WEIGHTED, bool PRE>
__global__ void __launch_bounds__(kCgBlock) k_cg_apply(Seg sg, const int* __restrict__ slice_pos,
                                                     const IDX* __restrict__ col, const int* __restrict__ anc,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ r, const T* __restrict__ u,
                                                     T* __restrict__ w, const double* __restrict__ dgl,
                                                     double* __restrict__ part, CgScalars* scs) {
#pragma clang fp contract(fast)
    __shared__ double red[3 * kCgBlock / 64];
    const BlockRows br = block_rows(sg);
    CgScalars* sc = scs + br.seg;
    if (block_uniform_done(sc)) return;
    if (br.v == 0 && threadIdx.x == 0 && !sc->frozen) sc->iter += 1;
    const T* x = PRE ? u : r;
    double g = 0.0, d = 0.0, rw = 0.0;
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) {
        const T rr = r[row];
        const T uu = PRE ? x[row] : rr;
        // the FP64 solve streams its matrix (see sell_col)
        const T acc = sell_row<T, IDX, sizeof(T) == 8>(slice_pos, col, anc, val, x, row, uu);
        w[row] = acc;
        const double r2 = (double) rr * (double) rr;
        g += (double) rr * (double) uu;
        d += (double) acc * (double) uu;
        rw += WEIGHTED ? dgl[row] * r2 : r2;
    }
    double v[3] = {g, d, rw};
    write_partials<3>(v, part, red);
}

// Row ranges of the CG step that hold whole level-1 cells (one segment):
// entry v <= nb is (first row, first cell) of range v, its start moved from
// v * per (per as in block_rows) to the first cell that starts there or
// later; cst1 holds the cells' first rows, cst1[n1] = rows.
__global__ void __launch_bounds__(kBlock) k_step_cells(int nb, int rows, int n1, const int* __restrict__ cst1,
                                                       int2* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > nb) return;
    const int per = ((rows + nb - 1) / nb + 63) & ~63;
    const long long s0 = (long long) v * per;
    const int start = s0 < rows ? (int) s0 : rows;
    int lo = 0, hi = n1;  // first cell K <= n1 with cst1[K] >= start
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cst1[mid] < start)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[v] = make_int2(cst1[lo], lo);
}

// step: convergence test on r, then the vector recurrences.  RST (one
// segment, additive multigrid): the block takes the whole cells of
// cells[lb] .. cells[lb + 1] and, once its rows of r are written, forms the
// next cycle's level-1 right-hand side b1[K] = sum of r over cell K.
template <typename T, bool WEIGHTED, bool PRE, bool RST>
__global__ void __launch_bounds__(kCgBlock) k_cg_step(Seg sg, T* __restrict__ y, T* __restrict__ r, T* __restrict__ p,
                                                    T* __restrict__ s, const T* __restrict__ w,
                                                    const T* __restrict__ u, const double* __restrict__ part,
                                                    CgScalars* scs, int par, const int2* __restrict__ cells,
                                                    const int* __restrict__ cst1, T* __restrict__ b1) {
#pragma clang fp contract(fast)
    __shared__ double red[3 * kCgBlock / 64];
    BlockRows br = block_rows(sg);
    int k0 = 0, k1 = 0;
    if (RST) {
        const int2 a = cells[br.lb], b = cells[br.lb + 1];
        br.r0 = a.x;
        br.r1 = b.x;
        k0 = a.y;
        k1 = b.y;
    }
    CgScalars* sc = scs + br.seg;
    if (block_uniform_done(sc)) return;
    // the first kStepPre rows' vectors are loaded before the partials are
    // reduced, so the loads overlap the reduction's round trips (a block has
    // about 4 rows a thread at 2M)
    T vr[kStepPre], vw[kStepPre], vu[kStepPre], vp[kStepPre], vs[kStepPre], vy[kStepPre];
#pragma unroll
    for (int k = 0; k < kStepPre; k++) {
        const int row = br.r0 + (int) threadIdx.x + k * kCgBlock;
        vr[k] = vw[k] = vu[k] = vp[k] = vs[k] = vy[k] = (T) 0;
        if (row < br.r1) {
            vr[k] = r[row];
            vw[k] = w[row];
            vu[k] = PRE ? u[row] : vr[k];
            vp[k] = p[row];
            vs[k] = s[row];
            vy[k] = y[row];
        }
    }
    double sums[3];
    reduce_partials<3>(part + (size_t) 3 * br.b0, br.nb, sums, red);
    const double g = sums[0], dd = sums[1], rho = sums[2];
    const bool lead = br.v == 0 && threadIdx.x == 0;
    const int it = sc->iter;  // this iteration's apply has run: it = k + 1
    const bool first = it == 1;
    T alpha, beta;
    if (sc->frozen) {
        alpha = (T) 0.0;  // y and r stay, p and s stay bounded
        beta = (T) 0.5;
    } else {
        const bool conv = rho <= sc->thr;
        if (conv || it > sc->max_iter || !(rho == rho)) {
            // iter itself stays: the other blocks of this launch read it
            // and must all see the same count to take the same branch
            if (lead) {
                sc->done = 1;
                sc->converged = conv ? 1 : 0;
                sc->rho2 = rho;
                sc->fin = it - 1;
            }
            return;
        }
        const double b = first ? 0.0 : g / sc->gam[par ^ 1];
        const double a = first ? g / dd : g / (dd - b * g / sc->alp[par ^ 1]);
        if (lead) {
            sc->gam[par] = g;
            sc->alp[par] = a;
            sc->rho2 = rho;
        }
        alpha = (T) a;
        beta = (T) b;
    }
#pragma unroll
    for (int k = 0; k < kStepPre; k++) {
        const int row = br.r0 + (int) threadIdx.x + k * kCgBlock;
        if (row < br.r1) {
            const T pp = first ? vu[k] : vu[k] + beta * vp[k];
            const T ss = first ? vw[k] : vw[k] + beta * vs[k];
            p[row] = pp;
            s[row] = ss;
            y[row] = vy[k] + alpha * pp;
            r[row] = vr[k] - alpha * ss;
        }
    }
    for (int row = br.r0 + (int) threadIdx.x + kStepPre * kCgBlock; row < br.r1; row += kCgBlock) {
        const T rr = r[row], ww = w[row];
        const T uu = PRE ? u[row] : rr;
        const T pp = first ? uu : uu + beta * p[row];
        const T ss = first ? ww : ww + beta * s[row];
        p[row] = pp;
        s[row] = ss;
        y[row] += alpha * pp;
        r[row] = rr - alpha * ss;
    }
    if (RST) {
        __syncthreads();  // the block's rows of r, written above, are visible to all its threads
        for (int K = k0 + (int) threadIdx.x; K < k1; K += kCgBlock) b1[K] = cell_sum<T>(r, cst1[K], cst1[K + 1]);
    }
}

// One block per segment: sums its NV-interleaved partials into sc[seg].red.
// Each segment's NV partial sums into red[OFF .. OFF + NV).
template <int NV, int OFF = 0>
__global__ void __launch_bounds__(kBlock) k_reduce_parts(Seg sg, const double* __restrict__ part, CgScalars* scs) {
    static_assert(OFF + NV <= 3, "three result slots");
    __shared__ double red[3 * kBlock / 64];
    const int sgi = blockIdx.x;
    const int b0 = sg.blk_off[sgi], nb = sg.blk_off[sgi + 1] - b0;
    double o[NV];
    reduce_partials<NV>(part + (size_t) NV * b0, nb, o, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) scs[sgi].red[OFF + k] = o[k];
}

// Inner-solve set-up on the device, so the host need not read ||r32||^2
// first: segment s's threshold is ||r32||^2 (k_inner_init's partials) times
// the reduction the host stored in thr; a segment whose residual is zero
// does not iterate.
__global__ void __launch_bounds__(kBlock) k_inner_prep(Seg sg, const double* __restrict__ part, CgScalars* scs) {
    __shared__ double red[3 * kBlock / 64];
    const int sgi = blockIdx.x;
    const int b0 = sg.blk_off[sgi], nb = sg.blk_off[sgi + 1] - b0;
    double o[1];
    reduce_partials<1>(part + b0, nb, o, red);
    if (threadIdx.x == 0) {
        CgScalars* sc = scs + sgi;
        sc->red[0] = o[0];
        sc->thr = o[0] * sc->thr;
        if (!(o[0] > 0.0)) sc->done = 1;
    }
}

// FP64 init: yhat = x0 sqrt(d), bhat = S b; partial ||b||^2.
__global__ void __launch_bounds__(kCgBlock) k_cg_init(Seg sg, const int* __restrict__ row_slot,
                                                    const int32_t* __restrict__ id, const double* __restrict__ b_s,
                                                    const double* __restrict__ pprev,
                                                    const double* __restrict__ pprev2, const double* __restrict__ sdiag,
                                                    double* __restrict__ y, double* __restrict__ bh,
                                                    double* __restrict__ part) {
    __shared__ double red[3 * kCgBlock / 64];
    const BlockRows br = block_rows(sg);
    double bb = 0.0;
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) {
        const int s = row_slot[row];
        const double b = b_s[s];
        const double si = sdiag[row];
        bh[row] = b * si;
        double x0 = 0.0;
        if (pprev) {
            // previous solution, or the line through the last two where both
            // steps solved for this particle
            const double a = pprev[id[s]];
            x0 = a;
            if (pprev2) {
                const double c = pprev2[id[s]];
                if (a != 0.0 && c != 0.0) x0 = 2.0 * a - c;
            }
        }
        y[row] = x0 / si;
        bb += b * b;
    }
    double v[1] = {bb};
    write_partials<1>(v, part, red);
}

// FP64 residual r = bhat - Ahat y (unit diagonal); partials r.r, sum d r^2.
template <typename IDX>
__global__ void __launch_bounds__(kCgBlock) k_residual64(Seg sg, const int* __restrict__ slice_pos,
                                                       const IDX* __restrict__ col, const int* __restrict__ anc,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ y, const double* __restrict__ bh,
                                                       const double* __restrict__ dgl, double* __restrict__ r,
                                                       double* __restrict__ part) {
#pragma clang fp contract(fast)
    __shared__ double red[3 * kCgBlock / 64];
    const BlockRows br = block_rows(sg);
    double rr = 0.0, rw = 0.0;
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) {
        const double acc = sell_row<double, IDX>(slice_pos, col, anc, val, y, row, y[row]);
        const double rv = bh[row] - acc;
        r[row] = rv;
        rr += rv * rv;
        rw += dgl[row] * (rv * rv);
    }
    double v[2] = {rr, rw};
    write_partials<2>(v, part, red);
}

// Inner solve setup: r32 = r64 (scaled residual), e = 0; partial r.r.
__global__ void __launch_bounds__(kCgBlock) k_inner_init(Seg sg, const double* __restrict__ r64,
                                                       float* __restrict__ r32, float* __restrict__ e32,
                                                       double* __restrict__ part) {
    __shared__ double red[3 * kCgBlock / 64];
    const BlockRows br = block_rows(sg);
    double rr = 0.0;
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) {
        const float rv = (float) r64[row];
        r32[row] = rv;
        e32[row] = 0.0f;
        rr += (double) rv * (double) rv;
    }
    double v[1] = {rr};
    write_partials<1>(v, part, red);
}

__global__ void __launch_bounds__(kCgBlock) k_add_correction(Seg sg, double* __restrict__ y, const float* __restrict__ e) {
    const BlockRows br = block_rows(sg);
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) y[row] += (double) e[row];
}

__global__ void __launch_bounds__(kCgBlock) k_cg_finish(Seg sg, const int* __restrict__ row_slot,
                                                      const double* __restrict__ y, const double* __restrict__ sdiag,
                                                      double* __restrict__ x_s) {
    const BlockRows br = block_rows(sg);
    for (int row = br.r0 + (int) threadIdx.x; row < br.r1; row += kCgBlock) x_s[row_slot[row]] = y[row] * sdiag[row];
}

// Per-segment first rows: row_off[s] = compact row of the first slot of scene s.
__global__ void k_seg_rows(int nseg, const int* __restrict__ scene_key0, const int* __restrict__ cs,
                           const int* __restrict__ row_of, int* __restrict__ row_off) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= nseg) row_off[s] = row_of[cs[scene_key0[s]]];
}

struct CgWork {
    int cap = 0;
    int nseg_cap = 0;
    int nseg = 1;
    int total_blocks = 1;     // sum of the segments' block counts
    double* y = nullptr;      // FP64 solution (scaled)
    double* bh = nullptr;     // scaled rhs
    double* r64 = nullptr;    // FP64 residual (scaled)
    double* p64 = nullptr;
    double* s64 = nullptr;
    double* w64 = nullptr;
    float* r32 = nullptr;
    float* p32 = nullptr;
    float* s32 = nullptr;
    float* w32 = nullptr;
    float* e32 = nullptr;     // inner correction
    double* part = nullptr;   // block partials (3 per block)
    size_t part_cap = 0;
    CgScalars* sc = nullptr;  // device scalars, one per segment
    CgScalars* host = nullptr;
    Mailbox mb;  // scalar reads
    int* blk_off = nullptr;   // device, nseg + 1
    int* row_off = nullptr;   // device, nseg + 1
    int* h_blk = nullptr;     // pinned host mirrors
    int* h_row = nullptr;
    double* u64 = nullptr;    // preconditioned residual M r
    float* u32 = nullptr;
    MgWork* mg = nullptr;     // multigrid preconditioner (single segment)
    int mg_on = 0;            // the current solve uses it
    // the step forms the cycle's level-1 right-hand side (additive fine
    // level, one segment): its cell-aligned row ranges and the hierarchy's
    // cell starts and b1; MPS_AMD_CG_FUSE=0 leaves it to the cycle
    bool fuse_ok = true;
    int fused = 0;
    int2* step_cells = nullptr;  // kMaxBlocks + 1
    const int* f_cst1 = nullptr;
    void* f_b1 = nullptr;
    // chunk graphs of the FP64 CG [0] and the FP32 inner CG [1], one per
    // chunk length (the first chunk of a solve is sized by prediction)
    // Chunk lengths are even and at most 64, so 32 slots hold every length a
    // solver can ask for: once captured, a length is never captured again
    // while the signature below holds (with 8 slots replaced in turn, the
    // mixed path's inner solves cycled through more lengths than that and
    // captured a graph again every few steps, ~1.3 ms each at 2M).
    static constexpr int kGraphSlots = 32;
    hipGraphExec_t graph[2][kGraphSlots] = {};
    int graph_chunk[2][kGraphSlots] = {};
    long long graph_used[2][kGraphSlots] = {};  // last use, for replacing the least recent
    long long graph_clock = 0;
    long long graph_captures = 0;  // chunk graphs captured
    // everything a captured chunk bakes in: every device pointer its kernels
    // receive, the launch shape and the multigrid hierarchy's generation.
    // Any difference (a buffer reallocated by a larger system, another
    // column width) drops the cached graphs of that solver.
    static constexpr int kSig = 24;
    uintptr_t graph_sig[2][kSig] = {};
    static constexpr int kPredSolves = 4;
    bool predict = true;  // MPS_AMD_CG_PREDICT=0: fixed chunks of the check interval
    // last step's iterations per solve of a step; bank 1 holds the adjoint
    // solves', whose counts (zero start) differ from the forward ones'
    int pred[2][2][kPredSolves] = {};
    int pred_bank = 0;
    long long graph_rebuilds = 0;  // times a solver's cached graphs were dropped
    long long vec_reallocs = 0;    // times the CG vectors were reallocated
    long long launched = 0;        // CG iterations enqueued, including those past convergence
    Seg seg() const { return Seg{nseg, blk_off, row_off}; }
};

void cg_prediction_bank(CgWork* cw, int bank) { cw->pred_bank = bank ? 1 : 0; }

CgWork* cg_create() {
    CgWork* cw = new CgWork();
    if (const char* v = getenv("MPS_AMD_CG_PREDICT")) cw->predict = atoi(v) != 0;
    if (const char* v = getenv("MPS_AMD_CG_FUSE")) cw->fuse_ok = atoi(v) != 0;
    return cw;
}

static void cg_free_vectors(CgWork* cw) {
    void* ptrs[] = {cw->y,   cw->bh,  cw->r64, cw->p64, cw->s64, cw->w64, cw->r32,
                    cw->p32, cw->s32, cw->w32, cw->e32, cw->u64, cw->u32};
    for (void* p : ptrs) hipFree(p);
}

void cg_destroy(CgWork* cw) {
    if (!cw) return;
    cg_free_vectors(cw);
    hipFree(cw->part);
    hipFree(cw->sc);
    hipFree(cw->blk_off);
    hipFree(cw->row_off);
    hipFree(cw->step_cells);
    hipHostFree(cw->host);
    hipHostFree(cw->h_blk);
    hipHostFree(cw->h_row);
    mailbox_free(cw->mb);
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < CgWork::kGraphSlots; j++)
            if (cw->graph[k][j]) hipGraphExecDestroy(cw->graph[k][j]);
    mg_destroy(cw->mg);
    delete cw;
}

static hipError_t cg_reserve_segs(CgWork* cw, int nseg) {
    if (nseg <= cw->nseg_cap) return hipSuccess;
    hipFree(cw->sc);
    hipFree(cw->blk_off);
    hipFree(cw->row_off);
    hipHostFree(cw->host);
    hipHostFree(cw->h_blk);
    hipHostFree(cw->h_row);
    hipError_t e;
    if ((e = hipMalloc(&cw->sc, sizeof(CgScalars) * (size_t) nseg))) return e;
    if ((e = hipMalloc(&cw->blk_off, sizeof(int) * (size_t) (nseg + 1)))) return e;
    if ((e = hipMalloc(&cw->row_off, sizeof(int) * (size_t) (nseg + 1)))) return e;
    if ((e = hipHostMalloc(&cw->host, sizeof(CgScalars) * (size_t) nseg, hipHostMallocDefault))) return e;
    if ((e = hipHostMalloc(&cw->h_blk, sizeof(int) * (size_t) (nseg + 1), hipHostMallocDefault))) return e;
    if ((e = hipHostMalloc(&cw->h_row, sizeof(int) * (size_t) (nseg + 1), hipHostMallocDefault))) return e;
    cw->nseg_cap = nseg;
    return hipSuccess;
}

bool tight_caps() {
    const char* v = getenv("MPS_AMD_TIGHT_CAPS");
    return v && atoi(v) != 0;
}

void cg_debug_counts(const CgWork* cw, long long out[4]) {
    out[0] = cw->graph_rebuilds;
    out[1] = cw->vec_reallocs;
    out[2] = cw->launched;
    out[3] = cw->graph_captures;
}

hipError_t cg_reserve(CgWork* cw, int rows) {
    if (rows <= cw->cap && cw->y) return hipSuccess;
    const int cap = tight_caps() ? rows : rows + rows / 8 + 1024;
    cw->vec_reallocs++;
    cg_free_vectors(cw);
    hipError_t e;
    const size_t d = (size_t) cap * sizeof(double), f = (size_t) cap * sizeof(float);
    if ((e = hipMalloc(&cw->y, d)) || (e = hipMalloc(&cw->bh, d)) || (e = hipMalloc(&cw->r64, d)) ||
        (e = hipMalloc(&cw->p64, d)) || (e = hipMalloc(&cw->s64, d)) || (e = hipMalloc(&cw->w64, d)) ||
        (e = hipMalloc(&cw->r32, f)) || (e = hipMalloc(&cw->p32, f)) || (e = hipMalloc(&cw->s32, f)) ||
        (e = hipMalloc(&cw->w32, f)) || (e = hipMalloc(&cw->e32, f)) || (e = hipMalloc(&cw->u64, d)) ||
        (e = hipMalloc(&cw->u32, f)))
        return e;
    cw->cap = cap;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < CgWork::kGraphSlots; j++) {
            if (cw->graph[k][j]) hipGraphExecDestroy(cw->graph[k][j]);
            cw->graph[k][j] = nullptr;
            cw->graph_chunk[k][j] = 0;
        }
    return hipSuccess;
}

// Segment layout of the current system: per-segment rows from the device,
// block counts from the live row counts, partial storage sized to match.
static hipError_t cg_layout(CgWork* cw, const Sell& A, hipStream_t st) {
    hipError_t e = cg_reserve_segs(cw, A.nseg);
    if (e) return e;
    cw->nseg = A.nseg;
    if (A.nseg == 1) {
        cw->h_row[0] = 0;
        cw->h_row[1] = A.rows;
    } else {
        hipLaunchKernelGGL(k_seg_rows, dim3((A.nseg + 1 + 255) / 256), dim3(256), 0, st, A.nseg, A.scene_key0, A.cs,
                           A.row_of, cw->row_off);
        const MailSrc src{cw->row_off, (int) sizeof(int) * (A.nseg + 1)};
        if ((e = mailbox_read(cw->mb, &src, 1, cw->h_row, st))) return e;
    }
    cw->h_blk[0] = 0;
    for (int s = 0; s < A.nseg; s++) cw->h_blk[s + 1] = cw->h_blk[s] + cg_blocks(cw->h_row[s + 1] - cw->h_row[s]);
    cw->total_blocks = cw->h_blk[A.nseg];
    const size_t need = (size_t) 3 * cw->total_blocks;
    if (need > cw->part_cap) {
        hipFree(cw->part);
        cw->part = nullptr;
        if ((e = hipMalloc(&cw->part, sizeof(double) * (need + 3 * 1024)))) return e;
        cw->part_cap = need + 3 * 1024;
    }
    if ((e = hipMemcpyAsync(cw->blk_off, cw->h_blk, sizeof(int) * (size_t) (A.nseg + 1), hipMemcpyHostToDevice, st)))
        return e;
    if (A.nseg == 1 &&
        (e = hipMemcpyAsync(cw->row_off, cw->h_row, sizeof(int) * 2, hipMemcpyHostToDevice, st)))
        return e;
    return hipSuccess;
}

// Enqueue `chunk` CG iterations (two kernels each); chunk is even, so the
// scalar parities line up from one chunk to the next.
template <typename T> static T* u_vec(CgWork* cw);
template <> double* u_vec<double>(CgWork* cw) { return cw->u64; }
template <> float* u_vec<float>(CgWork* cw) { return cw->u32; }

template <typename T, typename IDX, bool WEIGHTED>
static void enqueue_chunk(CgWork* cw, const Sell& A, const IDX* col, const T* val, T* y, T* r, T* p, T* s, T* w,
                          int chunk, hipStream_t st) {
    const dim3 gr(cw->total_blocks), bl(kCgBlock);
    const Seg sg = cw->seg();
    T* u = u_vec<T>(cw);
    for (int k = 0; k < chunk; k++) {
        if (cw->mg_on) {
            const MgDone dn{&cw->sc[0].done, (int) (sizeof(CgScalars) / sizeof(int)), cw->nseg};
            // after the chunk's first step the steps leave b1 in place
            mg_cycle<T, IDX>(cw->mg, A, col, val, r, u, dn, A.dim, cw->fused && k > 0, st);
            hipLaunchKernelGGL((k_cg_apply<T, IDX, WEIGHTED, true>), gr, bl, 0, st, sg, A.slice_off, col, A.anc, val, r,
                               u, w,
                               A.dgl, cw->part, cw->sc);
            if (cw->fused)
                hipLaunchKernelGGL((k_cg_step<T, WEIGHTED, true, true>), gr, bl, 0, st, sg, y, r, p, s, w, u, cw->part,
                                   cw->sc, k & 1, cw->step_cells, cw->f_cst1, (T*) cw->f_b1);
            else
                hipLaunchKernelGGL((k_cg_step<T, WEIGHTED, true, false>), gr, bl, 0, st, sg, y, r, p, s, w, u, cw->part,
                                   cw->sc, k & 1, nullptr, nullptr, nullptr);
        } else {
            hipLaunchKernelGGL((k_cg_apply<T, IDX, WEIGHTED, false>), gr, bl, 0, st, sg, A.slice_off, col, A.anc, val, r,
                               r,
                               w, A.dgl, cw->part, cw->sc);
            hipLaunchKernelGGL((k_cg_step<T, WEIGHTED, false, false>), gr, bl, 0, st, sg, y, r, p, s, w, r, cw->part,
                               cw->sc, k & 1, nullptr, nullptr, nullptr);
        }
    }
}

static bool all_done(const CgWork* cw) {
    for (int s = 0; s < cw->nseg; s++)
        if (!cw->host[s].done) return false;
    return true;
}

// Runs chunks of CG iterations (each a captured graph, or enqueued directly)
// until every segment is done; iterations past convergence return at once.
// pred > 0 (the iterations this solve took in the previous step): chunks
// cover the prediction rounded down to even (at most 64 each), then, once it
// is used up, 2, 4, 8, ... 64; else chunks of the check interval.
template <typename T, typename IDX, bool WEIGHTED>
static hipError_t run_cg_loop(CgWork* cw, int which, const Sell& A, const IDX* col, const T* val, T* y, T* r, T* p,
                              T* s, T* w, const mps_solver_options& o, int pred, hipStream_t st) {
    const int interval = std::max(2, (o.check_interval + 1) & ~1);
    const int idx16 = sizeof(IDX) == 4 ? 0 : (static_cast<IDX>(-1) < 0 ? 1 : 2);
    const size_t sbytes = sizeof(CgScalars) * (size_t) cw->nseg;
    hipError_t e;
    if (o.use_graph) {
        const uintptr_t sig[CgWork::kSig] = {
            (uintptr_t) val,         (uintptr_t) col,          (uintptr_t) A.anc,       (uintptr_t) A.slice_off,
            (uintptr_t) A.dgl,       (uintptr_t) y,            (uintptr_t) r,           (uintptr_t) p,
            (uintptr_t) s,           (uintptr_t) w,            (uintptr_t) u_vec<T>(cw), (uintptr_t) cw->part,
            (uintptr_t) cw->sc,      (uintptr_t) cw->blk_off,  (uintptr_t) cw->row_off, (uintptr_t) A.rows_dev,
            (uintptr_t) cw->nseg,    (uintptr_t) cw->total_blocks, (uintptr_t) idx16,
            (uintptr_t) (cw->mg_on ? mg_generation(cw->mg) + 1 : 0), (uintptr_t) cw->mg_on,
            (uintptr_t) (cw->mg_on && cw->fused), (uintptr_t) cw->step_cells, 0};
        if (std::memcmp(sig, cw->graph_sig[which], sizeof(sig)) != 0) {
            for (int j = 0; j < CgWork::kGraphSlots; j++) {
                if (cw->graph[which][j]) hipGraphExecDestroy(cw->graph[which][j]);
                cw->graph[which][j] = nullptr;
                cw->graph_chunk[which][j] = 0;
            }
            std::memcpy(cw->graph_sig[which], sig, sizeof(sig));
            cw->graph_rebuilds++;
        }
    }
    int launched = 0, grow = 2;
    for (;;) {
        int chunk = interval;
        if (pred > 0) {
            const int rem = (pred - launched) & ~1;  // the prediction, rounded down to even
            if (rem >= 2) {
                chunk = rem < 64 ? rem : 64;
            } else {
                chunk = grow;
                grow = grow < 64 ? 2 * grow : 64;
            }
        }
        launched += chunk;
        cw->launched += chunk;
        if (o.use_graph) {
            int slot = -1;
            for (int j = 0; j < CgWork::kGraphSlots; j++)
                if (cw->graph[which][j] && cw->graph_chunk[which][j] == chunk) slot = j;
            if (slot < 0) {
                // a free slot, else the least recently used one
                slot = 0;
                for (int j = 0; j < CgWork::kGraphSlots; j++) {
                    if (!cw->graph[which][j]) {
                        slot = j;
                        break;
                    }
                    if (cw->graph_used[which][j] < cw->graph_used[which][slot]) slot = j;
                }
                if (cw->graph[which][slot]) hipGraphExecDestroy(cw->graph[which][slot]);
                cw->graph[which][slot] = nullptr;
                hipGraph_t gph;
                if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) return e;
                enqueue_chunk<T, IDX, WEIGHTED>(cw, A, col, val, y, r, p, s, w, chunk, st);
                if ((e = hipStreamEndCapture(st, &gph))) return e;
                e = hipGraphInstantiate(&cw->graph[which][slot], gph, nullptr, nullptr, 0);
                hipGraphDestroy(gph);
                if (e) return e;
                cw->graph_chunk[which][slot] = chunk;
                cw->graph_captures++;
            }
            cw->graph_used[which][slot] = ++cw->graph_clock;
            if ((e = hipGraphLaunch(cw->graph[which][slot], st))) return e;
        } else {
            enqueue_chunk<T, IDX, WEIGHTED>(cw, A, col, val, y, r, p, s, w, chunk, st);
        }
        const MailSrc src{cw->sc, (int) sbytes};
        if ((e = mailbox_read(cw->mb, &src, 1, cw->host, st))) return e;
        if (all_done(cw)) return hipSuccess;
    }
}

static hipError_t upload_scalars(CgWork* cw, hipStream_t st) {
    return hipMemcpyAsync(cw->sc, cw->host, sizeof(CgScalars) * (size_t) cw->nseg, hipMemcpyHostToDevice, st);
}

// Reduces nv-interleaved partials of every segment into sc[s].red and reads
// all scalars back.
static hipError_t reduce_and_read(CgWork* cw, int nv, hipStream_t st) {
    const Seg sg = cw->seg();
    if (nv == 1)
        hipLaunchKernelGGL(k_reduce_parts<1>, dim3(cw->nseg), dim3(kBlock), 0, st, sg, cw->part, cw->sc);
    else
        hipLaunchKernelGGL(k_reduce_parts<2>, dim3(cw->nseg), dim3(kBlock), 0, st, sg, cw->part, cw->sc);
    const MailSrc src{cw->sc, (int) (sizeof(CgScalars) * (size_t) cw->nseg)};
    return mailbox_read(cw->mb, &src, 1, cw->host, st);
}

static void launch_residual(CgWork* cw, const Sell& A, hipStream_t st) {
    const dim3 gr(cw->total_blocks), bl(kCgBlock);
    const Seg sg = cw->seg();
    if (A.use16 && A.dim == 3)
        hipLaunchKernelGGL(k_residual64<unsigned short>, gr, bl, 0, st, sg, A.slice_off, (const unsigned short*) A.dcol,
                           A.anc, A.val64, cw->y, cw->bh, A.dgl, cw->r64, cw->part);
    else if (A.use16)
        hipLaunchKernelGGL(k_residual64<short>, gr, bl, 0, st, sg, A.slice_off, A.dcol, A.anc, A.val64, cw->y, cw->bh,
                           A.dgl, cw->r64, cw->part);
    else
        hipLaunchKernelGGL(k_residual64<int>, gr, bl, 0, st, sg, A.slice_off, A.col, A.anc, A.val64, cw->y, cw->bh,
                           A.dgl, cw->r64, cw->part);
}

// `solve`: index of this solve within the step (0 for the FP64 CG, the
// refinement for the inner CG); the last step's iteration count of the same
// solve sizes the chunks.
template <typename T, bool WEIGHTED>
static hipError_t run_cg(CgWork* cw, int which, int solve, const Sell& A, const T* val, T* y, T* r, T* p, T* s, T* w,
                         const mps_solver_options& o, hipStream_t st) {
    const int slot = solve < CgWork::kPredSolves ? solve : CgWork::kPredSolves - 1;
    int* const last = &cw->pred[cw->pred_bank][which][slot];
    const int pred = cw->predict ? *last : 0;
    hipError_t e;
    if (A.use16 && A.dim == 3)
        e = run_cg_loop<T, unsigned short, WEIGHTED>(cw, which, A, (const unsigned short*) A.dcol, val, y, r, p, s, w, o,
                                                     pred, st);
    else if (A.use16)
        e = run_cg_loop<T, short, WEIGHTED>(cw, which, A, A.dcol, val, y, r, p, s, w, o, pred, st);
    else
        e = run_cg_loop<T, int, WEIGHTED>(cw, which, A, A.col, val, y, r, p, s, w, o, pred, st);
    if (e) return e;
    int it = 0;  // the segment that took longest sets the chunk
    for (int k = 0; k < cw->nseg; k++) it = std::max(it, cw->host[k].fin);
    *last = it;
    return hipSuccess;
}

hipError_t cg_solve(CgWork* cw, const Sell& A, const Hot& h, int n, const double* b_s, const double* pprev_id,
                    const double* pprev2_id,
                    double* x_s, const mps_solver_options& o, mps_solver_stats* stats, hipStream_t st,
                    std::string& msg) {
    (void) n;
    const int rows = A.rows;
    const int nseg = A.nseg;
    for (int s = 0; s < nseg; s++) {
        stats[s] = mps_solver_stats{};
        stats[s].converged = 1;
    }
    if (rows == 0) return hipSuccess;
    hipError_t e = cg_reserve(cw, rows);
    if (e) return e;
    if ((e = cg_layout(cw, A, st))) return e;
    const dim3 gr(cw->total_blocks), bl(kCgBlock);
    const Seg sg = cw->seg();
    // multigrid preconditioner (each segment its own hierarchy), built in the
    // precision of the CG it serves; if it cannot be built the solve runs
    // with the Jacobi scaling alone
    cw->mg_on = 0;
    if (o.preconditioner == MPS_PRECOND_MULTIGRID && rows >= kMgMinRows) {
        if (!cw->mg) cw->mg = mg_create();
        const hipError_t me = mg_setup(cw->mg, A, h, A.dim, o.precision == MPS_PREC_MIXED, st);
        if (me == hipSuccess)
            cw->mg_on = 1;
        else if (me != hipErrorInvalidValue)
            return me;
    }
    cw->fused = 0;
    if (cw->mg_on && cw->fuse_ok && nseg == 1) {
        const int* cst1 = nullptr;
        int n1 = 0;
        void* b1 = nullptr;
        if (mg_fine_restriction(cw->mg, &cst1, &n1, &b1)) {
            if (!cw->step_cells && (e = hipMalloc(&cw->step_cells, sizeof(int2) * (kMaxBlocks + 1)))) return e;
            hipLaunchKernelGGL(k_step_cells, dim3((cw->total_blocks + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                               cw->total_blocks, rows, n1, cst1, cw->step_cells);
            cw->fused = 1;
            cw->f_cst1 = cst1;
            cw->f_b1 = b1;
        }
    }

    hipLaunchKernelGGL(k_cg_init, gr, bl, 0, st, sg, A.row_slot, h.id, b_s, o.warm_start ? pprev_id : nullptr,
                       o.warm_start == 2 ? pprev2_id : nullptr, A.sdiag, cw->y, cw->bh, cw->part);
    // ||b||^2 into red[2], then the initial residual: one read for both
    // (and for the multigrid setup's stencil check)
    hipLaunchKernelGGL((k_reduce_parts<1, 2>), dim3(nseg), dim3(kBlock), 0, st, sg, cw->part, cw->sc);
    launch_residual(cw, A, st);
    hipLaunchKernelGGL(k_reduce_parts<2>, dim3(nseg), dim3(kBlock), 0, st, sg, cw->part, cw->sc);
    {
        int mg_bad = 0;
        const MailSrc src[2] = {{cw->sc, (int) (sizeof(CgScalars) * (size_t) nseg)},
                                {cw->mg_on ? (const void*) mg_error_flag(cw->mg) : (const void*) cw->sc,
                                 (int) sizeof(int)}};
        std::vector<CgScalars> buf((size_t) nseg + 1);
        if ((e = mailbox_read(cw->mb, src, 2, buf.data(), st))) return e;
        std::memcpy(cw->host, buf.data(), sizeof(CgScalars) * (size_t) nseg);
        std::memcpy(&mg_bad, (const char*) buf.data() + sizeof(CgScalars) * (size_t) nseg, sizeof(int));
        if (cw->mg_on && mg_bad) cw->mg_on = 0;  // a coarse row outside the stencil: Jacobi alone
    }
    std::vector<double> bn2(nseg), thr(nseg), rho2(nseg);
    std::vector<int> segrows(nseg), conv(nseg), total_it(nseg, 0), refinements(nseg, 0), max_it(nseg),
        restarts(nseg, 0);
    for (int s = 0; s < nseg; s++) {
        segrows[s] = cw->h_row[s + 1] - cw->h_row[s];
        bn2[s] = cw->host[s].red[2];
        rho2[s] = cw->host[s].red[1];
        stats[s].rows = segrows[s];
        stats[s].rhs_norm = sqrt(bn2[s]);
        thr[s] = o.tolerance * o.tolerance * bn2[s];
        max_it[s] = o.max_iterations > 0 ? o.max_iterations : 2 * segrows[s];
    }
    // Eigen: a zero right-hand side gives x = 0 (no iteration); warm starts
    // are discarded for such segments
    bool any_zero = false;
    for (int s = 0; s < nseg; s++) any_zero |= bn2[s] == 0.0;
    if (any_zero) {
        if (nseg == 1) {
            if ((e = fill_double(cw->y, 0.0, (size_t) rows, st))) return e;
        } else {
            for (int s = 0; s < nseg; s++)
                if (bn2[s] == 0.0 && segrows[s] > 0)
                    if ((e = fill_double(cw->y + cw->h_row[s], 0.0, (size_t) segrows[s], st))) return e;
        }
        launch_residual(cw, A, st);
        if ((e = reduce_and_read(cw, 2, st))) return e;
        for (int s = 0; s < nseg; s++) rho2[s] = cw->host[s].red[1];
    }
    for (int s = 0; s < nseg; s++) {
        conv[s] = bn2[s] == 0.0 || rho2[s] <= thr[s];
        stats[s].initial_residual = bn2[s] > 0 ? sqrt(rho2[s] / bn2[s]) : 0.0;
    }

    if (o.precision == MPS_PREC_FP64) {
        // CG from the current iterate (r64 holds its true residual), then the
        // true residual of the result decides convergence: the recurrence's
        // residual drifts from it by rounding, so a segment whose true
        // residual misses the tolerance restarts from its iterate while
        // iterations remain and restarts still halve it (a tolerance below
        // the rounding floor of A x stops after one restart)
        std::vector<int> stalled(nseg, 0);
        for (int pass = 0; pass <= kMaxRestarts; pass++) {
            std::vector<int> ran(nseg);
            bool any = false;
            for (int s = 0; s < nseg; s++) {
                ran[s] = !conv[s] && !stalled[s] && total_it[s] < max_it[s];
                any = any || ran[s];
                CgScalars z{};
                z.thr = thr[s];
                z.max_iter = max_it[s] - total_it[s];
                z.done = !ran[s];
                cw->host[s] = z;
            }
            if (!any) break;
            if ((e = upload_scalars(cw, st))) return e;
            if ((e = run_cg<double, true>(cw, 0, pass, A, A.val64, cw->y, cw->r64, cw->p64, cw->s64, cw->w64, o,
                                          st)))
                return e;
            for (int s = 0; s < nseg; s++)
                if (ran[s]) total_it[s] += cw->host[s].fin;
            launch_residual(cw, A, st);
            if ((e = reduce_and_read(cw, 2, st))) return e;
            for (int s = 0; s < nseg; s++) {
                if (!ran[s]) continue;
                const double before = rho2[s];
                rho2[s] = cw->host[s].red[1];
                conv[s] = rho2[s] <= thr[s];
                restarts[s] += pass > 0;
                stalled[s] = pass > 0 && !(rho2[s] <= 0.25 * before);
            }
        }
    } else {
        const int max_ref = o.max_refinements > 0 ? o.max_refinements : 20;
        const double eta = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-4;
        for (int round = 0;; round++) {
            bool active = false;
            for (int s = 0; s < nseg; s++)
                active = active || (!conv[s] && refinements[s] < max_ref && total_it[s] < max_it[s]);
            if (!active) break;
            hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, sg, cw->r64, cw->r32, cw->e32, cw->part);
            std::vector<int> run(nseg, 0);
            for (int s = 0; s < nseg; s++) {
                CgScalars z{};
                run[s] = !conv[s] && refinements[s] < max_ref && total_it[s] < max_it[s];
                // reduce the inner residual by eta, or by what the outer loop
                // still needs when that is less
                double red2 = eta * eta;
                if (run[s]) {
                    const double need = thr[s] / rho2[s];
                    if (need > red2) red2 = need * 0.25;
                }
                z.thr = red2;  // times ||r32||^2 in k_inner_prep
                z.max_iter = max_it[s] - total_it[s];
                z.done = run[s] ? 0 : 1;
                cw->host[s] = z;
            }
            if ((e = upload_scalars(cw, st))) return e;
            hipLaunchKernelGGL(k_inner_prep, dim3(nseg), dim3(kBlock), 0, st, sg, cw->part, cw->sc);
            if ((e = run_cg<float, false>(cw, 1, round, A, A.val32, cw->e32, cw->r32, cw->p32, cw->s32, cw->w32, o,
                                          st)))
                return e;
            for (int s = 0; s < nseg; s++)
                if (run[s]) total_it[s] += cw->host[s].fin;
            // segments that did not run keep e = 0 from k_inner_init
            hipLaunchKernelGGL(k_add_correction, gr, bl, 0, st, sg, cw->y, cw->e32);
            launch_residual(cw, A, st);
            if ((e = reduce_and_read(cw, 2, st))) return e;
            for (int s = 0; s < nseg; s++) {
                if (!run[s]) continue;
                rho2[s] = cw->host[s].red[1];
                conv[s] = rho2[s] <= thr[s];
                refinements[s]++;
            }
        }
    }
    // rho2 now holds the true FP64 residual of the final iterate: both
    // branches (and a solve that needed no iteration) computed it from it
    bool all_ok = true;
    for (int s = 0; s < nseg; s++) {
        stats[s].iterations = total_it[s];
        stats[s].refinements = o.precision == MPS_PREC_FP64 ? restarts[s] : refinements[s];
        stats[s].rel_residual = bn2[s] > 0 ? sqrt(rho2[s] / bn2[s]) : 0.0;
        stats[s].converged = conv[s] ? 1 : 0;
        all_ok = all_ok && conv[s];
    }
    hipLaunchKernelGGL(k_cg_finish, gr, bl, 0, st, sg, A.row_slot, cw->y, A.sdiag, x_s);
    if (!all_ok) msg = "pressure solve stopped above its tolerance";
    return hipGetLastError();
}

// The variants the solver itself runs: PRE when the last solve used the
// multigrid preconditioner (the apply gathers u = M r), else plain.
template <typename IDX, bool PRE>
static void time_launch(CgWork* cw, const Sell& A, const IDX* col, int kind, hipStream_t st) {
    const dim3 gr(cw->total_blocks), bl(kCgBlock);
    const Seg sg = cw->seg();
    switch (kind) {
        case 0:
            hipLaunchKernelGGL((k_cg_apply<double, IDX, true, PRE>), gr, bl, 0, st, sg, A.slice_off, col, A.anc,
                               A.val64, cw->r64, PRE ? cw->u64 : cw->r64, cw->w64, A.dgl, cw->part, cw->sc);
            break;
        case 1:
            hipLaunchKernelGGL((k_cg_step<double, true, PRE, false>), gr, bl, 0, st, sg, cw->y, cw->r64, cw->p64,
                               cw->s64, cw->w64, PRE ? cw->u64 : cw->r64, cw->part, cw->sc, 0, nullptr, nullptr,
                               nullptr);
            break;
        case 2:
            hipLaunchKernelGGL((k_cg_apply<float, IDX, false, PRE>), gr, bl, 0, st, sg, A.slice_off, col, A.anc,
                               A.val32, cw->r32, PRE ? cw->u32 : cw->r32, cw->w32, A.dgl, cw->part, cw->sc);
            break;
        default:
            hipLaunchKernelGGL((k_cg_step<float, false, PRE, false>), gr, bl, 0, st, sg, cw->e32, cw->r32, cw->p32,
                               cw->s32, cw->w32, PRE ? cw->u32 : cw->r32, cw->part, cw->sc, 0, nullptr, nullptr,
                               nullptr);
            break;
    }
}

template <typename IDX>
static void time_launch(CgWork* cw, const Sell& A, const IDX* col, int kind, bool pre, hipStream_t st) {
    if (pre)
        time_launch<IDX, true>(cw, A, col, kind, st);
    else
        time_launch<IDX, false>(cw, A, col, kind, st);
}

hipError_t cg_time_kernels(CgWork* cw, const Sell& A, int reps, long long nnz, hipStream_t st, double ms[4],
                           long long bytes[4]) {
    for (int k = 0; k < 4; k++) {
        ms[k] = 0.0;
        bytes[k] = 0;
    }
    const int rows = A.rows;
    if (rows == 0 || reps <= 0) return hipSuccess;
    hipError_t e = cg_reserve(cw, rows);
    if (e) return e;
    if ((e = cg_layout(cw, A, st))) return e;
    const dim3 gr(cw->total_blocks), bl(kCgBlock);
    const Seg sg = cw->seg();
    // finite vector contents
    const size_t R = (size_t) rows;
    const bool pre = cw->mg_on != 0;
    if ((e = fill_double(cw->r64, 1e-3, R, st)) || (e = fill_double(cw->p64, 0.0, R, st)) ||
        (e = fill_double(cw->s64, 0.0, R, st)) || (e = fill_double(cw->w64, 0.0, R, st)) ||
        (e = fill_double(cw->y, 0.0, R, st)) || (e = fill_double(cw->u64, 1e-3, R, st)))
        return e;
    hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, sg, cw->r64, cw->u32, cw->e32, cw->part);
    hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, sg, cw->r64, cw->r32, cw->e32, cw->part);
    hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, sg, cw->p64, cw->p32, cw->s32, cw->part);
    hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, sg, cw->p64, cw->w32, cw->e32, cw->part);
    if ((e = fill_double(cw->part, 1.0, 3 * (size_t) cw->total_blocks, st))) return e;
    for (int s = 0; s < cw->nseg; s++) {
        CgScalars z{};
        z.frozen = 1;
        z.iter = 2;
        z.max_iter = 0x7fffffff;
        z.thr = -1.0;
        cw->host[s] = z;
    }
    if ((e = upload_scalars(cw, st))) return e;
    hipEvent_t a, b;
    if ((e = hipEventCreate(&a))) return e;
    if ((e = hipEventCreate(&b))) return e;
    for (int kind = 0; kind < 4; kind++) {
        // one untimed launch, then reps timed launches
        for (int rep = -1; rep < reps; rep++) {
            if (rep == 0) hipEventRecord(a, st);
            if (A.use16 && A.dim == 3)
                time_launch<unsigned short>(cw, A, (const unsigned short*) A.dcol, kind, pre, st);
            else if (A.use16)
                time_launch<short>(cw, A, A.dcol, kind, pre, st);
            else
                time_launch<int>(cw, A, A.col, kind, pre, st);
        }
        hipEventRecord(b, st);
        if ((e = hipEventSynchronize(b))) break;
        float t = 0.f;
        hipEventElapsedTime(&t, a, b);
        ms[kind] = (double) t / reps;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (e) return e;
    const long long Rl = rows;
    const long long ib = A.use16 ? 2 : 4;
    const long long ab = A.use16 && A.dim == 3 ? 8 : 0;  // the row's two anchors
    const long long pu = pre ? 1 : 0;  // u = M r read besides r
    // apply: matrix entries + r (+ u) read + w written (+ d when weighted)
    bytes[0] = nnz * (ib + 8) + Rl * (8 * (3 + pu) + ab);
    // step: r, w, (u,) p, s, y read; p, s, y, r written
    bytes[1] = Rl * (9 + pu) * 8;
    bytes[2] = nnz * (ib + 4) + Rl * (4 * (2 + pu) + ab);
    bytes[3] = Rl * (9 + pu) * 4;
    return hipGetLastError();
}

}  // namespace mps
