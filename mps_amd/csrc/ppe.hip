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

#include <algorithm>
#include <vector>

#include "mps_internal.hpp"

namespace mps {

static inline dim3 grid_for(int n) { return dim3((unsigned) ((n + kBlock - 1) / kBlock)); }

// blockIdx -> logical block so that each of the 8 XCDs (dispatched round
// robin) works on one contiguous range of rows and keeps its halo in its L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b & 7, idx = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

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

__global__ void __launch_bounds__(kBlock) k_rows(int n, const int* __restrict__ inner_flag,
                                                 const int* __restrict__ row_of, const int* __restrict__ nnz_s,
                                                 int* __restrict__ row_slot, int* __restrict__ slice_w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || !inner_flag[s]) return;
    const int r = row_of[s];
    row_slot[r] = s;
    atomicMax(&slice_w[r >> 6], nnz_s[s]);
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_fill_sell(GridView g, Consts c, Hot h, const int* __restrict__ rows_p,
                                                      const int* __restrict__ row_slot, const int* __restrict__ row_of,
                                                      const int8_t* __restrict__ bc_s,
                                                      const double* __restrict__ diag_s,
                                                      const int* __restrict__ slice_pos, int* __restrict__ col,
                                                      double* __restrict__ val64, float* __restrict__ val32,
                                                      double* __restrict__ sdiag, double* __restrict__ dgl,
                                                      const int* __restrict__ raw_off, int* __restrict__ raw_col,
                                                      double* __restrict__ raw_val, double coefA) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= *rows_p) return;
    const int s = row_slot[r];
    const double di = diag_s[s];
    const double si = 1.0 / sqrt(di);
    sdiag[r] = si;
    dgl[r] = di;
    const int sl = r >> 6, lane = r & 63;
    const size_t base = (size_t) slice_pos[sl] * 64;
    const int width = slice_pos[sl + 1] - slice_pos[sl];
    int k = 0;
    int ro = raw_off ? raw_off[s] : 0;
    const double re = c.re_p;
    const double zi = DIM == 3 ? h.pz[s] : 0.0;
    for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                       [&](int q, double dist, double, double, double) {
                           if (bc_s[q] != MPS_BC_INNER) return;
                           // the off-diagonal value of src/mps.cpp:343-348
                           const double a = coefA * wfun(dist, re) / c.rho;
                           const double v = -a;
                           const double sj = 1.0 / sqrt(diag_s[q]);
                           const double sv = v * (si * sj);
                           const size_t idx = base + (size_t) k * 64 + lane;
                           col[idx] = row_of[q];
                           val64[idx] = sv;
                           if (val32) val32[idx] = (float) sv;
                           if (raw_col) {
                               raw_col[ro] = q;
                               raw_val[ro] = v;
                               ro++;
                           }
                           k++;
                       });
    for (; k < width; k++) {
        const size_t idx = base + (size_t) k * 64 + lane;
        col[idx] = r;
        val64[idx] = 0.0;
        if (val32) val32[idx] = 0.0f;
    }
}

hipError_t assemble_sell(int dim, const GridView& g, int n, const Consts& c, const Hot& h, const StepScratch& w,
                         Sell& A, int want_val32, int want_raw, int* err, int* host_rows, size_t* host_entries,
                         hipStream_t st) {
    (void) err;
    hipError_t e;
    const int nsl_max = (n + 63) / 64;
    // compact Inner rows in slot order
    e = scan_exclusive(w.inner_flag, w.row_of, n, w.scan_tmp, st);
    if (e != hipSuccess) return e;
    e = fill_int(A.slice_w, 0, (size_t) nsl_max + 1, st);
    if (e != hipSuccess) return e;
    if (n > 0)
        hipLaunchKernelGGL(k_rows, grid_for(n), dim3(kBlock), 0, st, n, w.inner_flag, w.row_of, w.nnz_s, A.row_slot,
                           A.slice_w);
    e = scan_exclusive(A.slice_w, A.slice_off, nsl_max, w.scan_tmp, st);
    if (e != hipSuccess) return e;
    int hv[2];
    e = hipMemcpyAsync(&hv[0], w.row_of + n, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(&hv[1], A.slice_off + nsl_max, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    const int rows = hv[0];
    const size_t entries = (size_t) hv[1] * 64;
    *host_rows = rows;
    *host_entries = entries;
    A.rows = rows;
    A.rows_dev = w.row_of + n;
    A.nslices = (rows + 63) / 64;
    if (entries > A.cap) {
        size_t cap = entries + entries / 4 + 64 * 64;
        hipFree(A.col);
        hipFree(A.val64);
        hipFree(A.val32);
        A.col = nullptr;
        A.val64 = nullptr;
        A.val32 = nullptr;
        A.cap = 0;
        if ((e = hipMalloc(&A.col, cap * sizeof(int))) != hipSuccess) return e;
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
    if (rows == 0) return hipSuccess;
    const double D = (double) c.dim;
    const double coefA = (2.0 * D / (c.lam_p * c.n0_p));
    const int* rows_p = w.row_of + n;
    if (dim == 2)
        hipLaunchKernelGGL(k_fill_sell<2>, grid_for(rows), dim3(kBlock), 0, st, g, c, h, rows_p, A.row_slot, w.row_of,
                           w.bc_s, w.diag_s, A.slice_off, A.col, A.val64, want_val32 ? A.val32 : nullptr, A.sdiag,
                           A.dgl, want_raw ? A.raw_off : nullptr, want_raw ? A.raw_col : nullptr,
                           want_raw ? A.raw_val : nullptr, coefA);
    else
        hipLaunchKernelGGL(k_fill_sell<3>, grid_for(rows), dim3(kBlock), 0, st, g, c, h, rows_p, A.row_slot, w.row_of,
                           w.bc_s, w.diag_s, A.slice_off, A.col, A.val64, want_val32 ? A.val32 : nullptr, A.sdiag,
                           A.dgl, want_raw ? A.raw_off : nullptr, want_raw ? A.raw_col : nullptr,
                           want_raw ? A.raw_val : nullptr, coefA);
    return hipGetLastError();
}

// ---------------- CG ----------------
//
// Every CG kernel launches the same grid of nb <= kMaxBlocks blocks; block b
// owns a contiguous range of rows (a multiple of 64, so whole SELL slices)
// and writes its partial sums to part[b * NV + k].  The next kernel starts by
// reducing all nb partials in one fixed order, redundantly in every block:
// the kernel boundary makes the partials visible on every XCD, no fence or
// atomic is needed, and every block derives the same scalars bit for bit.

constexpr int kMaxBlocks = 1024;

static int cg_blocks(int rows) {
    int nb = (rows + kBlock - 1) / kBlock;
    if (nb > kMaxBlocks) nb = kMaxBlocks;
    return nb < 1 ? 1 : nb;
}

struct CgScalars {
    double gam[2];     // r.r by iteration parity
    double thr;        // stop when rho2 <= thr
    double bnorm2;     // ||b||^2
    double rho2;       // latest residual^2 (unscaled sum a_ii r_i^2, or r.r for inner solves)
    double init_rho2;  // residual^2 at the start of an inner solve
    double red[2];     // results of k_reduce_pairs
    int iter;
    int max_iter;
    int done;
    int converged;
    int first;
    int frozen;        // timing mode: kernels leave every scalar as it is
};

// Fixed-order sum of nv interleaved partials (part[b*nv + k], b < nb) in every
// block; result broadcast to all threads.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nb, double (&out)[NV],
                                                double* red) {
    for (int k = 0; k < NV; k++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += part[(size_t) NV * b + k];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        double t = 0.0;
        for (int w = 0; w < (int) (blockDim.x >> 6); w++) t += red[w];
        out[k] = t;
        __syncthreads();
    }
}

template <int NV>
__device__ __forceinline__ void write_partials(const double (&v)[NV], double* __restrict__ part, double* red) {
    for (int k = 0; k < NV; k++) {
        double t = block_sum(v[k], red);
        if (threadIdx.x == 0) part[(size_t) NV * blockIdx.x + k] = t;
    }
}

// Row range of this block: contiguous, whole slices, XCD-grouped.
__device__ __forceinline__ void block_rows(int rows, int& r0, int& r1) {
    const int nb = gridDim.x;
    const int per = ((rows + nb - 1) / nb + 63) & ~63;
    const int lb = xcd_block(blockIdx.x, nb);
    r0 = lb * per;
    r1 = r0 + per < rows ? r0 + per : rows;
}

template <typename T>
__device__ __forceinline__ T pval(const T* __restrict__ r, const T* __restrict__ pold, T beta, int j, bool first) {
    return first ? r[j] : r[j] + beta * pold[j];
}

// spmv: p = r + beta p_old (own row and each neighbour's, on the fly),
// q = Ahat p, partial p.q.  Prologue: r.r and rho2 of the previous update.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(const int* __restrict__ rows_p, const int* __restrict__ slice_pos,
                                                    const int* __restrict__ col, const T* __restrict__ val,
                                                    const T* __restrict__ r, const T* __restrict__ pold,
                                                    T* __restrict__ pnew, T* __restrict__ q,
                                                    const double* __restrict__ part_u, double* __restrict__ part_s,
                                                    CgScalars* sc, int par) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    // block-uniform view of the done flag: block 0 of this very launch may
    // raise it while other blocks start
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = *(volatile int*) &sc->done;
    __syncthreads();
    if (s_done) return;
    double gr[2];
    reduce_partials<2>(part_u, gridDim.x, gr, red);
    const double g = gr[0], rho = gr[1];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const bool first = sc->first != 0;
    if (!sc->frozen) {
        const bool conv = rho <= sc->thr;
        if (conv || sc->iter >= sc->max_iter || !(rho == rho)) {
            if (lead) {
                sc->done = 1;
                sc->converged = conv ? 1 : 0;
                sc->rho2 = rho;
            }
            return;
        }
        if (lead) {
            sc->gam[par] = g;
            sc->rho2 = rho;
        }
    }
    const T beta = (T) (first ? 0.0 : (sc->frozen ? 0.5 : g / sc->gam[par ^ 1]));
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    double pq = 0.0;
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) {
        const T pr = pval(r, pold, beta, row, first);
        T acc = pr;
        const int sl = row >> 6, lane = row & 63;
        const int s0 = slice_pos[sl], w = slice_pos[sl + 1] - s0;
        const size_t base = (size_t) s0 * 64 + lane;
        for (int k = 0; k < w; k++) {
            const size_t idx = base + (size_t) k * 64;
            acc += val[idx] * pval(r, pold, beta, col[idx], first);
        }
        pnew[row] = pr;
        q[row] = acc;
        pq += (double) pr * (double) acc;
    }
    double v[1] = {pq};
    write_partials<1>(v, part_s, red);
}

// update: y += alpha p, r -= alpha q; partials r.r and rho2 (sum a_ii r^2 when
// WEIGHTED).  Prologue: p.q of the spmv.
template <typename T, bool WEIGHTED>
__global__ void __launch_bounds__(kBlock) k_cg_update(const int* __restrict__ rows_p, T* __restrict__ y,
                                                      T* __restrict__ r, const T* __restrict__ p,
                                                      const T* __restrict__ q, const double* __restrict__ dgl,
                                                      const double* __restrict__ part_s, double* __restrict__ part_u,
                                                      CgScalars* sc, int par) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = *(volatile int*) &sc->done;
    __syncthreads();
    if (s_done) return;
    double d[1];
    reduce_partials<1>(part_s, gridDim.x, d, red);
    const T alpha = (T) (sc->frozen ? 0.0 : sc->gam[par] / d[0]);
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    double rr = 0.0, rw = 0.0;
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) {
        y[row] += alpha * p[row];
        const T rn = r[row] - alpha * q[row];
        r[row] = rn;
        const double r2 = (double) rn * (double) rn;
        rr += r2;
        rw += WEIGHTED ? dgl[row] * r2 : r2;
    }
    double v[2] = {rr, rw};
    write_partials<2>(v, part_u, red);
    if (blockIdx.x == 0 && threadIdx.x == 0 && !sc->frozen) {
        sc->iter += 1;
        sc->first = 0;
    }
}

// One block: sums nv-interleaved partials into sc->red[0..nv-1].
__global__ void __launch_bounds__(kBlock) k_reduce_pairs(const double* __restrict__ part, int nb, int nv,
                                                         CgScalars* sc) {
    __shared__ double red[kBlock / 64];
    double o2[2];
    if (nv == 2) {
        reduce_partials<2>(part, nb, o2, red);
    } else {
        double o1[1];
        reduce_partials<1>(part, nb, o1, red);
        o2[0] = o1[0];
        o2[1] = 0.0;
    }
    if (threadIdx.x == 0) {
        sc->red[0] = o2[0];
        sc->red[1] = o2[1];
    }
}

// FP64 init: yhat = x0 sqrt(d), bhat = S b; partial ||b||^2 (and 0).
__global__ void __launch_bounds__(kBlock) k_cg_init(const int* __restrict__ rows_p, const int* __restrict__ row_slot,
                                                    const int32_t* __restrict__ id, const double* __restrict__ b_s,
                                                    const double* __restrict__ pprev, const double* __restrict__ sdiag,
                                                    double* __restrict__ y, double* __restrict__ bh,
                                                    double* __restrict__ part) {
    __shared__ double red[kBlock / 64];
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    double bb = 0.0;
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) {
        const int s = row_slot[row];
        const double b = b_s[s];
        const double si = sdiag[row];
        bh[row] = b * si;
        y[row] = pprev ? pprev[id[s]] / si : 0.0;
        bb += b * b;
    }
    double v[2] = {bb, 0.0};
    write_partials<2>(v, part, red);
}

// FP64 residual r = bhat - Ahat y (unit diagonal); optionally zeroes p;
// partials r.r and sum d r^2.
__global__ void __launch_bounds__(kBlock) k_residual64(const int* __restrict__ rows_p, const int* __restrict__ slice_pos,
                                                       const int* __restrict__ col, const double* __restrict__ val,
                                                       const double* __restrict__ y, const double* __restrict__ bh,
                                                       const double* __restrict__ dgl, double* __restrict__ r,
                                                       double* __restrict__ p0, double* __restrict__ part) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    double rr = 0.0, rw = 0.0;
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) {
        double acc = y[row];
        const int sl = row >> 6, lane = row & 63;
        const int s0 = slice_pos[sl], w = slice_pos[sl + 1] - s0;
        const size_t base = (size_t) s0 * 64 + lane;
        for (int k = 0; k < w; k++) {
            const size_t idx = base + (size_t) k * 64;
            acc += val[idx] * y[col[idx]];
        }
        const double rv = bh[row] - acc;
        r[row] = rv;
        if (p0) p0[row] = 0.0;
        rr += rv * rv;
        rw += dgl[row] * (rv * rv);
    }
    double v[2] = {rr, rw};
    write_partials<2>(v, part, red);
}

// Inner solve setup: r32 = r64 (scaled residual), e = 0, p = 0; partials r.r twice.
__global__ void __launch_bounds__(kBlock) k_inner_init(const int* __restrict__ rows_p, const double* __restrict__ r64,
                                                       float* __restrict__ r32, float* __restrict__ e32,
                                                       float* __restrict__ p32, double* __restrict__ part) {
    __shared__ double red[kBlock / 64];
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    double rr = 0.0;
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) {
        const float rv = (float) r64[row];
        r32[row] = rv;
        e32[row] = 0.0f;
        p32[row] = 0.0f;
        rr += (double) rv * (double) rv;
    }
    double v[2] = {rr, rr};
    write_partials<2>(v, part, red);
}

__global__ void __launch_bounds__(kBlock) k_add_correction(const int* __restrict__ rows_p, double* __restrict__ y,
                                                           const float* __restrict__ e) {
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) y[row] += (double) e[row];
}

__global__ void __launch_bounds__(kBlock) k_cg_finish(const int* __restrict__ rows_p, const int* __restrict__ row_slot,
                                                      const double* __restrict__ y, const double* __restrict__ sdiag,
                                                      double* __restrict__ x_s) {
    int r0, r1;
    block_rows(*rows_p, r0, r1);
    for (int row = r0 + (int) threadIdx.x; row < r1; row += kBlock) x_s[row_slot[row]] = y[row] * sdiag[row];
}

struct CgWork {
    int cap = 0;
    int nb = 1;               // blocks of every CG launch (fixed for the capacity)
    double* y = nullptr;      // FP64 solution (scaled)
    double* bh = nullptr;     // scaled rhs
    double* r64 = nullptr;    // FP64 residual (scaled)
    double* p64[2] = {nullptr, nullptr};
    double* q64 = nullptr;
    float* r32 = nullptr;
    float* p32[2] = {nullptr, nullptr};
    float* q32 = nullptr;
    float* e32 = nullptr;     // inner correction
    double* part_s = nullptr; // spmv partials (1 per block)
    double* part_u = nullptr; // update / residual partials (2 per block)
    CgScalars* sc = nullptr;  // device scalars
    CgScalars* host = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // chunk graphs: FP64 CG, FP32 inner CG
    int graph_chunk[2] = {0, 0};
    const void* graph_key[2] = {nullptr, nullptr};
};

CgWork* cg_create() { return new CgWork(); }

static void cg_free_vectors(CgWork* cw) {
    void* ptrs[] = {cw->y, cw->bh, cw->r64, cw->p64[0], cw->p64[1], cw->q64, cw->r32, cw->p32[0], cw->p32[1], cw->q32,
                    cw->e32, cw->part_s, cw->part_u};
    for (void* p : ptrs) hipFree(p);
}

void cg_destroy(CgWork* cw) {
    if (!cw) return;
    cg_free_vectors(cw);
    hipFree(cw->sc);
    hipHostFree(cw->host);
    for (int k = 0; k < 2; k++)
        if (cw->graph[k]) hipGraphExecDestroy(cw->graph[k]);
    delete cw;
}

hipError_t cg_reserve(CgWork* cw, int rows) {
    if (rows <= cw->cap && cw->sc) return hipSuccess;
    const int cap = rows + rows / 8 + 1024;
    cg_free_vectors(cw);
    hipError_t e;
    const size_t d = (size_t) cap * sizeof(double), f = (size_t) cap * sizeof(float);
    if ((e = hipMalloc(&cw->y, d)) || (e = hipMalloc(&cw->bh, d)) || (e = hipMalloc(&cw->r64, d)) ||
        (e = hipMalloc(&cw->p64[0], d)) || (e = hipMalloc(&cw->p64[1], d)) || (e = hipMalloc(&cw->q64, d)) ||
        (e = hipMalloc(&cw->r32, f)) || (e = hipMalloc(&cw->p32[0], f)) || (e = hipMalloc(&cw->p32[1], f)) ||
        (e = hipMalloc(&cw->q32, f)) || (e = hipMalloc(&cw->e32, f)))
        return e;
    cw->nb = cg_blocks(cap);
    if ((e = hipMalloc(&cw->part_s, sizeof(double) * (size_t) cw->nb))) return e;
    if ((e = hipMalloc(&cw->part_u, sizeof(double) * 2 * (size_t) cw->nb))) return e;
    if (!cw->sc) {
        if ((e = hipMalloc(&cw->sc, sizeof(CgScalars)))) return e;
        if ((e = hipHostMalloc(&cw->host, sizeof(CgScalars), hipHostMallocDefault))) return e;
    }
    cw->cap = cap;
    for (int k = 0; k < 2; k++) {
        if (cw->graph[k]) hipGraphExecDestroy(cw->graph[k]);
        cw->graph[k] = nullptr;
    }
    return hipSuccess;
}

// Enqueue `chunk` CG iterations (two kernels each); chunk is even, so the p
// buffer and scalar parities line up from one chunk to the next.
template <typename T, bool WEIGHTED>
static void enqueue_chunk(CgWork* cw, const Sell& A, const T* val, T* y, T* r, T* const* pbuf, T* q, int chunk,
                          hipStream_t st) {
    const dim3 gr(cw->nb), bl(kBlock);
    for (int k = 0; k < chunk; k++) {
        const int par = k & 1;
        T* pold = pbuf[par];
        T* pnew = pbuf[par ^ 1];
        hipLaunchKernelGGL((k_cg_spmv<T>), gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, val, r, pold, pnew, q,
                           cw->part_u, cw->part_s, cw->sc, par);
        hipLaunchKernelGGL((k_cg_update<T, WEIGHTED>), gr, bl, 0, st, A.rows_dev, y, r, pnew, q, A.dgl, cw->part_s,
                           cw->part_u, cw->sc, par);
    }
}

template <typename T, bool WEIGHTED>
static hipError_t run_cg_loop(CgWork* cw, int which, const Sell& A, const T* val, T* y, T* r, T* const* pbuf, T* q,
                              const mps_solver_options& o, hipStream_t st) {
    const int chunk = std::max(2, (o.check_interval + 1) & ~1);
    hipError_t e;
    for (;;) {
        if (o.use_graph) {
            if (!cw->graph[which] || cw->graph_chunk[which] != chunk || cw->graph_key[which] != (const void*) val) {
                if (cw->graph[which]) hipGraphExecDestroy(cw->graph[which]);
                cw->graph[which] = nullptr;
                hipGraph_t gph;
                if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) return e;
                enqueue_chunk<T, WEIGHTED>(cw, A, val, y, r, pbuf, q, chunk, st);
                if ((e = hipStreamEndCapture(st, &gph))) return e;
                e = hipGraphInstantiate(&cw->graph[which], gph, nullptr, nullptr, 0);
                hipGraphDestroy(gph);
                if (e) return e;
                cw->graph_chunk[which] = chunk;
                cw->graph_key[which] = val;
            }
            if ((e = hipGraphLaunch(cw->graph[which], st))) return e;
        } else {
            enqueue_chunk<T, WEIGHTED>(cw, A, val, y, r, pbuf, q, chunk, st);
        }
        if ((e = hipMemcpyAsync(cw->host, cw->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st))) return e;
        if ((e = hipStreamSynchronize(st))) return e;
        if (cw->host->done) return hipSuccess;
    }
}

static hipError_t set_scalars(CgWork* cw, const CgScalars& s, hipStream_t st) {
    *cw->host = s;
    return hipMemcpyAsync(cw->sc, cw->host, sizeof(CgScalars), hipMemcpyHostToDevice, st);
}

static hipError_t read_scalars(CgWork* cw, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(cw->host, cw->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st);
    if (e) return e;
    return hipStreamSynchronize(st);
}

// Reduces nv-interleaved partials into sc->red and reads the scalars back.
static hipError_t reduce_and_read(CgWork* cw, const double* part, int nv, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(kBlock), 0, st, part, cw->nb, nv, cw->sc);
    return read_scalars(cw, st);
}

hipError_t cg_solve(CgWork* cw, const Sell& A, const Hot& h, int n, const double* b_s, const double* pprev_id,
                    double* x_s, const mps_solver_options& o, mps_solver_stats& stats, hipStream_t st,
                    std::string& msg) {
    (void) n;
    const int rows = A.rows;
    stats = mps_solver_stats{};
    stats.rows = rows;
    stats.converged = 1;
    if (rows == 0) return hipSuccess;
    hipError_t e = cg_reserve(cw, rows);
    if (e) return e;
    const dim3 gr(cw->nb), bl(kBlock);
    const int max_it = o.max_iterations > 0 ? o.max_iterations : 2 * rows;
    const double tol = o.tolerance;

    CgScalars s0{};
    if ((e = set_scalars(cw, s0, st))) return e;
    hipLaunchKernelGGL(k_cg_init, gr, bl, 0, st, A.rows_dev, A.row_slot, h.id, b_s, o.warm_start ? pprev_id : nullptr,
                       A.sdiag, cw->y, cw->bh, cw->part_u);
    if ((e = reduce_and_read(cw, cw->part_u, 2, st))) return e;
    const double bn2 = cw->host->red[0];
    stats.rhs_norm = sqrt(bn2);
    if (bn2 == 0.0) {
        // Eigen: a zero right-hand side gives x = 0 (no iteration)
        if ((e = fill_double(cw->y, 0.0, (size_t) rows, st))) return e;
        hipLaunchKernelGGL(k_cg_finish, gr, bl, 0, st, A.rows_dev, A.row_slot, cw->y, A.sdiag, x_s);
        stats.rel_residual = 0.0;
        return hipGetLastError();
    }
    const double thr = tol * tol * bn2;
    hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, A.val64, cw->y, cw->bh, A.dgl,
                       cw->r64, cw->p64[0], cw->part_u);
    if ((e = reduce_and_read(cw, cw->part_u, 2, st))) return e;
    double rho2 = cw->host->red[1];
    int total_it = 0, refinements = 0;
    bool conv = rho2 <= thr;

    if (o.precision == MPS_PREC_FP64) {
        if (!conv) {
            // part_u already holds the initial r.r and rho2 for the first spmv
            CgScalars s{};
            s.thr = thr;
            s.bnorm2 = bn2;
            s.max_iter = max_it;
            s.first = 1;
            if ((e = set_scalars(cw, s, st))) return e;
            if ((e = run_cg_loop<double, true>(cw, 0, A, A.val64, cw->y, cw->r64, cw->p64, cw->q64, o, st))) return e;
            total_it = cw->host->iter;
            conv = cw->host->converged != 0;
        }
    } else {
        const int max_ref = o.max_refinements > 0 ? o.max_refinements : 20;
        const double eta = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-4;
        while (!conv && refinements < max_ref && total_it < max_it) {
            hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, A.rows_dev, cw->r64, cw->r32, cw->e32, cw->p32[0],
                               cw->part_u);
            if ((e = reduce_and_read(cw, cw->part_u, 2, st))) return e;
            const double init = cw->host->red[0];
            // reduce the inner residual by eta, or by what the outer loop
            // still needs when that is less
            const double need = thr / rho2;
            double red2 = eta * eta;
            if (need > red2) red2 = need * 0.25;
            CgScalars s{};
            s.thr = init * red2;
            s.init_rho2 = init;
            s.max_iter = max_it - total_it;
            s.first = 1;
            s.done = init == 0.0 ? 1 : 0;
            if ((e = set_scalars(cw, s, st))) return e;
            if (!s.done) {
                if ((e = run_cg_loop<float, false>(cw, 1, A, A.val32, cw->e32, cw->r32, cw->p32, cw->q32, o, st)))
                    return e;
                total_it += cw->host->iter;
            }
            hipLaunchKernelGGL(k_add_correction, gr, bl, 0, st, A.rows_dev, cw->y, cw->e32);
            hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, A.val64, cw->y, cw->bh,
                               A.dgl, cw->r64, nullptr, cw->part_u);
            if ((e = reduce_and_read(cw, cw->part_u, 2, st))) return e;
            rho2 = cw->host->red[1];
            conv = rho2 <= thr;
            refinements++;
        }
    }
    // true FP64 residual of the final iterate
    hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, A.val64, cw->y, cw->bh, A.dgl,
                       cw->r64, nullptr, cw->part_u);
    if ((e = reduce_and_read(cw, cw->part_u, 2, st))) return e;
    stats.iterations = total_it;
    stats.refinements = refinements;
    stats.rel_residual = sqrt(cw->host->red[1] / bn2);
    stats.converged = conv ? 1 : 0;
    hipLaunchKernelGGL(k_cg_finish, gr, bl, 0, st, A.rows_dev, A.row_slot, cw->y, A.sdiag, x_s);
    if (!conv) msg = "pressure solve stopped above its tolerance";
    return hipGetLastError();
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
    const dim3 gr(cw->nb), bl(kBlock);
    // finite vector contents: r = 1e-3, p = 0, y = 0, and their FP32 copies
    if ((e = fill_double(cw->r64, 1e-3, (size_t) rows, st))) return e;
    if ((e = fill_double(cw->p64[0], 0.0, (size_t) rows, st))) return e;
    if ((e = fill_double(cw->y, 0.0, (size_t) rows, st))) return e;
    hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, A.rows_dev, cw->r64, cw->r32, cw->e32, cw->p32[0], cw->part_u);
    if ((e = fill_double(cw->part_s, 1.0, (size_t) cw->nb, st))) return e;
    CgScalars s{};
    s.frozen = 1;  // alpha = 0: y and r stay as they are; p = r + p/2 stays bounded
    s.max_iter = 0x7fffffff;
    s.thr = -1.0;
    s.gam[0] = s.gam[1] = 1.0;
    if ((e = set_scalars(cw, s, st))) return e;
    hipEvent_t a, b;
    if ((e = hipEventCreate(&a))) return e;
    if ((e = hipEventCreate(&b))) return e;
    for (int kind = 0; kind < 4; kind++) {
        // one untimed launch, then reps timed launches
        for (int rep = -1; rep < reps; rep++) {
            if (rep == 0) hipEventRecord(a, st);
            switch (kind) {
                case 0:
                    hipLaunchKernelGGL((k_cg_spmv<double>), gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, A.val64,
                                       cw->r64, cw->p64[0], cw->p64[1], cw->q64, cw->part_u, cw->part_s, cw->sc, 0);
                    break;
                case 1:
                    hipLaunchKernelGGL((k_cg_update<double, true>), gr, bl, 0, st, A.rows_dev, cw->y, cw->r64,
                                       cw->p64[1], cw->q64, A.dgl, cw->part_s, cw->part_u, cw->sc, 0);
                    break;
                case 2:
                    hipLaunchKernelGGL((k_cg_spmv<float>), gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, A.val32,
                                       cw->r32, cw->p32[0], cw->p32[1], cw->q32, cw->part_u, cw->part_s, cw->sc, 0);
                    break;
                default:
                    hipLaunchKernelGGL((k_cg_update<float, false>), gr, bl, 0, st, A.rows_dev, cw->e32, cw->r32,
                                       cw->p32[1], cw->q32, A.dgl, cw->part_s, cw->part_u, cw->sc, 0);
                    break;
            }
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
    const long long R = rows;
    bytes[0] = R * 4 * 8 + nnz * (4 + 8);
    bytes[1] = R * (6 * 8 + 8);
    bytes[2] = R * 4 * 4 + nnz * (4 + 4);
    bytes[3] = R * (6 * 4);
    return hipGetLastError();
}

}  // namespace mps
