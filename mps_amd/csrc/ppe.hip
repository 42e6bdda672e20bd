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

struct CgScalars {
    double gamma;     // r.r (scaled system)
    double alpha;
    double beta;
    double rho2;      // sum a_ii r_i^2 = ||b - A x||^2 (unscaled), or r.r for inner solves
    double thr;       // stop when rho2 <= thr
    double bnorm2;    // ||b||^2
    double init_rho2; // residual^2 at the start of the solve
    int iter;
    int max_iter;
    int done;
    int converged;
    unsigned int ticket;
    int pad;
};

struct CgWork {
    int cap = 0;
    int nb_cap = 0;
    double* y = nullptr;      // FP64 solution (scaled)
    double* bh = nullptr;     // scaled rhs
    double* r64 = nullptr;    // FP64 residual (scaled)
    double* p64[2] = {nullptr, nullptr};
    double* q64 = nullptr;
    float* r32 = nullptr;
    float* p32[2] = {nullptr, nullptr};
    float* q32 = nullptr;
    float* e32 = nullptr;     // inner correction
    double* part = nullptr;   // block partials, 3 per block
    CgScalars* sc = nullptr;  // device scalars
    CgScalars* host = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // chunk graphs: FP64 CG, FP32 inner CG
    int graph_rows[2] = {-1, -1};
    int graph_chunk[2] = {0, 0};
    const void* graph_key[2] = {nullptr, nullptr};
};

CgWork* cg_create() { return new CgWork(); }

void cg_destroy(CgWork* cw) {
    if (!cw) return;
    hipFree(cw->y);
    hipFree(cw->bh);
    hipFree(cw->r64);
    hipFree(cw->p64[0]);
    hipFree(cw->p64[1]);
    hipFree(cw->q64);
    hipFree(cw->r32);
    hipFree(cw->p32[0]);
    hipFree(cw->p32[1]);
    hipFree(cw->q32);
    hipFree(cw->e32);
    hipFree(cw->part);
    hipFree(cw->sc);
    hipHostFree(cw->host);
    for (int k = 0; k < 2; k++)
        if (cw->graph[k]) hipGraphExecDestroy(cw->graph[k]);
    delete cw;
}

hipError_t cg_reserve(CgWork* cw, int rows) {
    if (rows <= cw->cap && cw->sc) return hipSuccess;
    int cap = rows + rows / 8 + 1024;
    hipFree(cw->y);
    hipFree(cw->bh);
    hipFree(cw->r64);
    hipFree(cw->p64[0]);
    hipFree(cw->p64[1]);
    hipFree(cw->q64);
    hipFree(cw->r32);
    hipFree(cw->p32[0]);
    hipFree(cw->p32[1]);
    hipFree(cw->q32);
    hipFree(cw->e32);
    hipFree(cw->part);
    hipError_t e;
    size_t d = (size_t) cap * sizeof(double), f = (size_t) cap * sizeof(float);
    if ((e = hipMalloc(&cw->y, d)) || (e = hipMalloc(&cw->bh, d)) || (e = hipMalloc(&cw->r64, d)) ||
        (e = hipMalloc(&cw->p64[0], d)) || (e = hipMalloc(&cw->p64[1], d)) || (e = hipMalloc(&cw->q64, d)) ||
        (e = hipMalloc(&cw->r32, f)) || (e = hipMalloc(&cw->p32[0], f)) || (e = hipMalloc(&cw->p32[1], f)) ||
        (e = hipMalloc(&cw->q32, f)) || (e = hipMalloc(&cw->e32, f)))
        return e;
    cw->nb_cap = (cap + kBlock - 1) / kBlock + 8;
    if ((e = hipMalloc(&cw->part, sizeof(double) * 3 * (size_t) cw->nb_cap))) return e;
    if (!cw->sc) {
        if ((e = hipMalloc(&cw->sc, sizeof(CgScalars)))) return e;
        if ((e = hipHostMalloc(&cw->host, sizeof(CgScalars), hipHostMallocDefault))) return e;
    }
    cw->cap = cap;
    for (int k = 0; k < 2; k++) {
        if (cw->graph[k]) hipGraphExecDestroy(cw->graph[k]);
        cw->graph[k] = nullptr;
        cw->graph_rows[k] = -1;
    }
    return hipSuccess;
}

// Last-block reduction helper: each block stores nv partial sums; returns true
// in the block that finishes last, after which out[k] holds the ordered sums.
template <int NV>
__device__ __forceinline__ bool reduce_last(const double (&v)[NV], double* part, unsigned int* ticket, double (&out)[NV],
                                            double* red, int nb) {
    __shared__ int am_last;
    for (int k = 0; k < NV; k++) {
        double t = block_sum(v[k], red);
        if (threadIdx.x == 0) part[(size_t) NV * blockIdx.x + k] = t;
    }
    __threadfence();
    if (threadIdx.x == 0) am_last = atomicAdd(ticket, 1u) == (unsigned) (nb - 1);
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    volatile double* vp = part;
    for (int k = 0; k < NV; k++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += vp[(size_t) NV * b + k];
        double t = block_sum(acc, red);
        if (threadIdx.x == 0) out[k] = t;
    }
    return true;
}

// SpMV row: q_r = p_r + sum_k Ahat[r,k] p_col, p from (r, pold, beta) on the fly.
template <typename T>
__device__ __forceinline__ T pval(const T* __restrict__ r, const T* __restrict__ pold, T beta, int j, bool first) {
    return first ? r[j] : r[j] + beta * pold[j];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(const int* __restrict__ rows_p, const int* __restrict__ slice_pos,
                                                    const int* __restrict__ col, const T* __restrict__ val,
                                                    const T* __restrict__ r, const T* __restrict__ pold,
                                                    T* __restrict__ pnew, T* __restrict__ q, double* part,
                                                    CgScalars* sc) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    if (*(volatile int*) &sc->done) return;
    const int rows = *rows_p;
    const int nb = gridDim.x;
    const int lb = xcd_block(blockIdx.x, nb);
    const int row = lb * kBlock + threadIdx.x;
    const bool first = sc->iter == 0;
    const T beta = (T) sc->beta;
    double pq = 0.0;
    if (row < rows) {
        const T pr = pval(r, pold, beta, row, first);
        T acc = pr;
        const int sl = row >> 6, lane = row & 63;
        const int p0 = slice_pos[sl], w = slice_pos[sl + 1] - p0;
        const size_t base = (size_t) p0 * 64 + lane;
        for (int k = 0; k < w; k++) {
            const size_t idx = base + (size_t) k * 64;
            const int j = col[idx];
            acc += val[idx] * pval(r, pold, beta, j, first);
        }
        pnew[row] = pr;
        q[row] = acc;
        pq = (double) pr * (double) acc;
    }
    double v[1] = {pq}, out[1];
    if (reduce_last<1>(v, part, &sc->ticket, out, red, nb)) {
        if (threadIdx.x == 0) {
            sc->alpha = sc->gamma / out[0];
            sc->ticket = 0;
        }
    }
}

template <typename T, bool WEIGHTED>
__global__ void __launch_bounds__(kBlock) k_cg_update(const int* __restrict__ rows_p, T* __restrict__ y, T* __restrict__ r,
                                                      const T* __restrict__ p, const T* __restrict__ q,
                                                      const double* __restrict__ dgl, double* part, CgScalars* sc) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    if (*(volatile int*) &sc->done) return;
    const int rows = *rows_p;
    const int nb = gridDim.x;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const T alpha = (T) sc->alpha;
    double rr = 0.0, rw = 0.0;
    if (row < rows) {
        y[row] += alpha * p[row];
        const T rn = r[row] - alpha * q[row];
        r[row] = rn;
        rr = (double) rn * (double) rn;
        rw = WEIGHTED ? dgl[row] * rr : rr;
    }
    double v[2] = {rr, rw}, out[2];
    if (reduce_last<2>(v, part, &sc->ticket, out, red, nb)) {
        if (threadIdx.x == 0) {
            sc->beta = out[0] / sc->gamma;
            sc->gamma = out[0];
            sc->rho2 = out[1];
            sc->iter += 1;
            if (out[1] <= sc->thr) {
                sc->done = 1;
                sc->converged = 1;
            } else if (sc->iter >= sc->max_iter || !(out[1] == out[1])) {
                sc->done = 1;
            }
            sc->ticket = 0;
        }
    }
}

// FP64 init: yhat = x0 sqrt(d), bhat = S b; partial ||b||^2.
__global__ void __launch_bounds__(kBlock) k_cg_init(int rows, const int* __restrict__ row_slot,
                                                    const int32_t* __restrict__ id, const double* __restrict__ b_s,
                                                    const double* __restrict__ pprev, const double* __restrict__ sdiag,
                                                    double* __restrict__ y, double* __restrict__ bh, double* part,
                                                    CgScalars* sc) {
    __shared__ double red[kBlock / 64];
    const int nb = gridDim.x;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    double bb = 0.0;
    if (row < rows) {
        const int s = row_slot[row];
        const double b = b_s[s];
        const double si = sdiag[row];
        bh[row] = b * si;
        y[row] = pprev ? pprev[id[s]] / si : 0.0;
        bb = b * b;
    }
    double v[1] = {bb}, out[1];
    if (reduce_last<1>(v, part, &sc->ticket, out, red, nb)) {
        if (threadIdx.x == 0) {
            sc->bnorm2 = out[0];
            sc->ticket = 0;
        }
    }
}

// FP64 residual r = bhat - Ahat y (unit diagonal); sets p = 0 and, for the
// outer loop, sums gamma = r.r and rho2 = sum d r^2.
__global__ void __launch_bounds__(kBlock) k_residual64(int rows, const int* __restrict__ slice_pos,
                                                       const int* __restrict__ col, const double* __restrict__ val,
                                                       const double* __restrict__ y, const double* __restrict__ bh,
                                                       const double* __restrict__ dgl, double* __restrict__ r,
                                                       double* __restrict__ p0, double* part, CgScalars* sc) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    const int nb = gridDim.x;
    const int lb = xcd_block(blockIdx.x, nb);
    const int row = lb * kBlock + threadIdx.x;
    double rr = 0.0, rw = 0.0;
    if (row < rows) {
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
        rr = rv * rv;
        rw = dgl[row] * rr;
    }
    double v[2] = {rr, rw}, out[2];
    if (reduce_last<2>(v, part, &sc->ticket, out, red, nb)) {
        if (threadIdx.x == 0) {
            sc->gamma = out[0];
            sc->rho2 = out[1];
            sc->ticket = 0;
        }
    }
}

// Inner solve setup: r32 = r64 (scaled residual), e = 0, p = 0.
__global__ void __launch_bounds__(kBlock) k_inner_init(int rows, const double* __restrict__ r64, float* __restrict__ r32,
                                                       float* __restrict__ e32, float* __restrict__ p32,
                                                       double* part, CgScalars* sc) {
#pragma clang fp contract(fast)
    __shared__ double red[kBlock / 64];
    const int nb = gridDim.x;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    double rr = 0.0;
    if (row < rows) {
        const float rv = (float) r64[row];
        r32[row] = rv;
        e32[row] = 0.0f;
        p32[row] = 0.0f;
        rr = (double) rv * (double) rv;
    }
    double v[1] = {rr}, out[1];
    if (reduce_last<1>(v, part, &sc->ticket, out, red, nb)) {
        if (threadIdx.x == 0) {
            sc->gamma = out[0];
            sc->rho2 = out[0];
            sc->init_rho2 = out[0];
            sc->ticket = 0;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_add_correction(int rows, double* __restrict__ y, const float* __restrict__ e) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row < rows) y[row] += (double) e[row];
}

__global__ void __launch_bounds__(kBlock) k_cg_finish(int rows, const int* __restrict__ row_slot,
                                                      const double* __restrict__ y, const double* __restrict__ sdiag,
                                                      double* __restrict__ x_s) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row < rows) x_s[row_slot[row]] = y[row] * sdiag[row];
}

// Enqueue `chunk` CG iterations (two kernels each).  `par` is the parity of
// the first iteration's p buffer; chunk is even so parity is preserved.
template <typename T, bool WEIGHTED>
static void enqueue_chunk(int grid_rows, const Sell& A, const T* val, T* y, T* r, T* const* pbuf, T* q,
                          const double* dgl, double* part, CgScalars* sc, int chunk, hipStream_t st) {
    const dim3 gr = grid_for(grid_rows), bl(kBlock);
    for (int k = 0; k < chunk; k++) {
        T* pold = pbuf[k & 1];
        T* pnew = pbuf[(k + 1) & 1];
        hipLaunchKernelGGL((k_cg_spmv<T>), gr, bl, 0, st, A.rows_dev, A.slice_off, A.col, val, r, pold, pnew, q, part,
                           sc);
        hipLaunchKernelGGL((k_cg_update<T, WEIGHTED>), gr, bl, 0, st, A.rows_dev, y, r, pnew, q, dgl, part, sc);
    }
}

template <typename T, bool WEIGHTED>
static hipError_t run_cg_loop(CgWork* cw, int which, int rows, const Sell& A, const T* val, T* y, T* r, T* const* pbuf,
                              T* q, const mps_solver_options& o, hipStream_t st) {
    const int chunk = std::max(2, (o.check_interval + 1) & ~1);
    // launch over the whole capacity: kernels read the live row count, so a
    // captured chunk stays valid from step to step
    const int grid_rows = o.use_graph ? cw->cap : rows;
    hipError_t e;
    for (;;) {
        if (o.use_graph) {
            if (!cw->graph[which] || cw->graph_rows[which] != grid_rows || cw->graph_chunk[which] != chunk ||
                cw->graph_key[which] != (const void*) val) {
                if (cw->graph[which]) hipGraphExecDestroy(cw->graph[which]);
                cw->graph[which] = nullptr;
                hipGraph_t gph;
                if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) return e;
                enqueue_chunk<T, WEIGHTED>(grid_rows, A, val, y, r, pbuf, q, A.dgl, cw->part, cw->sc, chunk, st);
                if ((e = hipStreamEndCapture(st, &gph))) return e;
                e = hipGraphInstantiate(&cw->graph[which], gph, nullptr, nullptr, 0);
                hipGraphDestroy(gph);
                if (e) return e;
                cw->graph_rows[which] = grid_rows;
                cw->graph_chunk[which] = chunk;
                cw->graph_key[which] = val;
            }
            if ((e = hipGraphLaunch(cw->graph[which], st))) return e;
        } else {
            enqueue_chunk<T, WEIGHTED>(grid_rows, A, val, y, r, pbuf, q, A.dgl, cw->part, cw->sc, chunk, st);
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
    const dim3 gr = grid_for(rows), bl(kBlock);
    const int max_it = o.max_iterations > 0 ? o.max_iterations : 2 * rows;
    const double tol = o.tolerance;

    CgScalars s0{};
    if ((e = set_scalars(cw, s0, st))) return e;
    hipLaunchKernelGGL(k_cg_init, gr, bl, 0, st, rows, A.row_slot, h.id, b_s, o.warm_start ? pprev_id : nullptr,
                       A.sdiag, cw->y, cw->bh, cw->part, cw->sc);
    hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, rows, A.slice_off, A.col, A.val64, cw->y, cw->bh, A.dgl, cw->r64,
                       cw->p64[0], cw->part, cw->sc);
    if ((e = read_scalars(cw, st))) return e;
    const double bn2 = cw->host->bnorm2;
    stats.rhs_norm = sqrt(bn2);
    if (bn2 == 0.0) {
        // Eigen: a zero right-hand side gives x = 0 (no iteration)
        if ((e = fill_double(cw->y, 0.0, (size_t) rows, st))) return e;
        hipLaunchKernelGGL(k_cg_finish, gr, bl, 0, st, rows, A.row_slot, cw->y, A.sdiag, x_s);
        stats.rel_residual = 0.0;
        return hipGetLastError();
    }
    const double thr = tol * tol * bn2;
    double rho2 = cw->host->rho2;
    int total_it = 0, refinements = 0;

    if (o.precision == MPS_PREC_FP64) {
        if (rho2 > thr) {
            CgScalars s = *cw->host;
            s.thr = thr;
            s.iter = 0;
            s.max_iter = max_it;
            s.done = 0;
            s.converged = 0;
            s.ticket = 0;
            s.beta = 0.0;
            if ((e = set_scalars(cw, s, st))) return e;
            if ((e = run_cg_loop<double, true>(cw, 0, rows, A, A.val64, cw->y, cw->r64, cw->p64, cw->q64, o, st)))
                return e;
            total_it = cw->host->iter;
        }
    } else {
        const int max_ref = o.max_refinements > 0 ? o.max_refinements : 20;
        const double eta = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-4;
        while (rho2 > thr && refinements < max_ref && total_it < max_it) {
            hipLaunchKernelGGL(k_inner_init, gr, bl, 0, st, rows, cw->r64, cw->r32, cw->e32, cw->p32[0], cw->part,
                               cw->sc);
            if ((e = read_scalars(cw, st))) return e;
            CgScalars s = *cw->host;
            // stop the inner solve at eta * its start, and no deeper than FP32
            // can resolve relative to what the outer loop still needs
            const double need = thr / rho2;  // remaining reduction asked of ||r||^2
            double red2 = eta * eta;
            if (need > red2) red2 = need * 0.25;
            s.thr = s.init_rho2 * red2;
            s.iter = 0;
            s.max_iter = max_it - total_it;
            s.done = s.init_rho2 == 0.0 ? 1 : 0;
            s.converged = 0;
            s.ticket = 0;
            s.beta = 0.0;
            if ((e = set_scalars(cw, s, st))) return e;
            if (!s.done) {
                if ((e = run_cg_loop<float, false>(cw, 1, rows, A, A.val32, cw->e32, cw->r32, cw->p32, cw->q32, o, st)))
                    return e;
                total_it += cw->host->iter;
            }
            hipLaunchKernelGGL(k_add_correction, gr, bl, 0, st, rows, cw->y, cw->e32);
            hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, rows, A.slice_off, A.col, A.val64, cw->y, cw->bh, A.dgl,
                               cw->r64, nullptr, cw->part, cw->sc);
            if ((e = read_scalars(cw, st))) return e;
            rho2 = cw->host->rho2;
            refinements++;
        }
    }
    // true FP64 residual of the final iterate
    hipLaunchKernelGGL(k_residual64, gr, bl, 0, st, rows, A.slice_off, A.col, A.val64, cw->y, cw->bh, A.dgl, cw->r64,
                       nullptr, cw->part, cw->sc);
    if ((e = read_scalars(cw, st))) return e;
    stats.iterations = total_it;
    stats.refinements = refinements;
    stats.rel_residual = sqrt(cw->host->rho2 / bn2);
    stats.converged = cw->host->rho2 <= thr ? 1 : 0;
    hipLaunchKernelGGL(k_cg_finish, gr, bl, 0, st, rows, A.row_slot, cw->y, A.sdiag, x_s);
    if (!stats.converged) msg = "pressure solve stopped above its tolerance";
    return hipGetLastError();
}

}  // namespace mps
