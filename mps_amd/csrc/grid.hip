// grid.hip — bucket grid as a GPU counting sort by cell key.
//
// Replaces Bucket::storeParticles (src/bucket.cpp:24-51).  The reference keeps
// first/last per cell and next per particle, appending particles in id order,
// so each cell's list ascends by id.  Here one build is:
//   1. key + count   : cell key per slot (src/bucket.cpp:40-43), atomic count
//   2. scan          : exclusive scan of counts -> cell starts cs[]
//   3. bin           : ids binned into their cell's range (arbitrary order)
//   4. rank + permute: each particle's rank within its cell = number of ids
//                      in the cell smaller than its own; the hot arrays are
//                      copied to their new slot.
// The result is deterministic: cells ordered by key, ids ascending inside.
// Contraction off: distances and weights must round as the reference's
// scalar FP64 code does (no fused multiply-add).
#pragma clang fp contract(off)

#include <chrono>
#include <cstring>

#include "mps_internal.hpp"

namespace mps {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 elements per block
constexpr int kScanTopThreads = 1024;

// Exclusive scan within a block of kBlock threads; returns the exclusive prefix
// of v and the block total in *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; w++) {
            int t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[nw] = acc;
    }
    __syncthreads();
    int res = x - v + lds[wid];
    *total = lds[nw];
    __syncthreads();
    return res;
}

__global__ void __launch_bounds__(kBlock) k_scan_partials(const int* __restrict__ in, int m,
                                                          int* __restrict__ part) {
    __shared__ int lds[kBlock / 64 + 1];
    const int base = blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int i = base + k * kBlock + threadIdx.x;
        if (i < m) s += in[i];
    }
    int tot;
    block_exclusive_scan(s, lds, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanTopThreads) k_scan_top(int* __restrict__ part, int nb) {
    __shared__ int lds[kScanTopThreads / 64 + 1];
    // each thread owns a contiguous run of ceil(nb / threads) partials
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; k++)
        if (b0 + k < nb) s += part[b0 + k];
    int tot;
    int ex = block_exclusive_scan(s, lds, &tot);
    for (int k = 0; k < per; k++) {
        if (b0 + k < nb) {
            int t = part[b0 + k];
            part[b0 + k] = ex;
            ex += t;
        }
    }
    if (threadIdx.x == 0) part[nb] = tot;
}

__global__ void __launch_bounds__(kBlock) k_scan_final(const int* __restrict__ in, int m,
                                                       const int* __restrict__ part,
                                                       int* __restrict__ out, int nb) {
    __shared__ int tile[kScanTile];
    __shared__ int lds[kBlock / 64 + 1];
    const int base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int i = base + k * kBlock + threadIdx.x;
        tile[k * kBlock + threadIdx.x] = i < m ? in[i] : 0;
    }
    __syncthreads();
    int loc[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        loc[k] = s;
        s += tile[threadIdx.x * kScanItems + k];
    }
    int tot;
    int ex = block_exclusive_scan(s, lds, &tot) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) tile[threadIdx.x * kScanItems + k] = ex + loc[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int i = base + k * kBlock + threadIdx.x;
        if (i < m) out[i] = tile[k * kBlock + threadIdx.x];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[m] = part[nb];
}

size_t scan_tmp_ints(int m) { return (size_t) ((m + kScanTile - 1) / kScanTile) + 2; }

// out[0..m] = exclusive prefix sums of in[0..m-1], out[m] = total.
hipError_t scan_exclusive(const int* in, int* out, int m, int* tmp, hipStream_t st) {
    if (m <= 0) return fill_int(out, 0, 1, st);
    const int nb = (m + kScanTile - 1) / kScanTile;
    if (nb > kScanTopThreads * 64) return hipErrorInvalidValue;  // > 134M elements
    hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(kBlock), 0, st, in, m, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, st, tmp, nb);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kBlock), 0, st, in, m, tmp, out, nb);
    return hipGetLastError();
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, size_t n) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static dim3 fill_grid(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((unsigned) b);
}

hipError_t fill_int(int* p, int v, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill<int>, fill_grid(n), dim3(kBlock), 0, st, p, v, n);
    return hipGetLastError();
}
hipError_t fill_double(double* p, double v, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill<double>, fill_grid(n), dim3(kBlock), 0, st, p, v, n);
    return hipGetLastError();
}

// ---------------- host reads ----------------

constexpr int kMailSrcMax = 4;
struct MailArgs {
    const uint32_t* p[kMailSrcMax];
    int words[kMailSrcMax];
    int n;
};

__global__ void __launch_bounds__(kBlock) k_mail(MailArgs a, uint32_t* __restrict__ box, uint32_t seq) {
    int o = 1;
    for (int k = 0; k < a.n; k++) {
        for (int i = threadIdx.x; i < a.words[k]; i += kBlock) box[o + i] = a.p[k][i];
        o += a.words[k];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(box, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void mailbox_free(Mailbox& mb) {
    if (mb.host) hipHostFree(mb.host);
    mb = Mailbox{};
}

hipError_t mailbox_read(Mailbox& mb, const MailSrc* src, int nsrc, void* out, hipStream_t st) {
    if (nsrc > kMailSrcMax) return hipErrorInvalidValue;
    MailArgs a{};
    int words = 0;
    for (int k = 0; k < nsrc; k++) {
        a.p[k] = (const uint32_t*) src[k].p;
        a.words[k] = src[k].bytes / 4;
        words += a.words[k];
    }
    a.n = nsrc;
    hipError_t e;
    if (words > mb.cap) {
        // grow only while the stream has no read in flight (the caller waits
        // for each read before the next)
        if (mb.host) hipHostFree(mb.host);
        mb.host = nullptr;
        const int cap = words < 256 ? 256 : words;
        if ((e = hipHostMalloc((void**) &mb.host, sizeof(uint32_t) * (size_t) (cap + 1),
                               hipHostMallocMapped | hipHostMallocCoherent)))
            return e;
        if ((e = hipHostGetDevicePointer((void**) &mb.dev, mb.host, 0))) return e;
        mb.cap = cap;
        __atomic_store_n(&mb.host[0], mb.seq, __ATOMIC_RELEASE);
    }
    const uint32_t want = ++mb.seq;
    hipLaunchKernelGGL(k_mail, dim3(1), dim3(kBlock), 0, st, a, mb.dev, want);
    if ((e = hipGetLastError())) return e;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0; __atomic_load_n(&mb.host[0], __ATOMIC_ACQUIRE) != want; spin++) {
        if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            if ((e = hipStreamSynchronize(st))) return e;
            if (__atomic_load_n(&mb.host[0], __ATOMIC_ACQUIRE) != want) return hipErrorUnknown;
            break;
        }
    }
    std::memcpy(out, mb.host + 1, sizeof(uint32_t) * (size_t) words);
    return hipSuccess;
}

// ---------------- build ----------------

// Runs of equal keys along a wavefront's lanes (slots are sorted by the
// previous build's key, so most neighbouring lanes share one): the run's
// first lane and, for that lane, the run's length.  One atomic per run
// replaces one per particle.  Every lane of the wavefront must call it.
__device__ __forceinline__ void wave_key_run(int key, bool act, int& head_lane, int& len) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const int prev_act = __shfl_up(act ? 1 : 0, 1, 64);
    const bool head = act && (lane == 0 || !prev_act || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long ends = heads | ~__ballot(act);
    const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
    const unsigned long long hb = heads & upto;
    head_lane = hb ? 63 - __clzll((long long) hb) : lane;
    const unsigned long long after = lane == 63 ? 0ull : ends >> (lane + 1);
    len = after ? __ffsll((long long) after) : 64 - lane;
}

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_key_count(GridView g0, SceneTab tab, int n, Hot cur,
                                                             int* __restrict__ cnt, int* __restrict__ err) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < n;
    int key = -1;
    if (act) {
        GridView g = g0;
        Consts c;
        scene_of<MULTI>(tab, cur.id[s], g, c);
        key = g.ghost_key;
        if (cur.type[s] != MPS_GHOST) {
            int ix, iy, iz = 1;
            bool ok = cell_coord(cur.px[s], g.xmin, g.len, g.nx, ix) &&
                      cell_coord(cur.py[s], g.ymin, g.len, g.ny, iy);
            if (DIM == 3) ok = ok && cell_coord(cur.pz[s], g.zmin, g.len, g.nz, iz);
            if (ok && g.flat && iz != 1) ok = false;
            if (ok) {
                key = g.key0 + (g.flat ? ix + iy * g.nx : ix + iy * g.nx + iz * g.nx * g.ny);
            } else {
                atomicOr(err, kErrRange);
            }
        }
        cur.key[s] = key;
    }
    int head, len;
    wave_key_run(key, act, head, len);
    if (act && head == (int) (threadIdx.x & 63)) atomicAdd(&cnt[key], len);
}

__global__ void __launch_bounds__(kMaxTuneBlock) k_bin(int n, Hot cur, int* __restrict__ cnt, const int* __restrict__ cs,
                                                int* __restrict__ bin_id, int* __restrict__ dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < n;
    const int key = act ? cur.key[s] : -1;
    int head, len;
    wave_key_run(key, act, head, len);
    const int lane = threadIdx.x & 63;
    int top = 0;
    if (act && head == lane) top = atomicSub(&cnt[key], len);  // leaves cnt at zero for the next build
    top = __shfl(top, head, 64);
    if (!act) return;
    const int pos = top - 1 - (lane - head);
    const int d = cs[key] + pos;
    bin_id[d] = cur.id[s];
    dst[s] = d;
}

template <int DIM>
__global__ void __launch_bounds__(kMaxTuneBlock) k_rank_permute(int ghost_key, const int* __restrict__ cs, int n,
                                                                Hot cur, Hot nxt,
                                                         const int* __restrict__ bin_id,
                                                         const int* __restrict__ dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int key = cur.key[s];
    const int myid = cur.id[s];
    int t;
    if (key != ghost_key) {
        const int b = cs[key], e = cs[key + 1];
        int r = 0;
        // eight loads in flight per round (indices past the cell re-read
        // its last entry and are not counted)
        for (int q = b; q < e; q += 8) {
            int v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bin_id[q + k < e ? q + k : e - 1];
#pragma unroll
            for (int k = 0; k < 8; k++) r += (q + k < e && v[k] < myid) ? 1 : 0;
        }
        t = b + r;
    } else {
        t = dst[s];  // ghosts: order inside the ghost segment is immaterial
    }
    nxt.px[t] = cur.px[s];
    nxt.py[t] = cur.py[s];
    nxt.vx[t] = cur.vx[s];
    nxt.vy[t] = cur.vy[s];
    if (DIM == 3) {
        nxt.pz[t] = cur.pz[s];
        nxt.vz[t] = cur.vz[s];
    }
    nxt.id[t] = myid;
    nxt.type[t] = cur.type[s];
    nxt.key[t] = key;
}

// id -> slot from the slots' ids (the builds do not keep this table: only
// mps_upload reads it).
__global__ void __launch_bounds__(kBlock) k_slot_of_ids(int n, const int* __restrict__ id, int* __restrict__ slot_of) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) slot_of[id[s]] = s;
}

hipError_t slot_of_ids(int n, const int* id, int* slot_of, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_slot_of_ids, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, id, slot_of);
    return hipGetLastError();
}

hipError_t build_grid(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& cur, Hot& nxt,
                      const BuildBufs& b, int* err, hipStream_t st, int bs) {
    const dim3 grid((n + bs - 1) / bs), blk(bs);
    const int keys = g.ghost_key + 1;  // every scene's cells and the ghost bin
    if (n == 0) return scan_exclusive(b.cnt, b.cs, keys, b.scan_tmp, st);
    const bool multi = tab.g != nullptr;
    if (dim == 2) {
        if (multi)
            hipLaunchKernelGGL((k_key_count<2, true>), grid, blk, 0, st, g, tab, n, cur, b.cnt, err);
        else
            hipLaunchKernelGGL((k_key_count<2, false>), grid, blk, 0, st, g, tab, n, cur, b.cnt, err);
    } else {
        if (multi)
            hipLaunchKernelGGL((k_key_count<3, true>), grid, blk, 0, st, g, tab, n, cur, b.cnt, err);
        else
            hipLaunchKernelGGL((k_key_count<3, false>), grid, blk, 0, st, g, tab, n, cur, b.cnt, err);
    }
    hipError_t e = scan_exclusive(b.cnt, b.cs, keys, b.scan_tmp, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin, grid, blk, 0, st, n, cur, b.cnt, b.cs, b.bin_id, b.dst);
    if (dim == 2)
        hipLaunchKernelGGL(k_rank_permute<2>, grid, blk, 0, st, g.ghost_key, b.cs, n, cur, nxt, b.bin_id, b.dst);
    else
        hipLaunchKernelGGL(k_rank_permute<3>, grid, blk, 0, st, g.ghost_key, b.cs, n, cur, nxt, b.bin_id, b.dst);
    return hipGetLastError();
}

// ---------------- neighbour dumps (test entry points) ----------------

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_nb_count(GridView g0, SceneTab tab, int n, Hot h,
                                                            int* __restrict__ cnt_by_id, double re0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    GridView g = g0;
    Consts cc;
    scene_of<MULTI>(tab, h.id[s], g, cc);
    const double re = MULTI ? cc.re_max : re0;
    int c = 0;
    if (h.type[s] != MPS_GHOST) {
        const double zi = DIM == 3 ? h.pz[s] : 0.0;
        for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                           [&](int, double, double, double, double) { c++; });
    }
    cnt_by_id[h.id[s]] = c;
}

template <int DIM, bool MULTI>
__global__ void __launch_bounds__(kMaxTuneBlock) k_nb_list(GridView g0, SceneTab tab, int n, Hot h,
                                                           const int* __restrict__ off_by_id, int* __restrict__ ids,
                                                           double* __restrict__ dists, double re0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (h.type[s] == MPS_GHOST) return;
    GridView g = g0;
    Consts cc;
    scene_of<MULTI>(tab, h.id[s], g, cc);
    const double re = MULTI ? cc.re_max : re0;
    int o = off_by_id[h.id[s]];
    const double zi = DIM == 3 ? h.pz[s] : 0.0;
    for_neighbors<DIM>(g, h.px, h.py, h.pz, h.key[s], s, h.px[s], h.py[s], zi, re,
                       [&](int q, double dist, double, double, double) {
                           ids[o] = h.id[q];
                           dists[o] = dist;
                           o++;
                       });
}

hipError_t dump_neighbor_counts(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& h,
                                int* cnt_by_id, double re, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kBlock - 1) / kBlock), blk(kBlock);
    const bool m = tab.g != nullptr;
    if (dim == 2 && m)
        hipLaunchKernelGGL((k_nb_count<2, true>), grid, blk, 0, st, g, tab, n, h, cnt_by_id, re);
    else if (dim == 2)
        hipLaunchKernelGGL((k_nb_count<2, false>), grid, blk, 0, st, g, tab, n, h, cnt_by_id, re);
    else if (m)
        hipLaunchKernelGGL((k_nb_count<3, true>), grid, blk, 0, st, g, tab, n, h, cnt_by_id, re);
    else
        hipLaunchKernelGGL((k_nb_count<3, false>), grid, blk, 0, st, g, tab, n, h, cnt_by_id, re);
    return hipGetLastError();
}

hipError_t dump_neighbor_lists(int dim, const GridView& g, const SceneTab& tab, int n, const Hot& h,
                               const int* off_by_id, int* ids, double* dists, double re, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kBlock - 1) / kBlock), blk(kBlock);
    const bool m = tab.g != nullptr;
    if (dim == 2 && m)
        hipLaunchKernelGGL((k_nb_list<2, true>), grid, blk, 0, st, g, tab, n, h, off_by_id, ids, dists, re);
    else if (dim == 2)
        hipLaunchKernelGGL((k_nb_list<2, false>), grid, blk, 0, st, g, tab, n, h, off_by_id, ids, dists, re);
    else if (m)
        hipLaunchKernelGGL((k_nb_list<3, true>), grid, blk, 0, st, g, tab, n, h, off_by_id, ids, dists, re);
    else
        hipLaunchKernelGGL((k_nb_list<3, false>), grid, blk, 0, st, g, tab, n, h, off_by_id, ids, dists, re);
    return hipGetLastError();
}

}  // namespace mps
