// capi.hip — the C ABI of libmps_amd.so (include/mps_amd.h): handle, device
// allocation, input validation and the orchestration of one MPS step
// (MPS::stepForward, src/mps.cpp:111-130) on one HIP stream.
// Contraction off: distances and weights must round as the reference's
// scalar FP64 code does (no fused multiply-add).
#pragma clang fp contract(off)

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mps_internal.hpp"

using namespace mps;

// One independent scene of a handle (a handle from mps_create holds one).
struct SceneInfo {
    mps_settings s{};
    Consts c{};
    GridView g{};
    int iz0 = 1;  // 2D: the common z layer of the reference's key
    int id0 = 0;  // first particle id of the scene
    int n = 0;
};

// Device memory of the reverse-mode step (allocated on first use).
struct AdjStore {
    int steps = 0, interval = 0, ready = 0;
    mps_solver_options opt{};          // solver options of the recorded run
    std::vector<double*> cpf;          // checkpoints: state before steps 0, M, 2M, ...
    std::vector<int8_t*> cpt;
    double* finf = nullptr;            // state after the recorded run
    int8_t* fint = nullptr;
    std::vector<double*> segf;         // step starts of the segment being reversed
    std::vector<int8_t*> segt;
    bool bufs = false;
    AdjBufs a{};
    int8_t* type_g = nullptr;          // slot-order types at the pressure-gradient stage
    CollTape coll{};
    int* ident = nullptr;              // 0..n-1
    int32_t* key0 = nullptr;           // keys of the start-state view
    Hot hs{};                          // start state rebuilt into slots
    std::vector<void*> owned;          // everything above, for release
};

// Pinned host copy of the export fields (id order), filled asynchronously.
struct ExportSlot {
    int8_t* type = nullptr;
    double* pos = nullptr;
    double* vel = nullptr;
    double* p = nullptr;
    double* ndr = nullptr;
    int8_t* bc = nullptr;
    hipEvent_t done = nullptr;
    double* dpos = nullptr;   // device staging (id order)
    double* dvel = nullptr;
    int8_t* dtype = nullptr;
    bool pending = false;
};

struct mps_handle {
    std::string err;
    mps_settings s{};  // scene 0
    Consts c{};        // scene 0 (kernel arguments of single-scene handles)
    GridView g{};      // scene 0, ghost key of the whole key space
    int n = 0;
    int dim = 2;
    int iz0 = 1;
    int nscene = 1;
    std::vector<SceneInfo> scenes;
    std::vector<GridView> sgrids;   // each scene's bucket grid (host)
    GridView* d_g = nullptr;   // device scene table (ensembles)
    Consts* d_c = nullptr;
    int16_t* d_sid = nullptr;
    int* d_key0 = nullptr;     // nscene + 1 first keys
    std::vector<mps_solver_stats> sstats;
    std::vector<double> scourant;
    hipStream_t st = nullptr;
    // hot sets
    double* P[2][6] = {};  // px py pz vx vy vz
    int32_t* ID[2] = {};
    int8_t* TY[2] = {};
    int32_t* KEY[2] = {};
    int pv = 0, it = 0;
    BuildBufs bb{};
    Cold cold{};
    StepScratch w{};
    double* x_s = nullptr;
    int32_t* dbg_id = nullptr;  // slot -> id at the pressure stage of the last step
    double* snap_pos = nullptr;
    double* snap_vel = nullptr;
    double* tmp3a = nullptr;  // 3n scratch for downloads
    double* tmp3b = nullptr;
    int8_t* tmp8 = nullptr;
    Sell A{};
    CgWork* cg = nullptr;
    mps_solver_options opt{};
    Events ev{};
    int* dev_counters = nullptr;  // [0] ghost, [1] unconnected, [2] err
    int ev_cap = 0;
    std::vector<int32_t> ghost_ids, unconn_ids;
    int64_t counts[3] = {0, 0, 0};
    int* pinned_i = nullptr;  // host pinned small buffer
    Mailbox mb;               // small device-to-host reads of the step
    double* vmax_dst = nullptr;  // where the next drain puts the scenes' largest speeds
    double* pinned_d = nullptr;
    double* vmax = nullptr;
    std::string solve_msg;  // set when a pressure solve of the last step stopped above its tolerance
    int dev = 0;            // device the handle was created on (its scratch copy)
    int want_raw = 0;
    int want_snap = 0;
    int have_system = 0;
    int built_once = 0;
    // launch block sizes per kernel family (see Fam); results do not depend
    // on them, the autotuner only changes speed
    int bs[9] = {256, 256, 256, 256, 256, 256, 256, 256, 256};
    bool timing = false;
    hipEvent_t tev[16][2] = {};
    int tfam[16] = {};
    int nsites = 0;
    double tacc[9] = {};
    AdjStore adj;
    ExportSlot xs[2];
};

// Kernel families the autotuner sizes independently.
enum Fam { F_BUILD, F_VISC, F_COLL, F_BCSRC, F_DIAGCC, F_FILL, F_POUT, F_GRAD, F_NDR, F_COUNT };

static void tbeg(mps_handle* h, int fam) {
    if (!h->timing || h->nsites >= 16) return;
    h->tfam[h->nsites] = fam;
    hipEventRecord(h->tev[h->nsites][0], h->st);
}
static void tend(mps_handle* h) {
    if (!h->timing || h->nsites >= 16) return;
    hipEventRecord(h->tev[h->nsites][1], h->st);
    h->nsites++;
}

static std::string g_create_error = "no error";

static int fail(mps_handle* h, int code, const std::string& m) {
    if (h) h->err = m;
    return code;
}

static int hip_fail(mps_handle* h, hipError_t e, const char* where) {
    return fail(h, MPS_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(h, expr)                                       \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return hip_fail((h), _e, #expr); \
    } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
    if (count == 0) count = 1;
    return hipMalloc((void**) p, count * sizeof(T));
}

static Hot hot_view(mps_handle* h, int pv, int it) {
    Hot v;
    v.px = h->P[pv][0];
    v.py = h->P[pv][1];
    v.pz = h->P[pv][2];
    v.vx = h->P[pv][3];
    v.vy = h->P[pv][4];
    v.vz = h->P[pv][5];
    v.id = h->ID[it];
    v.type = h->TY[it];
    v.key = h->KEY[it];
    return v;
}

static Hot cur(mps_handle* h) { return hot_view(h, h->pv, h->it); }

static SceneTab tab(const mps_handle* h) {
    if (h->nscene > 1) return SceneTab{h->d_g, h->d_c, h->d_sid};
    return SceneTab{nullptr, nullptr, nullptr};
}

static void ref_values(int dim, double l0, double re, double* n0_out, double* lam_out) {
    // RefValues (src/mps.cpp:132-163), host-side, FP64 in the reference's order
    int iz0 = dim == 2 ? 0 : -4, iz1 = dim == 2 ? 1 : 5;
    double n0 = 0.0, lam = 0.0;
    for (int ix = -4; ix < 5; ix++)
        for (int iy = -4; iy < 5; iy++)
            for (int iz = iz0; iz < iz1; iz++) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                const double rx = l0 * (double) ix, ry = l0 * (double) iy, rz = l0 * (double) iz;
                const double dist = sqrt((rx * rx + ry * ry) + rz * rz);
                const double dist2 = dist * dist;
                double w = 0.0;
                if (dist < re) w = (re / dist) - 1.0;
                n0 += w;
                lam += dist2 * w;
            }
    lam /= n0;
    *n0_out = n0;
    *lam_out = lam;
}

// Bucket::Bucket (src/bucket.cpp:10-22) for one scene.
static bool scene_grid(const mps_settings& s, int dim, GridView& g, std::string& err) {
    g.xmin = s.domain_min[0];
    g.ymin = s.domain_min[1];
    g.zmin = s.domain_min[2];
    g.len = s.re_max;
    const double lx = s.domain_max[0] - s.domain_min[0];
    const double ly = s.domain_max[1] - s.domain_min[1];
    const double lz = s.domain_max[2] - s.domain_min[2];
    const double fx = lx / g.len, fy = ly / g.len, fz = lz / g.len;
    if (!(fx >= 0 && fy >= 0 && fz >= 0) || fx > 1e8 || fy > 1e8 || fz > 1e8) {
        err = "domain range is not a valid bucket grid";
        return false;
    }
    g.nx = (int) fx + 3;
    g.ny = (int) fy + 3;
    g.nz = (int) fz + 3;
    const double cells = (double) g.nx * g.ny * (dim == 2 ? 1 : g.nz);
    if (cells > 1.0e9) {
        err = "bucket grid too large";
        return false;
    }
    g.flat = (dim == 2 || g.nz == 3) ? 1 : 0;
    g.ncell = g.flat ? g.nx * g.ny : g.nx * g.ny * g.nz;
    g.key0 = 0;
    g.ghost_key = g.ncell;
    g.cs = nullptr;
    // MPS_AMD_STAGE_CAP lowers the walks' LDS staging capacity (0: every walk
    // reads global memory); results do not depend on it
    g.stage_cap = 1 << 30;
    if (const char* e = getenv("MPS_AMD_STAGE_CAP")) g.stage_cap = std::max(atoi(e), 0);
    g.stage_stats = 0;
    if (const char* e = getenv("MPS_AMD_STAGE_STATS")) g.stage_stats = atoi(e) != 0;
    return true;
}

// Settings-derived constants (src/settings.cpp:43-61) and RefValues
// (src/mps.cpp:21-35) of one scene.
static void scene_consts(const mps_settings& s, Consts& c) {
    c = Consts{};
    c.dim = s.dim;
    c.dt = s.dt;
    c.l0 = s.particle_distance;
    c.rho = s.density;
    c.nu = s.kinematic_viscosity;
    for (int k = 0; k < 3; k++) {
        c.g[k] = s.gravity[k];
        c.dmin[k] = s.domain_min[k];
        c.dmax[k] = s.domain_max[k];
    }
    c.re_p = s.re_pressure;
    c.re_v = s.re_viscosity;
    c.re_s = s.re_surface;
    c.re_max = s.re_max;
    ref_values(s.dim, s.particle_distance, s.re_pressure, &c.n0_p, &c.lam_p);
    ref_values(s.dim, s.particle_distance, s.re_viscosity, &c.n0_v, &c.lam_v);
    ref_values(s.dim, s.particle_distance, s.re_surface, &c.n0_s, &c.lam_s);
    c.beta = s.surface_threshold;
    c.relax = s.relaxation;
    c.qc_on = s.qc_on;
    c.alpha = s.qc_on ? s.compressibility : 0.0;
    c.hos_on = s.hos_on;
    c.hos_gamma = s.hos_gamma;
    c.dc = s.collision_distance;
    c.restitution = s.restitution;
    c.cfl = s.cfl_condition;
    const double dt = c.dt;
    const double D = (double) c.dim;
    c.avisc = (c.nu) * (2.0 * D) / (c.n0_v * c.lam_v);
    c.src_scale = c.relax * (1.0 / (dt * dt));
    c.hos_scale = c.hos_gamma * (1.0 / (dt * dt));
    c.hos_coef = (1.0 / (c.n0_p * dt));
    c.coefA = (2.0 * D / (c.lam_p * c.n0_p));
    c.qc_add = c.alpha / (dt * dt);
    c.gscale = D / c.n0_p;
}

static int alloc_all(mps_handle* h) {
    const size_t n = (size_t) h->n;
    const int d3 = h->dim == 3;
    const size_t ns = (size_t) h->nscene;
    for (int k = 0; k < 2; k++) {
        for (int f = 0; f < 6; f++) {
            if (!d3 && (f == 2 || f == 5)) continue;
            HIPCHK(h, dalloc(&h->P[k][f], n));
        }
        HIPCHK(h, dalloc(&h->ID[k], n));
        HIPCHK(h, dalloc(&h->TY[k], n));
        HIPCHK(h, dalloc(&h->KEY[k], n));
    }
    const size_t nc = (size_t) h->g.ghost_key;  // every scene's cells
    HIPCHK(h, dalloc(&h->bb.cnt, nc + 1));
    HIPCHK(h, dalloc(&h->bb.cs, nc + 2));
    HIPCHK(h, dalloc(&h->bb.bin_id, n));
    HIPCHK(h, dalloc(&h->bb.dst, n));
    const size_t scan_m = std::max(nc + 1, n + 1);
    HIPCHK(h, dalloc(&h->bb.scan_tmp, scan_tmp_ints((int) scan_m) + 16));
    HIPCHK(h, dalloc(&h->bb.slot_of, n));
    HIPCHK(h, dalloc(&h->cold.p, n));
    HIPCHK(h, dalloc(&h->cold.pmin, n));
    HIPCHK(h, dalloc(&h->cold.ndr, n));
    HIPCHK(h, dalloc(&h->cold.bc, n));
    HIPCHK(h, dalloc(&h->cold.pprev, n));
    HIPCHK(h, dalloc(&h->cold.pprev2, n));
    HIPCHK(h, dalloc(&h->cold.ghost_n, n));
    HIPCHK(h, dalloc(&h->cold.z, n));
    HIPCHK(h, dalloc(&h->cold.vz, n));
    StepScratch& w = h->w;
    HIPCHK(h, dalloc(&w.bc_s, n));
    HIPCHK(h, dalloc(&w.b_s, n));
    HIPCHK(h, dalloc(&w.diag_s, n));
    HIPCHK(h, dalloc(&w.isq_s, n));
    HIPCHK(h, dalloc(&w.nnz_s, n));
    HIPCHK(h, dalloc(&w.parent, n));
    HIPCHK(h, dalloc(&w.surf, n + 8));
    HIPCHK(h, dalloc(&w.inner_flag, n));
    HIPCHK(h, dalloc(&w.row_of, n + 1));
    HIPCHK(h, dalloc(&w.p_s, n));
    HIPCHK(h, dalloc(&w.pmin_s, n));
    HIPCHK(h, dalloc(&w.vmax, ns));
    HIPCHK(h, dalloc(&h->d_key0, ns + 1));
    if (ns > 1) {
        HIPCHK(h, dalloc(&h->d_g, ns));
        HIPCHK(h, dalloc(&h->d_c, ns));
        HIPCHK(h, dalloc(&h->d_sid, n));
    }
    HIPCHK(h, dalloc(&w.coll_cnt, n));
    HIPCHK(h, dalloc(&w.coll_off, n + 1));
    w.coll_cap = (int) std::min<size_t>(std::max<size_t>(4 * n, 4096), (size_t) 1 << 26);
    HIPCHK(h, dalloc(&w.coll_pairs, 2 * (size_t) w.coll_cap));
    HIPCHK(h, dalloc(&w.coll_dist, (size_t) w.coll_cap));
    HIPCHK(h, dalloc(&w.coll_level, (size_t) w.coll_cap));
    w.coll_n2 = 1;
    while (w.coll_n2 < w.coll_cap) w.coll_n2 <<= 1;
    HIPCHK(h, dalloc(&w.coll_key, (size_t) w.coll_n2));
    HIPCHK(h, dalloc(&w.coll_prev, (size_t) w.coll_cap));
    w.scan_tmp = h->bb.scan_tmp;
    // pressure-stage neighbour lists: room for 1.5x the neighbours a lattice
    // of spacing l0 has within max(re_s, re_p), plus 8 (slots beyond it are
    // walked again by the readers), at most 256 per slot and 24 GB in all
    {
        double est = 0.0;
        for (const SceneInfo& sc : h->scenes) {
            const double r = std::max(sc.c.re_s, sc.c.re_p) / sc.c.l0;
            const double v = h->dim == 3 ? 4.0 / 3.0 * M_PI * r * r * r : M_PI * r * r;
            est = std::max(est, v);
        }
        size_t cap = (size_t) (1.5 * est) + 8;
        cap = std::min<size_t>(cap, 256);
        const size_t budget = (size_t) 24 << 30;
        if (n > 0 && cap * n * 12 > budget) cap = budget / (12 * n);
        // MPS_AMD_NB_CAP overrides the capacity (0: every reader walks the cells)
        if (const char* e = getenv("MPS_AMD_NB_CAP")) cap = std::min<size_t>((size_t) std::max(atoi(e), 0), cap);
        w.nb.cap = (int) cap;
        w.nb.stride = n;
        HIPCHK(h, dalloc(&w.nb.cnt, n));
        HIPCHK(h, dalloc(&w.nb.q, std::max<size_t>(cap * n, 1)));
        HIPCHK(h, dalloc(&w.nb.d, std::max<size_t>(cap * n, 1)));
    }
    HIPCHK(h, dalloc(&h->x_s, n));
    HIPCHK(h, dalloc(&h->dbg_id, n));
    HIPCHK(h, dalloc(&h->tmp3a, 3 * n));
    HIPCHK(h, dalloc(&h->tmp3b, 3 * n));
    HIPCHK(h, dalloc(&h->tmp8, n));
    const size_t nsl = (n + 63) / 64 + 2;
    HIPCHK(h, dalloc(&h->A.slice_off, nsl + 1));
    HIPCHK(h, dalloc(&h->A.slice_w, nsl + 1));
    HIPCHK(h, dalloc(&h->A.row_slot, n));
    HIPCHK(h, dalloc(&h->A.sdiag, n));
    HIPCHK(h, dalloc(&h->A.dgl, n));
    HIPCHK(h, dalloc(&h->A.raw_off, n + 1));
    HIPCHK(h, dalloc(&h->A.maxdelta, 1));
    h->ev_cap = (int) std::max<size_t>(n, 1024);
    HIPCHK(h, dalloc(&h->dev_counters, 4));
    HIPCHK(h, dalloc(&h->ev.ghost_ids, (size_t) h->ev_cap));
    HIPCHK(h, dalloc(&h->ev.unconn_ids, (size_t) h->ev_cap));
    h->ev.n_ghost = h->dev_counters;
    h->ev.n_unconn = h->dev_counters + 1;
    h->ev.err = h->dev_counters + 2;
    h->ev.cap = h->ev_cap;
    HIPCHK(h, hipHostMalloc((void**) &h->pinned_i, 16 * sizeof(int), hipHostMallocDefault));
    HIPCHK(h, hipHostMalloc((void**) &h->pinned_d, (16 + ns) * sizeof(double), hipHostMallocDefault));
    HIPCHK(h, fill_int(h->bb.cnt, 0, nc + 1, h->st));
    HIPCHK(h, fill_int(h->dev_counters, 0, 4, h->st));
    HIPCHK(h, fill_double(h->cold.p, 0.0, n, h->st));
    HIPCHK(h, fill_double(h->cold.pmin, 0.0, n, h->st));
    HIPCHK(h, fill_double(h->cold.ndr, 0.0, n, h->st));
    HIPCHK(h, fill_double(h->cold.pprev, 0.0, n, h->st));
    HIPCHK(h, fill_double(h->cold.pprev2, 0.0, n, h->st));
    HIPCHK(h, fill_double(h->cold.ghost_n, 0.0, n, h->st));
    HIPCHK(h, hipMemsetAsync(h->cold.bc, 0, n, h->st));  // Ignored
    h->cg = cg_create();
    return MPS_OK;
}

static void free_all(mps_handle* h) {
    for (int k = 0; k < 16; k++)
        for (int j = 0; j < 2; j++)
            if (h->tev[k][j]) hipEventDestroy(h->tev[k][j]);
    for (int k = 0; k < 2; k++) {
        for (int f = 0; f < 6; f++) hipFree(h->P[k][f]);
        hipFree(h->ID[k]);
        hipFree(h->TY[k]);
        hipFree(h->KEY[k]);
    }
    void* ptrs[] = {h->bb.cnt,      h->bb.cs,         h->bb.bin_id,  h->bb.dst,       h->bb.scan_tmp, h->bb.slot_of,
                    h->cold.p,      h->cold.pmin,     h->cold.ndr,   h->cold.bc,      h->cold.pprev,  h->cold.pprev2, h->cold.ghost_n,
                    h->cold.z,      h->cold.vz,       h->w.bc_s,     h->w.b_s,        h->w.diag_s,    h->w.isq_s,    h->w.nnz_s,
                    h->w.parent,    h->w.surf,        h->w.inner_flag, h->w.row_of,   h->w.p_s,       h->w.pmin_s,
                    h->w.vmax,      h->w.coll_cnt,    h->w.coll_off, h->w.coll_pairs, h->w.coll_dist, h->w.coll_level,
                    h->w.coll_key,  h->w.coll_prev,   h->x_s,           h->dbg_id,     h->tmp3a,        h->tmp3b,       h->tmp8,
                    h->A.slice_off, h->A.slice_w,     h->A.row_slot, h->A.sdiag,      h->A.dgl,       h->A.raw_off,
                    h->A.col,       h->A.dcol,        h->A.anc,         h->A.maxdelta, h->A.val64,       h->A.val32,    h->A.raw_col,    h->A.raw_val,   h->dev_counters,
                    h->ev.ghost_ids, h->ev.unconn_ids, h->snap_pos,  h->snap_vel,     h->d_g,         h->d_c,
                    h->d_sid,       h->d_key0,        h->w.nb.cnt,   h->w.nb.q,       h->w.nb.d};
    for (void* p : ptrs) hipFree(p);
    for (void* p : h->adj.owned) hipFree(p);
    h->adj.owned.clear();
    for (ExportSlot& x : h->xs) {
        if (x.done) hipEventDestroy(x.done);
        void* hp[] = {x.type, x.pos, x.vel, x.p, x.ndr, x.bc};
        for (void* q : hp)
            if (q) hipHostFree(q);
        hipFree(x.dpos);
        hipFree(x.dvel);
        hipFree(x.dtype);
    }
    if (h->pinned_i) hipHostFree(h->pinned_i);
    mailbox_free(h->mb);
    if (h->pinned_d) hipHostFree(h->pinned_d);
    if (h->cg) cg_destroy(h->cg);
    if (h->st) hipStreamDestroy(h->st);
}

// ---------------- allocation-free inspection: static device scratch ----------------
//
// mps_dump_neighbors_noalloc stages the neighbour lists in this file-scope
// device array instead of buffers it allocates.  Each device holds its own
// copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (the array's name in
// host code is a host-side stand-in, not a device address).
constexpr size_t kNoAllocMaxEntries = (size_t) 1 << 22;  // 4,194,304 neighbour entries
constexpr size_t kNoAllocScratchBytes = kNoAllocMaxEntries * (sizeof(int32_t) + sizeof(double));  // 48 MiB
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Looks up the scratch array's address on the current device once per
// device (the handle's set-up calls it with its device current).
static hipError_t scratch_setup(int* dev_out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return hipSuccess;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_dev[dev] = ptr;
    return hipSuccess;
}

// ---------------- internal stage drivers ----------------

static int do_build(mps_handle* h) {
    Hot c = cur(h);
    Hot nx = hot_view(h, 1 - h->pv, 1 - h->it);
    tbeg(h, F_BUILD);
    HIPCHK(h, build_grid(h->dim, h->g, tab(h), h->n, c, nx, h->bb, h->ev.err, h->st, h->bs[F_BUILD]));
    tend(h);
    h->pv ^= 1;
    h->it ^= 1;
    h->built_once = 1;
    return MPS_OK;
}

static GridView gview(mps_handle* h) {
    GridView g = h->g;
    g.cs = h->bb.cs;
    return g;
}

// Reads back the device error mask and event counters, and the scenes'
// largest speeds when the step asked for them (waits for the stream's work
// before the read).
static int drain(mps_handle* h) {
    {
        MailSrc src[2] = {{h->dev_counters, 4 * (int) sizeof(int)}, {h->w.vmax, (int) sizeof(double) * h->nscene}};
        std::vector<double> buf(2 + (size_t) h->nscene);
        HIPCHK(h, mailbox_read(h->mb, src, h->vmax_dst ? 2 : 1, buf.data(), h->st));
        std::memcpy(h->pinned_i, buf.data(), 4 * sizeof(int));
        if (h->vmax_dst) std::memcpy(h->vmax_dst, buf.data() + 2, sizeof(double) * (size_t) h->nscene);
        h->vmax_dst = nullptr;
    }
    const int ng = h->pinned_i[0], nu = h->pinned_i[1], er = h->pinned_i[2];
    if (ng > 0 || nu > 0) {
        std::vector<int32_t> tmp((size_t) std::max(std::min(ng, h->ev_cap), std::min(nu, h->ev_cap)));
        if (ng > 0) {
            const int m = std::min(ng, h->ev_cap);
            HIPCHK(h, hipMemcpy(tmp.data(), h->ev.ghost_ids, sizeof(int32_t) * (size_t) m, hipMemcpyDeviceToHost));
            h->ghost_ids.insert(h->ghost_ids.end(), tmp.begin(), tmp.begin() + m);
        }
        if (nu > 0) {
            const int m = std::min(nu, h->ev_cap);
            HIPCHK(h, hipMemcpy(tmp.data(), h->ev.unconn_ids, sizeof(int32_t) * (size_t) m, hipMemcpyDeviceToHost));
            h->unconn_ids.insert(h->unconn_ids.end(), tmp.begin(), tmp.begin() + m);
        }
        h->counts[0] += ng;
        h->counts[1] += nu;
        HIPCHK(h, fill_int(h->dev_counters, 0, 2, h->st));
    }
    if (er) {
        HIPCHK(h, fill_int(h->dev_counters + 2, 0, 1, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
        if (er & kErrRange) return fail(h, MPS_E_RANGE, "a particle fell outside the bucket grid");
        if (er & kErrCollCap) return fail(h, MPS_E_INVALID, "collision pair buffer overflow");
        return fail(h, MPS_E_INVALID, "device error flag set");
    }
    return MPS_OK;
}

static int number_density(mps_handle* h) {
    int rc = do_build(h);
    if (rc) return rc;
    GridView g = gview(h);
    tbeg(h, F_NDR);
    HIPCHK(h, launch_ndr(h->dim, g, tab(h), h->n, h->c, cur(h), h->cold, h->st, h->bs[F_NDR]));
    tend(h);
    return MPS_OK;
}

// Hooks of a step run for the reverse sweep: the collision record, the
// slot-order types at the pressure-gradient stage (before that stage ghosts
// anything) and the hot view of the pressure stage's bucket build.
struct StepTap {
    const CollTape* coll;
    int8_t* type_g;
    Hot c3;
};

static int step_core(mps_handle* h, int is_export, double* vmax_out, StepTap* tap = nullptr) {
    int rc;
    const int n = h->n;
    // setNeighbors #1, gravity, viscosity, move
    if ((rc = do_build(h))) return rc;
    {
        GridView g = gview(h);
        Hot c = cur(h);
        Hot o = hot_view(h, 1 - h->pv, h->it);
        tbeg(h, F_VISC);
        HIPCHK(h, launch_visc_move(h->dim, g, tab(h), n, h->c, c, o, h->cold, h->ev, h->st, h->bs[F_VISC]));
        tend(h);
        h->pv ^= 1;
    }
    // setNeighbors #2, collision
    if ((rc = do_build(h))) return rc;
    {
        GridView g = gview(h);
        Hot c = cur(h);
        tbeg(h, F_COLL);
        HIPCHK(h, launch_collision(h->dim, g, tab(h), n, h->c, c, h->w, h->ev.err, h->st, h->bs[F_COLL],
                                   tap ? tap->coll : nullptr));
        tend(h);
    }
    // setNeighbors #3, pressure
    if ((rc = do_build(h))) return rc;
    GridView g = gview(h);
    Hot c = cur(h);
    if (tap) tap->c3 = c;
    if (h->want_snap) HIPCHK(h, launch_snapshot(h->dim, n, c, h->cold, h->snap_pos, h->snap_vel, h->st));
    if (h->want_raw) HIPCHK(h, hipMemcpyAsync(h->dbg_id, c.id, sizeof(int32_t) * (size_t) n, hipMemcpyDeviceToDevice, h->st));
    tbeg(h, F_BCSRC);
    HIPCHK(h, launch_bc_source(h->dim, g, tab(h), n, h->c, c, h->cold, h->w, h->st, h->bs[F_BCSRC]));
    tend(h);
    tbeg(h, F_DIAGCC);
    HIPCHK(h, launch_diag_cc(h->dim, g, tab(h), n, h->c, c, h->w, h->ev, h->st, h->bs[F_DIAGCC]));
    tend(h);
    int rows = 0;
    size_t entries = 0;
    tbeg(h, F_FILL);
    HIPCHK(h, assemble_sell(h->dim, g, tab(h), n, h->c, c, h->w, h->A, h->opt.precision == MPS_PREC_MIXED, h->want_raw,
                            h->ev.err, &rows, &entries, h->mb, h->st, h->bs[F_FILL]));
    tend(h);
    std::string msg;
    HIPCHK(h, cg_solve(h->cg, h->A, c, n, h->w.b_s, h->cold.pprev, h->cold.pprev2, h->x_s, h->opt, h->sstats.data(), h->st, msg));
    h->solve_msg = msg;
    tbeg(h, F_POUT);
    HIPCHK(h, launch_pressure_out(h->dim, g, tab(h), n, h->c, c, h->x_s, h->cold, h->w, h->st, h->bs[F_POUT]));
    tend(h);
    h->have_system = h->want_raw;
    // pressure gradient + move
    if (tap && n > 0) HIPCHK(h, hipMemcpyAsync(tap->type_g, c.type, (size_t) n, hipMemcpyDeviceToDevice, h->st));
    {
        Hot o = hot_view(h, 1 - h->pv, h->it);
        tbeg(h, F_GRAD);
        HIPCHK(h, launch_grad_move(h->dim, g, tab(h), n, h->c, c, o, h->cold, h->w, h->ev, h->st, h->bs[F_GRAD]));
        tend(h);
        h->pv ^= 1;
    }
    if (is_export || h->c.hos_on) {
        if ((rc = number_density(h))) return rc;
    }
    HIPCHK(h, launch_courant(h->dim, n, cur(h), tab(h), h->nscene, h->w.vmax, h->st));
    h->vmax_dst = vmax_out;  // read by the drain that follows
    return MPS_OK;
}

// src/mps.cpp:554-565 per scene; returns the largest Courant number.
static double courant_of(mps_handle* h, const double* vmax) {
    double mx = 0.0;
    for (int k = 0; k < h->nscene; k++) {
        const mps_settings& s = h->scenes[k].s;
        const double cn = (vmax[k] * s.dt) / s.particle_distance;
        if (cn > s.cfl_condition) h->counts[2]++;
        h->scourant[k] = cn;
        mx = (mx < cn) ? cn : mx;
    }
    return mx;
}


// ---------------- reverse-mode step: device memory and one step's sweep ----------------

template <typename T>
static int adj_alloc(mps_handle* h, T** p, size_t count) {
    HIPCHK(h, dalloc(p, count));
    h->adj.owned.push_back((void*) *p);
    return MPS_OK;
}

static void adj_free(mps_handle* h, void* p) {
    if (!p) return;
    hipFree(p);
    auto& o = h->adj.owned;
    o.erase(std::remove(o.begin(), o.end(), p), o.end());
}

static int adj_alloc_state(mps_handle* h, double** f, int8_t** t) {
    const size_t N = (size_t) h->n;
    int rc = adj_alloc(h, f, (size_t) kStateDoubles * N);
    if (!rc) rc = adj_alloc(h, t, (size_t) kStateBytes * N);
    return rc;
}

static void adj_free_states(mps_handle* h, std::vector<double*>& f, std::vector<int8_t*>& t) {
    for (double* p : f) adj_free(h, p);
    for (int8_t* p : t) adj_free(h, p);
    f.clear();
    t.clear();
}

static int adj_alloc_bufs(mps_handle* h) {
    AdjStore& S = h->adj;
    if (S.bufs) return MPS_OK;
    const size_t N = (size_t) h->n;
    const size_t cap = (size_t) h->w.coll_cap;
    AdjBufs& a = S.a;
    int rc = MPS_OK;
    for (int d = 0; d < 3 && !rc; d++) {
        rc = adj_alloc(h, &a.xb[d], N);
        if (!rc) rc = adj_alloc(h, &a.vb[d], N);
        if (!rc) rc = adj_alloc(h, &a.gb[d], N);
        if (!rc) rc = adj_alloc(h, &a.ab[d], N);
    }
    if (!rc) rc = adj_alloc(h, &a.pb, N);
    if (!rc) rc = adj_alloc(h, &a.phb, N);
    if (!rc) rc = adj_alloc(h, &a.amin, N);
    if (!rc) rc = adj_alloc(h, &a.rhs, N);
    if (!rc) rc = adj_alloc(h, &a.lam, N);
    if (!rc) rc = adj_alloc(h, &a.cnu, N);
    if (!rc) rc = adj_alloc(h, &a.crelax, N);
    if (!rc) rc = adj_alloc(h, &a.cpair, adj_pair_scratch((int) cap));
    if (!rc) rc = adj_alloc(h, &a.part, adj_partials(h->n));
    if (!rc) rc = adj_alloc(h, &a.acc, 3);
    if (!rc) rc = adj_alloc(h, &S.type_g, N);
    if (!rc) rc = adj_alloc(h, &S.coll.pid, 2 * cap);
    if (!rc) rc = adj_alloc(h, &S.coll.flags, cap);
    if (!rc) rc = adj_alloc(h, &S.coll.pre, 12 * cap);
    if (!rc) rc = adj_alloc(h, &S.coll.dir, 3 * cap);
    if (!rc) rc = adj_alloc(h, &S.coll.nlev, 1);
    if (!rc) rc = adj_alloc(h, &S.ident, N);
    if (!rc) rc = adj_alloc(h, &S.key0, N);
    const int d3 = h->dim == 3;
    if (!rc) rc = adj_alloc(h, &S.hs.px, N);
    if (!rc) rc = adj_alloc(h, &S.hs.py, N);
    if (!rc) rc = adj_alloc(h, &S.hs.vx, N);
    if (!rc) rc = adj_alloc(h, &S.hs.vy, N);
    if (!rc && d3) rc = adj_alloc(h, &S.hs.pz, N);
    if (!rc && d3) rc = adj_alloc(h, &S.hs.vz, N);
    if (!rc) rc = adj_alloc(h, &S.hs.id, N);
    if (!rc) rc = adj_alloc(h, &S.hs.type, N);
    if (!rc) rc = adj_alloc(h, &S.hs.key, N);
    if (rc) return rc;
    if (N) {
        std::vector<int> ids(N);
        for (size_t i = 0; i < N; i++) ids[i] = (int) i;
        HIPCHK(h, hipMemcpy(S.ident, ids.data(), N * sizeof(int), hipMemcpyHostToDevice));
    }
    S.bufs = true;
    return MPS_OK;
}

static int adj_save(mps_handle* h, double* f, int8_t* t) {
    HIPCHK(h, save_state(h->dim, h->n, cur(h), h->cold, f, t, h->st));
    return MPS_OK;
}

static int adj_restore(mps_handle* h, const double* f, const int8_t* t) {
    Hot c = cur(h);
    HIPCHK(h, restore_state(h->dim, h->n, c, h->cold, h->bb.slot_of, f, t, h->st));
    return MPS_OK;
}

// Reverse of one step.  The step has just been run from the saved start
// state (f0, t0) with its hooks on, so the pressure stage's slots, boundary
// conditions, matrix, solution and collision record are those of the forward
// step.  Adds the step's transpose to the adjoint accumulators.
static int adj_reverse_step(mps_handle* h, const StepTap& tap, const double* f0, const int8_t* t0,
                            const mps_solver_options& oa, mps_solver_stats& agg) {
    AdjStore& S = h->adj;
    AdjBufs& a = S.a;
    const int n = h->n, dim = h->dim;
    const size_t N = (size_t) n;
    GridView g = gview(h);
    HIPCHK(h, adj_pressure_stage(dim, g, n, h->c, tap.c3, S.type_g, h->w, h->x_s, a, h->st));
    mps_solver_stats ss{};
    std::string msg;
    cg_prediction_bank(h->cg, 1);
    const hipError_t ce = cg_solve(h->cg, h->A, tap.c3, n, a.rhs, nullptr, nullptr, a.lam, oa, &ss, h->st, msg);
    cg_prediction_bank(h->cg, 0);
    if (ce != hipSuccess) return hip_fail(h, ce, "cg_solve (adjoint)");
    agg.rows = std::max(agg.rows, ss.rows);
    agg.iterations += ss.iterations;
    agg.refinements += ss.refinements;
    agg.converged = agg.converged && ss.converged;
    agg.rel_residual = std::max(agg.rel_residual, ss.rel_residual);
    agg.initial_residual = std::max(agg.initial_residual, ss.initial_residual);
    agg.rhs_norm = std::max(agg.rhs_norm, ss.rhs_norm);
    HIPCHK(h, adj_ppe_stage(dim, g, n, h->c, tap.c3, S.type_g, h->w, h->x_s, a, h->st));
    HIPCHK(h, adj_collision_stage(dim, n, h->c, h->w, S.coll, a, h->st));
    // the start state, rebuilt into slots, for the viscosity stage
    Hot h0{};
    h0.px = const_cast<double*>(f0);
    h0.py = const_cast<double*>(f0) + N;
    h0.vx = const_cast<double*>(f0) + 3 * N;
    h0.vy = const_cast<double*>(f0) + 4 * N;
    if (dim == 3) {
        h0.pz = const_cast<double*>(f0) + 2 * N;
        h0.vz = const_cast<double*>(f0) + 5 * N;
    }
    h0.id = S.ident;
    h0.type = const_cast<int8_t*>(t0);
    h0.key = S.key0;
    HIPCHK(h, build_grid(dim, h->g, tab(h), n, h0, S.hs, h->bb, h->ev.err, h->st, h->bs[F_BUILD]));
    HIPCHK(h, adj_explicit_stage(dim, gview(h), n, h->c, t0, S.hs, a, h->st));
    return MPS_OK;
}

static int adj_backward_impl(mps_handle* h, const double* pos_bar, const double* vel_bar, mps_solver_stats& agg) {
    AdjStore& S = h->adj;
    const int n = h->n;
    const size_t N = (size_t) n;
    const int M = S.interval;
    int rc = adj_alloc_bufs(h);
    if (rc) return rc;
    if ((int) S.segf.size() < M) {
        adj_free_states(h, S.segf, S.segt);
        for (int k = 0; k < M; k++) {
            double* f;
            int8_t* t;
            if ((rc = adj_alloc_state(h, &f, &t))) return rc;
            S.segf.push_back(f);
            S.segt.push_back(t);
        }
    }
    // seeds dL/dx(T), dL/dv(T)
    std::vector<double> comp(N);
    for (int d = 0; d < 3; d++) {
        for (size_t i = 0; i < N; i++) comp[i] = pos_bar ? pos_bar[3 * i + d] : 0.0;
        if (N) HIPCHK(h, hipMemcpy(S.a.xb[d], comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        for (size_t i = 0; i < N; i++) comp[i] = vel_bar ? vel_bar[3 * i + d] : 0.0;
        if (N) HIPCHK(h, hipMemcpy(S.a.vb[d], comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(h, fill_double(S.a.acc, 0.0, 3, h->st));
    mps_solver_options oa = S.opt;
    oa.warm_start = 0;
    agg = mps_solver_stats{};
    agg.converged = 1;
    const int nseg = (S.steps + M - 1) / M;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int k0 = sg * M, k1 = std::min(S.steps, k0 + M);
        if ((rc = adj_restore(h, S.cpf[sg], S.cpt[sg]))) return rc;
        for (int k = k0; k < k1; k++) {
            if ((rc = adj_save(h, S.segf[k - k0], S.segt[k - k0]))) return rc;
            if (k + 1 < k1) {
                if ((rc = step_core(h, 0, h->pinned_d))) return rc;
                if ((rc = drain(h))) return rc;
            }
        }
        for (int k = k1 - 1; k >= k0; k--) {
            if ((rc = adj_restore(h, S.segf[k - k0], S.segt[k - k0]))) return rc;
            StepTap tap{&S.coll, S.type_g, Hot{}};
            if ((rc = step_core(h, 0, h->pinned_d, &tap))) return rc;
            if ((rc = drain(h))) return rc;
            if ((rc = adj_reverse_step(h, tap, S.segf[k - k0], S.segt[k - k0], oa, agg))) return rc;
        }
    }
    return MPS_OK;
}

// ---------------- C ABI ----------------

extern "C" {

const char* mps_last_error(const mps_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static bool valid_settings(const mps_settings* s, std::string& err) {
    if (s->dim != 2 && s->dim != 3) {
        err = "dim must be 2 or 3";
        return false;
    }
    if (!(s->re_max > 0) || s->re_max < s->re_pressure || s->re_max < s->re_viscosity || s->re_max < s->re_surface ||
        !(s->re_pressure > 0) || !(s->re_viscosity > 0) || !(s->re_surface > 0) || !(s->particle_distance > 0) ||
        !(s->dt > 0) || !(s->density > 0)) {
        err = "invalid settings (radii, particle distance, dt or density)";
        return false;
    }
    return true;
}

// Creates a handle for nscene independent scenes packed one after another:
// scene k holds particles [off_k, off_k + counts[k]) with local ids 0..n_k-1
// in order; ids[] carries the local ids.
static int create_impl(int32_t nscene, const mps_settings* sset, const int32_t* counts, const int32_t* id,
                       const int8_t* type, const double* pos, const double* vel, mps_handle** out) {
    if (!out) {
        g_create_error = "out is NULL";
        return MPS_E_INVALID;
    }
    *out = nullptr;
    if (nscene < 1 || nscene > 32767 || !sset || !counts) {
        g_create_error = "invalid arguments";
        return MPS_E_INVALID;
    }
    int64_t ntot = 0;
    for (int k = 0; k < nscene; k++) {
        if (counts[k] < 0) {
            g_create_error = "invalid arguments";
            return MPS_E_INVALID;
        }
        ntot += counts[k];
    }
    if (ntot > 0x7fffffff / 4) {  // byte offsets of 8-byte fields must fit 32 bits
        g_create_error = "too many particles for one handle (limit 536870911)";
        return MPS_E_INVALID;
    }
    if (ntot > 0 && (!id || !type || !pos || !vel)) {
        g_create_error = "invalid arguments";
        return MPS_E_INVALID;
    }
    const int n = (int) ntot;
    for (int k = 0; k < nscene; k++) {
        std::string err;
        if (!valid_settings(&sset[k], err)) {
            g_create_error = err;
            return MPS_E_INVALID;
        }
        if (sset[k].dim != sset[0].dim) {
            g_create_error = "every scene of a handle must have the same dim";
            return MPS_E_INVALID;
        }
    }
    mps_handle* h = new mps_handle();
    h->n = n;
    h->dim = sset[0].dim;
    h->nscene = nscene;
    h->scenes.resize(nscene);
    h->sstats.assign(nscene, mps_solver_stats{});
    h->scourant.assign(nscene, 0.0);
    int off = 0, key0 = 0;
    for (int k = 0; k < nscene; k++) {
        SceneInfo& sc = h->scenes[k];
        const mps_settings* s = &sset[k];
        sc.s = *s;
        sc.id0 = off;
        sc.n = counts[k];
        for (int32_t i = 0; i < sc.n; i++) {
            const size_t gi = (size_t) off + i;
            if (id[gi] != i) {
                g_create_error = "particle ids must be 0..N-1 in order within each scene";
                delete h;
                return MPS_E_INVALID;
            }
            if (type[gi] < -1 || type[gi] > 2) {
                g_create_error = "particle type out of range";
                delete h;
                return MPS_E_INVALID;
            }
        }
        if (h->dim == 2) {
            // In 2D every non-ghost particle must sit on one z plane with zero
            // z velocity and zero z gravity: the reference would ghost any other
            // particle at its first move (src/mps.cpp:651-653).
            if (s->gravity[2] != 0.0) {
                g_create_error = "2D requires gravity z = 0";
                delete h;
                return MPS_E_INVALID;
            }
            bool have = false;
            double z0 = 0.0;
            for (int32_t i = 0; i < sc.n; i++) {
                const size_t gi = (size_t) off + i;
                if (type[gi] == MPS_GHOST) continue;
                if (vel[3 * gi + 2] != 0.0 || (have && pos[3 * gi + 2] != z0)) {
                    g_create_error = "2D requires every non-ghost particle on one z plane with zero z velocity";
                    delete h;
                    return MPS_E_INVALID;
                }
                z0 = pos[3 * gi + 2];
                have = true;
            }
            if (have) {
                const double q = (z0 - s->domain_min[2]) / s->re_max;
                if (!(q > -1.0 && q < 1e8)) {
                    g_create_error = "z plane outside the bucket grid";
                    delete h;
                    return MPS_E_INVALID;
                }
                sc.iz0 = (int) q + 1;
            }
        }
        scene_consts(*s, sc.c);
        std::string err;
        if (!scene_grid(*s, h->dim, sc.g, err)) {
            g_create_error = err;
            delete h;
            return MPS_E_INVALID;
        }
        if (h->dim == 2 && (sc.iz0 < 1 || sc.iz0 > sc.g.nz - 2)) {
            g_create_error = "z plane outside the bucket grid";
            delete h;
            return MPS_E_INVALID;
        }
        sc.g.key0 = key0;
        if ((int64_t) key0 + sc.g.ncell > 0x7fffffff / 2) {
            g_create_error = "bucket grids too large";
            delete h;
            return MPS_E_INVALID;
        }
        key0 += sc.g.ncell;
        off += sc.n;
    }
    for (auto& sc : h->scenes) sc.g.ghost_key = key0;
    h->s = h->scenes[0].s;
    h->c = h->scenes[0].c;
    h->g = h->scenes[0].g;
    h->iz0 = h->scenes[0].iz0;
    if (nscene > 1) {
        // the launch-level flags cover every scene; kernels read each scene's own
        for (auto& sc : h->scenes) h->c.hos_on |= sc.c.hos_on;
    }
    hipError_t e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete h;
        return MPS_E_HIP;
    }
    if ((e = scratch_setup(&h->dev)) != hipSuccess) {
        g_create_error = std::string("device scratch lookup (hipGetSymbolAddress): ") + hipGetErrorString(e);
        hipStreamDestroy(h->st);
        delete h;
        return MPS_E_HIP;
    }
    int rc = alloc_all(h);
    if (rc) {
        g_create_error = h->err;
        free_all(h);
        delete h;
        return rc;
    }
    // scene tables
    {
        std::vector<int> k0(nscene + 1);
        for (int k = 0; k < nscene; k++) k0[k] = h->scenes[k].g.key0;
        k0[nscene] = key0;
        e = hipMemcpy(h->d_key0, k0.data(), sizeof(int) * (size_t) (nscene + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess && nscene > 1) {
            std::vector<GridView> gs(nscene);
            std::vector<Consts> cs(nscene);
            for (int k = 0; k < nscene; k++) {
                gs[k] = h->scenes[k].g;
                gs[k].cs = h->bb.cs;
                cs[k] = h->scenes[k].c;
            }
            e = hipMemcpy(h->d_g, gs.data(), sizeof(GridView) * (size_t) nscene, hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(h->d_c, cs.data(), sizeof(Consts) * (size_t) nscene, hipMemcpyHostToDevice);
            std::vector<int16_t> sid((size_t) n);
            for (int k = 0; k < nscene; k++)
                for (int i = 0; i < h->scenes[k].n; i++) sid[(size_t) h->scenes[k].id0 + i] = (int16_t) k;
            if (e == hipSuccess && n > 0)
                e = hipMemcpy(h->d_sid, sid.data(), sizeof(int16_t) * (size_t) n, hipMemcpyHostToDevice);
        }
    }
    h->A.nseg = nscene;
    h->A.scene_key0 = h->d_key0;
    h->A.cs = h->bb.cs;
    h->A.row_of = h->w.row_of;
    // upload (slots = global ids until the first build)
    const size_t N = (size_t) n;
    std::vector<double> comp(N);
    const int fields = h->dim == 3 ? 3 : 2;
    for (int f = 0; f < fields && e == hipSuccess; f++) {
        for (size_t i = 0; i < N; i++) comp[i] = pos[3 * i + f];
        e = hipMemcpy(h->P[0][f], comp.data(), N * sizeof(double), hipMemcpyHostToDevice);
        for (size_t i = 0; i < N && e == hipSuccess; i++) comp[i] = vel[3 * i + f];
        if (e == hipSuccess) e = hipMemcpy(h->P[0][3 + f], comp.data(), N * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        for (size_t i = 0; i < N; i++) comp[i] = pos[3 * i + 2];
        e = hipMemcpy(h->cold.z, comp.data(), N * sizeof(double), hipMemcpyHostToDevice);
        for (size_t i = 0; i < N && e == hipSuccess; i++) comp[i] = vel[3 * i + 2];
        if (e == hipSuccess) e = hipMemcpy(h->cold.vz, comp.data(), N * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n > 0) {
        std::vector<int32_t> ids(N);
        for (size_t i = 0; i < N; i++) ids[i] = (int32_t) i;
        e = hipMemcpy(h->ID[0], ids.data(), N * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->bb.slot_of, ids.data(), N * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->TY[0], type, N, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        g_create_error = std::string("upload: ") + hipGetErrorString(e);
        free_all(h);
        delete h;
        return MPS_E_HIP;
    }
    // reference-faithful solver defaults: FP64, tolerance = machine epsilon,
    // at most 2 x rows iterations, zero initial guess (Eigen defaults)
    h->opt.precision = MPS_PREC_FP64;
    h->opt.warm_start = 0;
    h->opt.max_iterations = 0;
    h->opt.max_refinements = 0;
    h->opt.tolerance = DBL_EPSILON;
    h->opt.inner_tolerance = 1e-4;
    h->opt.check_interval = 32;
    h->opt.use_graph = 1;
    h->opt.preconditioner = MPS_PRECOND_JACOBI;
    h->A.dim = h->dim;
    h->sgrids.resize((size_t) nscene);
    for (int k = 0; k < nscene; k++) h->sgrids[k] = h->scenes[k].g;
    h->A.scene_grids = h->sgrids.data();
    *out = h;
    return MPS_OK;
}

int mps_create(const mps_settings* s, int32_t n, const int32_t* id, const int8_t* type, const double* pos,
               const double* vel, mps_handle** out) {
    if (n < 0) {
        g_create_error = "invalid arguments";
        return MPS_E_INVALID;
    }
    return create_impl(1, s, &n, id, type, pos, vel, out);
}

int mps_create_ensemble(int32_t nscene, const mps_settings* s, const int32_t* counts, const int32_t* id,
                        const int8_t* type, const double* pos, const double* vel, mps_handle** out) {
    return create_impl(nscene, s, counts, id, type, pos, vel, out);
}

int mps_scene_count(const mps_handle* h) { return h ? h->nscene : 0; }

int mps_ensemble_stats(const mps_handle* h, mps_solver_stats* out, double* courant) {
    if (!h) return MPS_E_INVALID;
    for (int k = 0; k < h->nscene; k++) {
        if (out) out[k] = h->sstats[k];
        if (courant) courant[k] = h->scourant[k];
    }
    return MPS_OK;
}

int mps_destroy(mps_handle* h) {
    if (!h) return MPS_OK;
    if (h->st) hipStreamSynchronize(h->st);
    free_all(h);
    delete h;
    return MPS_OK;
}

int mps_set_solver(mps_handle* h, const mps_solver_options* o) {
    if (!h || !o) return MPS_E_INVALID;
    if (o->precision != MPS_PREC_FP64 && o->precision != MPS_PREC_MIXED)
        return fail(h, MPS_E_INVALID, "unknown precision");
    if (!(o->tolerance > 0)) return fail(h, MPS_E_INVALID, "tolerance must be positive");
    if (o->preconditioner != MPS_PRECOND_JACOBI && o->preconditioner != MPS_PRECOND_MULTIGRID)
        return fail(h, MPS_E_INVALID, "unknown preconditioner");
    if (o->warm_start < 0 || o->warm_start > 2) return fail(h, MPS_E_INVALID, "warm_start must be 0, 1 or 2");
    h->opt = *o;
    if (h->opt.check_interval < 1) h->opt.check_interval = 1;
    return MPS_OK;
}

int mps_get_solver_stats(const mps_handle* h, mps_solver_stats* out) {
    if (!h || !out) return MPS_E_INVALID;
    // one scene: its statistics; several: iterations and refinements summed,
    // residuals the worst over the scenes
    mps_solver_stats t = h->sstats[0];
    for (int k = 1; k < h->nscene; k++) {
        const mps_solver_stats& x = h->sstats[k];
        t.rows += x.rows;
        t.iterations += x.iterations;
        t.refinements += x.refinements;
        t.converged = t.converged && x.converged;
        t.rhs_norm = sqrt(t.rhs_norm * t.rhs_norm + x.rhs_norm * x.rhs_norm);
        t.rel_residual = t.rel_residual > x.rel_residual ? t.rel_residual : x.rel_residual;
        t.initial_residual = t.initial_residual > x.initial_residual ? t.initial_residual : x.initial_residual;
    }
    *out = t;
    return MPS_OK;
}

int mps_load_initial_state(mps_handle* h) {
    if (!h) return MPS_E_INVALID;
    int rc = number_density(h);
    if (rc) return rc;
    return drain(h);
}

// Message of a solve that stopped above its tolerance, naming the scenes.
static std::string solve_failure(const mps_handle* h) {
    std::string m = h->solve_msg;
    if (h->nscene > 1) {
        m += " (scenes";
        for (int k = 0; k < h->nscene; k++)
            if (!h->sstats[k].converged) m += " " + std::to_string(k);
        m += ")";
    }
    char buf[96];
    double worst = 0.0;
    for (int k = 0; k < h->nscene; k++) worst = std::max(worst, h->sstats[k].rel_residual);
    std::snprintf(buf, sizeof buf, ": relative residual %.3e > %.3e", worst, h->opt.tolerance);
    return m + buf;
}

int mps_step(mps_handle* h, int32_t is_export, double* courant) {
    if (!h) return MPS_E_INVALID;
    int rc = step_core(h, is_export, h->pinned_d);
    if (rc) return rc;
    rc = drain(h);
    if (rc) return rc;
    const double cn = courant_of(h, h->pinned_d);
    if (courant) *courant = cn;
    if (!h->solve_msg.empty()) return fail(h, MPS_E_SOLVER, solve_failure(h));
    return MPS_OK;
}

int mps_run_steps(mps_handle* h, int32_t n_steps, double* courants) {
    if (!h || n_steps < 0) return MPS_E_INVALID;
    int first_bad = -1;
    std::string bad;
    for (int32_t k = 0; k < n_steps; k++) {
        int rc = step_core(h, 0, h->pinned_d);
        if (rc) return rc;
        if ((rc = drain(h))) return rc;
        const double cn = courant_of(h, h->pinned_d);
        if (courants) courants[k] = cn;
        if (!h->solve_msg.empty() && first_bad < 0) {
            first_bad = k;
            bad = solve_failure(h);
        }
    }
    if (first_bad >= 0) return fail(h, MPS_E_SOLVER, "step " + std::to_string(first_bad) + ": " + bad);
    return MPS_OK;
}

static int gather_hot(const mps_handle* hc, int8_t* type, double* pos, double* vel) {
    mps_handle* h = const_cast<mps_handle*>(hc);
    const size_t N = (size_t) h->n;
    if (N == 0) return MPS_OK;
    Hot c = cur(h);
    if (pos || vel) {
        HIPCHK(h, launch_snapshot(h->dim, h->n, c, h->cold, h->tmp3a, h->tmp3b, h->st));
        if (pos) HIPCHK(h, hipMemcpyAsync(pos, h->tmp3a, 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        if (vel) HIPCHK(h, hipMemcpyAsync(vel, h->tmp3b, 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->st));
    }
    if (type) {
        std::vector<int8_t> ts(N);
        std::vector<int32_t> ids(N);
        HIPCHK(h, hipMemcpyAsync(ts.data(), c.type, N, hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(ids.data(), c.id, N * sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
        for (size_t s = 0; s < N; s++) type[ids[s]] = ts[s];
    }
    HIPCHK(h, hipStreamSynchronize(h->st));
    return MPS_OK;
}

int mps_download(const mps_handle* hc, int8_t* type, double* pos, double* vel, double* pressure, double* ndr,
                 int8_t* bc) {
    if (!hc) return MPS_E_INVALID;
    mps_handle* h = const_cast<mps_handle*>(hc);
    int rc = gather_hot(h, type, pos, vel);
    if (rc) return rc;
    const size_t N = (size_t) h->n;
    if (N == 0) return MPS_OK;
    if (pressure) HIPCHK(h, hipMemcpyAsync(pressure, h->cold.p, N * sizeof(double), hipMemcpyDeviceToHost, h->st));
    if (ndr) HIPCHK(h, hipMemcpyAsync(ndr, h->cold.ndr, N * sizeof(double), hipMemcpyDeviceToHost, h->st));
    if (bc) HIPCHK(h, hipMemcpyAsync(bc, h->cold.bc, N, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return MPS_OK;
}

int mps_download_min_pressure(const mps_handle* hc, double* pmin) {
    if (!hc || !pmin) return MPS_E_INVALID;
    mps_handle* h = const_cast<mps_handle*>(hc);
    if (h->n == 0) return MPS_OK;
    HIPCHK(h, hipMemcpyAsync(pmin, h->cold.pmin, (size_t) h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    return MPS_OK;
}

int mps_upload(mps_handle* h, const int8_t* type, const double* pos, const double* vel, const double* pressure,
               const double* ndr, const int8_t* bc) {
    if (!h) return MPS_E_INVALID;
    const size_t N = (size_t) h->n;
    if (N == 0) return MPS_OK;
    // host copies of slot_of and the current hot arrays, rewritten in slot order
    std::vector<int32_t> slot(N);
    Hot c = cur(h);
    HIPCHK(h, slot_of_ids(h->n, c.id, h->bb.slot_of, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(slot.data(), h->bb.slot_of, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (type) {
        for (size_t i = 0; i < N; i++)
            if (type[i] < -1 || type[i] > 2) return fail(h, MPS_E_INVALID, "particle type out of range");
        std::vector<int8_t> ts(N);
        for (size_t i = 0; i < N; i++) ts[slot[i]] = type[i];
        HIPCHK(h, hipMemcpy(c.type, ts.data(), N, hipMemcpyHostToDevice));
    }
    if (h->dim == 2 && (pos || vel)) {
        // keep the 2D plane invariant (see mps_create)
        std::vector<int8_t> ts(N);
        std::vector<int8_t> tid(N);
        HIPCHK(h, hipMemcpy(ts.data(), c.type, N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) tid[i] = ts[slot[i]];
        std::vector<double> zc(N);
        HIPCHK(h, hipMemcpy(zc.data(), h->cold.z, N * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) {
            if (tid[i] == MPS_GHOST) continue;
            if ((pos && pos[3 * i + 2] != zc[i]) || (vel && vel[3 * i + 2] != 0.0))
                return fail(h, MPS_E_INVALID, "2D upload must keep z and z velocity of non-ghost particles");
        }
    }
    std::vector<double> comp(N);
    const int fields = h->dim == 3 ? 3 : 2;
    double* dstp[6] = {c.px, c.py, c.pz, c.vx, c.vy, c.vz};
    for (int f = 0; f < fields; f++) {
        if (pos) {
            for (size_t i = 0; i < N; i++) comp[slot[i]] = pos[3 * i + f];
            HIPCHK(h, hipMemcpy(dstp[f], comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        }
        if (vel) {
            for (size_t i = 0; i < N; i++) comp[slot[i]] = vel[3 * i + f];
            HIPCHK(h, hipMemcpy(dstp[3 + f], comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    if (h->dim == 2) {
        if (pos) {
            for (size_t i = 0; i < N; i++) comp[i] = pos[3 * i + 2];
            HIPCHK(h, hipMemcpy(h->cold.z, comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        }
        if (vel) {
            for (size_t i = 0; i < N; i++) comp[i] = vel[3 * i + 2];
            HIPCHK(h, hipMemcpy(h->cold.vz, comp.data(), N * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    if (pressure) {
        HIPCHK(h, hipMemcpy(h->cold.p, pressure, N * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->cold.pprev, pressure, N * sizeof(double), hipMemcpyHostToDevice));
        // no earlier solution to extrapolate from
        HIPCHK(h, fill_double(h->cold.pprev2, 0.0, N, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
    }
    if (ndr) HIPCHK(h, hipMemcpy(h->cold.ndr, ndr, N * sizeof(double), hipMemcpyHostToDevice));
    if (bc) {
        for (size_t i = 0; i < N; i++)
            if (bc[i] < 0 || bc[i] > 2) return fail(h, MPS_E_INVALID, "boundary condition out of range");
        HIPCHK(h, hipMemcpy(h->cold.bc, bc, N, hipMemcpyHostToDevice));
    }
    return MPS_OK;
}

int mps_debug_events(const mps_handle* hc, int64_t counts[3], int32_t* ghost_ids, int32_t* unconnected_ids,
                     int32_t cap) {
    if (!hc || !counts) return MPS_E_INVALID;
    mps_handle* h = const_cast<mps_handle*>(hc);
    int rc = drain(h);
    if (rc) return rc;
    for (int k = 0; k < 3; k++) counts[k] = h->counts[k];
    std::vector<int32_t> a = h->ghost_ids, b = h->unconn_ids;
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    if (ghost_ids)
        for (size_t k = 0; k < a.size() && (int32_t) k < cap; k++) ghost_ids[k] = a[k];
    if (unconnected_ids)
        for (size_t k = 0; k < b.size() && (int32_t) k < cap; k++) unconnected_ids[k] = b[k];
    return MPS_OK;
}

int mps_debug_counters(const mps_handle* h, int64_t* out, int32_t cap) {
    if (!h || (cap > 0 && !out)) return MPS_E_INVALID;
    long long cg[4], stg[4] = {0, 0, 0, 0};
    cg_debug_counts(h->cg, cg);
    if (cap > 3) {
        const hipError_t e = stage_stats_read(stg);
        if (e != hipSuccess) return hip_fail(const_cast<mps_handle*>(h), e, "mps_debug_counters");
    }
    const int64_t v[9] = {cg[0], cg[1], h->A.grows, stg[0], stg[1], stg[2], stg[3], cg[2], cg[3]};
    for (int k = 0; k < cap && k < 9; k++) out[k] = v[k];
    return MPS_OK;
}

int mps_clear_debug_events(mps_handle* h) {
    if (!h) return MPS_E_INVALID;
    int rc = drain(h);
    if (rc) return rc;
    h->ghost_ids.clear();
    h->unconn_ids.clear();
    h->counts[0] = h->counts[1] = h->counts[2] = 0;
    return MPS_OK;
}

int mps_bucket_shape(const mps_handle* h, int32_t dims[4], double* length) {
    if (!h || !dims) return MPS_E_INVALID;
    dims[0] = h->g.nx * h->g.ny * h->g.nz;
    dims[1] = h->g.nx;
    dims[2] = h->g.ny;
    dims[3] = h->g.nz;
    if (length) *length = h->g.len;
    return MPS_OK;
}

int mps_dump_bucket(mps_handle* h, int32_t* first, int32_t* last, int32_t* next) {
    if (!h || !first || !last || !next) return MPS_E_INVALID;
    if (h->nscene != 1) return fail(h, MPS_E_INVALID, "mps_dump_bucket takes a single-scene handle");
    int rc = do_build(h);
    if (rc) return rc;
    if ((rc = drain(h))) return rc;
    const size_t N = (size_t) h->n;
    const size_t nc = (size_t) h->g.ncell;
    std::vector<int> cs(nc + 2);
    std::vector<int32_t> ids(N);
    HIPCHK(h, hipMemcpy(cs.data(), h->bb.cs, (nc + 2) * sizeof(int), hipMemcpyDeviceToHost));
    if (N) HIPCHK(h, hipMemcpy(ids.data(), cur(h).id, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    const size_t num = (size_t) h->g.nx * h->g.ny * h->g.nz;
    for (size_t k = 0; k < num; k++) first[k] = last[k] = -1;
    for (size_t i = 0; i < N; i++) next[i] = -1;
    const size_t layer = h->g.flat && h->dim == 2 ? (size_t) h->iz0 * h->g.nx * h->g.ny
                                                  : (h->g.flat ? (size_t) h->g.nx * h->g.ny : 0);
    for (size_t k = 0; k < nc; k++) {
        const int b = cs[k], e = cs[k + 1];
        if (b == e) continue;
        const size_t rk = k + layer;  // the reference's key: ix + iy*numX + iz*numX*numY
        first[rk] = ids[b];
        last[rk] = ids[e - 1];
        for (int q = b; q + 1 < e; q++) next[ids[q]] = ids[q + 1];
    }
    return MPS_OK;
}

int mps_dump_neighbors(mps_handle* h, int32_t* offsets, int32_t* ids, double* dists) {
    if (!h || !offsets) return MPS_E_INVALID;
    int rc = do_build(h);
    if (rc) return rc;
    if ((rc = drain(h))) return rc;
    const int n = h->n;
    GridView g = gview(h);
    Hot c = cur(h);
    int* cnt = h->w.coll_cnt;
    int* off = h->w.coll_off;
    HIPCHK(h, dump_neighbor_counts(h->dim, g, tab(h), n, c, cnt, h->s.re_max, h->st));
    HIPCHK(h, scan_exclusive(cnt, off, n, h->bb.scan_tmp, h->st));
    HIPCHK(h, hipMemcpyAsync(offsets, off, ((size_t) n + 1) * sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (!ids && !dists) return MPS_OK;
    const size_t tot = (size_t) offsets[n];
    int32_t* d_ids = nullptr;
    double* d_dist = nullptr;
    HIPCHK(h, dalloc(&d_ids, tot));
    hipError_t e = dalloc(&d_dist, tot);
    if (e == hipSuccess) e = dump_neighbor_lists(h->dim, g, tab(h), n, c, off, d_ids, d_dist, h->s.re_max, h->st);
    if (e == hipSuccess && ids && tot)
        e = hipMemcpyAsync(ids, d_ids, tot * sizeof(int32_t), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess && dists && tot)
        e = hipMemcpyAsync(dists, d_dist, tot * sizeof(double), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    hipFree(d_ids);
    hipFree(d_dist);
    if (e != hipSuccess) return hip_fail(h, e, "mps_dump_neighbors");
    return MPS_OK;
}

int mps_dump_neighbors_noalloc(mps_handle* h, int32_t* offsets, int32_t* ids, double* dists) {
    if (!h || !offsets) return MPS_E_INVALID;
    void* const base = h->dev >= 0 && h->dev < kMaxDevices ? g_scratch_dev[h->dev] : nullptr;
    if (!base) return fail(h, MPS_E_INVALID, "device scratch was not set up for this handle's device");
    int rc = do_build(h);
    if (rc) return rc;
    if ((rc = drain(h))) return rc;
    const int n = h->n;
    GridView g = gview(h);
    Hot c = cur(h);
    int* cnt = h->w.coll_cnt;
    int* off = h->w.coll_off;
    HIPCHK(h, dump_neighbor_counts(h->dim, g, tab(h), n, c, cnt, h->s.re_max, h->st));
    HIPCHK(h, scan_exclusive(cnt, off, n, h->bb.scan_tmp, h->st));
    HIPCHK(h, hipMemcpyAsync(offsets, off, ((size_t) n + 1) * sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (!ids && !dists) return MPS_OK;
    const size_t tot = (size_t) offsets[n];
    if (tot > kNoAllocMaxEntries)
        return fail(h, MPS_E_INVALID,
                    "mps_dump_neighbors_noalloc: " + std::to_string(tot) + " neighbour entries exceed the " +
                        std::to_string(kNoAllocScratchBytes >> 20) + " MiB device scratch (at most " +
                        std::to_string(kNoAllocMaxEntries) + " entries); use mps_dump_neighbors");
    // ids in the first 16 MiB, distances in the 32 MiB after them
    int32_t* d_ids = reinterpret_cast<int32_t*>(base);
    double* d_dist = reinterpret_cast<double*>(static_cast<char*>(base) + kNoAllocMaxEntries * sizeof(int32_t));
    hipError_t e = dump_neighbor_lists(h->dim, g, tab(h), n, c, off, d_ids, d_dist, h->s.re_max, h->st);
    if (e == hipSuccess && ids && tot)
        e = hipMemcpyAsync(ids, d_ids, tot * sizeof(int32_t), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess && dists && tot)
        e = hipMemcpyAsync(dists, d_dist, tot * sizeof(double), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) return hip_fail(h, e, "mps_dump_neighbors_noalloc");
    return MPS_OK;
}

int mps_get_pressure_system(const mps_handle* hc, double* b, double* diag, int32_t* row_ptr, int32_t* col,
                            double* val, double* x) {
    if (!hc) return MPS_E_INVALID;
    mps_handle* h = const_cast<mps_handle*>(hc);
    if (!h->have_system) return fail(h, MPS_E_INVALID, "no captured pressure system: enable capture and step");
    const size_t N = (size_t) h->n;
    std::vector<int32_t> sid(N);
    std::vector<double> bs(N), ds(N), xs(N);
    std::vector<int8_t> bcs(N);
    std::vector<int> ro(N + 1);
    HIPCHK(h, hipMemcpy(sid.data(), h->dbg_id, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(bs.data(), h->w.b_s, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ds.data(), h->w.diag_s, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(xs.data(), h->x_s, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(bcs.data(), h->w.bc_s, N, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ro.data(), h->A.raw_off, (N + 1) * sizeof(int), hipMemcpyDeviceToHost));
    const size_t nnz = (size_t) ro[N];
    std::vector<int> rc(nnz);
    std::vector<double> rv(nnz);
    if (nnz) {
        HIPCHK(h, hipMemcpy(rc.data(), h->A.raw_col, nnz * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(rv.data(), h->A.raw_val, nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<int32_t> cnt(N, 0);
    for (size_t s = 0; s < N; s++) {
        const int32_t i = sid[s];
        if (b) b[i] = bs[s];
        if (diag) diag[i] = ds[s];
        if (x) x[i] = bcs[s] == MPS_BC_INNER ? xs[s] : 0.0;
        cnt[i] = ro[s + 1] - ro[s];
    }
    if (row_ptr) {
        std::vector<int32_t> rp(N + 1, 0);
        for (size_t i = 0; i < N; i++) rp[i + 1] = rp[i] + cnt[i];
        memcpy(row_ptr, rp.data(), (N + 1) * sizeof(int32_t));
        if (col || val) {
            std::vector<std::pair<int32_t, double>> row;
            for (size_t s = 0; s < N; s++) {
                const int32_t i = sid[s];
                row.clear();
                for (int k = ro[s]; k < ro[s + 1]; k++) row.emplace_back(sid[rc[k]], rv[k]);
                std::sort(row.begin(), row.end(),
                          [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b2) {
                              return a.first < b2.first;
                          });
                for (size_t k = 0; k < row.size(); k++) {
                    if (col) col[rp[i] + k] = row[k].first;
                    if (val) val[rp[i] + k] = row[k].second;
                }
            }
        }
    }
    return MPS_OK;
}

int mps_set_snapshot(mps_handle* h, int32_t on) {
    if (!h) return MPS_E_INVALID;
    const size_t N = (size_t) h->n;
    if (on && !h->snap_pos) {
        HIPCHK(h, dalloc(&h->snap_pos, 3 * N));
        HIPCHK(h, dalloc(&h->snap_vel, 3 * N));
    }
    h->want_snap = on ? 1 : 0;
    h->want_raw = on ? 1 : 0;
    return MPS_OK;
}

int mps_get_snapshot(const mps_handle* hc, double* pos, double* vel) {
    if (!hc) return MPS_E_INVALID;
    mps_handle* h = const_cast<mps_handle*>(hc);
    if (!h->snap_pos) return fail(h, MPS_E_INVALID, "snapshot not enabled");
    const size_t N = (size_t) h->n;
    if (pos) HIPCHK(h, hipMemcpy(pos, h->snap_pos, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    if (vel) HIPCHK(h, hipMemcpy(vel, h->snap_vel, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    return MPS_OK;
}

// Project-written autotuner: for each candidate block size, run real steps
// with every per-particle kernel family at that size and HIP events around
// each family; then give each family its fastest size.  The state advances
// by candidates x steps_per_candidate steps exactly as untuned steps would:
// no result depends on these block sizes.
int mps_autotune(mps_handle* h, int32_t steps_per_candidate, int32_t block_sizes_out[9]) {
    if (!h || steps_per_candidate < 1) return MPS_E_INVALID;
    const int cand[4] = {64, 128, 256, 512};
    double best_t[F_COUNT];
    int best_b[F_COUNT];
    for (int f = 0; f < F_COUNT; f++) {
        best_t[f] = 1e300;
        best_b[f] = 256;
    }
    for (int k = 0; k < 16; k++)
        for (int j = 0; j < 2; j++)
            if (!h->tev[k][j]) HIPCHK(h, hipEventCreate(&h->tev[k][j]));
    for (int ci = 0; ci < 4; ci++) {
        for (int f = 0; f < F_COUNT; f++) h->bs[f] = cand[ci];
        double acc[F_COUNT] = {};
        for (int k = 0; k < steps_per_candidate; k++) {
            h->timing = true;
            h->nsites = 0;
            int rc = step_core(h, 0, h->pinned_d);
            h->timing = false;
            if (rc) return rc;
            if ((rc = drain(h))) return rc;
            courant_of(h, h->pinned_d);
            for (int sidx = 0; sidx < h->nsites; sidx++) {
                float ms = 0.f;
                HIPCHK(h, hipEventElapsedTime(&ms, h->tev[sidx][0], h->tev[sidx][1]));
                acc[h->tfam[sidx]] += ms;
            }
        }
        for (int f = 0; f < F_COUNT; f++)
            if (acc[f] > 0 && acc[f] < best_t[f]) {
                best_t[f] = acc[f];
                best_b[f] = cand[ci];
            }
    }
    for (int f = 0; f < F_COUNT; f++) {
        h->bs[f] = best_b[f];
        if (block_sizes_out) block_sizes_out[f] = best_b[f];
    }
    return MPS_OK;
}

int mps_set_block_sizes(mps_handle* h, const int32_t block_sizes[9]) {
    if (!h || !block_sizes) return MPS_E_INVALID;
    for (int f = 0; f < F_COUNT; f++)
        if (block_sizes[f] < 64 || block_sizes[f] > kMaxTuneBlock || block_sizes[f] % 64)
            return fail(h, MPS_E_INVALID, "block sizes must be multiples of 64 in [64, 512]");
    for (int f = 0; f < F_COUNT; f++) h->bs[f] = block_sizes[f];
    return MPS_OK;
}

int mps_time_pressure_kernels(mps_handle* h, int32_t reps, double out_ms[4], int64_t out_bytes[4]) {
    if (!h || !out_ms || !out_bytes || reps < 1) return MPS_E_INVALID;
    // off-diagonal count of the system: scan of the per-slot counts
    HIPCHK(h, scan_exclusive(h->w.nnz_s, h->w.coll_off, h->n, h->bb.scan_tmp, h->st));
    int nnz = 0;
    HIPCHK(h, hipMemcpyAsync(h->pinned_i + 8, h->w.coll_off + h->n, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    nnz = h->pinned_i[8];
    long long bytes[4];
    HIPCHK(h, cg_time_kernels(h->cg, h->A, reps, nnz, h->st, out_ms, bytes));
    for (int k = 0; k < 4; k++) out_bytes[k] = bytes[k];
    return MPS_OK;
}

int mps_adjoint_record(mps_handle* h, int32_t steps, int32_t checkpoint_interval, double* courants) {
    if (!h || steps < 1 || checkpoint_interval < 1) return MPS_E_INVALID;
    if (h->nscene != 1) return fail(h, MPS_E_INVALID, "the adjoint takes a single-scene handle");
    if (h->c.hos_on) return fail(h, MPS_E_INVALID, "the adjoint does not cover the higher-order source term");
    AdjStore& S = h->adj;
    S.ready = 0;
    adj_free_states(h, S.cpf, S.cpt);
    adj_free(h, S.finf);
    adj_free(h, S.fint);
    S.finf = nullptr;
    S.fint = nullptr;
    const int ncp = (steps + checkpoint_interval - 1) / checkpoint_interval;
    int rc;
    for (int k = 0; k < ncp; k++) {
        double* f;
        int8_t* t;
        if ((rc = adj_alloc_state(h, &f, &t))) return rc;
        S.cpf.push_back(f);
        S.cpt.push_back(t);
    }
    if ((rc = adj_alloc_state(h, &S.finf, &S.fint))) return rc;
    S.steps = steps;
    S.interval = checkpoint_interval;
    S.opt = h->opt;
    for (int k = 0; k < steps; k++) {
        if (k % checkpoint_interval == 0)
            if ((rc = adj_save(h, S.cpf[k / checkpoint_interval], S.cpt[k / checkpoint_interval]))) return rc;
        if ((rc = step_core(h, 0, h->pinned_d))) return rc;
        if ((rc = drain(h))) return rc;
        const double cn = courant_of(h, h->pinned_d);
        if (courants) courants[k] = cn;
    }
    if ((rc = adj_save(h, S.finf, S.fint))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->st));
    S.ready = 1;
    return MPS_OK;
}

int mps_adjoint_backward(mps_handle* h, const double* pos_bar, const double* vel_bar, double* pos_bar0,
                         double* vel_bar0, double* param_bar, mps_solver_stats* adjoint_stats) {
    if (!h || !pos_bar0 || !vel_bar0 || !param_bar) return MPS_E_INVALID;
    AdjStore& S = h->adj;
    if (!S.ready) return fail(h, MPS_E_INVALID, "no recorded run: call mps_adjoint_record first");
    int rc = drain(h);
    if (rc) return rc;
    // the recomputation repeats the recorded steps: keep the handle's events,
    // statistics and options as they were
    const int64_t counts0[3] = {h->counts[0], h->counts[1], h->counts[2]};
    const std::vector<int32_t> ghost0 = h->ghost_ids, unconn0 = h->unconn_ids;
    const std::vector<mps_solver_stats> sstats0 = h->sstats;
    const std::vector<double> scourant0 = h->scourant;
    const mps_solver_options opt0 = h->opt;
    h->opt = S.opt;
    mps_solver_stats agg{};
    rc = adj_backward_impl(h, pos_bar, vel_bar, agg);
    const std::string err = h->err;
    h->opt = opt0;
    int rc2 = adj_restore(h, S.finf, S.fint);
    if (!rc2) rc2 = drain(h);
    h->counts[0] = counts0[0];
    h->counts[1] = counts0[1];
    h->counts[2] = counts0[2];
    h->ghost_ids = ghost0;
    h->unconn_ids = unconn0;
    h->sstats = sstats0;
    h->scourant = scourant0;
    if (rc) {
        h->err = err;
        return rc;
    }
    if (rc2) return rc2;
    const size_t N = (size_t) h->n;
    std::vector<double> comp(N);
    for (int d = 0; d < 3; d++) {
        if (N) HIPCHK(h, hipMemcpy(comp.data(), S.a.xb[d], N * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) pos_bar0[3 * i + d] = comp[i];
        if (N) HIPCHK(h, hipMemcpy(comp.data(), S.a.vb[d], N * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) vel_bar0[3 * i + d] = comp[i];
    }
    HIPCHK(h, hipMemcpy(param_bar, S.a.acc, 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (adjoint_stats) *adjoint_stats = agg;
    return MPS_OK;
}

int mps_snapshot_begin(mps_handle* h, int32_t slot) {
    if (!h || slot < 0 || slot > 1) return MPS_E_INVALID;
    ExportSlot& x = h->xs[slot];
    const size_t N = (size_t) h->n;
    if (!x.done) {
        HIPCHK(h, hipEventCreateWithFlags(&x.done, hipEventDisableTiming));
        const size_t n1 = N ? N : 1;
        HIPCHK(h, hipHostMalloc((void**) &x.type, n1, hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void**) &x.bc, n1, hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void**) &x.pos, 3 * n1 * sizeof(double), hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void**) &x.vel, 3 * n1 * sizeof(double), hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void**) &x.p, n1 * sizeof(double), hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void**) &x.ndr, n1 * sizeof(double), hipHostMallocDefault));
        HIPCHK(h, dalloc(&x.dpos, 3 * n1));
        HIPCHK(h, dalloc(&x.dvel, 3 * n1));
        HIPCHK(h, dalloc(&x.dtype, n1));
    }
    if (N) {
        const Hot c = cur(h);
        HIPCHK(h, launch_snapshot(h->dim, h->n, c, h->cold, x.dpos, x.dvel, h->st));
        HIPCHK(h, launch_gather_type(h->n, c, x.dtype, h->st));
        HIPCHK(h, hipMemcpyAsync(x.pos, x.dpos, 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(x.vel, x.dvel, 3 * N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(x.type, x.dtype, N, hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(x.p, h->cold.p, N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(x.ndr, h->cold.ndr, N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipMemcpyAsync(x.bc, h->cold.bc, N, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(h, hipEventRecord(x.done, h->st));
    x.pending = true;
    return MPS_OK;
}

int mps_snapshot_end(mps_handle* h, int32_t slot, const int8_t** type, const double** pos, const double** vel,
                     const double** p, const double** ndr, const int8_t** bc) {
    if (!h || slot < 0 || slot > 1) return MPS_E_INVALID;
    // may run on another host thread while the owner steps: touches only the slot
    ExportSlot& x = h->xs[slot];
    if (!x.pending) return MPS_E_INVALID;
    if (hipEventSynchronize(x.done) != hipSuccess) return MPS_E_HIP;
    if (type) *type = x.type;
    if (pos) *pos = x.pos;
    if (vel) *vel = x.vel;
    if (p) *p = x.p;
    if (ndr) *ndr = x.ndr;
    if (bc) *bc = x.bc;
    return MPS_OK;
}

// ---- lossless checkpoints ----
// Image: 64-byte header {magic "MPSAMDCK", version 2, n, dim, nscene, sizes,
// a hash of every scene's particle count and settings, 0...},
// then the saved state (kStateDoubles x n doubles, kStateBytes x n bytes) in
// the layout of save_state: per particle id x y z vx vy vz, warm-start
// solution, pressure, minimum pressure, number density ratio, a ghost's
// stale number density; then type and boundary condition.
struct CkHeader {
    char magic[8];
    int32_t version, n, dim, nscene;
    int64_t doubles, bytes;
    uint64_t shape_hash;  // ck_shape_hash of the handle that saved it
    char pad[16];
};
static_assert(sizeof(CkHeader) == 64, "checkpoint header is 64 bytes");
constexpr int32_t kCkVersion = 2;

// FNV-1a over each scene's particle count and the settings the step depends
// on (all but the finish time, the output interval and the CFL limit, so a
// run can be resumed towards a later finish): an image loads only into a
// handle whose scenes hold as many particles each and step the same way.
static uint64_t ck_shape_hash(const mps_handle* h) {
    uint64_t x = 1469598103934665603ull;
    auto mix = [&x](const void* p, size_t len) {
        const unsigned char* b = (const unsigned char*) p;
        for (size_t i = 0; i < len; i++) {
            x ^= b[i];
            x *= 1099511628211ull;
        }
    };
    for (int k = 0; k < h->nscene; k++) {
        const SceneInfo& sc = h->scenes[k];
        const int32_t ints[4] = {sc.n, sc.s.dim, sc.s.qc_on, sc.s.hos_on};
        mix(ints, sizeof ints);
        mix(&sc.s.particle_distance, sizeof(double));
        mix(&sc.s.dt, sizeof(double));
        const char* d0 = (const char*) &sc.s.domain_min[0];
        mix(d0, (size_t) ((const char*) (&sc.s.restitution + 1) - d0));
    }
    return x;
}

int64_t mps_checkpoint_size(const mps_handle* h) {
    if (!h) return -1;
    return (int64_t) sizeof(CkHeader) + (int64_t) h->n * (kStateDoubles * (int64_t) sizeof(double) + kStateBytes);
}

static int ck_staging(mps_handle* h, double** f, int8_t** t) {
    const size_t N = (size_t) h->n;
    HIPCHK(h, dalloc(f, (size_t) kStateDoubles * N));
    if (hipError_t e = dalloc(t, (size_t) kStateBytes * N)) {
        hipFree(*f);
        return hip_fail(h, e, "checkpoint staging");
    }
    return MPS_OK;
}

int mps_checkpoint_save(mps_handle* h, void* buf, int64_t cap) {
    if (!h || !buf) return MPS_E_INVALID;
    const int64_t need = mps_checkpoint_size(h);
    if (cap < need) return fail(h, MPS_E_INVALID, "checkpoint buffer too small");
    int rc = drain(h);
    if (rc) return rc;
    CkHeader hd{};
    memcpy(hd.magic, "MPSAMDCK", 8);
    hd.version = kCkVersion;
    hd.shape_hash = ck_shape_hash(h);
    hd.n = h->n;
    hd.dim = h->dim;
    hd.nscene = h->nscene;
    hd.doubles = (int64_t) kStateDoubles * h->n;
    hd.bytes = (int64_t) kStateBytes * h->n;
    memcpy(buf, &hd, sizeof hd);
    if (h->n == 0) return MPS_OK;
    double* f = nullptr;
    int8_t* t = nullptr;
    if ((rc = ck_staging(h, &f, &t))) return rc;
    hipError_t e = save_state(h->dim, h->n, cur(h), h->cold, f, t, h->st);
    char* out = (char*) buf + sizeof hd;
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, f, (size_t) hd.doubles * sizeof(double), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out + hd.doubles * sizeof(double), t, (size_t) hd.bytes, hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    hipFree(f);
    hipFree(t);
    if (e != hipSuccess) return hip_fail(h, e, "mps_checkpoint_save");
    return MPS_OK;
}

int mps_checkpoint_load(mps_handle* h, const void* buf, int64_t size) {
    if (!h || !buf) return MPS_E_INVALID;
    if (size < (int64_t) sizeof(CkHeader)) return fail(h, MPS_E_INVALID, "checkpoint image too short");
    CkHeader hd;
    memcpy(&hd, buf, sizeof hd);
    if (memcmp(hd.magic, "MPSAMDCK", 8) != 0 || hd.version != kCkVersion)
        return fail(h, MPS_E_INVALID, "not a checkpoint image of this library");
    if (hd.n != h->n || hd.dim != h->dim || hd.nscene != h->nscene)
        return fail(h, MPS_E_INVALID, "checkpoint was taken from a handle of another shape");
    if (hd.shape_hash != ck_shape_hash(h))
        return fail(h, MPS_E_INVALID,
                    "checkpoint was taken from a handle with other per-scene particle counts or settings");
    if (size < mps_checkpoint_size(h) || hd.doubles != (int64_t) kStateDoubles * h->n ||
        hd.bytes != (int64_t) kStateBytes * h->n)
        return fail(h, MPS_E_INVALID, "checkpoint image truncated");
    const int8_t* tb = (const int8_t*) buf + sizeof hd + hd.doubles * (int64_t) sizeof(double);
    for (int64_t i = 0; i < h->n; i++) {
        if (tb[i] < -1 || tb[i] > 2) return fail(h, MPS_E_INVALID, "checkpoint holds an invalid particle type");
        if (tb[h->n + i] < 0 || tb[h->n + i] > 2)
            return fail(h, MPS_E_INVALID, "checkpoint holds an invalid boundary condition");
    }
    int rc = drain(h);
    if (rc) return rc;
    if (h->n == 0) return MPS_OK;
    double* f = nullptr;
    int8_t* t = nullptr;
    if ((rc = ck_staging(h, &f, &t))) return rc;
    const char* in = (const char*) buf + sizeof hd;
    hipError_t e = hipMemcpyAsync(f, in, (size_t) hd.doubles * sizeof(double), hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t, in + hd.doubles * sizeof(double), (size_t) hd.bytes, hipMemcpyHostToDevice, h->st);
    Hot c = cur(h);
    if (e == hipSuccess) e = restore_state(h->dim, h->n, c, h->cold, h->bb.slot_of, f, t, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    hipFree(f);
    hipFree(t);
    if (e != hipSuccess) return hip_fail(h, e, "mps_checkpoint_load");
    return MPS_OK;
}

}  // extern "C"
