"""ctypes binding of libmps_amd.so (include/mps_amd.h).

The library is the product; this module only marshals numpy arrays across
the C ABI.  Loading fails loudly if the shared library has not been built:
there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .configs import MpsSettings, Scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmps_amd.so")

MPS_PREC_FP64 = 0
MPS_PREC_MIXED = 1
MPS_PRECOND_JACOBI = 0
MPS_PRECOND_MULTIGRID = 1

MPS_E_SOLVER = 5
_ERRORS = {1: "MPS_E_INVALID", 2: "MPS_E_HIP", 3: "MPS_E_NOMEM", 4: "MPS_E_RANGE", 5: "MPS_E_SOLVER"}

EXPORTED = [
    "mps_last_error", "mps_create", "mps_destroy", "mps_set_solver", "mps_get_solver_stats",
    "mps_load_initial_state", "mps_step", "mps_run_steps", "mps_download", "mps_download_min_pressure",
    "mps_upload", "mps_debug_events", "mps_clear_debug_events", "mps_bucket_shape", "mps_dump_bucket",
    "mps_dump_neighbors", "mps_get_pressure_system", "mps_set_snapshot", "mps_get_snapshot",
    "mps_time_pressure_kernels", "mps_set_block_sizes", "mps_autotune",
    "mps_create_ensemble", "mps_scene_count", "mps_ensemble_stats",
    "mps_adjoint_record", "mps_adjoint_backward", "mps_snapshot_begin", "mps_snapshot_end",
    "mps_checkpoint_size", "mps_checkpoint_save", "mps_checkpoint_load", "mps_debug_counters",
    "mps_dump_neighbors_noalloc",
]


class MpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("max_refinements", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("inner_tolerance", ctypes.c_double),
        ("check_interval", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("preconditioner", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SolverStats(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("refinements", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("rhs_norm", ctypes.c_double),
        ("rel_residual", ctypes.c_double),
        ("initial_residual", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libmps_amd.so; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m mps_amd.build` (no fallback path exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    L.mps_last_error.restype = ctypes.c_char_p
    L.mps_last_error.argtypes = [vp]
    L.mps_create.restype = i32
    L.mps_create.argtypes = [ctypes.POINTER(MpsSettings), i32, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.mps_destroy.argtypes = [vp]
    L.mps_set_solver.argtypes = [vp, ctypes.POINTER(SolverOptions)]
    L.mps_get_solver_stats.argtypes = [vp, ctypes.POINTER(SolverStats)]
    L.mps_load_initial_state.argtypes = [vp]
    L.mps_step.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double)]
    L.mps_run_steps.argtypes = [vp, i32, vp]
    L.mps_download.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.mps_download_min_pressure.argtypes = [vp, vp]
    L.mps_upload.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.mps_debug_events.argtypes = [vp, vp, vp, vp, i32]
    L.mps_clear_debug_events.argtypes = [vp]
    L.mps_bucket_shape.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_double)]
    L.mps_dump_bucket.argtypes = [vp, vp, vp, vp]
    L.mps_dump_neighbors.argtypes = [vp, vp, vp, vp]
    L.mps_get_pressure_system.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.mps_set_snapshot.argtypes = [vp, i32]
    L.mps_get_snapshot.argtypes = [vp, vp, vp]
    L.mps_time_pressure_kernels.argtypes = [vp, i32, vp, vp]
    L.mps_set_block_sizes.argtypes = [vp, vp]
    L.mps_autotune.argtypes = [vp, i32, vp]
    L.mps_create_ensemble.argtypes = [i32, vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.mps_scene_count.argtypes = [vp]
    L.mps_ensemble_stats.argtypes = [vp, vp, vp]
    L.mps_adjoint_record.argtypes = [vp, i32, i32, vp]
    L.mps_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.POINTER(SolverStats)]
    L.mps_snapshot_begin.argtypes = [vp, i32]
    L.mps_snapshot_end.argtypes = [vp, i32] + [ctypes.POINTER(vp)] * 6
    L.mps_checkpoint_size.restype = ctypes.c_int64
    L.mps_checkpoint_size.argtypes = [vp]
    L.mps_checkpoint_save.argtypes = [vp, vp, ctypes.c_int64]
    L.mps_checkpoint_load.argtypes = [vp, vp, ctypes.c_int64]
    L.mps_debug_counters.argtypes = [vp, vp, i32]
    L.mps_dump_neighbors_noalloc.argtypes = [vp, vp, vp, vp]
    for name in EXPORTED:
        if name not in ("mps_last_error", "mps_checkpoint_size"):
            getattr(L, name).restype = i32
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Simulation:
    """One particle system on the GPU (the C ABI handle).  With a list of
    (settings, scene) pairs it is an ensemble: independent scenes stepped
    together in one handle, particles packed scene after scene."""

    def __init__(self, settings, scene: Scene = None):
        self.L = lib()
        if scene is None:
            members = list(settings)
            cs = (MpsSettings * len(members))(*[m[0].to_c() for m in members])
            scenes = [m[1] for m in members]
            counts = np.array([sc.n for sc in scenes], np.int32)
            ids = np.concatenate([np.arange(sc.n, dtype=np.int32) for sc in scenes])
            t = np.ascontiguousarray(np.concatenate([sc.type for sc in scenes]), np.int8)
            pos = np.ascontiguousarray(np.concatenate([sc.pos for sc in scenes]), np.float64)
            vel = np.ascontiguousarray(np.concatenate([sc.vel for sc in scenes]), np.float64)
            self.n = int(counts.sum())
            self.offsets = np.concatenate([[0], np.cumsum(counts)])
            self._cs = cs
            h = ctypes.c_void_p()
            rc = self.L.mps_create_ensemble(len(members), ctypes.cast(cs, ctypes.c_void_p), _p(counts), _p(ids),
                                            _p(t), _p(pos), _p(vel), ctypes.byref(h))
        else:
            self.n = scene.n
            self.offsets = np.array([0, self.n])
            self._cs = settings.to_c() if hasattr(settings, "to_c") else settings
            ids = np.arange(self.n, dtype=np.int32)
            t = np.ascontiguousarray(scene.type, np.int8)
            pos = np.ascontiguousarray(scene.pos, np.float64)
            vel = np.ascontiguousarray(scene.vel, np.float64)
            h = ctypes.c_void_p()
            rc = self.L.mps_create(ctypes.byref(self._cs), self.n, _p(ids), _p(t), _p(pos), _p(vel), ctypes.byref(h))
        if rc:
            raise MpsError(rc, self.L.mps_last_error(None).decode())
        self.h = h
        self.nscene = len(self.offsets) - 1
        self.strict_solver = False  # True: a step whose solve missed its tolerance raises
        self.solver_warning = None  # message of the last such step

    def scene_stats(self):
        """Per-scene solver statistics and Courant numbers of the last step."""
        st = (SolverStats * self.nscene)()
        cn = np.zeros(self.nscene)
        self._check(self.L.mps_ensemble_stats(self.h, ctypes.cast(st, ctypes.c_void_p), _p(cn)))
        return [x.as_dict() for x in st], cn

    def adjoint_record(self, steps, checkpoint_interval=50):
        """Runs `steps` steps keeping device checkpoints for the reverse sweep;
        returns the Courant numbers."""
        out = np.zeros(max(steps, 1))
        self._check(self.L.mps_adjoint_record(self.h, int(steps), int(checkpoint_interval), _p(out)))
        return out[:steps]

    def adjoint_backward(self, pos_bar, vel_bar=None):
        """dL/dx, dL/dv at the recorded run's start and dL/d(nu, restitution,
        relaxation) from dL/dx, dL/dv at its end (N x 3 arrays, id order)."""
        pb = np.ascontiguousarray(pos_bar, np.float64).reshape(self.n, 3)
        vb = None if vel_bar is None else np.ascontiguousarray(vel_bar, np.float64).reshape(self.n, 3)
        pos0 = np.zeros((self.n, 3))
        vel0 = np.zeros((self.n, 3))
        par = np.zeros(3)
        st = SolverStats()
        self._check(self.L.mps_adjoint_backward(self.h, _p(pb), _p(vb), _p(pos0), _p(vel0), _p(par),
                                                ctypes.byref(st)))
        return pos0, vel0, par, st.as_dict()

    def checkpoint(self) -> bytes:
        """Lossless image of the full state (mps_checkpoint_save)."""
        size = self.L.mps_checkpoint_size(self.h)
        buf = ctypes.create_string_buffer(size)
        self._check(self.L.mps_checkpoint_save(self.h, ctypes.cast(buf, ctypes.c_void_p), size))
        return buf.raw

    def restore(self, image: bytes):
        buf = ctypes.create_string_buffer(image, len(image))
        self._check(self.L.mps_checkpoint_load(self.h, ctypes.cast(buf, ctypes.c_void_p), len(image)))

    def scene_state(self, k, state=None):
        """The fields of scene k (local ids) out of a state() dict."""
        state = self.state() if state is None else state
        a, b = int(self.offsets[k]), int(self.offsets[k + 1])
        return {key: v[a:b] for key, v in state.items()}

    def _check(self, rc):
        if rc:
            raise MpsError(rc, self.L.mps_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.mps_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solver(self, precision=MPS_PREC_FP64, tolerance=None, warm_start=False, max_iterations=0,
                   max_refinements=0, inner_tolerance=1e-4, check_interval=32, use_graph=True,
                   preconditioner=0):
        o = SolverOptions()
        o.precision = precision
        o.warm_start = 2 if warm_start == 2 else int(bool(warm_start))
        o.max_iterations = max_iterations
        o.max_refinements = max_refinements
        o.tolerance = float(np.finfo(np.float64).eps) if tolerance is None else float(tolerance)
        o.inner_tolerance = inner_tolerance
        o.check_interval = check_interval
        o.use_graph = int(bool(use_graph))
        o.preconditioner = int(preconditioner)
        self._check(self.L.mps_set_solver(self.h, ctypes.byref(o)))

    def debug_counters(self):
        """Internal counters: graph rebuilds, CG and matrix reallocations, the
        LDS staging counts of the walks (MPS_AMD_STAGE_STATS=1 handles) and
        the CG iterations enqueued (including a chunk's past convergence) and
        the CG chunk graphs captured."""
        out = np.zeros(9, np.int64)
        self._check(self.L.mps_debug_counters(self.h, _p(out), 9))
        keys = ("graph_rebuilds", "cg_reallocs", "matrix_reallocs", "staged_blocks", "staged_full",
                "staged_partial", "staged_partial_gap", "cg_launched", "graph_captures")
        return {k: int(v) for k, v in zip(keys, out)}

    def solver_stats(self):
        s = SolverStats()
        self._check(self.L.mps_get_solver_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def load_initial_state(self):
        self._check(self.L.mps_load_initial_state(self.h))

    def _check_step(self, rc):
        # MPS_E_SOLVER: the step completed with a solve that stopped above its
        # tolerance; like the reference (which ignores the solver's status,
        # src/mps.cpp:457-460) the run goes on, and the message is kept
        if rc == MPS_E_SOLVER:
            self.solver_warning = self.L.mps_last_error(self.h).decode()
            if self.strict_solver:
                raise MpsError(rc, self.solver_warning)
            return
        self._check(rc)

    def step(self, is_export=False):
        c = ctypes.c_double()
        self._check_step(self.L.mps_step(self.h, int(bool(is_export)), ctypes.byref(c)))
        return c.value

    def run_steps(self, k):
        out = np.zeros(max(k, 1))
        self._check_step(self.L.mps_run_steps(self.h, k, _p(out)))
        return out[:k]

    def state(self):
        n = self.n
        t = np.zeros(n, np.int8)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        p = np.zeros(n)
        ndr = np.zeros(n)
        bc = np.zeros(n, np.int8)
        self._check(self.L.mps_download(self.h, _p(t), _p(pos), _p(vel), _p(p), _p(ndr), _p(bc)))
        pmin = np.zeros(n)
        self._check(self.L.mps_download_min_pressure(self.h, _p(pmin)))
        return dict(type=t, pos=pos, vel=vel, p=p, ndr=ndr, bc=bc, pmin=pmin)

    def upload(self, type=None, pos=None, vel=None, p=None, ndr=None, bc=None):
        def c(a, dt, shape=None):
            if a is None:
                return None
            return np.ascontiguousarray(a, dt)
        arrs = [c(type, np.int8), c(pos, np.float64), c(vel, np.float64), c(p, np.float64), c(ndr, np.float64),
                c(bc, np.int8)]
        self._check(self.L.mps_upload(self.h, *[_p(a) for a in arrs]))

    def debug_events(self):
        counts = np.zeros(3, np.int64)
        cap = self.n + 16
        g = np.zeros(cap, np.int32)
        u = np.zeros(cap, np.int32)
        self._check(self.L.mps_debug_events(self.h, _p(counts), _p(g), _p(u), cap))
        return counts, g[:min(int(counts[0]), cap)], u[:min(int(counts[1]), cap)]

    def clear_debug_events(self):
        self._check(self.L.mps_clear_debug_events(self.h))

    def bucket(self):
        dims = np.zeros(4, np.int32)
        ln = ctypes.c_double()
        self._check(self.L.mps_bucket_shape(self.h, _p(dims), ctypes.byref(ln)))
        first = np.zeros(dims[0], np.int32)
        last = np.zeros(dims[0], np.int32)
        nxt = np.zeros(max(self.n, 1), np.int32)
        self._check(self.L.mps_dump_bucket(self.h, _p(first), _p(last), _p(nxt)))
        return dict(dims=dims, length=ln.value, first=first, last=last, next=nxt[:self.n])

    def neighbors(self):
        off = np.zeros(self.n + 1, np.int32)
        self._check(self.L.mps_dump_neighbors(self.h, _p(off), None, None))
        tot = int(off[-1])
        ids = np.zeros(max(tot, 1), np.int32)
        d = np.zeros(max(tot, 1))
        self._check(self.L.mps_dump_neighbors(self.h, _p(off), _p(ids), _p(d)))
        return off, ids[:tot], d[:tot]

    def set_snapshot(self, on=True):
        self._check(self.L.mps_set_snapshot(self.h, int(bool(on))))

    def snapshot(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        self._check(self.L.mps_get_snapshot(self.h, _p(pos), _p(vel)))
        return pos, vel

    def autotune(self, steps_per_candidate=2):
        """Runs 4 x steps_per_candidate steps, keeps the fastest block size per
        kernel family; returns the 9 chosen sizes."""
        out = np.zeros(9, np.int32)
        self._check(self.L.mps_autotune(self.h, steps_per_candidate, _p(out)))
        return out

    def set_block_sizes(self, sizes):
        a = np.ascontiguousarray(sizes, np.int32)
        self._check(self.L.mps_set_block_sizes(self.h, _p(a)))

    def time_pressure_kernels(self, reps=50):
        """Average launch time (ms) and algorithmic bytes of the CG kernels on
        the last step's system: keys spmv64, update64, spmv32, update32."""
        ms = np.zeros(4)
        by = np.zeros(4, np.int64)
        self._check(self.L.mps_time_pressure_kernels(self.h, reps, _p(ms), _p(by)))
        names = ["spmv64", "update64", "spmv32", "update32"]
        return {k: (float(ms[i]), int(by[i])) for i, k in enumerate(names)}

    def pressure_system(self):
        n = self.n
        rp = np.zeros(n + 1, np.int32)
        self._check(self.L.mps_get_pressure_system(self.h, None, None, _p(rp), None, None, None))
        nnz = int(rp[-1])
        b = np.zeros(n)
        diag = np.zeros(n)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1))
        x = np.zeros(n)
        self._check(self.L.mps_get_pressure_system(self.h, _p(b), _p(diag), _p(rp), _p(col), _p(val), _p(x)))
        return dict(b=b, diag=diag, row_ptr=rp, col=col[:nnz], val=val[:nnz], x=x)
