"""ctypes binding of the FP64 CPU oracle (oracle/libmps_oracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libmps_oracle.so")

_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build()
    L = ctypes.CDLL(ORACLE_SO)
    vp = ctypes.c_void_p
    L.oracle_ref_values.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.oracle_weight.restype = ctypes.c_double
    L.oracle_weight.argtypes = [ctypes.c_double, ctypes.c_double]
    L.oracle_create.restype = vp
    L.oracle_create.argtypes = [vp, ctypes.c_int32, _i32p, _f64p, _f64p]
    L.oracle_destroy.argtypes = [vp]
    L.oracle_set_solver.argtypes = [vp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]
    L.oracle_load_initial_state.argtypes = [vp]
    L.oracle_step.restype = ctypes.c_double
    L.oracle_step.argtypes = [vp, ctypes.c_int32]
    L.oracle_get_state.argtypes = [vp, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p]
    L.oracle_set_state.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.oracle_debug_counts.argtypes = [vp, _i64p]
    L.oracle_event_ids.restype = ctypes.c_int32
    L.oracle_event_ids.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32]
    L.oracle_set_neighbors.argtypes = [vp]
    L.oracle_bucket_shape.argtypes = [vp, _i32p, ctypes.POINTER(ctypes.c_double)]
    L.oracle_dump_bucket.argtypes = [vp, _i32p, _i32p, _i32p]
    L.oracle_neighbor_total.restype = ctypes.c_int64
    L.oracle_neighbor_total.argtypes = [vp]
    L.oracle_dump_neighbors.argtypes = [vp, _i32p, _i32p, _f64p]
    L.oracle_offdiag_nnz.restype = ctypes.c_int64
    L.oracle_offdiag_nnz.argtypes = [vp]
    L.oracle_get_pressure_system.argtypes = [vp, _f64p, _f64p, _i32p, _i32p, _f64p, _f64p]
    L.oracle_get_snapshot.argtypes = [vp, _f64p, _f64p]
    L.oracle_cg_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]
    L.oracle_stage_seconds.argtypes = [vp, _f64p]
    L.oracle_pcg.restype = ctypes.c_int32
    L.oracle_pcg.argtypes = [ctypes.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, ctypes.c_double,
                             ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def ref_values(dim, l0, re):
    n0, lam = ctypes.c_double(), ctypes.c_double()
    lib().oracle_ref_values(dim, l0, re, ctypes.byref(n0), ctypes.byref(lam))
    return n0.value, lam.value


def pcg(row_ptr, col, val, b, tol, max_iter=0):
    n = b.shape[0]
    x = np.zeros(n)
    rr = ctypes.c_double()
    it = lib().oracle_pcg(n, np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32),
                          np.ascontiguousarray(val, np.float64), np.ascontiguousarray(b, np.float64), x,
                          tol, max_iter, ctypes.byref(rr))
    return x, it, rr.value


class Oracle:
    """One oracle simulation (the reference's MPS object restated in C)."""

    def __init__(self, settings, scene, tol=None, max_iter=0, duplicate=False):
        self.L = lib()
        self.n = scene.n
        self._cs = settings.to_c() if hasattr(settings, "to_c") else settings
        self.h = self.L.oracle_create(ctypes.byref(self._cs), self.n,
                                      np.ascontiguousarray(scene.type, np.int32),
                                      np.ascontiguousarray(scene.pos, np.float64),
                                      np.ascontiguousarray(scene.vel, np.float64))
        self.L.oracle_set_solver(self.h, float(np.finfo(np.float64).eps) if tol is None else tol,
                                 max_iter, int(duplicate))

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_initial_state(self):
        self.L.oracle_load_initial_state(self.h)

    def step(self, is_export=False):
        return self.L.oracle_step(self.h, int(is_export))

    def state(self):
        n = self.n
        t = np.zeros(n, np.int32)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        p = np.zeros(n)
        pmin = np.zeros(n)
        ndr = np.zeros(n)
        bc = np.zeros(n, np.int32)
        self.L.oracle_get_state(self.h, t, pos, vel, p, pmin, ndr, bc)
        return dict(type=t, pos=pos, vel=vel, p=p, pmin=pmin, ndr=ndr, bc=bc)

    def set_state(self, type=None, pos=None, vel=None, p=None, ndr=None, bc=None):
        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            self._keep.append(a)
            return a.ctypes.data
        self._keep = []
        self.L.oracle_set_state(self.h, ptr(type, np.int32), ptr(pos, np.float64), ptr(vel, np.float64),
                                ptr(p, np.float64), ptr(ndr, np.float64), ptr(bc, np.int32))

    def debug_counts(self):
        c = np.zeros(3, np.int64)
        self.L.oracle_debug_counts(self.h, c)
        return c

    def event_ids(self, kind):
        m = self.L.oracle_event_ids(self.h, kind, None, 0)
        ids = np.zeros(max(m, 1), np.int32)
        self.L.oracle_event_ids(self.h, kind, ids.ctypes.data, m)
        return ids[:m]

    def set_neighbors(self):
        self.L.oracle_set_neighbors(self.h)

    def bucket(self):
        dims = np.zeros(4, np.int32)
        ln = ctypes.c_double()
        self.L.oracle_bucket_shape(self.h, dims, ctypes.byref(ln))
        first = np.zeros(dims[0], np.int32)
        last = np.zeros(dims[0], np.int32)
        nxt = np.zeros(self.n, np.int32)
        self.L.oracle_dump_bucket(self.h, first, last, nxt)
        return dict(dims=dims, length=ln.value, first=first, last=last, next=nxt)

    def neighbors(self):
        tot = self.L.oracle_neighbor_total(self.h)
        off = np.zeros(self.n + 1, np.int32)
        ids = np.zeros(max(tot, 1), np.int32)
        d = np.zeros(max(tot, 1))
        self.L.oracle_dump_neighbors(self.h, off, ids, d)
        return off, ids[:tot], d[:tot]

    def pressure_system(self):
        n = self.n
        nnz = self.L.oracle_offdiag_nnz(self.h)
        b = np.zeros(n)
        diag = np.zeros(n)
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1))
        x = np.zeros(n)
        self.L.oracle_get_pressure_system(self.h, b, diag, rp, col, val, x)
        return dict(b=b, diag=diag, row_ptr=rp, col=col[:nnz], val=val[:nnz], x=x)

    def snapshot(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        self.L.oracle_get_snapshot(self.h, pos, vel)
        return pos, vel

    def cg_stats(self):
        it = ctypes.c_int32()
        rr = ctypes.c_double()
        self.L.oracle_cg_stats(self.h, ctypes.byref(it), ctypes.byref(rr))
        return it.value, rr.value

    def stage_seconds(self):
        out = np.zeros(6)
        self.L.oracle_stage_seconds(self.h, out)
        return out
