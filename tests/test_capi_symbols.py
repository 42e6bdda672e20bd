"""The C-ABI library loads on a host without a GPU and exports every entry
point include/mps_amd.h declares (no compute call is made here)."""
import ctypes
import os
import re

from mps_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mps_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mps_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = _declared()
    assert "mps_step" in names and "mps_create" in names
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTED)


def test_create_rejects_bad_input_without_gpu_work():
    L = capi.lib()
    h = ctypes.c_void_p()
    rc = L.mps_create(None, 0, None, None, None, None, ctypes.byref(h))
    assert rc == 1 and h.value is None
    assert b"invalid" in L.mps_last_error(None)
