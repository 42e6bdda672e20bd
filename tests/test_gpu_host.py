"""The C++ class surface on the GPU: the `mps` command line (Simulation ->
MPS -> C ABI) against the oracle, and Bucket::storeParticles against the
reference's linked lists."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mps_amd import capi, configs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
CLI = os.path.join(ROOT, "mps_amd", "mps")


def _settings_text(csv, finish, interval):
    src = open(os.path.join(GOLD, "tutorials", "dambreak", "settings.yml")).read()
    src = src.replace("finish time: 2.0", f"finish time: {finish}")
    src = src.replace("output interval: 0.05", f"output interval: {interval}")
    return src.replace("inputCsvPath: inputs/dambreak/input.csv", f"inputCsvPath: {csv}")


def test_cli_runs_tutorial_and_matches_oracle(tmp_path):
    csv = os.path.join(GOLD, "inputs", "dambreak", "input.csv")
    yml = tmp_path / "settings.yml"
    yml.write_text(_settings_text(csv, "0.0105", "0.005"))
    r = subprocess.run([CLI, "-s", str(yml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = tmp_path / "result"
    outs = sorted(os.listdir(res / "csv"))
    assert outs == ["output_0000.csv", "output_0001.csv", "output_0002.csv"]
    assert sorted(os.listdir(res / "vtu")) == [o.replace(".csv", ".vtu") for o in outs]
    rep = open(res / "time_step_report.csv").read().strip().split("\n")
    assert len(rep) == 1 + 10
    assert (res / "execution.log").exists() and (res / "settings.yml").exists()
    assert "START SIMULATION" in r.stdout and "END SIMULATION" in r.stdout
    # the oracle over the same 10 steps, exporting at steps 5 and 10
    s = configs.Settings()
    _, sc, _, _ = configs.read_csv(csv)
    o = Oracle(s, sc)
    o.load_initial_state()
    for k in range(1, 11):
        o.step(k in (5, 10))
    st = o.state()
    t, got, p, ndr = configs.read_csv(str(res / "csv" / "output_0002.csv"))
    assert abs(t - 0.01) < 1e-9
    alive = st["type"] != configs.GHOST
    np.testing.assert_array_equal(got.type, st["type"][alive])
    assert np.abs(got.pos - st["pos"][alive]).max() <= 1.5e-6
    assert np.abs(got.vel - st["vel"][alive]).max() <= 1.5e-6
    assert np.abs(p - st["p"][alive]).max() <= 1e-5
    assert np.abs(ndr - st["ndr"][alive]).max() <= 1.5e-6


def test_cli_rejects_missing_settings(tmp_path):
    r = subprocess.run([CLI, "-s", str(tmp_path / "nope.yml")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot be found" in r.stdout
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "required" in r.stdout


def test_bucket_class_matches_reference_lists():
    capi.lib()
    L = ctypes.CDLL(os.path.join(ROOT, "mps_amd", "libmps_host.so"))
    L.mps_host_bucket_from_csv.restype = ctypes.c_int
    f64 = np.ctypeslib.ndpointer(np.float64)
    i32 = np.ctypeslib.ndpointer(np.int32)
    L.mps_host_bucket_from_csv.argtypes = [ctypes.c_char_p, ctypes.c_double, f64, f64, i32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
    csv = os.path.join(GOLD, "inputs", "dambreak_fine", "input.csv")
    s = configs.Settings(particle_distance=0.004, domain_x=(-0.1, 1.1), domain_y=(-0.1, 0.5))
    c = s.to_c()
    dmin = np.array(list(c.domain_min))
    dmax = np.array(list(c.domain_max))
    dims = np.zeros(4, np.int32)
    assert L.mps_host_bucket_from_csv(csv.encode(), c.re_max, dmin, dmax, dims, None, None, None) == 0
    _, sc, _, _ = configs.read_csv(csv)
    first = np.zeros(dims[0], np.int32)
    last = np.zeros(dims[0], np.int32)
    nxt = np.zeros(sc.n, np.int32)
    assert L.mps_host_bucket_from_csv(csv.encode(), c.re_max, dmin, dmax, dims, first.ctypes.data,
                                      last.ctypes.data, nxt.ctypes.data) == 0
    o = Oracle(s, sc)
    o.set_neighbors()
    ob = o.bucket()
    np.testing.assert_array_equal(dims, ob["dims"])
    np.testing.assert_array_equal(first, ob["first"])
    np.testing.assert_array_equal(last, ob["last"])
    np.testing.assert_array_equal(nxt, ob["next"])


def test_async_and_sync_export_write_the_same_files(tmp_path):
    """The CLI's default export (asynchronous snapshot, threaded formatting,
    background writes) and the reference's synchronous path
    (MPS_AMD_SYNC_EXPORT=1) produce byte-identical result files."""
    csv = os.path.join(GOLD, "inputs", "dambreak_fine", "input.csv")
    outs = {}
    for mode in ("async", "sync"):
        d = tmp_path / mode
        d.mkdir()
        yml = d / "settings.yml"
        src = open(os.path.join(GOLD, "tutorials", "dambreak_fine", "settings.yml")).read()
        src = src.replace("finish time: 2.0", "finish time: 0.0105")
        src = src.replace("output interval: 0.025", "output interval: 0.002")
        src = src.replace("inputCsvPath: inputs/dambreak_fine/input.csv", f"inputCsvPath: {csv}")
        yml.write_text(src)
        env = dict(os.environ)
        env.pop("MPS_AMD_SYNC_EXPORT", None)
        if mode == "sync":
            env["MPS_AMD_SYNC_EXPORT"] = "1"
        r = subprocess.run([CLI, "-s", str(yml)], cwd=d, capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        res = d / "result"
        files = sorted(os.listdir(res / "csv")) + sorted(os.listdir(res / "vtu"))
        outs[mode] = {f: ((res / "csv" / f) if f.endswith(".csv") else (res / "vtu" / f)).read_bytes()
                      for f in files}
    assert len(outs["async"]) == 12  # t = 0 and every 2 ms to 10 ms, CSV and VTU
    assert outs["async"].keys() == outs["sync"].keys()
    for f in outs["async"]:
        assert outs["async"][f] == outs["sync"][f], f
