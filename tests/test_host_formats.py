"""The C++ class surface's host-side formats, without a GPU: the YAML
settings reader against the reference's tutorial files, and the CSV / VTU
writers against the reference's committed input.csv / input.vtu."""
import ctypes
import os

import numpy as np
import pytest
import yaml

from mps_amd import capi, configs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
HOST_SO = os.path.join(os.path.dirname(HERE), "mps_amd", "libmps_host.so")


def _host():
    capi.lib()  # libmps_host.so links against libmps_amd.so
    L = ctypes.CDLL(HOST_SO)
    L.mps_host_load_settings.restype = ctypes.c_int
    L.mps_host_load_settings.argtypes = [ctypes.c_char_p, ctypes.POINTER(configs.MpsSettings), ctypes.c_char_p,
                                         ctypes.c_int]
    L.mps_host_convert_csv.restype = ctypes.c_int
    L.mps_host_convert_csv.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
    return L


TUTORIALS = sorted(os.listdir(os.path.join(GOLD, "tutorials")))


@pytest.mark.parametrize("name", TUTORIALS)
def test_settings_reader_matches_yaml(name):
    path = os.path.join(GOLD, "tutorials", name, "settings.yml")
    y = yaml.safe_load(open(path))
    s = configs.MpsSettings()
    buf = ctypes.create_string_buffer(512)
    assert _host().mps_host_load_settings(path.encode(), ctypes.byref(s), buf, 512) == 0
    l0 = y["particle distance"]
    assert s.dim == y["dim"] and s.particle_distance == l0 and s.dt == y["dt"]
    assert s.finish_time == y["finish time"] and s.output_interval == y["output interval"]
    assert s.cfl_condition == y["CFL condition"]
    assert list(s.domain_min) == [y["domain range X"][0], y["domain range Y"][0], y["domain range Z"][0]]
    assert list(s.domain_max) == [y["domain range X"][1], y["domain range Y"][1], y["domain range Z"][1]]
    assert list(s.gravity) == y["gravity"]
    er = y["effective radius ratio"]
    # src/settings.cpp:43-61: ratio times l0, and the largest of the three
    assert s.re_pressure == er["pressure"] * l0 and s.re_viscosity == er["viscosity"] * l0
    assert s.re_surface == er["surface detection"] * l0
    assert s.re_max == max(s.re_pressure, s.re_viscosity, s.re_surface)
    ps = y["pressure stabilization"]

    def on(d):  # YAML 1.1 (PyYAML) reads the key `on` itself as True
        return d["on"] if "on" in d else d[True]

    assert bool(s.qc_on) == on(ps["quasi compressibility"])
    assert bool(s.hos_on) == on(ps["higher order source term"])
    if s.qc_on:
        assert s.compressibility == ps["quasi compressibility"]["compressibility"]
    assert s.relaxation == y["relaxation coefficient for pressure"]
    assert s.collision_distance == y["collision distance ratio"] * l0
    assert s.restitution == y["coefficient of restitution"]
    assert buf.value.decode() == y["inputCsvPath"]


def test_settings_reader_rejects_integer_where_a_double_is_required(tmp_path):
    src = open(os.path.join(GOLD, "tutorials", "dambreak", "settings.yml")).read()
    bad = tmp_path / "bad.yml"
    bad.write_text(src.replace("density: 1000.0", "density: 1000"))
    s = configs.MpsSettings()
    assert _host().mps_host_load_settings(str(bad).encode(), ctypes.byref(s), None, 0) == 1


@pytest.mark.parametrize("name", ["dambreak", "hydrostatic", "dambreak_fine"])
def test_csv_writer_reproduces_reference_file(name, tmp_path):
    src = os.path.join(GOLD, "inputs", name, "input.csv")
    out = tmp_path / "out.csv"
    assert _host().mps_host_convert_csv(src.encode(), 1000.0, str(out).encode(), None) == 0
    assert out.read_bytes() == open(src, "rb").read()


def _blocks(text):
    """VTU text split into (header line, lines) data blocks."""
    lines = text.split("\n")
    out, cur, head = [], [], None
    for ln in lines:
        if ln.startswith("<DataArray"):
            head, cur = ln, []
        elif ln.startswith("</DataArray>") and head is not None:
            out.append((head, cur))
            head = None
        elif head is not None:
            cur.append(ln)
    return lines, out


def test_vtu_writer_reproduces_reference_file(tmp_path):
    src = os.path.join(GOLD, "inputs", "dambreak", "input.csv")
    ref = open(os.path.join(GOLD, "inputs", "dambreak", "input.vtu")).read()
    out = tmp_path / "out.vtu"
    assert _host().mps_host_convert_csv(src.encode(), 1000.0, None, str(out).encode()) == 0
    mine = out.read_text()
    ml, mb = _blocks(mine)
    rl, rb = _blocks(ref)
    assert len(ml) == len(rl)
    # every structural line identical
    assert [x for x in ml if not x[:1].isdigit() and not x[:1] == "-"] == \
           [x for x in rl if not x[:1].isdigit() and not x[:1] == "-"]
    assert len(mb) == len(rb)
    for (mh, mv), (rh, rv) in zip(mb, rb):
        assert mh == rh
        if "Name='Pressure'" in mh or "Name='Boundary Condition'" in mh:
            # the reference printed uninitialised memory here (1.061e-312, 2);
            # the loader here defines pressure 0 and BC Ignored (0)
            assert set(mv) == {"0"}
            assert len(mv) == len(rv)
        else:
            assert mv == rv


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_snapshot_writers_match_the_writers(threads, tmp_path):
    """The asynchronous export path formats the snapshot on several threads
    (ranges joined in order); its files must be byte-identical to the
    Exporter's for the same particles, ghosts (skipped in the CSV rows,
    counted in its header, kept in the VTU) and negative values included."""
    rng = np.random.default_rng(threads)
    n = 1001
    t = rng.choice([-1, 0, 1, 2], size=n, p=[0.05, 0.6, 0.2, 0.15]).astype(np.int8)
    pos = rng.normal(0, 0.3, size=(n, 3))
    vel = rng.normal(0, 2.0, size=(n, 3))
    p = rng.normal(0, 3000.0, size=n)
    ndr = rng.uniform(0.5, 1.2, size=n)
    bc = rng.integers(0, 3, size=n).astype(np.int8)
    L = _host()
    L.mps_host_export_both.restype = ctypes.c_int
    L.mps_host_export_both.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_double, ctypes.c_char_p,
                                                                                ctypes.c_int]
    arrs = [np.ascontiguousarray(a) for a in (t, pos, vel, p, ndr, bc)]
    pre = str(tmp_path) + "/"
    assert L.mps_host_export_both(n, *[a.ctypes.data for a in arrs], 0.125, pre.encode(), threads) == 0
    for ext in ("csv", "vtu"):
        a = (tmp_path / f"sync.{ext}").read_bytes()
        b = (tmp_path / f"async.{ext}").read_bytes()
        assert len(a) > 1000 and a == b, ext
