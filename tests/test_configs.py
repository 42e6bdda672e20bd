"""Scene generators against the reference's committed inputs and the
particle counts of the benchmark workloads."""
import os

import numpy as np
import pytest

from mps_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


def _fmt(a):
    # csv-parser writes doubles with 6 decimals (inputs/*/input.csv:4)
    return np.array(["%f" % v for v in np.ravel(a)])


@pytest.mark.parametrize("name,gen,counts", [
    ("dambreak", configs.tutorial_dambreak, (200, 82, 176)),
    ("dambreak_fine", configs.tutorial_dambreak_fine, (5000, 378, 768)),
    ("hydrostatic", configs.tutorial_hydrostatic, (2400, 182, 376)),
])
def test_generator_reproduces_reference_inputs(name, gen, counts):
    t0, ref, p, ndr = configs.read_csv(os.path.join(GOLD, name, "input.csv"))
    sc = gen()
    assert t0 == 0.0
    assert sc.n == ref.n
    np.testing.assert_array_equal(sc.type, ref.type)
    # positions as the CSV prints them, character for character
    assert (_fmt(sc.pos) == _fmt(ref.pos)).all()
    assert (_fmt(sc.vel) == _fmt(ref.vel)).all()
    c = sc.counts()
    assert (c[configs.FLUID], c[configs.WALL], c[configs.DUMMY_WALL]) == counts


def test_vtu_fixture_positions_match_generator():
    sc = configs.tutorial_dambreak()
    with open(os.path.join(GOLD, "dambreak", "input.vtu")) as f:
        lines = f.read().splitlines()
    k = lines.index("<DataArray NumberOfComponents='3' type='Float64' Name='position' format='ascii'>")
    pts = np.array([[float(v) for v in lines[k + 1 + i].split()] for i in range(sc.n)])
    np.testing.assert_allclose(pts, sc.pos, rtol=0, atol=1e-12)


def test_workload_counts_small():
    s, sc, steps = configs.workload("dambreak2d_584")
    c = sc.counts()
    assert sc.n == 584 and c[configs.FLUID] == 200 and steps == 1000
    assert configs.steps_between(0.0, s.dt, s.finish_time) == 1000
    s, sc, steps = configs.workload("dambreak2d_35k")
    c = sc.counts()
    assert sc.n == 35682 and c[configs.FLUID] == 31250
    assert configs.steps_between(0.0, s.dt, s.finish_time) == 2000


def test_workload_counts_2m():
    s, sc, steps = configs.workload("dambreak2d_2M")
    c = sc.counts()
    assert sc.n == 2035232
    assert (c[configs.FLUID], c[configs.WALL], c[configs.DUMMY_WALL]) == (2000000, 17608, 17624)
    assert configs.steps_between(0.0, s.dt, s.finish_time) == 500


def test_step_count_rule():
    # src/simulation.cpp:39-43 accumulates time in FP64
    assert configs.steps_between(0.0, 1e-3, 1.0) == 999
    assert configs.steps_between(0.0, 1e-5, 0.003) == 299
    assert configs.steps_between(0.0, 1e-5, 0.003005) == 300
    assert configs.steps_between(0.0, 1e-4, 0.2) == 2000
