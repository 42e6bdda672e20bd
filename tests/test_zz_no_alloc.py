"""mps_dump_neighbors_noalloc: the neighbour dump staged in the library's
static device scratch instead of buffers it allocates, for callers that
cannot allocate device memory.  Its output must equal mps_dump_neighbors'
bit for bit on two calls in a row, the second on a smaller system (the
scratch keeps the first call's contents), and a system whose lists exceed
the scratch is refused with a clear error."""
import ctypes

import numpy as np
import pytest

from mps_amd import capi, configs

pytestmark = pytest.mark.gpu


def _noalloc(sim):
    L = sim.L
    L.mps_dump_neighbors_noalloc.argtypes = [ctypes.c_void_p] * 4
    L.mps_dump_neighbors_noalloc.restype = ctypes.c_int32
    off = np.zeros(sim.n + 1, np.int32)
    assert L.mps_dump_neighbors_noalloc(sim.h, capi._p(off), None, None) == 0
    tot = int(off[-1])
    ids = np.full(max(tot, 1), -7, np.int32)
    d = np.full(max(tot, 1), -1.0)
    rc = L.mps_dump_neighbors_noalloc(sim.h, capi._p(off), capi._p(ids), capi._p(d))
    return rc, off, ids[:tot], d[:tot]


def _jittered(s, sc, seed):
    rng = np.random.default_rng(seed)
    f = sc.type == configs.FLUID
    sc.pos[f, :2] += rng.uniform(-0.2, 0.2, size=(int(f.sum()), 2)) * s.particle_distance
    return sc


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_dump_matches_allocating_dump(dim):
    if dim == 2:
        s, big, _ = configs.workload("dambreak2d_35k")
        big = _jittered(s, big, 1)
        small = _jittered(s, configs.lattice_2d(s.particle_distance, (1.0, 0.6), (0.0, 0.1), (0.0, 0.2), 2, 2), 2)
        s2 = s
    else:
        s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
        s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
        big = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
        small = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.06), (0.0, 0.1), (0.0, 0.06), 1, 2)
        s2 = s
    assert small.n < big.n
    for s_, sc in ((s, big), (s2, small)):
        sim = capi.Simulation(s_, sc)
        sim.load_initial_state()
        sim.step()  # off the initial lattice
        rc, off, ids, d = _noalloc(sim)
        assert rc == 0, sim.L.mps_last_error(sim.h)
        go, gi, gd = sim.neighbors()
        np.testing.assert_array_equal(off, go)
        np.testing.assert_array_equal(ids, gi)
        assert (d.view(np.uint64) == gd.view(np.uint64)).all()
        assert ids.size > 0
        sim.close()


def test_noalloc_dump_refuses_lists_beyond_the_scratch():
    s, sc, _ = configs.workload("dambreak2d_2M")  # ~57M entries: far beyond 4M
    sim = capi.Simulation(s, sc)
    rc, off, _, _ = _noalloc(sim)
    assert rc == 1
    msg = sim.L.mps_last_error(sim.h).decode()
    assert "exceed" in msg and "MiB" in msg
    assert int(off[-1]) > 4_000_000
    sim.close()
