"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Bit-exact: bucket linked lists, neighbour ids and distances, boundary
conditions, the source term, the matrix (diagonal and off-diagonals) and the
state entering the pressure stage.  Within stated FP64 tolerances: the
pressure solution (two different CG summation orders, both run to near
machine precision) and the state after the step.
"""
import ctypes
import dataclasses

import numpy as np
import pytest
from scipy.spatial import cKDTree

from mps_amd import capi, configs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

# tolerance of the GPU CG in parity tests: ||b-Ax|| <= 1e-14 ||b||; the oracle
# runs Eigen's rule (machine epsilon on the recursive residual)
GPU_TOL = 1e-14
# pressure parity: relative 2-norm error of the solution
P_RTOL = 1e-9


def _jitter(settings, scene, seed, amp=0.2, vamp=0.05):
    rng = np.random.default_rng(seed)
    f = scene.type == configs.FLUID
    d = settings.dim if hasattr(settings, "dim") else 2
    scene.pos[f, :d] += rng.uniform(-amp, amp, size=(int(f.sum()), d)) * settings.particle_distance
    scene.vel[f, :d] = rng.normal(0, vamp, size=(int(f.sum()), d))
    return scene


def _pair(settings, scene, tol=GPU_TOL):
    o = Oracle(settings, scene)
    g = capi.Simulation(settings, scene)
    g.set_solver(precision=capi.MPS_PREC_FP64, tolerance=tol, warm_start=False)
    return o, g


def _cases():
    out = []
    s = configs.Settings()  # tutorials/dambreak/settings.yml
    out.append(("dambreak_tutorial_jitter", s, _jitter(s, configs.tutorial_dambreak(), 1)))
    s2, sc2, _ = configs.workload("dambreak2d_584")
    out.append(("dambreak2d_584_jitter", s2, _jitter(s2, sc2, 2)))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,s,sc", CASES, ids=[c[0] for c in CASES])
def test_bucket_bit_exact(name, s, sc):
    o, g = _pair(s, sc)
    o.set_neighbors()
    ob = o.bucket()
    gb = g.bucket()
    np.testing.assert_array_equal(gb["dims"], ob["dims"])
    assert gb["length"] == ob["length"]
    np.testing.assert_array_equal(gb["first"], ob["first"])
    np.testing.assert_array_equal(gb["last"], ob["last"])
    np.testing.assert_array_equal(gb["next"], ob["next"])


@pytest.mark.parametrize("name,s,sc", CASES, ids=[c[0] for c in CASES])
def test_neighbors_bit_exact(name, s, sc):
    o, g = _pair(s, sc)
    o.set_neighbors()
    oo, oi, od = o.neighbors()
    go, gi, gd = g.neighbors()
    np.testing.assert_array_equal(go, oo)
    np.testing.assert_array_equal(gi, oi)
    assert (gd.view(np.uint64) == od.view(np.uint64)).all()  # distances bit for bit


def _step_both(s, sc, steps=1, export_last=True):
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    g.set_snapshot(True)
    co = cg = None
    for k in range(steps):
        ex = export_last and k == steps - 1
        co = o.step(ex)
        cg = g.step(ex)
    return o, g, co, cg


@pytest.mark.parametrize("name,s,sc", CASES, ids=[c[0] for c in CASES])
def test_initial_number_density_bit_exact(name, s, sc):
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    assert (g.state()["ndr"] == o.state()["ndr"]).all()


@pytest.mark.parametrize("name,s,sc", CASES, ids=[c[0] for c in CASES])
def test_one_step_pre_pressure_state_and_system_bit_exact(name, s, sc):
    o, g, co, cg = _step_both(s, sc)
    opos, ovel = o.snapshot()
    gpos, gvel = g.snapshot()
    assert (gpos == opos).all() and (gvel == ovel).all()
    os_, gs = o.state(), g.state()
    np.testing.assert_array_equal(gs["bc"], os_["bc"])
    op, gp = o.pressure_system(), g.pressure_system()
    assert (gp["b"] == op["b"]).all()
    assert (gp["diag"] == op["diag"]).all()
    np.testing.assert_array_equal(gp["row_ptr"], op["row_ptr"])
    np.testing.assert_array_equal(gp["col"], op["col"])
    assert (gp["val"] == op["val"]).all()
    # the solutions agree to the conditioning of A times the solver tolerances
    nx = np.linalg.norm(op["x"])
    assert nx > 0
    assert np.linalg.norm(gp["x"] - op["x"]) <= P_RTOL * nx
    st = g.solver_stats()
    assert st["rel_residual"] <= 1e-13 and st["rows"] == int((os_["bc"] == configs.BC_INNER).sum())


@pytest.mark.parametrize("name,s,sc", CASES, ids=[c[0] for c in CASES])
def test_one_step_final_state(name, s, sc):
    o, g, co, cg = _step_both(s, sc)
    os_, gs = o.state(), g.state()
    np.testing.assert_array_equal(gs["type"], os_["type"])
    pscale = np.abs(os_["p"]).max()
    assert np.abs(gs["p"] - os_["p"]).max() <= P_RTOL * pscale
    vscale = np.abs(os_["vel"]).max()
    assert np.abs(gs["vel"] - os_["vel"]).max() <= 1e-9 * vscale
    assert np.abs(gs["pos"] - os_["pos"]).max() <= 1e-15
    assert np.abs(gs["ndr"] - os_["ndr"]).max() <= 1e-14
    assert cg == pytest.approx(co, rel=1e-9)


def test_dambreak_584_fifty_steps():
    s, sc, _ = configs.workload("dambreak2d_584")
    o, g, co, cg = _step_both(s, sc, steps=50)
    os_, gs = o.state(), g.state()
    np.testing.assert_array_equal(gs["type"], os_["type"])
    np.testing.assert_array_equal(gs["bc"], os_["bc"])
    assert np.abs(gs["pos"] - os_["pos"]).max() <= 1e-12
    vs = np.abs(os_["vel"]).max()
    assert np.abs(gs["vel"] - os_["vel"]).max() <= 1e-8 * vs
    ps = np.abs(os_["p"]).max()
    assert np.abs(gs["p"] - os_["p"]).max() <= 1e-8 * ps


def _small(types, pos, vel, **kw):
    s = configs.Settings(**kw)
    return s, configs.Scene(np.array(types, np.int8), np.array(pos, float), np.array(vel, float))


COLLISION_CASES = {
    "head_on": _small([0, 0], [[0.10, 0.1, 0.0], [0.104, 0.1, 0.0]], [[0.5, 0, 0], [-0.5, 0, 0]],
                      gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0),
    "wall": _small([0, 1], [[0.10, 0.1, 0.0], [0.103, 0.1, 0.0]], [[0.3, 0, 0], [0, 0, 0]],
                   gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0),
    "chain": _small([0, 0, 0], [[0.10, 0.1, 0.0], [0.104, 0.1, 0.0], [0.108, 0.1, 0.0]],
                    [[0, 0, 0], [-0.2, 0, 0], [-0.6, 0, 0]], gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0),
    "out_of_domain": _small([0], [[0.399, 0.1, 0.0]], [[5.0, 0, 0]]),
}


@pytest.mark.parametrize("name", sorted(COLLISION_CASES))
def test_small_scenes_bit_exact(name):
    s, sc = COLLISION_CASES[name]
    o, g, co, cg = _step_both(s, sc, steps=3)
    os_, gs = o.state(), g.state()
    assert (gs["pos"] == os_["pos"]).all() and (gs["vel"] == os_["vel"]).all()
    np.testing.assert_array_equal(gs["type"], os_["type"])
    assert cg == co
    gc, gg, gu = g.debug_events()
    oc = o.debug_counts()
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gg, np.sort(o.event_ids(0)))


def test_dense_collisions_match_sequential_order():
    # many overlapping fluid pairs: the level-scheduled GPU resolution must
    # equal the reference's sequential loop bit for bit
    s = configs.Settings(gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0)
    sc = configs.tutorial_dambreak()
    rng = np.random.default_rng(7)
    f = np.flatnonzero(sc.type == configs.FLUID)
    sc.pos[f, :2] += rng.uniform(-0.35, 0.35, size=(f.size, 2)) * s.particle_distance
    sc.vel[f, :2] = rng.normal(0, 0.3, size=(f.size, 2))
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    g.set_snapshot(True)
    o.step()
    g.step()
    opos, ovel = o.snapshot()
    gpos, gvel = g.snapshot()
    assert (gpos == opos).all() and (gvel == ovel).all()


def _packed_block(nx, ny, spacing, seed, domain=None):
    """A free fluid block at spacing * l0 < the collision distance, jittered,
    with random velocities: most neighbours collide, in long chains.  domain
    (upper x and y) makes room for blocks larger than the tutorial's domain
    holds (about 70 x 60)."""
    s = configs.Settings(gravity=(0.0, 0.0, 0.0), kinematic_viscosity=0.0)
    l0 = s.particle_distance
    if domain is not None:
        s.domain_x = (s.domain_x[0], domain)
        s.domain_y = (s.domain_y[0], domain)
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pos = np.zeros((nx * ny, 3))
    pos[:, 0] = 0.05 + gx.ravel() * spacing * l0
    pos[:, 1] = 0.05 + gy.ravel() * spacing * l0
    pos[:, :2] += rng.uniform(-0.05, 0.05, size=(nx * ny, 2)) * l0
    vel = np.zeros_like(pos)
    vel[:, :2] = rng.normal(0, 0.3, size=(nx * ny, 2))
    sc = configs.Scene(np.full(nx * ny, configs.FLUID, np.int8), pos, vel)
    # candidate pairs as the reference forms them (after the explicit move,
    # which is x += v dt here): fluid pairs once, dist < 0.5 l0
    x = pos + vel * s.dt
    pairs = len(cKDTree(x).query_pairs(0.5 * l0))  # the count only sizes the case
    return s, sc, pairs


@pytest.mark.parametrize("nx,ny,lo,hi,domain", [(40, 40, 1500, 4096, None), (64, 64, 4097, 10 ** 6, None),
                                                (128, 128, 16385, 10 ** 6, 0.7)],
                         ids=["lds_levels", "global_levels", "global_levels_large"])
def test_packed_collisions_match_sequential_order(nx, ny, lo, hi, domain):
    """Thousands of colliding pairs in long chains: pair levels assigned in
    parallel out of LDS (up to 4096 pairs) or out of global scratch (more)
    must both reproduce the reference's sequential resolution bit for bit."""
    s, sc, pairs = _packed_block(nx, ny, 0.45, 9, domain)
    assert lo <= pairs <= hi, pairs
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    g.set_snapshot(True)
    o.step()
    g.step()
    opos, ovel = o.snapshot()
    gpos, gvel = g.snapshot()
    assert (gpos == opos).all() and (gvel == ovel).all()


def test_enclosed_fluid_unconnected_events():
    l0 = 0.01
    full = configs.lattice_2d(l0, (0.1, 0.1), (0.0, 0.1), (0.0, 0.1), 2, 2)
    top = full.pos[full.pos[:, 1] < 0].copy()
    top_t = full.type[full.pos[:, 1] < 0].copy()
    top[:, 1] = 0.1 - top[:, 1]
    sc = configs.Scene(np.concatenate([full.type, top_t]), np.concatenate([full.pos, top]),
                       np.zeros((full.n + top.shape[0], 3)))
    s = configs.Settings(gravity=(0.0, 0.0, 0.0), domain_x=(-0.1, 0.2), domain_y=(-0.1, 0.2))
    o, g, co, cg = _step_both(s, sc)
    gc, gg, gu = g.debug_events()
    np.testing.assert_array_equal(gu, np.sort(o.event_ids(1)))
    assert gc[1] == o.debug_counts()[1] > 0
    assert (g.pressure_system()["diag"] == o.pressure_system()["diag"]).all()


@pytest.mark.parametrize("jumps", ["0,1", "1,1", "12,2"])
def test_connectivity_jumps_change_nothing(jumps, monkeypatch):
    """Pointer jumping before the union-find hooking shortens the chains the
    finds chase; the components, and with them the doubled diagonals and
    the unconnected-particle events, are the hooking's alone.  The enclosed
    pocket of the events test and a 2D dam break compared bit for bit
    against the default rounds."""
    l0 = 0.01
    full = configs.lattice_2d(l0, (0.1, 0.1), (0.0, 0.1), (0.0, 0.1), 2, 2)
    top = full.pos[full.pos[:, 1] < 0].copy()
    top_t = full.type[full.pos[:, 1] < 0].copy()
    top[:, 1] = 0.1 - top[:, 1]
    pocket = configs.Scene(np.concatenate([full.type, top_t]), np.concatenate([full.pos, top]),
                           np.zeros((full.n + top.shape[0], 3)))
    sp = configs.Settings(gravity=(0.0, 0.0, 0.0), domain_x=(-0.1, 0.2), domain_y=(-0.1, 0.2))
    sd, sc, _ = configs.workload("dambreak2d_35k")

    def run(settings, scene, steps):
        g = capi.Simulation(settings, scene)
        g.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-9, warm_start=2)
        g.load_initial_state()
        g.set_snapshot(True)  # keep the assembled system for pressure_system()
        for _ in range(steps):
            g.step()
        out = (g.pressure_system()["diag"], g.debug_events(), g.state())
        g.close()
        return out

    base = [run(sp, pocket, 1), run(sd, sc, 3)]
    monkeypatch.setenv("MPS_AMD_CC_JUMPS", jumps)
    other = [run(sp, pocket, 1), run(sd, sc, 3)]
    for (d0, e0, s0), (d1, e1, s1) in zip(base, other):
        assert (d0 == d1).all()
        assert list(e0[0]) == list(e1[0])
        np.testing.assert_array_equal(e0[2], e1[2])
        for k in ("p", "pos", "vel"):
            assert (s0[k] == s1[k]).all(), k
    assert base[0][1][0][1] > 0  # the pocket is unconnected


def test_3d_one_step():
    s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
    sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
    s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
    sc = _jitter(s, sc, 5, amp=0.1)
    o, g = _pair(s, sc)
    o.set_neighbors()
    oo, oi, od = o.neighbors()
    go, gi, gd = g.neighbors()
    np.testing.assert_array_equal(go, oo)
    np.testing.assert_array_equal(gi, oi)
    assert (gd == od).all()
    o2, g2, co, cg = _step_both(s, sc)
    op, gp = o2.pressure_system(), g2.pressure_system()
    assert (gp["b"] == op["b"]).all() and (gp["diag"] == op["diag"]).all() and (gp["val"] == op["val"]).all()
    assert np.linalg.norm(gp["x"] - op["x"]) <= P_RTOL * np.linalg.norm(op["x"])


def _crowd(s, sc, centre, count, seed):
    """count extra fluid particles scattered over the bucket cell (side
    re_max) that holds centre, so that cell has more than 64 candidates."""
    d = s.dim
    l0 = s.particle_distance
    side = max(s.ratio_pressure, s.ratio_viscosity, s.ratio_surface) * l0
    lo = np.array([r[0] for r in (s.domain_x, s.domain_y, s.domain_z)][:d])
    c0 = lo + np.floor((np.asarray(centre[:d]) - lo) / side) * side
    rng = np.random.default_rng(seed)
    pos = np.zeros((count, 3))
    pos[:, :d] = c0 + rng.uniform(0.05, 0.95, size=(count, d)) * side
    return configs.Scene(np.concatenate([sc.type, np.full(count, configs.FLUID, dtype=sc.type.dtype)]),
                         np.concatenate([sc.pos, pos]), np.concatenate([sc.vel, np.zeros((count, 3))]))


def _crowded_case(dim):
    if dim == 2:
        s, sc, _ = configs.workload("dambreak2d_584")
        return s, _crowd(s, _jitter(s, sc, 3), (0.12, 0.2, 0.0), 100, 4)
    s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
    sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
    s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
    return s, _crowd(s, _jitter(s, sc, 5, amp=0.1), (0.05, 0.1, 0.05), 100, 6)


@pytest.mark.parametrize("dim", [2, 3])
def test_crowded_cell_bit_exact(dim):
    """One cell holding over 100 particles: the walks test candidates in
    chunks of 64 and their lists overflow, and the state entering the
    pressure stage (viscosity, move, collisions), the boundary conditions
    and the pressure system must still equal the oracle's bit for bit."""
    s, sc = _crowded_case(dim)
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    most = int(np.diff(g.neighbors()[0]).max())
    assert most > 64, most
    g.set_snapshot(True)
    o.step(True)
    g.step(True)
    opos, ovel = o.snapshot()
    gpos, gvel = g.snapshot()
    assert (gpos == opos).all() and (gvel == ovel).all()
    np.testing.assert_array_equal(g.state()["bc"], o.state()["bc"])
    op, gp = o.pressure_system(), g.pressure_system()
    assert (gp["b"] == op["b"]).all() and (gp["diag"] == op["diag"]).all()
    np.testing.assert_array_equal(gp["row_ptr"], op["row_ptr"])
    np.testing.assert_array_equal(gp["col"], op["col"])
    assert (gp["val"] == op["val"]).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_list_capacity_changes_nothing(dim, monkeypatch):
    """The pressure stages read the neighbour lists the BC pass records;
    slots beyond the list capacity are walked again.  Capacity 0 (every slot
    walked), 5 (most slots walked) and the default (none walked) must give the
    same bits, with the pressure stabilisation options on as well."""
    if dim == 2:
        s, sc, _ = configs.workload("dambreak2d_584")
        s = dataclasses.replace(s, hos_on=True)
        sc = _jitter(s, sc, 21, vamp=0.3)
    else:
        s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
        sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
        s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
        sc = _jitter(s, sc, 22, amp=0.1, vamp=0.3)
    res = []
    for cap in (None, "5", "0"):
        if cap is None:
            monkeypatch.delenv("MPS_AMD_NB_CAP", raising=False)
        else:
            monkeypatch.setenv("MPS_AMD_NB_CAP", cap)
        g = capi.Simulation(s, sc)
        g.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-12, warm_start=True)
        g.load_initial_state()
        for _ in range(6):
            g.step()
        res.append((g.state(), g.debug_events()))
        g.close()
    monkeypatch.delenv("MPS_AMD_NB_CAP", raising=False)
    for st, ev in res[1:]:
        for k in ("type", "pos", "vel", "p", "ndr", "bc", "pmin"):
            assert (st[k] == res[0][0][k]).all(), k
        for a, b in zip(ev, res[0][1]):
            assert (np.asarray(a) == np.asarray(b)).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_stage_capacity_changes_nothing(dim, monkeypatch):
    """The walks read a block's candidates from LDS when they fit the staging
    capacity; a 3D block stages the stencil rows that fit and reads the others
    from global memory.  The default capacity (every block staged here), 900
    slots (3D: two to six of a block's nine rows staged, some after a row
    left out) and 0 (every walk in global memory) must give the same bits
    over 6 steps."""
    if dim == 2:
        s, sc, _ = configs.workload("dambreak2d_584")
        sc = _jitter(s, sc, 23, vamp=0.3)
    else:
        s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
        sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
        s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
        sc = _jitter(s, sc, 24, amp=0.1, vamp=0.3)
    res = []
    monkeypatch.setenv("MPS_AMD_STAGE_STATS", "1")
    for cap in (None, "900", "0"):
        if cap is None:
            monkeypatch.delenv("MPS_AMD_STAGE_CAP", raising=False)
        else:
            monkeypatch.setenv("MPS_AMD_STAGE_CAP", cap)
        g = capi.Simulation(s, sc)
        g.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-12, warm_start=True)
        before = g.debug_counters()
        g.load_initial_state()
        for _ in range(6):
            g.step()
        after = g.debug_counters()
        stats = {k: after[k] - before[k] for k in after if k.startswith("staged")}
        print(f"[staging dim {dim} cap {cap}] {stats}", flush=True)
        if dim == 3 and cap == "900":
            # the capacity splits blocks: some stage part of their rows, and
            # some of those leave a row out before a later one that fits
            assert stats["staged_partial"] > 0 and stats["staged_partial_gap"] > 0, stats
        if cap == "0":  # (3D: a block may still "stage" the empty stencil rows at the grid's edge)
            assert stats["staged_full"] == 0, stats
        if cap is None:
            assert stats["staged_full"] > 0, stats
        res.append((g.state(), g.debug_events()))
        g.close()
    monkeypatch.delenv("MPS_AMD_STAGE_CAP", raising=False)
    monkeypatch.delenv("MPS_AMD_STAGE_STATS", raising=False)
    for st, ev in res[1:]:
        for k in ("type", "pos", "vel", "p", "ndr", "bc", "pmin"):
            assert (st[k] == res[0][0][k]).all(), k
        for a, b in zip(ev, res[0][1]):
            assert (np.asarray(a) == np.asarray(b)).all()


def test_graph_and_stream_paths_identical():
    s, sc, _ = configs.workload("dambreak2d_584")
    res = []
    for use_graph in (True, False):
        g = capi.Simulation(s, sc)
        g.set_solver(tolerance=1e-12, use_graph=use_graph, check_interval=8)
        g.load_initial_state()
        for _ in range(5):
            g.step()
        res.append(g.state())
    for k in ("pos", "vel", "p"):
        assert (res[0][k] == res[1][k]).all()


@pytest.mark.parametrize("case", ["3d_jacobi", "2d_multigrid_mixed"])
def test_graph_replay_after_storage_grows(case, monkeypatch):
    """With every capacity exactly the last system's size
    (MPS_AMD_TIGHT_CAPS=1), a larger system reallocates the matrix storage
    (3D: also the row anchors) and the CG vectors.  The cached CG graphs bake
    those pointers in, so a reallocation that moves them must drop and capture
    them again.  The run starts with a third of the fluid as ghosts, then
    uploads the whole fluid, so the system grows under cached graphs; graph
    replay and direct launches must give the same bits."""
    monkeypatch.setenv("MPS_AMD_TIGHT_CAPS", "1")
    if case == "3d_jacobi":
        s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
        sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
        s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
        sc = _jitter(s, sc, 25, amp=0.1, vamp=0.3)
        opts = dict(precision=capi.MPS_PREC_FP64, tolerance=1e-12, warm_start=True)
    else:
        s, sc, _ = configs.workload("dambreak2d_35k")
        sc = _jitter(s, sc, 26, amp=0.1, vamp=0.3)
        opts = dict(precision=capi.MPS_PREC_MIXED, tolerance=1e-9, warm_start=2,
                    preconditioner=capi.MPS_PRECOND_MULTIGRID)
    part = configs.Scene(sc.type.copy(), sc.pos.copy(), sc.vel.copy())
    fl = np.flatnonzero(sc.type == configs.FLUID)
    part.type[fl[fl.size * 2 // 3:]] = configs.GHOST  # the column's right-hand third (ids: x outer)
    res, counters = [], []
    for use_graph in (True, False):
        g = capi.Simulation(s, part)
        g.set_solver(use_graph=use_graph, check_interval=4, **opts)
        g.load_initial_state()
        for _ in range(4):
            g.step()
        before = g.debug_counters()
        g.upload(type=sc.type, pos=sc.pos, vel=sc.vel)
        for _ in range(6):
            g.step()
        res.append(g.state())
        counters.append((before, g.debug_counters()))
        g.close()
    monkeypatch.delenv("MPS_AMD_TIGHT_CAPS", raising=False)
    for k in ("type", "pos", "vel", "p", "ndr", "bc", "pmin"):
        assert (res[0][k] == res[1][k]).all(), k
    b, a = counters[0]
    print(f"[graph growth {case}] before upload {b}, after {a}", flush=True)
    # storage grew while graphs were cached (whether the allocator moved it,
    # and so whether the graphs had to be captured again, is its choice)
    assert a["matrix_reallocs"] > b["matrix_reallocs"] and a["cg_reallocs"] > b["cg_reallocs"], (b, a)
    assert b["graph_rebuilds"] >= 1


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED])
def test_capped_solve_reports_solver_status(precision):
    """A solve capped below what its tolerance needs completes the step and
    returns MPS_E_SOLVER with the residual in the message; the statistics say
    it did not converge, and an uncapped solve of the same step converges."""
    s, sc, _ = configs.workload("dambreak2d_584")
    sc = _jitter(s, sc, 27)
    g = capi.Simulation(s, sc)
    g.set_solver(precision=precision, tolerance=1e-12, max_iterations=3)
    g.load_initial_state()
    c = ctypes.c_double()
    rc = g.L.mps_step(g.h, 0, ctypes.byref(c))
    assert rc == capi.MPS_E_SOLVER
    msg = g.L.mps_last_error(g.h).decode()
    assert "tolerance" in msg and "relative residual" in msg
    st = g.solver_stats()
    assert st["converged"] == 0 and st["rel_residual"] > 1e-12 and st["iterations"] <= 3
    assert c.value > 0.0  # the step ran to its Courant number
    g.strict_solver = True
    with pytest.raises(capi.MpsError, match="MPS_E_SOLVER"):
        g.step()
    g.strict_solver = False
    g.step()  # not strict: the run goes on, the message is kept
    assert g.solver_warning and "tolerance" in g.solver_warning
    g.set_solver(precision=precision, tolerance=1e-12)
    g.solver_warning = None
    g.step()
    st = g.solver_stats()
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-12 and g.solver_warning is None


def test_autotuner_changes_speed_not_results():
    s, sc, _ = configs.workload("dambreak2d_584")
    sc = _jitter(s, sc, 11)
    runs = []
    for mode in ("default", "tuned", "odd"):
        g = capi.Simulation(s, sc)
        g.set_solver(tolerance=1e-12)
        g.load_initial_state()
        if mode == "tuned":
            sizes = g.autotune(2)  # 8 real steps
            assert all(int(x) in (64, 128, 256, 512) for x in sizes)
        else:
            if mode == "odd":
                g.set_block_sizes([64, 128, 512, 64, 512, 128, 64, 512, 128])
            for _ in range(8):
                g.step()
        runs.append(g.state())
    for k in ("pos", "vel", "p", "ndr", "bc", "type"):
        assert (runs[0][k] == runs[1][k]).all() and (runs[0][k] == runs[2][k]).all(), k


def _f3_case(**kw):
    s, sc, _ = configs.workload("dambreak2d_584")
    s = dataclasses.replace(s, **kw)
    return s, _jitter(s, sc, 21)


@pytest.mark.parametrize("opts", [dict(hos_on=True, hos_gamma=1e-3), dict(qc_on=False),
                                  dict(hos_on=True, hos_gamma=5e-3, qc_on=False)],
                         ids=["hos", "qc_off", "hos_qc_off"])
def test_pressure_stabilisation_options(opts):
    """The higher-order source term (src/mps.cpp:288-308, which reads the
    previous step's number density ratio and the neighbours' velocities) and
    the quasi-compressibility-off branch (src/mps.cpp:352-355): the system of
    every step is bit-exact, the state within the solver tolerance."""
    s, sc = _f3_case(**opts)
    o, g = _pair(s, sc)
    o.load_initial_state()
    g.load_initial_state()
    g.set_snapshot(True)
    for k in range(4):
        o.step(False)
        g.step(False)
        op, gp = o.pressure_system(), g.pressure_system()
        assert (gp["b"] == op["b"]).all(), k
        assert (gp["diag"] == op["diag"]).all(), k
        assert (gp["val"] == op["val"]).all(), k
        nx = np.linalg.norm(op["x"])
        assert np.linalg.norm(gp["x"] - op["x"]) <= P_RTOL * max(nx, 1e-300), k
        # keep the two runs on the same state so later steps compare the same system
        st = o.state()
        g.upload(pos=st["pos"], vel=st["vel"], p=st["p"], ndr=st["ndr"])


@pytest.mark.parametrize("warm", [1, 2])
def test_checkpoint_resume_is_bit_identical(warm):
    """A run resumed from mps_checkpoint_save's image (in a fresh handle built
    from the initial scene) continues bit for bit, with either warm start (the
    extrapolated one needs the last two solutions from the image); a
    corrupted or foreign image is refused."""
    s, sc, _ = configs.workload("dambreak2d_584")
    sc = _jitter(s, sc, 31, vamp=0.5)
    a = capi.Simulation(s, sc)
    a.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-12, warm_start=warm)
    a.load_initial_state()
    for _ in range(7):
        a.step()
    img = a.checkpoint()
    for _ in range(9):
        a.step()
    b = capi.Simulation(s, sc)
    b.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-12, warm_start=warm)
    b.restore(img)
    for _ in range(9):
        b.step()
    sa, sb = a.state(), b.state()
    for k in sa:
        assert (sa[k] == sb[k]).all(), k
    bad = bytearray(img)
    bad[0:8] = b"NOTACKPT"
    with pytest.raises(capi.MpsError):
        b.restore(bytes(bad))
    with pytest.raises(capi.MpsError):
        b.restore(img[:-10])
    other = capi.Simulation(*configs.workload("dambreak2d_35k")[:2])
    with pytest.raises(capi.MpsError):
        other.restore(img)
    # same particle count, dim and scene count, other settings: refused
    s2 = dataclasses.replace(s, dt=s.dt * 0.5)
    with pytest.raises(capi.MpsError, match="settings"):
        capi.Simulation(s2, sc).restore(img)
    # a later finish time (resuming towards a longer run) is accepted
    capi.Simulation(dataclasses.replace(s, finish_time=2 * s.finish_time), sc).restore(img)
    # ensembles: the same totals with other per-scene counts are refused
    small = configs.lattice_2d(s.particle_distance, (1.0, 0.6), (0.0, 0.15), (0.0, 0.25), 2, 2)
    assert small.n != sc.n
    e1 = capi.Simulation([(s, sc), (s, small)])
    e1.load_initial_state()
    e1.step()
    eimg = e1.checkpoint()
    with pytest.raises(capi.MpsError, match="per-scene"):
        capi.Simulation([(s, small), (s, sc)]).restore(eimg)
    e2 = capi.Simulation([(s, sc), (s, small)])
    e2.restore(eimg)
    se1, se2 = e1.state(), e2.state()
    for k in se1:
        assert (se1[k] == se2[k]).all(), k
