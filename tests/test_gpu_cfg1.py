"""Parity at the full 2M dam break (2,035,232 particles: l0 0.25 mm, 2 wall
and 2 dummy layers): the headline workload compared with the FP64 CPU
oracle at its own size, where the kernels' size-dependent paths run (16-bit
SELL column offsets near their limit, 32-bit byte offsets of the walks, LDS
staging at full occupancy, the union-find over 2M particles).

Bit-exact: the bucket lists (src/bucket.cpp:24-51), the neighbour ids and
FP64 distances (src/mps.cpp:585-619), the state entering the pressure stage,
the boundary conditions, the source term, the matrix diagonal and
off-diagonals (src/mps.cpp:262-415) and the debug events.  The pressure
solution is compared through the residual of the GPU's solution in the
oracle's own matrix (a bound that does not depend on the conditioning) and
against the oracle's solution; the state after the step within the
tolerances below.  Once from t = 0 and once from the state 300 GPU steps
in (t = 3 ms), where the flow has moved off the lattice.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from mps_amd import capi, configs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13          # both CGs: ||b - A x|| <= 1e-13 ||b||
RES_BOUND = 1e-12    # the GPU solution's residual in the oracle's matrix
X_RTOL = 1e-6        # ||x_gpu - x_oracle|| / ||x_oracle|| (two CGs at 1e-13, cond(A) ~ 1e4-1e5)
P_RTOL = 1e-6        # max |p_gpu - p_oracle| / max |p|
V_RTOL = 1e-6        # max |v_gpu - v_oracle| / max |v|
X_ATOL = 1e-14       # positions (m): the pressure move is dt^2 a


@pytest.fixture(scope="module")
def cfg1():
    s, sc, _ = configs.workload("dambreak2d_2M")
    assert sc.n == 2035232
    return s, sc


def test_cfg1_bucket_and_neighbours_bit_exact(cfg1):
    s, sc = cfg1
    print(f"[cfg1] bucket + neighbours, N={sc.n}", flush=True)
    o = Oracle(s, sc, tol=TOL)
    o.set_neighbors()
    g = capi.Simulation(s, sc)
    ob, gb = o.bucket(), g.bucket()
    np.testing.assert_array_equal(gb["dims"], ob["dims"])
    assert gb["length"] == ob["length"]
    np.testing.assert_array_equal(gb["first"], ob["first"])
    np.testing.assert_array_equal(gb["last"], ob["last"])
    np.testing.assert_array_equal(gb["next"], ob["next"])
    del ob, gb
    oo, oi, od = o.neighbors()
    go, gi, gd = g.neighbors()
    assert oi.size > 50_000_000
    np.testing.assert_array_equal(go, oo)
    np.testing.assert_array_equal(gi, oi)
    assert (gd.view(np.uint64) == od.view(np.uint64)).all()
    g.close()
    o.close()


def _one_step(s, sc):
    o = Oracle(s, sc, tol=TOL)
    g = capi.Simulation(s, sc)
    g.set_solver(precision=capi.MPS_PREC_FP64, tolerance=TOL, warm_start=False)
    g.set_snapshot(True)
    o.load_initial_state()
    g.load_initial_state()
    o.step()
    g.step()
    return o, g


def _check_step(o, g, label):
    opos, ovel = o.snapshot()
    gpos, gvel = g.snapshot()
    assert (gpos == opos).all() and (gvel == ovel).all(), "state entering the pressure stage"
    del opos, ovel, gpos, gvel
    os_, gs = o.state(), g.state()
    np.testing.assert_array_equal(gs["bc"], os_["bc"])
    op, gp = o.pressure_system(), g.pressure_system()
    assert (gp["b"] == op["b"]).all(), "source term"
    assert (gp["diag"] == op["diag"]).all(), "diagonal"
    np.testing.assert_array_equal(gp["row_ptr"], op["row_ptr"])
    np.testing.assert_array_equal(gp["col"], op["col"])
    assert (gp["val"] == op["val"]).all(), "off-diagonals"
    n = op["b"].shape[0]
    rows = np.repeat(np.arange(n), np.diff(op["row_ptr"]))
    A = sp.csr_matrix((op["val"], (rows, op["col"])), shape=(n, n)) + sp.diags(op["diag"])
    b = op["b"]
    res = np.linalg.norm(b - A @ gp["x"]) / np.linalg.norm(b)
    xerr = np.linalg.norm(gp["x"] - op["x"]) / np.linalg.norm(op["x"])
    perr = np.abs(gs["p"] - os_["p"]).max() / np.abs(os_["p"]).max()
    verr = np.abs(gs["vel"] - os_["vel"]).max() / np.abs(os_["vel"]).max()
    xdiff = np.abs(gs["pos"] - os_["pos"]).max()
    st = g.solver_stats()
    it, orr = o.cg_stats()
    print(f"[cfg1 {label}] rows={st['rows']} gpu iters={st['iterations']} rel_res={st['rel_residual']:.2e} "
          f"oracle iters={it} | residual of x_gpu in the oracle's A {res:.2e}, x err {xerr:.2e}, "
          f"p err {perr:.2e}, v err {verr:.2e}, pos diff {xdiff:.2e}", flush=True)
    assert st["rel_residual"] <= TOL and st["converged"] == 1
    assert res <= RES_BOUND
    assert xerr <= X_RTOL
    np.testing.assert_array_equal(gs["type"], os_["type"])
    assert perr <= P_RTOL
    assert verr <= V_RTOL
    assert xdiff <= X_ATOL
    gc, gg, gu = g.debug_events()
    np.testing.assert_array_equal(gc, o.debug_counts())
    np.testing.assert_array_equal(np.sort(gg), np.sort(o.event_ids(0)))
    np.testing.assert_array_equal(np.sort(gu), np.sort(o.event_ids(1)))


def test_cfg1_one_step_from_t0(cfg1):
    s, sc = cfg1
    print("[cfg1] one FP64 step from t = 0", flush=True)
    o, g = _one_step(s, sc)
    _check_step(o, g, "t=0")
    g.close()
    o.close()


def test_cfg1_one_step_from_mid_run_state(cfg1):
    s, sc = cfg1
    print("[cfg1] 300 mixed multigrid steps, then one FP64 step from that state", flush=True)
    run = capi.Simulation(s, sc)
    run.set_solver(precision=capi.MPS_PREC_MIXED, tolerance=1e-9, warm_start=2,
                   preconditioner=capi.MPS_PRECOND_MULTIGRID)
    run.load_initial_state()
    run.run_steps(300)
    st = run.state()
    run.close()
    mid = configs.Scene(st["type"].copy(), st["pos"].copy(), st["vel"].copy())
    assert np.abs(mid.pos - sc.pos).max() > 1e-5  # the flow has moved off the lattice
    o, g = _one_step(s, mid)
    _check_step(o, g, "t=3ms")
    g.close()
    o.close()
