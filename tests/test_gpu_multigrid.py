"""Multigrid-preconditioned pressure CG (MPS_PRECOND_MULTIGRID).

The preconditioner changes how many iterations the CG takes, not what it
solves: at the same tolerance the solution must match the Jacobi path's to
the accuracy the tolerance fixes, the true FP64 residual must meet it, and
the path must be deterministic (graph replay and stream launches identical,
a rerun bit-identical)."""
import numpy as np
import pytest

from mps_amd import capi, configs

pytestmark = pytest.mark.gpu


def _run(name, precision, precond, steps, tol=1e-9, use_graph=True, check_interval=8):
    s, sc, _ = configs.workload(name)
    sim = capi.Simulation(s, sc)
    sim.set_solver(precision=precision, tolerance=tol, warm_start=True, preconditioner=precond,
                   use_graph=use_graph, check_interval=check_interval)
    sim.load_initial_state()
    stats = []
    for _ in range(steps):
        sim.step()
        stats.append(sim.solver_stats())
    st = sim.state()
    sim.close()
    return st, stats


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED])
def test_multigrid_solves_the_same_system(precision):
    # 35k particles: 32k Inner rows, a three-level hierarchy
    jac, sj = _run("dambreak2d_35k", precision, capi.MPS_PRECOND_JACOBI, 5)
    mg, sm = _run("dambreak2d_35k", precision, capi.MPS_PRECOND_MULTIGRID, 5)
    for a, b in zip(sj, sm):
        assert b["converged"] == 1 and b["rel_residual"] <= 1e-9
        assert b["iterations"] * 5 < a["iterations"]  # the point of it
    # after five steps the pressures differ only by what the 1e-9 residual
    # leaves free (cond ~1e4) and its effect on the trajectory
    p_j, p_m = jac["p"], mg["p"]
    assert np.abs(p_j - p_m).max() <= 1e-5 * np.abs(p_j).max()


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED])
def test_both_fine_level_forms_solve_the_same_system(precision, monkeypatch):
    """MPS_AMD_MG_FINE forces the V-cycle's fine level: additive (no fine
    SpMV, the default for large single scenes) or multiplicative (the
    symmetric V(1,1), the default below).  Both are SPD preconditioners of
    the same CG: same tolerance met, same solution to what it leaves free;
    the additive one takes more iterations for its cheaper cycle."""
    jac, _ = _run("dambreak2d_35k", precision, capi.MPS_PRECOND_JACOBI, 4)
    its = {}
    for form in ("additive", "multiplicative"):
        monkeypatch.setenv("MPS_AMD_MG_FINE", form)
        st, stats = _run("dambreak2d_35k", precision, capi.MPS_PRECOND_MULTIGRID, 4)
        for s in stats:
            assert s["converged"] == 1 and s["rel_residual"] <= 1e-9, (form, s)
        assert np.abs(st["p"] - jac["p"]).max() <= 1e-5 * np.abs(jac["p"]).max(), form
        its[form] = sum(s["iterations"] for s in stats)
    assert its["multiplicative"] < its["additive"] < 3 * its["multiplicative"], its


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED])
@pytest.mark.parametrize("use_graph", [True, False])
def test_step_formed_restriction_changes_nothing(precision, use_graph, monkeypatch):
    """With the additive fine level the CG step forms the next cycle's
    level-1 right-hand side where it writes r, over row ranges that hold
    whole cells; MPS_AMD_CG_FUSE=0 leaves it to the cycle's restriction
    kernel.  Both sum each cell's rows in row order: the trajectories are
    bit-identical, across chunk boundaries (check interval 3) too."""
    monkeypatch.setenv("MPS_AMD_MG_FINE", "additive")
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("MPS_AMD_CG_FUSE", fuse)
        out[fuse], stats = _run("dambreak2d_35k", precision, capi.MPS_PRECOND_MULTIGRID, 4, use_graph=use_graph,
                                check_interval=3)
        for st in stats:
            assert st["converged"] == 1 and st["rel_residual"] <= 1e-9, (fuse, st)
    for k in ("p", "pos", "vel"):
        assert (out["1"][k] == out["0"][k]).all(), k


def test_multigrid_is_deterministic():
    a, _ = _run("dambreak2d_35k", capi.MPS_PREC_MIXED, capi.MPS_PRECOND_MULTIGRID, 4)
    b, _ = _run("dambreak2d_35k", capi.MPS_PREC_MIXED, capi.MPS_PRECOND_MULTIGRID, 4)
    c, _ = _run("dambreak2d_35k", capi.MPS_PREC_MIXED, capi.MPS_PRECOND_MULTIGRID, 4, use_graph=False,
                check_interval=3)
    for k in ("p", "pos", "vel"):
        assert (a[k] == b[k]).all() and (a[k] == c[k]).all(), k


def test_multigrid_3d():
    s = configs.dambreak_settings(0.01, 1e-4, 1.0, dim=3)
    s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.5), (-0.1, 0.5), (-0.1, 0.3)
    sc = configs.lattice_3d(0.01, (0.4, 0.4, 0.2), (0.0, 0.15), (0.0, 0.3), (0.0, 0.2), 2, 2)
    out = {}
    for pre in (capi.MPS_PRECOND_JACOBI, capi.MPS_PRECOND_MULTIGRID):
        sim = capi.Simulation(s, sc)
        sim.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-9, warm_start=True, preconditioner=pre)
        sim.load_initial_state()
        sim.step()
        out[pre] = (sim.state()["p"], sim.solver_stats())
        sim.close()
    (pj, sj), (pm, sm) = out[0], out[1]
    assert sm["rows"] >= 8192 and sm["rel_residual"] <= 1e-9
    assert sm["iterations"] * 3 < sj["iterations"]
    assert np.abs(pj - pm).max() <= 1e-5 * np.abs(pj).max()


def test_small_systems_keep_jacobi():
    # below the size threshold the multigrid option changes nothing
    a, _ = _run("dambreak2d_584", capi.MPS_PREC_FP64, capi.MPS_PRECOND_JACOBI, 3)
    b, _ = _run("dambreak2d_584", capi.MPS_PREC_FP64, capi.MPS_PRECOND_MULTIGRID, 3)
    assert (a["p"] == b["p"]).all()


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED])
def test_extrapolated_warm_start_solves_the_same_system(precision):
    """warm_start = 2 starts the CG from 2 p_n - p_(n-1): fewer iterations over
    a run, the same solutions to what the tolerance leaves free."""
    s, sc, _ = configs.workload("dambreak2d_35k")
    out = {}
    for warm in (1, 2):
        sim = capi.Simulation(s, sc)
        sim.set_solver(precision=precision, tolerance=1e-9, warm_start=warm,
                       preconditioner=capi.MPS_PRECOND_MULTIGRID)
        sim.load_initial_state()
        its = 0
        for _ in range(12):
            sim.step()
            st = sim.solver_stats()
            assert st["converged"] == 1 and st["rel_residual"] <= 1e-9
            its += st["iterations"]
        out[warm] = (sim.state(), its)
        sim.close()
    assert out[2][1] <= out[1][1]
    p1, p2 = out[1][0]["p"], out[2][0]["p"]
    assert np.abs(p1 - p2).max() <= 1e-5 * np.abs(p1).max()


def test_interleaved_handles_match_solo_runs():
    # two handles stepped alternately on one stream each: every handle keeps
    # its own host-read mailbox, CG graphs, chunk predictions and hierarchy,
    # so each trajectory is bit-identical to the same handle stepped alone
    solo2, _ = _run("dambreak2d_35k", capi.MPS_PREC_MIXED, capi.MPS_PRECOND_MULTIGRID, 4)
    s3 = configs.dambreak_settings(0.01, 1e-4, 1.0, dim=3)
    s3.domain_x, s3.domain_y, s3.domain_z = (-0.1, 0.5), (-0.1, 0.5), (-0.1, 0.3)
    sc3 = configs.lattice_3d(0.01, (0.4, 0.4, 0.2), (0.0, 0.15), (0.0, 0.3), (0.0, 0.2), 2, 2)

    def make3():
        sim = capi.Simulation(s3, sc3)
        sim.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-9, warm_start=2,
                       preconditioner=capi.MPS_PRECOND_MULTIGRID)
        sim.load_initial_state()
        return sim

    a = make3()
    for _ in range(4):
        a.step()
    solo3 = a.state()
    a.close()
    s2, sc2, _ = configs.workload("dambreak2d_35k")
    h2 = capi.Simulation(s2, sc2)
    h2.set_solver(precision=capi.MPS_PREC_MIXED, tolerance=1e-9, warm_start=True,
                  preconditioner=capi.MPS_PRECOND_MULTIGRID)
    h2.load_initial_state()
    h3 = make3()
    for _ in range(4):
        h2.step()
        h3.step()
    both2, both3 = h2.state(), h3.state()
    h2.close()
    h3.close()
    for k in ("p", "pos", "vel"):
        assert (both2[k] == solo2[k]).all(), k
        assert (both3[k] == solo3[k]).all(), k


def test_enqueued_iterations_cover_every_solve():
    """mps_debug_counters' out[7]: CG iterations enqueued in chunks.  They
    cover every iteration the solves report; what exceeds them is the tail
    of a chunk that returns at once after convergence, bounded by the
    chunk policy (a chunk of the last step's count, then 2, 4, 8, ...)."""
    s, sc, _ = configs.workload("dambreak2d_35k")
    for precision in (capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED):
        sim = capi.Simulation(s, sc)
        sim.set_solver(precision=precision, tolerance=1e-9, warm_start=2, preconditioner=capi.MPS_PRECOND_MULTIGRID)
        sim.load_initial_state()
        start = sim.debug_counters()["cg_launched"]
        its = solves = 0
        for _ in range(6):
            sim.step()
            st = sim.solver_stats()
            its += st["iterations"]
            # mixed: one inner solve per refinement; FP64: one plus its restarts
            solves += max(1, st["refinements"]) if precision == capi.MPS_PREC_MIXED else 1 + st["refinements"]
        launched = sim.debug_counters()["cg_launched"] - start
        sim.close()
        assert its <= launched <= 2 * its + 4 * solves, (precision, its, launched, solves)


def test_chunk_graphs_are_captured_once_per_length():
    """Every CG chunk graph is kept for the chunk length it was captured
    for: a stretch of steps run a second time from the same state (the same
    chunk lengths, the same predictions at its start) replays every chunk
    without capturing one, and gives the same bits."""
    s, sc, _ = configs.workload("dambreak2d_35k")
    sim = capi.Simulation(s, sc)
    sim.set_solver(precision=capi.MPS_PREC_MIXED, tolerance=1e-9, warm_start=2,
                   preconditioner=capi.MPS_PRECOND_MULTIGRID, check_interval=8)
    sim.load_initial_state()
    ck = sim.checkpoint()
    runs = []
    for _ in range(3):
        sim.restore(ck)
        c0 = sim.debug_counters()
        sim.run_steps(12)
        c1 = sim.debug_counters()
        runs.append((c1["graph_captures"] - c0["graph_captures"], c1["graph_rebuilds"] - c0["graph_rebuilds"],
                     sim.state()))
    sim.close()
    (n1, _, _), (n2, _, st2), (n3, r3, st3) = runs
    assert n1 > 0 and n3 == 0, (n1, n2, n3, r3)
    for k in ("p", "pos", "vel"):
        assert (st2[k] == st3[k]).all(), k
