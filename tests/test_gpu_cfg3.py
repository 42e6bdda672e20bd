"""The 64-scene ensemble at full size: SplitMix64 seed 2024, half dam breaks
and half hydrostatic tanks, l0 in {4, 2, 1} mm, nu log-uniform in
[1e-6, 1e-3], dt = 0.2 l0 / 3.5 (3,203,508 particles in one handle; the
reference runs one scene per process, src/simulation.cpp:29-64).

* With the multigrid preconditioner, on both precisions, every scene's true
  FP64 relative residual meets 1e-9 on every step (per-scene convergence).
* With the FP64 Jacobi solve, three members (the largest, the smallest and
  one between) are bit-identical to the same scenes run alone.
* The smallest member after one step agrees with the FP64 CPU oracle run on
  that scene alone: bit-exact up to the pressure stage, the state within the
  solver tolerance.
"""
import numpy as np
import pytest

from mps_amd import capi, configs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

MEMBERS = configs.ensemble_members(seed=2024, count=64)
SIZES = [m[1].n for m in MEMBERS]


def test_cfg3_workload_shape():
    assert len(MEMBERS) == 64 and sum(SIZES) == 3203508
    fluid = [int((m[1].type == configs.FLUID).sum()) for m in MEMBERS]
    assert 2000 <= min(fluid) and max(fluid) <= 215000


@pytest.mark.parametrize("precision", [capi.MPS_PREC_FP64, capi.MPS_PREC_MIXED], ids=["fp64", "mixed"])
def test_cfg3_multigrid_every_scene_meets_tolerance(precision):
    print(f"[cfg3] 64 scenes, multigrid, precision {precision}, 4 steps", flush=True)
    ens = capi.Simulation(MEMBERS)
    ens.set_solver(precision=precision, tolerance=1e-9, warm_start=2, preconditioner=capi.MPS_PRECOND_MULTIGRID)
    ens.load_initial_state()
    for k in range(4):
        ens.step()
        stats, cn = ens.scene_stats()
        worst = max(x["rel_residual"] for x in stats)
        print(f"[cfg3] step {k}: worst residual {worst:.2e}, max iterations "
              f"{max(x['iterations'] for x in stats)}", flush=True)
        assert all(x["converged"] == 1 and x["rel_residual"] <= 1e-9 for x in stats)
        assert all(x["rows"] > 0 for x in stats)
        assert (cn > 0).all()
    ens.close()


def test_cfg3_members_bit_identical_to_solo_runs():
    steps = 3
    print(f"[cfg3] 64 scenes, FP64 Jacobi, {steps} steps, then three solo runs", flush=True)
    ens = capi.Simulation(MEMBERS)
    ens.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-9, warm_start=True)
    ens.load_initial_state()
    for _ in range(steps):
        ens.step()
    stats, cn = ens.scene_stats()
    est = ens.state()
    ens.close()
    order = np.argsort(SIZES)
    picks = [int(order[-1]), int(order[0]), int(order[len(order) // 2])]
    for i in picks:
        s, sc, kind = MEMBERS[i]
        solo = capi.Simulation(s, sc)
        solo.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-9, warm_start=True)
        solo.load_initial_state()
        c = None
        for _ in range(steps):
            c = solo.step()
        ss = solo.state()
        a, b = int(np.sum(SIZES[:i])), int(np.sum(SIZES[:i + 1]))
        mine = {k: v[a:b] for k, v in est.items()}
        for key in ("pos", "vel", "p", "ndr", "bc", "type", "pmin"):
            assert (mine[key] == ss[key]).all(), (i, kind, key)
        assert cn[i] == c
        assert stats[i]["iterations"] == solo.solver_stats()["iterations"]
        solo.close()


def test_cfg3_smallest_member_matches_oracle():
    i = int(np.argmin(SIZES))
    s, sc, kind = MEMBERS[i]
    print(f"[cfg3] member {i} ({kind}, N={sc.n}) inside the ensemble vs the oracle, one step", flush=True)
    ens = capi.Simulation(MEMBERS)
    ens.set_solver(precision=capi.MPS_PREC_FP64, tolerance=1e-13, warm_start=False)
    ens.set_snapshot(True)
    ens.load_initial_state()
    ens.step()
    est = ens.state()
    a, b = int(np.sum(SIZES[:i])), int(np.sum(SIZES[:i + 1]))
    gpos, gvel = ens.snapshot()
    ens.close()
    o = Oracle(s, sc, tol=1e-13)
    o.load_initial_state()
    o.step()
    opos, ovel = o.snapshot()
    assert (gpos[a:b] == opos).all() and (gvel[a:b] == ovel).all()
    os_ = o.state()
    mine = {k: v[a:b] for k, v in est.items()}
    np.testing.assert_array_equal(mine["bc"], os_["bc"])
    np.testing.assert_array_equal(mine["type"], os_["type"])
    perr = np.abs(mine["p"] - os_["p"]).max() / np.abs(os_["p"]).max()
    verr = np.abs(mine["vel"] - os_["vel"]).max() / np.abs(os_["vel"]).max()
    print(f"[cfg3] member {i}: p err {perr:.2e}, v err {verr:.2e}", flush=True)
    assert perr <= 1e-9 and verr <= 1e-9
    assert np.abs(mine["pos"] - os_["pos"]).max() <= 1e-15
