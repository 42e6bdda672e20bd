"""Reverse-mode step (mps_adjoint_record / mps_adjoint_backward).

The adjoint holds every discrete decision of the recorded run fixed, so it is
the derivative of the piecewise-smooth map on the piece the run lies in.
Central differences see the same derivative when no decision flips between
theta - h d and theta + h d; h is chosen small enough for that (larger h is
where the pieces show).  Bar: FP64 central differences on 8 seeded random
directions in (v0, nu, restitution, relaxation) within 1e-4 relative.
"""
import dataclasses

import numpy as np
import pytest
import torch

from mps_amd import capi, configs, gradcheck

from adjoint_function import rollout
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)
REL = 1e-4


def _assert_close(r, h):
    rel = r["fd"][f"{h:.0e}"]["rel"]
    assert max(rel) <= REL, (rel, r["adjoint"], r["fd"][f"{h:.0e}"]["fd"])


def test_adjoint_matches_central_differences_2d():
    r = gradcheck.check("dambreak2d_584", steps=20, interval=7, ndir=8, hs=(1e-5,), tol=EPS)
    _assert_close(r, 1e-5)
    # the adjoint solves run with the recorded run's tolerance, machine
    # epsilon, which the true FP64 residual of A x cannot reach (converged
    # reports the true residual against the tolerance): they stop at the
    # rounding floor, far below what the FD check resolves
    assert r["adjoint_stats"]["rel_residual"] < 1e-13


def test_adjoint_through_collisions():
    # v0 ~ N(0, 1 m/s) makes pairs close below 0.5 l0, so the collision pass
    # fires and dL/d(restitution) is nonzero
    r = gradcheck.check("dambreak2d_584", steps=20, interval=6, ndir=8, hs=(1e-6,), tol=EPS, vscale=1.0)
    assert abs(r["param_bar"][1]) > 0.0
    _assert_close(r, 1e-6)


def test_adjoint_longer_horizon():
    r = gradcheck.check("dambreak2d_584", steps=200, interval=50, ndir=8, hs=(1e-7,), tol=EPS)
    _assert_close(r, 1e-7)


def test_adjoint_35k_workload():
    """The inverse dam break at its own size (31,250 fluid particles, nu 1e-3,
    dt 1e-4, loss against a target run with nu 2e-3 and restitution 0.25):
    50 steps with checkpoints every 10, CG to 1e-14, 8 seeded directions in
    (v0, nu, e, relaxation) against FP64 central differences with h = 1e-8
    (at h = 1e-6 a discrete decision of the step flips inside the stencil of
    five of the eight directions, and their differences jump by the same
    1.3e-13 / 2h; at the full 2000 steps no step size resolves the pieces the
    loss is made of; see the design notes on the reverse mode)."""
    r = gradcheck.check("dambreak2d_35k", steps=50, interval=10, ndir=8, hs=(1e-8,), tol=1e-14)
    assert r["n_fluid"] == 31250
    print(f"[adjoint 35k] max rel {max(r['fd']['1e-08']['rel']):.2e}", flush=True)
    _assert_close(r, 1e-8)


def test_adjoint_3d():
    s = configs.dambreak_settings(0.02, 1e-3, 1.0, dim=3)
    sc = configs.lattice_3d(0.02, (0.3, 0.3, 0.1), (0.0, 0.1), (0.0, 0.2), (0.0, 0.1), 1, 2)
    s.domain_x, s.domain_y, s.domain_z = (-0.1, 0.4), (-0.1, 0.4), (-0.1, 0.2)
    # the velocity draw covers x and y; z starts at rest and is moved by the flow
    r = gradcheck.check(case=(s, sc), steps=15, interval=4, ndir=8, hs=(1e-6,), tol=EPS, vscale=0.3)
    _assert_close(r, 1e-6)


def test_adjoint_against_oracle_differences():
    """Small-N reference gradient: central differences of the CPU oracle (an
    independent FP64 implementation of the forward step) against the GPU
    adjoint."""
    s, sc, _ = configs.workload("dambreak2d_584")
    steps = 20
    fl = sc.type == configs.FLUID
    v0 = gradcheck.initial_velocity(sc, 3, 0.05)
    target = gradcheck.run_positions(gradcheck.with_params(s, 2e-3, 0.25, s.relaxation), sc, v0, steps, EPS)
    grad = gradcheck.adjoint_gradient(s, sc, v0, target, steps, 8, EPS)
    dirs = gradcheck.directions(s, sc, 8, 99)
    ad = gradcheck.directional(grad, dirs, fl)
    h = 1e-5

    def oracle_loss(settings, vel):
        o = Oracle(settings, configs.Scene(sc.type.copy(), sc.pos.copy(), vel))
        o.load_initial_state()
        for _ in range(steps):
            o.step()
        L = gradcheck.loss_and_seed(o.state()["pos"], target, fl)[0]
        o.close()
        return L

    for (dv, dp), a in zip(dirs, ad):
        Ls = []
        for sign in (1.0, -1.0):
            ss = gradcheck.with_params(s, s.kinematic_viscosity + sign * h * dp[0], s.restitution + sign * h * dp[1],
                                       s.relaxation + sign * h * dp[2])
            Ls.append(oracle_loss(ss, v0 + sign * h * dv))
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(fd - a) <= REL * abs(fd), (fd, a)


def test_checkpoint_interval_changes_nothing_and_state_is_restored():
    s, sc, _ = configs.workload("dambreak2d_584")
    v0 = gradcheck.initial_velocity(sc, 5, 1.0)
    seed = np.random.default_rng(1).standard_normal((sc.n, 3))
    out = []
    for interval in (3, 25):
        sim = gradcheck._sim(s, gradcheck._scene(sc, v0), EPS)
        sim.adjoint_record(25, interval)
        after = sim.state()
        ev = sim.debug_events()
        st = sim.solver_stats()
        res = sim.adjoint_backward(seed, 0.5 * seed)
        back = sim.state()
        for k in after:
            assert (after[k] == back[k]).all(), k
        ev2 = sim.debug_events()
        assert (ev[0] == ev2[0]).all() and (ev[1] == ev2[1]).all() and (ev[2] == ev2[2]).all()
        assert sim.solver_stats() == st
        # the handle steps on from the recorded end exactly as an unrecorded run
        sim.step()
        ref = gradcheck._sim(s, gradcheck._scene(sc, v0), EPS)
        ref.run_steps(26)
        assert (ref.state()["pos"] == sim.state()["pos"]).all()
        out.append(res)
    for a, b in zip(out[0][:3], out[1][:3]):
        assert (a == b).all()


def test_adjoint_rejects_what_it_does_not_cover():
    s, sc, _ = configs.workload("dambreak2d_584")
    sim = gradcheck._sim(s, sc, 1e-10)
    with pytest.raises(capi.MpsError):
        sim.adjoint_backward(np.zeros((sc.n, 3)))  # nothing recorded
    s_hos = dataclasses.replace(s, hos_on=True)
    sim = gradcheck._sim(s_hos, sc, 1e-10)
    with pytest.raises(capi.MpsError):
        sim.adjoint_record(2, 1)
    ens = capi.Simulation([(s, sc), (s, sc)])
    ens.load_initial_state()
    with pytest.raises(capi.MpsError):
        ens.adjoint_record(2, 1)


def test_torch_function_gives_the_adjoint_gradient():
    s, sc, _ = configs.workload("dambreak2d_584")
    steps = 20
    fl = sc.type == configs.FLUID
    v0 = gradcheck.initial_velocity(sc, 4, 0.05)
    target = gradcheck.run_positions(gradcheck.with_params(s, 2e-3, 0.25, s.relaxation), sc, v0, steps, EPS)
    v = torch.tensor(v0[fl, :2], dtype=torch.float64, requires_grad=True)
    prm = torch.tensor([s.kinematic_viscosity, s.restitution, s.relaxation], dtype=torch.float64,
                       requires_grad=True)
    x = rollout(v, prm, s, sc, steps, 8, EPS)
    tgt = torch.from_numpy(target[fl, :2].copy())
    loss = ((x - tgt) ** 2).sum() / fl.sum()
    loss.backward()
    grad = gradcheck.adjoint_gradient(s, sc, v0, target, steps, 8, EPS)
    np.testing.assert_allclose(v.grad.numpy(), grad["vel_bar0"][fl, :2], rtol=1e-12, atol=0)
    np.testing.assert_allclose(prm.grad.numpy(), grad["param_bar"], rtol=1e-12, atol=0)
    # and a central difference through the same function
    d = torch.randn(v.shape, generator=torch.Generator().manual_seed(0), dtype=torch.float64) * 0.05
    h = 1e-5
    with torch.no_grad():
        lp = (((rollout(v + h * d, prm, s, sc, steps, 8, EPS) - tgt) ** 2).sum() / fl.sum()).item()
        lm = (((rollout(v - h * d, prm, s, sc, steps, 8, EPS) - tgt) ** 2).sum() / fl.sum()).item()
    fd = (lp - lm) / (2 * h)
    assert abs(fd - float((v.grad * d).sum())) <= REL * abs(fd)
