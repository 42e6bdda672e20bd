"""Central finite-difference check of the reverse-mode step (mps_adjoint_*).

The inverse problem: a 2D dam break whose fluid particles start with seeded
random velocities v0 ~ N(0, 0.05 m/s) (drawn in FP32, as an optimiser would
hold them), run for `steps` steps.  The loss is the mean squared distance of
the fluid positions at the end to those of a target run from the same start
with kinematic viscosity 2e-3 and coefficient of restitution 0.25:

    L = (1 / N_f) sum_{i fluid} |x_i(T) - x_i^target(T)|^2      (x and y)

Gradients are taken with respect to v0 (fluid, x and y), nu, the coefficient
of restitution and the relaxation coefficient for pressure.  They are checked
on seeded random directions d = (dv0, dnu, de, drelax) against FP64 central
differences (L(theta + h d) - L(theta - h d)) / (2 h).  The 2 x directions
perturbed runs of one step size h go through the GPU as one ensemble handle
(each member is bit-identical to the same run alone).

    python -m mps_amd.gradcheck --workload dambreak2d_35k --steps 2000 --interval 50 --dirs 8 --h 1e-7
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import time

import numpy as np

from . import capi, configs

TARGET_NU = 2.0e-3
TARGET_RESTITUTION = 0.25


def initial_velocity(scene: configs.Scene, seed: int, scale: float = 0.05) -> np.ndarray:
    """N x 3 velocities: fluid x, y ~ N(0, scale) rounded to FP32, the rest 0."""
    rng = np.random.default_rng(seed)
    fl = scene.type == configs.FLUID
    v0 = np.zeros((scene.n, 3))
    draw = rng.standard_normal((int(fl.sum()), 2)).astype(np.float32) * np.float32(scale)
    v0[fl, :2] = draw.astype(np.float64)
    return v0


def with_params(s: configs.Settings, nu: float, restitution: float, relaxation: float) -> configs.Settings:
    return dataclasses.replace(s, kinematic_viscosity=nu, restitution=restitution, relaxation=relaxation)


def _scene(scene: configs.Scene, v0: np.ndarray) -> configs.Scene:
    return configs.Scene(scene.type.copy(), scene.pos.copy(), np.ascontiguousarray(v0, np.float64))


def _sim(members_or_settings, scene=None, tol=1e-14):
    sim = capi.Simulation(members_or_settings, scene)
    sim.set_solver(precision=capi.MPS_PREC_FP64, tolerance=tol)
    sim.load_initial_state()
    return sim


def run_positions(settings, scene, v0, steps, tol):
    sim = _sim(settings, _scene(scene, v0), tol)
    sim.run_steps(steps)
    pos = sim.state()["pos"]
    sim.close()
    return pos


def loss_and_seed(pos: np.ndarray, target: np.ndarray, fluid: np.ndarray):
    d = pos[fluid, :2] - target[fluid, :2]
    nf = int(fluid.sum())
    seed = np.zeros_like(pos)
    seed[fluid, :2] = 2.0 * d / nf
    return float((d * d).sum() / nf), seed


def adjoint_gradient(settings, scene, v0, target, steps, interval, tol):
    """Loss, dL/dv0 (N x 3) and dL/d(nu, e, relax) from one recorded run and
    its reverse sweep."""
    fl = scene.type == configs.FLUID
    sim = _sim(settings, _scene(scene, v0), tol)
    t0 = time.perf_counter()
    sim.adjoint_record(steps, interval)
    t1 = time.perf_counter()
    L, seed = loss_and_seed(sim.state()["pos"], target, fl)
    _, vel_bar0, par, stats = sim.adjoint_backward(seed)
    t2 = time.perf_counter()
    sim.close()
    return dict(loss=L, vel_bar0=vel_bar0, param_bar=par, stats=stats, record_s=t1 - t0, backward_s=t2 - t1)


def directions(settings, scene, count, seed, vscale=0.05):
    """Seeded random directions in (v0 of fluid x/y, nu, e, relax), each
    component scaled by its natural size."""
    rng = np.random.default_rng(seed)
    fl = scene.type == configs.FLUID
    nf = int(fl.sum())
    scales = np.array([settings.kinematic_viscosity, settings.restitution, settings.relaxation])
    out = []
    for _ in range(count):
        dv = np.zeros((scene.n, 3))
        dv[fl, :2] = rng.standard_normal((nf, 2)) * vscale
        dp = rng.standard_normal(3) * scales
        out.append((dv, dp))
    return out


def directional(grad, dirs, fluid):
    return [float((grad["vel_bar0"][fluid, :2] * dv[fluid, :2]).sum() + (grad["param_bar"] * dp).sum())
            for dv, dp in dirs]


def finite_differences(settings, scene, v0, target, steps, dirs, h, tol):
    """Central differences of the loss along each direction; the 2 x len(dirs)
    perturbed runs share one ensemble handle."""
    fl = scene.type == configs.FLUID
    members = []
    for dv, dp in dirs:
        for sign in (1.0, -1.0):
            s = with_params(settings, settings.kinematic_viscosity + sign * h * dp[0],
                            settings.restitution + sign * h * dp[1], settings.relaxation + sign * h * dp[2])
            members.append((s, _scene(scene, v0 + sign * h * dv)))
    sim = _sim(members, None, tol)
    sim.run_steps(steps)
    st = sim.state()
    losses = [loss_and_seed(sim.scene_state(k, st)["pos"], target, fl)[0] for k in range(len(members))]
    sim.close()
    return [(losses[2 * k] - losses[2 * k + 1]) / (2.0 * h) for k in range(len(dirs))]


def check(workload="dambreak2d_35k", steps=None, interval=50, ndir=8, hs=(1e-7,), tol=1e-14, seed=2024,
          dir_seed=7, log=None, vscale=0.05, case=None):
    """Adjoint vs central differences on `ndir` directions for each step size
    in `hs`.  `case` = (settings, scene) replaces the named workload."""
    if case is None:
        settings, scene, default_steps = configs.workload(workload)
    else:
        (settings, scene), default_steps, workload = case, 20, "custom"
    steps = default_steps if steps is None else steps
    fl = scene.type == configs.FLUID
    v0 = initial_velocity(scene, seed, vscale)
    t0 = time.perf_counter()
    target = run_positions(with_params(settings, TARGET_NU, TARGET_RESTITUTION, settings.relaxation), scene, v0,
                           steps, tol)
    t_target = time.perf_counter() - t0
    if log:
        log(f"target run {t_target:.1f} s")
    grad = adjoint_gradient(settings, scene, v0, target, steps, interval, tol)
    if log:
        log(f"adjoint: record {grad['record_s']:.1f} s, backward {grad['backward_s']:.1f} s, loss {grad['loss']:.6e}")
    dirs = directions(settings, scene, ndir, dir_seed, vscale)
    ad = directional(grad, dirs, fl)
    fds = {}
    for h in hs:
        t0 = time.perf_counter()
        fd = finite_differences(settings, scene, v0, target, steps, dirs, h, tol)
        rel = [abs(f - a) / max(abs(f), 1e-300) for f, a in zip(fd, ad)]
        fds[h] = dict(fd=fd, rel=rel, seconds=time.perf_counter() - t0)
        if log:
            log(f"h={h:.0e}: max rel {max(rel):.3e} ({fds[h]['seconds']:.1f} s)")
    return dict(workload=workload, n=scene.n, n_fluid=int(fl.sum()), steps=steps, interval=interval, tol=tol,
                loss=grad["loss"], param_bar=grad["param_bar"].tolist(), adjoint=ad, adjoint_stats=grad["stats"],
                record_s=grad["record_s"], backward_s=grad["backward_s"],
                fd={f"{h:.0e}": v for h, v in fds.items()})


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="dambreak2d_35k")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--interval", type=int, default=50)
    ap.add_argument("--dirs", type=int, default=8)
    ap.add_argument("--h", type=float, nargs="+", default=[1e-7])
    ap.add_argument("--tolerance", type=float, default=1e-14)
    a = ap.parse_args()
    r = check(a.workload, a.steps, a.interval, a.dirs, tuple(a.h), a.tolerance,
              log=lambda m: print(m, flush=True))
    print(json.dumps(r))


if __name__ == "__main__":
    main()
