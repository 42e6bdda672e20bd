"""Benchmark: particle-steps/s of the MPS step on the 2M-particle 2D dam break,
mixed-precision vs FP64 pressure solve at an equal 1e-9 relative residual.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One GPU.  The timed region is K full MPS steps per path, inputs resident in
HBM; every step's true FP64 relative residual ||b-Ax||/||b|| is checked
against the tolerance.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mps_amd import capi, configs  # noqa: E402

METRIC = "particle-steps/s @2M 2D dambreak, mixed vs FP64 PPE at equal 1e-9 residual"
HBM_PEAK_GBS = 8000.0
TOL = 1e-9


def run_path(settings, scene, precision, steps, warmup, check_interval, reps, use_graph=True, precond=1, warm=1,
             inner_tol=1e-4, state_at=0):
    """settings, scene: one scene; settings = list of (settings, scene), scene =
    None: an ensemble in one handle.  state_at > 0: after the timed steps the
    handle steps on (untimed) to that many steps from t = 0 and returns that
    state as well (the CPU baseline's mid-run starting point)."""
    sim = capi.Simulation(settings, scene)
    sim.set_solver(precision=precision, tolerance=TOL, warm_start=warm, check_interval=check_interval,
                   inner_tolerance=inner_tol, use_graph=use_graph, preconditioner=precond)
    sim.load_initial_state()
    for _ in range(warmup):
        sim.step()
    its, res, refs = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()  # returns after the step's stream has drained
        if sim.nscene > 1:  # worst scene: the segmented CG runs until the last one converges
            per, _ = sim.scene_stats()
            its.append(max(x["iterations"] for x in per))
            res.append(max(x["rel_residual"] for x in per))
            refs.append(max(x["refinements"] for x in per))
            continue
        st = sim.solver_stats()
        its.append(st["iterations"])
        res.append(st["rel_residual"])
        refs.append(st["refinements"])
    t1 = time.perf_counter()
    kern = sim.time_pressure_kernels(reps)
    state = sim.state()
    rows = sim.solver_stats()["rows"]
    mid = None
    if state_at > warmup + steps:
        sim.run_steps(state_at - warmup - steps)
        mid = sim.state()
    sim.close()
    return dict(seconds=t1 - t0, iterations=its, residuals=res, refinements=refs, kernels=kern, state=state,
                rows=rows, mid_state=mid)


def host_cpu():
    """The host's CPU model, logical CPUs, the CPUs this process may run on and
    its cgroup CPU quota (the share a container gets)."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    return dict(model=model, logical_cpus=os.cpu_count(), allowed_cpus=allowed, cgroup_cpu_quota=quota)


def cpu_threads(info):
    """Threads for the CPU baseline: every CPU the process may use, capped by
    the cgroup quota and by OMP_NUM_THREADS when the environment sets it (the
    share of the host this process is given)."""
    t = info["allowed_cpus"]
    if info["cgroup_cpu_quota"]:
        t = min(t, max(1, int(info["cgroup_cpu_quota"])))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        t = min(t, int(env))
    return max(1, t)


def _oracle_steps(settings, scene, steps, tol, duplicate):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    o = Oracle(settings, scene, tol=tol, duplicate=duplicate)
    o.load_initial_state()
    t0 = time.perf_counter()
    its = []
    for k in range(steps):
        o.step()
        its.append(o.cg_stats()[0])
        print(f"[bench] cpu step {k + 1}/{steps} (N={scene.n}, tol {tol:g}{', solved twice' if duplicate else ''}): "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    o.close()
    return dt, its


def cpu_baseline(settings, scene, n_fluid, mid_state, full=False, workload="dambreak2d_2M"):
    """The FP64 CPU oracle (the reference's algorithm restated in C, OpenMP on
    the stage loops the reference parallelises and on the SpMV) on this
    host, timed on the same workload.  Default (a bounded sample): one step
    from t = 0 and one from the mid-run state the GPU reached, CG to the same
    1e-9 tolerance with a single solve, plus the reference's own 584-particle
    dam break for its full 1000 steps with the reference's solver (tolerance
    machine epsilon, solved twice per step, src/mps.cpp:457-460).  full: 3 +
    3 steps, and the same on the 2M workload with the reference's solver."""
    info = host_cpu()
    threads = cpu_threads(info)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    eps = float(np.finfo(np.float64).eps)
    k = 3 if full else 1
    t_a, it_a = _oracle_steps(settings, scene, k, TOL, False)
    parts = [f"{k} step(s) from t=0: {t_a / k:.1f} s/step, CG {it_a} iterations"]
    t_tot, n_steps = t_a, k
    if mid_state is not None:
        mid = configs.Scene(mid_state["type"].copy(), mid_state["pos"].copy(), mid_state["vel"].copy())
        t_b, it_b = _oracle_steps(settings, mid, k, TOL, False)
        parts.append(f"{k} step(s) from the GPU state at step {mid_state['step']}: {t_b / k:.1f} s/step, "
                     f"CG {it_b} iterations")
        t_tot += t_b
        n_steps += k
    out = dict(value=n_fluid * n_steps / t_tot, unit="particle-steps/s", cores=threads, kind="port",
               sample=f"{workload} (N={scene.n}), FP64 oracle, CG to {TOL:g} relative residual, single solve: "
                      + "; ".join(parts) + f"; on {threads} threads of {info['model']}",
               host=info)
    s0, sc0, steps0 = configs.workload("dambreak2d_584")
    t0, it0 = _oracle_steps(s0, sc0, steps0, eps, True)
    f0 = int((sc0.type == configs.FLUID).sum())
    out["reference_faithful_cfg0"] = dict(
        value=f0 * steps0 / t0, unit="particle-steps/s", seconds=t0, steps=steps0,
        sample=f"dambreak2d_584 (N={sc0.n}) full {steps0} steps, CG to machine epsilon on the recursive residual, "
               f"solved twice per step (src/mps.cpp:457-460), mean {np.mean(it0):.0f} iterations per solve")
    if full:
        t_r, it_r = _oracle_steps(settings, scene, k, eps, True)
        out["reference_faithful_2M"] = dict(
            value=n_fluid * k / t_r, unit="particle-steps/s", seconds_per_step=t_r / k,
            sample=f"{k} steps of {workload} from t=0, CG to machine epsilon solved twice per step, "
                   f"iterations per solve {it_r}")
    return out


def step_bytes(dim, s_bytes, c_bytes, cells_per_particle):
    """Algorithmic HBM bytes per particle of one step's fixed stages and of one
    CG iteration / one FP64 refinement residual: each array a stage needs read
    once, each output written once, neighbour reuse through LDS or cache not
    counted (s: bytes of a state scalar, c: of a CG scalar; the cell table
    costs 12 B per cell and build)."""
    D = dim
    build = D * s_bytes + 17 + 12 * cells_per_particle
    fixed = (3 * build                       # three bucket builds + neighbour walks
             + 4 * D * s_bytes + 2           # gravity, viscosity, move
             + 2 * D * s_bytes + 1           # collision detection
             + D * s_bytes + 10              # boundary condition + source term (b in FP64)
             + D * s_bytes + 2               # Dirichlet connectivity
             + D * s_bytes + 1 + c_bytes     # matrix diagonal
             + D * s_bytes + 10 + 2 * s_bytes  # clamp + minimum pressure
             + 4 * D * s_bytes + 2 + 2 * s_bytes  # pressure gradient + move
             + D * s_bytes + 1)              # Courant number
    per_iter = 9 * c_bytes + D * s_bytes + 1
    per_refine = 32 + D * s_bytes + 1
    return fixed, per_iter, per_refine


def step_roofline(settings, scene, n_total, steps, fp64, mixed):
    """Whole-step HBM roofline: sum over the timed steps of the algorithmic
    bytes B_step = N (fixed + k per_iter [+ m per_refine]) with each step's
    own iteration counts (FP64: state and CG scalars of 8 bytes; mixed: 4),
    over the wall time, against the 8 TB/s peak."""
    if scene is None:  # ensemble: every scene is 2D; cells over all scenes
        dim = 2
        cells = 0
        for s, sc, _ in settings:
            c = s.to_c()
            cells += (int((c.domain_max[0] - c.domain_min[0]) / c.re_max) + 3) * \
                     (int((c.domain_max[1] - c.domain_min[1]) / c.re_max) + 3)
    else:
        dim = settings.dim
        c = settings.to_c()
        cells = 1
        for d in range(dim):
            cells *= int((c.domain_max[d] - c.domain_min[d]) / c.re_max) + 3
    cpp = cells / n_total
    out = {"peak": HBM_PEAK_GBS, "unit": "GB/s", "cells_per_particle": cpp}
    for name, r, sb in (("fp64", fp64, 8), ("mixed", mixed, 4)):
        fixed, per_it, per_ref = step_bytes(dim, sb, sb, cpp)
        tot = 0.0
        for k, m in zip(r["iterations"], r["refinements"]):
            tot += n_total * (fixed + k * per_it + (m * per_ref if name == "mixed" else 0.0))
        ach = tot / r["seconds"] / 1e9
        out[name] = {"bytes_per_step": tot / steps, "achieved": ach, "frac": ach / HBM_PEAK_GBS,
                     "bytes_per_particle_fixed": fixed, "bytes_per_particle_iteration": per_it}
    out["note"] = ("per-unit bytes of a matrix-free step with FP32 state on the mixed path; the solver here stores "
                   "the matrix and keeps the particle state in FP64 on both paths, so its real traffic is higher")
    return out


def copy_probe(mib):
    """A plain device copy of known size (one elementwise kernel reading and
    writing mib MiB per launch), run under the same HBM counters as the
    solver kernels to turn the counters' units into bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_probe_mib": mib, "launches": 4,
                      "read_write_gbs": 4 * 2 * n * 4 / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dambreak2d_2M")
    ap.add_argument("--check-interval", type=int, default=8)
    ap.add_argument("--reps", type=int, default=30, help="launches per kernel for the event timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="enqueue CG chunks directly (for profilers)")
    ap.add_argument("--precond", choices=("multigrid", "jacobi"), default="multigrid",
                    help="CG preconditioner of both paths")
    ap.add_argument("--paths", default="fp64,mixed", help="paths to run (mixed sets value)")
    ap.add_argument("--warm-start", type=int, default=2, choices=(0, 1, 2),
                    help="CG start: 0 zero, 1 the last solution, 2 extrapolated from the last two")
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="mixed path: residual reduction asked of each FP32 inner solve")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline: 3 + 3 steps and the reference's own solver on the 2M workload too "
                         "(several minutes of CPU time)")
    ap.add_argument("--mid-state-step", type=int, default=250,
                    help="step (from t = 0) of the GPU state the CPU baseline's second sample starts from")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB between two device buffers 4 times and exit: the known-size "
                         "copy that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    a = ap.parse_args()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    precond = capi.MPS_PRECOND_MULTIGRID if a.precond == "multigrid" else capi.MPS_PRECOND_JACOBI
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {a.gpus} is not supported (the MPS step here is a single-GPU "
              f"path; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "ensemble64":
        # 64 independent 2D scenes (half dam breaks, half hydrostatic tanks,
        # l0 in {4, 2, 1} mm, nu log-uniform in [1e-6, 1e-3], dt = 0.2 l0 / 3.5)
        # drawn from SplitMix64 seed 2024, stepped together in one handle
        members = configs.ensemble_members(seed=2024, count=64)
        settings, scene = members, None
        n_total = sum(m[1].n for m in members)
        n_fluid = sum(int((m[1].type == configs.FLUID).sum()) for m in members)
        a.no_cpu = True
    else:
        settings, scene, _ = configs.workload(a.workload)
        n_total = scene.n
        n_fluid = int((scene.type == configs.FLUID).sum())
    print(f"[bench] workload {a.workload}: N={n_total} fluid={n_fluid} steps={a.steps} warmup={a.warmup}",
          file=sys.stderr, flush=True)

    out = {}
    for name, prec in (("fp64", capi.MPS_PREC_FP64), ("mixed", capi.MPS_PREC_MIXED)):
        if name not in a.paths.split(","):
            continue
        want_mid = name == "fp64" and not a.no_cpu
        r = run_path(settings, scene, prec, a.steps, a.warmup, a.check_interval, a.reps, not a.no_graph, precond,
                     a.warm_start, a.inner_tol, state_at=a.mid_state_step if want_mid else 0)
        worst = max(r["residuals"]) if r["residuals"] else 0.0
        if not worst <= TOL:
            print(f"[bench] {name}: residual {worst:.3e} above {TOL:g}", file=sys.stderr)
            sys.exit(1)
        out[name] = r
        print(f"[bench] {name}: {1e3 * r['seconds'] / a.steps:.2f} ms/step, mean CG iterations "
              f"{np.mean(r['iterations']):.1f}, refinements {np.mean(r['refinements']):.2f}, worst residual {worst:.2e}",
              file=sys.stderr, flush=True)

    if "mixed" not in out or "fp64" not in out:
        print(json.dumps({k: dict(ms_per_step=1e3 * v["seconds"] / a.steps,
                                  mean_iterations=float(np.mean(v["iterations"]))) for k, v in out.items()}))
        return
    mixed, fp64 = out["mixed"], out["fp64"]
    ms_mixed = 1e3 * mixed["seconds"] / a.steps
    ms_fp64 = 1e3 * fp64["seconds"] / a.steps
    value = n_fluid * a.steps / mixed["seconds"]
    value_fp64 = n_fluid * a.steps / fp64["seconds"]
    # error of the mixed path against the FP64 path after the same steps
    err = {}
    for k in ("p", "vel", "pos"):
        d = np.abs(mixed["state"][k] - fp64["state"][k]).max()
        s = np.abs(fp64["state"][k]).max()
        err[k] = float(d / s) if s > 0 else float(d)
    # dominant kernel: the FP32 SpMV of the mixed path's inner CG, the
    # variant the solver runs (with the multigrid: the apply gathering M r)
    ms, by = mixed["kernels"]["spmv32"]
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    # HBM bytes per launch from the committed, copy-calibrated FETCH_SIZE /
    # WRITE_SIZE runs of the same kernel on the same workload, scaled to this
    # run's row count (the ensemble's 2D scenes use the 2M dam break's)
    traffic = None
    pmc_of = "dambreak2d_2M" if a.workload == "ensemble64" else a.workload
    pmc = os.path.join(ROOT, "profiles", f"pmc_cg_apply_{pmc_of}.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            t = json.load(f)
        traffic = (t["fetch_bytes"] + t["write_bytes"]) * mixed["rows"] / t["rows"]
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": a.workload,
            "n_particles": n_total,
            "n_fluid": n_fluid,
            "dt": settings.dt if scene is not None else "per scene: 0.2 l0 / 3.5",
            "tolerance": TOL,
            "warm_start": {0: "zero", 1: "last solution", 2: "extrapolated from the last two solutions"}[a.warm_start],
            "preconditioner": a.precond,
            "inner_tolerance": a.inner_tol,
            "per_particle_value_basis": "n_fluid",
        },
        "fp64": {
            "value": value_fp64,
            "ms_per_step": ms_fp64,
            "mean_cg_iterations": float(np.mean(fp64["iterations"])),
            "worst_rel_residual": max(fp64["residuals"]),
        },
        "mixed": {
            "mean_inner_iterations": float(np.mean(mixed["iterations"])),
            "mean_refinements": float(np.mean(mixed["refinements"])),
            "worst_rel_residual": max(mixed["residuals"]),
            "rel_error_vs_fp64": err,
        },
        "speedup_mixed_vs_fp64": value / value_fp64,
        "value_n_total": n_total * a.steps / mixed["seconds"],
        "step_roofline": step_roofline(settings, scene, n_total, a.steps, fp64, mixed),
        "roofline": {
            "bound": "hbm",
            # 16-bit column codes: offsets (2D), layer codes (3D)
            "kernel": "k_cg_apply<float, %s, false, %s>" % (
                "unsigned short" if scene is not None and settings.dim == 3 else "short",
                "true" if a.precond == "multigrid" else "false"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": by,
            "avg_launch_ms": ms,
            "rows": mixed["rows"],
        },
        "kernels_ms": {k: v[0] for k, v in mixed["kernels"].items()},
    }
    if scene is None:
        res["config"]["scenes"] = len(settings)
        res["metric"] = "fluid particle-steps/s, 64-scene 2D ensemble in one handle, mixed vs FP64 PPE at 1e-9"
        res["traffic_note"] = "roofline.traffic is scaled from the 2M dam break's counters"
    if not a.no_cpu:
        mid = fp64.get("mid_state")
        if mid is not None:
            mid = dict(mid, step=a.mid_state_step)
        print("[bench] cpu baseline (FP64 oracle on the host)", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(settings, scene, n_fluid, mid, full=a.cpu_full, workload=a.workload)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
