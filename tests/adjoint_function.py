"""torch.autograd.Function over the reverse-mode step (test harness only).

forward(v0, params) runs the recorded steps from a fixed scene with the given
fluid initial velocities (N_f x 2) and (nu, restitution, relaxation), and
returns the fluid positions at the end (N_f x 2).  backward() feeds
dL/d(positions) to mps_adjoint_backward and returns dL/dv0 and dL/dparams.
Tensors stay on the CPU: the C ABI copies host buffers in and out.
"""
from __future__ import annotations

import numpy as np
import torch

from mps_amd import capi, configs, gradcheck


class MpsRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, params, settings, scene, steps, interval, tol):
        fl = scene.type == configs.FLUID
        vel = np.zeros((scene.n, 3))
        vel[fl, :2] = v0.detach().cpu().numpy()
        p = params.detach().cpu().numpy()
        s = gradcheck.with_params(settings, float(p[0]), float(p[1]), float(p[2]))
        sim = capi.Simulation(s, configs.Scene(scene.type.copy(), scene.pos.copy(), vel))
        sim.set_solver(precision=capi.MPS_PREC_FP64, tolerance=tol)
        sim.load_initial_state()
        sim.adjoint_record(steps, interval)
        pos = sim.state()["pos"]
        ctx.sim, ctx.fl, ctx.n = sim, fl, scene.n
        return torch.from_numpy(pos[fl, :2].copy())

    @staticmethod
    def backward(ctx, grad_pos):
        seed = np.zeros((ctx.n, 3))
        seed[ctx.fl, :2] = grad_pos.detach().cpu().numpy()
        _, vel_bar0, par, _ = ctx.sim.adjoint_backward(seed)
        ctx.sim.close()
        gv = torch.from_numpy(vel_bar0[ctx.fl, :2].copy())
        gp = torch.from_numpy(par.copy())
        return gv, gp, None, None, None, None, None


def rollout(v0, params, settings, scene, steps, interval=50, tol=1e-14):
    return MpsRollout.apply(v0, params, settings, scene, steps, interval, tol)
