"""torch.autograd binding of the P-SV misfit (a test binding over the C ABI).

``PSVMisfit.apply(lam, mu, rho, problem)`` returns the FP64 misfit
sum_shots sum_rec sum_t 0.5 dt (d - d_obs)^2 of a batch of shots for the
Lame parameters and density on the padded grid (the averaged mu_zx, rho_zp,
rho_xp are formed as average_parameters does, CUDA_week2/fd_medium.cu:
128-159).  Its backward depends on ``problem.gradient``:

- "discrete": the exact derivative of that misfit (the plan's discrete
  adjoint, an extension), dJ/dlam, dJ/dmu, dJ/drho on the whole grid, times
  the incoming gradient.  Central finite differences of the misfit agree with
  it to rounding.
- "reference": the NEGATED lambda / mu / rho gradients that
  adjoint_kernel_PSV accumulates (CUDA_week2/adjoint_kernel_PSV_GPU.cu:
  830-863), scattered from the FWI window onto the grid (zero outside it).
  That gradient correlates the adjoint and forward wavefields; it is not the
  derivative of the discrete misfit, and it points against it (the reference
  updates the model with m + s g, CUDA_week2/fd_medium.cu:161-240).  Negated,
  it is a descent direction: on a 64 x 64 random-smooth model its
  correlation with central differences of the misfit is about 0.56 (lambda),
  0.60 (mu) and 0.98 (rho), with an arbitrary scale.

The host code (C++) stays the product; this module only moves tensors to
and from the C ABI.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import psvfwi as P


@dataclass
class Problem:
    """Geometry and settings of a batch: grid and CPML, shots (as for
    psvfwi.Plan.set_shots), observed traces (one [nrec, nt] array per shot)
    and the FWI window."""
    nzt: int
    nxt: int
    fdorder: int
    npml: int
    fsurf: int
    dt: float
    dx: float
    dz: float
    hc: np.ndarray
    pml: np.ndarray
    shots: list
    observed: list
    window: dict
    precision: str = "fp64"
    tile: int = 0
    gradient: str = "reference"   # or "discrete"
    segment: int = 64             # discrete: steps between checkpoints
    plan: object = field(default=None, repr=False)

    def get_plan(self):
        if self.plan is None:
            p = P.Plan(nzt=self.nzt, nxt=self.nxt, fdorder=self.fdorder, npml=self.npml, fsurf=self.fsurf,
                       dt=self.dt, dx=self.dx, dz=self.dz, hc=self.hc, precision=self.precision)
            p.set_pml(self.pml)
            p.set_window(self.window)
            p.set_shots(self.shots)
            p.set_gradient_mode(self.gradient, self.segment)
            if self.tile:
                p.set_tile(self.tile)
            self.plan = p
        return self.plan


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class PSVMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lam, mu, rho, problem: Problem):
        plan = problem.get_plan()
        l, m, r = _np(lam), _np(mu), _np(rho)
        plan.set_model(l, m, *P.average_parameters(m, r))
        plan.forward(energy=False, gradient=True)
        mis = plan.misfit(problem.observed)
        ctx.problem = problem
        ctx.like = lam
        return torch.tensor(float(np.sum(mis)), dtype=torch.float64, device=lam.device)

    @staticmethod
    def backward(ctx, grad_out):
        problem = ctx.problem
        plan = problem.get_plan()
        plan.adjoint(energy=False, gradient=True)
        if problem.gradient == "discrete":
            return tuple(torch.from_numpy(g).to(ctx.like.device, ctx.like.dtype) * grad_out
                         for g in plan.model_gradient(-1)) + (None,)
        gl, gm, gr = plan.gradient(-1)
        w = problem.window
        full = []
        for g in (gl, gm, gr):
            a = np.zeros((problem.nzt, problem.nxt))
            nzw = (w["z2"] - w["z1"] + w["dz"] - 1) // w["dz"]
            nxw = (w["x2"] - w["x1"] + w["dx"] - 1) // w["dx"]
            a[w["z1"]:w["z1"] + nzw * w["dz"]:w["dz"], w["x1"]:w["x1"] + nxw * w["dx"]:w["dx"]] = g[:nzw, :nxw]
            full.append(torch.from_numpy(-a).to(ctx.like.device, ctx.like.dtype) * grad_out)
        return full[0], full[1], full[2], None
