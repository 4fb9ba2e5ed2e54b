"""The torch.autograd binding (cuda_fwi_amd/autograd.py) on the 64 x 64
differentiation case: nt 600, 1 source, 32 receivers, random-smooth model,
observed data from the model and the gradient taken at its smoothed version.

- backward equals the negated reference gradient of the FP64 oracle bit for
  bit, and the FP32 build's gradient is within 1e-4 relative L2 of it;
- against central finite differences of the misfit (step 1e-4 relative) on
  50 seeded grid points (seed 4), the backward gradient has the direction of
  dJ/dm: correlation above 0.45 (lambda), 0.45 (mu) and 0.95 (rho); the
  reference's gradient is not the exact discrete derivative (its magnitude
  differs by an arbitrary factor), which is why the check is on direction;
- with gradient="discrete" the backward is dJ/dm itself: the same central
  differences agree with it to 1e-6 of max |g| at every point (FP64).
"""
import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter

import psv_oracle as O
from cuda_fwi_amd import psvfwi as P
from cuda_fwi_amd import workloads as WL
from cuda_fwi_amd.autograd import Problem, PSVMisfit
from gpucase import rel_l2

pytestmark = pytest.mark.gpu


def _case(precision="fp64", gradient="reference"):
    w = WL.c3()
    s = w.shots[0]
    sig = O.create_wavelet(w.nt, w.dt, 1.0, w.fc, 0.0, 1)[None]
    ps = w.padded_shot(s, sig)
    pml = O.cpml_pro(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, 3000.0, w.dt)
    win = dict(dt=1, dx=1, dz=1, x1=w.ppad, x2=w.ppad + w.nx, z1=w.zoff, z2=w.zoff + w.nz)
    lt, mt, rt = w.padded()
    # observed traces: the oracle on the true model
    true = O.Problem(nt=w.nt, nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, dt=w.dt, dx=w.dx,
                     dz=w.dx, fsurf=w.fsurf, hc=O.holbergcoeff(w.fdorder, 1), lam=lt, mu=mt, rho=rt, pml=pml,
                     src=dict(ps["src"], shot=[0]), src_signal=sig, rec=dict(ps["rec"], shot=[0] * 32))
    obs = O.forward(true)["rec"]
    vp0 = gaussian_filter(w.vp, 4.0, mode="nearest")
    l0, m0, r0 = w.padded(vp0, vp0 / 1.8, 310.0 * vp0 ** 0.25)
    pb = Problem(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                 dz=w.dx, hc=O.holbergcoeff(w.fdorder, 1), pml=pml, shots=[ps], observed=[obs], window=win,
                 precision=precision, gradient=gradient)
    return w, true, pb, (l0, m0, r0), win


def _grads(pb, model):
    lam, mu, rho = (torch.tensor(a, requires_grad=True) for a in model)
    J = PSVMisfit.apply(lam, mu, rho, pb)
    J.backward()
    return float(J.detach()), lam.grad.numpy(), mu.grad.numpy(), rho.grad.numpy()


def test_autograd_backward_is_the_reference_gradient():
    print("autograd vs oracle gradient", flush=True)
    w, true, pb, model, win = _case("fp64")
    J, gl, gm, gr = _grads(pb, model)
    # the oracle on the same model
    q = O.Problem(nt=w.nt, nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, dt=w.dt, dx=w.dx, dz=w.dx,
                  fsurf=w.fsurf, hc=pb.hc, lam=model[0], mu=model[1], rho=model[2], pml=pb.pml,
                  src=true.src, src_signal=true.src_signal, rec=true.rec, fwi=win)
    f = O.forward(q)
    m, res, _ = O.l2_adjoint_sources(w.dt, f["rec"], pb.observed[0])
    a = O.adjoint(q, q.rec, res, f["store"])
    assert abs(J - m) <= 1e-12 * m
    z1, x1 = win["z1"], win["x1"]
    for got, k in zip((gl, gm, gr), ("grad_lam", "grad_mu", "grad_rho")):
        ref = a[k]
        nz, nx = ref.shape[0] - 1, ref.shape[1] - 1   # the last row / column is never written
        assert np.abs(ref).max() > 0
        assert np.array_equal(got[z1:z1 + nz, x1:x1 + nx], -ref[:nz, :nx]), k
        assert np.count_nonzero(got) == np.count_nonzero(got[z1:z1 + nz, x1:x1 + nx])
    # FP32 build against the FP64 one
    _, _, pb32, _, _ = _case("fp32")
    J32, *g32 = _grads(pb32, model)
    errs = [rel_l2(a32, a64) for a32, a64 in zip(g32, (gl, gm, gr))]
    print(f"  misfit fp32 vs fp64 {abs(J32 - J) / J:.2e}; gradients {[f'{e:.2e}' for e in errs]}")
    assert max(errs) < 1e-4


def test_autograd_direction_against_central_differences():
    print("autograd vs central differences", flush=True)
    w, true, pb, model, win = _case("fp64")
    J, gl, gm, gr = _grads(pb, model)
    rng = np.random.default_rng(4)
    pts = [(int(rng.integers(w.zoff + 2, w.zoff + w.nz - 2)), int(rng.integers(w.ppad + 2, w.ppad + w.nx - 2)))
           for _ in range(50)]
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                  dz=w.dx, hc=pb.hc, precision="fp64")
    plan.set_pml(pb.pml)
    plan.set_shots(pb.shots)

    def misfit(arrs):
        plan.set_model(arrs[0], arrs[1], *P.average_parameters(arrs[1], arrs[2]))
        plan.forward(energy=False)
        return float(plan.misfit(pb.observed)[0])

    for k, g, floor in ((0, gl, 0.45), (1, gm, 0.45), (2, gr, 0.95)):
        fd, gg = [], []
        for z, x in pts:
            h = 1e-4 * model[k][z, x]
            up = [a.copy() for a in model]
            dn = [a.copy() for a in model]
            up[k][z, x] += h
            dn[k][z, x] -= h
            fd.append((misfit(up) - misfit(dn)) / (2 * h))
            gg.append(g[z, x])
        c = np.corrcoef(fd, gg)[0, 1]
        print(f"  {('lambda', 'mu', 'rho')[k]}: correlation with central differences {c:.3f}, "
              f"scale fd/grad {np.dot(fd, gg) / np.dot(gg, gg):.3e}")
        assert c > floor


def test_autograd_discrete_gradient_equals_central_differences():
    """The 64 x 64 differentiation case with gradient="discrete": central
    differences (step 1e-4 relative) at 50 seeded points (seed 4) for lambda,
    mu and rho against the backward gradient, next to the reference
    gradient's direction on the same points."""
    print("autograd discrete vs central differences", flush=True)
    w, true, pb, model, win = _case("fp64", "discrete")
    J, gl, gm, gr = _grads(pb, model)
    _, _, pbr, _, _ = _case("fp64")
    _, rl, rm, rr = _grads(pbr, model)
    rng = np.random.default_rng(4)
    pts = [(int(rng.integers(w.zoff + 2, w.zoff + w.nz - 2)), int(rng.integers(w.ppad + 2, w.ppad + w.nx - 2)))
           for _ in range(50)]
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                  dz=w.dx, hc=pb.hc, precision="fp64")
    plan.set_pml(pb.pml)
    plan.set_shots(pb.shots)

    def misfit(arrs):
        plan.set_model(arrs[0], arrs[1], *P.average_parameters(arrs[1], arrs[2]))
        plan.forward(energy=False)
        return float(plan.misfit(pb.observed)[0])

    for k, g, gref in ((0, gl, rl), (1, gm, rm), (2, gr, rr)):
        fd, gg, gr_ = [], [], []
        for z, x in pts:
            h = 1e-4 * model[k][z, x]
            up = [a.copy() for a in model]
            dn = [a.copy() for a in model]
            up[k][z, x] += h
            dn[k][z, x] -= h
            fd.append((misfit(up) - misfit(dn)) / (2 * h))
            gg.append(g[z, x])
            gr_.append(gref[z, x])
        fd, gg = np.array(fd), np.array(gg)
        err = np.abs(fd - gg).max() / np.abs(g).max()
        print(f"  {('lambda', 'mu', 'rho')[k]}: max |fd - g| / max |g| = {err:.2e}; correlation "
              f"discrete {np.corrcoef(fd, gg)[0, 1]:.6f}, reference {np.corrcoef(fd, gr_)[0, 1]:.3f}")
        assert err < 1e-6
