"""Random-boundary reconstruction (an extension; PSVFWI_STORE_RANDOM_BOUNDARY):
with no CPML and a random boundary layer in the model, the adjoint rebuilds
the forward state by running the scheme backward from the forward's final
state.  Its gradient equals the stored-field gradient of the same model up to
the rounding of the reverse steps: measured about 4e-16 relative L2 in FP64
and 2e-7 in FP32 over 500 steps (bounds 1e-12 and 1e-5)."""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from cuda_fwi_amd import psvfwi as P
from gpucase import plan_of, rel_l2

pytestmark = pytest.mark.gpu


def _problem(fdorder, fsurf, nt=500, width=20):
    pb = RC.small_problem(nx=110, nz=90, npml=0, nt=nt, fdorder=fdorder, fsurf=fsurf, seed=3, nrec=7, nsrc=2)
    lam, mu, rho = P.random_boundary(pb.lam, pb.mu, pb.rho, width, fsurf, seed=5, max_increase=0.4)
    pb.lam, pb.mu, pb.rho = lam, mu, rho
    pb.mu_zx, pb.rho_zp, pb.rho_xp = O.average_parameters(mu, rho)
    # sources and receivers inside the layer-free part of the grid
    top = pb.fpad + (0 if fsurf else width)
    pb.src = dict(x=[45, 75], z=[top + 3, top + 5], comp=[2, 2], shot=[0, 0])
    xr = list(np.linspace(width + 5, pb.nxt - width - 6, 7).astype(int))
    pb.rec = dict(x=xr, z=[top + 1] * 7, comp=[2] * 6 + [1], shot=[0] * 7)
    pb.fwi = dict(dt=2, dx=1, dz=1, x1=pb.fpad, x2=pb.nxt - pb.fpad, z1=pb.fpad, z2=pb.nzt - pb.fpad)
    obs = np.random.default_rng(4).standard_normal((7, nt)) * 1e-3
    return pb, obs


def _gradient(pb, obs, prec, mode):
    plan = plan_of(pb, prec)
    plan.set_store_mode(mode, 0)
    plan.forward(energy=True, gradient=True)
    mis = plan.misfit([obs])
    plan.adjoint(energy=True, gradient=True)
    return plan, mis, plan.gradient(0)


CASES = [dict(fdorder=4, fsurf=0), dict(fdorder=4, fsurf=1), dict(fdorder=2, fsurf=0)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"o{c['fdorder']}fs{c['fsurf']}")
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_random_boundary_gradient_equals_stored(case, prec, tol):
    print(f"random boundary {case} {prec}", flush=True)
    pb, obs = _problem(**case)
    _, m_full, g_full = _gradient(pb, obs, prec, "full")
    _, m_rb, g_rb = _gradient(pb, obs, prec, "random_boundary")
    assert m_rb[0] == m_full[0]
    errs = [rel_l2(a, b) for a, b in zip(g_rb, g_full)]
    print(f"  reconstructed vs stored gradient rel L2: {[f'{e:.2e}' for e in errs]}")
    for g in g_full:
        assert np.abs(g).max() > 0
    assert max(errs) < tol


def test_random_boundary_batched_uneven_shots():
    """Shots of different lengths in one batch: each shot's reconstructed
    gradient matches its stored one."""
    print("random boundary batched", flush=True)
    pb, _ = _problem(4, 0, nt=10)
    rng = np.random.default_rng(6)
    shots, obs = [], []
    for i, nt in enumerate((420, 300, 180)):
        sig = O.create_wavelet(nt, pb.dt, 1.0 + i, 25.0, 0.0)[None]
        shots.append(dict(nt=nt, src=dict(x=[40 + 15 * i], z=[26], comp=[2]), src_signal=sig,
                          rec=dict(x=[30, 55, 80], z=[22, 22, 22], comp=[2, 2, 2])))
        obs.append(rng.standard_normal((3, nt)) * 1e-3)
    out = []
    for mode in ("full", "random_boundary"):
        plan = plan_of(pb, "fp64", shots=shots)
        plan.set_store_mode(mode, 0)
        plan.forward(gradient=True)
        plan.misfit(obs)
        plan.adjoint(gradient=True)
        out.append([plan.gradient(i) for i in range(3)])
    for i in range(3):
        errs = [rel_l2(a, b) for a, b in zip(out[1][i], out[0][i])]
        assert max(errs) < 1e-12, (i, errs)


def test_random_boundary_refusals():
    print("random boundary refusals", flush=True)
    pb, obs = _problem(2, 1, nt=20)
    plan = plan_of(pb, "fp64")
    plan.set_store_mode("random_boundary", 0)
    with pytest.raises(P.PsvfwiError, match="order-2 free surface"):
        plan.forward(gradient=True)
    pb, obs = _problem(4, 0, nt=20)
    plan = plan_of(pb, "fp64", tz=-8)
    plan.set_store_mode("random_boundary", 0)
    with pytest.raises(P.PsvfwiError, match="split step"):
        plan.forward(gradient=True)
    pb = RC.small_problem(nx=40, nz=30, npml=6, nt=20, fdorder=4, fsurf=0)
    pb.fwi = RC.full_window(pb)
    plan = plan_of(pb, "fp64")
    plan.set_store_mode("random_boundary", 0)
    with pytest.raises(P.PsvfwiError, match="npml = 0"):
        plan.forward(gradient=True)
