"""The host FWI loop over the C ABI (cuda_fwi_amd/fwi.py) on a small
layered / faulted model with a body (the c5 model shrunk to 240 x 80):
observed traces from the truth, start = the truth smoothed; with the
discrete gradient the misfit falls at every iteration and vp moves toward
the truth."""
import numpy as np
import pytest

from cuda_fwi_amd import psvfwi as P
from cuda_fwi_amd import workloads as WL
from cuda_fwi_amd.fwi import Inversion

pytestmark = pytest.mark.gpu


def _setup():
    w = WL.c5(nshots=6, nrec=60, nt=900, nx=240, nz=80)
    w.npml = 20
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                  dz=w.dx, hc=P.holberg(4, 1), precision="fp64")
    plan.set_pml(P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt))
    shots = [w.padded_shot(s, P.ricker(s["nt"], w.dt, s["amp"], w.fc)[None]) for s in w.shots]
    plan.set_shots(shots)
    truth = w.padded()
    plan.set_model(truth[0], truth[1], *P.average_parameters(truth[1], truth[2]))
    plan.forward(energy=False)
    observed = [plan.records(i) for i in range(len(w.shots))]
    win = (w.zoff, w.zoff + w.nz, w.ppad, w.ppad + w.nx)
    return w, plan, shots, observed, truth, win


def test_fwi_discrete_gradient_reduces_misfit_and_model_error():
    print("host FWI loop, 240 x 80, 6 shots, 4 iterations", flush=True)
    w, plan, shots, observed, truth, win = _setup()
    inv = Inversion(plan, shots, observed, w.padded(*w.background), win, segment=50)
    e0 = inv.model_error(truth)
    h = inv.run(4, truth=truth, log=print)
    print(f"  misfit {h.misfit}; vp error {e0:.5f} -> {h.model_error}")
    assert all(b < a for a, b in zip(h.misfit, h.misfit[1:]))
    assert h.model_error[-1] < e0


def test_fwi_shot_groups_equal_one_batch():
    """Running the shots in groups of 4 and 1 gives the same iterations as
    one batch (misfits and model to FP64 summation order)."""
    print("host FWI loop in shot groups", flush=True)
    runs = []
    for group in (None, 4, 1):
        w, plan, shots, observed, truth, win = _setup()
        inv = Inversion(plan, shots, observed, w.padded(*w.background), win, segment=50, group=group)
        h = inv.run(2)
        runs.append((h.misfit, inv.vp()))
    for mis, vp in runs[1:]:
        assert np.allclose(mis, runs[0][0], rtol=1e-12, atol=0)
        assert np.allclose(vp, runs[0][1], rtol=1e-9, atol=0)


def test_receiver_energy_run_leaves_the_discrete_gradient_unchanged():
    """precond="source_receiver" runs an adjoint without the gradient (for the
    receiver side's energy) between the forward and the discrete adjoint: the
    gradient and misfit equal those of the plain loop bit for bit, the receiver
    energy is positive, and it differs from the source energy."""
    print("receiver-side energy between forward and discrete adjoint", flush=True)
    out = []
    for precond in ("source", "source_receiver"):
        w, plan, shots, observed, truth, win = _setup()
        inv = Inversion(plan, shots, observed, w.padded(*w.background), win, segment=50, group=4,
                        precond=precond)
        L2, ill, g = inv._gradient()
        out.append((L2, ill, g, inv.adjoints))
    (L2a, ws, ga, na), (L2b, wsr, gb, nb) = out
    assert L2a == L2b
    for a, b in zip(ga, gb):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)
    assert (na, nb) == (2, 4)
    z1, z2, x1, x2 = win
    assert np.all(wsr[z1:z2, x1:x2] > 0) and not np.allclose(wsr, ws)


def test_fwi_direction_options_reduce_misfit_and_model_error():
    """The direction options together (sqrt(Ws Wr) illumination, smoothing, depth gain, mute
    and taper) on the same small model: the misfit falls at every iteration, the model error
    falls, and the muted rows keep the start model."""
    print("host FWI loop with the direction options, 240 x 80, 6 shots, 3 iterations", flush=True)
    w, plan, shots, observed, truth, win = _setup()
    start = w.padded(*w.background)
    inv = Inversion(plan, shots, observed, start, win, segment=50, precond="source_receiver", smooth=2.0,
                    depth_power=2.0, mute=3, taper=6)
    e0 = inv.model_error(truth)
    h = inv.run(3, truth=truth, log=print)
    print(f"  misfit {h.misfit}; vp error {e0:.5f} -> {h.model_error}")
    assert all(b < a for a, b in zip(h.misfit, h.misfit[1:]))
    assert h.model_error[-1] < e0
    z1, z2, x1, x2 = win
    for k, i in (("lam", 0), ("mu", 1), ("rho", 2)):
        assert np.array_equal(inv.m[k][z1:z1 + 3, x1:x2], start[i][z1:z1 + 3, x1:x2])
