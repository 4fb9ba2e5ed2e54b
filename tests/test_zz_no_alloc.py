"""psvfwi_misfit_noalloc, the misfit entry point that allocates no device memory, against
psvfwi_misfit on the same plan: per-shot misfits and the residuals it leaves as adjoint
sources (seen through the adjoint's gradient) equal bit for bit, on two calls in a row, the
second on a smaller batch (fewer, shorter shots) so that the static scratch still holds the
first call's values; and the batch-size limit is refused."""
import ctypes as C

import numpy as np
import pytest

import refcase as RC
from cuda_fwi_amd import psvfwi as P
from gpucase import plan_of, shots_of

pytestmark = pytest.mark.gpu


def _noalloc(plan, n):
    out = np.zeros(n)
    st = P.lib().psvfwi_misfit_noalloc(plan._h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out


def _case(nshots, nt, seed, precision):
    pb = RC.small_problem(nx=48, nz=32, npml=6, nt=nt, fdorder=4, fsurf=1, seed=seed, nsrc=1, nrec=5)
    pb.fwi = RC.full_window(pb, dt=2)
    shots = []
    for k in range(nshots):
        s = shots_of(pb)
        s["nt"] = nt - 17 * k          # uneven record lengths
        s["src_signal"] = s["src_signal"][:, :s["nt"]] * (1.0 + 0.25 * k)
        shots.append(s)
    rng = np.random.default_rng(seed)
    obs = [rng.standard_normal((5, s["nt"])) * 1e-9 for s in shots]
    plan = plan_of(pb, precision, shots=shots)
    plan.set_observed(obs)

    plan.forward(energy=False, gradient=True)
    m_ref = plan.misfit()
    plan.adjoint(energy=False, gradient=True)
    g_ref = [plan.gradient(i) for i in range(nshots)]

    plan.forward(energy=False, gradient=True)
    st, m = _noalloc(plan, nshots)
    assert st == 0, P.lib().psvfwi_last_error().decode()
    plan.adjoint(energy=False, gradient=True)
    g = [plan.gradient(i) for i in range(nshots)]
    plan.close()
    assert np.all(m_ref > 0)
    assert np.array_equal(m, m_ref)
    for a, b in zip(g, g_ref):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_misfit_equals_misfit_twice_then_smaller(precision):
    print(f"no-alloc misfit {precision}: 4 shots nt 220, then 2 shots nt 120", flush=True)
    _case(4, 220, 3, precision)
    _case(2, 120, 4, precision)


def test_noalloc_misfit_refuses_batches_beyond_its_scratch():
    print("no-alloc misfit: batch of 1025 shots refused", flush=True)
    pb = RC.small_problem(nx=12, nz=10, npml=0, nt=3, fdorder=2, fsurf=0, seed=1, nsrc=1, nrec=1)
    pb.fwi = None
    plan = plan_of(pb, "fp32", shots=[shots_of(pb)] * 1025)
    plan.forward(energy=False)
    plan.set_observed([np.zeros((1, 3))] * 1025)
    st, _ = _noalloc(plan, 1025)
    assert st == -1, st
    assert "PSVFWI_NOALLOC_MAX_SHOTS" in P.lib().psvfwi_last_error().decode()
    plan.close()
