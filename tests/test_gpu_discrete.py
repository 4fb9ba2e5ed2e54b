"""Discrete gradient (an extension: the exact gradient of the misfit of the
scheme) on the GPU against the oracle's discrete adjoint
(oracle/psv_oracle_discrete.inc), which central differences of the misfit
pin (test_oracle_golden.py).

The kernels evaluate the oracle's per-cell expressions in the same order, so
the five coefficient gradients agree bit for bit in FP64 and in FP32.
"""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from gpucase import plan_of, rel_l2, shots_of

pytestmark = pytest.mark.gpu

COEF = ("g_lam", "g_mu", "g_mu_zx", "g_rho_zp", "g_rho_xp")


def _observed(pb, seed=7):
    rng = np.random.default_rng(seed)
    return O.forward(pb)["rec"] * 0.7 + 1e-3 * rng.standard_normal((len(pb.rec["x"]), pb.nt))


def _device_gradient(pb, obs, prec, segment, tz=None, shots=None):
    plan = plan_of(pb, prec, tz=tz, shots=shots)
    plan.set_gradient_mode("discrete", segment)
    plan.forward(energy=False, gradient=True)
    mis = plan.misfit(obs if isinstance(obs, list) else [obs])
    plan.adjoint(energy=False, gradient=True)
    return plan, mis


CASES = [
    dict(fdorder=2, fsurf=1, npml=8), dict(fdorder=4, fsurf=0, npml=8),
    dict(fdorder=4, fsurf=1, npml=6), dict(fdorder=2, fsurf=0, npml=0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"o{c['fdorder']}fs{c['fsurf']}p{c['npml']}")
@pytest.mark.parametrize("segment,tz", [(7, 8), (64, -16), (500, 16), (1, 8), (1, -16)])
def test_discrete_gradient_fp64_bit_exact(case, segment, tz):
    print(f"discrete fp64 {case} segment {segment} tile {tz}", flush=True)
    pb = RC.small_problem(nx=50, nz=36, nt=160, seed=4, nrec=6, **case)
    obs = _observed(pb)
    f = O.forward(pb)
    m, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"].copy(), obs)
    ref = O.discrete_gradient(pb, res)
    plan, mis = _device_gradient(pb, obs, "fp64", segment, tz)
    assert abs(mis[0] - m) <= 1e-12 * m
    g = plan.coefficient_gradient(0)
    for k in COEF:
        assert np.abs(ref[k]).max() > 0, k
        assert np.array_equal(g[k], ref[k]), k
    ours = plan.model_gradient(0)
    want = O.chain_gradient(pb.mu, pb.mu_zx, pb.rho_zp, pb.rho_xp, ref)
    for a, b in zip(ours, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fdorder", [2, 4])
def test_discrete_gradient_fp32_bit_exact_and_tolerance(fdorder):
    print(f"discrete fp32 order {fdorder}", flush=True)
    pb = RC.small_problem(nx=60, nz=40, nt=300, seed=5, nrec=7, fdorder=fdorder, fsurf=1, npml=8)
    obs = _observed(pb, 3)
    f32 = O.forward(pb, prec="f32")
    res32 = (f32["rec"].astype(np.float64) - obs).astype(np.float32)
    ref32 = O.discrete_gradient(pb, res32, prec="f32")
    plan, _ = _device_gradient(pb, obs, "fp32", 50)
    g = plan.coefficient_gradient(0)
    for k in COEF:
        assert np.array_equal(g[k], ref32[k].astype(np.float64)), k
    # the FP32 path against the FP64 discrete gradient
    f64 = O.forward(pb)
    _, res64, _ = O.l2_adjoint_sources(pb.dt, f64["rec"].copy(), obs)
    ref64 = O.discrete_gradient(pb, res64)
    err = {k: rel_l2(g[k], ref64[k]) for k in COEF}
    print("  fp32 vs fp64 discrete gradient rel L2:", err)
    assert max(err.values()) < 1e-4


@pytest.mark.parametrize("segment", [33, 1])
def test_discrete_gradient_batched_uneven_shots(segment):
    """Shots of different nt and receiver counts in one batch: each shot's
    gradient is its own oracle gradient; the batch sum is their FP64 sum.
    Segment 1: the forward keeps every state (a shot's final state while the
    longer shots go on) and the adjoint reads them without a replay."""
    print(f"discrete batched, segment {segment}", flush=True)
    base = RC.small_problem(nx=56, nz=40, npml=8, nt=10, fdorder=4, fsurf=1, seed=2)
    rng = np.random.default_rng(12)
    shots, obs, refs = [], [], []
    for i, nt in enumerate((130, 200, 75)):
        nrec = int(rng.integers(3, 9))
        xs = int(rng.integers(base.ppad + 4, base.nxt - base.ppad - 4))
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        rec = dict(x=list(np.linspace(base.ppad, base.nxt - base.ppad - 1, nrec).astype(int)),
                   z=[base.fpad + 1] * nrec, comp=[2] * (nrec - 1) + [1])
        s = dict(nt=nt, src=dict(x=[xs], z=[base.fpad + 3], comp=[2]), src_signal=sig, rec=rec)
        shots.append(s)
        o = rng.standard_normal((nrec, nt)) * 1e-3
        obs.append(o)
        pb = RC.small_problem(nx=56, nz=40, npml=8, nt=nt, fdorder=4, fsurf=1, seed=2)
        pb.src = dict(x=[xs], z=[base.fpad + 3], comp=[2], shot=[0])
        pb.src_signal = sig
        pb.rec = dict(rec, shot=[0] * nrec)
        _, res, _ = O.l2_adjoint_sources(pb.dt, O.forward(pb)["rec"].copy(), o)
        refs.append(O.discrete_gradient(pb, res))
    plan, _ = _device_gradient(base, obs, "fp64", segment, shots=shots)
    for i in range(len(shots)):
        g = plan.coefficient_gradient(i)
        for k in COEF:
            assert np.array_equal(g[k], refs[i][k]), (i, k)
    tot = plan.coefficient_gradient(-1)
    for k in COEF:
        want = refs[0][k] + refs[1][k] + refs[2][k]
        assert np.allclose(tot[k], want, rtol=1e-13, atol=0), k


def test_discrete_gradient_many_workgroups():
    """A batch large enough that the reverse-step kernels run in several
    rounds of workgroups (4 shots x 730 workgroups of 256 threads per
    launch), with nonzero values in almost every workgroup's cells: a kernel
    that read a value another workgroup of the same launch writes would
    differ from the oracle here."""
    print("discrete, several rounds of workgroups", flush=True)
    O.set_threaded(True)
    try:
        base = RC.small_problem(nx=620, nz=280, npml=8, nt=10, fdorder=4, fsurf=1, seed=4)
        rng = np.random.default_rng(21)
        xr = list(np.linspace(base.ppad, base.nxt - base.ppad - 1, 32).astype(int))
        rec = dict(x=xr * 2, z=[base.fpad + 1] * 32 + [base.nzt // 2] * 32, comp=[2] * 63 + [1])
        shots, obs, refs = [], [], []
        for i in range(4):
            nt = 120 - 7 * i
            xs = [100 + 11 * i, 300 - 13 * i, 500 + 5 * i]
            zs = [base.nzt // 3 + i] * 3
            sig = np.vstack([O.create_wavelet(nt, base.dt, 1.0 + 0.2 * k, 25.0, 0.0) for k in range(3)])
            shots.append(dict(nt=nt, src=dict(x=xs, z=zs, comp=[2, 2, 1]), src_signal=sig, rec=rec))
            pb = RC.small_problem(nx=620, nz=280, npml=8, nt=nt, fdorder=4, fsurf=1, seed=4)
            pb.src = dict(x=xs, z=zs, comp=[2, 2, 1], shot=[0] * 3)
            pb.src_signal = sig
            pb.rec = dict(rec, shot=[0] * 64)
            d = O.forward(pb)["rec"]
            o = d * 0.5 + 1e-3 * rng.standard_normal(d.shape)
            obs.append(o)
            _, res, _ = O.l2_adjoint_sources(pb.dt, d.copy(), o)
            refs.append(O.discrete_gradient(pb, res))
    finally:
        O.set_threaded(False)
    plan, _ = _device_gradient(base, obs, "fp64", 16, shots=shots)
    for i in range(len(shots)):
        g = plan.coefficient_gradient(i)
        for k in COEF:
            assert (refs[i][k] != 0).mean() > 0.5, (i, k)
            assert np.array_equal(g[k], refs[i][k]), (i, k)


def test_discrete_mode_errors():
    """Explosive sources and a CPML profile with a != 0 at the inner edge are
    refused with a message; the reference-mode getters stay unavailable."""
    from cuda_fwi_amd import psvfwi as P
    print("discrete errors", flush=True)
    pb = RC.small_problem(nx=40, nz=30, nt=40, npml=6, fdorder=2, fsurf=1, src_comp=4)
    obs = np.zeros((len(pb.rec["x"]), pb.nt))
    plan = plan_of(pb, "fp64")
    plan.set_gradient_mode("discrete", 16)
    plan.forward(gradient=True)
    plan.misfit([obs])
    with pytest.raises(P.PsvfwiError, match="vx / vz point sources"):
        plan.adjoint(gradient=True)
    pb = RC.small_problem(nx=40, nz=30, nt=40, npml=6, fdorder=2, fsurf=1)
    plan = plan_of(pb, "fp64")
    prof = pb.pml.copy()
    prof[0, pb.npml] = -1e-3
    plan.set_pml(prof)
    plan.set_gradient_mode("discrete", 16)
    plan.forward(gradient=True)
    plan.misfit([obs])
    with pytest.raises(P.PsvfwiError, match="CPML profile"):
        plan.adjoint(gradient=True)
    with pytest.raises(P.PsvfwiError):
        plan.gradient(0)
    with pytest.raises(P.PsvfwiError):
        plan.set_gradient_mode("discrete", 0)
