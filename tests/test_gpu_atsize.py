"""Parity of the configurations at their full sizes.

- 1000 x 500, nt 5000, 400 receivers, decimated store fwi_dt 5 on both sides: the HIP path
  against the oracle itself (its row-parallel build, the same arithmetic): FP64 traces,
  misfit and gradients bit for bit; FP32 traces within 1e-5 and gradients within 1e-4
  relative L2 of the FP64 oracle.
- The 64-shot 1500 x 600 batch at its full record lengths (nt 3000-8000, 50-400 receivers):
  all 64 shots in one plan with checkpoints every 200 steps; three shots (the longest, the
  shortest and one between) equal single-shot plans with the full decimated store bit for
  bit (traces, misfit, gradients).
- The 2000 x 1000, nt 8000 memory workload: the gradient from checkpoints every 200 steps
  equals the one from the every-step store on a 96 x 96 window bit for bit, and the whole-model
  gradient with checkpoints stays within 20 GB of device memory, where the every-step store
  alone would need 320 GB.
- One full-size group of the 2400 x 800 FWI configuration (8 shots, nt 6000, 600 receivers,
  8 Hz): FP32 traces within 1e-5 and gradients within 1e-4 relative L2 of the FP64 kernels,
  and linearity in the source (twice the signal: twice the traces, four times the gradient
  against zero observed data, to FP32 rounding).
"""
import ctypes

import numpy as np
import pytest

import psv_oracle as O
from cuda_fwi_amd import psvfwi as P
from cuda_fwi_amd import workloads as WL
from gpucase import rel_l2

pytestmark = pytest.mark.gpu


def _plan(w, prec, shots, window, store="full", interval=0, tile=8, scale=1.0):
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                  dz=w.dx, hc=P.holberg(w.fdorder, w.max_relative_error), precision=prec)
    plan.set_pml(P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt))
    plan.set_window(window)
    plan.set_store_mode(store, interval)
    plan.set_shots([w.padded_shot(s, scale * P.ricker(s["nt"], w.dt, s["amp"], w.fc)[None]) for s in shots])
    plan.set_tile(tile)
    return plan


def _models(w):
    lt, mt, rt = w.padded()
    lam, mu, rho = w.padded(*w.background)
    return (lt, mt, *P.average_parameters(mt, rt)), (lam, mu, *P.average_parameters(mu, rho))


def _observed(w, truth, shots, window):
    p = _plan(w, "fp64", shots, window)
    p.set_model(*truth)
    p.forward(energy=False)
    obs = [p.records(i) for i in range(len(shots))]
    p.close()
    return obs


def _gradient_run(plan, model, obs):
    plan.set_model(*model)
    plan.forward(energy=False, gradient=True)
    n = len(obs)
    rec = [plan.records(i) for i in range(n)]
    mis = plan.misfit(obs)
    plan.adjoint(energy=False, gradient=True)
    return rec, mis, [plan.gradient(i) for i in range(n)]


def test_c2_full_size_against_the_threaded_oracle():
    print("1000 x 500, nt 5000: against the oracle (row-parallel build)", flush=True)
    w = WL.c2()
    window = w.model_window(dt=5)
    truth, start = _models(w)
    obs = _observed(w, truth, w.shots, window)
    # the oracle on the starting model, same geometry, fwi_dt 5
    s = w.shots[0]
    ps = w.padded_shot(s, P.ricker(s["nt"], w.dt, s["amp"], w.fc)[None])
    src = dict(ps["src"], shot=[0] * len(ps["src"]["x"]))
    rec = dict(ps["rec"], shot=[0] * len(ps["rec"]["x"]))
    lam, mu, mu_zx, rho_zp, rho_xp = start
    pb = O.Problem(nt=w.nt, nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, dt=w.dt, dx=w.dx, dz=w.dx,
                   fsurf=w.fsurf, hc=P.holberg(w.fdorder, w.max_relative_error), lam=lam, mu=mu, mu_zx=mu_zx,
                   rho_zp=rho_zp, rho_xp=rho_xp,
                   pml=P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt),
                   src=src, src_signal=ps["src_signal"], rec=rec, fwi=window)
    O.set_threaded(True)
    try:
        f = O.forward(pb)
        m, res, _ = O.l2_adjoint_sources(w.dt, f["rec"].copy(), obs[0])
        print("  oracle forward done", flush=True)
        a = O.adjoint(pb, rec, res, f["store"])
    finally:
        O.set_threaded(False)
    del f["store"]
    print("  oracle adjoint done", flush=True)
    grads = [a[k] for k in ("grad_lam", "grad_mu", "grad_rho")]
    p64 = _plan(w, "fp64", w.shots, window)
    r64, m64, g64 = _gradient_run(p64, start, obs)
    p64.close()
    assert np.array_equal(r64[0], f["rec"])
    assert abs(m64[0] - m) <= 1e-12 * m
    for got, ref in zip(g64[0], grads):
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref)
    p32 = _plan(w, "fp32", w.shots, window)
    r32, m32, g32 = _gradient_run(p32, start, obs)
    p32.close()
    es = rel_l2(r32[0], f["rec"])
    eg = [rel_l2(x, y) for x, y in zip(g32[0], grads)]
    print(f"  fp32 vs the fp64 oracle: seismogram {es:.2e}, gradients {[f'{e:.2e}' for e in eg]}, "
          f"misfit {abs(m32[0] - m) / m:.2e}", flush=True)
    assert es < 1e-5 and max(eg) < 1e-4


def test_c4_full_batch_sample_equals_single_shots():
    print("64 shots of 1500 x 600 at full nt: batch (checkpoints) against single shots", flush=True)
    w = WL.c4(nshots=64)
    window = w.model_window(dt=5)
    truth, start = _models(w)
    nts = [s["nt"] for s in w.shots]
    pick = [int(np.argmax(nts)), int(np.argmin(nts)), 17]
    obs_pick = _observed(w, truth, [w.shots[i] for i in pick], window)
    # the batch needs observed traces for all shots: zero traces except the picked ones
    obs = [np.zeros((len(s["rec"]["x"]), s["nt"])) for s in w.shots]
    for i, o in zip(pick, obs_pick):
        obs[i] = o
    batch = _plan(w, "fp32", w.shots, window, store="checkpoint", interval=200)
    rec_b, mis_b, grad_b = _gradient_run(batch, start, obs)
    ms_f = batch.timing()[0]
    batch.close()
    print(f"  batch done (last adjoint {ms_f / 1e3:.1f} s)", flush=True)
    for i, o in zip(pick, obs_pick):
        one = _plan(w, "fp32", [w.shots[i]], window)
        rec, mis, grad = _gradient_run(one, start, [o])
        one.close()
        print(f"  shot {i} (nt {nts[i]}, {len(w.shots[i]['rec']['x'])} receivers)", flush=True)
        assert np.array_equal(rec[0], rec_b[i]), i
        assert mis[0] == mis_b[i], i
        for a, b in zip(grad[0], grad_b[i]):
            assert np.abs(b).max() > 0
            assert np.array_equal(a, b), i


def _device_used():
    """Bytes of device memory in use (hipMemGetInfo)."""
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return total.value - free.value


def test_c3_memory_workload_checkpoints():
    print("2000 x 1000, nt 8000: checkpoints against the every-step store on a window", flush=True)
    w = WL.c3mem()
    _, start = _models(w)
    rng = np.random.default_rng(3)
    obs = [1e-3 * rng.standard_normal((len(w.shots[0]["rec"]["x"]), w.nt))]
    x1, z1 = w.ppad + 950, w.zoff + 450
    win = dict(dt=1, dx=1, dz=1, x1=x1, x2=x1 + 96, z1=z1, z2=z1 + 96)
    out = []
    for store, k in (("full", 0), ("checkpoint", 200)):
        plan = _plan(w, "fp32", w.shots, win, store=store, interval=k)
        rec, mis, grad = _gradient_run(plan, start, obs)
        plan.close()
        out.append(grad[0])
        print(f"  {store} done", flush=True)
    for a, b in zip(*out):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)
    base = _device_used()
    plan = _plan(w, "fp32", w.shots, w.model_window(dt=1), store="checkpoint", interval=200)
    plan.set_model(*start)
    plan.forward(energy=False, gradient=True)
    used = _device_used() - base
    plan.misfit(obs)
    plan.adjoint(energy=False, gradient=True)
    used = max(used, _device_used() - base)
    assert np.abs(plan.gradient(0)[0]).max() > 0
    plan.close()
    every_step = 5 * w.nx * w.nz * w.nt * 4
    print(f"  device memory {used / 1e9:.1f} GB with checkpoints; every-step store {every_step / 1e9:.0f} GB",
          flush=True)
    assert used < 20e9 and every_step > 288e9


def test_c5_full_size_group_fp32_and_linearity():
    print("2400 x 800, 8 shots, nt 6000: FP32 against FP64, linearity", flush=True)
    w = WL.c5()
    shots = w.shots[::16]
    window = w.model_window(dt=5)
    truth, start = _models(w)
    obs = _observed(w, truth, shots, window)
    zero = [np.zeros_like(o) for o in obs]
    runs = {}
    for prec, scale, ob in (("fp64", 1.0, obs), ("fp32", 1.0, obs), ("fp32", 2.0, zero), ("fp32", 1.0, zero)):
        plan = _plan(w, prec, shots, window, store="checkpoint", interval=200, scale=scale)
        runs[(prec, scale, ob is zero)] = _gradient_run(plan, start, ob)
        plan.close()
        print(f"  {prec} x{scale} {'zero observed' if ob is zero else 'observed'} done", flush=True)
    r64, _, g64 = runs[("fp64", 1.0, False)]
    r32, _, g32 = runs[("fp32", 1.0, False)]
    es = max(rel_l2(a, b) for a, b in zip(r32, r64))
    eg = max(rel_l2(a, b) for ga, gb in zip(g32, g64) for a, b in zip(ga, gb))
    ra, _, ga = runs[("fp32", 1.0, True)]
    rb, _, gb = runs[("fp32", 2.0, True)]
    lin_r = max(rel_l2(b, 2.0 * a) for a, b in zip(ra, rb))
    lin_g = max(rel_l2(b, 4.0 * a) for x, y in zip(ga, gb) for a, b in zip(x, y))
    print(f"  fp32 vs fp64: seismograms {es:.2e}, gradients {eg:.2e}; linearity: traces {lin_r:.2e}, "
          f"gradients {lin_g:.2e}", flush=True)
    assert np.abs(r64[0]).max() > 0
    assert es < 1e-5 and eg < 1e-4
    assert lin_r < 1e-5 and lin_g < 1e-5
