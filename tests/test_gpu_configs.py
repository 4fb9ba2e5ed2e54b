"""The 201 x 201 configuration (the reference's CPU-runnable case size) end to
end against the oracle: 201 x 201 cells, dx 5 m, npml 20, free surface, a
+10 % Gaussian anomaly (20 m) in vp / vs / rho at the centre, a 15 Hz Ricker
at (100, 2), 101 receivers along the surface, dt 0.5 ms, nt 2000.

- order 4 with the explosive source: the forward of the anomaly model, FP64
  bit for bit; FP32 and compensated FP32 (fp32x2) within 1e-5 relative L2 of
  the FP64 oracle;
- observed = forward of the anomaly model, synthetic = forward of the
  homogeneous background; one forward + misfit + adjoint + gradient over the
  model window (time decimation 10), orders 4 (explosive source) and 2 (vz
  force, the reference's own source type): FP64 bit for bit, FP32 bit for bit
  against the FP32 oracle, and compensated FP32 within the stated tolerances
  of the FP64 oracle: seismograms 1e-5 and gradients 1e-4 relative L2.
- Plain FP32 does not meet the gradient tolerance on this case, and the test
  that asserts it is an expected failure: the anomaly is small (20 m), its
  scattered field is 1e-3 to 1e-4 of the direct wave, so FP32 synthetics
  resolve the residual only to about 3e-3, and the rounding noise of the
  direct wave correlates with the strong forward field near the source and
  receivers (measured 1.7e-2 to 3.7e-2 relative L2).  The kernels are not the
  cause (they equal the FP32 oracle bit for bit); compensated FP32 removes it.
"""
import numpy as np
import pytest

import psv_oracle as O
from cuda_fwi_amd import workloads as WL
from gpucase import plan_of, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vz", "sxx", "szx", "szz", "We")


def _problem(fdorder=4, comp=4, background=False, fwi_dt=None):
    w = WL.c1()
    w.fdorder = fdorder
    if background:
        lam, mu, rho = w.padded(*w.background)
    else:
        lam, mu, rho = w.padded()
    s = w.shots[0]
    sig = O.create_wavelet(w.nt, w.dt, s["amp"], w.fc, 0.0, 1)[None]
    src = dict(x=[s["src"]["x"][0] + w.ppad], z=[s["src"]["z"][0] + w.zoff], comp=[comp], shot=[0])
    rec = dict(x=[x + w.ppad for x in s["rec"]["x"]], z=[z + w.zoff for z in s["rec"]["z"]],
               comp=list(s["rec"]["comp"]), shot=[0] * len(s["rec"]["x"]))
    pml = O.cpml_pro(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt)
    fwi = w.model_window(dt=fwi_dt) if fwi_dt else None
    return O.Problem(nt=w.nt, nzt=w.nzt, nxt=w.nxt, fdorder=fdorder, npml=w.npml, dt=w.dt, dx=w.dx, dz=w.dx,
                     fsurf=w.fsurf, hc=O.holbergcoeff(fdorder, 1), lam=lam, mu=mu, rho=rho, pml=pml, src=src,
                     src_signal=sig, rec=rec, fwi=fwi)


def test_c1_forward_explosive_order4():
    print("201 x 201 forward, explosive source, order 4", flush=True)
    pb = _problem()
    ref = O.forward(pb)
    assert np.abs(ref["rec"]).max() > 0
    p64 = plan_of(pb, "fp64")
    p64.forward(energy=True)
    assert np.array_equal(p64.records(), ref["rec"])
    for k in FIELDS:
        assert np.array_equal(p64.field(k), ref[k]), k
    for prec in ("fp32", "fp32x2"):
        p = plan_of(pb, prec)
        p.forward(energy=True)
        e = rel_l2(p.records(), ref["rec"])
        print(f"  {prec} seismogram rel L2 vs fp64 oracle: {e:.2e}")
        assert e < 1e-5


def _c1_gradients(fdorder, comp, precs):
    """Oracle FP64 (forward, residual, gradient) and each precision's misfit, traces and
    gradient on the GPU."""
    obs = O.forward(_problem(fdorder, comp))["rec"]
    pb = _problem(fdorder, comp, background=True, fwi_dt=10)
    f = O.forward(pb)
    m, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"].copy(), obs)
    a = O.adjoint(pb, pb.rec, res, f["store"])
    out = {}
    for prec in precs:
        p = plan_of(pb, prec)
        p.forward(gradient=True)
        mis = p.misfit([obs])[0]
        rec = p.records()
        p.adjoint(gradient=True)
        out[prec] = dict(misfit=mis, rec=rec, grad=p.gradient(0))
        p.close()
    return pb, obs, f, m, res, a, out


GRADS = ("grad_lam", "grad_mu", "grad_rho")


@pytest.mark.parametrize("fdorder,comp", [(4, 4), (2, 2)])
def test_c1_fwd_adj_gradient(fdorder, comp):
    print(f"201 x 201 fwd + adj + gradient, order {fdorder}, source comp {comp}", flush=True)
    pb, obs, f, m, res, a, out = _c1_gradients(fdorder, comp, ("fp64", "fp32", "fp32x2"))
    o64 = out["fp64"]
    assert abs(o64["misfit"] - m) <= 1e-12 * m
    for got, k in zip(o64["grad"], GRADS):
        assert np.abs(a[k]).max() > 0
        assert np.array_equal(got, a[k]), k
    # plain FP32 against the FP32 oracle: bit for bit
    f32 = O.forward(pb, prec="f32")
    res32 = (f32["rec"].astype(np.float64) - obs).astype(np.float32)
    a32 = O.adjoint(pb, pb.rec, res32, f32["store"], prec="f32")
    for got, k in zip(out["fp32"]["grad"], GRADS):
        assert np.array_equal(got, a32[k].astype(np.float64)), k
    for prec in ("fp32", "fp32x2"):
        o = out[prec]
        es = rel_l2(o["rec"], f["rec"])
        eg = [rel_l2(x, a[k]) for x, k in zip(o["grad"], GRADS)]
        em = abs(o["misfit"] - m) / m
        print(f"  {prec} vs fp64 oracle: seismogram rel L2 {es:.2e}, misfit rel {em:.2e}, "
              f"gradients rel L2 {[f'{e:.2e}' for e in eg]}")
    # compensated FP32: the stated tolerances
    o = out["fp32x2"]
    assert rel_l2(o["rec"], f["rec"]) < 1e-5
    assert max(rel_l2(x, a[k]) for x, k in zip(o["grad"], GRADS)) < 1e-4
    assert abs(o["misfit"] - m) < 1e-6 * m


@pytest.mark.xfail(strict=True, reason="plain FP32 resolves the residual of this 20 m anomaly (1e-3 to 1e-4 "
                                       "of the direct wave) only to about 3e-3: its gradient differs from "
                                       "the FP64 one by 1.7e-2 to 3.7e-2 relative L2; compensated FP32 "
                                       "(fp32x2) meets the 1e-4 tolerance")
def test_c1_plain_fp32_gradient_within_tolerance():
    print("201 x 201 gradient, plain FP32 against the 1e-4 tolerance", flush=True)
    pb, obs, f, m, res, a, out = _c1_gradients(4, 4, ("fp32",))
    errs = [rel_l2(x, a[k]) for x, k in zip(out["fp32"]["grad"], GRADS)]
    print(f"  fp32 gradients rel L2 vs fp64 oracle {[f'{e:.2e}' for e in errs]}")
    assert max(errs) < 1e-4
