"""The 201 x 201 configuration (the reference's CPU-runnable case size) end to
end against the oracle: 201 x 201 cells, dx 5 m, npml 20, free surface, a
+10 % Gaussian anomaly (20 m) in vp / vs / rho at the centre, a 15 Hz Ricker
at (100, 2), 101 receivers along the surface, dt 0.5 ms, nt 2000.

- order 4 with the explosive source: the forward of the anomaly model, FP64
  bit for bit and FP32 within 1e-5 relative L2 of the FP64 oracle;
- observed = forward of the anomaly model, synthetic = forward of the
  homogeneous background; one forward + misfit + adjoint + gradient over the
  model window (time decimation 10), orders 4 (explosive source) and 2 (vz
  force, the reference's own source type): gradients FP64 bit for bit and
  FP32 bit for bit against the FP32 oracle.  Against the FP64 gradient the
  FP32 one is limited by FP32 itself, not by the kernels: the anomaly is
  small (20 m), its scattered field is 1e-3 to 1e-4 of the direct wave, so
  FP32 traces resolve the residual only to about 3e-3, and the rounding noise
  of the direct wave correlates with the strong forward field near the source
  and receivers.  Measured: 1.7e-2 to 3.7e-2 relative L2 with the explosive
  source; the test bounds it at 0.1 and the direction (correlation) at
  0.999.
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
    p32 = plan_of(pb, "fp32")
    p32.forward(energy=True)
    e = rel_l2(p32.records(), ref["rec"])
    print(f"  fp32 seismogram rel L2 vs fp64 oracle: {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("fdorder,comp", [(4, 4), (2, 2)])
def test_c1_fwd_adj_gradient(fdorder, comp):
    print(f"201 x 201 fwd + adj + gradient, order {fdorder}, source comp {comp}", flush=True)
    obs = O.forward(_problem(fdorder, comp))["rec"]
    pb = _problem(fdorder, comp, background=True, fwi_dt=10)
    f = O.forward(pb)
    m, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"].copy(), obs)
    a = O.adjoint(pb, pb.rec, res, f["store"])
    p64 = plan_of(pb, "fp64")
    p64.forward(gradient=True)
    mis = p64.misfit([obs])
    assert abs(mis[0] - m) <= 1e-12 * m
    p64.adjoint(gradient=True)
    g64 = p64.gradient(0)
    for got, k in zip(g64, ("grad_lam", "grad_mu", "grad_rho")):
        assert np.abs(a[k]).max() > 0
        assert np.array_equal(got, a[k]), k
    p32 = plan_of(pb, "fp32")
    p32.forward(gradient=True)
    p32.misfit([obs])
    p32.adjoint(gradient=True)
    g32 = p32.gradient(0)
    f32 = O.forward(pb, prec="f32")
    res32 = (f32["rec"].astype(np.float64) - obs).astype(np.float32)
    a32 = O.adjoint(pb, pb.rec, res32, f32["store"], prec="f32")
    for got, k in zip(g32, ("grad_lam", "grad_mu", "grad_rho")):
        assert np.array_equal(got, a32[k].astype(np.float64)), k
    eres = rel_l2(res32, res)
    errs = [rel_l2(x, y) for x, y in zip(g32, g64)]
    corr = [float(np.corrcoef(x.ravel(), y.ravel())[0, 1]) for x, y in zip(g32, g64)]
    print(f"  fp32 residual rel L2 {eres:.2e}; fp32 gradients vs fp64: rel L2 {[f'{e:.2e}' for e in errs]}, "
          f"correlation {[f'{c:.6f}' for c in corr]}")
    assert max(errs) < 0.1 and min(corr) > 0.999
