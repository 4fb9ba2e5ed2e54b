"""Shared builders for test problems (oracle layout, FP64 numpy)."""
import json
import os

import numpy as np

import psv_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def reference_run(nt=None):
    """The reference's shipped forward run (tests/golden/ref_run.json)."""
    r = json.load(open(os.path.join(HERE, "golden", "ref_run.json")))
    fdorder, npml = r["fdorder"], r["npml"]
    fpad = fdorder // 2
    ppad = npml + fpad
    nxt, nzt = r["nx"] + 2 * ppad, r["nz"] + ppad + fpad
    nt = r["nt"] if nt is None else nt
    lam0, mu0 = O.velocity_to_lame(r["scalar_vp"], r["scalar_vs"], r["scalar_rho"])
    assert r["model"].startswith("homogeneous")
    rho, lam, mu = (np.full((nzt, nxt), v) for v in (r["scalar_rho"], lam0, mu0))
    pml = O.cpml_pro(npml, r["dh"], r["fpml"], r["npower"], r["k_max_PML"], r["rcoef"],
                     r["damp_v_PML"], r["dt"])
    hc = O.holbergcoeff(fdorder, r["max_relative_error"])
    sig = O.create_wavelet(nt, r["dt"], r["source_amp"], r["source_fc"], 0.0, 1)[None]
    src = dict(x=[r["source_x"] + ppad], z=[r["source_z"] + fpad], comp=[2], shot=[0])
    rec = dict(x=[x + ppad for x in r["receiver_x"]], z=[z + fpad for z in r["receiver_z"]],
               comp=[2] * 3, shot=[0] * 3)
    return O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, dt=r["dt"],
                     dx=r["dx"], dz=r["dz"], fsurf=r["free_surface"], hc=hc, lam=lam, mu=mu,
                     rho=rho, pml=pml, src=src, src_signal=sig, rec=rec)


def smooth_random(shape, seed, sigma=4.0):
    from scipy.ndimage import gaussian_filter
    g = np.random.default_rng(seed).standard_normal(shape)
    g = gaussian_filter(g, sigma)
    return g / (np.abs(g).max() + 1e-30)


def small_problem(nx=60, nz=40, npml=8, nt=300, fsurf=1, fdorder=2, seed=0, hetero=True,
                  fwi=None, nsrc=2, nrec=5, dx=5.0, dt=5e-4, fc=25.0, maxrelerr=1,
                  src_comp=2):
    """A small heterogeneous problem in the spirit of the reference INI files."""
    fpad = fdorder // 2
    ppad = npml + fpad
    nxt = nx + 2 * ppad
    nzt = nz + ppad + fpad if fsurf else nz + 2 * ppad
    vp = np.full((nzt, nxt), 2500.0)
    vs = np.full((nzt, nxt), 1443.0)
    rho = np.full((nzt, nxt), 2000.0)
    if hetero:
        p = smooth_random((nzt, nxt), seed)
        vp *= 1 + 0.1 * p
        vs *= 1 + 0.1 * p
        rho *= 1 + 0.05 * p
    mu = vs ** 2 * rho
    lam = vp ** 2 * rho - 2 * mu
    pml = O.cpml_pro(npml, dx, fc, 2.0, 1.0, 1e-3, 2750.0, dt)
    hc = O.holbergcoeff(fdorder, maxrelerr)
    zoff = fpad if fsurf else ppad
    xs = np.linspace(nx * 0.3, nx * 0.7, nsrc).astype(int) + ppad
    src = dict(x=xs, z=[3 + zoff] * nsrc, comp=[src_comp] * nsrc, shot=[0] * nsrc)
    sig = np.stack([O.create_wavelet(nt, dt, 1.0 + 0.5 * i, fc, 0.002 * i, 1) for i in range(nsrc)])
    xr = np.linspace(2, nx - 3, nrec).astype(int) + ppad
    rec = dict(x=xr, z=[1 + zoff] * nrec, comp=[2] * nrec, shot=[0] * nrec)
    return O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, dt=dt, dx=dx, dz=dx,
                     fsurf=fsurf, hc=hc, lam=lam, mu=mu, rho=rho, pml=pml, src=src,
                     src_signal=sig, rec=rec, fwi=fwi)


def full_window(pb, dt=1, dx=1, dz=1):
    """FWI window over the whole interior (clamped inside the grid)."""
    return dict(dt=dt, dx=dx, dz=dz, x1=pb.fpad, x2=pb.nxt - pb.fpad, z1=pb.fpad,
                z2=pb.nzt - pb.fpad)
