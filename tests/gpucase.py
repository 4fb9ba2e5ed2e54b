"""Builds a device plan (cuda_fwi_amd.psvfwi) from an oracle Problem."""
import numpy as np

from cuda_fwi_amd import psvfwi as P


def shots_of(pb, ishot=0):
    """Sources/receivers of shot `ishot` in the reference's shot_to_fire sense."""
    s = pb.src
    keep = [i for i in range(len(s["x"])) if s["shot"][i] == ishot]
    src = {k: [int(s[k][i]) for i in keep] for k in ("x", "z", "comp")}
    sig = pb.src_signal[keep] if keep else np.zeros((0, pb.nt))
    rec = None
    if pb.rec is not None:
        r = pb.rec
        rk = [i for i in range(len(r["x"])) if r["shot"][i] == ishot]
        rec = {k: [int(r[k][i]) for i in rk] for k in ("x", "z", "comp")}
    return dict(nt=pb.nt, src=src, src_signal=sig, rec=rec)


def plan_of(pb, precision="fp64", tz=None, shots=None):
    p = P.Plan(nzt=pb.nzt, nxt=pb.nxt, fdorder=pb.fdorder, npml=pb.npml, fsurf=pb.fsurf,
               dt=pb.dt, dx=pb.dx, dz=pb.dz, hc=pb.hc, precision=precision)
    if tz:
        p.set_tile(tz)
    p.set_pml(pb.pml)
    p.set_model(pb.lam, pb.mu, pb.mu_zx, pb.rho_zp, pb.rho_xp)
    p.set_window(pb.fwi)
    p.set_shots(shots if shots is not None else [shots_of(pb)])
    return p


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
