"""ctypes binding of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module; it is the checker, never the product path.
Every array is a C-contiguous numpy array in the layouts documented in
``psv_oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("PSV_ORACLE_LIB") or os.path.join(_HERE, "libpsv_oracle.so")
_OMP_PATH = os.path.join(_HERE, "libpsv_oracle_omp.so")
_lib = None
_threaded = False   # set_threaded(True): the row-parallel build (same results, all host cores)


class Cfg(C.Structure):
    _fields_ = [
        ("nt", C.c_int), ("nzt", C.c_int), ("nxt", C.c_int), ("fpad", C.c_int), ("ppad", C.c_int),
        ("dt", C.c_double), ("dx", C.c_double), ("dz", C.c_double),
        ("fsurf", C.c_int), ("fdorder", C.c_int), ("hc", C.c_double * 7), ("npml", C.c_int),
        ("fwinv", C.c_int), ("fwi_dt", C.c_int), ("fwi_dx", C.c_int), ("fwi_dz", C.c_int),
        ("fwi_x1", C.c_int), ("fwi_x2", C.c_int), ("fwi_z1", C.c_int), ("fwi_z2", C.c_int),
    ]


class Seis(C.Structure):
    _fields_ = [("n", C.c_int), ("x", C.POINTER(C.c_int)), ("z", C.POINTER(C.c_int)),
                ("comp", C.POINTER(C.c_int)), ("shot", C.POINTER(C.c_int))]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        build()
    h = C.CDLL(path)
    d = C.POINTER(C.c_double)
    h.psvo_calc_opt_step.restype = C.c_double
    h.psvo_calc_opt_step.argtypes = [d, d]
    return h


_omp = None


def set_threaded(on: bool):
    """Route the propagators through the OpenMP build (rows of each sweep in
    parallel; bit-identical results) or the single-threaded one."""
    global _threaded
    _threaded = bool(on)


def lib():
    global _lib, _omp
    if _threaded:
        if _omp is None:
            _omp = _load(_OMP_PATH)
        return _omp
    if _lib is None:
        _lib = _load(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def holbergcoeff(fdorder: int, maxrelerr: int = 1) -> np.ndarray:
    hc = np.zeros(7)
    lib().psvo_holbergcoeff(C.c_int(fdorder), C.c_int(maxrelerr), _p(hc))
    return hc


def cpml_pro(npml, dh, fpml, npower, k_max, rcoef, damp_v, dt) -> np.ndarray:
    """Returns the (6, 2(npml+1)) array a, b, K, a_half, b_half, K_half."""
    tp = 2 * (npml + 1)
    out = np.zeros((6, tp))
    lib().psvo_cpml_pro(C.c_int(npml), C.c_double(dh), C.c_double(fpml), C.c_double(npower),
                        C.c_double(k_max), C.c_double(rcoef), C.c_double(damp_v), C.c_double(dt),
                        *[_p(out[i]) for i in range(6)])
    return out


def velocity_to_lame(vp, vs, rho):
    lam, mu = C.c_double(), C.c_double()
    lib().psvo_velocity_to_lame(C.c_double(vp), C.c_double(vs), C.c_double(rho),
                                C.byref(lam), C.byref(mu))
    return lam.value, mu.value


def average_parameters(mu: np.ndarray, rho: np.ndarray):
    nzt, nxt = mu.shape
    mu = np.ascontiguousarray(mu, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    out = [np.zeros((nzt - 1, nxt - 1)) for _ in range(3)]
    lib().psvo_average_parameters(C.c_int(nzt - 1), C.c_int(nxt - 1), _p(mu), _p(rho),
                                  *[_p(o) for o in out])
    return tuple(out)  # mu_zx, rho_zp, rho_xp


def stagger_over_grid_modify(nzt, nxt, rho, lam, mu):
    out = [np.zeros((nzt, nxt)) for _ in range(3)]
    lib().psvo_stagger_over_grid_modify(C.c_int(nzt), C.c_int(nxt), C.c_double(rho),
                                        C.c_double(lam), C.c_double(mu), *[_p(o) for o in out])
    return tuple(out)  # rho, lam, mu


def create_wavelet(nt, dt, amp, fc, ts=0.0, shape=1) -> np.ndarray:
    s = np.zeros(nt)
    lib().psvo_create_wavelet(_p(s), C.c_int(nt), C.c_double(dt), C.c_double(amp),
                              C.c_double(fc), C.c_double(ts), C.c_int(shape))
    return s


def l2_adjoint_sources(dt, signal: np.ndarray, meas: np.ndarray):
    """Returns (misfit, residual, status); residual = signal - meas."""
    sig = np.ascontiguousarray(signal, np.float64).copy()
    meas = np.ascontiguousarray(meas, np.float64)
    m = C.c_double()
    st = lib().psvo_l2_adjoint_sources(C.c_int(sig.shape[0]), C.c_int(sig.shape[1]),
                                       C.c_double(dt), _p(sig), _p(meas), C.byref(m))
    return m.value, sig, st


def calc_opt_step(L2, sl) -> float:
    a = (C.c_double * 3)(*L2)
    b = (C.c_double * 3)(*sl)
    return lib().psvo_calc_opt_step(a, b)


def medium_average(lam, mu, rho):
    nzt, nxt = lam.shape
    out = [C.c_double() for _ in range(3)]
    lib().psvo_medium_average(_p(np.ascontiguousarray(lam)), _p(np.ascontiguousarray(mu)),
                              _p(np.ascontiguousarray(rho)), C.c_int(nzt), C.c_int(nxt),
                              *[C.byref(o) for o in out])
    return tuple(o.value for o in out)  # C_lam, C_mu, C_rho


def scale_gradients_energy(Ws, Wr, C_rho, C_lam, C_mu, nfz, nfx, z1, x1, fdz, fdx, grads):
    g = [np.ascontiguousarray(x, np.float64).copy() for x in grads]
    lib().psvo_scale_gradients_energy(_p(np.ascontiguousarray(Ws)), _p(np.ascontiguousarray(Wr)),
                                      C.c_int(Ws.shape[1]), C.c_double(C_rho), C.c_double(C_lam),
                                      C.c_double(C_mu), C.c_int(nfz), C.c_int(nfx), C.c_int(z1),
                                      C.c_int(x1), C.c_int(fdz), C.c_int(fdx), *[_p(x) for x in g])
    return g


def fwi_grad_update(mat, grad, step, nfz, nfx, z1, x1, fdz, fdx, override, mat_type):
    mat = np.ascontiguousarray(mat, np.float64).copy()
    new = mat.copy()
    grad = np.ascontiguousarray(grad, np.float64)
    lib().psvo_fwi_grad_update(_p(mat), _p(new), _p(grad), C.c_int(mat.shape[1]), C.c_double(step),
                               C.c_int(nfz), C.c_int(nfx), C.c_int(z1), C.c_int(x1), C.c_int(fdz),
                               C.c_int(fdx), C.c_int(int(override)), C.c_int(mat_type),
                               C.c_int(grad.shape[1]))
    return mat, new


# --------------------------------------------------------------------------
class Problem:
    """Everything the propagators need, in oracle layout.

    grid: nt, nzt, nxt, fpad, ppad, dt, dx, dz, fsurf, fdorder, hc, npml
    model: lam, mu [nzt, nxt]; mu_zx, rho_zp, rho_xp [nzt-1, nxt-1]
    pml: (6, 2(npml+1)); src/rec geometry as int arrays (padded indices).
    """

    def __init__(self, *, nt, nzt, nxt, fdorder, npml, dt, dx, dz, fsurf, hc,
                 lam, mu, rho=None, mu_zx=None, rho_zp=None, rho_xp=None, pml=None,
                 src=None, src_signal=None, rec=None,
                 fwi=None):
        self.nt, self.nzt, self.nxt = nt, nzt, nxt
        self.fdorder, self.fpad = fdorder, fdorder // 2
        self.npml, self.ppad = npml, npml + fdorder // 2
        self.dt, self.dx, self.dz, self.fsurf = dt, dx, dz, int(fsurf)
        self.hc = np.asarray(hc, np.float64)
        self.lam = np.ascontiguousarray(lam, np.float64)
        self.mu = np.ascontiguousarray(mu, np.float64)
        self.rho = None if rho is None else np.ascontiguousarray(rho, np.float64)
        if mu_zx is None:
            mu_zx, rho_zp, rho_xp = average_parameters(self.mu, self.rho)
        self.mu_zx = np.ascontiguousarray(mu_zx, np.float64)
        self.rho_zp = np.ascontiguousarray(rho_zp, np.float64)
        self.rho_xp = np.ascontiguousarray(rho_xp, np.float64)
        tp = 2 * (npml + 1)
        self.pml = np.ascontiguousarray(pml if pml is not None else
                                        np.vstack([np.zeros(tp), np.ones(tp), np.ones(tp),
                                                   np.zeros(tp), np.ones(tp), np.ones(tp)]),
                                        np.float64)
        self.src = src            # dict x, z, comp, shot (int arrays)
        self.src_signal = None if src_signal is None else np.ascontiguousarray(src_signal, np.float64)
        self.rec = rec
        # fwi: dict dt, dx, dz, x1, x2, z1, z2 (padded indices)
        self.fwi = fwi

    def cfg(self):
        c = Cfg()
        c.nt, c.nzt, c.nxt, c.fpad, c.ppad = self.nt, self.nzt, self.nxt, self.fpad, self.ppad
        c.dt, c.dx, c.dz = self.dt, self.dx, self.dz
        c.fsurf, c.fdorder, c.npml = self.fsurf, self.fdorder, self.npml
        for i in range(7):
            c.hc[i] = float(self.hc[i])
        if self.fwi:
            c.fwinv = 1
            f = self.fwi
            c.fwi_dt, c.fwi_dx, c.fwi_dz = f["dt"], f["dx"], f["dz"]
            c.fwi_x1, c.fwi_x2, c.fwi_z1, c.fwi_z2 = f["x1"], f["x2"], f["z1"], f["z2"]
        else:
            c.fwinv, c.fwi_dt, c.fwi_dx, c.fwi_dz = 0, 1, 1, 1
        return c

    def fwi_dims(self):
        f = self.fwi
        nft = 1 + (self.nt - 1) // f["dt"]
        nfz = 1 + (f["z2"] - f["z1"]) // f["dz"]
        nfx = 1 + (f["x2"] - f["x1"]) // f["dx"]
        return nft, nfz, nfx

    @staticmethod
    def _seis(g, keep):
        if g is None:
            return None
        arrs = [np.ascontiguousarray(g[k], np.int32) for k in ("x", "z", "comp", "shot")]
        keep.extend(arrs)
        ip = C.POINTER(C.c_int)
        return Seis(len(arrs[0]), *[a.ctypes.data_as(ip) for a in arrs])


def _dtype(prec):
    return np.float64 if prec == "f64" else np.float32


def forward(pb: Problem, ishot=0, prec="f64"):
    """Runs the oracle forward propagator.  Returns dict with the final
    wavefield, energy We, rec (nrec, nt), store (5, nft, nfz, nfx) and mem."""
    dt_ = _dtype(prec)
    fn = lib().psvo_forward_d if prec == "f64" else lib().psvo_forward_f
    n = (pb.nzt, pb.nxt)
    f = {k: np.zeros(n, dt_) for k in ("vx", "vz", "sxx", "szx", "szz", "We")}
    tp = 2 * (pb.npml + 1)
    mem = np.zeros(4 * pb.nzt * tp + 4 * tp * pb.nxt, dt_)
    keep = []
    src = Problem._seis(pb.src, keep)
    rec = Problem._seis(pb.rec, keep)
    nrec = 0 if pb.rec is None else len(pb.rec["x"])
    rec_sig = np.zeros((max(nrec, 1), pb.nt), dt_)
    if pb.fwi:
        nft, nfz, nfx = pb.fwi_dims()
        store = np.zeros((5, nft, nfz, nfx), dt_)
    else:
        store = np.zeros(1, dt_)
    cast = lambda a: np.ascontiguousarray(a, dt_)
    mats = [cast(pb.lam), cast(pb.mu), cast(pb.mu_zx), cast(pb.rho_zp), cast(pb.rho_xp)]
    pml = cast(pb.pml)
    ssig = cast(pb.src_signal) if pb.src_signal is not None else np.zeros(1, dt_)
    cfg = pb.cfg()
    st = fn(C.byref(cfg), C.c_int(ishot), *[_p(f[k]) for k in ("vx", "vz", "sxx", "szx", "szz", "We")],
            *[_p(m) for m in mats], _p(pml), _p(mem),
            C.byref(src) if src else None, _p(ssig),
            C.byref(rec) if rec else None, _p(rec_sig), _p(store))
    if st != 0:
        raise ValueError(f"oracle forward rejected its arguments (status {st})")
    f["rec"] = rec_sig[:nrec]
    f["store"] = store
    f["mem"] = mem
    return f


def adjoint(pb: Problem, adj_src: dict, adj_signal: np.ndarray, store: np.ndarray,
            ishot=0, prec="f64"):
    """Runs the oracle adjoint propagator with adj_src as sources and
    adj_signal (nsrc, nt) as their signals.  Returns dict with the final
    wavefield, We and grad_lam/grad_mu/grad_rho (nfz, nfx)."""
    dt_ = _dtype(prec)
    fn = lib().psvo_adjoint_d if prec == "f64" else lib().psvo_adjoint_f
    n = (pb.nzt, pb.nxt)
    f = {k: np.zeros(n, dt_) for k in ("vx", "vz", "sxx", "szx", "szz", "We")}
    tp = 2 * (pb.npml + 1)
    mem = np.zeros(4 * pb.nzt * tp + 4 * tp * pb.nxt, dt_)
    keep = []
    src = Problem._seis(adj_src, keep)
    nft, nfz, nfx = pb.fwi_dims()
    g = [np.zeros((nfz, nfx), dt_) for _ in range(3)]
    cast = lambda a: np.ascontiguousarray(a, dt_)
    mats = [cast(pb.lam), cast(pb.mu), cast(pb.mu_zx), cast(pb.rho_zp), cast(pb.rho_xp)]
    sig = cast(adj_signal)
    st_arr = cast(store)
    pml = cast(pb.pml)
    cfg = pb.cfg()
    st = fn(C.byref(cfg), C.c_int(ishot), *[_p(f[k]) for k in ("vx", "vz", "sxx", "szx", "szz", "We")],
            *[_p(m) for m in mats], _p(pml), _p(mem),
            C.byref(src) if src else None, _p(sig), _p(st_arr), *[_p(x) for x in g])
    if st != 0:
        raise ValueError(f"oracle adjoint rejected its arguments (status {st})")
    f["grad_lam"], f["grad_mu"], f["grad_rho"] = g
    f["mem"] = mem
    return f


def discrete_gradient(pb: Problem, res: np.ndarray, ishot=0, prec="f64"):
    """Exact gradient of the discrete misfit sum 0.5 dt (d - d_obs)^2 of shot
    `ishot` with respect to the coefficient arrays (discrete adjoint, an
    extension; psv_oracle_discrete.inc).  res [nrec, nt] = d - d_obs for the
    receivers of pb.rec.  Returns dict g_lam, g_mu [nzt, nxt], g_mu_zx,
    g_rho_zp, g_rho_xp [nzt-1, nxt-1] and adj (5, nzt, nxt): the adjoint of
    the initial vx, vz, sxx, szx, szz."""
    dt_ = _dtype(prec)
    fn = lib().psvo_discrete_gradient_d if prec == "f64" else lib().psvo_discrete_gradient_f
    n, na = (pb.nzt, pb.nxt), (pb.nzt - 1, pb.nxt - 1)
    g = {k: np.zeros(n, dt_) for k in ("g_lam", "g_mu")}
    g.update({k: np.zeros(na, dt_) for k in ("g_mu_zx", "g_rho_zp", "g_rho_xp")})
    adj = np.zeros((5,) + n, dt_)
    keep = []
    src = Problem._seis(pb.src, keep)
    rec = Problem._seis(pb.rec, keep)
    cast = lambda a: np.ascontiguousarray(a, dt_)
    mats = [cast(pb.lam), cast(pb.mu), cast(pb.mu_zx), cast(pb.rho_zp), cast(pb.rho_xp)]
    ssig = cast(pb.src_signal) if pb.src_signal is not None else np.zeros(1, dt_)
    r = cast(res) if np.size(res) else np.zeros(1, dt_)
    cfg = pb.cfg()
    st = fn(C.byref(cfg), C.c_int(ishot), *[_p(m) for m in mats], _p(cast(pb.pml)),
            C.byref(src) if src else None, _p(ssig), C.byref(rec) if rec else None, _p(r),
            *[_p(g[k]) for k in ("g_lam", "g_mu", "g_mu_zx", "g_rho_zp", "g_rho_xp")], _p(adj))
    if st != 0:
        raise ValueError(f"oracle discrete gradient rejected its arguments (status {st})")
    g["adj"] = adj
    return g


def chain_gradient(mu, mu_zx, rho_zp, rho_xp, g):
    """lam / mu / rho gradients from the coefficient gradients g (as returned
    by discrete_gradient), for averages made by average_parameters."""
    nzt, nxt = mu.shape
    out = [np.zeros((nzt, nxt)) for _ in range(3)]
    a = lambda x: np.ascontiguousarray(x, np.float64)
    lib().psvo_chain_gradient(C.c_int(nzt), C.c_int(nxt), _p(a(mu)), _p(a(mu_zx)), _p(a(rho_zp)),
                              _p(a(rho_xp)), *[_p(a(g[k])) for k in ("g_lam", "g_mu", "g_mu_zx",
                                                                      "g_rho_zp", "g_rho_xp")],
                              *[_p(o) for o in out])
    return out
