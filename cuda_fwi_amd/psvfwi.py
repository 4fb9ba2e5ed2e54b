"""ctypes binding of the C ABI in include/psvfwi.h (libpsvfwi.so).

This is the thin Python layer the tests, the benchmark and the autograd
binding use; the product is the native library.  Loading fails loudly when
the library has not been built (``__graft_entry__.build()`` or
``make -C cuda_fwi_amd/csrc``): there is no CPU fallback.

Array conventions follow the reference (real_sim = double, row-major):
lam, mu [nzt, nxt]; mu_zx, rho_zp, rho_xp [nzt-1, nxt-1]; CPML profiles
[6, 2(npml+1)]; traces [nrec, nt]; forward store [5, nft, nfz, nfx];
gradients [nfz, nfx].
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpsvfwi.so")

FP32, FP64, FP32X2, FP32_FMA = 0, 1, 2, 3
PRECISIONS = {"fp32": FP32, "fp64": FP64, "fp32x2": FP32X2, "fp32fma": FP32_FMA}
ENERGY, GRADIENT = 1, 2
STORE_FULL, STORE_CHECKPOINT = 0, 1
FIELDS = {"vx": 0, "vz": 1, "sxx": 2, "szx": 3, "szz": 4, "We": 5, "ux": 6, "uz": 7,
          "mem_vx_x": 8, "mem_vz_x": 9, "mem_sxx_x": 10, "mem_szx_x": 11,
          "mem_vx_z": 12, "mem_vz_z": 13, "mem_szz_z": 14, "mem_szx_z": 15}
EXPORTS = ("psvfwi_last_error", "psvfwi_plan_create", "psvfwi_plan_destroy", "psvfwi_set_pml",
           "psvfwi_set_model", "psvfwi_set_window", "psvfwi_set_store_mode", "psvfwi_set_shots",
           "psvfwi_set_tile", "psvfwi_autotune", "psvfwi_forward", "psvfwi_set_observed", "psvfwi_misfit",
           "psvfwi_set_adjoint_sources",
           "psvfwi_set_store", "psvfwi_adjoint", "psvfwi_get_records", "psvfwi_get_field",
           "psvfwi_get_store", "psvfwi_get_gradient", "psvfwi_last_timing", "psvfwi_kernel_timing",
           "psvfwi_parabola_step", "psvfwi_holberg", "psvfwi_cpml_profiles",
           "psvfwi_average_parameters", "psvfwi_ricker", "psvfwi_set_snapshots",
           "psvfwi_snapshot_count", "psvfwi_get_snapshot", "psvfwi_set_graph_steps",
           "psvfwi_set_gradient_mode", "psvfwi_get_coefficient_gradient", "psvfwi_get_model_gradient",
           "psvfwi_chain_gradient", "psvfwi_random_boundary", "psvfwi_set_shot_order",
           "psvfwi_get_shot_order", "psvfwi_set_launch_group", "psvfwi_get_launch_group",
           "psvfwi_misfit_noalloc")


class PsvfwiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"psvfwi status {status}: {msg}")
        self.status = status


class GridDesc(C.Structure):
    _fields_ = [("nzt", C.c_int), ("nxt", C.c_int), ("fdorder", C.c_int), ("npml", C.c_int),
                ("fsurf", C.c_int), ("dt", C.c_double), ("dx", C.c_double), ("dz", C.c_double),
                ("hc", C.c_double * 7)]


class WindowDesc(C.Structure):
    _fields_ = [("dt", C.c_int), ("dx", C.c_int), ("dz", C.c_int), ("x1", C.c_int),
                ("x2", C.c_int), ("z1", C.c_int), ("z2", C.c_int)]


_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class ShotDesc(C.Structure):
    _fields_ = [("nt", C.c_int), ("nsrc", C.c_int), ("src_x", _ip), ("src_z", _ip),
                ("src_comp", _ip), ("src_signal", _dp), ("nrec", C.c_int), ("rec_x", _ip),
                ("rec_z", _ip), ("rec_comp", _ip)]


_lib = None


def lib():
    """The loaded native library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _lib.psvfwi_last_error.restype = C.c_char_p
        _lib.psvfwi_plan_create.argtypes = [C.POINTER(GridDesc), C.c_int, C.POINTER(C.c_void_p)]
        for name in EXPORTS[2:]:
            getattr(_lib, name).restype = C.c_int
    return _lib


def _check(st):
    if st != 0:
        raise PsvfwiError(st, lib().psvfwi_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape, what):
    """a as a C-contiguous FP64 array of exactly this shape: the library copies
    shape-product elements from the pointer, so a mismatch is refused here."""
    a = np.ascontiguousarray(a, np.float64)
    if a.shape != tuple(shape):
        raise ValueError(f"{what}: expected shape {tuple(shape)}, got {a.shape}")
    return a


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Plan:
    """One device plan: a grid, a model, a batch of shots."""

    def __init__(self, *, nzt, nxt, fdorder, npml, fsurf, dt, dx, dz, hc, precision="fp32"):
        g = GridDesc(nzt, nxt, fdorder, npml, int(fsurf), dt, dx, dz)
        for i in range(7):
            g.hc[i] = float(hc[i])
        self.grid = dict(nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, fsurf=int(fsurf),
                         dt=dt, dx=dx, dz=dz)
        self._h = C.c_void_p()
        if precision not in PRECISIONS and precision not in PRECISIONS.values():
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.prec = PRECISIONS.get(precision, precision)
        _check(lib().psvfwi_plan_create(C.byref(g), self.prec, C.byref(self._h)))
        self._shots = []
        self.window = None

    def close(self):
        if self._h:
            lib().psvfwi_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pml(self, profiles):
        prof = _f64(profiles, (6, 2 * (self.grid["npml"] + 1)), "set_pml")
        _check(lib().psvfwi_set_pml(self._h, _dptr(prof)))

    def set_model(self, lam, mu, mu_zx, rho_zp, rho_xp):
        nzt, nxt = self.grid["nzt"], self.grid["nxt"]
        arrs = [_f64(a, (nzt, nxt) if k < 2 else (nzt - 1, nxt - 1), "set_model")
                for k, a in enumerate((lam, mu, mu_zx, rho_zp, rho_xp))]
        _check(lib().psvfwi_set_model(self._h, *[_dptr(a) for a in arrs]))

    def set_window(self, win):
        self.window = win
        if win is None:
            _check(lib().psvfwi_set_window(self._h, None))
            return
        w = WindowDesc(win["dt"], win["dx"], win["dz"], win["x1"], win["x2"], win["z1"], win["z2"])
        _check(lib().psvfwi_set_window(self._h, C.byref(w)))

    def set_store_mode(self, mode, interval=0):
        m = {"full": STORE_FULL, "checkpoint": STORE_CHECKPOINT, "random_boundary": 2}.get(mode, mode)
        _check(lib().psvfwi_set_store_mode(self._h, m, int(interval)))

    def set_tile(self, tz):
        _check(lib().psvfwi_set_tile(self._h, int(tz)))

    def autotune(self, steps=50):
        best = C.c_int()
        _check(lib().psvfwi_autotune(self._h, int(steps), C.byref(best)))
        return best.value

    def set_shots(self, shots):
        """shots: list of dicts nt, src (x, z, comp), src_signal [nsrc, nt], rec (x, z, comp)."""
        keep, descs = [], []
        for s in shots:
            nt = int(s["nt"])
            src = s.get("src") or dict(x=[], z=[], comp=[])
            rec = s.get("rec") or dict(x=[], z=[], comp=[])
            sx, sz, sc = (np.ascontiguousarray(src[k], np.int32) for k in ("x", "z", "comp"))
            rx, rz, rc = (np.ascontiguousarray(rec[k], np.int32) for k in ("x", "z", "comp"))
            sig = np.ascontiguousarray(s.get("src_signal", np.zeros((len(sx), nt))), np.float64)
            if sig.shape != (len(sx), nt):
                raise ValueError("src_signal must be [nsrc, nt]")
            keep += [sx, sz, sc, rx, rz, rc, sig]
            descs.append(ShotDesc(nt, len(sx), _iptr(sx), _iptr(sz), _iptr(sc), _dptr(sig),
                                  len(rx), _iptr(rx), _iptr(rz), _iptr(rc)))
        arr = (ShotDesc * len(descs))(*descs)
        _check(lib().psvfwi_set_shots(self._h, len(descs), arr))
        self._shots = [dict(nt=int(s["nt"]), nrec=len((s.get("rec") or {"x": []})["x"]))
                       for s in shots]

    def set_shot_order(self, band):
        """0: shot-major workgroups; k > 0: bands of k tile rows, shots together within a band."""
        _check(lib().psvfwi_set_shot_order(self._h, int(band)))

    def shot_order(self):
        b = C.c_int()
        _check(lib().psvfwi_get_shot_order(self._h, C.byref(b)))
        return b.value

    def set_launch_group(self, shots):
        """Run the batch `shots` slots at a time, every step of a group before the next (0: all)."""
        _check(lib().psvfwi_set_launch_group(self._h, int(shots)))

    def launch_group(self):
        k = C.c_int()
        _check(lib().psvfwi_get_launch_group(self._h, C.byref(k)))
        return k.value

    def set_graph_steps(self, steps):
        """Replay time steps as captured HIP graphs of `steps` launches (0: off)."""
        _check(lib().psvfwi_set_graph_steps(self._h, int(steps)))

    def set_snapshots(self, interval, field="szz"):
        _check(lib().psvfwi_set_snapshots(self._h, int(interval), FIELDS[field] if isinstance(field, str) else int(field)))

    def snapshots(self, shot=0):
        n = C.c_int()
        _check(lib().psvfwi_snapshot_count(self._h, int(shot), C.byref(n)))
        out = np.zeros((n.value, self.grid["nzt"], self.grid["nxt"]))
        for k in range(n.value):
            _check(lib().psvfwi_get_snapshot(self._h, int(shot), k, _dptr(out[k])))
        return out

    def forward(self, energy=True, gradient=False):
        _check(lib().psvfwi_forward(self._h, (ENERGY if energy else 0) | (GRADIENT if gradient else 0)))

    def _concat(self, observed):
        if len(observed) != len(self._shots):
            raise ValueError(f"observed: expected {len(self._shots)} shots, got {len(observed)}")
        parts = [_f64(o, (s["nrec"], s["nt"]), "observed").ravel() for o, s in zip(observed, self._shots)]
        return np.concatenate(parts) if parts and sum(p.size for p in parts) else np.zeros(1)

    def set_observed(self, observed):
        """observed: list (one [nrec, nt] array per shot), kept on the device."""
        obs = self._concat(observed)
        _check(lib().psvfwi_set_observed(self._h, _dptr(obs)))

    def misfit(self, observed=None):
        """Per-shot misfits; observed as in set_observed, or None for the resident traces."""
        out = np.zeros(len(self._shots))
        if observed is None:
            _check(lib().psvfwi_misfit(self._h, None, _dptr(out)))
        else:
            obs = self._concat(observed)
            _check(lib().psvfwi_misfit(self._h, _dptr(obs), _dptr(out)))
        return out

    def misfit_noalloc(self):
        """misfit() on the resident observed traces with no device allocation (the
        library's static scratch; calls that overlap on one GPU race)."""
        out = np.zeros(len(self._shots))
        _check(lib().psvfwi_misfit_noalloc(self._h, _dptr(out)))
        return out

    def set_adjoint_sources(self, shot, signal):
        s = self._shots[shot]
        sig = _f64(signal, (s["nrec"], s["nt"]), "set_adjoint_sources")
        _check(lib().psvfwi_set_adjoint_sources(self._h, int(shot), _dptr(sig)))

    def set_store(self, shot, store):
        st = _f64(store, (5,) + self._fwi_dims(shot), "set_store")
        _check(lib().psvfwi_set_store(self._h, int(shot), _dptr(st)))

    def adjoint(self, energy=True, gradient=True):
        _check(lib().psvfwi_adjoint(self._h, (ENERGY if energy else 0) | (GRADIENT if gradient else 0)))

    def records(self, shot=0):
        s = self._shots[shot]
        out = np.zeros((s["nrec"], s["nt"]))
        _check(lib().psvfwi_get_records(self._h, int(shot), _dptr(out)))
        return out

    def field(self, name, shot=0):
        k = FIELDS[name]
        nzt, nxt, npml = self.grid["nzt"], self.grid["nxt"], self.grid["npml"]
        tp = 2 * (npml + 1)
        if k <= 7:
            shape = (nzt, nxt)
        elif k <= 11:
            shape = (nzt, tp)
        else:
            shape = (tp, nxt)
        out = np.zeros(shape)
        _check(lib().psvfwi_get_field(self._h, int(shot), k, _dptr(out)))
        return out

    def _fwi_dims(self, shot):
        w = self.window
        if w is None:
            raise PsvfwiError(-5, "no FWI window set")
        nft = 1 + (self._shots[shot]["nt"] - 1) // w["dt"]
        nfz = 1 + (w["z2"] - w["z1"]) // w["dz"]
        nfx = 1 + (w["x2"] - w["x1"]) // w["dx"]
        return nft, nfz, nfx

    def store(self, shot=0):
        nft, nfz, nfx = self._fwi_dims(shot)
        out = np.zeros((5, nft, nfz, nfx))
        _check(lib().psvfwi_get_store(self._h, int(shot), _dptr(out)))
        return out

    def gradient(self, shot=-1):
        _, nfz, nfx = self._fwi_dims(max(shot, 0))
        g = [np.zeros((nfz, nfx)) for _ in range(3)]
        _check(lib().psvfwi_get_gradient(self._h, int(shot), *[_dptr(a) for a in g]))
        return g  # lam, mu, rho

    def set_gradient_mode(self, mode, segment=64):
        """'reference' (the correlation gradient of adjoint_kernel_PSV) or
        'discrete' (the exact gradient of the misfit of this scheme, with
        checkpoints every `segment` steps; segment 1 keeps every state of the
        forward and the adjoint replays nothing)."""
        m = {"reference": 0, "discrete": 1}[mode] if isinstance(mode, str) else int(mode)
        _check(lib().psvfwi_set_gradient_mode(self._h, m, int(segment)))

    def coefficient_gradient(self, shot=-1):
        """Discrete mode: dJ/d lam, mu [nzt, nxt] and dJ/d mu_zx, rho_zp,
        rho_xp [nzt-1, nxt-1] (dict)."""
        nzt, nxt = self.grid["nzt"], self.grid["nxt"]
        g = dict(g_lam=np.zeros((nzt, nxt)), g_mu=np.zeros((nzt, nxt)),
                 g_mu_zx=np.zeros((nzt - 1, nxt - 1)), g_rho_zp=np.zeros((nzt - 1, nxt - 1)),
                 g_rho_xp=np.zeros((nzt - 1, nxt - 1)))
        _check(lib().psvfwi_get_coefficient_gradient(self._h, int(shot), *[_dptr(a) for a in g.values()]))
        return g

    def model_gradient(self, shot=-1):
        """Discrete mode: dJ/d lam, dJ/d mu, dJ/d rho [nzt, nxt]."""
        nzt, nxt = self.grid["nzt"], self.grid["nxt"]
        g = [np.zeros((nzt, nxt)) for _ in range(3)]
        _check(lib().psvfwi_get_model_gradient(self._h, int(shot), *[_dptr(a) for a in g]))
        return g

    def timing(self):
        ms, n = C.c_double(), C.c_long()
        _check(lib().psvfwi_last_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_timing(self, kind):
        """Mean us of the stress and velocity kernels sampled in the last run:
        kind 'forward', 'adjoint' or 'replay'."""
        k = {"forward": 0, "adjoint": 1, "replay": 2}[kind]
        a, b, n = C.c_double(), C.c_double(), C.c_int()
        _check(lib().psvfwi_kernel_timing(self._h, k, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value


def parabola_step(L2, steps):
    a = (C.c_double * 3)(*L2)
    b = (C.c_double * 3)(*steps)
    out = C.c_double()
    _check(lib().psvfwi_parabola_step(a, b, C.byref(out)))
    return out.value


# ---- host-side setup functions of the library (FP64) ----
def holberg(fdorder, max_relative_error=1):
    hc = np.zeros(7)
    _check(lib().psvfwi_holberg(int(fdorder), int(max_relative_error), _dptr(hc)))
    return hc


def cpml_profiles(npml, dh, fpml, npower, k_max, rcoef, damp_v, dt):
    out = np.zeros((6, 2 * (npml + 1)))
    _check(lib().psvfwi_cpml_profiles(int(npml), C.c_double(dh), C.c_double(fpml),
                                      C.c_double(npower), C.c_double(k_max), C.c_double(rcoef),
                                      C.c_double(damp_v), C.c_double(dt), _dptr(out)))
    return out


def average_parameters(mu, rho):
    mu = np.ascontiguousarray(mu, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    nzt, nxt = mu.shape
    out = [np.zeros((nzt - 1, nxt - 1)) for _ in range(3)]
    _check(lib().psvfwi_average_parameters(int(nzt), int(nxt), _dptr(mu), _dptr(rho),
                                           *[_dptr(o) for o in out]))
    return tuple(out)  # mu_zx, rho_zp, rho_xp


def chain_gradient(mu, mu_zx, rho_zp, rho_xp, g):
    """lam / mu / rho gradients from coefficient gradients g (dict as
    returned by Plan.coefficient_gradient), on the host."""
    a = lambda x: np.ascontiguousarray(x, np.float64)
    mu = a(mu)
    nzt, nxt = mu.shape
    out = [np.zeros((nzt, nxt)) for _ in range(3)]
    arrs = [mu, a(mu_zx), a(rho_zp), a(rho_xp)] + [a(g[k]) for k in ("g_lam", "g_mu", "g_mu_zx",
                                                                       "g_rho_zp", "g_rho_xp")]
    _check(lib().psvfwi_chain_gradient(int(nzt), int(nxt), *[_dptr(x) for x in arrs],
                                       *[_dptr(o) for o in out]))
    return out


def random_boundary(lam, mu, rho, width, fsurf, seed=1, max_increase=0.5):
    """Copies of lam, mu, rho with a random boundary layer (psvfwi_random_boundary)."""
    lam, mu, rho = (np.array(a, np.float64, copy=True, order="C") for a in (lam, mu, rho))
    nzt, nxt = lam.shape
    _check(lib().psvfwi_random_boundary(int(nzt), int(nxt), int(width), int(fsurf), C.c_ulonglong(seed),
                                        C.c_double(max_increase), _dptr(lam), _dptr(mu), _dptr(rho)))
    return lam, mu, rho


def ricker(nt, dt, amp, fc, ts=0.0):
    s = np.zeros(int(nt))
    _check(lib().psvfwi_ricker(int(nt), C.c_double(dt), C.c_double(amp), C.c_double(fc),
                               C.c_double(ts), _dptr(s)))
    return s
