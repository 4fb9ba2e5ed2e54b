"""Synthetic models and acquisition geometries of the benchmark workloads.

All models are seeded and built from closed-form layers, faults and smoothed
Gaussian noise (no external data).  Velocities are on the unpadded nz x nx
grid; ``Workload.padded()`` extends them into the CPML / free-surface padding
by edge replication and converts to Lame parameters (mu = vs^2 rho,
lam = vp^2 rho - 2 mu, as FDMedium::velocity_to_lame).

  c1  201 x 201, dx 5 m, order 4, npml 20, free surface, homogeneous
      vp 2500 / vs 1443 / rho 2000 with a +10 % Gaussian anomaly (20 m) at the
      centre, explosive 15 Hz Ricker at (100, 2), 101 receivers, dt 0.5 ms,
      nt 2000
  c2  1000 x 500, dx 5 m, order 4, npml 40, 8 flat layers vp 1800-4500,
      vs = vp/1.73, Gardner density, 3 % Gaussian perturbation (50 m),
      12 Hz Ricker at (500, 2), 400 receivers, dt 0.4 ms, nt 5000
  c3  64 x 64, nt 600, 1 source, 32 receivers, random smooth model
      (differentiation check); c3mem 2000 x 1000, nt 8000, 8 flat layers,
      start = layers smoothed over 50 m (memory test)
  c4  1500 x 600, 64 shots: nt U{3000..8000}, receivers U{50..400},
      source x U[100, 1400], layered with a 60 degree dipping fault
  c5  2400 x 800, 128 shots x 600 receivers, nt 6000, 8 Hz, layers + two
      faults + a vp 4500 irregular body; start = truth smoothed (150 m)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter


@dataclass
class Workload:
    name: str
    nx: int
    nz: int
    dx: float
    dt: float
    nt: int
    npml: int
    fdorder: int
    fsurf: int
    fc: float
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    shots: list = field(default_factory=list)   # dicts: nt, src (x, z, comp), rec (x, z, comp), amp
    max_relative_error: int = 1

    @property
    def fpad(self):
        return self.fdorder // 2

    @property
    def ppad(self):
        return self.npml + self.fpad

    @property
    def nxt(self):
        return self.nx + 2 * self.ppad

    @property
    def nzt(self):
        return self.nz + (self.ppad + self.fpad if self.fsurf else 2 * self.ppad)

    @property
    def zoff(self):
        return self.fpad if self.fsurf else self.ppad

    def interior_cells(self):
        return (self.nzt - 2 * self.fpad) * (self.nxt - 2 * self.fpad)

    def pad(self, a):
        top = self.zoff
        bottom = self.nzt - self.nz - top
        return np.pad(a, ((top, bottom), (self.ppad, self.ppad)), mode="edge")

    def padded(self, vp=None, vs=None, rho=None):
        """lam, mu, rho on the padded [nzt, nxt] grid."""
        vp = self.vp if vp is None else vp
        vs = self.vs if vs is None else vs
        rho = self.rho if rho is None else rho
        vp, vs, rho = self.pad(vp), self.pad(vs), self.pad(rho)
        mu = vs ** 2 * rho
        lam = vp ** 2 * rho - 2.0 * mu
        return lam, mu, rho

    def padded_shot(self, s, signal):
        """A shot in padded indices for cuda_fwi_amd.psvfwi.Plan.set_shots."""
        src = dict(x=[x + self.ppad for x in s["src"]["x"]], z=[z + self.zoff for z in s["src"]["z"]],
                   comp=list(s["src"]["comp"]))
        rec = dict(x=[x + self.ppad for x in s["rec"]["x"]], z=[z + self.zoff for z in s["rec"]["z"]],
                   comp=list(s["rec"]["comp"]))
        return dict(nt=s["nt"], src=src, src_signal=signal, rec=rec)

    def model_window(self, dt=1):
        """FWI window over the unpadded model (padded indices)."""
        return dict(dt=dt, dx=1, dz=1, x1=self.ppad, x2=self.ppad + self.nx, z1=self.zoff,
                    z2=self.zoff + self.nz)


def _gardner(vp):
    return 310.0 * vp ** 0.25


def _noise(shape, seed, sigma_cells, rel):
    g = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma_cells)
    return g * (rel / g.std())


def _line(nx, n, z=0, comp=2):
    x = np.linspace(0, nx - 1, n).round().astype(int)
    return dict(x=list(map(int, x)), z=[z] * n, comp=[comp] * n)


def c1():
    nx = nz = 201
    dx = 5.0
    zz, xx = np.mgrid[0:nz, 0:nx] * dx
    bump = 1.0 + 0.1 * np.exp(-((zz - 100 * dx) ** 2 + (xx - 100 * dx) ** 2) / 20.0 ** 2)
    vp, vs, rho = 2500.0 * bump, 1443.0 * bump, 2000.0 * bump
    shots = [dict(nt=2000, src=dict(x=[100], z=[2], comp=[4]), rec=_line(nx, 101), amp=1.0)]
    w = Workload("c1", nx, nz, dx, 5e-4, 2000, 20, 4, 1, 15.0, vp, vs, rho, shots)
    w.background = (np.full_like(vp, 2500.0), np.full_like(vs, 1443.0), np.full_like(rho, 2000.0))
    return w


def c2_layers(nx=1000, nz=500):
    vps = np.linspace(1800.0, 4500.0, 8)
    layer = np.minimum(np.arange(nz) * 8 // nz, 7)
    return np.repeat(vps[layer][:, None], nx, axis=1)


def c2():
    nx, nz, dx = 1000, 500, 5.0
    base = c2_layers(nx, nz)
    vp = base * (1.0 + _noise((nz, nx), 2, 10.0, 0.03))
    vs = vp / 1.73
    shots = [dict(nt=5000, src=dict(x=[500], z=[2], comp=[2]), rec=_line(nx, 400), amp=1.0)]
    w = Workload("c2", nx, nz, dx, 4e-4, 5000, 40, 4, 0, 12.0, vp, vs, _gardner(vp), shots)
    w.background = (base, base / 1.73, _gardner(base))
    return w


def c3(nx=64, nz=64, nt=600, npml=10, seed=3):
    dx = 5.0
    g = gaussian_filter(np.random.default_rng(seed).standard_normal((nz, nx)), 6.0)
    g = (g - g.min()) / (g.max() - g.min())
    vp = 2000.0 + 1000.0 * g
    vs = vp / 1.8
    rec = dict(x=list(map(int, np.linspace(8, nx - 9, 32).round())), z=[4] * 32, comp=[2] * 32)
    shots = [dict(nt=nt, src=dict(x=[nx // 2], z=[4], comp=[2]), rec=rec, amp=1.0)]
    return Workload("c3", nx, nz, dx, 5e-4, nt, npml, 4, 0, 25.0, vp, vs, _gardner(vp), shots)


def c3mem():
    nx, nz, dx = 2000, 1000, 5.0
    vp = c2_layers(nx, nz)
    vs = vp / 1.73
    shots = [dict(nt=8000, src=dict(x=[1000], z=[2], comp=[2]), rec=_line(nx, 400), amp=1.0)]
    w = Workload("c3mem", nx, nz, dx, 4e-4, 8000, 40, 4, 0, 12.0, vp, vs, _gardner(vp), shots)
    sm = gaussian_filter(vp, 10.0, mode="nearest")
    w.background = (sm, sm / 1.73, _gardner(sm))
    return w


def _layered_with_fault(nx, nz, dx, seed, nlayer=8, vmin=1800.0, vmax=4200.0, offset_m=50.0,
                        dip_deg=60.0, x_centre=None):
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.uniform(0.05, 0.95, nlayer - 1)) * nz
    vps = np.sort(rng.uniform(vmin, vmax, nlayer))
    zz, xx = np.mgrid[0:nz, 0:nx].astype(float)
    xc = nx / 2 if x_centre is None else x_centre
    # fault plane through (xc, nz/2) dipping at dip_deg; hanging wall shifted down by offset
    hanging = (xx - xc) * np.tan(np.radians(dip_deg)) > (zz - nz / 2)
    zshift = np.where(hanging, zz - offset_m / dx, zz)
    layer = np.searchsorted(tops, zshift)
    return vps[np.clip(layer, 0, nlayer - 1)]


def c4(nshots=64, seed_geom=6, nx=1500, nz=600):
    """nx, nz below 1500 x 600 give the same kind of model and geometry on a smaller grid
    (source positions and receiver counts scaled with nx)."""
    dx = 5.0
    vp = _layered_with_fault(nx, nz, dx, 5)
    vs = vp / 1.73
    rng = np.random.default_rng(seed_geom)
    sc = nx / 1500.0
    shots = []
    for _ in range(nshots):
        nt = int(rng.integers(3000, 8001))
        nrec = max(2, int(round(int(rng.integers(50, 401)) * min(1.0, sc))))
        xs = int(round(int(rng.integers(100, 1401)) * sc))
        shots.append(dict(nt=nt, src=dict(x=[xs], z=[2], comp=[2]), rec=_line(nx, nrec), amp=1.0))
    ntmax = max(s["nt"] for s in shots)
    w = Workload("c4", nx, nz, dx, 5e-4, ntmax, 40, 4, 0, 12.0, vp, vs, _gardner(vp), shots)
    sm = gaussian_filter(vp, 10.0, mode="nearest")
    w.background = (sm, sm / 1.73, _gardner(sm))
    return w


def c5(nshots=128, nrec=600, nt=6000, nx=2400, nz=800):
    """nx, nz below 2400 x 800 give the same model shrunk (body radius and the
    150 m smoothing of the start model scale with nz / 800)."""
    dx = 5.0
    sc = nz / 800.0
    vp = _layered_with_fault(nx, nz, dx, 7, nlayer=10, vmax=4000.0, x_centre=nx * 0.35)
    vp2 = _layered_with_fault(nx, nz, dx, 7, nlayer=10, vmax=4000.0, offset_m=80.0, dip_deg=-55.0,
                              x_centre=nx * 0.7)
    vp = np.where(np.arange(nx)[None, :] > nx * 0.55, vp2, vp)
    zz, xx = np.mgrid[0:nz, 0:nx].astype(float)
    ang = np.arctan2(zz - 0.6 * nz, xx - 0.5 * nx)
    r = np.hypot((zz - 0.6 * nz) / 0.6, xx - 0.5 * nx)
    body = r < 120.0 * sc * (1.0 + 0.25 * np.sin(3 * ang) + 0.1 * np.cos(5 * ang))
    vp = np.where(body, 4500.0, vp)
    vs = vp / 1.73
    xs = np.linspace(20, nx - 21, nshots).round().astype(int)
    shots = [dict(nt=nt, src=dict(x=[int(x)], z=[2], comp=[2]), rec=_line(nx, nrec), amp=1.0) for x in xs]
    w = Workload("c5", nx, nz, dx, 5e-4, nt, 40, 4, 0, 8.0, vp, vs, _gardner(vp), shots)
    sm = gaussian_filter(vp, 30.0 * sc, mode="nearest")
    w.background = (sm, sm / 1.73, _gardner(sm))
    return w


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c3mem": c3mem, "c4": c4, "c5": c5}
