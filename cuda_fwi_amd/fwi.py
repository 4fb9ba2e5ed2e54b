"""Multi-shot FWI on one device plan: the host loop over the C ABI for models
that the reference's INI cannot describe (layers, faults, bodies).

It follows FDWaveModel::fwi_simulate (CUDA_week2/fd_wave_model.cu:237-968)
with the batched, window-correct choices of the C++ driver
(cuda_fwi_amd/csrc/host/fd_wave_model.cpp, ``[mi355x] gradient = discrete``):

- per iteration one batched forward with the energy weights, the FP64 misfit
  sum_shots sum_rec sum_t 0.5 dt (d - d_obs)^2 and the discrete adjoint (or,
  with gradient="reference", the reference's adjoint with the correlation
  gradient over the model window, forward states from checkpoints);
- descent direction -dJ/dm, divided by the summed source illumination
  sum_shots Ws (or sqrt(sum Ws * sum Wr) with the receiver side's energy,
  precond="source_receiver") plus 0.005 of its maximum over the window; optionally zero in
  the first `mute` rows of the window, where the sources and receivers sit
  and the gradient peaks (a common FWI practice; 0 keeps every row);
  optionally smoothed by a Gaussian of `smooth` cells (0: not smoothed),
  scaled by a depth gain ((row + 1) / rows)^depth_power, and ramped up by a
  cosine over the `taper` rows below the muted ones;
- the reference's bracketing step search (fwi_step_length_estimation.cu:
  11-224: halve until the misfit drops, then grow by 1.5 until it rises, at
  most 7 tries each) with batched trial forwards, and the vertex of the
  parabola through the three (step, misfit) points in FP64;
- update m + s d on the window with s = step max(m) / max|d| and the bounds
  of fwi_grad_update (fd_medium.cu:161-240), rho with half the step; the
  padding is refilled by edge replication of the model region.

Everything that touches the wavefields runs on the GPU through the plan; this
module only moves the model (three grids per iteration) and scalars.  Shots
can run in groups (one batched launch per group, the gradients and
illuminations summed over groups) when the discrete adjoint's checkpoints and
segment states of all shots do not fit in HBM together.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter

from . import psvfwi as P

BOUNDS = {"rho": (1.5, 3000.0), "lam": (0.0, 48.0e9), "mu": (0.0, 27.0e9)}


class StepSearchFailed(RuntimeError):
    pass


@dataclass
class History:
    misfit: list = field(default_factory=list)
    step: list = field(default_factory=list)
    trials: list = field(default_factory=list)
    seconds: list = field(default_factory=list)
    model_error: list = field(default_factory=list)


def _update(m, d, step, win, bounds):
    """m + s d on the window cells, s = step max(m) / max|d| (window maxima)."""
    z1, z2, x1, x2 = win
    out = m.copy()
    w_m, w_d = m[z1:z2, x1:x2], d[z1:z2, x1:x2]
    dmax = np.abs(w_d).max()
    if dmax == 0.0:
        return out
    s = step * w_m.max() / dmax
    out[z1:z2, x1:x2] = np.clip(w_m + s * w_d, *bounds)
    return out


def _repad(a, win):
    z1, z2, x1, x2 = win
    inner = a[z1:z2, x1:x2]
    return np.pad(inner, ((z1, a.shape[0] - z2), (x1, a.shape[1] - x2)), mode="edge")


def _parabola(L2, sl):
    try:
        s = P.parabola_step(L2, sl)
    except P.PsvfwiError:
        return None
    return s if np.isfinite(s) and s > 0 else None


class Inversion:
    """plan: a device plan with pml set; shots: the shot dicts of
    psvfwi.Plan.set_shots; observed: one [nrec, nt] array per shot; model:
    lam, mu, rho on the padded grid; win: (z1, z2, x1, x2), the cells the
    inversion updates; group: shots per batched launch (None: all)."""

    def __init__(self, plan: P.Plan, shots, observed, model, win, segment=100, step=0.01, mute=0, group=None,
                 smooth=0.0, precond="source", gradient="discrete", fwi_dt=5, taper=0, depth_power=0.0):
        self.plan = plan
        self.win = win
        self.step = step
        self.mute = int(mute)
        self.smooth = float(smooth)
        self.taper = int(taper)
        self.depth_power = float(depth_power)
        if precond not in ("source", "source_receiver"):
            raise ValueError("precond: 'source' or 'source_receiver'")
        self.precond = precond
        self.diag = []   # per iteration: where the largest direction component sits, and how peaked it is
        self.m = {"lam": model[0].copy(), "mu": model[1].copy(), "rho": model[2].copy()}
        self.shots = list(shots)
        self.observed = list(observed)
        n = len(self.shots)
        g = n if not group else int(group)
        self.groups = [list(range(i, min(i + g, n))) for i in range(0, n, g)]
        self.current = None
        self.log = None   # progress callback: one line per group of a gradient or trial
        if gradient not in ("discrete", "reference"):
            raise ValueError("gradient: 'discrete' or 'reference'")
        self.gradient = gradient
        if gradient == "discrete":
            plan.set_gradient_mode("discrete", segment)
        else:
            # the reference's correlation gradient over the model window, every fwi_dt-th
            # step, the forward state from checkpoints every `segment` steps
            z1, z2, x1, x2 = win
            plan.set_gradient_mode("reference")
            plan.set_store_mode("checkpoint", segment)
            plan.set_window(dict(dt=int(fwi_dt), dx=1, dz=1, x1=x1, x2=x2, z1=z1, z2=z2))
        self.forwards = 0
        self.adjoints = 0

    def _use(self, gi):
        if self.current != gi:
            ids = self.groups[gi]
            self.plan.set_shots([self.shots[i] for i in ids])
            self.plan.set_observed([self.observed[i] for i in ids])
            self.current = gi

    def _set(self, m):
        self.plan.set_model(m["lam"], m["mu"], *P.average_parameters(m["mu"], m["rho"]))

    def _misfit(self, m):
        self._set(m)
        total = 0.0
        for gi in range(len(self.groups)):
            self._use(gi)
            self.plan.forward(energy=False, gradient=False)
            self.forwards += 1
            total += float(np.sum(self.plan.misfit()))
            if self.log and len(self.groups) > 1:
                self.log(f"  trial forward, group {gi + 1}/{len(self.groups)}")
        return total

    def _gradient(self):
        """Misfit, illumination and dJ/d(lam, mu, rho) of the current model.  The
        illumination is the summed source energy sum_shots Ws, or with
        precond="source_receiver" sqrt(sum Ws * sum Wr), Wr the energy of the
        adjoint wavefield (the reference's sqrt(Ws Wr) of fd_wave_model.cu:1004-1052
        on the summed energies; with the discrete gradient one more adjoint run per
        group, without the gradient, since the discrete adjoint keeps no energy)."""
        self._set(self.m)
        L2, ws, wr, g = 0.0, 0.0, 0.0, None
        for gi, ids in enumerate(self.groups):
            self._use(gi)
            self.plan.forward(energy=True, gradient=True)
            self.forwards += 1
            L2 += float(np.sum(self.plan.misfit()))
            ws = ws + sum(self.plan.field("We", i) for i in range(len(ids)))
            sr = self.precond == "source_receiver"
            if sr and self.gradient == "discrete":
                # the discrete adjoint keeps no energy: an adjoint run without the gradient
                self.plan.adjoint(energy=True, gradient=False)
                self.adjoints += 1
                wr = wr + sum(self.plan.field("We", i) for i in range(len(ids)))
            self.plan.adjoint(energy=sr and self.gradient == "reference", gradient=True)
            self.adjoints += 1
            if sr and self.gradient == "reference":
                wr = wr + sum(self.plan.field("We", i) for i in range(len(ids)))
            gg = self._model_gradient()
            g = gg if g is None else [a + b for a, b in zip(g, gg)]
            if self.log and len(self.groups) > 1:
                self.log(f"  gradient, group {gi + 1}/{len(self.groups)}")
        ill = np.sqrt(ws * wr) if self.precond == "source_receiver" else ws
        return L2, ill, g

    def _model_gradient(self):
        """dJ/d(lam, mu, rho) of the current group on the padded grid.  The reference's
        correlation gradient approximates -dJ/dm on the window (its [nfz][nfx] arrays
        carry one more row and column, never written): negated and placed on the grid,
        zero outside the window."""
        if self.gradient == "discrete":
            return self.plan.model_gradient(-1)
        z1, z2, x1, x2 = self.win
        out = []
        for gw in self.plan.gradient(-1):
            a = np.zeros_like(self.m["lam"])
            a[z1:z2, x1:x2] = -gw[:z2 - z1, :x2 - x1]
            out.append(a)
        return out

    def _trial(self, d, est):
        m = self.m
        t = {k: _update(m[k], d[k], est * (0.5 if k == "rho" else 1.0), self.win, BOUNDS[k]) for k in m}
        return {k: _repad(v, self.win) for k, v in t.items()}

    def iterate(self):
        """One iteration; returns (misfit before the update, step taken, trials)."""
        L2, ill, g = self._gradient()
        z1, z2, x1, x2 = self.win
        w = ill + 0.005 * ill[z1:z2, x1:x2].max()
        d = {k: -gk / w for k, gk in zip(("lam", "mu", "rho"), g)}
        if self.smooth > 0:
            # Gaussian smoothing of the direction over the window (sigma in cells)
            for k, v in d.items():
                v[z1:z2, x1:x2] = gaussian_filter(v[z1:z2, x1:x2], self.smooth, mode="nearest")
        if self.depth_power > 0:
            # depth gain: row r of the window (r = 0 at its top) times ((r + 1) / rows)^p
            g = ((np.arange(z2 - z1) + 1.0) / (z2 - z1)) ** self.depth_power
            for v in d.values():
                v[z1:z2, :] *= g[:, None]
        if self.mute > 0:
            for v in d.values():
                v[z1:z1 + self.mute, :] = 0.0
        if self.taper > 0:
            # a cosine ramp from 0 to 1 over the `taper` rows below the muted ones
            r = np.arange(self.taper)
            g = 0.5 * (1.0 - np.cos(np.pi * (r + 0.5) / self.taper))
            a0 = z1 + self.mute
            n = min(self.taper, z2 - a0)
            for v in d.values():
                v[a0:a0 + n, :] *= g[:n, None]
        a = np.abs(d["lam"][z1:z2, x1:x2])
        iz, ix = np.unravel_index(int(np.argmax(a)), a.shape)
        self.diag.append({"lam_max_row": int(iz), "lam_max_col": int(ix),
                          "lam_max_over_p99": float(a.max() / max(np.percentile(a, 99.0), 1e-300))})
        # the reference's bracketing search
        L2t, sl = [L2, 0.0, 0.0], [0.0, 0.0, 0.0]
        est, step1, step2, itest, count, ntrial = self.step, False, False, 1, 0, 0
        while not (step1 and step2):
            L2t[itest] = self._misfit(self._trial(d, est))
            sl[itest] = est
            ntrial += 1
            if step1 and L2t[1] < L2t[2]:
                sl[2], step2 = est, True
            elif step1:
                sl[2] = est
                est += est / 2.0
                count += 1
            elif L2t[0] > L2t[1]:
                sl[1], step1, itest, count = est, True, 2, 0
                est += est / 2.0
            else:
                est /= 2.0
                count += 1
            if not step1 and count > 6:
                raise StepSearchFailed("Step length estimation failed")
            if step1 and count > 6:
                step2 = True
        opt = _parabola(L2t, sl)
        if opt is None:   # a parabola open downwards: keep the best tried step
            opt = sl[int(np.argmin(L2t[1:])) + 1]
        self.m = self._trial(d, opt)
        self.step = opt
        return L2, opt, ntrial

    def run(self, iters, truth=None, log=None, first=0):
        """iters iterations; first numbers the first one in the log lines."""
        self.log = log
        h = History()
        for i in range(iters):
            t0 = time.perf_counter()
            L2, s, n = self.iterate()
            h.seconds.append(time.perf_counter() - t0)
            h.misfit.append(L2)
            h.step.append(s)
            h.trials.append(n)
            if truth is not None:
                h.model_error.append(self.model_error(truth))
            if log:
                log(f"iteration {first + i}: misfit {L2:.6e}, step {s:.4g}, {n} trials, {h.seconds[-1]:.1f} s"
                    + (f", vp error {h.model_error[-1]:.4f}" if truth is not None else ""))
        return h

    def vp(self, m=None):
        m = self.m if m is None else m
        return np.sqrt((m["lam"] + 2.0 * m["mu"]) / m["rho"])

    def model_error(self, truth, sigma=0.0, m=None):
        """Relative L2 error of vp over the window against truth (lam, mu, rho); with
        sigma > 0 both vp fields are first low-passed by a Gaussian of sigma cells,
        which keeps the scales a band-limited inversion can resolve."""
        z1, z2, x1, x2 = self.win
        vt = np.sqrt((truth[0] + 2.0 * truth[1]) / truth[2])[z1:z2, x1:x2]
        v = self.vp(m)[z1:z2, x1:x2]
        if sigma > 0:
            vt = gaussian_filter(vt, sigma, mode="nearest")
            v = gaussian_filter(v, sigma, mode="nearest")
        return float(np.linalg.norm(v - vt) / np.linalg.norm(vt))
