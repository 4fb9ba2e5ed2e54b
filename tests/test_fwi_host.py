"""The multi-shot FWI loop of cuda_fwi_amd/fwi.py on the CPU, against a
stand-in plan whose misfit is a known quadratic of the model: the reference's
bracketing step search (fwi_step_length_estimation.cu:11-224), the parabola
vertex, the update bounds of fwi_grad_update (fd_medium.cu:161-240), the
edge-replicated padding, the muted rows and the shot groups.  The device
path of the same loop runs in test_gpu_fwi.py."""
import numpy as np
import pytest

from cuda_fwi_amd import fwi as F
from cuda_fwi_amd import psvfwi as P


class QuadraticPlan:
    """Stands in for psvfwi.Plan: shot s has misfit
    J_s = 0.5 c_s sum_window [((lam - lam*) / 1e9)^2 + ((mu - mu*) / 1e9)^2],
    its source illumination is 1 and its receiver illumination 4 everywhere,
    and model_gradient(-1) returns the
    exact dJ/d(lam, mu, rho) summed over the current shots (rho: 0).
    `flip` negates the gradient (a search that can only fail)."""

    def __init__(self, truth, win, flip=False):
        self.truth = truth
        self.win = win
        self.flip = flip
        self.mode = None
        self.calls = {"forward": 0, "adjoint": 0}

    def set_gradient_mode(self, mode, segment=64):
        self.mode = (mode, segment)

    def set_store_mode(self, mode, interval=0):
        self.store = (mode, interval)

    def set_window(self, win):
        self.window = win

    def gradient(self, shot=-1):
        """The reference's window gradient: -dJ/dm on [nfz][nfx] = window + 1 row and column
        (that row and column never written)."""
        z1, z2, x1, x2 = self.win
        g = self.model_gradient(shot)
        out = []
        for a in g:
            w = np.zeros((z2 - z1 + 1, x2 - x1 + 1))
            w[:-1, :-1] = -a[z1:z2, x1:x2]
            out.append(w)
        return out

    def set_shots(self, shots):
        self.shots = list(shots)

    def set_observed(self, observed):
        self.observed = list(observed)

    def set_model(self, lam, mu, mu_zx, rho_zp, rho_xp):
        assert mu_zx.shape == (lam.shape[0] - 1, lam.shape[1] - 1)
        self.lam, self.mu = lam.copy(), mu.copy()

    def _terms(self):
        z1, z2, x1, x2 = self.win
        dl = (self.lam - self.truth[0])[z1:z2, x1:x2] / 1e9
        dm = (self.mu - self.truth[1])[z1:z2, x1:x2] / 1e9
        return dl, dm

    def forward(self, energy=True, gradient=False):
        self.calls["forward"] += 1
        self.energy = 1.0
        dl, dm = self._terms()
        self.J = np.array([0.5 * s["c"] * (np.sum(dl * dl) + np.sum(dm * dm)) for s in self.shots])

    def misfit(self):
        return self.J.copy()

    def field(self, name, shot):
        assert name == "We"
        return np.full_like(self.lam, self.energy)

    def adjoint(self, energy=False, gradient=True):
        self.calls["adjoint"] += 1
        if energy:
            self.energy = 4.0   # the adjoint wavefield's energy: 4 per shot everywhere

    def model_gradient(self, shot=-1):
        z1, z2, x1, x2 = self.win
        dl, dm = self._terms()
        c = sum(s["c"] for s in self.shots) * (-1.0 if self.flip else 1.0)
        g = [np.zeros_like(self.lam) for _ in range(3)]
        g[0][z1:z2, x1:x2] = c * dl / 1e9
        g[1][z1:z2, x1:x2] = c * dm / 1e9
        return g


def _case(nz=14, nx=18, pad=3, seed=1):
    rng = np.random.default_rng(seed)
    shape = (nz + 2 * pad, nx + 2 * pad)
    win = (pad, pad + nz, pad, pad + nx)
    truth = [np.full(shape, 9.0e9), np.full(shape, 4.0e9), np.full(shape, 2200.0)]
    truth[0][win[0]:win[1], win[2]:win[3]] += 1.0e9 * rng.random((nz, nx))
    truth[1][win[0]:win[1], win[2]:win[3]] += 0.5e9 * rng.random((nz, nx))
    start = [np.full(shape, 9.5e9), np.full(shape, 4.2e9), np.full(shape, 2200.0)]
    shots = [dict(c=1.0 + 0.25 * i) for i in range(5)]
    observed = [np.zeros((2, 3)) for _ in shots]
    return truth, start, win, shots, observed


def _inversion(plan, start, win, shots, observed, **kw):
    return F.Inversion(plan, shots, observed, start, win, segment=40, step=0.01, **kw)


def test_misfit_falls_and_model_approaches_the_truth():
    truth, start, win, shots, observed = _case()
    plan = QuadraticPlan(truth, win)
    inv = _inversion(plan, start, win, shots, observed)
    assert plan.mode == ("discrete", 40)
    z1, z2, x1, x2 = win
    err0 = np.abs(inv.m["lam"] - truth[0])[z1:z2, x1:x2].max()
    h = inv.run(6)
    assert all(b < a for a, b in zip(h.misfit, h.misfit[1:]))
    assert all(s > 0 for s in h.step)
    assert all(2 <= t <= 16 for t in h.trials)
    assert np.abs(inv.m["lam"] - truth[0])[z1:z2, x1:x2].max() < err0
    # rho has a zero gradient: untouched
    assert np.array_equal(inv.m["rho"], start[2])
    # the padding is edge replication of the window
    for k in ("lam", "mu"):
        m = inv.m[k]
        assert np.array_equal(m, F._repad(m, win))
    # one gradient (forward + adjoint) and the trial forwards per iteration
    assert plan.calls["adjoint"] == 6
    assert plan.calls["forward"] == 6 + sum(h.trials)


def test_parabola_vertex_on_a_quadratic():
    """On an exactly quadratic misfit the parabola through the bracketing
    points has its vertex at the minimum along the search direction: the
    accepted step beats every step of a scan around it."""
    truth, start, win, shots, observed = _case()
    inv = _inversion(QuadraticPlan(truth, win), start, win, shots, observed)
    _, ill, g = inv._gradient()
    z1, z2, x1, x2 = win
    w = ill + 0.005 * ill[z1:z2, x1:x2].max()
    d = {k: -gk / w for k, gk in zip(("lam", "mu", "rho"), g)}
    L2, step, _ = inv.iterate()
    ref = _inversion(QuadraticPlan(truth, win), start, win, shots, observed)  # still at the start model
    ref.plan.set_shots(shots)

    def J(s):
        t = ref._trial(d, s)
        ref.plan.set_model(t["lam"], t["mu"], *P.average_parameters(t["mu"], t["rho"]))
        ref.plan.forward()
        return float(np.sum(ref.plan.misfit()))

    j_opt = J(step)
    assert j_opt < L2
    assert all(j_opt <= J(s) * (1 + 1e-9) for s in np.linspace(0.5, 1.5, 21) * step)


def test_shot_groups_equal_one_batch():
    truth, start, win, shots, observed = _case()
    a = _inversion(QuadraticPlan(truth, win), start, win, shots, observed)
    b = _inversion(QuadraticPlan(truth, win), start, win, shots, observed, group=2)
    assert [len(g) for g in b.groups] == [2, 2, 1]
    ha, hb = a.run(3), b.run(3)
    assert np.allclose(ha.misfit, hb.misfit, rtol=1e-12, atol=0)
    assert np.allclose(ha.step, hb.step, rtol=1e-9, atol=0)
    for k in a.m:
        assert np.allclose(a.m[k], b.m[k], rtol=1e-9, atol=0)


def test_muted_rows_and_bounds():
    truth, start, win, shots, observed = _case()
    z1, z2, x1, x2 = win
    inv = _inversion(QuadraticPlan(truth, win), start, win, shots, observed, mute=4)
    inv.run(2)
    for k in ("lam", "mu"):
        assert np.array_equal(inv.m[k][z1:z1 + 4, x1:x2], start[{"lam": 0, "mu": 1}[k]][z1:z1 + 4, x1:x2])
    # _update clips to the fwi_grad_update bounds and scales by max(m) / max|d|
    m = np.full((6, 6), 10.0e9)
    d = np.zeros((6, 6))
    d[2, 3] = -5.0
    d[3, 3] = 2.5
    out = F._update(m, d, 0.1, (1, 5, 1, 5), F.BOUNDS["lam"])
    assert out[2, 3] == pytest.approx(10.0e9 - 0.1 * 10.0e9)
    assert out[3, 3] == pytest.approx(10.0e9 + 0.05 * 10.0e9)
    big = F._update(m, d, 10.0, (1, 5, 1, 5), F.BOUNDS["lam"])
    assert big[2, 3] == F.BOUNDS["lam"][0] and big[3, 3] == F.BOUNDS["lam"][1]
    assert np.array_equal(F._update(m, np.zeros_like(m), 0.1, (1, 5, 1, 5), F.BOUNDS["lam"]), m)


def test_search_that_cannot_descend_raises():
    """A direction along which the misfit only grows: the search halves the
    step until its try count runs out and raises, as the reference exits."""
    truth, start, win, shots, observed = _case()
    inv = _inversion(QuadraticPlan(truth, win, flip=True), start, win, shots, observed)
    with pytest.raises(F.StepSearchFailed):
        inv.iterate()


def test_source_receiver_illumination():
    """precond="source_receiver": one more adjoint run per group for the
    receiver side's energy; the illumination is sqrt(sum Ws * sum Wr)."""
    truth, start, win, shots, observed = _case()
    plan = QuadraticPlan(truth, win)
    inv = _inversion(plan, start, win, shots, observed, precond="source_receiver", group=2)
    _, ill, _ = inv._gradient()
    n = len(shots)
    assert np.allclose(ill, np.sqrt(n * 1.0 * n * 4.0), rtol=1e-15)
    assert plan.calls["adjoint"] == 2 * len(inv.groups)
    plain = QuadraticPlan(truth, win)
    _, ill1, _ = _inversion(plain, start, win, shots, observed)._gradient()
    assert np.array_equal(ill1, np.full_like(ill1, n)) and plain.calls["adjoint"] == 1
    with pytest.raises(ValueError):
        _inversion(QuadraticPlan(truth, win), start, win, shots, observed, precond="hessian")


def test_smoothed_direction_keeps_the_mute_and_converges():
    """smooth > 0: the direction is low-passed over the window before the muted rows are
    zeroed; the loop still descends, and the muted rows stay at the start model."""
    truth, start, win, shots, observed = _case(nz=20, nx=24)
    z1, z2, x1, x2 = win
    inv = _inversion(QuadraticPlan(truth, win), start, win, shots, observed, smooth=1.5, mute=3)
    h = inv.run(4)
    assert all(b < a for a, b in zip(h.misfit, h.misfit[1:]))
    for k, i in (("lam", 0), ("mu", 1)):
        assert np.array_equal(inv.m[k][z1:z1 + 3, x1:x2], start[i][z1:z1 + 3, x1:x2])
    # the direction's peak is recorded per iteration, inside the window and below the mute
    assert len(inv.diag) == 4
    for d in inv.diag:
        assert 3 <= d["lam_max_row"] < z2 - z1 and 0 <= d["lam_max_col"] < x2 - x1
        assert d["lam_max_over_p99"] >= 1.0


def test_low_pass_model_error():
    """model_error(truth, sigma): a checkerboard difference (the scale a band-limited
    inversion cannot resolve) vanishes after the low-pass, a smooth one survives it."""
    truth, start, win, shots, observed = _case(nz=32, nx=32)
    z1, z2, x1, x2 = win
    inv = _inversion(QuadraticPlan(truth, win), start, win, shots, observed)
    lam, mu, rho = (np.full_like(truth[0], 9.0e9), np.full_like(truth[0], 4.0e9), np.full_like(truth[0], 2200.0))
    ref = (lam, mu, rho)
    zz, xx = np.mgrid[0:lam.shape[0], 0:lam.shape[1]]
    checker = {"lam": lam * (1.0 + 0.02 * (-1.0) ** (zz + xx)), "mu": mu, "rho": rho}
    smooth = {"lam": lam * (1.0 + 0.02 * np.sin(zz / 40.0)), "mu": mu, "rho": rho}
    e_c0, e_c4 = inv.model_error(ref, 0.0, checker), inv.model_error(ref, 4.0, checker)
    e_s0, e_s4 = inv.model_error(ref, 0.0, smooth), inv.model_error(ref, 4.0, smooth)
    assert e_c0 > 1e-3 and e_c4 < 0.05 * e_c0   # what survives: the sqrt of vp is not linear in lam
    assert e_s0 > 1e-4 and e_s4 > 0.8 * e_s0
    assert inv.model_error(ref, 0.0, {"lam": lam, "mu": mu, "rho": rho}) == 0.0


def test_reference_gradient_mode_equals_the_discrete_loop_on_the_same_gradient():
    """gradient="reference": the window gradient is negated and placed on the grid; with a
    plan whose reference gradient is exactly -dJ/dm the loop takes the discrete loop's
    iterations, and the plan is put into the checkpoint store mode over the model window."""
    truth, start, win, shots, observed = _case()
    a = _inversion(QuadraticPlan(truth, win), start, win, shots, observed)
    pb = QuadraticPlan(truth, win)
    b = _inversion(pb, start, win, shots, observed, gradient="reference", fwi_dt=3)
    z1, z2, x1, x2 = win
    assert pb.mode[0] == "reference" and pb.store == ("checkpoint", 40)
    assert pb.window == dict(dt=3, dx=1, dz=1, x1=x1, x2=x2, z1=z1, z2=z2)
    ha, hb = a.run(3), b.run(3)
    assert ha.misfit == hb.misfit and ha.step == hb.step
    for k in a.m:
        assert np.array_equal(a.m[k], b.m[k])
    with pytest.raises(ValueError):
        _inversion(QuadraticPlan(truth, win), start, win, shots, observed, gradient="adjoint")


def _first_update(**kw):
    truth, start, win, shots, observed = _case(nz=24, nx=20)
    inv = _inversion(QuadraticPlan(truth, win), start, win, shots, observed, **kw)
    inv.iterate()
    return inv.m["lam"] - start[0], win


def test_depth_gain_and_taper_scale_the_rows():
    """depth_power p multiplies row r of the window (r = 0 at its top) by ((r + 1) / rows)^p,
    and taper t ramps the t rows below the muted ones by 0.5 (1 - cos(pi (k + 0.5) / t)):
    between two rows of one update the ratio of the model changes is the unscaled ratio
    times the ratio of those weights, whatever step the search takes."""
    d0, win = _first_update()
    z1, z2, x1, x2 = win
    rows = z2 - z1
    c = x1 + 7
    for kw, weight in (
        (dict(depth_power=2.0), lambda r: ((r + 1.0) / rows) ** 2),
        (dict(mute=3, taper=6), lambda r: 0.0 if r < 3 else (0.5 * (1 - np.cos(np.pi * (r - 3 + 0.5) / 6))
                                                              if r < 9 else 1.0)),
    ):
        d1, _ = _first_update(**kw)
        ra, rb = 12, 5
        assert abs(d0[z1 + ra, c]) > 0 and abs(d0[z1 + rb, c]) > 0
        got = (d1[z1 + rb, c] / d1[z1 + ra, c]) / (d0[z1 + rb, c] / d0[z1 + ra, c])
        assert got == pytest.approx(weight(rb) / weight(ra), rel=1e-9), kw
    # muted rows stay, and a taper row moves less than the rows below the taper
    d2, _ = _first_update(mute=3, taper=6)
    assert np.all(d2[z1:z1 + 3, x1:x2] == 0.0)
