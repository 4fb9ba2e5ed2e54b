"""Pins the CPU oracle against the reference's own output files and against
closed forms, library routines and invariants.

Golden data (tests/golden/ref_forward_run.npz) are the reference's shipped
outputs of one forward run, written with 6 significant digits, so the
comparison tolerance is half a unit in the 6th digit (5e-6 relative) and
every exact zero must be reproduced (the numerical light cone).
"""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC

SIX_DIGITS = 5.0e-6


def _match_6_digits(ours, ref):
    ref = np.asarray(ref)
    assert np.array_equal(ours == 0, ref == 0), "zero pattern (light cone) differs"
    nz = ref != 0
    rel = np.abs(ours[nz] - ref[nz]) / np.abs(ref[nz])
    return rel.max() if rel.size else 0.0


def test_rec_signal_matches_reference_output(golden):
    """rec_signal.csv: 3 receivers x 334 steps incl. 1e-149 precursors."""
    out = O.forward(RC.reference_run())
    assert _match_6_digits(out["rec"], golden["rec_signal"]) <= SIX_DIGITS


@pytest.mark.parametrize("k", [39, 52, 66])
def test_vz_snapshots_match_reference_output(golden, k):
    """vz_snap<k>.csv = vz after 5k+1 updates; snapshot 66 reaches the x CPML strips."""
    out = O.forward(RC.reference_run(nt=5 * k + 1))
    rows, cols = golden["snap_rows"], golden["snap_cols"]
    assert _match_6_digits(out["vz"][np.ix_(rows, cols)], golden[f"snap{k}"]) <= SIX_DIGITS
    strips = out["vz"][:, np.r_[0:12, 611:623]]
    assert _match_6_digits(strips, golden[f"snap{k}_strips"]) <= SIX_DIGITS
    if k == 66:
        assert np.count_nonzero(golden["snap66_strips"]) > 500


def test_wavelet_matches_reference_output(golden):
    s = O.create_wavelet(2500, 0.2e-3, 1000.0, 15.0, 0.0, 1)
    assert _match_6_digits(s, golden["src_signal_row"]) <= SIX_DIGITS


def test_wavelet_closed_form():
    nt, dt, fc, ts, amp = 400, 1e-3, 12.0, 0.01, 3.0
    t = np.arange(nt) * dt
    tau = 3.14159265 * (t - 1.5 / fc - ts) / (1.5 / fc)
    ref = amp * (1 - 2 * tau ** 2) * np.exp(-2 * tau ** 2)
    assert np.allclose(O.create_wavelet(nt, dt, amp, fc, ts), ref, rtol=1e-13, atol=1e-15)
    # peak at t0 = 1.5/fc + ts
    assert abs(np.argmax(O.create_wavelet(nt, dt, amp, fc, ts)) * dt - (1.5 / fc + ts)) <= dt


def test_holberg_taylor_rows_are_textbook_weights():
    assert np.allclose(O.holbergcoeff(2, 0)[1:2], [1.0])
    assert np.allclose(O.holbergcoeff(4, 0)[1:3], [9 / 8, -1 / 24])
    assert np.allclose(O.holbergcoeff(6, 0)[1:4], [75 / 64, -25 / 384, 3 / 640])
    # staggered first-derivative consistency: sum (2k-1) c_k = 1 for Taylor
    for order in (2, 4, 6, 8, 10, 12):
        hc = O.holbergcoeff(order, 0)
        k = np.arange(1, order // 2 + 1)
        assert abs(np.sum((2 * k - 1) * hc[1:order // 2 + 1]) - 1) < 1e-12
    # Holberg rows deviate from exact consistency by their error class only
    hc = O.holbergcoeff(4, 1)
    assert np.allclose(hc[1:3], [1.1382, -0.046414])
    assert abs(hc[1] + 3 * hc[2] - 1) < 1e-2


def test_cpml_profile_invariants():
    npml, dh, fpml, dt = 12, 5.0, 15.0, 5e-4
    p = O.cpml_pro(npml, dh, fpml, 2.0, 1.0, 1e-3, 2750.0, dt)
    a, b, K, ah, bh, Kh = p
    assert np.all((b > 0) & (b <= 1)) and np.all((bh > 0) & (bh <= 1))
    assert np.all(a <= 0) and np.all(ah <= 0)
    assert np.all(K == 1) and np.all(Kh == 1)        # k_max_PML = 1
    # damping is strongest at the outer edge and vanishes at the inner edge
    assert a[0] < a[npml // 2] < 0 and a[npml] == 0
    assert a[2 * npml + 1] < a[npml + npml // 2] < 0 and a[npml + 1] == 0
    # the right side mirrors the left: a[npml+1+k] == a[npml-k]
    assert np.allclose(p[0, npml + 1:], p[0, npml::-1], rtol=1e-12)
    # closed form at the outer edge: d = d0, alpha = 0, b = exp(-d0 dt), a = b - 1
    d0 = -3 * 2750.0 * np.log(1e-3) / (2 * npml * dh)
    assert np.isclose(b[0], np.exp(-d0 * dt), rtol=1e-14)
    assert np.isclose(a[0], b[0] - 1, rtol=1e-12)
    # K_max > 1 grows K towards the edge
    p2 = O.cpml_pro(npml, dh, fpml, 2.0, 3.0, 1e-3, 2750.0, dt)
    assert np.isclose(p2[2, 0], 3.0) and p2[2, npml] == 1.0


def test_lame_and_averaging_closed_forms():
    lam, mu = O.velocity_to_lame(2500.0, 1443.0, 2000.0)
    assert np.isclose(mu, 1443.0 ** 2 * 2000.0) and np.isclose(lam, 2500.0 ** 2 * 2000 - 2 * mu)
    from scipy.stats import hmean
    rng = np.random.default_rng(1)
    mu = rng.uniform(1e9, 5e9, (7, 9))
    rho = rng.uniform(1500, 2500, (7, 9))
    mu_zx, rho_zp, rho_xp = O.average_parameters(mu, rho)
    for iz in range(6):
        for ix in range(8):
            q = [mu[iz, ix], mu[iz, ix + 1], mu[iz + 1, ix], mu[iz + 1, ix + 1]]
            assert np.isclose(mu_zx[iz, ix], hmean(q), rtol=1e-13)
            assert np.isclose(rho_zp[iz, ix], 1 / np.mean([rho[iz, ix], rho[iz + 1, ix]]), rtol=1e-14)
            assert np.isclose(rho_xp[iz, ix], 1 / np.mean([rho[iz, ix], rho[iz, ix + 1]]), rtol=1e-14)
    mu[3, 4] = 0.0
    mu_zx, _, _ = O.average_parameters(mu, rho)
    assert mu_zx[2, 3] == 0 and mu_zx[3, 4] == 0 and mu_zx[2, 4] == 0 and mu_zx[3, 3] == 0


def test_misfit_closed_form():
    rng = np.random.default_rng(2)
    syn, obs = rng.standard_normal((4, 50)), rng.standard_normal((4, 50))
    m, res, st = O.l2_adjoint_sources(1e-3, syn, obs)
    assert st == 0 and np.array_equal(res, syn - obs)
    assert np.isclose(m, 0.5 * 1e-3 * np.sum((syn - obs) ** 2), rtol=1e-13)
    syn[2, 7] = 1e11
    assert O.l2_adjoint_sources(1e-3, syn, obs)[2] == 1


def test_parabola_vertex_matches_polyfit():
    rng = np.random.default_rng(3)
    for _ in range(20):
        sl = np.sort(rng.uniform(0, 1, 3))
        c = rng.uniform(0.5, 2.0), rng.uniform(-2, 2), rng.uniform(0.1, 3)
        L2 = c[0] * sl ** 2 + c[1] * sl + c[2] + rng.normal(0, 0.05, 3)
        p = np.polyfit(sl, L2, 2)
        assert np.isclose(O.calc_opt_step(L2, sl), -p[1] / (2 * p[0]), rtol=1e-8)


def test_energy_scaling_special_cases():
    """scale_gradients_with_energy_weights (CUDA_week2/fd_wave_model.cu:1004-1052):
    with Ws = Wr = w everywhere the weight is w(1 + 0.005), so each gradient
    is divided by w(1.005) C^2; cells outside the loop range [z1, nfz) x
    [x1, nfx) are untouched; Ws Wr = 4 at one cell and 1 elsewhere gives
    weights 2.01 and 1.01 there (max sqrt = 2); off-stride cells are skipped."""
    nfz, nfx = 6, 8
    W = np.full((9, 11), 3.0)
    g = [np.full((nfz, nfx), v) for v in (1.0, 2.0, 4.0)]
    out = O.scale_gradients_energy(W, W, 2.0, 3.0, 5.0, nfz, nfx, 1, 2, 1, 1, g)
    w = 3.0 + 0.005 * 3.0
    for o, v, C in zip(out, (1.0, 2.0, 4.0), (3.0, 5.0, 2.0)):
        assert np.allclose(o[1:, 2:], v / (w * C * C), rtol=1e-15)
        assert np.all(o[0, :] == v) and np.all(o[:, :2] == v)
    Ws = np.ones((9, 11))
    Ws[3, 4] = 4.0
    out = O.scale_gradients_energy(Ws, np.ones((9, 11)), 1.0, 1.0, 1.0, nfz, nfx, 0, 0, 2, 2,
                                   [np.ones((nfz, nfx))] * 3)
    # (3, 4) is off the stride-2 loop: its energy is not seen, (2, 3) is not scaled
    assert np.isclose(out[0][2, 2], 1 / 1.005) and out[0][2, 3] == 1.0
    Ws[2, 2] = 4.0
    out = O.scale_gradients_energy(Ws, np.ones((9, 11)), 1.0, 1.0, 1.0, nfz, nfx, 0, 0, 2, 2,
                                   [np.ones((nfz, nfx))] * 3)
    assert np.isclose(out[0][2, 2], 1 / 2.01) and np.isclose(out[0][0, 0], 1 / 1.01)


def test_material_update_special_cases():
    """fwi_grad_update (CUDA_week2/fd_medium.cu:161-240): s = step max(mat) /
    grad_max, where the scan keeps the signed value of an entry whose
    magnitude exceeds the current grad_max (so once grad_max is negative the
    next entry replaces it); with the largest-magnitude entry negative and
    scanned last, the cell of that entry moves by exactly step max(mat) in the
    +s grad direction (it rises); results are clipped to the material bounds;
    without override the old model is kept."""
    mat = np.full((5, 6), 2.0e9)
    mat[1, 1] = 4.0e9
    grad = np.zeros((4, 5))
    grad[3, 4] = -8.0                      # largest magnitude, negative, scanned last
    grad[1, 2] = 2.0
    old, new = O.fwi_grad_update(mat, grad, 0.01, 4, 5, 0, 0, 1, 1, False, 2)
    assert np.array_equal(old, mat)
    s = 0.01 * 4.0e9 / -8.0
    assert new[3, 4] == 2.0e9 + s * -8.0 and new[3, 4] > 2.0e9
    assert new[1, 2] == 2.0e9 + s * 2.0 and new[1, 2] < 2.0e9
    g2 = grad.copy()
    g2[3, 4], g2[2, 3] = 0.0, -8.0         # a later zero entry then becomes grad_max
    _, bad = O.fwi_grad_update(mat, g2, 0.01, 4, 5, 0, 0, 1, 1, False, 2)
    assert bad[2, 3] == 0.0 and not np.isfinite(bad[0, 0])   # s = inf: -inf clipped, 0 * inf
    assert new[4, 5] == mat[4, 5]           # outside the loop range: unchanged
    old, new = O.fwi_grad_update(mat, grad, 100.0, 4, 5, 0, 0, 1, 1, True, 2)
    assert new[3, 4] == 48.0e9 and new[1, 2] == 0.0   # clipped to [0, 48e9]
    assert np.array_equal(old[:4, :5], new[:4, :5])   # override copies the window
    rho = np.full((5, 6), 2000.0)
    _, new = O.fwi_grad_update(rho, grad, 10.0, 4, 5, 0, 0, 1, 1, False, 1)
    assert new[3, 4] == 3000.0 and new[1, 2] == 1.5   # rho bounds [1.5, 3000]
    # full-grid loop indices address the gradient directly: cell (1, 2) with
    # z1 = 1, x1 = 2 uses grad[1][2]; max(mat) is taken over the loop's cells
    # only, which exclude mat[1, 1]
    _, new = O.fwi_grad_update(mat, grad, 0.01, 4, 5, 1, 2, 1, 1, False, 2)
    assert new[1, 2] == 2.0e9 + (0.01 * 2.0e9 / -8.0) * 2.0 and new[0, 0] == mat[0, 0]


def test_forward_is_linear_in_source_amplitude():
    """Doubling every source doubles every output bit for bit (FP64)."""
    pb = RC.small_problem(nt=150)
    a = O.forward(pb)
    pb.src_signal = pb.src_signal * 2.0
    b = O.forward(pb)
    for k in ("rec", "vx", "vz", "sxx", "szx", "szz"):
        assert np.array_equal(b[k], 2.0 * a[k])


def test_forward_reciprocity_homogeneous():
    """vz force at A recorded at B equals vz force at B recorded at A."""
    pb = RC.small_problem(nx=50, nz=40, nt=250, fsurf=0, hetero=False, nsrc=1, nrec=1)
    A, B = (pb.ppad + 12, pb.ppad + 10), (pb.ppad + 25, pb.ppad + 37)
    pb.src = dict(x=[A[1]], z=[A[0]], comp=[2], shot=[0])
    pb.rec = dict(x=[B[1]], z=[B[0]], comp=[2], shot=[0])
    ab = O.forward(pb)["rec"][0]
    pb.src, pb.rec = dict(pb.rec), dict(pb.src)
    ba = O.forward(pb)["rec"][0]
    # exact in the interior stencil; the CPML strips break it at the 1e-7 level
    assert np.linalg.norm(ab - ba) / np.linalg.norm(ab) < 1e-5


def test_cpml_absorbs_energy():
    """With damping the late-time energy is far below the undamped run."""
    pb = RC.small_problem(nx=40, nz=40, nt=900, fsurf=0, hetero=False, nsrc=1, nrec=1)
    e_damp = sum(np.sum(O.forward(pb)[k] ** 2) for k in ("vx", "vz"))
    pb.pml = np.vstack([np.zeros(pb.pml.shape[1]), np.ones(pb.pml.shape[1]),
                        np.ones(pb.pml.shape[1])] * 2)
    e_free = sum(np.sum(O.forward(pb)[k] ** 2) for k in ("vx", "vz"))
    assert e_damp < 1e-3 * e_free


@pytest.mark.parametrize("fsurf", [0, 1])
@pytest.mark.parametrize("fdorder", [2, 4])
def test_adjoint_propagation_is_time_reversed_forward(fsurf, fdorder):
    """adjoint_kernel_PSV runs the forward operator with the injected signal
    read at it = nt-1 ... 0: its final wavefield equals a forward run with the
    reversed signal, bit for bit."""
    pb = RC.small_problem(nt=120, fsurf=fsurf, fdorder=fdorder)
    pb.fwi = RC.full_window(pb)
    f = O.forward(pb)
    sig = np.random.default_rng(5).standard_normal((len(pb.rec["x"]), pb.nt)) * 1e-3
    adj = O.adjoint(pb, pb.rec, sig, f["store"])
    pb2 = RC.small_problem(nt=120, fsurf=fsurf, fdorder=fdorder)
    pb2.src, pb2.src_signal, pb2.rec = pb.rec, sig[:, ::-1].copy(), None
    fw = O.forward(pb2)
    for k in ("vx", "vz", "sxx", "szx", "szz", "We"):
        assert np.array_equal(adj[k], fw[k]), k


def _compliance_pairings(L, M, fwd, adj):
    """Per cell, -adj^T (dS/dlam) fwd and the three parts of -adj^T (dS/dmu) fwd along
    the eigenmodes of S (pressure xx = zz, deviatoric xx = -zz, shear), where S = C^-1 is
    the compliance of 2-D plane-strain isotropic elasticity, C = [[lam+2mu, lam, 0],
    [lam, lam+2mu, 0], [0, 0, mu]] on Voigt vectors (sxx, szz, sxz).  S comes from
    numpy's inverse and its derivatives from central differences (step 1e-4 relative),
    so nothing here restates the reference's correlation formulas.  fwd, adj: (..., 3)."""
    def S(lam, mu):
        c = np.zeros(lam.shape + (3, 3))
        c[..., 0, 0] = c[..., 1, 1] = lam + 2 * mu
        c[..., 0, 1] = c[..., 1, 0] = lam
        c[..., 2, 2] = mu
        return np.linalg.inv(c)
    hl, hm = (1e-4 * L)[..., None, None], (1e-4 * M)[..., None, None]
    dSl = (S(L + hl[..., 0, 0], M) - S(L - hl[..., 0, 0], M)) / (2 * hl)
    dSm = (S(L, M + hm[..., 0, 0]) - S(L, M - hm[..., 0, 0])) / (2 * hm)
    form = lambda D, a, b: -np.einsum("...i,...ij,...j->...", a, D, b)
    parts = []
    for e in (np.array([1.0, 1.0, 0.0]) / np.sqrt(2), np.array([1.0, -1.0, 0.0]) / np.sqrt(2),
              np.array([0.0, 0.0, 1.0])):
        parts.append(form(dSm, (adj @ e)[..., None] * e, (fwd @ e)[..., None] * e))
    return form(dSl, adj, fwd), parts


def test_gradient_brute_force_tiny():
    """Re-evaluates the reference correlation (adjoint_kernel_PSV_GPU.cu:818-863)
    from independently produced wavefields: the adjoint state before step it is
    the forward state after nt-1-it steps driven by the reversed residual, the
    adjoint displacement is the running dt-weighted sum of those velocities.
    Checks the decimated window, the ordering (displacement updated before the
    correlation), the v/dt store and every term's sign.  The lambda and mu terms
    come from the derivative of the plane-strain compliance (_compliance_pairings),
    not from the reference's formulas: lambda's is -adj (dS/dlam) fwd; mu's is the
    same derivative with its deviatoric part taken 4 times, the reference's s2."""
    nt = 23
    pb = RC.small_problem(nx=10, nz=9, npml=3, nt=nt, fsurf=1, nsrc=1, nrec=3, dt=6e-4, fc=60.0)
    pb.fwi = dict(dt=3, dx=2, dz=2, x1=pb.ppad - 1, x2=pb.nxt - 2, z1=pb.fpad, z2=pb.nzt - 3)
    nft, nfz, nfx = pb.fwi_dims()
    f = O.forward(pb)
    res = np.random.default_rng(7).standard_normal((3, nt))
    adj = O.adjoint(pb, pb.rec, res, f["store"])

    # forward states before each step (independent runs, no store)
    fw_states = []
    for k in range(nt):
        q = RC.small_problem(nx=10, nz=9, npml=3, nt=k, fsurf=1, nsrc=1, nrec=3, dt=6e-4, fc=60.0) \
            if k > 0 else None
        if k == 0:
            fw_states.append({n: np.zeros((pb.nzt, pb.nxt)) for n in ("vx", "vz", "sxx", "szx", "szz")})
            continue
        q.src_signal = pb.src_signal[:, :k].copy()
        o = O.forward(q)
        fw_states.append(o)
    # adjoint states before step it: forward with reversed residual for nt-1-it steps
    adj_states = {}
    for it in range(nt - 1, -1, -1):
        k = nt - 1 - it
        if k == 0:
            adj_states[it] = {n: np.zeros((pb.nzt, pb.nxt)) for n in ("vx", "vz", "sxx", "szx", "szz")}
            continue
        q = RC.small_problem(nx=10, nz=9, npml=3, nt=k, fsurf=1, nsrc=1, nrec=3, dt=6e-4, fc=60.0)
        q.src, q.rec = pb.rec, None
        q.src_signal = res[:, ::-1][:, :k].copy()
        adj_states[it] = O.forward(q)
    dt = pb.dt
    fpad = pb.fpad
    inner = (slice(fpad, pb.nzt - fpad), slice(fpad, pb.nxt - fpad))
    ux = np.zeros((pb.nzt, pb.nxt))
    uz = np.zeros((pb.nzt, pb.nxt))
    gl = np.zeros((nfz, nfx)); gm = np.zeros((nfz, nfx)); gr = np.zeros((nfz, nfx))
    gmx = np.zeros((nfz, nfx))
    lam, mu = pb.lam, pb.mu
    w = pb.fwi
    zs = np.arange(w["z1"], w["z2"], w["dz"])
    xs = np.arange(w["x1"], w["x2"], w["dx"])
    for it in range(nt - 1, -1, -1):
        a = adj_states[it]
        ux[inner] += dt * a["vx"][inner]
        uz[inner] += dt * a["vz"][inner]
        if it % w["dt"] == 0:
            F = fw_states[it]
            sel = np.ix_(zs, xs)
            fsxx, fszz, fszx = F["sxx"][sel], F["szz"][sel], F["szx"][sel]
            fvx, fvz = F["vx"][sel] / dt, F["vz"][sel] / dt
            fwd = np.stack([fsxx, fszz, fszx], -1)
            adj_s = np.stack([a["sxx"][sel], a["szz"][sel], a["szx"][sel]], -1)
            dlam, (dmu_p, dmu_q, dmu_s) = _compliance_pairings(lam[sel], mu[sel], fwd, adj_s)
            # kinetic pairing: adjoint displacement against the stored forward v / dt
            s4 = np.einsum("i...,i...->...", np.stack([ux[sel], uz[sel]]), np.stack([fvx, fvz]))
            gl[:len(zs), :len(xs)] += w["dt"] * dt * dlam
            gm[:len(zs), :len(xs)] += w["dt"] * dt * (dmu_p + 4.0 * dmu_q + dmu_s)
            gmx[:len(zs), :len(xs)] += w["dt"] * dt * (dmu_p + dmu_q + dmu_s)
            gr[:len(zs), :len(xs)] -= w["dt"] * dt * s4
    for ours, ref in ((adj["grad_lam"], gl), (adj["grad_mu"], gm), (adj["grad_rho"], gr)):
        assert np.abs(ref).max() > 0
        assert np.allclose(ours, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())
    # the reference's mu kernel is not the compliance derivative: its deviatoric term
    # (adjoint_kernel_PSV_GPU.cu:846-847) has no factor 1/4, so it is 4x the derivative's
    # (replicated); the exact derivative differs visibly
    assert np.abs(adj["grad_mu"] - gmx).max() > 1e-3 * np.abs(gmx).max()
    # the last decimated row/column is never written (strict < in the window loops)
    assert nfz > len(zs) or nfx > len(xs)
    assert np.all(adj["grad_lam"][len(zs):, :] == 0) and np.all(adj["grad_lam"][:, len(xs):] == 0)


def _lag(a, b):
    """Delay of trace b behind trace a in samples (cross-correlation peak,
    refined by a parabola through its three top samples)."""
    c = np.correlate(b, a, "full")
    k = int(np.argmax(c))
    y0, y1, y2 = c[k - 1], c[k], c[k + 1]
    return k - (len(a) - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)


@pytest.mark.parametrize("fdorder,tol", [(2, 0.03), (4, 0.01)])
def test_homogeneous_p_and_s_speeds(fdorder, tol):
    """A vertical point force in a homogeneous medium radiates P waves along
    the vertical and S waves along the horizontal.  The move-out between two
    receivers on each axis gives vp = sqrt((lam + 2 mu) / rho) and
    vs = sqrt(mu / rho).  dx != dz, so a swapped grid step fails; a dropped
    2 mu term (vp -> 1444 m/s) or a dropped or sign-flipped second Holberg
    coefficient (speeds x 9/8 or x 5/4) fail too.  The tolerance covers the
    grid dispersion of each order at 4.6 points per shortest S wavelength."""
    npml, n = 20, 200
    fpad = fdorder // 2
    ppad = npml + fpad
    nxt = nzt = n + 2 * ppad
    dx, dz, dt, fc, nt = 5.0, 4.0, 4e-4, 25.0, 1400
    vp, vs, rho = 2500.0, 1443.0, 2000.0
    lam0, mu0 = O.velocity_to_lame(vp, vs, rho)
    lam, mu, rh = (np.full((nzt, nxt), v) for v in (lam0, mu0, rho))
    z0 = x0 = ppad + 60
    r1, r2 = 40, 100
    rec = dict(x=[x0, x0, x0 + r1, x0 + r2], z=[z0 + r1, z0 + r2, z0, z0], comp=[2] * 4, shot=[0] * 4)
    pb = O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, dt=dt, dx=dx, dz=dz, fsurf=0,
                   hc=O.holbergcoeff(fdorder, 1), lam=lam, mu=mu, rho=rh,
                   pml=O.cpml_pro(npml, dx, fc, 2.0, 1.0, 1e-3, 2750.0, dt),
                   src=dict(x=[x0], z=[z0], comp=[2], shot=[0]),
                   src_signal=O.create_wavelet(nt, dt, 1.0, fc, 0.0, 1)[None], rec=rec)
    d = O.forward(pb)["rec"]
    cp = (r2 - r1) * dz / (_lag(d[0], d[1]) * dt)
    cs = (r2 - r1) * dx / (_lag(d[2], d[3]) * dt)
    assert abs(cp / vp - 1) < tol, cp
    assert abs(cs / vs - 1) < tol, cs


def _misfit_of(pb, lam, mu, rho, obs, coeff=None):
    pb.lam, pb.mu = lam, mu
    pb.mu_zx, pb.rho_zp, pb.rho_xp = O.average_parameters(mu, rho)
    if coeff:
        name, z, x, h = coeff
        arr = getattr(pb, name).copy()
        arr[z, x] += h
        setattr(pb, name, arr)
    return O.l2_adjoint_sources(pb.dt, O.forward(pb)["rec"].copy(), obs)[0]


@pytest.mark.parametrize("fdorder,fsurf,npml", [(2, 1, 8), (4, 0, 8), (4, 1, 6), (2, 0, 0)])
def test_discrete_gradient_matches_central_differences(fdorder, fsurf, npml):
    """The discrete adjoint (an extension) is the exact derivative of the
    misfit: central differences with a 1e-4 relative step at seeded cells,
    the free-surface row, a CPML corner and the last interior cell, for lam,
    mu and rho (through the averages) and for the averaged coefficients
    themselves.  Differences stay below 1e-7 of max |g| (about 1e-10 is seen;
    the truncation error of the difference quotient is ~1e-8 relative)."""
    pb = RC.small_problem(nx=30, nz=24, npml=npml, nt=160, fdorder=fdorder, fsurf=fsurf, seed=4, nrec=6)
    lam, mu, rho = pb.lam.copy(), pb.mu.copy(), pb.rho.copy()
    rng = np.random.default_rng(7)
    obs = O.forward(pb)["rec"] * 0.7 + 1e-3 * rng.standard_normal((len(pb.rec["x"]), pb.nt))
    _, res, _ = O.l2_adjoint_sources(pb.dt, O.forward(pb)["rec"].copy(), obs)
    g = O.discrete_gradient(pb, res)
    mzx, rzp, rxp = pb.mu_zx.copy(), pb.rho_zp.copy(), pb.rho_xp.copy()
    gl, gm, gr = O.chain_gradient(mu, mzx, rzp, rxp, g)
    pts = list(zip(rng.integers(pb.fpad, pb.nzt - pb.fpad, 8), rng.integers(pb.fpad, pb.nxt - pb.fpad, 8)))
    pts += [(pb.fpad, pb.nxt // 2), (pb.fpad + 1, pb.fpad + 2), (pb.nzt - pb.fpad - 1, pb.nxt - pb.fpad - 1)]
    checked = 0
    for name, arr, G in (("lam", lam, gl), ("mu", mu, gm), ("rho", rho, gr)):
        scale = np.abs(G).max()
        assert scale > 0
        for z, x in pts:
            h = 1e-4 * arr[z, x]
            args = dict(lam=lam, mu=mu, rho=rho)
            up, dn = arr.copy(), arr.copy()
            up[z, x] += h
            dn[z, x] -= h
            args[name] = up
            jp = _misfit_of(pb, args["lam"], args["mu"], args["rho"], obs)
            args[name] = dn
            jm = _misfit_of(pb, args["lam"], args["mu"], args["rho"], obs)
            assert abs((jp - jm) / (2 * h) - G[z, x]) < 1e-7 * scale, (name, z, x)
            checked += 1
    for name, arr, key in (("mu_zx", mzx, "g_mu_zx"), ("rho_zp", rzp, "g_rho_zp"), ("rho_xp", rxp, "g_rho_xp")):
        scale = np.abs(g[key]).max()
        for z, x in pts[:4]:
            h = 1e-4 * arr[z, x]
            jp = _misfit_of(pb, lam, mu, rho, obs, (name, z, x, h))
            jm = _misfit_of(pb, lam, mu, rho, obs, (name, z, x, -h))
            assert abs((jp - jm) / (2 * h) - g[key][z, x]) < 1e-7 * scale, (name, z, x)
    assert checked == 33


def test_discrete_adjoint_dot_product():
    """<L dm, r> = <dm, L^T r>: the linearised traces for a model
    perturbation dm (central differences of the forward, which is linear in
    the traces' dependence on dm up to O(h^2)) against the gradient of the
    linear functional sum r * d, which the discrete adjoint returns when the
    residual is r / dt."""
    pb = RC.small_problem(nx=26, nz=20, npml=6, nt=140, fdorder=4, fsurf=1, seed=6, nrec=5)
    rng = np.random.default_rng(8)
    lam, mu, rho = pb.lam.copy(), pb.mu.copy(), pb.rho.copy()
    dlam = RC.smooth_random(lam.shape, 9, 2.0) * lam * 1e-4
    r = rng.standard_normal((len(pb.rec["x"]), pb.nt))
    g = O.discrete_gradient(pb, r / pb.dt)
    lhs = np.sum(g["g_lam"] * dlam)

    def traces(l):
        pb.lam = l
        return O.forward(pb)["rec"]
    rhs = np.sum(r * (traces(lam + dlam) - traces(lam - dlam)) / 2)
    assert abs(lhs - rhs) < 1e-6 * abs(rhs)


def test_threaded_oracle_is_bit_identical():
    """The row-parallel OpenMP build (the multi-core CPU baseline) gives the
    single-threaded oracle's bits: forward fields, traces, store, gradients."""
    import os
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    pb = RC.small_problem(nx=50, nz=40, npml=8, nt=150, fdorder=4, fsurf=0, seed=2)
    pb.fwi = RC.full_window(pb, dt=2)
    res = np.random.default_rng(1).standard_normal((len(pb.rec["x"]), pb.nt)) * 1e-3
    out = []
    for threaded in (False, True):
        O.set_threaded(threaded)
        try:
            f = O.forward(pb)
            a = O.adjoint(pb, pb.rec, res, f["store"])
        finally:
            O.set_threaded(False)
        out.append((f, a))
    (f0, a0), (f1, a1) = out
    for k in ("rec", "vx", "vz", "sxx", "szx", "szz", "We", "store", "mem"):
        assert np.array_equal(f0[k], f1[k]), k
    for k in ("grad_lam", "grad_mu", "grad_rho", "vx", "szz"):
        assert np.array_equal(a0[k], a1[k]), k


def test_free_surface_long_records_stay_bounded():
    """3000 steps with a free surface: the late traces decay at orders 2 and
    4 (the reference's order-2 surface treatment carried to order 4 grows
    without bound; the order-4 extension integrates sxx and images szx about
    the surface instead)."""
    for fdorder in (2, 4):
        pb = RC.small_problem(nx=80, nz=60, npml=20, nt=3000, fdorder=fdorder, fsurf=1, hetero=False,
                              nsrc=1, dt=5e-4, fc=15.0)
        r = O.forward(pb)["rec"]
        early, late = np.abs(r[:, 300:600]).max(), np.abs(r[:, 2700:]).max()
        assert late < 1e-2 * early, (fdorder, early, late)


@pytest.mark.parametrize("fdorder,tol", [(2, 0.06), (4, 0.02)])
def test_rayleigh_wave_speed(fdorder, tol):
    """Surface vz of a vertical force at the free surface: the move-out of the
    dominant (Rayleigh) arrival between two receivers 600 m apart is
    c_R = 0.9194 vs for a Poisson solid (vp / vs = sqrt(3)).  The reference's
    order-2 surface (szx forced to zero half a cell down, sxx set to its
    increment) runs about 4.6 % slow at 17 cells per Rayleigh wavelength; the
    order-4 surface is within 1 %."""
    npml = 20
    fpad = fdorder // 2
    ppad = npml + fpad
    nx, nz = 400, 120
    nxt, nzt = nx + 2 * ppad, nz + ppad + fpad
    dx, dt, fc, nt = 5.0, 4e-4, 15.0, 2600
    vs = 2500.0 / np.sqrt(3.0)
    lam0, mu0 = O.velocity_to_lame(2500.0, vs, 2000.0)
    lam, mu, rho = (np.full((nzt, nxt), v) for v in (lam0, mu0, 2000.0))
    x0, r1, r2 = ppad + 40, 100, 220
    pb = O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, dt=dt, dx=dx, dz=dx, fsurf=1,
                   hc=O.holbergcoeff(fdorder, 1), lam=lam, mu=mu, rho=rho,
                   pml=O.cpml_pro(npml, dx, fc, 2.0, 1.0, 1e-3, 2750.0, dt),
                   src=dict(x=[x0], z=[fpad + 1], comp=[2], shot=[0]),
                   src_signal=O.create_wavelet(nt, dt, 1.0, fc, 0.0, 1)[None],
                   rec=dict(x=[x0 + r1, x0 + r2], z=[fpad, fpad], comp=[2, 2], shot=[0, 0]))
    d = O.forward(pb)["rec"]
    t1, t2 = int(np.argmax(np.abs(d[0]))), int(np.argmax(np.abs(d[1])))
    w = 150
    c = (r2 - r1) * dx / ((t2 - t1 + _lag(d[0][t1 - w:t1 + w], d[1][t2 - w:t2 + w])) * dt)
    assert abs(c / (0.9194 * vs) - 1) < tol, c
