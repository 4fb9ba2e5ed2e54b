"""Parity of the HIP path with the oracle (CPU restatement of the reference).

Both sides evaluate every multiply and add separately rounded in the same
order (no FMA contraction), so the FP64 build and the FP32 build of the
kernels reproduce the FP64 / FP32 oracle bit for bit.  The precision claim
of the FP32 path is then stated against the FP64 oracle (the reference's own
precision): seismograms within 1e-5 and gradients within 1e-4 relative L2.
"""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from gpucase import plan_of, rel_l2, shots_of

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vz", "sxx", "szx", "szz", "We")
MEMS = ("mem_vx_x", "mem_vz_x", "mem_sxx_x", "mem_szx_x", "mem_vx_z", "mem_vz_z",
        "mem_szz_z", "mem_szx_z")


def _oracle_mems(pb, mem):
    tp = 2 * (pb.npml + 1)
    nx = pb.nzt * tp
    nz = tp * pb.nxt
    x = [mem[i * nx:(i + 1) * nx].reshape(pb.nzt, tp) for i in range(4)]
    z = [mem[4 * nx + i * nz:4 * nx + (i + 1) * nz].reshape(tp, pb.nxt) for i in range(4)]
    return dict(zip(MEMS, x + z))


CASES = [
    dict(fdorder=2, fsurf=1), dict(fdorder=2, fsurf=0),
    dict(fdorder=4, fsurf=1), dict(fdorder=4, fsurf=0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"o{c['fdorder']}fs{c['fsurf']}")
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_forward_bit_exact(case, prec):
    print(f"forward bit-exact {case} {prec}", flush=True)
    pb = RC.small_problem(nx=90, nz=50, npml=8, nt=260, **case)
    ref = O.forward(pb, prec="f64" if prec == "fp64" else "f32")
    plan = plan_of(pb, prec)
    plan.forward(energy=True)
    assert np.abs(ref["rec"]).max() > 0
    assert np.array_equal(plan.records(), ref["rec"].astype(np.float64))
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k].astype(np.float64)), k
    om = _oracle_mems(pb, ref["mem"])
    for k in MEMS:
        assert np.array_equal(plan.field(k), om[k].astype(np.float64)), k


def test_reference_run_fp32_within_tolerance(golden):
    """The reference's shipped run (613 x 623, 334 steps): FP32 traces against
    the reference's own output and against the FP64 oracle."""
    print("reference run fp32", flush=True)
    pb = RC.reference_run()
    plan = plan_of(pb, "fp32")
    plan.forward(energy=False)
    rec = plan.records()
    ref64 = O.forward(pb)["rec"]
    e64 = rel_l2(rec, ref64)
    egold = rel_l2(rec, golden["rec_signal"])
    print(f"  rel L2 fp32 vs fp64 oracle {e64:.3e}, vs reference csv {egold:.3e}")
    assert e64 < 1e-5 and egold < 1e-5
    plan64 = plan_of(pb, "fp64")
    plan64.forward(energy=False)
    assert np.array_equal(plan64.records(), ref64)


def _fwi_problem(fdorder=2, fsurf=1, nt=300, fwi=None, hetero=True):
    pb_true = RC.small_problem(nx=70, nz=44, npml=8, nt=nt, fdorder=fdorder, fsurf=fsurf, seed=11,
                               nrec=9, nsrc=1, hetero=True)
    obs = O.forward(pb_true)["rec"]
    pb = RC.small_problem(nx=70, nz=44, npml=8, nt=nt, fdorder=fdorder, fsurf=fsurf, seed=11,
                          nrec=9, nsrc=1, hetero=False)
    pb.fwi = fwi(pb) if fwi else RC.full_window(pb)
    return pb, obs


def _oracle_gradient(pb, obs, prec="f64"):
    f = O.forward(pb, prec=prec)
    m, res, st = O.l2_adjoint_sources(pb.dt, f["rec"].astype(np.float64), obs)
    a = O.adjoint(pb, pb.rec, res, f["store"], prec=prec)
    return f, m, res, a


WINDOWS = {
    "full": None,
    "decimated": lambda pb: dict(dt=3, dx=2, dz=3, x1=pb.ppad - 3, x2=pb.nxt - 5, z1=pb.fpad + 1,
                                 z2=pb.nzt - 4),
}


@pytest.mark.parametrize("win", list(WINDOWS))
@pytest.mark.parametrize("fdorder", [2, 4])
def test_gradient_fp64_bit_exact(fdorder, win):
    print(f"gradient fp64 order {fdorder} window {win}", flush=True)
    pb, obs = _fwi_problem(fdorder=fdorder, fwi=WINDOWS[win])
    f, m, res, a = _oracle_gradient(pb, obs)
    plan = plan_of(pb, "fp64")
    plan.forward(energy=True, gradient=True)
    assert np.array_equal(plan.store(), f["store"])
    mis = plan.misfit([obs])
    assert abs(mis[0] - m) <= 1e-12 * m
    plan.adjoint(energy=True, gradient=True)
    gl, gm, gr = plan.gradient(0)
    assert np.abs(a["grad_lam"]).max() > 0
    assert np.array_equal(gl, a["grad_lam"])
    assert np.array_equal(gm, a["grad_mu"])
    assert np.array_equal(gr, a["grad_rho"])
    for k in FIELDS:
        assert np.array_equal(plan.field(k), a[k]), k


@pytest.mark.parametrize("fdorder", [2, 4])
def test_gradient_fp32_bit_exact_and_tolerance(fdorder):
    print(f"gradient fp32 order {fdorder}", flush=True)
    pb, obs = _fwi_problem(fdorder=fdorder, nt=400)
    # FP32 oracle, residual formed from FP32 traces exactly as the device does
    f32 = O.forward(pb, prec="f32")
    plan = plan_of(pb, "fp32")
    plan.forward(energy=True, gradient=True)
    assert np.array_equal(plan.records(), f32["rec"].astype(np.float64))
    plan.misfit([obs])
    res32 = (f32["rec"].astype(np.float64) - obs).astype(np.float32)
    a32 = O.adjoint(pb, pb.rec, res32, f32["store"], prec="f32")
    plan.adjoint(energy=True, gradient=True)
    g = plan.gradient(0)
    for got, k in zip(g, ("grad_lam", "grad_mu", "grad_rho")):
        assert np.array_equal(got, a32[k].astype(np.float64)), k
    # against the FP64 reference path
    f64, m64, res64, a64 = _oracle_gradient(pb, obs)
    es = rel_l2(plan.records(), f64["rec"])
    eg = [rel_l2(got, a64[k]) for got, k in zip(g, ("grad_lam", "grad_mu", "grad_rho"))]
    print(f"  fp32 vs fp64: seismogram {es:.2e}, gradients {eg}")
    assert es < 1e-5
    assert max(eg) < 1e-4


@pytest.mark.parametrize("tz", [8, -16, -64])
@pytest.mark.parametrize("prec", ["fp32fma", "fp32x2"])
@pytest.mark.parametrize("fdorder", [2, 4])
def test_gradient_other_precisions_within_tolerance(fdorder, prec, tz):
    """FP32 with contracted multiply-adds and compensated FP32, on the split, fused-tile and
    z-marching steps, against the FP64 oracle at the stated tolerances (seismograms 1e-5,
    gradients 1e-4); compensated FP32 within 1e-9 of it."""
    print(f"gradient {prec} order {fdorder} tile {tz}", flush=True)
    pb, obs = _fwi_problem(fdorder=fdorder, nt=400)
    f64, m64, res64, a64 = _oracle_gradient(pb, obs)
    plan = plan_of(pb, prec, tz=tz)
    plan.forward(energy=True, gradient=True)
    mis = plan.misfit([obs])[0]
    plan.adjoint(energy=True, gradient=True)
    es = rel_l2(plan.records(), f64["rec"])
    eg = [rel_l2(got, a64[k]) for got, k in zip(plan.gradient(0), ("grad_lam", "grad_mu", "grad_rho"))]
    print(f"  {prec} vs fp64: seismogram {es:.2e}, gradients {eg}, misfit {abs(mis - m64) / m64:.2e}")
    assert es < 1e-5 and max(eg) < 1e-4
    if prec == "fp32x2":
        assert es < 1e-9 and max(eg) < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_checkpoint_gradient_equals_full_store(prec):
    print(f"checkpoint vs store {prec}", flush=True)
    pb, obs = _fwi_problem(nt=333, fwi=WINDOWS["decimated"])
    plan = plan_of(pb, prec)
    plan.forward(gradient=True)
    plan.misfit([obs])
    plan.adjoint(gradient=True)
    g_full = plan.gradient(0)
    for k in (1, 7, 50, 333, 400):
        plan.set_store_mode("checkpoint", k)
        plan.forward(gradient=True)
        plan.misfit([obs])
        plan.adjoint(gradient=True)
        g_ck = plan.gradient(0)
        for a, b in zip(g_full, g_ck):
            assert np.array_equal(a, b), k


@pytest.mark.parametrize("tz", [4, 16, 32])
def test_tile_heights_agree(tz):
    print(f"tile {tz}", flush=True)
    pb = RC.small_problem(nx=100, nz=70, npml=10, nt=200, fdorder=4, fsurf=1)
    a = plan_of(pb, "fp32")
    a.forward()
    b = plan_of(pb, "fp32", tz=tz)
    b.forward()
    assert np.array_equal(a.records(), b.records())
    for k in FIELDS:
        assert np.array_equal(a.field(k), b.field(k))


def test_autotune_picks_a_tile_and_keeps_results():
    print("autotune", flush=True)
    pb = RC.small_problem(nx=120, nz=60, npml=10, nt=150, fdorder=4, fsurf=0)
    ref = O.forward(pb, prec="f32")
    plan = plan_of(pb, "fp32")
    best = plan.autotune(30)
    assert best in (4, 8, 16, 32, -8, -16)
    plan.forward()
    assert np.array_equal(plan.records(), ref["rec"].astype(np.float64))


def test_batched_uneven_shots_equal_single_shots():
    """Shots of different nt and receiver counts in one batch give exactly
    what each shot gives alone (records, gradients)."""
    print("batched uneven shots", flush=True)
    base = RC.small_problem(nx=80, nz=48, npml=8, nt=10, fdorder=4, fsurf=1, hetero=True)
    rng = np.random.default_rng(9)
    shots, obs = [], []
    for i, nt in enumerate((150, 230, 90, 230)):
        nrec = int(rng.integers(3, 12))
        xs = int(rng.integers(base.ppad + 5, base.nxt - base.ppad - 5))
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        rec = dict(x=list(np.linspace(base.ppad, base.nxt - base.ppad - 1, nrec).astype(int)),
                   z=[base.fpad + 1] * nrec, comp=[2] * nrec)
        shots.append(dict(nt=nt, src=dict(x=[xs], z=[base.fpad + 3], comp=[2]), src_signal=sig,
                          rec=rec))
        obs.append(rng.standard_normal((nrec, nt)) * 1e-3)
    win = RC.full_window(base, dt=2)
    batch = plan_of(base, "fp32", shots=shots)
    batch.set_window(win)
    batch.forward(gradient=True)
    mb = batch.misfit(obs)
    batch.adjoint(gradient=True)
    for i, s in enumerate(shots):
        one = plan_of(base, "fp32", shots=[s])
        one.set_window(win)
        one.forward(gradient=True)
        assert np.array_equal(one.records(), batch.records(i))
        m1 = one.misfit([obs[i]])
        assert m1[0] == mb[i]
        one.adjoint(gradient=True)
        for a, b in zip(one.gradient(0), batch.gradient(i)):
            assert np.array_equal(a, b)
    total = batch.gradient(-1)
    parts = [batch.gradient(i) for i in range(len(shots))]
    for k in range(3):
        assert np.allclose(total[k], sum(p[k] for p in parts), rtol=1e-12, atol=0)


def test_explosive_and_vx_sources_bit_exact():
    print("explosive / vx sources", flush=True)
    pb = RC.small_problem(nx=60, nz=60, npml=8, nt=200, fdorder=4, fsurf=0, nsrc=3, src_comp=4)
    pb.src = dict(pb.src)
    pb.src["comp"] = [4, 1, 2]
    pb.src["x"] = [pb.src["x"][0], pb.src["x"][0], pb.src["x"][2]]   # co-located 4 and 1
    pb.src["z"] = list(pb.src["z"])
    ref = O.forward(pb)
    plan = plan_of(pb, "fp64")
    plan.forward()
    assert np.array_equal(plan.records(), ref["rec"])
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k])


def test_colocated_sources_and_boundary_geometry():
    """Two sources on one cell add in index order; a receiver on the ghost row
    and one on a boundary column; no CPML."""
    print("co-located sources, boundary receivers, npml 0", flush=True)
    pb = RC.small_problem(nx=40, nz=30, npml=0, nt=120, fdorder=2, fsurf=1, nsrc=2)
    pb.src = dict(x=[20, 20], z=[5, 5], comp=[2, 2], shot=[0, 0])
    pb.rec = dict(x=[0, 15, 41], z=[0, 0, 7], comp=[2, 1, 2], shot=[0, 0, 0])
    ref = O.forward(pb)
    plan = plan_of(pb, "fp64")
    plan.forward()
    assert np.array_equal(plan.records(), ref["rec"])
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k])


def test_snapshots_match_reference_output(golden):
    """Snapshots every 5 steps of vz (the field the shipped vz_snap<k>.csv
    hold): FP64 snapshot k equals the oracle after 5k+1 steps bit for bit,
    and matches the reference's CSV to its 6 printed digits."""
    print("snapshots fp64 vz / 5", flush=True)
    from test_oracle_golden import SIX_DIGITS, _match_6_digits
    pb = RC.reference_run()
    plan = plan_of(pb, "fp64")
    plan.set_snapshots(5, "vz")
    plan.forward(energy=False)
    snaps = plan.snapshots(0)
    assert snaps.shape == ((pb.nt - 1) // 5 + 1, pb.nzt, pb.nxt)
    rows, cols = golden["snap_rows"], golden["snap_cols"]
    for k in (39, 52, 66):
        assert _match_6_digits(snaps[k][np.ix_(rows, cols)], golden[f"snap{k}"]) <= SIX_DIGITS
        strips = snaps[k][:, np.r_[0:12, 611:623]]
        assert _match_6_digits(strips, golden[f"snap{k}_strips"]) <= SIX_DIGITS
    ref = O.forward(RC.reference_run(nt=5 * 52 + 1))
    assert np.array_equal(snaps[52], ref["vz"])
