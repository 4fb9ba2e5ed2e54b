"""The fused step kernels (one launch per time step, ping-pong state, set
with a negative tile height: -8 / -16 the tile kernel, -32 / -64 the
z-marching segment kernel) against the oracle, bit for bit, on the
same cases as the split kernels: orders 2/4 with and without free surface
(wavefields and CPML memory), gradients with full and decimated windows,
checkpointed gradients, uneven batches (final state of shots that end
early), explosive / vx sources, boundary geometry and snapshots."""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from gpucase import plan_of, shots_of
from test_gpu_parity import CASES, FIELDS, MEMS, WINDOWS, _fwi_problem, _oracle_gradient, _oracle_mems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"o{c['fdorder']}fs{c['fsurf']}")
@pytest.mark.parametrize("prec,tz", [("fp64", -8), ("fp32", -8), ("fp32", -16), ("fp64", -16),
                                     ("fp64", -32), ("fp32", -32), ("fp64", -64), ("fp32", -64)])
def test_fused_forward_bit_exact(case, prec, tz):
    print(f"fused forward {case} {prec} {tz}", flush=True)
    pb = RC.small_problem(nx=90, nz=50, npml=8, nt=260, **case)
    ref = O.forward(pb, prec="f64" if prec == "fp64" else "f32")
    plan = plan_of(pb, prec, tz=tz)
    plan.forward(energy=True)
    assert np.array_equal(plan.records(), ref["rec"].astype(np.float64))
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k].astype(np.float64)), k
    om = _oracle_mems(pb, ref["mem"])
    for k in MEMS:
        assert np.array_equal(plan.field(k), om[k].astype(np.float64)), k


@pytest.mark.parametrize("tz", [-16, -32])
@pytest.mark.parametrize("win", list(WINDOWS))
@pytest.mark.parametrize("fdorder", [2, 4])
def test_fused_gradient_fp64_bit_exact(fdorder, win, tz):
    print(f"fused gradient fp64 order {fdorder} window {win} tile {tz}", flush=True)
    pb, obs = _fwi_problem(fdorder=fdorder, fwi=WINDOWS[win])
    f, m, res, a = _oracle_gradient(pb, obs)
    plan = plan_of(pb, "fp64", tz=tz)
    plan.forward(energy=True, gradient=True)
    assert np.array_equal(plan.store(), f["store"])
    plan.misfit([obs])
    plan.adjoint(energy=True, gradient=True)
    for got, k in zip(plan.gradient(0), ("grad_lam", "grad_mu", "grad_rho")):
        assert np.array_equal(got, a[k]), k
    for k in FIELDS:
        assert np.array_equal(plan.field(k), a[k]), k


@pytest.mark.parametrize("tz", [-8, -64])
def test_fused_checkpoint_gradient_equals_full_store(tz):
    print(f"fused checkpoint vs store, tile {tz}", flush=True)
    pb, obs = _fwi_problem(nt=333, fwi=WINDOWS["decimated"])
    plan = plan_of(pb, "fp32", tz=tz)
    plan.forward(gradient=True)
    plan.misfit([obs])
    plan.adjoint(gradient=True)
    g_full = plan.gradient(0)
    split = plan_of(pb, "fp32", tz=8)
    split.forward(gradient=True)
    split.misfit([obs])
    split.adjoint(gradient=True)
    for a, b in zip(g_full, split.gradient(0)):
        assert np.array_equal(a, b)
    for k in (7, 50, 333):
        plan.set_store_mode("checkpoint", k)
        plan.forward(gradient=True)
        plan.misfit([obs])
        plan.adjoint(gradient=True)
        for a, b in zip(g_full, plan.gradient(0)):
            assert np.array_equal(a, b), k


@pytest.mark.parametrize("tz", [-16, -8, -32, -64])
def test_batched_uneven_shots_variants(tz):
    """Batches with the fused kernel equal single-shot plans of the split
    kernels: shots that end before the longest one keep their own final
    state (which buffer set holds it depends on nt)."""
    print(f"batched uneven shots, tile {tz}", flush=True)
    base = RC.small_problem(nx=80, nz=48, npml=8, nt=10, fdorder=4, fsurf=1, hetero=True)
    rng = np.random.default_rng(9)
    shots, obs = [], []
    for i, nt in enumerate((150, 231, 90, 230, 177)):
        nrec = int(rng.integers(3, 12))
        xs = int(rng.integers(base.ppad + 5, base.nxt - base.ppad - 5))
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        rec = dict(x=list(np.linspace(base.ppad, base.nxt - base.ppad - 1, nrec).astype(int)),
                   z=[base.fpad + 1] * nrec, comp=[2] * nrec)
        shots.append(dict(nt=nt, src=dict(x=[xs], z=[base.fpad + 3], comp=[2]), src_signal=sig,
                          rec=rec))
        obs.append(rng.standard_normal((nrec, nt)) * 1e-3)
    win = RC.full_window(base, dt=2)
    batch = plan_of(base, "fp32", shots=shots, tz=tz)
    batch.set_window(win)
    batch.forward(gradient=True)
    fwd_fields = [{k: batch.field(k, i) for k in FIELDS + MEMS} for i in range(len(shots))]
    mb = batch.misfit(obs)
    batch.adjoint(gradient=True)
    for i, s in enumerate(shots):
        one = plan_of(base, "fp32", shots=[s], tz=8)
        one.set_window(win)
        one.forward(gradient=True)
        assert np.array_equal(one.records(), batch.records(i))
        for k in FIELDS + MEMS:
            assert np.array_equal(one.field(k), fwd_fields[i][k]), (i, k)
        assert one.misfit([obs[i]])[0] == mb[i]
        one.adjoint(gradient=True)
        for a, b in zip(one.gradient(0), batch.gradient(i)):
            assert np.array_equal(a, b)
        for k in FIELDS:
            assert np.array_equal(one.field(k), batch.field(k, i)), (i, k)


@pytest.mark.parametrize("tzs", [(-8, -16), (-32, -64), (-64, -32)])
def test_fused_explosive_vx_sources_and_boundary_geometry(tzs):
    print(f"fused explosive / vx sources, boundary geometry, tiles {tzs}", flush=True)
    pb = RC.small_problem(nx=60, nz=60, npml=8, nt=200, fdorder=4, fsurf=0, nsrc=3, src_comp=4)
    pb.src = dict(pb.src)
    pb.src["comp"] = [4, 1, 2]
    pb.src["x"] = [pb.src["x"][0], pb.src["x"][0], pb.src["x"][2]]
    pb.src["z"] = list(pb.src["z"])
    ref = O.forward(pb)
    plan = plan_of(pb, "fp64", tz=tzs[0])
    plan.forward()
    assert np.array_equal(plan.records(), ref["rec"])
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k])
    pb = RC.small_problem(nx=40, nz=30, npml=0, nt=120, fdorder=2, fsurf=1, nsrc=2)
    pb.src = dict(x=[20, 20, 0], z=[5, 5, 0], comp=[2, 2, 2], shot=[0, 0, 0])
    pb.src_signal = np.vstack([pb.src_signal, pb.src_signal[:1] * 0.5])
    pb.rec = dict(x=[0, 15, 41], z=[0, 0, 7], comp=[2, 1, 2], shot=[0, 0, 0])
    ref = O.forward(pb)
    plan = plan_of(pb, "fp64", tz=tzs[1])
    plan.forward()
    assert np.array_equal(plan.records(), ref["rec"])
    for k in FIELDS:
        assert np.array_equal(plan.field(k), ref[k])


def test_fused_snapshots_and_autotune():
    print("fused snapshots, autotune", flush=True)
    pb = RC.small_problem(nx=70, nz=40, npml=8, nt=121, fdorder=4, fsurf=1)
    plan = plan_of(pb, "fp64", tz=-16)
    plan.set_snapshots(40, "szz")
    plan.forward()
    snaps = plan.snapshots(0)
    ref = O.forward(RC.small_problem(nx=70, nz=40, npml=8, nt=81, fdorder=4, fsurf=1))
    assert np.array_equal(snaps[2], ref["szz"])
    p32 = plan_of(pb, "fp32")
    best = p32.autotune(20)
    assert best in (4, 8, 16, 32, -8, -16, -32, -64)
    p32.forward()
    assert np.array_equal(p32.records(), O.forward(pb, prec="f32")["rec"].astype(np.float64))


@pytest.mark.parametrize("tz,steps", [(8, 16), (-16, 8), (16, 2), (-64, 8)])
def test_graph_time_loop_equals_direct_launches(tz, steps):
    """Forward and adjoint replayed as captured graphs (uneven batch: the
    active shots change between graph chunks; the chunk length does not
    divide the runs) give what direct launches give, bit for bit."""
    print(f"graph time loop, tile {tz}, {steps} steps per graph", flush=True)
    base = RC.small_problem(nx=70, nz=40, npml=8, nt=10, fdorder=4, fsurf=1, hetero=True)
    rng = np.random.default_rng(4)
    shots, obs = [], []
    for i, nt in enumerate((123, 200, 77)):
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        rec = dict(x=list(np.linspace(base.ppad, base.nxt - base.ppad - 1, 5).astype(int)),
                   z=[base.fpad + 1] * 5, comp=[2] * 5)
        shots.append(dict(nt=nt, src=dict(x=[base.ppad + 10 + 15 * i], z=[base.fpad + 3], comp=[2]),
                          src_signal=sig, rec=rec))
        obs.append(rng.standard_normal((5, nt)) * 1e-3)
    win = RC.full_window(base, dt=3)
    out = []
    for graph in (0, steps):
        plan = plan_of(base, "fp32", shots=shots, tz=tz)
        plan.set_window(win)
        plan.set_graph_steps(graph)
        plan.forward(gradient=True)
        f = [plan.records(i) for i in range(3)] + [plan.field(k, i) for i in range(3) for k in FIELDS]
        plan.misfit(obs)
        plan.adjoint(gradient=True)
        g = [x for i in range(3) for x in plan.gradient(i)] + [plan.field(k, 1) for k in FIELDS]
        out.append(f + g)
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_shot_orders_give_identical_results():
    """Shot-major and banded workgroup orders of a batch (set by hand and by
    the autotuner) give the same traces and gradients bit for bit."""
    print("shot orders", flush=True)
    base = RC.small_problem(nx=150, nz=70, npml=8, nt=10, fdorder=4, fsurf=0, hetero=True)
    shots = []
    for i, nt in enumerate((160, 120, 200)):
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        shots.append(dict(nt=nt, src=dict(x=[30 + 40 * i], z=[base.ppad + 4], comp=[2]), src_signal=sig,
                          rec=dict(x=[20, 70, 140], z=[base.ppad + 2] * 3, comp=[2, 2, 1])))
    obs = [np.full((3, s["nt"]), 1e-4) for s in shots]
    out = []
    for band in (0, 1, 4, "auto"):
        plan = plan_of(base, "fp32", shots=shots)
        plan.set_window(RC.full_window(base, dt=2))
        if band == "auto":
            plan.autotune(20)
            assert plan.shot_order() in (0, 1, 4)
        else:
            plan.set_shot_order(band)
        plan.forward(gradient=True)
        plan.misfit(obs)
        plan.adjoint(gradient=True)
        out.append(([plan.records(i) for i in range(3)], [plan.gradient(i) for i in range(3)]))
    for rec, grad in out[1:]:
        for a, b in zip(rec, out[0][0]):
            assert np.array_equal(a, b)
        for ga, gb in zip(grad, out[0][1]):
            for a, b in zip(ga, gb):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("fsurf", [0, 1])
@pytest.mark.parametrize("fdorder", [2, 4])
def test_march_tall_grid_segments_bit_exact(fdorder, fsurf):
    """The z-marching kernel on a grid of several 32- and 64-row segments with a short last
    one, CPML on every side, a source near a segment edge and receivers on the free surface
    and at the bottom: traces, wavefields and CPML memory against the oracle, bit for bit,
    and the FP64 gradient of the same case."""
    print(f"march tall grid order {fdorder} fsurf {fsurf}", flush=True)
    pb = RC.small_problem(nx=100, nz=150, npml=10, nt=300, fdorder=fdorder, fsurf=fsurf, nsrc=2, nrec=6)
    pb.src = dict(pb.src, z=[63, 96])
    pb.rec = dict(pb.rec, z=[pb.fpad, pb.fpad, 64, 64, pb.nzt - pb.ppad - 2, pb.nzt - pb.fpad - 1])
    ref = O.forward(pb)
    om = _oracle_mems(pb, ref["mem"])
    for tz in (-32, -64):
        plan = plan_of(pb, "fp64", tz=tz)
        plan.forward(energy=True)
        assert np.array_equal(plan.records(), ref["rec"]), tz
        for k in FIELDS:
            assert np.array_equal(plan.field(k), ref[k]), (tz, k)
        for k in MEMS:
            assert np.array_equal(plan.field(k), om[k]), (tz, k)
