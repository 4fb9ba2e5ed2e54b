"""Edge cases of the HIP path against the oracle: shots without receivers or
sources inside a batch, records of one to four steps, the smallest grids the
stencil and the CPML allow, grids one cell past a tile, receivers on the last
row and column and on the grid's corners, and FWI windows on the grid's edge or empty.

Every comparison is bit for bit (the kernels and the oracle round every
operation the same way, see test_gpu_parity.py).
"""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from gpucase import plan_of, shots_of

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vz", "sxx", "szx", "szz", "We")
GRADS = ("grad_lam", "grad_mu", "grad_rho")


def _assert_fields(plan, ref, shot=0):
    for k in FIELDS:
        assert np.array_equal(plan.field(k, shot), ref[k].astype(np.float64)), k


def _grid_problem(nzt, nxt, fdorder, npml, fsurf, nt, seed=3, src_at=None):
    """Homogeneous-plus-noise model on an explicit padded grid, one vz source
    (by default in the interior's first cell) and receivers on the grid's
    corners and on the interior's last cell."""
    fpad = fdorder // 2
    rng = np.random.default_rng(seed)
    rho = 2000.0 * (1 + 0.05 * rng.random((nzt, nxt)))
    mu = 1443.0 ** 2 * rho
    lam = 2500.0 ** 2 * rho - 2 * mu
    dt, dx = 4e-4, 5.0
    pml = O.cpml_pro(npml, dx, 25.0, 2.0, 1.0, 1e-3, 2750.0, dt)
    rec = dict(x=[0, nxt - 1, nxt - fpad - 1, nxt - fpad - 1, 0],
               z=[0, nzt - 1, nzt - fpad - 1, fpad, nzt - fpad - 1],
               comp=[1, 2, 1, 2, 2], shot=[0] * 5)
    sz, sx = src_at if src_at else (fpad, fpad)
    src = dict(x=[sx], z=[sz], comp=[2], shot=[0])
    sig = rng.standard_normal((1, nt))
    return O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=fdorder, npml=npml, dt=dt, dx=dx, dz=dx,
                     fsurf=fsurf, hc=O.holbergcoeff(fdorder, 1), lam=lam, mu=mu, rho=rho, pml=pml,
                     src=src, src_signal=sig, rec=rec)


def test_batch_with_a_shot_without_receivers_and_one_without_sources():
    """A shot with no receivers still propagates; a shot with no sources stays
    zero; neither disturbs its neighbours; their gradients are zero."""
    print("batch: empty receivers / empty sources", flush=True)
    pb = RC.small_problem(nx=50, nz=40, npml=6, nt=140, fdorder=4, fsurf=1)
    pb.fwi = RC.full_window(pb)
    s0 = shots_of(pb)
    s_norec = dict(s0, rec=None)
    s_nosrc = dict(s0, src=dict(x=[], z=[], comp=[]), src_signal=np.zeros((0, pb.nt)))
    plan = plan_of(pb, "fp64", shots=[s0, s_norec, s_nosrc])
    plan.forward(energy=True, gradient=True)
    ref = O.forward(pb)
    for shot in (0, 1):
        _assert_fields(plan, ref, shot)
    assert np.array_equal(plan.records(0), ref["rec"])
    assert plan.records(1).shape == (0, pb.nt)
    assert not np.any(plan.records(2))
    for k in FIELDS:
        assert not np.any(plan.field(k, 2)), k
    obs = np.random.default_rng(4).standard_normal(ref["rec"].shape) * 1e-3
    mis = plan.misfit([obs, np.zeros((0, pb.nt)), obs])
    m0, res, _ = O.l2_adjoint_sources(pb.dt, ref["rec"], obs)
    assert abs(mis[0] - m0) <= 1e-12 * m0 and mis[1] == 0.0 and mis[2] > 0
    plan.adjoint(energy=True, gradient=True)
    a = O.adjoint(pb, pb.rec, res, ref["store"])
    g0 = plan.gradient(0)
    for got, k in zip(g0, GRADS):
        assert np.array_equal(got, a[k]), k
    for shot in (1, 2):
        assert not any(np.any(g) for g in plan.gradient(shot)), shot
    for got, want in zip(plan.gradient(-1), g0):
        assert np.array_equal(got, want)


@pytest.mark.parametrize("nt", [1, 2, 3, 4])
@pytest.mark.parametrize("tz", [8, -8])
def test_records_of_one_to_four_steps(nt, tz):
    """Receivers on and beside the first source, so that the records, and
    from three steps on the gradients, are not zero."""
    print(f"nt {nt} tile {tz}", flush=True)
    pb = RC.small_problem(nx=40, nz=30, npml=6, nt=nt, fdorder=4, fsurf=1)
    pb.src_signal = np.random.default_rng(nt).standard_normal(pb.src_signal.shape)
    xs, zs = int(pb.src["x"][0]), int(pb.src["z"][0])
    pb.rec = dict(x=[xs, xs + 1, xs, xs - 1], z=[zs, zs, zs + 1, zs - 1], comp=[2, 2, 1, 1], shot=[0] * 4)
    pb.fwi = RC.full_window(pb)
    obs = np.random.default_rng(10 + nt).standard_normal((len(pb.rec["x"]), nt))
    f = O.forward(pb)
    plan = plan_of(pb, "fp64", tz=tz)
    plan.forward(energy=True, gradient=True)
    assert np.array_equal(plan.records(), f["rec"])
    assert np.array_equal(plan.store(), f["store"])
    _assert_fields(plan, f)
    plan.misfit([obs])
    _, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"], obs)
    plan.adjoint(energy=True, gradient=True)
    a = O.adjoint(pb, pb.rec, res, f["store"])
    if nt >= 3:
        assert np.abs(a["grad_rho"]).max() > 0
    for got, k in zip(plan.gradient(0), GRADS):
        assert np.array_equal(got, a[k]), k
    _assert_fields(plan, a)


SMALL_GRIDS = [
    # (fdorder, npml, fsurf, nzt, nxt): the smallest each check in plan_create allows
    (2, 0, 1, 3, 3), (2, 0, 0, 3, 3), (4, 0, 1, 5, 5), (4, 0, 0, 5, 5),
    (2, 4, 1, 6, 10), (2, 4, 0, 10, 10), (4, 4, 1, 8, 12), (4, 4, 0, 12, 12),
]


@pytest.mark.parametrize("fdorder,npml,fsurf,nzt,nxt", SMALL_GRIDS)
@pytest.mark.parametrize("tz", [4, -16])
def test_smallest_grids(fdorder, npml, fsurf, nzt, nxt, tz):
    print(f"small grid o{fdorder} npml{npml} fs{fsurf} {nzt}x{nxt} tile {tz}", flush=True)
    pb = _grid_problem(nzt, nxt, fdorder, npml, fsurf, nt=40)
    ref = O.forward(pb)
    plan = plan_of(pb, "fp64", tz=tz)
    plan.forward(energy=True)
    assert np.array_equal(plan.records(), ref["rec"])
    _assert_fields(plan, ref)


@pytest.mark.parametrize("tz", [4, 8, 16, 32, -8, -16])
def test_grid_one_cell_past_a_tile(tz):
    """65 columns and 33 rows: the last tile in x and in z holds one cell."""
    print(f"65 x 33 tile {tz}", flush=True)
    pb = _grid_problem(33, 65, 4, 8, 0, nt=120, src_at=(16, 40))
    ref = O.forward(pb, prec="f32")
    plan = plan_of(pb, "fp32", tz=tz)
    plan.forward(energy=True)
    assert np.abs(ref["rec"]).max() > 0
    assert np.array_equal(plan.records(), ref["rec"].astype(np.float64))
    _assert_fields(plan, ref)


EDGE_WINDOWS = {
    "whole_grid_strided": lambda pb: dict(dt=4, dx=7, dz=5, x1=0, x2=pb.nxt, z1=0, z2=pb.nzt),
    "last_row_and_column": lambda pb: dict(dt=1, dx=1, dz=1, x1=pb.nxt - 1, x2=pb.nxt,
                                           z1=pb.nzt - 1, z2=pb.nzt),
    "ghost_rows": lambda pb: dict(dt=2, dx=3, dz=1, x1=0, x2=pb.nxt, z1=0, z2=pb.fpad + 1),
}


@pytest.mark.parametrize("win", list(EDGE_WINDOWS))
@pytest.mark.parametrize("tz", [8, -16])
def test_windows_on_the_grid_edge(win, tz):
    print(f"window {win} tile {tz}", flush=True)
    pb = RC.small_problem(nx=60, nz=36, npml=6, nt=150, fdorder=4, fsurf=1)
    pb.fwi = EDGE_WINDOWS[win](pb)
    obs = np.random.default_rng(2).standard_normal((len(pb.rec["x"]), pb.nt)) * 1e-3
    f = O.forward(pb)
    plan = plan_of(pb, "fp64", tz=tz)
    plan.forward(gradient=True)
    assert np.array_equal(plan.store(), f["store"])
    plan.misfit([obs])
    _, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"], obs)
    plan.adjoint(gradient=True)
    a = O.adjoint(pb, pb.rec, res, f["store"])
    for got, k in zip(plan.gradient(0), GRADS):
        assert np.array_equal(got, a[k]), k


def test_empty_window_gives_zero_store_and_gradient():
    """z1 == z2: no cell is stored or correlated; the single row the reference
    allocates stays zero."""
    print("empty window", flush=True)
    pb = RC.small_problem(nx=40, nz=30, npml=6, nt=60, fdorder=2, fsurf=0)
    pb.fwi = dict(dt=1, dx=1, dz=1, x1=pb.fpad, x2=pb.nxt - pb.fpad, z1=10, z2=10)
    obs = np.zeros((len(pb.rec["x"]), pb.nt))
    plan = plan_of(pb, "fp64")
    plan.forward(gradient=True)
    assert plan.store().shape == (5, pb.nt, 1, pb.nxt - 2 * pb.fpad + 1)
    assert not np.any(plan.store())
    plan.misfit([obs])
    plan.adjoint(gradient=True)
    assert not any(np.any(g) for g in plan.gradient(0))


def test_tile_heights_refused():
    """Tile heights outside the instantiated kernels are refused with a
    message; the 64 x 32 split tile is FP32 only."""
    from cuda_fwi_amd import psvfwi as P
    print("tile heights refused", flush=True)
    pb = _grid_problem(33, 65, 4, 8, 0, nt=10, src_at=(16, 40))
    for prec, bad in (("fp32", 12), ("fp32", -48), ("fp32", -4), ("fp32", -24), ("fp64", 32),
                      ("fp32", -128)):
        plan = plan_of(pb, prec)
        with pytest.raises(P.PsvfwiError, match="tile height"):
            plan.set_tile(bad)


def test_results_dropped_when_geometry_changes():
    """A full-store forward followed by the autotuner (which runs its candidates in the
    plan's own state, store and gradient buffers), or by a new window or batch, leaves no
    result readable: the getters refuse with PSVFWI_E_STATE instead of returning the
    autotuner's values or reading past buffers sized for the old geometry."""
    from cuda_fwi_amd import psvfwi as P
    print("stale results: autotune / set_window / set_shots", flush=True)
    pb = RC.small_problem(nx=40, nz=30, npml=6, nt=60, fdorder=4, fsurf=1, seed=2, nsrc=1, nrec=4)
    pb.fwi = RC.full_window(pb, dt=2)

    def state_error(fn):
        with pytest.raises(P.PsvfwiError) as e:
            fn()
        assert e.value.status == -5, str(e.value)

    plan = plan_of(pb, "fp32")
    plan.forward(energy=False, gradient=True)
    plan.store(0)                        # held
    plan.autotune(4)
    state_error(lambda: plan.store(0))
    state_error(lambda: plan.records(0))
    state_error(lambda: plan.field("vx", 0))

    plan.forward(energy=False, gradient=True)
    plan.misfit([np.zeros((4, pb.nt))])
    plan.adjoint(energy=False, gradient=True)
    plan.gradient(0)                     # held
    w = dict(pb.fwi, x2=pb.fwi["x2"] - 7, z2=pb.fwi["z2"] - 5)
    plan.set_window(w)
    state_error(lambda: plan.gradient(0))
    state_error(lambda: plan.store(0))
    state_error(lambda: plan.adjoint(energy=False, gradient=True))

    plan.set_window(pb.fwi)
    plan.forward(energy=False, gradient=True)
    plan.set_shots([shots_of(pb), shots_of(pb)])
    state_error(lambda: plan.store(1))
    state_error(lambda: plan.records(1))
    plan.close()
