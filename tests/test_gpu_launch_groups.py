"""Launch groups (psvfwi_set_launch_group): a batch whose time loops run a
few shots at a time gives the same traces, fields, energy weights, snapshots
and gradients bit for bit as the whole batch stepped together, in every
gradient store mode and with graph-captured time loops."""
import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from cuda_fwi_amd import psvfwi as P
from gpucase import plan_of

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vz", "sxx", "szx", "szz")
NTS = (150, 231, 90, 230, 177)


def _shots(base, comps=(2, 4, 2, 1, 2), seed=9):
    rng = np.random.default_rng(seed)
    shots, obs = [], []
    for i, nt in enumerate(NTS):
        nrec = int(rng.integers(3, 9))
        xs = int(rng.integers(base.ppad + 5, base.nxt - base.ppad - 5))
        sig = O.create_wavelet(nt, base.dt, 1.0 + i, 25.0, 0.0)[None]
        rec = dict(x=list(np.linspace(base.ppad, base.nxt - base.ppad - 1, nrec).astype(int)),
                   z=[base.fpad + 1] * nrec, comp=[2] * (nrec - 1) + [1])
        shots.append(dict(nt=nt, src=dict(x=[xs], z=[base.fpad + 4], comp=[comps[i]]), src_signal=sig,
                          rec=rec))
        obs.append(rng.standard_normal((nrec, nt)) * 1e-3)
    return shots, obs


def _run(base, shots, obs, group, store="full", interval=0, gradient="reference", graph=0, snaps=0,
         prec="fp32", tz=8, segment=40):
    plan = plan_of(base, prec, shots=shots, tz=tz)
    plan.set_window(RC.full_window(base, dt=2))
    plan.set_store_mode(store, interval)
    if gradient == "discrete":
        plan.set_gradient_mode("discrete", segment)
    plan.set_graph_steps(graph)
    if snaps:
        plan.set_snapshots(snaps, "vz")
    plan.set_launch_group(group)
    assert plan.launch_group() == group
    n = len(shots)
    plan.forward(energy=True, gradient=True)
    out = [plan.records(i) for i in range(n)]
    if gradient != "discrete":
        out += [plan.field("We", i) for i in range(n)]
    if store != "random_boundary" and gradient != "discrete":
        out += [plan.field(k, i) for i in range(n) for k in FIELDS]
    if snaps:
        out += [s for i in range(n) for s in plan.snapshots(i)]
    out += list(plan.misfit(obs))
    plan.adjoint(energy=False, gradient=True)
    if gradient == "discrete":
        out += [g for i in range(n) for g in plan.model_gradient(i)]
    else:
        out += [g for i in range(n) for g in plan.gradient(i)]
    return out


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("group", [1, 2, 3, 4])
def test_launch_groups_full_store(group):
    print(f"launch groups of {group}, full store", flush=True)
    base = RC.small_problem(nx=80, nz=48, npml=8, nt=10, fdorder=4, fsurf=1, hetero=True)
    shots, obs = _shots(base)
    ref = _run(base, shots, obs, 0)
    assert np.abs(ref[0]).max() > 0
    _same(_run(base, shots, obs, group), ref)


def test_launch_groups_checkpoint_snapshots_graphs():
    print("launch groups: checkpoints, snapshots, graph-captured loops", flush=True)
    base = RC.small_problem(nx=80, nz=48, npml=8, nt=10, fdorder=4, fsurf=0, hetero=True)
    shots, obs = _shots(base)
    ref = _run(base, shots, obs, 0, store="checkpoint", interval=37, snaps=25)
    _same(_run(base, shots, obs, 2, store="checkpoint", interval=37, snaps=25), ref)
    ref = _run(base, shots, obs, 0, graph=8)
    _same(_run(base, shots, obs, 3, graph=8), ref)
    _same(_run(base, shots, obs, 2), ref)


def test_launch_groups_discrete_gradient():
    print("launch groups: discrete gradient", flush=True)
    base = RC.small_problem(nx=70, nz=40, npml=6, nt=10, fdorder=4, fsurf=1, hetero=True)
    shots, obs = _shots(base, comps=(2, 2, 1, 2, 2))
    ref = _run(base, shots, obs, 0, gradient="discrete", prec="fp64")
    assert np.abs(ref[-1]).max() > 0
    _same(_run(base, shots, obs, 2, gradient="discrete", prec="fp64"), ref)
    # segments of one step: every state kept by the forward, no replay
    _same(_run(base, shots, obs, 3, gradient="discrete", prec="fp64", segment=1), ref)


def test_launch_groups_random_boundary():
    print("launch groups: random-boundary reconstruction", flush=True)
    base = RC.small_problem(nx=90, nz=60, npml=0, nt=10, fdorder=4, fsurf=0, hetero=True)
    base.lam, base.mu, base.rho = P.random_boundary(base.lam, base.mu, base.rho, 12, 0, seed=5,
                                                    max_increase=0.4)
    base.mu_zx, base.rho_zp, base.rho_xp = O.average_parameters(base.mu, base.rho)
    shots, obs = _shots(base, comps=(2, 4, 2, 2, 1))
    ref = _run(base, shots, obs, 0, store="random_boundary", prec="fp64")
    _same(_run(base, shots, obs, 2, store="random_boundary", prec="fp64"), ref)


def test_autotune_launch_group_and_fused_ignores_it():
    print("launch groups: autotune, fused step", flush=True)
    base = RC.small_problem(nx=80, nz=48, npml=8, nt=10, fdorder=4, fsurf=1, hetero=True)
    shots, obs = _shots(base)
    ref = _run(base, shots, obs, 0)
    plan = plan_of(base, "fp32", shots=shots)
    plan.set_window(RC.full_window(base, dt=2))
    plan.autotune(20)
    assert plan.launch_group() in (0, 16, 12, 8, 6, 4, 2)
    plan.forward(energy=True, gradient=True)
    rec = [plan.records(i) for i in range(len(shots))]
    _same(rec, ref[:len(shots)])
    # the fused step runs the batch as one group whatever the setting
    fa = _run(base, shots, obs, 0, tz=-8)
    _same(_run(base, shots, obs, 2, tz=-8), fa)
