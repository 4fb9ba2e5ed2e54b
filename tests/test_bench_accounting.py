"""The benchmark's accounting on the CPU: the algorithmic bytes behind the
roofline figure, re-derived here from the per-cell figures of the
measurement definition (5 state fields read and written plus 5 model arrays
read: 60 B per interior cell and step; CPML memory: 32 B per strip cell;
adjoint displacement: 16 B per cell; gradient correlation: 44 B and forward
store: 20 B per window cell of each decimated step), the workload sizes they
apply to, and the grouping of shots by store memory."""
import numpy as np

import bench
from cuda_fwi_amd import workloads as WL


def test_interior_cells_of_the_workloads():
    # padded grids of the measurement definition: 1000 x 500 with npml 40 -> 584 x 1084
    w = WL.c2()
    assert (w.nzt, w.nxt) == (584, 1084)
    assert w.interior_cells() == 626_400
    w4 = WL.c4(nshots=2)
    assert (w4.nzt, w4.nxt) == (684, 1584)
    # 201 x 201 with npml 20 and a free surface: no strip on top, fpad ghost rows instead
    w1 = WL.c1()
    assert (w1.nzt, w1.nxt) == (201 + 22 + 2, 245)


def test_algorithmic_bytes_per_step():
    w = WL.c2()
    n = w.interior_cells()
    rows, cols = w.nzt - 2 * w.fpad, w.nxt - 2 * w.fpad
    strip = 2 * w.npml * rows + 2 * w.npml * cols   # no free surface: four strips
    nwin, fwi_dt = w.nx * w.nz, 5
    kb = bench.alg_bytes(w, nwin, fwi_dt, 4)
    fwd = kb["forward_stress"] + kb["forward_velocity"]
    adj = kb["adjoint_stress"] + kb["adjoint_velocity"]
    assert fwd == 60 * n + 32 * strip + 20 * nwin / fwi_dt
    assert adj == 60 * n + 16 * n + 32 * strip + 44 * nwin / fwi_dt
    # the dominant kernel's bytes per launch that the benchmark line reports
    assert kb["adjoint_stress"] == 39_097_600
    # the velocity kernel moves the velocities (16 B) and the two densities (8 B)
    assert kb["forward_velocity"] == 24 * n + 16 * strip
    # FP64 doubles every term
    kb8 = bench.alg_bytes(w, nwin, fwi_dt, 8)
    for k in kb:
        assert kb8[k] == 2 * kb[k]
    # a free surface removes the top strip
    w1 = WL.c1()
    r1, c1 = w1.nzt - 2 * w1.fpad, w1.nxt - 2 * w1.fpad
    kb1 = bench.alg_bytes(w1, w1.nx * w1.nz, 1, 4)
    assert kb1["forward_velocity"] == 24 * w1.interior_cells() + 16 * (2 * w1.npml * r1 + w1.npml * c1)


def test_store_groups_fit_the_budget_and_follow_nt():
    w = WL.c4(nshots=64)
    nwin, fwi_dt, esize = w.nx * w.nz, 5, 4
    budget = 200e9
    groups = bench._groups_for_store(w, nwin, fwi_dt, esize, budget)
    flat = [i for g in groups for i in g]
    assert sorted(flat) == list(range(64))
    nts = [w.shots[i]["nt"] for i in flat]
    assert nts == sorted(nts, reverse=True)
    need = lambda i: 5 * nwin * (1 + (w.shots[i]["nt"] - 1) // fwi_dt) * esize
    for g in groups:
        assert sum(need(i) for i in g) <= budget or len(g) == 1
    # each group but the last is full: the next shot would not have fitted
    for g, nxt in zip(groups, groups[1:]):
        assert sum(need(i) for i in g) + need(nxt[0]) > budget
