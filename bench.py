#!/usr/bin/env python3
"""Benchmark of the MI355X P-SV FWI hot path: forward + adjoint + gradient.

Default workload: the batched one the metric is quoted on.  64 shots on a
1500 x 600 model (dx 5 m, order 4 Holberg, 40-cell CPML, seeded layers with a
dipping fault), nt uniform in 3000-8000, 50-400 receivers per shot, FP32
kernels, gradient window = the model, time decimation 5.  The shots run in
store groups whose decimated forward stores fit in HBM together; one "step"
= one group's batched forward (with the store) + FP64 misfit/residual +
adjoint with the fused gradient correlation, on device-resident models and
observed traces; steps cycle through the groups.  The same shots are then
run one shot per launch for the batched / per-shot comparison.  Extra keys:
the FP64 kernels' rate, the 1000 x 500 single-shot workload, the CPU oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c2|c3mem|c5]

Prints one JSON line.  Gcell-updates/s = 2 * interior cells * sum of nt / time
(forward and adjoint counted, checkpoint recomputation not counted).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s fwd+adj+grad, batched vs per-shot shots/s, % HBM peak"
KERNEL_NAMES = {"forward_stress": "psv_stress_kernel (forward)",
                "forward_velocity": "psv_velocity_kernel (forward)",
                "adjoint_stress": "psv_stress_kernel (adjoint: displacement + gradient correlation)",
                "adjoint_velocity": "psv_velocity_kernel (adjoint)"}
HBM_PEAK_GBS = 8000.0


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def make_plan(w, precision, window, store_mode, interval, tile, shots=None):
    """store_mode "random_boundary": the same padded grid with no CPML; its
    npml outer cells hold a random boundary layer (vp, vs up to +40 %) in
    both the true and the starting model, and the adjoint rebuilds the
    forward state by running the scheme backward."""
    from cuda_fwi_amd import psvfwi as P
    rb = store_mode == "random_boundary"
    vp0, vs0, rho0 = w.background
    lam, mu, rho = w.padded(vp0, vs0, rho0)
    if rb:
        lam, mu, rho = P.random_boundary(lam, mu, rho, w.npml, w.fsurf, seed=11, max_increase=0.4)
    mu_zx, rho_zp, rho_xp = P.average_parameters(mu, rho)
    hc = P.holberg(w.fdorder, w.max_relative_error)
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=0 if rb else w.npml, fsurf=w.fsurf, dt=w.dt,
                  dx=w.dx, dz=w.dx, hc=hc, precision=precision)
    vmax = float(np.max(w.vp))
    if not rb:
        plan.set_pml(P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, vmax, w.dt))
    plan.set_window(window)
    plan.set_store_mode(store_mode, interval)
    shots = w.shots if shots is None else shots
    plan.set_shots([w.padded_shot(s, P.ricker(s["nt"], w.dt, s["amp"], w.fc)[None]) for s in shots])
    # observed data from the true model, then switch to the starting model
    lt, mt, rt = w.padded()
    if rb:
        lt, mt, rt = P.random_boundary(lt, mt, rt, w.npml, w.fsurf, seed=11, max_increase=0.4)
    plan.set_model(lt, mt, *P.average_parameters(mt, rt))
    plan.forward(energy=False)
    plan.set_observed([plan.records(i) for i in range(len(shots))])
    plan.set_model(lam, mu, mu_zx, rho_zp, rho_xp)
    if tile != 0:
        plan.set_tile(tile)
        plan.tile = tile
    else:
        plan.tile = plan.autotune(min(200, w.nt))
    return plan


def alg_bytes(w, nwin, fwi_dt, esize):
    """Algorithmic HBM bytes per launch of each step kernel (one shot).

    The whole-step figures are the 60 B per interior cell of a wavefield
    update (5 fields read and written: 40 B, 5 model arrays read: 20 B),
    32 B per CPML strip cell (4 memory arrays read and written), 16 B per
    cell of adjoint displacement, 44 B per window cell per correlated step
    and 20 B per window cell per stored step (amortised over fwi_dt).  They
    are split between the two kernels of a step by what each one updates:
    stress kernel = stresses read+written (24) + lambda, mu, mu_zx (12) =
    36 B/cell, half of the CPML memory (16 B/strip cell), the displacement,
    the correlation and the store; velocity kernel = velocities read+written
    (16) + the two densities (8) = 24 B/cell and the other 16 B/strip cell.
    The reads of the other field group (velocities in the stress kernel,
    stresses in the velocity kernel) are not algorithmic bytes.  FP32
    figures, scaled by esize / 4."""
    s = esize / 4.0
    n = w.interior_cells()
    rows, cols = w.nzt - 2 * w.fpad, w.nxt - 2 * w.fpad
    strip = 2 * w.npml * rows + (1 if w.fsurf else 2) * w.npml * cols
    return {
        "forward_stress": s * (36 * n + 16 * strip + 20 * nwin / fwi_dt),
        "forward_velocity": s * (24 * n + 16 * strip),
        "adjoint_stress": s * (52 * n + 16 * strip + 44 * nwin / fwi_dt),
        "adjoint_velocity": s * (24 * n + 16 * strip),
    }


def cpu_baseline(w, window, nt_sample, threaded=False):
    """The oracle (FP64 CPU restatement of the reference path) on the same
    grid, model and geometry, for nt_sample steps: single-threaded (the
    reference CPU path has no threads), or its row-parallel OpenMP build over
    the host cores OMP_NUM_THREADS allows (same results)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import psv_oracle as O
    from cuda_fwi_amd import psvfwi as P
    vp0, vs0, rho0 = w.background
    lam, mu, rho = w.padded(vp0, vs0, rho0)
    s = w.shots[0]
    ps = w.padded_shot(s, P.ricker(nt_sample, w.dt, s["amp"], w.fc)[None])
    src = dict(ps["src"], shot=[0] * len(ps["src"]["x"]))
    rec = dict(ps["rec"], shot=[0] * len(ps["rec"]["x"]))
    pml = O.cpml_pro(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt)
    pb = O.Problem(nt=nt_sample, nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, dt=w.dt,
                   dx=w.dx, dz=w.dx, fsurf=w.fsurf, hc=O.holbergcoeff(w.fdorder, w.max_relative_error),
                   lam=lam, mu=mu, rho=rho, pml=pml, src=src, src_signal=ps["src_signal"], rec=rec,
                   fwi=window)
    obs = np.zeros((len(rec["x"]), nt_sample))
    threads = 1
    if threaded:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        os.environ["OMP_NUM_THREADS"] = str(threads)
    O.set_threaded(threaded)
    try:
        t0 = time.perf_counter()
        f = O.forward(pb)
        _, res, _ = O.l2_adjoint_sources(w.dt, f["rec"], obs)
        O.adjoint(pb, rec, res, f["store"])
        t = time.perf_counter() - t0
    finally:
        O.set_threaded(False)
    return dict(value=2.0 * w.interior_cells() * nt_sample / t / 1e9, unit="Gcell-updates/s", cores=threads,
                kind="port",
                sample=f"oracle FP64, {threads} thread{'s' if threads > 1 else ''}: same grid/model/geometry, "
                       f"forward+misfit+adjoint+gradient over nt={nt_sample} of {w.nt} steps ({t:.1f} s)")


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (torch elementwise copy kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "copy_ms": ms,
            "copy_GBps": 2 * n * 4 / ms / 1e6}


def run_c2(args):
    from cuda_fwi_amd import workloads as WL
    w = WL.c2()
    if args.nt:
        w.nt = args.nt
        w.shots[0]["nt"] = args.nt
    fwi_dt = args.fwi_dt
    window = w.model_window(dt=fwi_dt)
    plan = make_plan(w, args.precision, window, args.store, args.interval, args.tile)
    plan.set_graph_steps(args.graph_steps)
    _log("setup done (observed data computed)")
    esize = 8 if args.precision in ("fp64", "fp32x2") else 4

    def one_pass():
        plan.forward(energy=False, gradient=True)
        t_f = plan.timing() + plan.kernel_timing("forward")
        plan.misfit()
        plan.adjoint(energy=False, gradient=True)
        t_a = plan.timing() + plan.kernel_timing("adjoint")
        return t_f, t_a

    for i in range(args.warmup):
        one_pass()
        _log(f"warmup pass {i} done")
    _sync()
    t0 = time.perf_counter()
    tf_ms = ta_ms = 0.0
    nf = na = 0
    ksum = {"forward_stress": [], "forward_velocity": [], "adjoint_stress": [], "adjoint_velocity": []}
    for i in range(args.steps):
        (fms, fl, fs_us, fv_us, fn), (ams, al, as_us, av_us, an) = one_pass()
        _log(f"timed pass {i}: forward {fms:.1f} ms, adjoint {ams:.1f} ms; stress/velocity kernel "
             f"us: forward {fs_us:.2f}/{fv_us:.2f}, adjoint {as_us:.2f}/{av_us:.2f}")
        tf_ms += fms
        ta_ms += ams
        nf += fl
        na += al
        ksum["forward_stress"].append((fs_us, fn))
        ksum["forward_velocity"].append((fv_us, fn))
        ksum["adjoint_stress"].append((as_us, an))
        ksum["adjoint_velocity"].append((av_us, an))
    _sync()
    T = time.perf_counter() - t0

    nwin = w.nx * w.nz
    kb_all = alg_bytes(w, nwin, fwi_dt, esize)
    # Per-kernel average duration: the loop's time per step (HIP events around
    # the whole time loop, on the plan's stream) split between the two kernels
    # of a step in the ratio of their sampled event-bracketed durations (a
    # single bracketed kernel also carries the event's own dispatch overhead).
    sampled = {k: sum(u * n for u, n in v) / max(sum(n for _, n in v), 1) for k, v in ksum.items()}
    kus = {}
    for kind, ms, n in (("forward", tf_ms, nf), ("adjoint", ta_ms, na)):
        step_us = ms * 1e3 / n
        a, b = sampled[kind + "_stress"], sampled[kind + "_velocity"]
        kus[kind + "_stress"] = step_us * a / (a + b)
        kus[kind + "_velocity"] = step_us * b / (a + b)
    launches = {"forward_stress": nf, "forward_velocity": nf, "adjoint_stress": na, "adjoint_velocity": na}
    # dominant kernel: the largest share of the timed region (mean duration x launches)
    kname = max(kus, key=lambda k: kus[k] * launches[k])
    kb = kb_all[kname]
    per_launch_s = kus[kname] / 1e6
    achieved = kb / per_launch_s / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "pmc_traffic_c2.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get(kname)
    cells = w.interior_cells()
    value = 2.0 * cells * w.nt * args.steps / T / 1e9
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(T / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if esize == 4 else "f64",
        "data": "synthetic: seeded 8-layer model with 3% Gaussian perturbation (observed) and "
                "the unperturbed layers (synthetic)",
        "config": {"workload": "1000x500 dx5 order-4 npml40, nt 5000, dt 0.4 ms, 1 shot, "
                               "400 receivers, fwd+adj+grad",
                   "padded_grid": [w.nzt, w.nxt], "interior_cells": cells,
                   "gradient_store": {"full": "full", "random_boundary": "random-boundary reconstruction "
                                      "(no CPML: the outer 40 cells hold a random layer)"}.get(
                                          args.store, f"checkpoint every {args.interval}"),
                   "fwi_time_decimation": fwi_dt,
                   "tile": [64, abs(plan.tile)],
                   "step_kernel": "fused (one launch per step)" if plan.tile < 0 else "split (stress, velocity)",
                   "tile_choice": "autotuned" if args.tile == 0 else "fixed",
                   "graph_steps": args.graph_steps},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[kname],
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": int(kb), "avg_launch_us": round(kus[kname], 2),
                     "launches": launches[kname],
                     "kernel_us": {k: round(v, 2) for k, v in kus.items()},
                     "sampled_us_per_launch": {k: round(v, 2) for k, v in sampled.items()},
                     "duration_note": "avg_launch_us: the loop time per step (HIP events around each time "
                                      "loop) split between the two kernels in the ratio of their sampled "
                                      "durations; sampled_us_per_launch: event-bracketed durations of every "
                                      "64th launch, each with one event's dispatch overhead",
                     "step_alg_GBps": round(60.0 * (esize / 4) * cells * w.nt * 2 * args.steps
                                            / (tf_ms + ta_ms) * 1e3 / 1e9, 1)},
        "forward_ms_per_step": round(tf_ms / args.steps, 2),
        "adjoint_ms_per_step": round(ta_ms / args.steps, 2),
    }
    if not args.no_cpu_baseline:
        _log("cpu baseline (oracle) starting")
        line["cpu_baseline"] = cpu_baseline(w, window, args.cpu_nt)
        _log("cpu baseline (threaded oracle) starting")
        line["cpu_baseline_threads"] = cpu_baseline(w, window, args.cpu_nt, threaded=True)
    return line


def run_c2_discrete(args):
    """The default workload with the discrete gradient (an extension: the
    exact misfit gradient, checkpoints every --interval steps, each segment
    replayed and the transposed step run back through it).  Reports the
    throughput of forward + misfit + discrete adjoint and the split between
    forward and adjoint."""
    from cuda_fwi_amd import workloads as WL
    w = WL.c2()
    if args.nt:
        w.nt = args.nt
        w.shots[0]["nt"] = args.nt
    plan = make_plan(w, args.precision, None, "full", args.interval, args.tile)
    plan.set_gradient_mode("discrete", args.interval)
    _log("setup done (observed data computed)")

    def one_pass():
        plan.forward(energy=True, gradient=True)
        f = plan.timing()[0]
        plan.misfit()
        plan.adjoint(energy=False, gradient=True)
        return f, plan.timing()[0]

    for i in range(args.warmup):
        one_pass()
        _log(f"warmup pass {i} done")
    _sync()
    t0 = time.perf_counter()
    tf = ta = 0.0
    for i in range(args.steps):
        f, a = one_pass()
        _log(f"timed pass {i}: forward {f:.1f} ms, discrete adjoint {a:.1f} ms")
        tf += f
        ta += a
    _sync()
    T = time.perf_counter() - t0
    cells = w.interior_cells()
    esize = 8 if args.precision in ("fp64", "fp32x2") else 4
    return {
        "metric": METRIC, "value": round(2.0 * cells * w.nt * args.steps / T / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(T / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if esize == 4 else "f64",
        "data": "synthetic: seeded 8-layer model with 3% Gaussian perturbation (observed) and "
                "the unperturbed layers (synthetic)",
        "config": {"workload": f"1000x500 dx5 order-4 npml40, nt {w.nt}, dt 0.4 ms, 1 shot, 400 receivers, "
                               "fwd+adj+grad with the discrete (exact) gradient on the whole grid",
                   "checkpoint_interval": args.interval, "tile": [64, abs(plan.tile)]},
        "forward_ms_per_step": round(tf / args.steps, 2),
        "adjoint_ms_per_step": round(ta / args.steps, 2),
    }


def run_c5(args):
    """FWI on the layered / faulted / salt-body model: observed traces from the
    true model, start = the truth smoothed over 150 m, --iters iterations of
    cuda_fwi_amd.fwi.Inversion (batched forward, FP64 misfit, discrete
    adjoint, the reference's bracketing step search with batched trials).
    Sizes default to the 2400 x 800, 128-shot, nt 6000 configuration scaled by
    --c5-scale (lengths and shot count; nt by the same factor)."""
    from cuda_fwi_amd import psvfwi as P
    from cuda_fwi_amd import workloads as WL
    from cuda_fwi_amd.fwi import Inversion
    sc = args.c5_scale
    nx, nz = int(round(2400 * sc)), int(round(800 * sc))
    nshots, nrec, nt = max(1, int(round(128 * sc))), max(2, int(round(600 * sc))), int(round(6000 * sc))
    w = WL.c5(nshots=nshots, nrec=nrec, nt=nt, nx=nx, nz=nz)
    hc = P.holberg(w.fdorder, w.max_relative_error)
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                  dz=w.dx, hc=hc, precision=args.precision)
    plan.set_pml(P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt))
    shots = [w.padded_shot(s, P.ricker(s["nt"], w.dt, s["amp"], w.fc)[None]) for s in w.shots]
    truth = w.padded()
    plan.set_model(truth[0], truth[1], *P.average_parameters(truth[1], truth[2]))
    g = args.c5_group or nshots
    if args.tile:
        plan.set_tile(args.tile)
        plan.tile = args.tile
    else:
        plan.set_shots(shots[:g])
        plan.tile = plan.autotune(min(100, nt))
    if args.launch_group >= 0:
        plan.set_launch_group(args.launch_group)
    _log(f"tile {plan.tile}, shot order {plan.shot_order()}, launch group {plan.launch_group()}")
    observed = []
    for i in range(0, nshots, g):
        plan.set_shots(shots[i:i + g])
        plan.forward(energy=False)
        observed += [plan.records(k) for k in range(len(shots[i:i + g]))]
        _log(f"observed traces of shots {i}-{min(i + g, nshots) - 1}")
    _log(f"c5 at scale {sc}: {nx} x {nz}, {nshots} shots x {nrec} receivers, nt {nt}, groups of {g}; "
         "observed data done")
    start = w.padded(*w.background)
    win = (w.zoff, w.zoff + nz, w.ppad, w.ppad + nx)
    inv = Inversion(plan, shots, observed, start, win, segment=args.interval, mute=args.c5_mute,
                    group=args.c5_group or None, smooth=args.c5_smooth * sc, precond=args.c5_precond,
                    gradient=args.c5_gradient, fwi_dt=args.fwi_dt, taper=args.c5_taper,
                    depth_power=args.c5_depth_power)
    err0 = inv.model_error(truth)
    # vp errors after low-passing both models: 50 m and 150 m (the start model's smoothing) at full scale
    sigmas = (10.0 * sc, 30.0 * sc)
    low0 = [inv.model_error(truth, sig) for sig in sigmas]
    # a run longer than one job continues from the state a previous job saved:
    # the FP64 model over the window (the padding is edge replication of it),
    # the last step length and the per-iteration history
    prev = {"misfit": [], "step": [], "trials": [], "seconds": [], "model_error": []}
    if args.c5_resume:
        st = np.load(args.c5_resume, allow_pickle=False)
        z1, z2, x1, x2 = win
        for k in ("lam", "mu", "rho"):
            m = inv.m[k].copy()
            m[z1:z2, x1:x2] = st[k]
            inv.m[k] = np.pad(m[z1:z2, x1:x2], ((z1, m.shape[0] - z2), (x1, m.shape[1] - x2)), mode="edge")
        inv.step = float(st["step"])
        prev = json.loads(str(st["history"]))
        _log(f"resumed after {len(prev['misfit'])} iterations, step {inv.step:.4g}, "
             f"vp error {inv.model_error(truth):.5f}")
    _sync()
    t0 = time.perf_counter()
    h = inv.run(args.iters, truth=truth, log=_log, first=len(prev["misfit"]))
    _sync()
    T = time.perf_counter() - t0
    for k in prev:
        prev[k] = list(prev[k]) + [float(x) for x in getattr(h, k)]
    if args.c5_save:
        z1, z2, x1, x2 = win
        np.savez(args.c5_save, lam=inv.m["lam"][z1:z2, x1:x2], mu=inv.m["mu"][z1:z2, x1:x2],
                 rho=inv.m["rho"][z1:z2, x1:x2], step=np.float64(inv.step), history=json.dumps(prev))
        _log(f"state after {len(prev['misfit'])} iterations saved to {args.c5_save}")
    cells = w.interior_cells()
    upd = 2.0 * cells * nt * nshots * args.iters
    return {
        "metric": METRIC, "value": round(upd / T / 1e9, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.iters, "warmup": 0, "ms_per_step": round(T / max(args.iters, 1) * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic: 10 layers, two faults and a vp 4500 irregular body (seed 7); start model = "
                "truth smoothed over 150 m (scaled)",
        "config": {"workload": f"FWI {nx}x{nz} dx5 order-4 npml40, {nshots} shots x {nrec} receivers, nt {nt}, "
                               f"8 Hz, {args.iters} iterations, "
                               + ("discrete gradient" if args.c5_gradient == "discrete" else
                                  f"reference correlation gradient (fwi_dt {args.fwi_dt})")
                               + ", checkpoints every "
                               f"{args.interval} steps, top {args.c5_mute} model rows not updated"
                               + (f", direction smoothed over {args.c5_smooth:g} cells (full scale)"
                                  if args.c5_smooth > 0 else "")
                               + (", preconditioner sqrt(Ws Wr)" if args.c5_precond == "source_receiver" else "")
                               + (f", cosine taper over {args.c5_taper} rows" if args.c5_taper > 0 else "")
                               + (f", depth gain power {args.c5_depth_power:g}" if args.c5_depth_power > 0 else ""),
                   "scale_of_2400x800_128_shots": sc, "tile": plan.tile,
                   "launch_group": plan.launch_group() or "whole batch"},
        "fwi": {"misfit": prev["misfit"], "vp_rel_error_start": err0, "vp_rel_error": prev["model_error"],
                "vp_rel_error_lowpass": {f"sigma {sig:g} cells": {"start": a, "end": inv.model_error(truth, sig)}
                                         for sig, a in zip(sigmas, low0)},
                "direction_peak": inv.diag,
                "step": prev["step"], "trial_forwards": prev["trials"],
                "seconds": [round(x, 2) for x in prev["seconds"]], "iterations_this_run": args.iters,
                "note": "value counts forward + adjoint cell updates of each iteration; trial forwards "
                        "and replays are not counted"},
    }


def _groups_for_store(w, window_cells, fwi_dt, esize, budget_bytes):
    """Consecutive groups of shots (sorted by decreasing nt) whose full
    decimated stores fit in budget_bytes together."""
    order = sorted(range(len(w.shots)), key=lambda i: -w.shots[i]["nt"])
    groups, cur, used = [], [], 0
    for i in order:
        need = 5 * window_cells * (1 + (w.shots[i]["nt"] - 1) // fwi_dt) * esize
        if cur and used + need > budget_bytes:
            groups.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += need
    if cur:
        groups.append(cur)
    return groups


def _shot_descs(w, ids):
    """Padded shot descriptors with their Ricker signals, built once on the host."""
    from cuda_fwi_amd import psvfwi as P
    return [w.padded_shot(w.shots[i], P.ricker(w.shots[i]["nt"], w.dt, w.shots[i]["amp"], w.fc)[None])
            for i in ids]


def run_c4(args):
    """64 shots of 1500 x 600 cells with uneven nt (3000-8000) and receiver counts (50-400):
    forward + FP64 misfit + adjoint + gradient, batched, against the same kernels one shot
    per launch.

    The full decimated forward store of all 64 shots (1.5 TB at fwi_dt 5) exceeds the 288 GB
    of HBM, so the set runs in store groups (shots sorted by nt, each group's stores within
    --store-budget-gb).  One benchmark step is one group's batched fwd + misfit + adj +
    gradient; successive steps cycle through the groups, so --steps at least the number of
    groups covers all 64 shots.  The shot tables of a group are uploaded inside its step, as
    they are for each shot of the per-shot comparison.  With --store checkpoint all 64 shots
    form one group (checkpoints every --interval steps, recomputation not counted)."""
    from cuda_fwi_amd import psvfwi as P
    from cuda_fwi_amd import workloads as WL
    nx, nz = (int(v) for v in args.c4_grid.lower().split("x"))
    w = WL.c4(nshots=args.nshots, nx=nx, nz=nz)
    if args.nt:
        for sh in w.shots:
            sh["nt"] = min(sh["nt"], args.nt)
        w.nt = max(sh["nt"] for sh in w.shots)
    fwi_dt = args.fwi_dt
    window = w.model_window(dt=fwi_dt)
    prec = args.precision
    esize = 8 if prec in ("fp64", "fp32x2") else 4
    nwin = w.nx * w.nz
    if args.store == "full":
        groups = _groups_for_store(w, nwin, fwi_dt, esize, args.store_budget_gb * 1e9)
    else:
        groups = [sorted(range(len(w.shots)), key=lambda i: -w.shots[i]["nt"])]
    _log(f"{len(groups)} store group(s) of shots: {[len(g) for g in groups]}")
    descs = {i: d for i, d in zip(range(len(w.shots)), _shot_descs(w, range(len(w.shots))))}
    lt, mt, rt = w.padded()
    lam, mu, rho = w.padded(*w.background)
    avg0 = P.average_parameters(mu, rho)
    hc = P.holberg(w.fdorder, w.max_relative_error)
    pml = P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt)

    def new_plan(precision=prec):
        plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt,
                      dx=w.dx, dz=w.dx, hc=hc, precision=precision)
        plan.set_pml(pml)
        plan.set_window(window)
        plan.set_store_mode(args.store, args.interval)
        plan.set_graph_steps(args.graph_steps)
        return plan

    # observed traces from the true model (untimed)
    obs = {}
    plan = new_plan()
    plan.set_model(lt, mt, *P.average_parameters(mt, rt))
    for g in groups:
        plan.set_shots([descs[i] for i in g])
        plan.forward(energy=False)
        for k, i in enumerate(g):
            obs[i] = plan.records(k)
    plan.set_model(lam, mu, *avg0)
    # kernel, tile, shot order and launch group tuned once on the first (longest) group
    plan.set_shots([descs[i] for i in groups[0]])
    tile = args.tile if args.tile != 0 else plan.autotune(min(args.tune_steps, w.nt))
    plan.set_tile(tile)
    if args.shot_order >= 0:
        plan.set_shot_order(args.shot_order)
    band = plan.shot_order()
    if args.launch_group >= 0:
        plan.set_launch_group(args.launch_group)
    lgroup = plan.launch_group()
    _log(f"observed traces done; tile {tile}, shot order {band}, launch group {lgroup}")

    kinds = ("forward_stress", "forward_velocity", "adjoint_stress", "adjoint_velocity")

    def group_step(g):
        """One step: the group's fwd + misfit + adj + gradient; returns loop ms, launches and
        sampled kernel us of the forward and the adjoint."""
        plan.set_shots([descs[i] for i in g])
        plan.set_tile(tile)
        plan.set_shot_order(band)
        plan.set_observed([obs[i] for i in g])
        plan.forward(energy=False, gradient=True)
        tf = plan.timing() + plan.kernel_timing("forward")
        plan.misfit()
        plan.adjoint(energy=False, gradient=True)
        ta = plan.timing() + plan.kernel_timing("adjoint")
        return tf, ta

    # setup: one pass of the group with the largest store, so that the store allocation (which
    # only grows) is made before the timed steps
    def store_need(g):
        return len(g) * (1 + (max(w.shots[i]["nt"] for i in g) - 1) // fwi_dt)
    group_step(max(groups, key=store_need))
    for i in range(args.warmup):
        group_step(groups[i % len(groups)])
        _log(f"warmup step {i} done")
    _sync()
    t0 = time.perf_counter()
    shots_done = cell_steps = 0
    loop_ms = {"forward": 0.0, "adjoint": 0.0}
    nlaunch = {"forward": 0, "adjoint": 0}
    samples = {k: [] for k in kinds}
    alg_kernel = {k: 0.0 for k in kinds}
    kb = alg_bytes(w, nwin, fwi_dt, esize)
    for i in range(args.steps):
        g = groups[i % len(groups)]
        ts = time.perf_counter()
        (fms, fl, fs_us, fv_us, fn), (ams, al, as_us, av_us, an) = group_step(g)
        nts = sum(w.shots[j]["nt"] for j in g)
        shots_done += len(g)
        cell_steps += nts * w.interior_cells()
        loop_ms["forward"] += fms
        loop_ms["adjoint"] += ams
        nlaunch["forward"] += fl
        nlaunch["adjoint"] += al
        for k, v, n in zip(kinds, (fs_us, fv_us, as_us, av_us), (fn, fn, an, an)):
            samples[k].append((v, n))
            alg_kernel[k] += kb[k] * nts
        _log(f"timed step {i}: group of {len(g)} shots, forward {fms:.0f} ms, adjoint {ams:.0f} ms, "
             f"wall {(time.perf_counter() - ts) * 1e3:.0f} ms")
    _sync()
    T = time.perf_counter() - t0

    # dominant kernel: each kind's share of its loop time in the ratio of the sampled
    # event-bracketed durations; per launch = that time / launches, algorithmic bytes per
    # launch = the kernel's bytes summed over all shot-steps / launches
    # the sampled durations themselves: every 64th launch of each timed loop bracketed by
    # HIP events on the plan's stream, averaged over the samples of all timed steps
    direct = {k: sum(v * n for v, n in samples[k]) / max(sum(n for _, n in samples[k]), 1) for k in kinds}
    nsamp = {k: sum(n for _, n in samples[k]) for k in kinds}
    kus_total = {}
    for kind in ("forward", "adjoint"):
        a = direct[kind + "_stress"]
        b = direct[kind + "_velocity"]
        kus_total[kind + "_stress"] = loop_ms[kind] * 1e3 * a / (a + b)
        kus_total[kind + "_velocity"] = loop_ms[kind] * 1e3 * b / (a + b)
    launches = {k: nlaunch[k.split("_")[0]] for k in kinds}
    kname = max(kus_total, key=lambda k: kus_total[k])
    avg_us = kus_total[kname] / launches[kname]
    bytes_per_launch = alg_kernel[kname] / launches[kname]
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    alg_total = sum(alg_kernel.values())
    # measured HBM bytes of this kernel per shot-step (FETCH_SIZE + WRITE_SIZE, copy-calibrated,
    # from a separate rocprofv3 run), times the shots per launch of this run
    traffic = None
    prof = os.path.join(ROOT, "profiles", "pmc_traffic_c4.json")
    if os.path.exists(prof) and esize == 4:
        k = json.load(open(prof))["kernels"].get(kname)
        if k:
            traffic = int(k["traffic_bytes_per_shot_step"] * bytes_per_launch / kb[kname])
    batched_rate = shots_done / T
    batched_gcell = 2.0 * cell_steps / T / 1e9
    del plan
    _log(f"batched: {batched_rate:.3f} shots/s, {batched_gcell:.2f} Gcell-updates/s")

    # the same kernels, one shot per plan, over all shots (shot tables uploaded per shot)
    p1 = new_plan()
    p1.set_model(lam, mu, *avg0)
    p1.set_shots([descs[0]])
    p1.set_tile(tile)
    p1.set_observed([obs[0]])
    p1.forward(energy=False, gradient=True)   # warm-up
    p1.misfit()
    p1.adjoint(energy=False, gradient=True)
    per = list(range(len(w.shots))) if args.per_shot_sample <= 0 else list(range(min(args.per_shot_sample,
                                                                                      len(w.shots))))
    _sync()
    t1 = time.perf_counter()
    per_cells = 0
    for i in per:
        p1.set_shots([descs[i]])
        p1.set_tile(tile)
        p1.set_observed([obs[i]])
        p1.forward(energy=False, gradient=True)
        p1.misfit()
        p1.adjoint(energy=False, gradient=True)
        per_cells += w.shots[i]["nt"] * w.interior_cells()
    _sync()
    Tp = time.perf_counter() - t1
    del p1
    per_rate = len(per) / Tp
    per_gcell = 2.0 * per_cells / Tp / 1e9
    _log(f"per-shot: {per_rate:.3f} shots/s, {per_gcell:.2f} Gcell-updates/s over {len(per)} shots")

    line = {
        "metric": METRIC, "value": round(batched_gcell, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(T / args.steps * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32fma": "f32", "fp64": "f64", "fp32x2": "f32x2"}[prec],
        "data": "synthetic: seeded layered model with a dipping fault (observed), the same smoothed "
                "over 50 m (synthetic)",
        "config": {"workload": f"1500x600 dx5 order-4 npml40, {len(w.shots)} shots, nt U[3000,8000], "
                               "receivers U[50,400], dt 0.5 ms, fwd+adj+grad; one step = one store group "
                               "of the shot set, steps cycle through the groups",
                   "padded_grid": [w.nzt, w.nxt], "interior_cells": w.interior_cells(),
                   "store_groups": [len(g) for g in groups],
                   "gradient_store": (f"full decimated store, fwi_dt {fwi_dt}, groups sized to "
                                      f"{args.store_budget_gb} GB of store" if args.store == "full" else
                                      f"checkpoint every {args.interval} steps, segments recomputed "
                                      "(recomputation not counted)"),
                   "tile": tile, "step_kernel": ("split (stress, velocity)" if tile > 0 else
                                                 "fused tile" if tile > -32 else "z-marching segments"),
                   "shot_order": "shot-major" if band == 0 else f"bands of {band} tile rows",
                   "launch_group": lgroup or "whole group", "graph_steps": args.graph_steps},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[kname], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": ("stored profile, not measured in this run: rocprofv3 FETCH_SIZE + "
                                        "WRITE_SIZE of this kernel per shot-step, copy-calibrated "
                                        "(profiles/pmc_traffic_c4.json), times this run's shots per launch")
                                       if traffic else None,
                     "alg_bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(avg_us, 2),
                     "launches": launches[kname],
                     "kernel_us_per_launch": {k: round(kus_total[k] / launches[k], 2) for k in kinds},
                     "sampled_us_per_launch": {k: round(direct[k], 2) for k in kinds},
                     "samples": nsamp[kname],
                     "duration_note": "avg_launch_us: the loop time of the timed steps (HIP events around "
                                      "each time loop) split between the stress and velocity kernels in the "
                                      "ratio of their sampled durations, so it also carries the launch gaps; "
                                      "sampled_us_per_launch: the event-bracketed durations of every 64th "
                                      "launch, each with one event's overhead"},
        "hbm": {"alg_bytes": int(alg_total), "alg_GBps": round(alg_total / T / 1e9, 1),
                "frac_of_peak": round(alg_total / T / 1e9 / HBM_PEAK_GBS, 4),
                "in_time_loops": round((loop_ms["forward"] + loop_ms["adjoint"]) / 1e3 / T, 4),
                "frac_of_peak_in_time_loops": round(alg_total / ((loop_ms["forward"] + loop_ms["adjoint"]) / 1e3)
                                                    / 1e9 / HBM_PEAK_GBS, 4),
                "note": "algorithmic bytes of forward + adjoint + gradient of every timed step over the "
                        "wall time of the timed steps; in_time_loops: the share of that wall time inside "
                        "the forward and adjoint time loops (HIP events), the rest being shot tables, "
                        "observed-trace uploads and the misfit"},
        "batched_shots_per_s": round(batched_rate, 3),
        "per_shot_shots_per_s": round(per_rate, 3),
        "per_shot_gcell_updates_per_s": round(per_gcell, 3),
        "per_shot_shots": len(per),
        "batched_over_per_shot": round(batched_rate / per_rate, 3),
        "batched_over_per_shot_per_cell_step": round(batched_gcell / per_gcell, 3),
    }
    return line


def c4_extras(args, line):
    """Keys beside the batched line: the FP64 build's rate on one store group, the
    1000 x 500 single-shot workload, and the CPU baseline on a sample of a C4 shot."""
    from cuda_fwi_amd import psvfwi as P
    from cuda_fwi_amd import workloads as WL
    if args.extras:
        w = WL.c4(nshots=args.nshots)
        fwi_dt = args.fwi_dt
        window = w.model_window(dt=fwi_dt)
        g = max(_groups_for_store(w, w.nx * w.nz, fwi_dt, 8, args.store_budget_gb * 1e9), key=len)
        lam, mu, rho = w.padded(*w.background)
        plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=w.fdorder, npml=w.npml, fsurf=w.fsurf, dt=w.dt, dx=w.dx,
                      dz=w.dx, hc=P.holberg(w.fdorder, w.max_relative_error), precision="fp64")
        plan.set_pml(P.cpml_profiles(w.npml, w.dx, w.fc, 2.0, 1.0, 1e-3, float(np.max(w.vp)), w.dt))
        plan.set_window(window)
        plan.set_model(lam, mu, *P.average_parameters(mu, rho))
        plan.set_shots(_shot_descs(w, g))
        plan.set_tile(min(abs(line["config"]["tile"]), 16) if line["config"]["tile"] > 0 else 8)
        plan.set_observed([np.zeros((len(w.shots[i]["rec"]["x"]), w.shots[i]["nt"])) for i in g])
        _sync()
        t0 = time.perf_counter()
        plan.forward(energy=False, gradient=True)
        plan.misfit()
        plan.adjoint(energy=False, gradient=True)
        _sync()
        T = time.perf_counter() - t0
        cells = sum(w.shots[i]["nt"] for i in g) * w.interior_cells()
        line["fp64"] = {"value": round(2.0 * cells / T / 1e9, 3), "unit": "Gcell-updates/s",
                        "sample": f"FP64 kernels (the reference's arithmetic), one store group of {len(g)} "
                                  f"shots, fwd+adj+grad, {T:.2f} s"}
        del plan
        _log("fp64 group done")
        a2 = argparse.Namespace(**vars(args))
        a2.steps, a2.warmup, a2.nt, a2.tile, a2.no_cpu_baseline = 3, 1, 0, 0, True
        a2.precision = "fp32"
        c2 = run_c2(a2)
        line["c2_single_shot"] = {"value": c2["value"], "unit": c2["unit"], "ms_per_pass": c2["ms_per_step"],
                                  "workload": c2["config"]["workload"], "roofline_frac": c2["roofline"]["frac"],
                                  "tile": c2["config"]["tile"]}
        _log("1000 x 500 single shot done")
        # the same with the time loops replayed as HIP graphs of 64 steps (its launches are
        # short enough for the gaps between them to show)
        a2.graph_steps = 64
        c2g = run_c2(a2)
        line["c2_single_shot"]["value_graph_steps_64"] = c2g["value"]
        _log("1000 x 500 single shot with graphs done")
    if not args.no_cpu_baseline:
        w = WL.c4(nshots=1)
        window = w.model_window(dt=args.fwi_dt)
        _log("cpu baseline (oracle) starting")
        line["cpu_baseline"] = cpu_baseline(w, window, args.cpu_nt)
        line["cpu_baseline"]["sample"] += "; one C4 shot's grid (1584 x 684 padded), same model and geometry"
        _log("cpu baseline (threaded oracle) starting")
        line["cpu_baseline_threads"] = cpu_baseline(w, window, args.cpu_nt, threaded=True)
    return line


def run_c3mem(args):
    """Memory test: 2000 x 1000 cells, nt 8000, FP32, one shot, gradient with
    the source wavefield rebuilt from checkpoints every --interval steps.
    Reports the device memory in use (hipMemGetInfo through torch) and what
    the full decimated store would need."""
    import torch
    from cuda_fwi_amd import workloads as WL
    w = WL.c3mem()
    if args.nt:
        w.nt = args.nt
        w.shots[0]["nt"] = args.nt
    fwi_dt = args.fwi_dt
    window = w.model_window(dt=fwi_dt)
    torch.cuda.init()
    free0, total = torch.cuda.mem_get_info()
    mode = "random_boundary" if args.store == "random_boundary" else "checkpoint"
    plan = make_plan(w, args.precision, window, mode, args.interval, args.tile)
    esize = 8 if args.precision in ("fp64", "fp32x2") else 4
    used = []
    _sync()
    t0 = time.perf_counter()
    plan.forward(energy=False, gradient=True)
    used.append(total - torch.cuda.mem_get_info()[0])
    plan.misfit()
    plan.adjoint(energy=False, gradient=True)
    _sync()
    T = time.perf_counter() - t0
    used.append(total - torch.cuda.mem_get_info()[0])
    cells = w.interior_cells()
    full_store = 5 * w.nx * w.nz * (1 + (w.nt - 1) // 1) * esize
    return {
        "metric": METRIC, "value": round(2.0 * cells * w.nt / T / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(T * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if esize == 4 else "f64",
        "data": "synthetic: 8 flat layers", "config": {
            "workload": f"2000x1000 dx5 order-4, nt {w.nt}, 1 shot, fwd+adj+grad, " + (
                "random-boundary reconstruction (no CPML: a 40-cell random layer)" if mode == "random_boundary"
                else f"npml40, checkpoint every {args.interval} steps") + f", fwi_dt {fwi_dt}", "tile": plan.tile},
        "memory": {"device_used_before_GB": round((total - free0) / 1e9, 2),
                   "device_used_peak_GB": round(max(used) / 1e9, 2),
                   "device_total_GB": round(total / 1e9, 1),
                   "full_store_every_step_GB": round(full_store / 1e9, 1),
                   "note": "device memory in use after the forward and after the adjoint (hipMemGetInfo); "
                           "the store of every step (fwi_dt 1) would not fit"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c2", "c4", "c3mem", "c5"])
    ap.add_argument("--iters", type=int, default=20, help="c5: FWI iterations")
    ap.add_argument("--c5-group", type=int, default=0,
                    help="c5: shots per batched launch (0: all); the discrete adjoint of all 128 full-size "
                         "shots does not fit in HBM at once")
    ap.add_argument("--c5-mute", type=int, default=10,
                    help="c5: model rows below the surface with no update (sources and receivers at row 2)")
    ap.add_argument("--c5-smooth", type=float, default=0.0,
                    help="c5: Gaussian smoothing of the descent direction, sigma in cells at full scale (0: none)")
    ap.add_argument("--c5-precond", default="source", choices=["source", "source_receiver"],
                    help="c5: illumination that divides the gradient: source energy, or sqrt(source x receiver)")
    ap.add_argument("--c5-gradient", default="discrete", choices=["discrete", "reference"],
                    help="c5: the exact gradient of the scheme, or the reference's correlation gradient")
    ap.add_argument("--c5-taper", type=int, default=0,
                    help="c5: rows below the muted ones over which the direction ramps up (cosine)")
    ap.add_argument("--c5-depth-power", type=float, default=0.0,
                    help="c5: depth gain of the direction, ((row + 1) / rows)^p (0: none)")
    ap.add_argument("--c5-resume", default="",
                    help="c5: continue from a state file written by --c5-save")
    ap.add_argument("--c5-save", default="", help="c5: write the model, step and history here after the run")
    ap.add_argument("--c5-scale", type=float, default=0.5,
                    help="c5: lengths, shot and receiver counts and nt relative to 2400 x 800 / 128 / 600 / 6000")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32fma", "fp64", "fp32x2"])
    ap.add_argument("--store", default="full", choices=["full", "checkpoint", "random_boundary"])
    ap.add_argument("--interval", type=int, default=200)
    ap.add_argument("--fwi-dt", type=int, default=5)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--cpu-nt", type=int, default=200)
    ap.add_argument("--tune-steps", type=int, default=60, help="c4: steps per autotuner candidate")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="c4: skip the FP64 group and the 1000 x 500 line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nshots", type=int, default=64)
    ap.add_argument("--c4-grid", default="1500x600",
                    help="c4: model cells NXxNZ (smaller grids: the same kind of model and geometry)")
    ap.add_argument("--nt", type=int, default=0, help="override nt (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only time a 1 GiB device copy (counter calibration)")
    ap.add_argument("--per-shot-sample", type=int, default=0,
                    help="c4: shots timed one per launch (0: all of them)")
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="replay the time loop as captured HIP graphs of this many steps (0: direct launches)")
    ap.add_argument("--gradient", default="reference", choices=["reference", "discrete"],
                    help="default workload: the reference's correlation gradient or the discrete (exact) one")
    ap.add_argument("--shot-order", type=int, default=-1,
                    help="c4: block order, 0 shot-major or bands of this many tile rows (-1: autotuned)")
    ap.add_argument("--launch-group", type=int, default=6,
                    help="c4: shots per launch group (-1: the autotuner's choice, 0: whole group); 6 was the "
                         "fastest over whole runs (the autotuner times short runs)")
    ap.add_argument("--store-budget-gb", type=float, default=200.0,
                    help="batched workload with --store full: store memory per group of shots")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(shots shard across GPUs with one all-reduce of the gradient, see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.calibrate_copy:
        print(json.dumps(calibrate_copy()), flush=True)
        return
    run = {"c2": run_c2, "c4": run_c4, "c3mem": run_c3mem, "c5": run_c5}[args.workload]
    if args.workload == "c2" and args.gradient == "discrete":
        run = run_c2_discrete
    line = run(args)
    if args.workload == "c4":
        line = c4_extras(args, line)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
