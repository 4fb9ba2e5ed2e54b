"""The reference's host API on the GPU: fd_kernel.hpp entry points
(forward_kernel_PSV / adjoint_kernel_PSV with the reference's argument
lists) and the psvfwi_main driver (FDWaveModel::forward_simulate and
fwi_simulate from an INI file), against the oracle.  FP64 runs of the
kernels equal the oracle bit for bit; CSV outputs carry the 6 significant
digits the stream formatting prints."""
import os
import subprocess

import numpy as np
import pytest

import psv_oracle as O
import refcase as RC
from cuda_fwi_amd import psvfwi as P
from hostbuild import ROOT, build_harness, run

pytestmark = pytest.mark.gpu
SIX = 5e-6
MAIN = os.path.join(ROOT, "cuda_fwi_amd", "lib", "psvfwi_main")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return build_harness(tmp_path_factory.mktemp("hh"))


def _csv(path):
    rows = [l.rstrip().rstrip(",") for l in open(path) if l.strip()]
    return np.array([[float(v) for v in r.split(",")] for r in rows])


def _close6(ours_csv, exact):
    exact = np.asarray(exact, np.float64)
    assert ours_csv.shape == exact.shape
    assert np.array_equal(ours_csv == 0, exact == 0)
    nz = exact != 0
    rel = np.abs(ours_csv[nz] - exact[nz]) / np.abs(exact[nz])
    return rel.max() if rel.size else 0.0


def _write_case(d, pb, ishot, src_shot, rec_shot, adj, snap, fp64=1):
    w = pb.fwi
    s, r = pb.src, pb.rec
    head = [ishot, pb.nt, pb.nzt, pb.nxt, pb.fpad, pb.ppad, repr(pb.dt), repr(pb.dx), repr(pb.dz), snap,
            pb.fsurf, pb.fdorder, pb.npml, len(s["x"]), len(r["x"]), 1, w["dt"], w["dx"], w["dz"], w["x1"],
            w["x2"], w["z1"], w["z2"], fp64]
    lines = [" ".join(map(str, head)), " ".join(repr(float(h)) for h in pb.hc)]
    for arr in (s["x"], s["z"], s["comp"], src_shot, r["x"], r["z"], r["comp"], rec_shot):
        lines.append(" ".join(str(int(v)) for v in arr))
    (d / "in.txt").write_text("\n".join(lines) + "\n")
    for name, a in (("lam", pb.lam), ("mu", pb.mu), ("mu_zx", pb.mu_zx), ("rho_zp", pb.rho_zp),
                    ("rho_xp", pb.rho_xp), ("pml", pb.pml), ("src_signal", pb.src_signal), ("adj_signal", adj)):
        np.ascontiguousarray(a, np.float64).tofile(d / f"{name}.bin")


@pytest.mark.parametrize("entry", ["cpu_signature", "gpu_signature"])
@pytest.mark.parametrize("fdorder", [2, 4])
def test_dropin_kernels_fp64_bit_exact(harness, tmp_path, fdorder, entry):
    """forward_kernel_PSV then adjoint_kernel_PSV (or the _GPU entry points of
    CUDA_week2/fd_kernel.cuh:75-173, whose device pointers the plan replaces) for
    shot 1 of a two-shot geometry: traces, wavefields, CPML memory, FWI store,
    gradients.  The gradient arrays hold 7.0 on entry and must come back holding
    this shot's gradient alone (the reference zeroes them,
    CUDA_week2/adjoint_kernel_PSV_GPU.cu:799-806)."""
    print(f"drop-in kernels order {fdorder}, {entry}", flush=True)
    gpu = entry == "gpu_signature"
    # the _GPU signatures have no We argument (fd_kernel.cuh:75-120)
    fields = ("vx", "vz", "sxx", "szx", "szz") if gpu else ("vx", "vz", "sxx", "szx", "szz", "We")
    pb = RC.small_problem(nx=64, nz=40, npml=8, nt=240, fdorder=fdorder, fsurf=1, seed=5, nsrc=2, nrec=6)
    pb.fwi = dict(dt=2, dx=1, dz=2, x1=pb.ppad - 2, x2=pb.nxt - 3, z1=pb.fpad, z2=pb.nzt - 5)
    src_shot = [0, 1]
    rec_shot = [1, 0, 1, 1, 0, 1]
    pb.src = dict(pb.src, shot=src_shot)
    pb.rec = dict(pb.rec, shot=rec_shot)
    adj = np.random.default_rng(1).standard_normal((6, pb.nt)) * 1e-3
    _write_case(tmp_path, pb, 1, src_shot, rec_shot, adj, snap=0)
    run(harness, "dropin", tmp_path, *(["gpu"] if gpu else []))

    ref = O.forward(pb, ishot=1)
    n = pb.nzt * pb.nxt
    rec = np.fromfile(tmp_path / "out_rec.bin").reshape(6, pb.nt)
    mine = [i for i, s in enumerate(rec_shot) if s == 1]
    other = [i for i, s in enumerate(rec_shot) if s != 1]
    assert np.abs(ref["rec"][mine]).max() > 0
    assert np.array_equal(rec[mine], ref["rec"][mine])
    assert np.all(rec[other] == -1.0)           # rows of other shots untouched
    fl = np.fromfile(tmp_path / "out_fwd_fields.bin").reshape(6, pb.nzt, pb.nxt)
    for k, name in enumerate(fields):
        assert np.array_equal(fl[k], ref[name]), name
    assert np.array_equal(np.fromfile(tmp_path / "out_fwd_mem.bin"), ref["mem"])
    store = np.fromfile(tmp_path / "out_store.bin").reshape(ref["store"].shape)
    assert np.array_equal(store, ref["store"])

    adj_src = dict(x=[pb.rec["x"][i] for i in mine], z=[pb.rec["z"][i] for i in mine],
                   comp=[pb.rec["comp"][i] for i in mine], shot=[1] * len(mine))
    a = O.adjoint(pb, adj_src, adj[mine], ref["store"], ishot=1)
    nft, nfz, nfx = pb.fwi_dims()
    g = np.fromfile(tmp_path / "out_grad.bin").reshape(3, nfz, nfx)
    assert np.abs(a["grad_lam"]).max() > 0
    for k, name in enumerate(("grad_lam", "grad_mu", "grad_rho")):
        assert np.array_equal(g[k], a[name]), name
    fa = np.fromfile(tmp_path / "out_adj_fields.bin").reshape(6, pb.nzt, pb.nxt)
    for k, name in enumerate(fields):
        assert np.array_equal(fa[k], a[name]), name


def test_dropin_snapshot_csv(harness, tmp_path):
    """forward_kernel_PSV with fwinv = false writes szz every snap_interval
    steps to vz_snap<k>.csv (the reference's file name and row format)."""
    print("drop-in snapshots", flush=True)
    pb = RC.small_problem(nx=50, nz=30, npml=6, nt=61, fdorder=2, fsurf=1, seed=2, nsrc=1, nrec=3)
    pb.fwi = RC.full_window(pb)
    pb.src = dict(pb.src, shot=[0])
    pb.rec = dict(pb.rec, shot=[0, 0, 0])
    _write_case(tmp_path, pb, 0, [0], [0, 0, 0], np.zeros((3, pb.nt)), snap=20)
    run(harness, "dropin", tmp_path)
    files = sorted(os.listdir(tmp_path / "snap"))
    assert files == [f"vz_snap{k}.csv" for k in range(4)]
    for k in (1, 3):
        ref = O.forward(RC.small_problem(nx=50, nz=30, npml=6, nt=20 * k + 1, fdorder=2, fsurf=1, seed=2,
                                         nsrc=1, nrec=3))
        assert _close6(_csv(tmp_path / "snap" / f"vz_snap{k}.csv"), ref["szz"]) <= SIX


INI = """\
[grid]
nt = {nt}
nx = {nx}
nz = {nz}
dx = 5.0
dz = 5.0
dt = 0.5e-3
snap_interval = {snap}
free_surface = 1

[medium]
scalar_rho = 2000.0
scalar_vp = 2500.0
scalar_vs = 1443.0

[PML_boundary]
npml = 8
fpml = 25.0
dh = 5.0
npower = 2.0
k_max_PML = 1.0
rcoef = 0.001
damp_v_PML = 2750.0

[FDM]
fdorder = {order}
max_relative_error = 1
full_waveform_inversion = {fwi}
nshots = 1
max_iteration = {iters}
max_error = 1.0e-30

[source]
nseis = 1
x_seis = {{{sx}}}
z_seis = {{2}}
velocity_or_stress = {{0}}
component_type = {{2}}
shot_to_fire = {{0}}
shape = {{1}}
amplitude = {{1.0}}
peak_frequency = {{25.0}}
time_shift = {{0.0}}

[receiver]
nseis = 4
x_seis = {{5, 20, 35, 50}}
z_seis = {{1, 1, 1, 1}}
velocity_or_stress = {{0, 0, 0, 0}}
component_type = {{2, 2, 2, 2}}
shot_to_fire = {{0, 0, 0, 0}}
measurement_records = {mr}

[full_waveform_inversion]
time_interval = 2
grid_interval_x = 1
grid_interval_y = 1
left_grid = 0
right_grid = {nx}
top_grid = 0
bottom_grid = {nz}

[mi355x]
precision = fp64
observed = {obs}
"""


def _ini_problem(nt, nx, nz, order, sx, modified, fwi=None):
    fpad = order // 2
    ppad = 8 + fpad
    nxt, nzt = nx + 2 * ppad, nz + ppad + fpad
    lam0, mu0 = O.velocity_to_lame(2500.0, 1443.0, 2000.0)
    rho, lam, mu = (np.full((nzt, nxt), v) for v in (2000.0, lam0, mu0))
    if modified:
        rho, lam, mu = O.stagger_over_grid_modify(nzt, nxt, 2000.0, lam0, mu0)
    pml = O.cpml_pro(8, 5.0, 25.0, 2.0, 1.0, 0.001, 2750.0, 0.5e-3)
    sig = O.create_wavelet(nt, 0.5e-3, 1.0, 25.0, 0.0, 1)[None]
    src = dict(x=[sx + ppad], z=[2 + fpad], comp=[2], shot=[0])
    rec = dict(x=[x + ppad for x in (5, 20, 35, 50)], z=[1 + fpad] * 4, comp=[2] * 4, shot=[0] * 4)
    return O.Problem(nt=nt, nzt=nzt, nxt=nxt, fdorder=order, npml=8, dt=0.5e-3, dx=5.0, dz=5.0, fsurf=1,
                     hc=O.holbergcoeff(order, 1), lam=lam, mu=mu, rho=rho, pml=pml, src=src, src_signal=sig,
                     rec=rec, fwi=fwi)


def _main(tmp_path, text):
    (tmp_path / "run.ini").write_text(text)
    r = subprocess.run([MAIN, "run.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_main_forward_simulate(tmp_path):
    """psvfwi_main on a forward INI: rec_signal.csv and the szz snapshots of
    the modified (lambda disc) model."""
    print("psvfwi_main forward", flush=True)
    _main(tmp_path, INI.format(nt=201, nx=60, nz=40, order=4, fwi=0, iters=1, sx=30, snap=100, mr=0, obs=""))
    pb = _ini_problem(201, 60, 40, 4, 30, modified=True)
    ref = O.forward(pb)
    assert _close6(_csv(tmp_path / "io/receiver/rec_signal.csv"), ref["rec"]) <= SIX
    assert _close6(_csv(tmp_path / "io/snap_data/vz_snap2.csv"), ref["szz"]) <= SIX
    assert _close6(_csv(tmp_path / "io/snap_data/lamda_snap.csv"), pb.rho) == 0


def _oracle_fwi(nt, nx, nz, order, sx, iters, observed):
    """The reference's FWI loop (CPU branch, one shot) restated with the
    oracle's functions: CUDA_week2/fd_wave_model.cu:753-966 and
    CUDA_week2/fwi_step_length_estimation.cu:11-224.  Returns the misfit per
    iteration, the trial traces per (iteration, trial) and the model; a
    failed step-length search ends it like the reference's exit."""
    if observed == "modified":
        obs = O.forward(_ini_problem(nt, nx, nz, order, sx, modified=True))["rec"]
    else:
        obs = np.zeros((4, nt))       # measurement_records = 1: allocated, never filled
    pb = _ini_problem(nt, nx, nz, order, sx, modified=False)
    pb.rho = np.full_like(pb.lam, 2000.0)
    pb.fwi = dict(dt=2, dx=1, dz=1, x1=pb.ppad, x2=min(nx + pb.ppad, pb.nxt), z1=pb.fpad,
                  z2=min(nz + pb.fpad, pb.nzt))
    nft, nfz, nfx = pb.fwi_dims()
    w = pb.fwi
    old = dict(rho=pb.rho.copy(), lam=pb.lam.copy(), mu=pb.mu.copy())
    cur = dict(rho=pb.rho.copy(), lam=pb.lam.copy(), mu=pb.mu.copy())
    types = dict(rho=1, lam=2, mu=3)

    def set_model(m):
        pb.lam, pb.mu, pb.rho = m["lam"], m["mu"], m["rho"]
        pb.mu_zx, pb.rho_zp, pb.rho_xp = O.average_parameters(pb.mu, pb.rho)

    def update(step, grads, override, order):
        for k in order:
            f = 0.5 if k == "rho" else 1.0
            # cells outside the update loop hold the old values in both arrays
            o, cur[k] = O.fwi_grad_update(old[k], grads[k], f * step, nfz, nfx, w["z1"], w["x1"],
                                          w["dz"], w["dx"], override, types[k])
            if override:
                old[k] = o

    step, hist, trials = 0.01, [], {}
    for itr in range(1, iters + 1):
        set_model(cur)
        C_lam, C_mu, C_rho = O.medium_average(cur["lam"], cur["mu"], cur["rho"])
        f = O.forward(pb)
        L2, res, _ = O.l2_adjoint_sources(pb.dt, f["rec"], obs)
        hist.append(L2)
        a = O.adjoint(pb, pb.rec, res, f["store"])
        g = O.scale_gradients_energy(f["We"], a["We"], C_rho, C_lam, C_mu, nfz, nfx, w["z1"], w["x1"],
                                     w["dz"], w["dx"], [a["grad_lam"], a["grad_mu"], a["grad_rho"]])
        grads = dict(lam=g[0], mu=g[1], rho=g[2])
        L2t, sl = [L2, 0, 0, L2], [0.0, 0, 0, 0]
        step1 = step2 = cnt = 0
        it, est, check = 1, step, 0
        while not (step1 and step2):
            check += 1
            update(est, grads, False, ("lam", "mu", "rho"))
            set_model(cur)
            rec = O.forward(pb)["rec"]
            trials[(itr, check)] = rec
            L2t[it] = O.l2_adjoint_sources(pb.dt, rec, obs)[0]
            sl[it] = est
            if step1 and L2t[1] < L2t[2]:
                sl[2], step2 = est, 1
            elif step1:
                sl[2] = est
                est += est / 2
                cnt += 1
            elif L2t[0] > L2t[1]:
                sl[1], step1, it, cnt = est, 1, 2, 0
                est += est / 2
            else:
                est /= 2
                cnt += 1
            if not step1 and cnt > 6:
                return hist, trials, None
            if step1 and cnt > 6:
                step2 = 1
        step = O.calc_opt_step(L2t[:3], sl[:3])
        update(step, grads, True, ("rho", "mu", "lam"))
    return hist, trials, cur


def _printed_l2(out):
    return [float(l.split("=")[1].split(",")[0]) for l in out.splitlines() if l.startswith("L2 norm =")]


def test_main_fwi_simulate(tmp_path):
    """psvfwi_main on an FWI INI as the reference runs it (measurement_records
    = 1: observed traces allocated and left zero; homogeneous start; two
    iterations): misfits, every step-length trial's traces and the model
    after each iteration (density_snap / lam_snap / mu_snap CSVs), against
    the reference's loop restated over the oracle."""
    print("psvfwi_main fwi", flush=True)
    nt, nx, nz, order, sx = 240, 60, 40, 2, 20
    out = _main(tmp_path, INI.format(nt=nt, nx=nx, nz=nz, order=order, fwi=1, iters=2, sx=sx, snap=1000,
                                     mr=1, obs=""))
    hist, trials, model = _oracle_fwi(nt, nx, nz, order, sx, 2, "")
    assert model is not None
    assert np.allclose(_printed_l2(out), hist, rtol=1e-5, atol=0)
    for (itr, k), rec in trials.items():
        assert _close6(_csv(tmp_path / f"io/receiver/rec_signal_{itr}_{k}.csv"), rec) <= SIX, (itr, k)
    d = tmp_path / "io/density_snap"
    for k, f in (("rho", "density_snap2.csv"), ("lam", "lam_snap2.csv"), ("mu", "mu_snap2.csv")):
        assert _close6(_csv(d / f), model[k]) <= SIX, k
    assert model["lam"].max() > model["lam"].min()


def test_main_fwi_step_search_fails_like_reference(tmp_path):
    """Observed traces of the lambda-disc model: the reference's update
    direction raises the misfit at every trial step, so its search gives up
    after 7 halvings (the reference exits there).  The driver stops with an
    error after the same 7 trials, whose traces match the oracle's."""
    print("psvfwi_main fwi failing search", flush=True)
    nt, nx, nz, order, sx = 240, 60, 40, 2, 30
    (tmp_path / "run.ini").write_text(INI.format(nt=nt, nx=nx, nz=nz, order=order, fwi=1, iters=2, sx=sx,
                                                 snap=1000, mr=0, obs="modified"))
    r = subprocess.run([MAIN, "run.ini"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "Step length estimation failed" in r.stderr
    hist, trials, model = _oracle_fwi(nt, nx, nz, order, sx, 2, "modified")
    assert model is None and len(trials) == 7
    assert np.allclose(_printed_l2(r.stdout), hist, rtol=1e-5, atol=0)
    for (itr, k), rec in trials.items():
        assert _close6(_csv(tmp_path / f"io/receiver/rec_signal_{itr}_{k}.csv"), rec) <= SIX, (itr, k)


def test_main_fwi_discrete_gradient_converges(tmp_path):
    """psvfwi_main with [mi355x] gradient = discrete on the lambda-disc data
    that defeats the reference's gradient (previous test): the misfit falls
    at every iteration and lambda moves toward the disc."""
    print("psvfwi_main fwi discrete", flush=True)
    nt, nx, nz, order, sx, iters = 240, 60, 40, 2, 30, 4
    text = INI.format(nt=nt, nx=nx, nz=nz, order=order, fwi=1, iters=iters, sx=sx, snap=1000, mr=0,
                      obs="modified\ngradient = discrete\ngradient_segment = 50")
    out = _main(tmp_path, text)
    l2 = _printed_l2(out)
    print("  misfit per iteration:", l2)
    assert len(l2) == iters and all(b < a for a, b in zip(l2, l2[1:])), l2
    assert l2[-1] < 0.7 * l2[0]
    pb_true = _ini_problem(nt, nx, nz, order, sx, modified=True)
    pb0 = _ini_problem(nt, nx, nz, order, sx, modified=False)
    lam = _csv(tmp_path / f"io/density_snap/lam_snap{iters}.csv")
    e0 = np.linalg.norm(pb0.lam - pb_true.lam)
    e1 = np.linalg.norm(lam - pb_true.lam)
    print(f"  lambda error vs the disc model: start {e0:.4e}, after {iters} iterations {e1:.4e}")
    assert e1 < e0
    # gradient_segment = 1: the forward keeps every state and the adjoint
    # replays nothing; the same gradients, so the same run to the printed digit
    every = tmp_path / "every"
    every.mkdir()
    out1 = _main(every, text.replace("gradient_segment = 50", "gradient_segment = 1"))
    assert _printed_l2(out1) == l2
    assert np.array_equal(_csv(every / f"io/density_snap/lam_snap{iters}.csv"), lam)


MODEL_INI = """\
[grid]
nt = {nt}
nx = {nx}
nz = {nz}
dx = 5.0
dz = 5.0
dt = 0.5e-3
free_surface = 0

[medium]
scalar_rho = 2000.0
scalar_vp = 2500.0
scalar_vs = 1443.0

[PML_boundary]
npml = 10
fpml = 25.0
dh = 5.0
npower = 2.0
k_max_PML = 1.0
rcoef = 0.001
damp_v_PML = 4500.0

[FDM]
fdorder = 4
max_relative_error = 1
full_waveform_inversion = 1
nshots = {nshots}
max_iteration = {iters}
max_error = 1.0e-30

[source]
nseis = {nshots}
x_seis = {{{sx}}}
z_seis = {{{sz}}}
velocity_or_stress = {{{s0}}}
component_type = {{{s2}}}
shot_to_fire = {{{sshot}}}
shape = {{{s1}}}
amplitude = {{{samp}}}
peak_frequency = {{{sfc}}}
time_shift = {{{s0}}}

[receiver]
nseis = {nrec}
x_seis = {{{rx}}}
z_seis = {{{rz}}}
velocity_or_stress = {{{r0}}}
component_type = {{{r2}}}
shot_to_fire = {{{rshot}}}

[full_waveform_inversion]
time_interval = 2
grid_interval_x = 1
grid_interval_y = 1
left_grid = 0
right_grid = {nx}
top_grid = 0
bottom_grid = {nz}

[mi355x]
precision = fp64
observed = model
model_file = start.bin
true_model_file = true.bin
gradient = discrete
gradient_segment = 50
write_csv = false
"""


def test_main_fwi_model_files(tmp_path):
    """psvfwi_main runs the FWI of the layered / faulted / high-velocity-body configuration
    (shrunk to 160 x 60 cells, 4 shots x 24 receivers) from model files: observed traces from
    true_model_file, the start model (the truth smoothed) from model_file.  Its first misfit
    equals the misfit the C ABI computes for the same padded models and geometry (FP64), and
    the misfit falls at every iteration (discrete gradient)."""
    from cuda_fwi_amd import workloads as WL
    print("psvfwi_main FWI from model files", flush=True)
    nx, nz, nt, ns, nr = 160, 60, 500, 4, 24
    w = WL.c5(nshots=ns, nrec=nr, nt=nt, nx=nx, nz=nz)
    w.npml, w.fc, w.dt = 10, 25.0, 0.5e-3
    np.concatenate([w.vp.ravel(), w.vs.ravel(), w.rho.ravel()]).astype(np.float64).tofile(tmp_path / "true.bin")
    vp0, vs0, rho0 = w.background
    np.concatenate([vp0.ravel(), vs0.ravel(), rho0.ravel()]).astype(np.float64).tofile(tmp_path / "start.bin")
    sx = [s["src"]["x"][0] for s in w.shots]
    rx = [x for s in w.shots for x in s["rec"]["x"]]
    rz = [z + 2 for s in w.shots for z in s["rec"]["z"]]
    j = lambda v: ", ".join(str(a) for a in v)
    text = MODEL_INI.format(nt=nt, nx=nx, nz=nz, nshots=ns, iters=3, sx=j(sx), sz=j([2] * ns),
                            s0=j([0] * ns), s1=j([1] * ns), s2=j([2] * ns), sshot=j(range(ns)),
                            samp=j([1.0] * ns), sfc=j([25.0] * ns), nrec=len(rx), rx=j(rx), rz=j(rz),
                            r0=j([0] * len(rx)), r2=j([2] * len(rx)), rshot=j(s for s in range(ns) for _ in range(nr)))
    out = _main(tmp_path, text)
    hist = [float(l.split("=")[1].split(",")[0]) for l in out.splitlines() if l.startswith("L2 norm")]
    print(f"  misfit per iteration {hist}", flush=True)
    assert len(hist) == 3 and hist[2] < hist[1] < hist[0]
    # the same misfit through the C ABI: padded models, padded geometry
    shots = []
    for i, s in enumerate(w.shots):
        sig = P.ricker(nt, w.dt, 1.0, 25.0)[None]
        shots.append(dict(nt=nt, src=dict(x=[sx[i] + w.ppad], z=[2 + w.zoff], comp=[2]), src_signal=sig,
                          rec=dict(x=[x + w.ppad for x in s["rec"]["x"]], z=[z + 2 + w.zoff for z in s["rec"]["z"]],
                                   comp=[2] * nr)))
    plan = P.Plan(nzt=w.nzt, nxt=w.nxt, fdorder=4, npml=10, fsurf=0, dt=w.dt, dx=5.0, dz=5.0,
                  hc=P.holberg(4, 1), precision="fp64")
    plan.set_pml(P.cpml_profiles(10, 5.0, 25.0, 2.0, 1.0, 1e-3, 4500.0, w.dt))
    plan.set_shots(shots)
    lt, mt, rt = w.padded()
    plan.set_model(lt, mt, *P.average_parameters(mt, rt))
    plan.forward(energy=False)
    obs = [plan.records(i) for i in range(ns)]
    lam, mu, rho = w.padded(*w.background)
    plan.set_model(lam, mu, *P.average_parameters(mu, rho))
    plan.forward(energy=False)
    m0 = float(np.sum(plan.misfit(obs)))
    plan.close()
    # psvfwi_main prints the misfit with the stream's default 6 significant digits
    assert abs(hist[0] - m0) <= 1e-5 * m0, (hist[0], m0)


def test_main_fwi_model_files_shaped_direction(tmp_path):
    """The same model-file FWI with the direction options of the C++ loop (sqrt(Ws Wr)
    illumination, which runs one more adjoint per iteration, smoothing, depth gain, mute and
    taper): the first misfit equals the unshaped run's (the same forward), and the misfit falls
    at every iteration."""
    from cuda_fwi_amd import workloads as WL
    print("psvfwi_main FWI from model files, shaped direction", flush=True)
    nx, nz, nt, ns, nr = 160, 60, 500, 4, 24
    w = WL.c5(nshots=ns, nrec=nr, nt=nt, nx=nx, nz=nz)
    np.concatenate([w.vp.ravel(), w.vs.ravel(), w.rho.ravel()]).astype(np.float64).tofile(tmp_path / "true.bin")
    vp0, vs0, rho0 = w.background
    np.concatenate([vp0.ravel(), vs0.ravel(), rho0.ravel()]).astype(np.float64).tofile(tmp_path / "start.bin")
    sx = [s["src"]["x"][0] for s in w.shots]
    rx = [x for s in w.shots for x in s["rec"]["x"]]
    rz = [z + 2 for s in w.shots for z in s["rec"]["z"]]
    j = lambda v: ", ".join(str(a) for a in v)
    base = MODEL_INI.format(nt=nt, nx=nx, nz=nz, nshots=ns, iters=3, sx=j(sx), sz=j([2] * ns),
                            s0=j([0] * ns), s1=j([1] * ns), s2=j([2] * ns), sshot=j(range(ns)),
                            samp=j([1.0] * ns), sfc=j([25.0] * ns), nrec=len(rx), rx=j(rx), rz=j(rz),
                            r0=j([0] * len(rx)), r2=j([2] * len(rx)), rshot=j(s for s in range(ns) for _ in range(nr)))
    hists = []
    for extra in ("", "illumination = source_receiver\ndirection_smooth = 2\ndirection_depth_power = 2\n"
                      "direction_mute = 4\ndirection_taper = 8\n"):
        d = tmp_path / ("shaped" if extra else "plain")
        d.mkdir()
        for f in ("true.bin", "start.bin"):
            (d / f).write_bytes((tmp_path / f).read_bytes())
        out = _main(d, base + extra)
        hists.append([float(l.split("=")[1].split(",")[0]) for l in out.splitlines() if l.startswith("L2 norm")])
    print(f"  misfit per iteration: plain {hists[0]}, shaped {hists[1]}", flush=True)
    h = hists[1]
    assert len(h) == 3 and h[2] < h[1] < h[0]
    assert h[0] == hists[0][0]
