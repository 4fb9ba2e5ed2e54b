"""The C++ host layer (include/fd_host.hpp): INI syntax, configuration
parsing and the host-side setup it feeds the kernels, against the oracle.

The INI texts below follow the reference's own configuration file
(CUDA_week2/test_configuration.ini), including its unusual lines: values
with trailing text after ';' without a space ("0.5;    !! ..."), a section
header followed by "; comment", free-text lines without '=' and '!!' lines
with an '='.
"""
import numpy as np
import pytest

import psv_oracle as O
from hostbuild import build_harness, run

REF_LIKE_INI = """\
[grid]
time = 0.5;    !! Either time (float) or nt (int) to be given
nx = 601;                        int
nz = 601;                        int
dx = 4.0;                             float
dz = 4.0;                             float
dt = 0.6e-3;                           float
snap_interval = 50;
free_surface = 1;

[medium]; Default values for the simulated models if none are loaded
scalar_rho = 7100.0;    float
scalar_vp = 2955.0;     float
scalar_vs = 2362.0;      float

[PML_boundary]
npml = 10;
fpml = 17.0;
dh = 4.0;

npower = 2.0;
k_max_PML = 1.0;
rcoef = 0.001; reflection coefficient
damp_v_PML = 3300.0; damping velocity

[FDM]
fdorder = 2;
full_waveform_inversion = 1;
nshots = 1;
max_iteration = 500;
max_error = 1.0e-3;

[source]
nseis = 1;

x_seis = {300};
z_seis = {2};

!! 0 = velocity, 1 = stress
velocity_or_stress = {0};

!! 2 = vz component
component_type = {2};
verticle_angle_deg = {0};
shot_to_fire = {0};

shape = {1}; //Ricker
amplitude = {1.0};
peak_frequency = {17.0};
time_shift = {0.0};

seis_file =
signal_file =

[receiver]
nseis = 3;

x_seis = {100, 300, 400};
z_seis = {0, 0, 0};
velocity_or_stress = {0, 0, 0}
component_type = {2, 2, 2};
shot_to_fire = {0};

[full_waveform_inversion]
time_interval = 3;
grid_interval_x = 2;
grid_interval_z = 2;

left_grid = 0;
right_grid = 0;

top_grid = 0;
bottom_grid = 0;
"""


@pytest.fixture(scope="module", params=["plain", "asan+ubsan"])
def harness(tmp_path_factory, request):
    """The harness against the library, and (asan+ubsan) with the host layer's sources
    compiled into it under AddressSanitizer and UndefinedBehaviorSanitizer."""
    return build_harness(tmp_path_factory.mktemp("hh"), sanitize=request.param == "asan+ubsan")


def _ini(harness, tmp_path, text, keys, bom=False):
    p = tmp_path / "t.ini"
    p.write_bytes((b"\xef\xbb\xbf" if bom else b"") + text.encode())
    out = tmp_path / "o.txt"
    run(harness, "ini", p, out, *[x for k in keys for x in k])
    lines = out.read_text().split("\n")
    err = int(lines[0])
    vals = [l.split("\t") for l in lines[1:1 + len(keys)]]
    return err, {k: (v[0], int(v[1]), float(v[2]), int(v[3])) for k, v in zip(keys, vals)}


def test_ini_syntax(harness, tmp_path):
    text = """\
; comment line
# another comment
[Grid]          ; inline comment after whitespace
NX = 10;   int
nz = 20 ; trailing comment
dx : 2.5
name = first
   continued ; note
dup = a
dup = b
flag = Yes
off = off
hexv = 0x1F
empty =
this line has no separator
[sec2
[other]
k = v;v
"""
    keys = [("grid", "nx"), ("GRID", "NZ"), ("grid", "dx"), ("grid", "name"), ("grid", "dup"),
            ("grid", "flag"), ("grid", "off"), ("grid", "hexv"), ("grid", "empty"), ("grid", "missing"),
            ("other", "k")]
    err, v = _ini(harness, tmp_path, text, keys, bom=True)
    assert err == 15                                  # first bad line; parsing went on
    assert v[("grid", "nx")] == ("10;   int", 10, 10.0, -1)
    assert v[("GRID", "NZ")][:2] == ("20", 20)
    assert v[("grid", "dx")][2] == 2.5
    assert v[("grid", "name")][0] == "first|continued"
    assert v[("grid", "dup")][0] == "a|b"
    assert v[("grid", "flag")][3] == 1 and v[("grid", "off")][3] == 0
    assert v[("grid", "hexv")][1:3] == (31, 31.0)
    assert v[("grid", "empty")] == ("", -7, -7.5, -1)
    assert v[("grid", "missing")][0] == "<default>"
    assert v[("other", "k")][0] == "v;v"


def test_ini_missing_file(harness, tmp_path):
    out = tmp_path / "o.txt"
    run(harness, "ini", tmp_path / "nope.ini", out)
    assert out.read_text().split("\n")[0] == "-1"


def _config(harness, tmp_path, text):
    p = tmp_path / "c.ini"
    p.write_text(text)
    out = tmp_path / "c.txt"
    run(harness, "config", p, out, cwd=tmp_path)
    d = {}
    for line in out.read_text().splitlines():
        k, *rest = line.split(" ")
        d[k] = rest
    return d


def test_config_reference_like(harness, tmp_path):
    d = _config(harness, tmp_path, REF_LIKE_INI)
    g = d["grid"]
    # nt = ceil(0.5 / 0.6e-3); nxt = nx + 2 ppad; nzt = nz + ppad + fpad (free surface)
    assert [int(g[0]), int(g[1]), int(g[2])] == [834, 601, 601]
    assert [int(x) for x in g[6:12]] == [50, 1, 1, 11, 623, 613]
    lam, mu = O.velocity_to_lame(2955.0, 2362.0, 7100.0)
    assert [float(x) for x in d["medium"][3:5]] == [lam, mu]
    assert d["fdm"][:4] == ["2", "1", "1", "1"]
    assert np.array_equal(np.array(d["hc"], float), O.holbergcoeff(2, 1))
    # receivers and sources moved onto the padded grid; shot list padded with 0
    assert d["rec_x"] == ["111", "311", "411"] and d["rec_z"] == ["1", "1", "1"]
    assert d["rec_shot"] == ["0", "0", "0"]
    assert d["src_x"] == ["311"] and d["src_z"] == ["3"]
    # FWI window: grid_interval_y is the key read (grid_interval_z is not);
    # zero right/bottom bounds become nxt-1 / nzt-1 plus padding, clamped
    assert d["fwi"] == ["3", "2", "1", "11", "623", "1", "613"]
    assert d["dev"][:2] == ["fp32", "full"]
    sig = O.create_wavelet(834, 0.6e-3, 1.0, 17.0, 0.0, 1)
    assert np.array_equal(np.array(d["src_signal0"], float), sig)
    prof = O.cpml_pro(10, 4.0, 17.0, 2.0, 1.0, 0.001, 3300.0, 0.6e-3)
    assert np.array_equal(np.array(d["profiles"], float), np.asarray(prof).ravel())


def test_config_device_section(harness, tmp_path):
    text = REF_LIKE_INI + """
[mi355x]
precision = fp64
gradient_store = checkpoint
checkpoint_interval = 50
tile = 16
observed = modified
gradient_scaling = clean
illumination = source_receiver
direction_smooth = 2.5
direction_depth_power = 2
direction_mute = 10
direction_taper = 30
"""
    d = _config(harness, tmp_path, text)
    assert d["dev"] == ["fp64", "checkpoint", "50", "16", "modified", "clean"]
    assert d["direction"] == ["source_receiver", "2.5", "2", "10", "30"]
    assert _config(harness, tmp_path, REF_LIKE_INI)["direction"] == ["source", "0", "0", "0", "0"]


def test_config_errors(harness, tmp_path):
    bad = REF_LIKE_INI.replace("x_seis = {100, 300, 400};", "x_seis = {100};")
    p = tmp_path / "b.ini"
    p.write_text(bad)
    with pytest.raises(RuntimeError, match="fewer than nseis"):
        run(harness, "config", p, tmp_path / "o.txt", cwd=tmp_path)
    p.write_text(REF_LIKE_INI + "\n[mi355x]\nprecision = fp16\n")
    with pytest.raises(RuntimeError, match="precision"):
        run(harness, "config", p, tmp_path / "o.txt", cwd=tmp_path)
    for bad_key in ("illumination = receiver", "direction_taper = -1"):
        p.write_text(REF_LIKE_INI + "\n[mi355x]\n" + bad_key + "\n")
        with pytest.raises(RuntimeError, match="illumination|direction"):
            run(harness, "config", p, tmp_path / "o.txt", cwd=tmp_path)


@pytest.mark.parametrize("L2,steps", [((3.0, 2.0, 2.5), (0.0, 1.0, 2.0)),
                                      ((1.0, 0.7, 0.8), (0.0, 0.01, 0.03)),
                                      ((5.0, 4.0, 4.5), (0.0, 2.0, 3.0))])
def test_calc_opt_step_dropin(harness, tmp_path, L2, steps):
    """calc_opt_step (CUDA_week2/fd_kernel.cuh:72) through fd_kernel.hpp: the vertex of the
    parabola through the three trial misfits, against numpy's quadratic fit."""
    out = tmp_path / "opt.txt"
    run(harness, "opt", *[repr(v) for v in L2 + steps], out)
    c = np.polyfit(steps, L2, 2)
    assert np.isclose(float(out.read_text()), -c[1] / (2 * c[0]), rtol=1e-12)


@pytest.mark.parametrize("smooth,power,mute,taper", [(0.0, 0.0, 0, 0), (2.5, 0.0, 0, 0), (0.0, 2.0, 3, 5),
                                                     (1.5, 1.0, 2, 4), (3.0, 2.0, 0, 6)])
def test_shape_direction(harness, tmp_path, smooth, power, mute, taper):
    """shape_direction of the C++ FWI loop ([mi355x] direction_smooth, direction_depth_power,
    direction_mute, direction_taper): scipy's gaussian_filter (mode "nearest"), then row r
    times ((r + 1) / rows)^power, the first `mute` rows zeroed and the next `taper` rows times
    0.5 (1 - cos(pi (r - mute + 0.5) / taper)) -- the weights of cuda_fwi_amd/fwi.py."""
    from scipy.ndimage import gaussian_filter
    nfz, nfx = 23, 31
    rng = np.random.default_rng(11)
    a = rng.standard_normal((3, nfz, nfx))
    src, out = tmp_path / "in.bin", tmp_path / "out.bin"
    a.tofile(src)
    run(harness, "shape", src, nfz, nfx, repr(smooth), repr(power), mute, taper, out)
    got = np.fromfile(out).reshape(3, nfz, nfx)
    r = np.arange(nfz)
    g = ((r + 1.0) / nfz) ** power if power > 0 else np.ones(nfz)
    g = np.where(r < mute, 0.0, g)
    t = (r >= mute) & (r < mute + taper)
    g[t] *= 0.5 * (1.0 - np.cos(np.pi * (r[t] - mute + 0.5) / taper))
    for k in range(3):
        ref = gaussian_filter(a[k], smooth, mode="nearest") if smooth > 0 else a[k]
        assert np.allclose(got[k], ref * g[:, None], rtol=1e-12, atol=1e-14), k
    if smooth == 0 and power == 0 and mute == 0 and taper == 0:
        assert np.array_equal(got, a)
