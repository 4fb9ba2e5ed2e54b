"""Invariants that pin the oracle's explosive source (source comp 4: sxx and szz).

The reference has only a vz point force (CUDA_week2/forward_kernel_PSV_GPU.cu:1038-1040);
the explosive type used by the 201 x 201 configuration is an extension, so no reference
output covers it.  In a homogeneous medium, away from every boundary, the mathematics fixes
three properties of the discrete scheme that any mistake in the injection breaks:

* mirror symmetry about the source column and row (vz even and vx odd in x, vz odd in z,
  the normal stresses even, szx odd), on the staggered positions of each field;
* no S wave: a pure pressure source drives only the curl-free mode of the staggered
  operators (they are tensor-product stencils and commute, so the discrete curl of a
  discrete gradient vanishes), so curl v = D_z vx - D_x vz stays at rounding level while
  div v does not.  A vz force, or a source on sxx or szz alone, radiates S waves;
* linearity in the amplitude (exact for a power-of-two factor).
"""
import numpy as np
import pytest

import psv_oracle as O

N = 101           # odd: the source sits on the middle cell, the grid is symmetric about it
C0 = (N - 1) // 2


def _run(comp=4, amp=1.0, fdorder=4, nt=150):
    """Homogeneous 101 x 101 grid, no CPML, no free surface, one source at the centre; the
    wave stays well inside the grid for nt = 150 (40 Hz Ricker, dt 0.5 ms, dx 5 m)."""
    vp, vs, rho = 2500.0, 1443.0, 2000.0
    mu = np.full((N, N), vs * vs * rho)
    lam = np.full((N, N), vp * vp * rho) - 2 * mu
    dt = 5e-4
    sig = O.create_wavelet(nt, dt, amp, 40.0, 0.0, 1)[None]
    src = dict(x=[C0], z=[C0], comp=[comp], shot=[0])
    pb = O.Problem(nt=nt, nzt=N, nxt=N, fdorder=fdorder, npml=0, dt=dt, dx=5.0, dz=5.0, fsurf=0,
                   hc=O.holbergcoeff(fdorder, 1), lam=lam, mu=mu, rho=np.full((N, N), rho),
                   src=src, src_signal=sig, rec=None)
    return pb, O.forward(pb)


def _curl_div(pb, f):
    """Discrete curl at the szx points and divergence at the normal-stress points, with the
    scheme's own staggered stencils (order 2 uses hc[1] only)."""
    h1 = pb.hc[1]
    h2 = pb.hc[2] if pb.fdorder == 4 else 0.0
    vx, vz = f["vx"], f["vz"]
    # D_z vx and D_x vz at (iz + 1/2, ix + 1/2), iz, ix = 2 .. N-3
    vx_z = h1 * (vx[3:-1, :] - vx[2:-2, :]) + h2 * (vx[4:, :] - vx[1:-3, :])
    vz_x = h1 * (vz[:, 3:-1] - vz[:, 2:-2]) + h2 * (vz[:, 4:] - vz[:, 1:-3])
    curl = vx_z[:, 2:-2] - vz_x[2:-2, :]
    # D_x vx and D_z vz at (iz, ix)
    vx_x = h1 * (vx[:, 2:-2] - vx[:, 1:-3]) + h2 * (vx[:, 3:-1] - vx[:, :-4])
    vz_z = h1 * (vz[2:-2, :] - vz[1:-3, :]) + h2 * (vz[3:-1, :] - vz[:-4, :])
    div = vx_x[2:-2, :] + vz_z[:, 2:-2]
    return curl, div


@pytest.mark.parametrize("fdorder", [2, 4])
def test_explosive_source_mirror_symmetry(fdorder):
    _, f = _run(fdorder=fdorder)
    k = np.arange(1, C0 - 3)
    rel = lambda a, b, ref: np.abs(a - b).max() / np.abs(ref).max()
    vx, vz, sxx, szz, szx = (f[n] for n in ("vx", "vz", "sxx", "szz", "szx"))
    assert np.abs(vz).max() > 0 and np.abs(vx).max() > 0
    # x mirror about column C0: fields on integer x (vz, sxx, szz) even; fields half a cell
    # right of their index (vx, szx: ix + 1/2 mirrors to C0 - k - 1 + 1/2) odd
    assert rel(vz[:, C0 + k], vz[:, C0 - k], vz) < 1e-13
    assert rel(sxx[:, C0 + k], sxx[:, C0 - k], sxx) < 1e-13
    assert rel(szz[:, C0 + k], szz[:, C0 - k], szz) < 1e-13
    assert rel(vx[:, C0 + k], -vx[:, C0 - k - 1], vx) < 1e-13
    assert rel(szx[:, C0 + k], -szx[:, C0 - k - 1], szx) < 1e-13
    # z mirror: half-z fields (vz, szx) odd, integer-z fields even
    assert rel(vz[C0 + k, :], -vz[C0 - k - 1, :], vz) < 1e-13
    assert rel(vx[C0 + k, :], vx[C0 - k, :], vx) < 1e-13
    assert rel(sxx[C0 + k, :], sxx[C0 - k, :], sxx) < 1e-13
    # x <-> z: with dx = dz the transpose maps sxx to szz and vx to vz (both half-shifted)
    assert rel(sxx, szz.T, sxx) < 1e-13
    assert rel(vx, vz.T, vx) < 1e-13


@pytest.mark.parametrize("fdorder", [2, 4])
def test_explosive_source_radiates_no_s_wave(fdorder):
    pb, f = _run(fdorder=fdorder)
    curl, div = _curl_div(pb, f)
    assert np.abs(div).max() > 0
    assert np.abs(curl).max() < 1e-12 * np.abs(div).max()
    # the same measure sees the S wave of a vertical point force (the reference's source)
    pb2, f2 = _run(comp=2, fdorder=fdorder)
    curl2, div2 = _curl_div(pb2, f2)
    assert np.abs(curl2).max() > 0.1 * np.abs(div2).max()


def test_explosive_source_is_linear_in_amplitude():
    _, f1 = _run(amp=1.0)
    _, f2 = _run(amp=2.0)
    for k in ("vx", "vz", "sxx", "szx", "szz"):
        assert np.array_equal(2.0 * f1[k], f2[k]), k
