"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/*.h (no device calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(psvfwi_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    from cuda_fwi_amd import psvfwi as P
    if not os.path.exists(P.LIB_PATH):
        pytest.fail("libpsvfwi.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(P.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(P.EXPORTS) <= set(names)


def test_parabola_step_is_host_only_and_matches_polyfit():
    import numpy as np
    from cuda_fwi_amd import psvfwi as P
    sl = np.array([0.0, 0.01, 0.015])
    L2 = 3.0 * (sl - 0.012) ** 2 + 0.5
    assert abs(P.parabola_step(L2, sl) - 0.012) < 1e-12
    with pytest.raises(P.PsvfwiError):
        P.parabola_step([1.0, 1.0, 1.0], [0.0, 1.0, 2.0])


def test_library_setup_functions_equal_the_oracle():
    """The product's host setup (holbergcoeff, cpml_pro, average_parameters,
    create_wavelet) reproduces the oracle bit for bit."""
    import numpy as np
    import psv_oracle as O
    from cuda_fwi_amd import psvfwi as P
    for order in (2, 4, 6, 8, 10, 12):
        for err in range(5):
            assert np.array_equal(P.holberg(order, err), O.holbergcoeff(order, err))
    for args in [(10, 4.0, 17.0, 2.0, 1.0, 1e-3, 3300.0, 6e-4), (40, 5.0, 12.0, 2.0, 1.0, 1e-3, 4500.0, 4e-4),
                 (7, 2.5, 30.0, 3.0, 2.5, 1e-4, 2000.0, 1e-4)]:
        assert np.array_equal(P.cpml_profiles(*args), O.cpml_pro(*args))
    rng = np.random.default_rng(0)
    mu = rng.uniform(1e9, 4e9, (33, 47))
    rho = rng.uniform(1000, 3000, (33, 47))
    mu[4, 5] = 0.0
    rho[10, 10:13] = 1e-5
    rho[11, 10] = 1e-5
    for a, b in zip(P.average_parameters(mu, rho), O.average_parameters(mu, rho)):
        assert np.array_equal(a, b)
    assert np.array_equal(P.ricker(3000, 4e-4, 2.0, 12.0, 0.01), O.create_wavelet(3000, 4e-4, 2.0, 12.0, 0.01))


def test_plan_create_rejects_bad_grids_before_touching_a_device():
    """Argument errors come back as status codes with a message (the
    reference exits instead); these checks run before any device call."""
    from cuda_fwi_amd import psvfwi as P
    lib = ctypes.CDLL(P.LIB_PATH)
    lib.psvfwi_last_error.restype = ctypes.c_char_p

    def create(**kw):
        g = P.GridDesc()
        base = dict(nzt=60, nxt=80, fdorder=4, npml=8, fsurf=1, dt=5e-4, dx=5.0, dz=5.0)
        base.update(kw)
        for k, v in base.items():
            setattr(g, k, v)
        out = ctypes.c_void_p()
        st = lib.psvfwi_plan_create(ctypes.byref(g), kw.get("prec", 0), ctypes.byref(out))
        return st, lib.psvfwi_last_error().decode()

    assert create(fdorder=3)[0] == -3                       # PSVFWI_E_ORDER
    st, msg = create(nzt=4)
    assert st == -1 and "smaller than the stencil" in msg
    st, msg = create(npml=40)
    assert st == -1 and "CPML wider" in msg
    st, msg = create(dt=0.0)
    assert st == -1 and "dt, dx, dz" in msg
    st, msg = create(prec=7)
    assert st == -1 and "precision" in msg
    st, msg = create(nzt=30000, nxt=30000, npml=8)
    assert st == -1 and "too large" in msg
    assert lib.psvfwi_plan_create(None, 0, None) == -1


def test_chain_gradient_equals_the_oracle():
    """The library's host chain rule (coefficient gradients -> lam, mu, rho)
    equals the oracle's bit for bit, on a model with a zero-mu cell (its
    mu_zx averages are zero and contribute nothing)."""
    import numpy as np
    import psv_oracle as O
    from cuda_fwi_amd import psvfwi as P
    rng = np.random.default_rng(3)
    nzt, nxt = 13, 17
    mu = rng.uniform(1e9, 3e9, (nzt, nxt))
    mu[4, 5] = 0.0
    rho = rng.uniform(1500, 2500, (nzt, nxt))
    mzx, rzp, rxp = O.average_parameters(mu, rho)
    g = dict(g_lam=rng.standard_normal((nzt, nxt)), g_mu=rng.standard_normal((nzt, nxt)),
             g_mu_zx=rng.standard_normal((nzt - 1, nxt - 1)), g_rho_zp=rng.standard_normal((nzt - 1, nxt - 1)),
             g_rho_xp=rng.standard_normal((nzt - 1, nxt - 1)))
    ours = P.chain_gradient(mu, mzx, rzp, rxp, g)
    ref = O.chain_gradient(mu, mzx, rzp, rxp, g)
    for a, b in zip(ours, ref):
        assert np.array_equal(a, b)
    assert np.isfinite(ours[1]).all()
    # the chain rule is the derivative of average_parameters: a directional check
    dmu = rng.standard_normal((nzt, nxt)) * 1e5
    dmu[4, 5] = 0.0
    drho = rng.standard_normal((nzt, nxt))
    h = 1e-3
    up = O.average_parameters(mu + h * dmu, rho + h * drho)
    dn = O.average_parameters(mu - h * dmu, rho - h * drho)
    d = [(u - v) / (2 * h) for u, v in zip(up, dn)]
    lin = np.sum(g["g_mu_zx"] * d[0]) + np.sum(g["g_rho_zp"] * d[1]) + np.sum(g["g_rho_xp"] * d[2])
    ch = np.sum((ours[1] - g["g_mu"]) * dmu) + np.sum(ours[2] * drho)
    assert abs(ch - lin) < 1e-6 * abs(lin)
