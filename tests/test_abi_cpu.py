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
