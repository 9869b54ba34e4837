"""C-ABI library: loads, exports every entry point of include/chapsim_hip.h, and its host
set-up restatement (chs_channel_setup) agrees with the reference's set-up dump.  No GPU."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "chapsim_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chs_[a-z_0-9]+)\s*\(", txt)))


def test_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    from chapsim_amd import tg
    assert set(names) == set(tg.SIGNATURES), "ctypes signatures out of sync with the header"
    assert lib.chs_version().decode().startswith("chapsim_hip")


@pytest.mark.parametrize("name", ["tg_16x16x16.npz", "tg_20x18x30.npz"])
def test_channel_setup_matches_reference(pkg, golden, name):
    g = golden(name)
    L, M, N, STR2 = int(g["params"][0]), int(g["params"][1]), int(g["params"][2]), g["params"][3]
    c = pkg.channel_setup(L, M, N, STR2=STR2)
    for k, ref in (("dxi", "s_DXI"), ("dzi", "s_DZI"), ("dxqi", "s_DXQI"), ("dzqi", "s_DZQI")):
        assert c[k] == g[ref], k
    np.testing.assert_allclose(c["volm"], g["s_VOLM"], rtol=1e-13)
    np.testing.assert_allclose(c["dyfi"], g["s_DYFI"], rtol=1e-13)
    np.testing.assert_allclose(c["dyci"], g["s_DYCI"], rtol=1e-13)
    np.testing.assert_allclose(c["wff"], g["s_WFF"], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(c["wfb"], g["s_WFB"], rtol=1e-13, atol=1e-15)
    for k in ("amvr", "acvr", "apvr"):
        np.testing.assert_allclose(c[k].reshape(3, M + 1), g["s_" + k.upper()], rtol=1e-13, atol=1e-12)
    for k in ("amph", "acph", "apph", "ak1", "ak3"):
        np.testing.assert_allclose(c[k], g["s_" + k.upper()], rtol=1e-13, atol=1e-12)
    np.testing.assert_array_equal(c["tgam"], g["s_TGAM"])
    np.testing.assert_array_equal(c["troh"], g["s_TROH"])
    np.testing.assert_array_equal(c["talp"], g["s_TALP"])


def test_rejects_bad_sizes(pkg):
    with pytest.raises(pkg.ChsError):
        pkg.channel_setup(2, 16, 16)
