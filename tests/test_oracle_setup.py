"""Oracle set-up restatement (oracle/setup.py) against the reference's own set-up dump
(golden fixtures written by the compiled reference) and closed forms."""
import numpy as np
import pytest

from oracle import setup

CASES = ["tg_16x16x16.npz", "tg_20x18x30.npz"]


@pytest.mark.parametrize("name", CASES)
def test_setup_matches_reference(golden, name):
    g = golden(name)
    L, M, N, STR2 = (g["params"][i] for i in range(4))
    m = setup.Mesh(int(L), int(M), int(N), STR2=STR2)
    # index maps: bit-exact
    for k in ("IPV", "IMV", "KPV", "KMV", "KSYM", "JGPV", "JGMV"):
        np.testing.assert_array_equal(getattr(m, k), g["s_" + k], err_msg=k)
    np.testing.assert_array_equal(setup.TGAM, g["s_TGAM"])
    np.testing.assert_array_equal(setup.TROH, g["s_TROH"])
    np.testing.assert_array_equal(setup.TALP, g["s_TALP"])
    for k in ("DX", "DZ", "DXI", "DZI", "DXQI", "DZQI"):
        assert getattr(m, k) == g["s_" + k], k
    rt = 1e-13
    np.testing.assert_allclose(m.VOLM, g["s_VOLM"], rtol=rt)
    for k, sl in (("YND", slice(None)), ("YCC", slice(None)), ("DYFI", slice(None)),
                  ("DYCI", slice(1, None)), ("WFF", slice(1, None)), ("WFB", slice(1, None))):
        np.testing.assert_allclose(getattr(m, k)[sl], g["s_" + k], rtol=rt, atol=1e-15, err_msg=k)
    for k in ("AMVR", "ACVR", "APVR"):
        np.testing.assert_allclose(getattr(m, k)[:, 1:], g["s_" + k], rtol=rt, atol=1e-12, err_msg=k)
    for k in ("AMPH", "ACPH", "APPH", "AK1", "AK3"):
        np.testing.assert_allclose(getattr(m, k), g["s_" + k], rtol=rt, atol=1e-12, err_msg=k)
    np.testing.assert_allclose(m.VINI, g["s_VINI"], rtol=rt, atol=1e-15)


def test_setup_closed_forms():
    m = setup.Mesh(32, 24, 16, STR2=2.0)
    # tanh mesh spans the channel [-1, 1] and is antisymmetric about the centre line
    assert m.YND[0] == -1.0 and abs(m.YND[-1] - 1.0) < 1e-15
    np.testing.assert_allclose(m.YND, -m.YND[::-1], atol=1e-15)
    # cell widths telescope to the channel height; VOLM = L*N*height
    assert abs(np.sum(1.0 / m.DYFI) - 2.0) < 1e-14
    assert abs(m.VOLM - 32 * 16 * 2.0) < 1e-11
    # Neumann-folded Poisson stencil: rows sum to zero, ends decoupled from the ghosts
    np.testing.assert_allclose(m.AMPH + m.ACPH + m.APPH, 0.0, atol=1e-10)
    assert m.AMPH[0] == 0.0 and m.APPH[-1] == 0.0
    # interpolation weights partition unity
    np.testing.assert_allclose(m.WFF[1:] + m.WFB[1:], 1.0, atol=1e-15)
    # modified wavenumbers: 0 at k=0, 4/dx^2 at Nyquist, symmetric
    assert m.AK1[0] == 0.0 and abs(m.AK1[16] - 4.0 * m.DXQI) < 1e-12
    np.testing.assert_allclose(m.AK1[1:], m.AK1[1:][::-1], rtol=1e-13)
    # RK3 consistency: sum of alpha = 1 (one full step), gamma_1 + zeta_... per Williamson
    assert abs(np.sum(setup.TALP[1:]) - 1.0) < 1e-15
    # Poiseuille profile has unit bulk velocity to discretisation accuracy
    bulk = np.sum(m.VINI / m.DYFI) / 2.0
    assert abs(bulk - 1.0) < 5e-3   # midpoint rule, O(dy^2)
