"""x-z mean statistics of the TG field (PP_MEAN_ZX_FLOW_TG, processing/PP_MEAN_ZX.f90:14-307):
oracle and GPU against the reference's own output (tests/golden/meanzx.npz, written by the
reference routine through oracle/ref_harness.f90 for the fields after one RK3 step held in
tg_*.npz).  Tolerance: 1e-12 of the largest value of each statistic group (the velocity
derivative means are near zero by periodicity, so a relative bound per entry would be
meaningless there)."""
import numpy as np
import pytest

from oracle import oracle, setup

GROUPS = (("U1", 0, 4), ("UP", 4, 7), ("U2", 7, 13), ("U3", 13, 23), ("DVDL1", 23, 32),
          ("DVDLP", 32, 41), ("DVDL2", 41, 122))
NAMES = ("tg_16x16x16", "tg_20x18x30")


def check(st, sc, ref, rsc, tol=1e-12):
    for name, a, b in GROUPS:
        err = np.max(np.abs(st[:, a:b] - ref[:, a:b])) / np.max(np.abs(ref[:, a:b]))
        assert err < tol, (name, err)
    assert np.all(np.abs(sc - rsc) <= tol * np.abs(rsc))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference(golden, name):
    g, ref = golden(name + ".npz"), golden("meanzx.npz")
    L, M, N, STR2, REN, dt = (g["params"][i] for i in range(6))
    o = oracle.OracleTG(setup.Mesh(int(L), int(M), int(N), STR2=STR2), REN)
    o.set_fields(g["q_step1"], g["pr_step1"])
    st, sc = o.mean_zx()
    # the velocity moments are sums of the same terms in the same order: bit-exact
    np.testing.assert_array_equal(st[:, :23], ref[name + "_stats"][:, :23])
    check(st, sc, ref[name + "_stats"], ref[name + "_scal"], tol=1e-14)


def test_oracle_closed_forms():
    """A uniform flow u = U, v = 0, w = 0, p = P: U1 = (U, 0, 0, P), U2_11 = U^2,
    U3_111 = U^3, every derivative statistic zero."""
    L, M, N = 8, 6, 4
    o = oracle.OracleTG(setup.Mesh(L, M, N, STR2=1.5), 1000.0)
    q = np.zeros((3, N, M + 2, L))
    q[0] = 1.25
    p = np.full((N, M + 2, L), 0.5)
    o.set_fields(q, p)
    st, sc = o.mean_zx()
    assert np.allclose(st[:, 0], 1.25, rtol=0, atol=1e-15) and np.allclose(st[:, 3], 0.5, rtol=0, atol=1e-15)
    assert np.allclose(st[:, 7], 1.25 ** 2) and np.allclose(st[:, 13], 1.25 ** 3)
    assert np.allclose(st[:, 4], 1.25 * 0.5)
    # the y derivative of u uses the ghost rows (walls at zero): nonzero at J = 1 and M only
    inner = np.delete(st[1:M - 1, 23:], [1], axis=1)
    assert np.all(inner == 0.0)
    assert abs(sc[0] - 1.25) < 1e-14 and sc[1] == st[:, 0].max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_reference(pkg, golden, name):
    g, ref = golden(name + ".npz"), golden("meanzx.npz")
    L, M, N, STR2, REN, dt = (g["params"][i] for i in range(6))
    ch = pkg.ChannelTG(pkg.channel_setup(int(L), int(M), int(N), STR2=STR2), pkg.PLAIN_FP64)
    ch.upload(g["q_step1"], g["pr_step1"])
    st, sc = ch.PP_MEAN_ZX_FLOW_TG()
    check(st, sc, ref[name + "_stats"], ref[name + "_scal"])
    mon = ch.monitor(REN)
    assert mon["cf_bottom"] > 0 and mon["cf_top"] > 0
    ch.close()


@pytest.mark.gpu
def test_gpu_matches_oracle_large(pkg):
    L, M, N, STR2, REN = 256, 96, 128, 2.0, 6875.0
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    q, p = pkg.poiseuille_perturbed(co, seed=33, amp=0.2)
    p[:, 1:M + 1, :] = np.random.default_rng(4).standard_normal((N, M, L)) * 0.01
    o = oracle.OracleTG(setup.Mesh(L, M, N, STR2=STR2), REN)
    o.set_fields(q, p)
    st0, sc0 = o.mean_zx()
    ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    ch.upload(q, p)
    st, sc = ch.PP_MEAN_ZX_FLOW_TG()
    check(st, sc, st0, sc0)
    ch.close()
