"""Ensemble of independent channels of uneven size: one hipGraph per RK3 step of all channels
must give exactly what each channel gives on its own."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ensemble_matches_single_channels(pkg):
    shapes = [(16, 16, 16, 1.8, 2800.0), (30, 20, 24, 2.0, 4000.0), (48, 36, 40, 2.2, 6875.0)]
    cos = [pkg.channel_setup(L, M, N, STR2=s) for L, M, N, s, _ in shapes]
    dt = np.array([1e-3, 8e-4, 5e-4])
    cv = np.array([1.0 / r for *_, r in shapes])
    fields = [pkg.poiseuille_perturbed(c, 40 + i, 0.1) for i, c in enumerate(cos)]
    singles = []
    for co, (q, p), d, v in zip(cos, fields, dt, cv):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
        for _ in range(2):
            ch.rk3_step(d, v)
        singles.append(ch.download())
        ch.close()
    ens = pkg.Ensemble(cos, pkg.PLAIN_FP64)
    for ch, (q, p), d, v in zip(ens.channels, fields, dt, cv):
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
    for _ in range(2):
        ens.step(dt, cv)
    ens.sync()
    for ch, (qs, ps) in zip(ens.channels, singles):
        q, p = ch.download()
        np.testing.assert_array_equal(q, qs)
        np.testing.assert_array_equal(p, ps)
        assert ch.max_div() < 1e-10
    ens.close()


def test_graph_step_matches_launches(pkg):
    co = pkg.channel_setup(32, 24, 16, STR2=1.8)
    q, p = pkg.poiseuille_perturbed(co, 5, 0.1)
    out = []
    for graph in (False, True):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        ch.set_graph(graph)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, 1e-3, 1 / 2800.0)
        for _ in range(3):
            ch.rk3_step(1e-3, 1 / 2800.0)
        out.append(ch.download())
        ch.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_ensemble_more_channels_than_lanes(pkg):
    """Ten channels on the eight capture lanes (two lanes carry two channels each), replayed
    three times with eager work between the replays: every replay equals the channels run one
    by one."""
    shapes = [(16 + 4 * (i % 3), 12 + 2 * (i % 4), 16 + 8 * (i % 2), 1.8 + 0.1 * (i % 3), 2800.0 + 500 * i)
              for i in range(10)]
    cos = [pkg.channel_setup(L, M, N, STR2=s) for L, M, N, s, _ in shapes]
    dt = np.array([1e-3 - 5e-5 * i for i in range(10)])
    cv = np.array([1.0 / r for *_, r in shapes])
    fields = [pkg.poiseuille_perturbed(c, 70 + i, 0.1) for i, c in enumerate(cos)]
    singles = [[] for _ in shapes]
    for i, (co, (q, p), d, v) in enumerate(zip(cos, fields, dt, cv)):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
        for _ in range(3):
            ch.rk3_step(d, v)
            singles[i].append(ch.download())
        ch.close()
    ens = pkg.Ensemble(cos, pkg.PLAIN_FP64)
    for ch, (q, p), d, v in zip(ens.channels, fields, dt, cv):
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
    for step in range(3):
        ens.step(dt, cv)
        ens.sync()
        for i, ch in enumerate(ens.channels):
            q, p = ch.download()
            np.testing.assert_array_equal(q, singles[i][step][0])
            np.testing.assert_array_equal(p, singles[i][step][1])
            assert ch.max_div() < 1e-10
    ens.close()
