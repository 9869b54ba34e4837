"""Ensemble of independent channels of uneven size: one hipGraph per RK3 step of all channels
must give exactly what each channel gives on its own.  Batched (one launch per kernel kind
over all channels) the ensemble runs the x-solve in its own kernel and the LDS transforms for
every mesh, so it equals single channels tuned conv_fx 0, fft_reg 0; on capture lanes (each
channel's own launches) it equals single channels with their default shapes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BATCHED_SHAPES = dict(conv_fx=0, fft_reg=0, cnzy=0)


@pytest.mark.parametrize("batched", [True, False])
def test_ensemble_matches_single_channels(pkg, batched):
    shapes = [(16, 16, 16, 1.8, 2800.0), (30, 20, 24, 2.0, 4000.0), (48, 36, 40, 2.2, 6875.0)]
    cos = [pkg.channel_setup(L, M, N, STR2=s) for L, M, N, s, _ in shapes]
    dt = np.array([1e-3, 8e-4, 5e-4])
    cv = np.array([1.0 / r for *_, r in shapes])
    fields = [pkg.poiseuille_perturbed(c, 40 + i, 0.1) for i, c in enumerate(cos)]
    singles = []
    for co, (q, p), d, v in zip(cos, fields, dt, cv):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        if batched:
            ch.set_tuning(**BATCHED_SHAPES)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
        for _ in range(2):
            ch.rk3_step(d, v)
        singles.append(ch.download())
        ch.close()
    ens = pkg.Ensemble(cos, pkg.PLAIN_FP64)
    ens.set_batched(batched)
    for ch, (q, p), d, v in zip(ens.channels, fields, dt, cv):
        if batched:
            ch.set_tuning(**BATCHED_SHAPES)   # the NS = 0 substep runs the channel's own kernels
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


@pytest.mark.parametrize("batched", [True, False])
def test_ensemble_more_channels_than_lanes(pkg, batched):
    """Ten channels (on the eight capture lanes, two lanes carrying two channels each; or in the
    batched launches), replayed three times with eager work between the replays: every replay
    equals the channels run one by one."""
    shapes = [(16 + 4 * (i % 3), 12 + 2 * (i % 4), 16 + 8 * (i % 2), 1.8 + 0.1 * (i % 3), 2800.0 + 500 * i)
              for i in range(10)]
    cos = [pkg.channel_setup(L, M, N, STR2=s) for L, M, N, s, _ in shapes]
    dt = np.array([1e-3 - 5e-5 * i for i in range(10)])
    cv = np.array([1.0 / r for *_, r in shapes])
    fields = [pkg.poiseuille_perturbed(c, 70 + i, 0.1) for i, c in enumerate(cos)]
    singles = [[] for _ in shapes]
    for i, (co, (q, p), d, v) in enumerate(zip(cos, fields, dt, cv)):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        if batched:
            ch.set_tuning(**BATCHED_SHAPES)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, d, v)
        for _ in range(3):
            ch.rk3_step(d, v)
            singles[i].append(ch.download())
        ch.close()
    ens = pkg.Ensemble(cos, pkg.PLAIN_FP64)
    ens.set_batched(batched)
    for ch, (q, p), d, v in zip(ens.channels, fields, dt, cv):
        if batched:
            ch.set_tuning(**BATCHED_SHAPES)   # the NS = 0 substep runs the channel's own kernels
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


@pytest.mark.parametrize("batched", [True, False])
def test_config4_ensemble_mixed(pkg, batched):
    """The batched workload itself: 32 seeded channels of uneven size (64^3 .. 192 x 128 x 192,
    Re_tau 180 .. 395), mixed precision, one graph per RK3 step (batched launches, or each
    channel's own on capture lanes), three steps.  Each channel equals its own single-channel
    run bit for bit, keeps its divergence below 1e-9, its mass flux (1e-9 relative) and its
    wall values."""
    from chapsim_amd.init import ensemble_cases
    cases = ensemble_cases()
    print(f"config-4 ensemble: {len(cases)} channels, {sum(c['L'] * c['M'] * c['N'] for c in cases)} points",
          flush=True)
    cos = [pkg.channel_setup(c["L"], c["M"], c["N"], c["HX"], c["HZ"], c["STR2"]) for c in cases]
    dt = np.array([c["dt"] for c in cases])
    cv = np.array([1.0 / c["REN"] for c in cases])
    ens = pkg.Ensemble(cos, pkg.MIXED)
    ens.set_batched(batched)
    starts, bulk0 = [], []
    for i, (ch, co) in enumerate(zip(ens.channels, cos)):
        q, p = pkg.poiseuille_perturbed(co, 1000 + i, 0.1)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, dt[i], cv[i])
        qa, pa = ch.download()
        starts.append((qa, pa))
        bulk0.append(np.sum(qa[0][:, 1:co["M"] + 1] / np.asarray(co["dyfi"])[None, :, None]))
    for _ in range(3):
        ens.step(dt, cv)
    ens.sync()
    for i, (ch, co) in enumerate(zip(ens.channels, cos)):
        M = co["M"]
        q, p = ch.download()
        assert ch.max_div() < 1e-9, i
        b = np.sum(q[0][:, 1:M + 1] / np.asarray(co["dyfi"])[None, :, None])
        assert abs(b - bulk0[i]) / abs(bulk0[i]) < 1e-9, i
        assert np.all(q[:, :, 0] == 0) and np.all(q[:, :, M + 1] == 0) and np.all(q[1, :, 1] == 0)
        one = pkg.ChannelTG(co, pkg.MIXED)
        if batched:
            one.set_tuning(**BATCHED_SHAPES)
        one.upload(*starts[i])
        for _ in range(3):
            one.rk3_step(dt[i], cv[i])
        q1, p1 = one.download()
        one.close()
        np.testing.assert_array_equal(q, q1)
        np.testing.assert_array_equal(p, p1)
    ens.close()
