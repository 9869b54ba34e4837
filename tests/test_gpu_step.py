"""The step-level driver logic on the device: CFL_tg and DIVGCK_tg against the reference's
values, the CFL a step forms of its new field, NaN propagation, graphs that stay correct when
DT changes between calls, and SOLVE's time-step rule run on the device (chs_run_tg) against
the same rule run from the host and against the oracle."""
import numpy as np
import pytest

from oracle import oracle, setup

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", ["tg_16x16x16.npz", "tg_20x18x30.npz"])
def test_cfl_and_divergence_check_match_reference(pkg, golden, name):
    """chs_cfl_tg and chs_max_div on the reference's set-up arrays equal CFLMM_tg and
    MAXDIVGV_tg of the reference bit for bit (a field with O(1) divergence, the projected one
    and one RK3 step later)."""
    g = golden(name)
    ref = golden("cfl_div_tg.npz")
    L, M, N, STR2 = (g["params"][i] for i in range(4))
    L, M, N = int(L), int(M), int(N)
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    co["dyfi"] = np.asarray(g["s_DYFI"][:M])
    for tag in ("init", "ns0", "step1"):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        ch.upload(g[f"q_{tag}"], g[f"pr_{tag}"])
        c, d = ch.cfl(1.0), ch.max_div()
        print(f"  {name} {tag}: CFLMM {c!r} (reference {ref[f'{L}x{M}x{N}_{tag}_cfl'][0]!r}), "
              f"max div {d!r} (reference {ref[f'{L}x{M}x{N}_{tag}_div'][0]!r})")
        assert c == ref[f"{L}x{M}x{N}_{tag}_cfl"][0]
        assert d == ref[f"{L}x{M}x{N}_{tag}_div"][0]
        ch.close()


@pytest.mark.parametrize("graph", [False, True])
def test_step_forms_cfl_of_new_field(pkg, graph):
    """The CFLMM a step forms inside its last correction equals the standalone CFL_tg of the
    downloaded field bit for bit (same per-cell arithmetic, corrected neighbours recomputed)."""
    L, M, N, REN, dt = 32, 24, 16, 2800.0, 1e-3
    co = pkg.channel_setup(L, M, N, STR2=1.8)
    mesh = setup.Mesh(L, M, N, STR2=1.8)
    mesh.DYFI = np.asarray(co["dyfi"])
    q, pr = oracle.initial_field(mesh, 4)
    ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    ch.upload(q, pr)
    ch.set_graph(graph)
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    for _ in range(2):
        ch.rk3_step(dt, 1.0 / REN)
        in_step = ch.step_state()["cflmm"]
        qn, _ = ch.download()
        assert in_step == oracle.cfl_tg(mesh, qn), (in_step, oracle.cfl_tg(mesh, qn))
        assert ch.cfl(2.0) == 2.0 * in_step
    ch.close()


def test_nonfinite_field_is_reported(pkg):
    L, M, N = 32, 16, 16
    co = pkg.channel_setup(L, M, N, STR2=1.8)
    mesh = setup.Mesh(L, M, N, STR2=1.8)
    q, pr = oracle.initial_field(mesh, 5)
    q[0][3, 5, 7] = np.nan
    ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    ch.upload(q, pr)
    with pytest.raises(pkg.ChsError):
        ch.cfl(1.0)
    with pytest.raises(pkg.ChsError):
        ch.max_div()
    ch.close()


def test_graph_replay_after_dt_change(pkg):
    """A step graph captured at dt1 stays right after substeps at dt2 (which rebuild the
    Crank-Nicolson tables and the device DT) and a refinement-count change; compared bit for
    bit with the same sequence of ungraphed launches."""
    L, M, N, REN = 48, 36, 40, 4000.0
    dt1, dt2 = 8e-4, 3e-4
    co = pkg.channel_setup(L, M, N, STR2=2.0)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    q, pr = oracle.initial_field(mesh, 6)
    out = []
    for graph in (True, False):
        ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
        ch.upload(q, pr)
        ch.set_graph(graph)
        cv = 1.0 / REN
        ch.SOLVERRK3_MOM_tg(0, dt1, cv)
        ch.rk3_step(dt1, cv)
        for ns in (1, 2, 3):
            ch.SOLVERRK3_MOM_tg(ns, dt2, cv)
        ch.rk3_step(dt1, cv)
        ch.rk3_step(dt2, cv)
        ch.rk3_step(dt1, cv)
        out.append(ch.download())
        ch.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("graph", [False, True])
def test_device_time_step_rule(pkg, graph):
    """chs_run_tg (DT chosen on the device each step) against the rule run from the host
    (cfl -> next_dt -> rk3_step) and against the oracle stepped at the host's DTs: the DT
    sequence agrees to 1e-14 and the fields to rounding (the device run keeps the band widths
    of DT0)."""
    L, M, N, REN = 32, 32, 32, 2800.0
    dt0, cflgv, dtmin, nst = 1e-3, 0.0025, 1e-6, 6
    co = pkg.channel_setup(L, M, N, STR2=1.8)
    mesh = setup.Mesh(L, M, N, STR2=1.8)
    q, pr = oracle.initial_field(mesh, 7)
    cv = 1.0 / REN
    dev = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    dev.upload(q, pr)
    dev.SOLVERRK3_MOM_tg(0, dt0, cv)
    q0, p0 = dev.download()
    dev.set_graph(graph)
    dev.set_step_state(dt0, 0.0)
    dts_dev = []
    for _ in range(nst):
        dev.run(1, dt0, cflgv, dtmin, cv)
        st = dev.step_state()
        dts_dev.append(st["dt"])
    assert not st["stopped"]
    assert abs(st["phytime"] - sum(dts_dev)) < 1e-15
    host = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    host.upload(q0, p0)
    o = oracle.OracleTG(mesh, REN)
    o.set_fields(q0, p0)
    dt, dts_host = dt0, []
    for _ in range(nst):
        dt = pkg.next_dt(dt, dt0, cflgv, host.cfl(1.0), dtmin)
        dts_host.append(dt)
        host.rk3_step(dt, cv)
        o.step(dt)
    print(f"  DT sequence {dts_dev}")
    # the rule is the same arithmetic on the same CFLMM; the fields (and so CFLMM) may differ in
    # the last bit where the device run's wider bands add entries below 2^-64 of the diagonal
    assert np.allclose(dts_dev, dts_host, rtol=1e-14, atol=0.0), (dts_dev, dts_host)
    assert len(set(dts_dev)) > 2            # the CFL limit moved DT
    qd, pd = dev.download()
    qh, ph = host.download()
    e_h = max(rel(qd[c], qh[c]) for c in range(3))
    e_o = max(rel(qd[c], o.q[c]) for c in range(3))
    print(f"  device vs host rule: velocity {e_h:.2e}, pressure {rel(pd, ph):.2e}; vs oracle velocity "
          f"{e_o:.2e}, pressure {rel(pd, o.pr):.2e}")
    assert e_h < 1e-12 and rel(pd, ph) < 1e-10
    assert e_o < 1e-10 and rel(pd, o.pr) < 1e-10
    dev.close()
    host.close()


def test_device_rule_stops_like_the_reference(pkg):
    """DT / DT0 < 0.01 sets the stop flag (SOLVE.f90:199-207)."""
    L, M, N, REN = 32, 16, 16, 2800.0
    co = pkg.channel_setup(L, M, N, STR2=1.8)
    q, pr = pkg.poiseuille_perturbed(co, seed=3, amp=0.05)
    ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    ch.upload(q, pr)
    ch.set_step_state(1e-3, 0.0)
    ch.run(1, 1e-3, 1e-6, 1e-9, 1.0 / REN)
    st = ch.step_state()
    assert st["stopped"] and st["dt"] < 1e-5
    ch.close()
