"""Mixed-precision form (FP32 stencils, Crank-Nicolson stencils and FFTs; FP64 state,
reductions, pivots of the lowest modes, divergence residual) against the FP64 oracle and the
plain FP64 form on the same inputs.  Stated bars: per step, u, v, w max-norm error < 1e-4
relative, pressure < 1e-3, max divergence < 1e-9; over the window, the x-z mean u profile
< 1e-5 and Re_tau (from the wall shear of the mean profile) < 1e-5."""
import numpy as np
import pytest

from oracle import oracle, setup

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def mean_u(q, M):
    return q[0][:, 1:M + 1].mean(axis=(0, 2))


def re_tau(q, dyfi, REN, M):
    """Re_tau = REN sqrt(|tau_w|) from the x-z mean u of the first cell at each wall (the wall
    shear of PP_MONITOR_TG's Cf, processing/PP_SCN_MONITOR.f90:147-253, bulk velocity 1),
    averaged over the two walls."""
    um = mean_u(q, M)
    tb = um[0] * dyfi[0] * 2.0 / REN
    tt = um[M - 1] * dyfi[M - 1] * 2.0 / REN
    return REN * 0.5 * (np.sqrt(abs(tb)) + np.sqrt(abs(tt)))


@pytest.mark.parametrize("L,M,N,STR2,REN,dt,steps", [(64, 64, 64, 1.8, 2800.0, 1e-3, 10),
                                                    (48, 40, 32, 2.0, 6875.0, 5e-4, 6)])
def test_mixed_against_oracle_per_step(pkg, L, M, N, STR2, REN, dt, steps):
    print(f"mixed vs oracle {L}x{M}x{N}, {steps} steps", flush=True)
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    q0, p0 = pkg.poiseuille_perturbed(co, 9, 0.1)
    ch = pkg.ChannelTG(co, pkg.MIXED)
    ch.set_refinement(0, 2)
    ch.upload(q0, p0)
    o = oracle.OracleTG(mesh, REN)
    o.set_fields(q0, p0)
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    o.substep(0, dt)
    dyfi = np.asarray(co["dyfi"])
    worst = dict(u=0.0, v=0.0, w=0.0, p=0.0, div=0.0, mean=0.0, retau=0.0)
    for n in range(steps):
        ch.rk3_step(dt, cv)
        o.step(dt)
        q, p = ch.download()
        e = dict(u=rel(q[0], o.q[0]), v=rel(q[1], o.q[1]), w=rel(q[2], o.q[2]), p=rel(p, o.pr),
                 div=ch.max_div(), mean=rel(mean_u(q, M), mean_u(o.q, M)),
                 retau=abs(re_tau(q, dyfi, REN, M) / re_tau(o.q, dyfi, REN, M) - 1.0))
        print("  step %d: " % (n + 1) + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
        for k in worst:
            worst[k] = max(worst[k], e[k])
    print("  worst over the window: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    assert worst["u"] < 1e-4 and worst["v"] < 1e-4 and worst["w"] < 1e-4
    assert worst["p"] < 1e-3
    assert worst["div"] < 1e-9
    assert worst["mean"] < 1e-5
    assert worst["retau"] < 1e-5
    ch.close()


def run(pkg, mode, L, M, N, STR2, REN, dt, steps, seed=9, nref_step=0):
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    ch = pkg.ChannelTG(co, mode)
    if mode == pkg.MIXED:
        ch.set_refinement(nref_step, 2)
    q, p = pkg.poiseuille_perturbed(co, seed, 0.1)
    ch.upload(q, p)
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    divs = [ch.max_div()]
    for _ in range(steps):
        ch.rk3_step(dt, 1.0 / REN)
        divs.append(ch.max_div())
    q, p = ch.download()
    ch.close()
    return q, p, divs, co


def test_mixed_vs_plain_128(pkg):
    L, M, N, STR2, REN, dt = 128, 96, 128, 2.0, 6875.0, 5e-4
    qp, pp, dp, co = run(pkg, pkg.PLAIN_FP64, L, M, N, STR2, REN, dt, 5)
    qm, pm, dm, _ = run(pkg, pkg.MIXED, L, M, N, STR2, REN, dt, 5)
    e_mean = rel(mean_u(qm, M), mean_u(qp, M))
    e_vel = max(rel(qm[c], qp[c]) for c in range(3))
    e_p = rel(pm, pp)
    print(f"mixed vs plain {L}x{M}x{N}: mean-u {e_mean:.2e} vel {e_vel:.2e} p {e_p:.2e} "
          f"div plain {max(dp):.2e} mixed {max(dm):.2e}")
    assert e_mean < 1e-5 and e_vel < 1e-4 and e_p < 1e-3
    assert max(dm) < 1e-9


def test_mixed_refinement_sweep_lowers_divergence(pkg):
    _, _, d0, _ = run(pkg, pkg.MIXED, 64, 48, 64, 1.8, 2800.0, 1e-3, 3, nref_step=0)
    _, _, d1, _ = run(pkg, pkg.MIXED, 64, 48, 64, 1.8, 2800.0, 1e-3, 3, nref_step=1)
    print(f"mixed max div per step, 0 sweeps {d0}, 1 sweep {d1}")
    assert max(d1[1:]) < max(d0[1:])
    assert max(d1) < 1e-12
