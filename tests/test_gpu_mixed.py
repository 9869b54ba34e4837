"""Mixed-precision form (FP32 stencils, CN and FFTs; FP64 state, reductions, pivots) against
the plain FP64 form on the same inputs.  Bars: mean-velocity (x-z mean u) error < 1e-5,
velocity max-norm error < 1e-4 relative, max divergence < 1e-9 after each substep."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("L,M,N,STR2,REN,dt", [(64, 64, 64, 1.8, 2800.0, 1e-3),
                                              (128, 96, 128, 2.0, 6875.0, 5e-4)])
def test_mixed_vs_plain(pkg, L, M, N, STR2, REN, dt):
    qp, pp, dp, co = run(pkg, pkg.PLAIN_FP64, L, M, N, STR2, REN, dt, 5)
    qm, pm, dm, _ = run(pkg, pkg.MIXED, L, M, N, STR2, REN, dt, 5)
    umean_p = qp[0][:, 1:M + 1].mean(axis=(0, 2))
    umean_m = qm[0][:, 1:M + 1].mean(axis=(0, 2))
    e_mean = np.max(np.abs(umean_m - umean_p)) / np.max(np.abs(umean_p))
    e_vel = max(np.max(np.abs(qm[c] - qp[c])) / np.max(np.abs(qp[c])) for c in range(3))
    e_p = np.max(np.abs(pm - pp)) / np.max(np.abs(pp))
    print(f"mixed vs plain {L}x{M}x{N}: mean-u {e_mean:.2e} vel {e_vel:.2e} p {e_p:.2e} "
          f"div plain {max(dp):.2e} mixed {max(dm):.2e}")
    assert e_mean < 1e-5
    assert e_vel < 1e-4
    assert max(dm) < 1e-9


def test_mixed_refinement_sweep_lowers_divergence(pkg):
    _, _, d0, _ = run(pkg, pkg.MIXED, 64, 48, 64, 1.8, 2800.0, 1e-3, 3, nref_step=0)
    _, _, d1, _ = run(pkg, pkg.MIXED, 64, 48, 64, 1.8, 2800.0, 1e-3, 3, nref_step=1)
    print(f"mixed max div per step, 0 sweeps {d0}, 1 sweep {d1}")
    assert max(d1[1:]) < max(d0[1:])
    assert max(d1) < 1e-12
