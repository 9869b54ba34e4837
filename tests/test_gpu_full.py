"""The workloads at their full sizes, checked through properties that do not depend on a CPU
solution of the same size: divergence after projection, constant mass flux, wall
conditions, the Poisson residual and the mixed form's distance from the FP64 form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bulk(q, dyfi, M):
    return float(np.sum(q[0][:, 1:M + 1] / dyfi[None, :, None]))


def test_channel_256x192x256_plain_and_mixed(pkg):
    L, M, N, STR2, REN, dt = 256, 192, 256, 2.0, 6875.0, 5e-4
    print(f"channel {L}x{M}x{N}", flush=True)
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    dyfi = np.asarray(co["dyfi"])
    q0, p0 = pkg.poiseuille_perturbed(co, seed=395, amp=0.05)
    out = {}
    for mode in (pkg.PLAIN_FP64, pkg.MIXED):
        ch = pkg.ChannelTG(co, mode)
        if mode == pkg.MIXED:
            ch.set_refinement(0, 2)
        ch.upload(q0, p0)
        ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
        assert ch.max_div() < (1e-10 if mode == pkg.PLAIN_FP64 else 1e-9)
        qa, _ = ch.download()
        b0 = bulk(qa, dyfi, M)
        divs = []
        for _ in range(3):
            ch.rk3_step(dt, 1.0 / REN)
            divs.append(ch.max_div())
        q, p = ch.download()
        assert max(divs) < (1e-10 if mode == pkg.PLAIN_FP64 else 1e-9), divs
        assert abs(bulk(q, dyfi, M) - b0) / abs(b0) < 1e-9
        assert np.all(q[:, :, 0] == 0) and np.all(q[:, :, M + 1] == 0) and np.all(q[1, :, 1] == 0)
        np.testing.assert_array_equal(p[:, 0], p[:, 1])
        out[mode] = q
        ch.close()
    qp, qm = out[pkg.PLAIN_FP64], out[pkg.MIXED]
    um_p = qp[0][:, 1:M + 1].mean(axis=(0, 2))
    um_m = qm[0][:, 1:M + 1].mean(axis=(0, 2))
    e_mean = np.max(np.abs(um_m - um_p)) / np.max(np.abs(um_p))
    e_vel = max(np.max(np.abs(qm[c] - qp[c])) / np.max(np.abs(qp[c])) for c in range(3))
    print(f"mixed vs plain at full size after 3 steps: mean-u {e_mean:.2e}, velocity {e_vel:.2e}", flush=True)
    assert e_mean < 1e-5 and e_vel < 1e-4


def test_poisson_512x384x512_residual(pkg):
    L, M, N = 512, 384, 512
    print(f"poisson {L}x{M}x{N}", flush=True)
    co = pkg.channel_setup(L, M, N, 4 * np.pi, 2 * np.pi, 2.2)
    rng = np.random.Generator(np.random.Philox(key=3))
    r = rng.standard_normal((N, M, L))
    w = (1.0 / np.asarray(co["dyfi"]))[None, :, None]
    r -= np.sum(r * w) / (np.sum(w) * L * N)
    rmax = float(np.max(np.abs(r)))
    d_r = pkg.DeviceArray.from_host(r)
    del r
    d_phi = pkg.DeviceArray(d_r.shape)
    for mode, nref, tol in ((pkg.PLAIN_FP64, 0, 1e-12), (pkg.MIXED, 2, 1e-10)):
        ch = pkg.ChannelTG(co, mode)
        res = ch.poisson_device(d_r.ptr, d_phi.ptr, nref, True)
        print(f"mode {mode} sweeps {nref}: max residual / max rhs {res / rmax:.2e}", flush=True)
        assert res / rmax < tol
        ch.close()
    d_phi.free()
    d_r.free()


def test_heated_256x128x256_conserves_mass_flux(pkg):
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    init = import_module(pkg.__name__ + ".init")
    L, M, N, STR2, REN, dt, T0 = 256, 128, 256, 1.8, 2800.0, 1e-3, 665.0
    print(f"heated {L}x{M}x{N}", flush=True)
    t = thermo.table_setup(thermo.synthetic_table(), T0)
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    dyfi = np.asarray(co["dyfi"])
    dh, g, pr = init.heated_channel_initial(co, t, seed=180)
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    eng = thermo.EnergyTG(co, t, (1, 1), cthecd, t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1],
                          (0.115, -0.115))
    eng.upload(dh, g)
    f0 = eng.download()
    io = pkg.ChannelIO(co, REN, flow_driven=1)
    io.upload(G=g, PR=pr, D0=f0["D"], M0=f0["M"])
    b0 = None
    for step in range(3):
        for ns in (1, 2, 3):
            pkg.heated_substep(eng, io, ns, dt)
        f = io.download()
        b = bulk(f["G"], dyfi, M)
        b0 = b if b0 is None else b0
        assert abs(b - b0) / abs(b0) < 1e-9, (step, b, b0)
    fe = eng.download()
    assert np.all(f["G"][:, :, 0] == 0) and np.all(f["G"][1, :, 1] == 0)
    T = fe["T"][:, 1:M + 1]
    assert 0.95 < T.min() and T.max() < 1.05
    eng.close()
    io.close()
