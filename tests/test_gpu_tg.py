"""GPU parity of the plain FP64 substep against the oracle (and the reference's golden
vectors).  Tolerances: velocity and pressure 1e-10 relative (max-norm) per substep,
divergence after projection < 1e-10.  Each test prints the largest errors it saw."""
import numpy as np
import pytest

from oracle import oracle, setup

pytestmark = pytest.mark.gpu

MESHES = [  # L, M, N, STR2, REN, dt
    (16, 16, 16, 1.8, 2800.0, 1.0e-3),
    (20, 18, 30, 2.2, 6875.0, 5.0e-4),
    (64, 64, 64, 1.8, 2800.0, 1.0e-3),
    (48, 36, 40, 2.0, 4000.0, 8.0e-4),
    (256, 24, 256, 2.0, 6875.0, 5.0e-4),
]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(pkg, L, M, N, STR2, REN, seed=3):
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    q, pr = oracle.initial_field(mesh, seed)
    ch = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=STR2), pkg.PLAIN_FP64)
    ch.upload(q, pr)
    o = oracle.OracleTG(mesh, REN)
    o.set_fields(q, pr)
    return mesh, ch, o


@pytest.mark.parametrize("L,M,N,STR2,REN,dt", MESHES)
def test_projection_and_steps_match_oracle(pkg, L, M, N, STR2, REN, dt):
    print(f"case {L}x{M}x{N}", flush=True)
    mesh, ch, o = make(pkg, L, M, N, STR2, REN)
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    o.substep(0, dt)
    gq, gp = ch.download()
    eq = max(rel(gq[c], o.q[c]) for c in range(3))
    ep = rel(gp, o.pr)
    assert eq < 1e-10, ("ns0", eq)
    assert ep < 1e-10, ("ns0", ep)
    assert ch.max_div() < 1e-10
    for step in range(3):
        for ns in (1, 2, 3):
            ch.SOLVERRK3_MOM_tg(ns, dt, cv)
            o.substep(ns, dt)
            gq, gp, gh = ch.download(with_history=True)
            e = max(rel(gq[c], o.q[c]) for c in range(3))
            e_p = rel(gp, o.pr)
            eq, ep = max(eq, e), max(ep, e_p)
            assert e < 1e-10, (step, ns, e)
            assert e_p < 1e-10, (step, ns, e_p)
            assert rel(gh[0], o.convh0[0]) < 1e-10
    div = ch.max_div()
    print(f"  {L}x{M}x{N}: max rel error velocity {eq:.2e}, pressure {ep:.2e}; max div {div:.2e}")
    assert div < 1e-10
    ch.close()


@pytest.mark.parametrize("name", ["tg_16x16x16.npz", "tg_20x18x30.npz"])
def test_matches_reference_golden(pkg, golden, name):
    g = golden(name)
    L, M, N, STR2, REN, dt = (g["params"][i] for i in range(6))
    L, M, N = int(L), int(M), int(N)
    ch = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=STR2), pkg.PLAIN_FP64)
    ch.upload(g["q_init"], g["pr_init"])
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    gq, gp = ch.download()
    e0 = [max(rel(gq[c], g["q_ns0"][c]) for c in range(3)), rel(gp, g["pr_ns0"])]
    ch.rk3_step(dt, 1.0 / REN)
    gq, gp = ch.download()
    e1 = [max(rel(gq[c], g["q_step1"][c]) for c in range(3)), rel(gp, g["pr_step1"])]
    phi = ch.poisson(g["poisson_rhs"])
    e2 = rel(phi, g["poisson_phi"])
    print(f"  {name}: projection velocity {e0[0]:.2e} pressure {e0[1]:.2e}; RK3 step velocity "
          f"{e1[0]:.2e} pressure {e1[1]:.2e}; Poisson {e2:.2e}")
    assert e0[0] < 1e-10 and e1[0] < 1e-10
    assert e0[1] < 1e-10 and e1[1] < 1e-10
    assert e2 < 1e-10
    ch.close()


# line lengths 64 ... 1024 run the register FFTs (chs_fftr.h), the others the LDS Stockham
# transforms; fft_reg=0 forces the latter on planned lengths too
# transforms; fft_reg=0 forces the latter on planned lengths too.  ytc: rows per chunk of the
# chunked y-Thomas (-1 automatic, 0 the one-thread-per-line kernel; 4 and 8 give many chunks)
POISSON_MESHES = [(16, 16, 16, 1, -1), (16, 16, 16, 1, 0), (30, 20, 24, 1, 4), (128, 96, 128, 1, -1),
                  (128, 96, 128, 0, 8), (128, 96, 128, 1, 0), (64, 20, 96, 1, 4), (256, 12, 256, 1, -1),
                  (512, 8, 512, 1, -1), (1024, 6, 64, 1, -1), (96, 10, 1024, 1, 4), (64, 256, 64, 1, -1),
                  (64, 300, 32, 1, -1), (32, 61, 48, 1, 4)]


@pytest.mark.parametrize("L,M,N,reg,ytc", POISSON_MESHES)
def test_poisson_matches_oracle(pkg, L, M, N, reg, ytc):
    print(f"poisson {L}x{M}x{N} fft_reg={reg} ytdma_c={ytc}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.2)
    rng = np.random.default_rng(5)
    r = rng.standard_normal((N, M, L))
    w = (1.0 / mesh.DYFI)[None, :, None]
    r -= np.sum(r * w) / (np.sum(w) * L * N)
    o = oracle.OracleTG(mesh, 2800.0)
    ref = o.poisson(r)
    for mode, tol in ((pkg.PLAIN_FP64, 1e-10), (pkg.MIXED, 1e-5)):
        ch = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=2.2), mode)
        ch.set_tuning(fft_reg=reg, ytdma_c=ytc)
        phi = ch.poisson(r)
        e = rel(phi, ref)
        print(f"  {'plain' if mode == pkg.PLAIN_FP64 else 'mixed'}: rel error {e:.2e}")
        assert e < tol, (mode, e)
        ch.close()


def test_mass_flux_and_walls(pkg):
    L, M, N, REN, dt = 32, 32, 32, 2800.0, 1e-3
    mesh, ch, _ = make(pkg, L, M, N, 1.8, REN)
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    q0, _ = ch.download()
    bulk0 = np.sum(q0[0][:, 1:M + 1] / mesh.DYFI[None, :, None])
    for _ in range(3):
        ch.rk3_step(dt, 1.0 / REN)
    q, p = ch.download()
    bulk = np.sum(q[0][:, 1:M + 1] / mesh.DYFI[None, :, None])
    # constant mass flux (iFlowDriven = 1): the explicit increment has zero volume mean
    # (RHS_CvLpGpS_tg.f90:165-201); the wall-bounded CN y-solve then moves the bulk by O(beta)
    assert abs(bulk - bulk0) / abs(bulk0) < 1e-9
    assert np.all(q[:, :, 0, :] == 0) and np.all(q[:, :, M + 1, :] == 0)
    assert np.all(q[1, :, 1, :] == 0)                    # v on the bottom wall
    np.testing.assert_array_equal(p[:, 0, :], p[:, 1, :])
    np.testing.assert_array_equal(p[:, M + 1, :], p[:, M, :])
    ch.close()


@pytest.mark.parametrize("L,M,N", [(64, 48, 32), (30, 20, 24)])
def test_poisson_device_refinement(pkg, L, M, N):
    """Device Poisson: FP64 form matches the oracle; the mixed form (FP32 transforms) reaches
    FP64 accuracy with residual-driven FP64 refinement sweeps."""
    mesh = setup.Mesh(L, M, N, STR2=2.2)
    rng = np.random.default_rng(8)
    r = rng.standard_normal((N, M, L))
    w = (1.0 / mesh.DYFI)[None, :, None]
    r -= np.sum(r * w) / (np.sum(w) * L * N)
    ref = oracle.OracleTG(mesh, 2800.0).poisson(r)
    d_r = pkg.DeviceArray.from_host(r)
    d_phi = pkg.DeviceArray(r.shape)
    co = pkg.channel_setup(L, M, N, STR2=2.2)
    plain = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    res = plain.poisson_device(d_r.ptr, d_phi.ptr, 0)
    assert rel(d_phi.to_host(), ref) < 1e-10
    assert res < 1e-9 * np.max(np.abs(r))
    mixed = pkg.ChannelTG(co, pkg.MIXED)
    errs, gauge = [], []
    for nref in (0, 1, 2):
        mixed.poisson_device(d_r.ptr, d_phi.ptr, nref)
        phi = d_phi.to_host()
        errs.append(rel(phi, ref))
        gauge.append(abs(phi[:, 0, :].mean()) / np.max(np.abs(ref)))
    print(f"mixed Poisson {L}x{M}x{N} error vs FP64 after 0/1/2 sweeps: {errs}; "
          f"J=1 plane mean: {gauge}")
    assert errs[0] < 1e-3
    # the sweeps restore the reference's gauge (zero mean over J = 1) in FP64
    assert gauge[1] < 1e-14 and gauge[2] < 1e-14
    assert errs[1] < 1e-8 and errs[2] < 1e-10
    plain.close()
    mixed.close()
    d_r.free()
    d_phi.free()


SHAPES = [
    dict(conv_by=16, conv_kc=3, cnx_nlx=3, cnz_kz=5, cny_ky=7, fftx_tk=2, fftz_nl=1, nxcd=3, cn_xz=1,
         cnxz_by=2),
    dict(conv_by=1, conv_kc=1, cnx_nlx=1, cnz_kz=40, cny_ky=36, fftx_tk=8, fftz_nl=5, cn_xz=1),
    dict(conv_by=4, conv_bx=4, conv_kc=7, nxcd=8, cnxz_by=4, cnz_kz=9, cn_xz=1),
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_launch_shapes_keep_results(pkg, shape):
    """chs_set_tuning: other tilings of every kernel (ragged chunks included) give the
    oracle's result to the same tolerance as the default shapes."""
    L, M, N, STR2, REN, dt = 48, 36, 40, 2.0, 4000.0, 8.0e-4
    print(f"shape {SHAPES[shape]}", flush=True)
    mesh, ch, o = make(pkg, L, M, N, STR2, REN)
    ch.set_tuning(**SHAPES[shape])
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    o.substep(0, dt)
    for _ in range(2):
        ch.rk3_step(dt, cv)
        for ns in (1, 2, 3):
            o.substep(ns, dt)
    gq, gp = ch.download()
    e, ep = max(rel(gq[c], o.q[c]) for c in range(3)), rel(gp, o.pr)
    print(f"  velocity {e:.2e} pressure {ep:.2e}")
    assert e < 1e-10 and ep < 1e-10
    ch.close()


@pytest.mark.parametrize("L,M,N,fx", [(192, 24, 32, 1), (192, 24, 32, 2), (128, 32, 48, 2), (128, 32, 48, 0),
                                       (64, 48, 32, 2)])
def test_fused_x_solve_matches_oracle(pkg, L, M, N, fx):
    """The x-solve inside k_conv_rhs (conv_fx rows per block; 0: the separate x kernel) on
    whole-line tiles of one to three wavefronts (a 3-wave tile folds its mass-flux partials
    first), with the mass-flux correction applied after the z-solve."""
    STR2, REN, dt = 2.0, 4000.0, 8.0e-4
    print(f"fused x {L}x{M}x{N} conv_fx {fx}", flush=True)
    mesh, ch, o = make(pkg, L, M, N, STR2, REN)
    ch.set_tuning(conv_fx=fx)
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    o.substep(0, dt)
    for _ in range(2):
        ch.rk3_step(dt, cv)
        for ns in (1, 2, 3):
            o.substep(ns, dt)
    gq, gp = ch.download()
    e, ep = max(rel(gq[c], o.q[c]) for c in range(3)), rel(gp, o.pr)
    print(f"  velocity {e:.2e} pressure {ep:.2e}")
    assert e < 1e-10 and ep < 1e-10
    assert ch.max_div() < 1e-10
    ch.close()


@pytest.mark.parametrize("L,M,N,REN,dt,ky", [(64, 48, 32, 4000.0, 8.0e-4, 0), (64, 48, 32, 4000.0, 8.0e-4, 16),
                                              (128, 40, 48, 6875.0, 5.0e-4, 7), (64, 24, 16, 2800.0, 2.0e-3, 20),
                                              (192, 36, 64, 2800.0, 1.0e-3, 11), (96, 36, 40, 4000.0, 8.0e-4, 16),
                                              (20, 18, 30, 6875.0, 5.0e-4, 0)])
def test_fused_zy_solve_matches_two_kernels_and_oracle(pkg, L, M, N, REN, dt, ky):
    """k_cn_zy_march (the z and y Crank-Nicolson inverses in one pass, chunks of ky rows, 0 =
    whole lines, ragged last chunks; ragged x and z tiles where L is not a multiple of 64 or N
    of 16) gives the separate z and y kernels' result to rounding and
    the oracle's to 1e-10, in FP64 and in the mixed form (against the mixed two-kernel path)."""
    STR2 = 2.0
    print(f"fused zy {L}x{M}x{N} ky {ky}", flush=True)
    mesh, ch, o = make(pkg, L, M, N, STR2, REN)
    two = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=STR2), pkg.PLAIN_FP64)
    q0, p0 = oracle.initial_field(mesh, 3)
    two.upload(q0, p0)
    ch.set_tuning(cnzy=2, cnzy_ky=ky)
    two.set_tuning(cnzy=0)
    cv = 1.0 / REN
    for c in (ch, two):
        c.SOLVERRK3_MOM_tg(0, dt, cv)
        for _ in range(2):
            c.rk3_step(dt, cv)
    o.substep(0, dt)
    for _ in range(2):
        for ns in (1, 2, 3):
            o.substep(ns, dt)
    gq, gp = ch.download()
    tq, tp = two.download()
    dq, dp = max(rel(gq[c], tq[c]) for c in range(3)), rel(gp, tp)
    e, ep = max(rel(gq[c], o.q[c]) for c in range(3)), rel(gp, o.pr)
    print(f"  fused vs two kernels velocity {dq:.2e} pressure {dp:.2e}; vs oracle velocity {e:.2e} "
          f"pressure {ep:.2e}")
    # rounding only (contractions differ between the kernels); the pressure carries the
    # Poisson solve's amplification of it
    assert dq < 1e-13 and dp < 1e-10
    assert e < 1e-10 and ep < 1e-10
    assert ch.max_div() < 1e-10
    ch.close()
    two.close()
    mixed = [pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=STR2), pkg.MIXED) for _ in range(2)]
    for m, on in zip(mixed, (1, 0)):
        m.upload(q0, p0)
        m.set_tuning(cnzy=2 * on, cnzy_ky=ky)
        m.SOLVERRK3_MOM_tg(0, dt, cv)
        for _ in range(2):
            m.rk3_step(dt, cv)
    (aq, ap), (bq, bp) = mixed[0].download(), mixed[1].download()
    # the mixed form's error against FP64, fused and with two kernels: FP32 rounding of the
    # same size
    ea, eb = max(rel(aq[c], gq[c]) for c in range(3)), max(rel(bq[c], gq[c]) for c in range(3))
    pa, pb = rel(ap, gp), rel(bp, gp)
    print(f"  mixed vs FP64: fused velocity {ea:.2e} pressure {pa:.2e}; two kernels velocity {eb:.2e} "
          f"pressure {pb:.2e}")
    assert ea < 1e-5 and ea < 3 * eb + 1e-7
    assert pa < 3 * pb + 1e-6
    for m in mixed:
        m.close()


@pytest.mark.parametrize("L,M,N,mode", [(128, 40, 48, "plain"), (256, 24, 32, "plain"), (128, 20, 24, "mixed")])
def test_update_two_points_per_lane_bit_identical(pkg, L, M, N, mode):
    """k_update2 (two points per lane, neighbours across lanes and the stretch's edges loaded)
    gives k_update's fields and CFL bit for bit, with and without the CFL reduction."""
    print(f"update2 {L}x{M}x{N} {mode}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    q, pr = oracle.initial_field(mesh, 5)
    m = pkg.PLAIN_FP64 if mode == "plain" else pkg.MIXED
    out = []
    for w2 in (1, 0):
        ch = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=2.0), m)
        ch.set_tuning(upd_w2=w2)
        ch.upload(q, pr)
        ch.SOLVERRK3_MOM_tg(0, 5e-4, 1 / 4000.0)
        for _ in range(2):
            ch.rk3_step(5e-4, 1 / 4000.0)
        out.append((*ch.download(), ch.step_state()["cflmm"]))
        ch.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] and np.isfinite(out[0][2])


def test_tuning_rejects_bad_shapes(pkg):
    ch = pkg.ChannelTG(pkg.channel_setup(48, 36, 40, STR2=2.0), pkg.PLAIN_FP64)
    for bad in (dict(conv_by=3), dict(conv_bx=2), dict(conv_kc=0), dict(fft_reg=2), dict(nxcd=0), dict(cn_xz=2), dict(cnxz_by=3), dict(conv_kc=41), dict(fftx_tk=3),
                dict(fftx_tk=6), dict(fftz_nl=100), dict(cny_ky=37), dict(cnx_nlx=10**6), dict(conv_fx=3),
                dict(cnzy=3), dict(cnzy_ky=37), dict(cnzy_nxcd=0), dict(upd_w2=2),
                dict(ytdma_c=12), dict(ytdma_c=-2), dict(no_such_key=1)):
        with pytest.raises(pkg.ChsError):
            ch.set_tuning(**bad)
    ch.close()


def test_autotune_restores_state(pkg):
    """The autotuner leaves the fields as they were and picks valid shapes."""
    import importlib
    tune = importlib.import_module(pkg.__name__ + ".tune")
    L, M, N, STR2, REN, dt = 48, 36, 40, 2.0, 4000.0, 8.0e-4
    mesh, ch, o = make(pkg, L, M, N, STR2, REN)
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    before = ch.download(with_history=True)
    res = tune.autotune(ch, dt, cv, steps=1)
    after = ch.download(with_history=True)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    cands = tune.candidates(L, M, N)
    for k, v in res["best"].items():
        assert v in cands[k]
    o.substep(0, dt)
    ch.rk3_step(dt, cv)
    for ns in (1, 2, 3):
        o.substep(ns, dt)
    gq, _ = ch.download()
    for c in range(3):
        assert rel(gq[c], o.q[c]) < 1e-10
    ch.close()


def test_config1_100_steps_match_oracle(pkg):
    """The CPU-runnable configuration as configured: 64^3, Re_tau 180 (REN 2800), tanh beta 1.8,
    laminar profile + 10 % seeded perturbations projected by NS = 0, dt 1e-3, 100 RK3 steps,
    plain FP64, against the oracle after every 10 steps (velocity, pressure, x-z mean u,
    maximum divergence)."""
    L, M, N, STR2, REN, dt = 64, 64, 64, 1.8, 2800.0, 1.0e-3
    print(f"config 1 {L}x{M}x{N}, 100 steps", flush=True)
    mesh, ch, o = make(pkg, L, M, N, STR2, REN, seed=180)
    cv = 1.0 / REN
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    o.substep(0, dt)
    worst = dict(u=0.0, p=0.0, mean=0.0)
    for step in range(1, 101):
        ch.rk3_step(dt, cv)
        for ns in (1, 2, 3):
            o.substep(ns, dt)
        if step % 10 == 0:
            gq, gp = ch.download()
            e = max(rel(gq[c], o.q[c]) for c in range(3))
            e_p = rel(gp, o.pr)
            um_g = gq[0][:, 1:M + 1].mean(axis=(0, 2))
            um_o = o.q[0][:, 1:M + 1].mean(axis=(0, 2))
            e_m = float(np.max(np.abs(um_g - um_o)) / np.max(np.abs(um_o)))
            worst = dict(u=max(worst["u"], e), p=max(worst["p"], e_p), mean=max(worst["mean"], e_m))
            print(f"  step {step}: velocity {e:.2e}, pressure {e_p:.2e}, mean u {e_m:.2e}", flush=True)
    div = ch.max_div()
    print(f"  worst {worst}; max div {div:.2e}", flush=True)
    # measured on MI355X: velocity 7.6e-15, pressure 3.4e-12, mean u 4.4e-16, div 1.6e-15
    assert worst["u"] < 1e-13 and worst["mean"] < 1e-14, worst
    assert worst["p"] < 1e-10, worst
    assert div < 1e-12
    ch.close()
