"""Restart files of the TG field in the reference's format (WRT_INSTANT_VARS_TG,
processing/WRT_VARS_TG.f90:12-104; ReStart_INSTANT_VARS_TG, processing/RESTART_tg.f90:11-56):
round trip, and, where the compiled reference is present (oracle/_ref), byte equality with
the files the reference itself writes and reading each other's files."""
import os

import numpy as np
import pytest

from oracle import refrun


def field(golden):
    g = golden("tg_20x18x30.npz")
    L, M, N, STR2, REN, dt = (g["params"][i] for i in range(6))
    return int(L), int(M), int(N), float(STR2), float(REN), float(dt), g["q_step1"], g["pr_step1"]


def test_round_trip_and_walls(pkg, golden, tmp_path):
    L, M, N, STR2, REN, dt, q, p = field(golden)
    pkg.restart_write_tg(tmp_path, q, p, 7, 0.125, REN, dt)
    names = sorted(os.listdir(tmp_path))
    assert names == [f"DNS_perixz_INSTANT_T1.250000000E-01_{v}.D" for v in "PUVW"]
    size = os.path.getsize(tmp_path / names[0])
    assert size == 4 * 4 + 3 * 8 + 8 * L * M * N
    q2, p2, it, t = pkg.restart_read_tg(tmp_path, 0.125, L, M, N)
    assert it == 7 and t == 0.125
    np.testing.assert_array_equal(q2[:, :, 1:M + 1], q[:, :, 1:M + 1])
    np.testing.assert_array_equal(p2[:, 1:M + 1], p[:, 1:M + 1])
    assert np.all(q2[:, :, 0] == 0) and np.all(q2[:, :, M + 1] == 0) and np.all(q2[1, :, 1] == 0)
    np.testing.assert_array_equal(p2[:, 0], p2[:, 1])
    with pytest.raises(pkg.ChsError):
        pkg.restart_read_tg(tmp_path, 0.125, L, M + 2, N)
    with pytest.raises(pkg.ChsError):
        pkg.restart_read_tg(tmp_path, 0.5, L, M, N)


@pytest.mark.skipif(not refrun.available(), reason="compiled reference (oracle/_ref) absent")
def test_interchange_with_reference(pkg, golden):
    L, M, N, STR2, REN, dt, q, p = field(golden)
    h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt)
    try:
        h.write_fields("f.bin", q[0], q[1], q[2], p)
        rdir = h.path("1_instant_rawdata")
        h.run(["load f.bin", "rstwrite 0.25 11"])
        ref_names = sorted(n for n in os.listdir(rdir) if n.startswith("DNS_perixz_INSTANT_T2.5"))
        assert len(ref_names) == 4
        # the library writes the same bytes
        mine = h.path("mine")
        os.makedirs(mine)
        pkg.restart_write_tg(mine, q, p, 11, 0.25, REN, dt)
        for n in ref_names:
            with open(os.path.join(rdir, n), "rb") as a, open(os.path.join(mine, n), "rb") as b:
                assert a.read() == b.read(), n
        # the library reads the reference's files
        q2, p2, it, t = pkg.restart_read_tg(rdir, 0.25, L, M, N)
        assert it == 11 and t == 0.25
        np.testing.assert_array_equal(q2[:, :, 1:M + 1], q[:, :, 1:M + 1])
        # the reference reads the library's files (time 0.5), then the walls as FLOWSTART
        pkg.restart_write_tg(rdir, q2, p2, 12, 0.5, REN, dt)
        h.run(["rstread 0.5", "dump back"])
        u, v, w, pr = h.read_q("back")
        np.testing.assert_array_equal(np.stack([u, v, w]), q2)
        np.testing.assert_array_equal(pr, p2)
    finally:
        h.close()


@pytest.mark.gpu
def test_gpu_resume_from_restart_files(pkg, tmp_path):
    """Checkpoint from device state and resume: 2 RK3 steps on the GPU, the fields downloaded
    and written as the reference's restart files, read back into a fresh context (CONVH0
    zero, as after the reference's restart), 2 more steps; bit-identical to 4 uninterrupted
    steps (the first substep after a restart weights CONVH0 by TROH(1) = 0)."""
    L, M, N, REN, dt = 32, 24, 16, 2800.0, 2e-3
    print(f"resume {L}x{M}x{N}", flush=True)
    co = pkg.channel_setup(L, M, N, STR2=1.8)
    q0, p0 = pkg.poiseuille_perturbed(co, 5, 0.1)
    ref = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    ref.upload(q0, p0)
    ref.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    for _ in range(4):
        ref.rk3_step(dt, 1.0 / REN)
    qr, pr = ref.download()
    ref.close()
    a = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    a.upload(q0, p0)
    a.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    for _ in range(2):
        a.rk3_step(dt, 1.0 / REN)
    qa, pa = a.download()
    a.close()
    pkg.restart_write_tg(tmp_path, qa, pa, 2, 2 * dt, REN, dt)
    q2, p2, it, t = pkg.restart_read_tg(tmp_path, 2 * dt, L, M, N)
    assert it == 2 and t == 2 * dt
    b = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    b.upload(q2, p2)
    for _ in range(2):
        b.rk3_step(dt, 1.0 / REN)
    qb, pb = b.download()
    b.close()
    np.testing.assert_array_equal(qb, qr)
    np.testing.assert_array_equal(pb, pr)
