"""Restart files of the main domain in the reference's format (WRT_INSTANT_VARS_IO,
processing/WRT_VARS_IO.f90:16-227; ReStart_INSTANT_VARS_io, processing/RESTART_io.f90:16-131):
round trip, and, where the compiled reference is present (oracle/_ref), byte equality with the
nine files the reference itself writes (flow U, V, W, P and thermal T, D, H, M, K) and the
reference reading the library's files."""
import os

import numpy as np
import pytest

from oracle import oracle, refrun, setup

L, M, N, STR2, REN, DT, T0 = 16, 12, 8, 1.8, 2800.0, 1.0e-3, 665.0


def state(pkg):
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    t = thermo.table_setup(thermo.synthetic_table(), T0)
    dh, _ = oracle.eng_initial_field(mesh, t, 41)
    st = oracle.io_initial_field(mesh, 42)
    return thermo, mesh, t, dh, st


def test_round_trip(pkg, tmp_path):
    rng = np.random.default_rng(3)
    hs = (N, M + 2, L)
    q = rng.standard_normal((3,) + hs)
    pr = rng.standard_normal(hs)
    th = {k: 1.0 + 0.1 * rng.standard_normal(hs) for k in "TDHMK"}
    pkg.restart_write_io(tmp_path, q, pr, 5, 0.375, REN, DT, th)
    names = sorted(os.listdir(tmp_path))
    assert names == sorted(f"DNS_perixz_INSTANT_T3.750000000E-01_{v}.D" for v in "UVWPTDHMK")
    o = pkg.restart_read_io(tmp_path, 0.375, L, M, N, ini_field_type=0)
    assert o["ITERG"] == 5 and o["PhyTIME"] == 0.375
    np.testing.assert_array_equal(o["FLOW"][:, :, 1:M + 1][:, :, 1:], q[:, :, 2:M + 1])   # v row 1 is the wall
    np.testing.assert_array_equal(o["FLOW"][0][:, 1:M + 1], q[0][:, 1:M + 1])
    np.testing.assert_array_equal(o["PR"][:, 1:M + 1], pr[:, 1:M + 1])
    assert np.all(o["FLOW"][:, :, 0] == 0) and np.all(o["FLOW"][:, :, M + 1] == 0) and np.all(o["FLOW"][1, :, 1] == 0)
    np.testing.assert_array_equal(o["PR"][:, 0], o["PR"][:, 1])
    for k in "TDH":
        np.testing.assert_array_equal(o[k][:, 1:M + 1], th[k][:, 1:M + 1])
    np.testing.assert_array_equal(o["DH"][:, 1:M + 1], th["D"][:, 1:M + 1] * th["H"][:, 1:M + 1])
    # flow only, velocity form
    o1 = pkg.restart_read_io(tmp_path, 0.375, L, M, N, ini_field_type=1, thermal=False)
    np.testing.assert_array_equal(o1["FLOW"][0][:, 1:M + 1], q[0][:, 1:M + 1])
    with pytest.raises(pkg.ChsError):
        pkg.restart_read_io(tmp_path, 0.375, L, M, N, ini_field_type=1, thermal=True)
    with pytest.raises(pkg.ChsError):
        pkg.restart_read_io(tmp_path, 0.5, L, M, N)
    with pytest.raises(pkg.ChsError):
        pkg.restart_write_io(tmp_path, q, pr, 5, 0.5, REN, DT, {"T": th["T"], "D": None, "H": th["H"],
                                                                 "M": th["M"], "K": th["K"]})


@pytest.mark.skipif(not refrun.available(), reason="compiled reference (oracle/_ref) absent")
def test_interchange_with_reference(pkg):
    thermo, mesh, t, dh, st = state(pkg)
    h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=DT, iDomain=2, iThermoDynamics=1, T0=T0, Ti=T0,
                       iVisScheme=0)
    try:
        thermo.write_table_file(h.path("NIST_WATER_23.5MP.DAT"), thermo.synthetic_table())
        refrun.write_io(h, "io.bin", st["G"], st["G0"], st["PR"], st["D0"], st["M0"], st["EXPLT0"])
        h.write_eng("eng.bin", dh, st["G"])
        rdir = h.path("1_instant_rawdata")
        h.run(["thermoinitall", "loadio io.bin", "loadeng eng.bin", "ioprops 1", "rstwriteio 0.25 11",
               "dumpio d1", "dumpeng e1"])
        ref_names = sorted(n for n in os.listdir(rdir) if n.startswith("DNS_perixz_INSTANT_T2.5"))
        assert len(ref_names) == 9
        f, e = refrun.read_io(h, "d1"), h.read_eng("e1")
        # the library writes the same bytes from the same arrays
        mine = h.path("mine")
        os.makedirs(mine)
        pkg.restart_write_io(mine, f["Q"], f["PR"], 11, 0.25, REN, DT, {k: e[k] for k in "TDHMK"})
        for n in ref_names:
            with open(os.path.join(rdir, n), "rb") as a, open(os.path.join(mine, n), "rb") as b:
                assert a.read() == b.read(), n
        # the library reads the reference's files
        o = pkg.restart_read_io(rdir, 0.25, L, M, N)
        assert o["ITERG"] == 11 and o["PhyTIME"] == 0.25
        np.testing.assert_array_equal(o["FLOW"][0][:, 1:M + 1], f["Q"][0][:, 1:M + 1])
        # the reference reads the library's files (time 0.5): G_io takes the files' U, V, W
        # (iIniFieldType = 0), DH = DENSITY * ENTHALPY as the library computes it
        pkg.restart_write_io(rdir, f["Q"], f["PR"], 12, 0.5, REN, DT, {k: e[k] for k in "TDHMK"})
        h.run(["thermoinitall", "rstreadio 0.5", "dumpio d2", "dumpeng e2"])
        f2, e2 = refrun.read_io(h, "d2"), h.read_eng("e2")
        o2 = pkg.restart_read_io(rdir, 0.5, L, M, N)
        for c in range(3):
            j0 = 2 if c == 1 else 1
            np.testing.assert_array_equal(f2["G"][c][:, j0:M + 1], o2["FLOW"][c][:, j0:M + 1])
        np.testing.assert_array_equal(f2["PR"][:, 1:M + 1], o2["PR"][:, 1:M + 1])
        np.testing.assert_array_equal(e2["DH"][:, 1:M + 1], o2["DH"][:, 1:M + 1])
    finally:
        h.close()
