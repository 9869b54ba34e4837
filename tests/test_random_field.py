"""Random start of the turbulence-generator channel (random_FL_FLD_TG, INITIAL_FL_FLD_TG.f90:16-104)
and its NS = 0 projection (CALC_INITIALIZATION_tg, :107-140).

The sine field (iRandomType 3) is deterministic: it is compared with the reference's own routine
(run through oracle/ref_harness.f90) where the compiled reference is present.  The random types
draw from the compiler's RANDOM_NUMBER, which nothing else reproduces; for them the tests check
what the algorithm fixes whatever the generator: zero x-z means of v and w, unit bulk velocity,
row means of u equal to the laminar profile over the bulk velocity, the amplitude envelope
(VPERG, SVPERG near the walls), the wall row of v, and, for type 2, that the field depends on
I + K + J only (checked on the reference's field too)."""
import numpy as np
import pytest

from oracle import oracle, refrun, setup

L, M, N, STR2, VPER, SVPER = 16, 12, 8, 1.8, 0.6, 0.3


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def check_structure(q, mesh, vper, svper, m_over_d=1.0):
    """Invariants of random_FL_FLD_TG for any generator (rows 1..M; the halo rows wrap around the
    rank ring: row 0 = row M, row M+1 = row 1)."""
    dyfi, ycc = np.asarray(mesh.DYFI), np.asarray(mesh.YCC)
    assert np.array_equal(q[:, :, 0], q[:, :, M]) and np.array_equal(q[:, :, M + 1], q[:, :, 1])
    assert np.all(q[1, :, 1] == 0)
    u, v, w = (q[c][:, 1:M + 1] / m_over_d for c in range(3))
    assert np.max(np.abs(v.mean(axis=(0, 2)))) < 1e-15 and np.max(np.abs(w.mean(axis=(0, 2)))) < 1e-15
    bulk = np.sum(u / dyfi[None, :, None]) / (np.sum(1.0 / dyfi) * L * N)
    assert abs(bulk - 1.0) < 1e-14
    vini = 1.5 * (1.0 - ycc ** 2)
    um = np.sum(vini / dyfi) / np.sum(1.0 / dyfi)   # bulk of the laminar part alone
    np.testing.assert_allclose(u.mean(axis=(0, 2)), vini / um, rtol=0, atol=1e-14)
    # perturbation amplitude: |Q - <Q>_xz| <= 2 * VPERQ (mean removal of a [-a, a] draw)
    amp = np.where(1.0 - np.abs(ycc) < 0.25, svper, vper)
    for f in (v[:, 1:], w[:, 1:]):
        assert np.all(np.abs(f) <= 2 * amp[None, 1:, None] + 1e-15)


def shift_invariant(q):
    """Q(I+1, J, K) == Q(I, J, K+1) for every row: the field is a function of I + K + J."""
    a = q[:, :-1, 1:M + 1, 1:]
    b = q[:, 1:, 1:M + 1, :-1]
    return np.array_equal(a, b)


@pytest.mark.parametrize("rtype", [1, 2])
def test_random_types_structure(pkg, rtype):
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    q, pr = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=rtype, vper=VPER, svper=SVPER, seed=7)
    assert np.all(pr == 0)
    check_structure(q, mesh, VPER, SVPER)
    assert shift_invariant(q) == (rtype == 2)
    q2, _ = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=rtype, vper=VPER, svper=SVPER, seed=7)
    np.testing.assert_array_equal(q, q2)                       # a pure function of (seed, mesh)
    q3, _ = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=rtype, vper=VPER, svper=SVPER, seed=8)
    assert not np.array_equal(q, q3)
    q4, _ = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=rtype, vper=VPER, svper=SVPER, seed=7,
                                m_over_d=2.0)
    np.testing.assert_array_equal(q4, 2.0 * q)                 # Q * M_inlet / D_inlet


def test_bad_arguments_are_errors(pkg):
    for rtype in (0, 4):
        with pytest.raises(pkg.ChsError):
            pkg.random_field_tg(L, M, N, random_type=rtype)
    with pytest.raises(pkg.ChsError):
        pkg.random_field_tg(1, M, N)
    with pytest.raises(pkg.ChsError):
        pkg.random_field_tg(L, M, N, HX=0.0)


def test_sine_field_closed_form(pkg):
    """The sine field's x-z means vanish over the periodic box (HX = 4 pi, HZ = 2 pi), so after
    the mean removal u is sin(x) cos(y) cos(z) plus the laminar profile, over the bulk velocity."""
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    q, pr = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=3)
    check_structure(q, mesh, 1.0, 1.0)
    x = np.arange(L) * 4 * np.pi / L
    z = (np.arange(N) + 0.5) * 2 * np.pi / N
    ycc = np.asarray(mesh.YCC)
    vini = 1.5 * (1.0 - ycc ** 2)
    um = np.sum(vini / np.asarray(mesh.DYFI)) / np.sum(1.0 / np.asarray(mesh.DYFI))
    u = (np.sin(x)[None, None, :] * np.cos(ycc)[None, :, None] * np.cos(z)[:, None, None] + vini[None, :, None]) / um
    np.testing.assert_allclose(q[0][:, 1:M + 1], u, rtol=0, atol=1e-14)
    assert np.all(q[2] == 0)
    assert np.all(pr[:, 0] == 0) and np.all(pr[:, M + 1] == 0) and np.max(np.abs(pr)) <= 2.0 / 16.0


@pytest.mark.skipif(not refrun.available(), reason="compiled reference (oracle/_ref) absent")
def test_sine_field_and_projection_match_reference(pkg):
    dt, REN = 1.0e-3, 2800.0
    h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt, iRandomType=3)
    try:
        h.run(["randomtg", "dump f0"])
        f0 = h.read_q("f0")
        # the projection substep of that state; loaded, so every rank's halo rows are current
        # (the reference's own sequence leaves the pressure's rank-interface halo rows at 0,
        # IniField_FLOW_tg sets rows 1..N2DO only, which makes its result depend on the rank count)
        h.write_fields("f0.bin", *f0)
        h.run(["load f0.bin", "substep 0 0.001", "dump f1"])
        f1 = h.read_q("f1")
    finally:
        h.close()
    q, pr = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=3)
    # v, w, p follow the reference's operation order, halo rows included; u differs by the
    # rounding of the bulk velocity (a two-rank sum in the reference)
    np.testing.assert_array_equal(q[1], f0[1])
    np.testing.assert_array_equal(q[2], f0[2])
    np.testing.assert_array_equal(pr, f0[3])
    assert rel(q[0], f0[0]) < 1e-14
    # the projection of the library's field (oracle NS = 0 substep) against the reference's
    o = oracle.OracleTG(setup.Mesh(L, M, N, STR2=STR2), REN)
    o.set_fields(q, pr)
    o.substep(0, dt)
    for c in range(3):
        assert rel(o.q[c], f1[c]) < 1e-12, c
    assert rel(o.pr, f1[3]) < 1e-10, rel(o.pr, f1[3])


@pytest.mark.skipif(not refrun.available(), reason="compiled reference (oracle/_ref) absent")
def test_reference_random_field_structure():
    """The reference's own type-2 field has the properties the library's keeps."""
    h = refrun.Harness(L, M, N, STR2=STR2, iRandomType=2, VPER=VPER, SVPER=SVPER)
    try:
        h.run(["randomtg", "dump f0"])
        u, v, w, _ = h.read_q("f0")
    finally:
        h.close()
    q = np.stack([u, v, w])
    check_structure(q, setup.Mesh(L, M, N, STR2=STR2), VPER, SVPER)
    assert shift_invariant(q)


@pytest.mark.gpu
@pytest.mark.parametrize("rtype", [2, 3])
def test_gpu_projection_of_random_field(pkg, rtype):
    print(f"case random_type {rtype}", flush=True)
    dt, REN = 1.0e-3, 2800.0
    q, pr = pkg.random_field_tg(L, M, N, STR2=STR2, random_type=rtype, seed=5)
    ch = pkg.ChannelTG(pkg.channel_setup(L, M, N, STR2=STR2), pkg.PLAIN_FP64)
    ch.upload(q, pr)
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)
    gq, gp = ch.download()
    assert ch.max_div() < 1e-10
    ch.close()
    o = oracle.OracleTG(setup.Mesh(L, M, N, STR2=STR2), REN)
    o.set_fields(q, pr)
    o.substep(0, dt)
    for c in range(3):
        assert rel(gq[c], o.q[c]) < 1e-10, c
    assert rel(gp, o.pr) < 1e-10, rel(gp, o.pr)
