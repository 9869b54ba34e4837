"""Main-domain momentum substep with the Crank-Nicolson viscous scheme (SOLVERRK3_MOM_io with
iVisScheme = 1: VISCOUS_PAR_EXPLT_X/Y/Z_io, RHS_MOM_EXPLICIT_io, MOMFA_io, PRCALC_io's viscous
term): the oracle restatement and the GPU substep against the reference's own output
(tests/golden/io_cn.npz, written by the reference's routines through oracle/ref_harness.f90 on
two ranks, from the io.npz states).

The reference's y-solve (MOMFA3_Y_io) depends on the rank count in two ways, so the fixtures
are read where a two-rank run computes the one-rank formula the oracle and the GPU follow:
  * after TRASP23_Y2Z a rank holds z-planes KDSWT(MYID).., but the wall viscosities of the u, w
    lines are read from the y-slab arrays Viscousity / VISN with the slab-local z index
    (MOMFA_io.f90:342-352), and AMVR1 / APVRN, the wall coefficients, are set by TDMA_COEF on
    rank 0 only and never broadcast (TDMA_COEF.f90:81, :87, :338-358): rank 1 solves its
    z-planes with no wall term.  Hence the ``cn`` fixture (noisy current viscosity) is compared
    on rank 0's z-planes k < N/2 only, where all of it is the one-rank formula, and the test
    shows that rank 1's planes differ.
  * the ``s1..s3`` fixtures use a viscosity whose ghost rows are the negatives of the wall
    cells: the wall-row means are zero, both ranks' wall terms vanish alike, and every z-plane
    of three whole substeps compares.
DENSITY_IMPLICIT_add's DRHOI_STG is identically zero (it subtracts the reciprocal of D_STG's own
expression of the same DENSITY0), checked on the reference's dump; the GPU and the oracle omit
the terms it multiplies.  Tolerances (max-norm relative per field) as the explicit scheme's:
1e-12 for G, Q and the momentum terms, 1e-10 for pressure and its correction."""
import numpy as np
import pytest

from oracle import oracle, setup

TOL, PTOL = 1e-12, 1e-10


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def cn_viscosity(st, seed, walls_zero):
    """The current viscosity the fixtures were made with (tests/golden/make_golden.py)."""
    rng = np.random.Generator(np.random.Philox(key=seed + 1000))
    m1 = st["M0"] * (1.0 + 0.1 * rng.standard_normal(st["M0"].shape))
    if walls_zero:
        M = m1.shape[1] - 2
        m1[:, 0] = -m1[:, 1]
        m1[:, M + 1] = -m1[:, M]
    return m1


def cases(golden):
    g, gi = golden("io_cn.npz"), golden("io.npz")
    out = []
    for tag in sorted({k.split("_")[0] for k in g}):
        L, M, N, STR2, REN, dt, grav, fd, seed, T0 = g[tag + "_params"]
        f_a, volm, cvisc = gi[tag + "_par"]
        init = {k: gi[f"{tag}_init_{k}"] for k in ("G", "G0", "PR", "D0", "M0", "EXPLT0")}
        out.append(dict(tag=tag, L=int(L), M=int(M), N=int(N), STR2=float(STR2), REN=float(REN), dt=float(dt),
                        grav=int(grav), fd=int(fd), f_a=float(f_a), seed=int(seed), init=init,
                        m1=cn_viscosity(init, int(seed), False), m1w=cn_viscosity(init, int(seed), True), g=g))
    return out


def make_oracle(c, vis_scheme=1):
    mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
    return oracle.OracleIO(mesh, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"],
                           cf_given=8.18e-3, vis_scheme=vis_scheme)


def check_state(f, g, s, M, first=False):
    for c in range(3):
        assert rel(f["G"][c], g[f"{s}_G"][c]) < TOL, (s, c, rel(f["G"][c], g[f"{s}_G"][c]))
        j0 = 1 if c == 1 else 0        # v's row J = 1 is not an unknown
        assert rel(f["EXPLT0"][c][:, j0:], g[f"{s}_EXPLT0"][c][:, j0:]) < TOL, (s, c)
    assert rel(f["PR"], g[f"{s}_PR"]) < PTOL, (s, rel(f["PR"], g[f"{s}_PR"]))
    if first:
        assert rel(f["RHSP"], g[f"{s}_RHSP"]) < TOL
        assert rel(f["DPH"], g[f"{s}_DPH"]) < PTOL


def test_fixture_inputs(golden):
    """The viscosities are reproducible from the seeds, DRHOI_STG of the reference is zero
    everywhere, and the substep inputs have zero wall-row means."""
    for c in cases(golden):
        assert np.all(c["g"][c["tag"] + "_cn_DRHOI"] == 0.0)
        M = c["M"]
        assert np.all(c["m1w"][:, 0] + c["m1w"][:, 1] == 0.0)
        assert np.all(c["m1w"][:, M] + c["m1w"][:, M + 1] == 0.0)
        assert np.all(c["m1"] != c["init"]["M0"])


def test_oracle_cn_momentum_matches_reference(golden):
    for c in cases(golden):
        g, tag, M, N = c["g"], c["tag"], c["M"], c["N"]
        o = make_oracle(c)
        o.set_state(dict(c["init"], M1=c["m1"]))
        o.props(1)
        o.momentum_cn(1, c["dt"])
        # RHS_io as each MOMFA_io receives it and the new EXPLT0: every z-plane
        for comp in range(3):
            j0 = 1 if comp == 1 else 0
            assert rel(o.f["RKEEP"][comp][:, j0:], g[f"{tag}_cn_RHS"][comp][:, j0:]) < TOL, comp
            assert rel(o.f["EXPLT0"][comp][:, j0:], g[f"{tag}_cn_EXPLT0"][comp][:, j0:]) < TOL, comp
        # G after the three line solves: rank 0's z-planes
        h = N // 2
        for comp in range(3):
            gr = g[f"{tag}_cn_G"][comp]
            assert rel(o.f["G"][comp][:h, 1:M + 1], gr[:h, 1:M + 1]) < TOL, comp
        # ... and rank 1's planes of u and w carry the reference's rank-1 wall terms (none)
        d = np.abs(o.f["G"][0][h:] - g[f"{tag}_cn_G"][0][h:]).max()
        assert d > 1e-12 * np.abs(g[f"{tag}_cn_G"][0]).max()


def test_oracle_cn_substeps_match_reference(golden):
    for c in cases(golden):
        g, tag, M = c["g"], c["tag"], c["M"]
        o = make_oracle(c)
        o.set_state(dict(c["init"], M1=c["m1w"]))
        for ns in (1, 2, 3):
            o.props(ns)
            o.substep(ns, c["dt"])
            check_state(o.f, g, f"{tag}_s{ns}", M, first=ns == 1)


def test_oracle_cn_planted_mistakes_fail(golden):
    """The fixtures tell the schemes apart and pin that the Crank-Nicolson terms read the
    current viscosity: the explicit scheme, or Viscousity0 in place of Viscousity, moves G."""
    c = cases(golden)[0]
    g, tag = c["g"], c["tag"]
    for vis, st in ((0, dict(c["init"], M1=c["m1w"])), (1, dict(c["init"]))):
        o = make_oracle(c, vis_scheme=vis)
        o.set_state(st)
        o.props(1)
        o.substep(1, c["dt"])
        assert rel(o.f["G"][0], g[f"{tag}_s1_G"][0]) > 1e-9, vis


@pytest.mark.gpu
def test_gpu_cn_momentum_matches_reference(pkg, golden):
    for c in cases(golden):
        g, tag, M, N = c["g"], c["tag"], c["M"], c["N"]
        print(f"io cn momentum {tag}", flush=True)
        co = pkg.channel_setup(c["L"], M, N, STR2=c["STR2"])
        ch = pkg.ChannelIO(co, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"], cf_given=8.18e-3,
                           vis_scheme=1)
        ch.upload(**c["init"], M1=c["m1"])
        ch.ioprops(1)
        ch.momentum(1, c["dt"])
        f = ch.download()
        o = make_oracle(c)
        o.set_state(dict(c["init"], M1=c["m1"]))
        o.props(1)
        o.momentum_cn(1, c["dt"])
        h = N // 2
        for comp in range(3):
            assert rel(f["G"][comp][:h, 1:M + 1], g[f"{tag}_cn_G"][comp][:h, 1:M + 1]) < TOL, comp
            assert rel(f["G"][comp], o.f["G"][comp]) < TOL, comp
            j0 = 1 if comp == 1 else 0
            assert rel(f["EXPLT0"][comp][:, j0:], g[f"{tag}_cn_EXPLT0"][comp][:, j0:]) < TOL, comp
        ch.close()


@pytest.mark.gpu
def test_gpu_cn_substeps_match_reference(pkg, golden):
    for c in cases(golden):
        g, tag, M = c["g"], c["tag"], c["M"]
        print(f"io cn substeps {tag} gravity {c['grav']} driven {c['fd']}", flush=True)
        co = pkg.channel_setup(c["L"], M, c["N"], STR2=c["STR2"])
        ch = pkg.ChannelIO(co, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"], cf_given=8.18e-3,
                           vis_scheme=1)
        ch.upload(**c["init"], M1=c["m1w"])
        for ns in (1, 2, 3):
            ch.ioprops(ns)
            ch.SOLVERRK3_MOM_io(ns, c["dt"])
            check_state(ch.download(), g, f"{tag}_s{ns}", M, first=ns == 1)
        ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("L,M,N,steps,mom,cols", [(64, 48, 32, 3, {}, 16), (48, 36, 40, 3, {}, 16),
                                                  (128, 64, 96, 1, dict(mom_by=8, mom_kz=5), 16),
                                                  (64, 48, 32, 1, dict(mom_kz=0, mom_rows=2), 64),
                                                  (48, 37, 30, 1, dict(mom_kz=0, mom_rows=2), 8),
                                                  (48, 37, 30, 1, dict(mom_kz=0, mom_rows=1), 64),
                                                  (48, 37, 30, 1, dict(mom_by=8, mom_kz=7), 16),
                                                  (128, 64, 96, 1, dict(mom_kz=0, mom_rows=4), 64),
                                                  (64, 48, 32, 2, dict(cn_fastdiv=0), 16),
                                                  (64, 48, 32, 1, dict(cnz_p=4, cny_p=4), 16),
                                                  (48, 37, 30, 1, dict(cnz_p=4, cny_p=8), 8)])
def test_gpu_cn_matches_oracle(pkg, L, M, N, steps, mom, cols):
    """Larger seeded states with a noisy current viscosity against the oracle (gravity and
    constant mass flux on), with each momentum kernel and block shape (the LDS-staged z-march,
    mom_kz planes of mom_by rows, or the per-point kernel, mom_kz = 0, with mom_rows rows;
    ragged M, N), several line-solve block widths (x-columns per block) and points per thread
    (cnz_p, cny_p 4 or 8), and the z and y coefficients by IEEE division (cn_fastdiv = 0) as
    well as by Newton reciprocals."""
    print(f"io cn oracle {L}x{M}x{N} mom {mom} cols {cols}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 90 + L)
    m1 = cn_viscosity(st, 90 + L, False)
    REN, dt = 4000.0, 5e-4
    o = oracle.OracleIO(mesh, REN, gravity=1, f_a=0.02, flow_driven=1, vis_scheme=1)
    o.set_state(dict(st, M1=m1))
    ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), REN, gravity=1, f_a=0.02, flow_driven=1, vis_scheme=1)
    ch.set_tuning(cn_cols=cols, **mom)
    ch.upload(**st, M1=m1)
    for _ in range(steps):
        for ns in (1, 2, 3):
            o.props(ns)
            o.substep(ns, dt)
            ch.ioprops(ns)
            ch.SOLVERRK3_MOM_io(ns, dt)
    f = ch.download()
    for k in ("G", "Q"):
        for c in range(3):
            assert rel(f[k][c], o.f[k][c]) < 1e-11, (k, c, rel(f[k][c], o.f[k][c]))
    assert rel(f["PR"], o.f["PR"]) < 1e-10, rel(f["PR"], o.f["PR"])
    ch.close()


@pytest.mark.gpu
def test_gpu_heated_stage_cn_matches_oracle(pkg):
    """chs_heated_substep with the Crank-Nicolson momentum scheme: the momentum terms read the
    energy context's current viscosity in place."""
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    t = thermo.table_setup(thermo.synthetic_table(), 665.0)
    L, M, N, STR2, REN, dt = 32, 24, 16, 1.9, 2800.0, 5e-4
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    dh, _ = oracle.eng_initial_field(mesh, t, 83)
    st = oracle.io_initial_field(mesh, 84)
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    wf = (0.05, -0.05)
    oe = oracle.OracleEng(mesh, t, (1, 1), cthecd, dhmin, dhmax, wf)
    assert oe.init(dh, st["G"]) == 0
    oi = oracle.OracleIO(mesh, REN, flow_driven=1, vis_scheme=1)
    oi.set_state(dict(st, D0=oe.f["d"], M0=oe.f["m"]))
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    ge = thermo.EnergyTG(co, t, (1, 1), cthecd, dhmin, dhmax, wf)
    ge.upload(dh, st["G"])
    gi = pkg.ChannelIO(co, REN, flow_driven=1, vis_scheme=1)
    gi.upload(G=st["G"], G0=st["G0"], PR=st["PR"], D0=oe.f["d"], M0=oe.f["m"], EXPLT0=st["EXPLT0"])
    for ns in (1, 2, 3):
        oe.f["g"][...] = oi.f["G"]
        assert oe.substep(ns, dt) == 0
        oi.f["D0"][...] = oe.f["d0"]
        oi.f["M0"][...] = oe.f["m0"]
        oi.f["M1"][...] = oe.f["m"]
        oi.props(ns)
        oi.substep(ns, dt)
        pkg.heated_substep(ge, gi, ns, dt)
    f = gi.download()
    for c in range(3):
        assert rel(f["G"][c], oi.f["G"][c]) < 1e-11, (c, rel(f["G"][c], oi.f["G"][c]))
    assert rel(f["PR"], oi.f["PR"]) < 1e-10, rel(f["PR"], oi.f["PR"])
    ge.close()
    gi.close()


def test_oracle_cfl_io_matches_reference(golden):
    """CFL_io on the stage-1 starting velocity Q = G / D_STG, DENSITY = DENSITY0, Viscousity the
    noisy current viscosity (the ``cflio`` fixture); both terms matter on these meshes."""
    for c in cases(golden):
        g, tag = c["g"], c["tag"]
        mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
        o = make_oracle(c)
        o.set_state(dict(c["init"], M1=c["m1"]))
        o.props(1)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = o.f["G"] / o.f["DSTG"]
        q[:, :, 0] = 0.0
        q[:, :, c["M"] + 1] = 0.0
        q[1, :, 1] = 0.0
        v = oracle.cfl_io(mesh, q, c["init"]["D0"], c["m1"], 1.0 / c["REN"])
        assert abs(v - g[f"{tag}_cn_CFLMM"][0]) <= 1e-15 * abs(v), (v, g[f"{tag}_cn_CFLMM"][0])


def test_next_dt_rule(pkg):
    """BCAST_COMM_STEP's rule (SOLVE.f90:177-178): each bound wins in its regime."""
    assert pkg.next_dt(1e-3, 2e-3, 1.0, 10.0, 1e-6) == 2e-3            # DT0
    assert pkg.next_dt(1e-4, 2e-3, 1.0, 10.0, 1e-6) == 5e-4            # five times the last step
    assert pkg.next_dt(1e-3, 2e-3, 1.0, 2000.0, 1e-6) == 1.0 / 2000.0  # CFLGV / CFLMM
    assert pkg.next_dt(1e-3, 2e-3, 1.0, 1e9, 1e-6) == 1e-6             # DTMIN


@pytest.mark.gpu
def test_gpu_cfl_io_matches_reference(pkg, golden):
    for c in cases(golden):
        g, tag = c["g"], c["tag"]
        print(f"cfl_io {tag}", flush=True)
        co = pkg.channel_setup(c["L"], c["M"], c["N"], STR2=c["STR2"])
        ch = pkg.ChannelIO(co, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"], cf_given=8.18e-3,
                           vis_scheme=1)
        ch.upload(**c["init"], M1=c["m1"])
        ch.ioprops(1)
        ref = g[f"{tag}_cn_CFLMM"][0]
        assert abs(ch.cfl(1.0) - ref) <= 1e-14 * ref
        assert abs(ch.cfl(2.5e-4) - ref * 2.5e-4) <= 1e-14 * ref * 2.5e-4
        ch.close()


@pytest.mark.gpu
def test_gpu_heated_cfl_matches_oracle(pkg):
    """chs_heated_cfl reads the energy context's current DENSITY and Viscousity."""
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    t = thermo.table_setup(thermo.synthetic_table(), 665.0)
    L, M, N, STR2, REN, dt = 32, 24, 16, 1.9, 2800.0, 5e-4
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    dh, _ = oracle.eng_initial_field(mesh, t, 85)
    st = oracle.io_initial_field(mesh, 86)
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    ge = thermo.EnergyTG(co, t, (1, 1), cthecd, dhmin, dhmax, (0.05, -0.05))
    ge.upload(dh, st["G"])
    fe0 = ge.download()
    gi = pkg.ChannelIO(co, REN, flow_driven=1, vis_scheme=1)
    gi.upload(G=st["G"], G0=st["G0"], PR=st["PR"], D0=fe0["D"], M0=fe0["M"], EXPLT0=st["EXPLT0"])
    for ns in (1, 2, 3):
        pkg.heated_substep(ge, gi, ns, dt)
    f, fe = gi.download(), ge.download()
    ref = oracle.cfl_io(mesh, f["Q"], fe["D"], fe["M"], 1.0 / REN)
    assert abs(pkg.heated_cfl(ge, gi, 1.0) - ref) <= 1e-14 * ref
    ge.close()
    gi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("REN,dt", [(1.0, 0.5), (50.0, 0.02)])
def test_gpu_cn_line_solve_regimes(pkg, REN, dt):
    """The line kernels solve by Jacobi sweeps while a block's diagonal-dominance ratio rho is
    at most 1/4 and hand the block to the thread-per-line recurrences above it.  REN = 1,
    dt = 0.5: rho > 1/4 in x, z and y (the recurrences); REN = 50, dt = 0.02: rho ~ 1e-3 in x and
    z and ~ 4e-2 at the walls in y (a dozen sweeps).  Momentum part against the oracle."""
    L, M, N = 32, 16, 32
    print(f"io cn line regimes REN {REN} dt {dt}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 95)
    m1 = cn_viscosity(st, 95, False)
    o = oracle.OracleIO(mesh, REN, flow_driven=1, vis_scheme=1)
    o.set_state(dict(st, M1=m1))
    o.props(1)
    o.momentum_cn(1, dt)
    ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), REN, flow_driven=1, vis_scheme=1)
    ch.upload(**st, M1=m1)
    ch.ioprops(1)
    ch.momentum(1, dt)
    f = ch.download()
    for c in range(3):
        assert rel(f["G"][c], o.f["G"][c]) < TOL, (c, rel(f["G"][c], o.f["G"][c]))
    ch.close()


def test_oracle_max_div_io_matches_reference(golden):
    """DIVGCK_io on the stage-1 starting G with the stage's DrhoDtP (the ``divckio`` fixture);
    periodic x: the three maxima agree.  After a whole substep (projection) it is at solver
    precision, for both viscous schemes."""
    for c in cases(golden):
        g, tag = c["g"], c["tag"]
        mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
        ref = g[f"{tag}_cn_MAXDIV"]
        assert ref[0] == ref[1] == ref[2] and ref[0] > 0.1
        o = make_oracle(c)
        o.set_state(dict(c["init"], M1=c["m1"]))
        o.props(1)
        v = oracle.max_div_io(mesh, o.f["G"], o.f["DRHO"])
        assert abs(v - ref[0]) <= 1e-14 * ref[0], (v, ref[0])
        for vis in (0, 1):
            o = make_oracle(c, vis_scheme=vis)
            o.set_state(dict(c["init"], M1=c["m1w"]))
            o.props(1)
            o.substep(1, c["dt"])
            assert oracle.max_div_io(mesh, o.f["G"], o.f["DRHO"]) < 1e-9 * ref[0]


@pytest.mark.gpu
def test_gpu_max_div_io(pkg, golden):
    for c in cases(golden):
        g, tag = c["g"], c["tag"]
        print(f"max_div_io {tag}", flush=True)
        co = pkg.channel_setup(c["L"], c["M"], c["N"], STR2=c["STR2"])
        for vis in (0, 1):
            ch = pkg.ChannelIO(co, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"],
                               cf_given=8.18e-3, vis_scheme=vis)
            ch.upload(**c["init"], **({"M1": c["m1"]} if vis else {}))
            ch.ioprops(1)
            ref = g[f"{tag}_cn_MAXDIV"][0]
            assert abs(ch.max_div() - ref) <= 1e-14 * ref
            ch.SOLVERRK3_MOM_io(1, c["dt"])
            assert ch.max_div() < 1e-9 * ref
            ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [2, 1])
def test_gpu_explicit_momentum_block_shapes(pkg, rows):
    """Explicit viscous scheme: every momentum block shape gives the default shape's substeps
    (only the mass-flux partials' summation order differs)."""
    L, M, N = 48, 37, 30
    print(f"io explicit rows {rows}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 71)
    out = []
    for r in (4, rows):
        ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0, gravity=1, f_a=0.02, flow_driven=1,
                           vis_scheme=0)
        ch.set_tuning(mom_rows=r, mom_kz=0)
        ch.upload(**st)
        for ns in (1, 2, 3):
            ch.ioprops(ns)
            ch.SOLVERRK3_MOM_io(ns, 5e-4)
        out.append(ch.download())
        ch.close()
    for k in ("G", "Q"):
        for c in range(3):
            assert rel(out[1][k][c], out[0][k][c]) < 1e-13, (k, c)
    assert rel(out[1]["PR"], out[0]["PR"]) < 1e-11
    ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0)
    for bad in (dict(mom_rows=3), dict(mom_by=2), dict(mom_kz=-1), dict(mom_nxcd=0), dict(props_kc=-1)):
        with pytest.raises(pkg.ChsError):
            ch.set_tuning(**bad)
    ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("vis", [1, 0])
def test_gpu_momentum_lds_march_equals_per_point_kernel(pkg, vis):
    """The LDS-staged z-march momentum kernel (mom_kz > 0) and staggered-property kernel
    (props_kc > 0) against the per-point kernels (mom_kz = 0, props_kc = 0) under a constant Cf,
    where the mass-flux partials are not used: bit-identical
    substeps (the same expressions from the same values), for both block heights, march
    lengths 1, 3, 5, 8, N and beyond N, tiles grouped for 8, 3 or 1 XCDs, and ragged meshes (L below or not a multiple of 64, M not a
    multiple of the block height, N not a multiple of the march length).  With a constant mass
    flux only the partials' summation order differs."""
    for L, M, N in [(48, 37, 30), (80, 21, 10), (16, 9, 6)]:
        print(f"io momentum kernels {L}x{M}x{N} vis {vis}", flush=True)
        mesh = setup.Mesh(L, M, N, STR2=2.0)
        st = oracle.io_initial_field(mesh, 55 + L)
        if vis == 1:
            st = dict(st, M1=cn_viscosity(st, 55 + L, False))
        for fd in (2, 1):
            out = []
            for mom in (dict(mom_kz=0, props_kc=0), dict(mom_kz=8), dict(mom_kz=3, mom_by=8), dict(mom_kz=1),
                        dict(mom_kz=N, mom_by=8), dict(mom_kz=64), dict(mom_kz=8, mom_nxcd=1),
                        dict(mom_kz=5, mom_nxcd=3), dict(mom_kz=0, props_kc=3), dict(props_kc=1, mom_nxcd=1),
                        dict(props_kc=64)):
                ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0, gravity=2, f_a=0.02,
                                   flow_driven=fd, cf_given=8.18e-3, vis_scheme=vis)
                ch.set_tuning(**mom)
                ch.upload(**st)
                for ns in (1, 2, 3):
                    ch.ioprops(ns)
                    ch.SOLVERRK3_MOM_io(ns, 5e-4)
                out.append((mom, ch.download()))
                ch.close()
            ref = out[0][1]
            for mom, f in out[1:]:
                for k, v in ref.items():
                    if fd == 2:
                        assert np.array_equal(f[k], v), (mom, fd, k)
                    else:
                        assert rel(f[k], v) < 1e-11, (mom, fd, k, rel(f[k], v))


@pytest.mark.gpu
def test_gpu_cn_x_wave_equals_block_solve(pkg):
    """The wavefront-per-line Crank-Nicolson x-solve (cnx_wave = 1, lane l holding L/64
    consecutive points) against the block-per-line one (cnx_wave = 0): at L = 256 both take
    the same per-line sweep count and give the same substeps bit for bit; at L = 64, 128, 512
    the block kernel's count comes from several lines (or L = 512 from one 512-thread line) and
    the results agree to rounding."""
    for L, M, N in [(256, 12, 8), (128, 10, 6), (64, 9, 6), (512, 6, 4)]:
        print(f"io cn x wave {L}x{M}x{N}", flush=True)
        mesh = setup.Mesh(L, M, N, STR2=2.0)
        st = oracle.io_initial_field(mesh, 31 + L)
        st = dict(st, M1=cn_viscosity(st, 31 + L, False))
        out = []
        for w in (0, 1):
            ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0, gravity=1, f_a=0.02, flow_driven=1,
                               vis_scheme=1)
            ch.set_tuning(cnx_wave=w)
            ch.upload(**st)
            for ns in (1, 2, 3):
                ch.ioprops(ns)
                ch.SOLVERRK3_MOM_io(ns, 5e-4)
            out.append(ch.download())
            ch.close()
        for k, v in out[0].items():
            if L == 256:
                assert np.array_equal(out[1][k], v), (L, k)
            else:
                assert rel(out[1][k], v) < 1e-13, (L, k, rel(out[1][k], v))


@pytest.mark.gpu
@pytest.mark.parametrize("vis", [1, 0])
def test_gpu_io_mixed_form(pkg, vis):
    """Mixed form (chs_io_set_mixed: EXPLT0 and the momentum increment in FP32) against the
    plain FP64 form and the oracle over three RK3 steps from a seeded state with a noisy
    viscosity: the error is that of rounding the increments to FP32 (bounded here by 1e-6
    relative in G and 1e-4 in p); switching to the mixed form carries EXPLT0 over, and an
    upload into a mixed context reads the FP64 EXPLT0 given."""
    L, M, N = 64, 48, 32
    print(f"io mixed vis {vis}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 77)
    if vis == 1:
        st = dict(st, M1=cn_viscosity(st, 77, False))
    o = oracle.OracleIO(mesh, 4000.0, gravity=1, f_a=0.02, flow_driven=1, vis_scheme=vis)
    o.set_state(st)
    out = {}
    for mixed in (False, True):
        ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0, gravity=1, f_a=0.02, flow_driven=1,
                           vis_scheme=vis)
        ch.upload(**st)
        if mixed:
            ch.set_mixed(True)       # EXPLT0 carried over
        for step in range(3):
            for ns in (1, 2, 3):
                if not mixed:
                    o.props(ns)
                    o.substep(ns, 5e-4)
                ch.ioprops(ns)
                ch.SOLVERRK3_MOM_io(ns, 5e-4)
        out[mixed] = ch.download()
        ch.close()
    p, m = out[False], out[True]
    errs = {f"G{c}": rel(m["G"][c], p["G"][c]) for c in range(3)}
    errs["PR"] = rel(m["PR"], p["PR"])
    print("mixed vs plain", errs, flush=True)
    for c in range(3):
        assert rel(p["G"][c], o.f["G"][c]) < 1e-11
        assert errs[f"G{c}"] < 1e-6, errs
        assert rel(m["G"][c], o.f["G"][c]) < 1e-6
        assert errs[f"G{c}"] > 0.0   # the FP32 path ran
    assert errs["PR"] < 1e-4, errs
    # an upload into a context already in the mixed form
    ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0, gravity=1, f_a=0.02, flow_driven=1,
                       vis_scheme=vis)
    ch.set_mixed(True)
    ch.upload(**st)
    for ns in (1, 2, 3):
        ch.ioprops(ns)
        ch.SOLVERRK3_MOM_io(ns, 5e-4)
    f = ch.download()
    ch.set_mixed(False)
    ch.close()
    assert np.isfinite(f["G"]).all()
    ch2 = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), 4000.0)
    with pytest.raises(pkg.ChsError):
        ch2._check(pkg.lib().chs_io_set_mixed(ch2._h, 2), "chs_io_set_mixed")
    ch2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("vis", [0, 1])
def test_gpu_io_fishpack_gauge(pkg, vis):
    """chs_io_set_poisson(1): the substep's pressure solve with FISHPACK_POIS3D_SIMPLE's gauge of
    the (0,0) mode.  G and Q agree with the FFT99-gauge substeps to rounding and the pressure
    differs by a constant (the gauges fix only that mode's constant), and DPH equals what
    chs_poisson_fishpack (pinned against the reference's FISHPACK) gives for the substep's
    Poisson right-hand side."""
    L, M, N = 64, 48, 32
    print(f"io fishpack gauge vis {vis}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 61)
    if vis == 1:
        st = dict(st, M1=cn_viscosity(st, 61, False))
    co = pkg.channel_setup(L, M, N, STR2=2.0)
    out = {}
    for fish in (False, True):
        ch = pkg.ChannelIO(co, 4000.0, gravity=1, f_a=0.02, flow_driven=1, vis_scheme=vis)
        ch.set_poisson(fish)
        ch.upload(**st)
        for ns in (1, 2, 3):
            ch.ioprops(ns)
            ch.SOLVERRK3_MOM_io(ns, 5e-4)
        out[fish] = ch.download()
        ch.close()
    f0, f1 = out[False], out[True]
    for k in ("G", "Q"):
        for c in range(3):
            assert rel(f1[k][c], f0[k][c]) < 1e-12, (k, c, rel(f1[k][c], f0[k][c]))
    dp = f1["PR"] - f0["PR"]
    spread = float(dp.max() - dp.min()) / float(np.max(np.abs(f0["PR"])))
    print(f"  pressure offset {float(dp.mean()):.3e}, spread {spread:.2e}", flush=True)
    assert spread < 1e-10
    tg = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    phi = tg.FISHPACK_POIS3D_SIMPLE(f1["RHSP"])
    tg.close()
    assert rel(f1["DPH"][:, 1:M + 1], phi) < 1e-13, rel(f1["DPH"][:, 1:M + 1], phi)
    ch = pkg.ChannelIO(co, 4000.0)
    with pytest.raises(pkg.ChsError):
        ch._check(pkg.lib().chs_io_set_poisson(ch._h, 2), "chs_io_set_poisson")
    ch.close()
