"""Main-domain momentum substep with variable density and viscosity (SOLVERRK3_MOM_io,
explicit viscous scheme): the oracle restatement and the GPU substep against the
reference's own output (tests/golden/io.npz, written by the reference's routines through
oracle/ref_harness.f90 on two ranks).  Tolerances (max-norm relative per field): 1e-12 for
G, Q, the staggered properties and the momentum terms; 1e-10 for the pressure and its
correction (the (0,0)-mode Neumann sweep of the Poisson solve is ill-conditioned)."""
import numpy as np
import pytest

from oracle import oracle, setup

TOL, PTOL = 1e-12, 1e-10


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def cases(golden):
    g = golden("io.npz")
    out = []
    for tag in sorted({k.split("_")[0] for k in g}):
        L, M, N, STR2, REN, dt, grav, fd, seed, T0 = g[tag + "_params"]
        f_a, volm, cvisc = g[tag + "_par"]
        out.append(dict(tag=tag, L=int(L), M=int(M), N=int(N), STR2=float(STR2), REN=float(REN), dt=float(dt),
                        grav=int(grav), fd=int(fd), f_a=float(f_a), volm=float(volm), cvisc=float(cvisc),
                        init={k: g[f"{tag}_init_{k}"] for k in ("G", "G0", "PR", "D0", "M0", "EXPLT0")}, g=g))
    return out


def check_state(f, g, s, M):
    for k in ("G", "Q"):
        for c in range(3):
            assert rel(f[k][c], g[f"{s}_{k}"][c]) < TOL, (s, k, c, rel(f[k][c], g[f"{s}_{k}"][c]))
    for k in ("PR", "DPH"):
        assert rel(f[k], g[f"{s}_{k}"]) < PTOL, (s, k, rel(f[k], g[f"{s}_{k}"]))
    for k in ("DRHO", "RHSP"):
        assert rel(f[k], g[f"{s}_{k}"]) < TOL, (s, k, rel(f[k], g[f"{s}_{k}"]))
    for c in range(3):
        j0 = 1 if c == 1 else 0        # v's row J = 1 is not an unknown
        assert rel(f["EXPLT0"][c][:, j0:], g[f"{s}_EXPLT0"][c][:, j0:]) < TOL


def test_fixture_constants(golden):
    for c in cases(golden):
        mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
        assert abs(c["volm"] - mesh.VOLM) < 1e-13 * c["volm"]   # VOLM_io (COORDJR.f90:391-397)
        assert c["cvisc"] == 1.0 / c["REN"]
        assert (c["f_a"] != 0.0) == (c["grav"] != 0)


def test_oracle_matches_reference(golden):
    for c in cases(golden):
        g, tag, M = c["g"], c["tag"], c["M"]
        mesh = setup.Mesh(c["L"], M, c["N"], STR2=c["STR2"])
        o = oracle.OracleIO(mesh, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"],
                            cf_given=8.18e-3)
        o.set_state(c["init"])
        o.props(1)
        for comp in range(3):
            rows = slice(2, M + 1) if comp == 1 else slice(1, M + 1)
            assert rel(o.f["DSTG"][comp][:, rows], g[f"{tag}_p1_DSTG"][comp][:, rows]) < TOL
            rows = slice(1, M + 1) if comp == 1 else slice(1, M + 2)
            assert rel(o.f["MSTG"][comp][:, rows], g[f"{tag}_p1_MSTG"][comp][:, rows]) < TOL
        assert rel(o.f["DRHO"], g[f"{tag}_p1_DRHO"]) < TOL
        o.terms()
        for comp in range(3):
            j0 = 1 if comp == 1 else 0
            assert rel(o.f["H"][comp][:, j0:], g[f"{tag}_t1_H"][comp][:, j0:]) < TOL, comp
        assert rel(o.f["DIVU"], g[f"{tag}_t1_DIVU"]) < TOL
        for ns in (1, 2, 3):
            if ns > 1:
                o.props(ns)
            o.substep(ns, c["dt"])
            check_state(o.f, g, f"{tag}_s{ns}", M)


def test_oracle_planted_mistakes_fail(golden):
    """The fixtures pin the buoyancy term and the driving: dropping F_A, or switching the
    constant-Cf case to constant mass flux, moves G away from the reference."""
    c = [c for c in cases(golden) if c["grav"] != 0][0]
    g, tag = c["g"], c["tag"]
    mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
    for kw in (dict(gravity=c["grav"], f_a=0.0, flow_driven=c["fd"]),
               dict(gravity=c["grav"], f_a=c["f_a"], flow_driven=1)):
        o = oracle.OracleIO(mesh, c["REN"], cf_given=8.18e-3, **kw)
        o.set_state(c["init"])
        o.props(1)
        o.substep(1, c["dt"])
        assert rel(o.f["G"][0], g[f"{tag}_s1_G"][0]) > 1e-9, kw


@pytest.mark.gpu
def test_gpu_io_substeps_match_reference(pkg, golden):
    for c in cases(golden):
        g, tag, M = c["g"], c["tag"], c["M"]
        print(f"io {tag} gravity {c['grav']} driven {c['fd']}", flush=True)
        co = pkg.channel_setup(c["L"], M, c["N"], STR2=c["STR2"])
        ch = pkg.ChannelIO(co, c["REN"], gravity=c["grav"], f_a=c["f_a"], flow_driven=c["fd"], cf_given=8.18e-3)
        ch.upload(**c["init"])
        for ns in (1, 2, 3):
            ch.ioprops(ns)
            ch.SOLVERRK3_MOM_io(ns, c["dt"])
            check_state(ch.download(), g, f"{tag}_s{ns}", M)
        ch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("L,M,N", [(64, 48, 32), (48, 36, 40), (128, 64, 96)])
def test_gpu_io_matches_oracle(pkg, L, M, N):
    """Larger seeded states against the oracle over three RK3 steps (gravity and constant
    mass flux on)."""
    print(f"io oracle {L}x{M}x{N}", flush=True)
    mesh = setup.Mesh(L, M, N, STR2=2.0)
    st = oracle.io_initial_field(mesh, 70 + L)
    REN, dt = 4000.0, 5e-4
    o = oracle.OracleIO(mesh, REN, gravity=1, f_a=0.02, flow_driven=1)
    o.set_state(st)
    ch = pkg.ChannelIO(pkg.channel_setup(L, M, N, STR2=2.0), REN, gravity=1, f_a=0.02, flow_driven=1)
    ch.upload(**st)
    for _ in range(3):
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
def test_gpu_heated_stage_matches_oracle(pkg):
    """chs_heated_substep (energy substep, staggered properties, momentum substep, in SOLVE's
    order) against the same composition of the two pinned oracles, one RK3 step."""
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    t = thermo.table_setup(thermo.synthetic_table(), 665.0)
    L, M, N, STR2, REN, dt = 32, 24, 16, 1.9, 2800.0, 5e-4
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    dh, _ = oracle.eng_initial_field(mesh, t, 81)
    st = oracle.io_initial_field(mesh, 82)
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    wf = (0.05, -0.05)
    oe = oracle.OracleEng(mesh, t, (1, 1), cthecd, dhmin, dhmax, wf)
    assert oe.init(dh, st["G"]) == 0
    oi = oracle.OracleIO(mesh, REN, flow_driven=1)
    oi.set_state(dict(st, D0=oe.f["d"], M0=oe.f["m"]))
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    ge = thermo.EnergyTG(co, t, (1, 1), cthecd, dhmin, dhmax, wf)
    ge.upload(dh, st["G"])
    gi = pkg.ChannelIO(co, REN, flow_driven=1)
    gi.upload(G=st["G"], G0=st["G0"], PR=st["PR"], D0=oe.f["d"], M0=oe.f["m"], EXPLT0=st["EXPLT0"])
    for ns in (1, 2, 3):
        oe.f["g"][...] = oi.f["G"]
        assert oe.substep(ns, dt) == 0
        oi.f["D0"][...] = oe.f["d0"]
        oi.f["M0"][...] = oe.f["m0"]
        oi.props(ns)
        oi.substep(ns, dt)
        pkg.heated_substep(ge, gi, ns, dt)
    f, fe = gi.download(), ge.download()
    for c in range(3):
        assert rel(f["G"][c], oi.f["G"][c]) < 1e-11, (c, rel(f["G"][c], oi.f["G"][c]))
    assert rel(f["PR"], oi.f["PR"]) < 1e-10, rel(f["PR"], oi.f["PR"])
    for k, kk in (("dh", "DH"), ("t", "T"), ("d", "D")):
        assert rel(fe[kk], oe.f[k]) < 1e-11, (kk, rel(fe[kk], oe.f[k]))
    ge.close()
    gi.close()
