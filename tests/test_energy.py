"""Energy equation of the heated channel (SOLVERRK3_ENG_io, table properties): the oracle
restatement and the GPU substep against the reference's own output (tests/golden/eng.npz,
written by the reference's SOLVERRK3_ENG_io through oracle/ref_harness.f90 on two ranks).
Tolerance 1e-12 relative (max-norm) per field: both follow the reference's operation order
with FMA contraction off, so what remains is the rounding of the compilers' own choices."""
import numpy as np
import pytest

from oracle import oracle, setup

FIELDS = (("dh", "DH"), ("h", "H"), ("t", "T"), ("d", "D"), ("m", "M"), ("k", "K"), ("cp", "CP"))
TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def cases(golden):
    g = golden("eng.npz")
    tags = sorted({k.split("_")[0] for k in g})
    out = []
    for tag in tags:
        L, M, N, STR2, REN, dt, w0, w1, b0, b1, T0, seed = g[tag + "_params"]
        par = g[tag + "_par"]
        out.append(dict(tag=tag, L=int(L), M=int(M), N=int(N), STR2=float(STR2), dt=float(dt),
                        wt=(int(w0), int(w1)), T0=float(T0), cthecd=par[0], dhmin=par[1], dhmax=par[2],
                        wf=(par[3], par[4]), ws=par[5:15].reshape(5, 2).T, top_face=int(par[15]), g=g))
    return out


@pytest.fixture(scope="module")
def table(pkg):
    from importlib import import_module
    thermo = import_module(pkg.__name__ + ".thermo")
    return thermo, thermo.table_setup(thermo.synthetic_table(), 665.0)


def test_fixture_constants_follow_from_table(golden, table):
    """The reference's CTHECD = 1/(REN PRT0), PRT0 = M0 CP0 / K0, DHmin = DH(IMIN_DH),
    DHmax = DH(IMAX_DH) equal the host set-up's values bit for bit."""
    _, t = table
    for c in cases(golden):
        REN = float(c["g"][c["tag"] + "_params"][4])
        assert 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"]) == c["cthecd"]
        assert t["DH"][t["IMIN_DH"] - 1] == c["dhmin"] and t["DH"][t["IMAX_DH"] - 1] == c["dhmax"]
        # two ranks: the top wall's face index is the last rank's row count
        assert c["top_face"] == c["M"] // 2


def test_oracle_matches_reference(golden, table):
    _, t = table
    for c in cases(golden):
        g = c["g"]
        mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
        o = oracle.OracleEng(mesh, t, c["wt"], c["cthecd"], c["dhmin"], c["dhmax"], c["wf"],
                             wall_state=c["ws"], top_face=c["top_face"])
        assert o.init(g[c["tag"] + "_dh"], g[c["tag"] + "_g"]) == 0
        for ns in (0, 1, 2, 3):
            if ns:
                assert o.substep(ns, c["dt"]) == 0
            s = f"{c['tag']}_s{ns}"
            for k, kk in FIELDS:
                assert rel(o.f[k], g[f"{s}_{kk}"]) < TOL, (s, k, rel(o.f[k], g[f"{s}_{kk}"]))
            assert rel(o.f["rhs0"], g[f"{s}_RHS0"]) < TOL if ns else np.all(o.f["rhs0"] == 0)


def test_oracle_wall_face_index_matters(golden, table):
    """Planted mistake: with the one-rank face index (M) the iso-flux case's top ghost row
    departs from the two-rank reference output, so the fixture pins the index."""
    _, t = table
    c = [c for c in cases(golden) if c["wt"] == (1, 1)][0]
    g = c["g"]
    mesh = setup.Mesh(c["L"], c["M"], c["N"], STR2=c["STR2"])
    o = oracle.OracleEng(mesh, t, c["wt"], c["cthecd"], c["dhmin"], c["dhmax"], c["wf"], top_face=c["M"])
    o.init(g[c["tag"] + "_dh"], g[c["tag"] + "_g"])
    assert rel(o.f["dh"][:, -1, :], g[c["tag"] + "_s0_DH"][:, -1, :]) > 1e-6


@pytest.mark.gpu
def test_gpu_substeps_match_reference(pkg, golden, table):
    thermo, t = table
    for c in cases(golden):
        print(f"energy {c['tag']} walls {c['wt']}", flush=True)
        g = c["g"]
        co = pkg.channel_setup(c["L"], c["M"], c["N"], STR2=c["STR2"])
        e = thermo.EnergyTG(co, t, c["wt"], c["cthecd"], c["dhmin"], c["dhmax"], c["wf"],
                            wall_state=c["ws"], top_face=c["top_face"])
        e.upload(g[c["tag"] + "_dh"], g[c["tag"] + "_g"])
        for ns in (0, 1, 2, 3):
            if ns:
                e.SOLVERRK3_ENG_io(ns, c["dt"])
            out = e.download()
            s = f"{c['tag']}_s{ns}"
            for _, kk in FIELDS:
                assert rel(out[kk], g[f"{s}_{kk}"]) < TOL, (s, kk, rel(out[kk], g[f"{s}_{kk}"]))
            if ns:
                assert rel(out["RHS0"], g[f"{s}_RHS0"]) < TOL
        e.close()


@pytest.mark.gpu
def test_gpu_energy_matches_oracle_large(pkg, table):
    """A larger seeded state (64 x 48 x 32, iso-flux bottom, fixed-T top, one-rank face index)
    against the oracle over a full RK3 step."""
    thermo, t = table
    L, M, N, STR2, dt = 64, 48, 32, 1.9, 5e-4
    mesh = setup.Mesh(L, M, N, STR2=STR2)
    dh, gg = oracle.eng_initial_field(mesh, t, 77)
    REN = 2800.0
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    Tw = 675.0 / t["T0"]
    hw = float(np.interp(Tw, t["T"], t["H"]))
    ws = [[0] * 5, [Tw, hw, float(np.interp(hw, t["H"], t["D"])), float(np.interp(hw, t["H"], t["M"])),
                   float(np.interp(hw, t["H"], t["K"]))]]
    wf = (0.1, 0.0)
    o = oracle.OracleEng(mesh, t, (1, 2), cthecd, dhmin, dhmax, wf, wall_state=ws)
    assert o.init(dh, gg) == 0
    e = thermo.EnergyTG(pkg.channel_setup(L, M, N, STR2=STR2), t, (1, 2), cthecd, dhmin, dhmax, wf, wall_state=ws)
    e.upload(dh, gg)
    for ns in (1, 2, 3):
        assert o.substep(ns, dt) == 0
        e.SOLVERRK3_ENG_io(ns, dt)
    out = e.download()
    for k, kk in FIELDS:
        assert rel(out[kk], o.f[k]) < TOL, (kk, rel(out[kk], o.f[k]))
    e.close()


@pytest.mark.gpu
def test_gpu_cell_props_equal_lookup(pkg, table):
    """The energy context's property update (two cells per thread, searches in lockstep over the
    bucket index) gives bit for bit what the single-value lookup gives, for DH at the table's
    knots, between them, and at and beside the bucket edges."""
    thermo, t = table
    L, M, N, STR2 = 64, 48, 32, 1.9
    x = np.asarray(t["DH"])
    lo, hi = t["IMIN_DH"] - 1, t["IMAX_DH"] - 1
    nb = 2 * len(x)
    edges = x[0] + np.arange(nb) * ((x[-1] - x[0]) / nb)
    vals = np.concatenate([x[lo:hi + 1], 0.5 * (x[lo:hi] + x[lo + 1:hi + 1]), edges,
                           np.nextafter(edges, -np.inf), np.nextafter(edges, np.inf)])
    vals = vals[(vals >= x[lo]) & (vals <= x[hi])]
    cells = np.resize(vals, N * M * L).reshape(N, M, L)
    dh = np.zeros((N, M + 2, L))
    dh[:, 1:M + 1] = cells
    dh[:, 0], dh[:, M + 1] = cells[:, 0], cells[:, -1]
    REN = 2800.0
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    Tw = 675.0 / t["T0"]
    hw = float(np.interp(Tw, t["T"], t["H"]))
    ws = [[Tw, hw, float(np.interp(hw, t["H"], t["D"])), float(np.interp(hw, t["H"], t["M"])),
           float(np.interp(hw, t["H"], t["K"]))]] * 2
    e = thermo.EnergyTG(pkg.channel_setup(L, M, N, STR2=STR2), t, (2, 2), cthecd, x[lo], x[hi], (0.0, 0.0),
                        wall_state=ws)
    e.upload(dh, np.zeros((3, N, M + 2, L)))
    out = e.download()
    e.close()
    pt = thermo.PropTable(t)
    ref = pt.from_dh(cells.ravel())
    pt.close()
    for i, kk in enumerate(("H", "T", "D", "M", "K", "CP")):
        np.testing.assert_array_equal(out[kk][:, 1:M + 1].ravel(), ref[i], err_msg=kk)


@pytest.mark.gpu
def test_gpu_energy_rhs_lds_equals_per_point(pkg, table):
    """The energy right-hand side from LDS-staged planes (rhs_lds = 1) against the per-point
    kernel (rhs_lds = 0): bit-identical substeps (the same expressions from the same values),
    for ragged meshes (L below or not a multiple of 64, M not a multiple of 4, N not a
    multiple of the march), march lengths 1, 3, 32, beyond N, tiles grouped for 8, 3 or 1
    XCDs, both wall kinds on either side."""
    thermo, t = table
    REN, STR2, dt = 2800.0, 1.9, 5e-4
    cthecd = 1.0 / REN / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    for L, M, N in [(64, 48, 32), (48, 37, 30), (80, 21, 10), (16, 9, 6)]:
        print(f"energy rhs kernels {L}x{M}x{N}", flush=True)
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        dh, gg = oracle.eng_initial_field(mesh, t, 50 + L)
        Tw = 675.0 / t["T0"]
        hw = float(np.interp(Tw, t["T"], t["H"]))
        ws = [[Tw, hw, float(np.interp(hw, t["H"], t["D"])), float(np.interp(hw, t["H"], t["M"])),
               float(np.interp(hw, t["H"], t["K"]))]] * 2
        for wt in [(1, 2), (2, 1), (1, 1)]:
            out = []
            for shape in (dict(rhs_lds=0), dict(rhs_lds=1), dict(rhs_lds=1, rhs_kc=3, rhs_nxcd=3),
                          dict(rhs_lds=1, rhs_kc=1, rhs_nxcd=1), dict(rhs_lds=1, rhs_kc=64)):
                e = thermo.EnergyTG(pkg.channel_setup(L, M, N, STR2=STR2), t, wt, cthecd, dhmin, dhmax, (0.1, -0.05),
                                    wall_state=ws)
                e.set_tuning(**shape)
                e.upload(dh, gg)
                for ns in (1, 2, 3):
                    e.SOLVERRK3_ENG_io(ns, dt)
                out.append((shape, e.download()))
                e.close()
            for shape, f in out[1:]:
                for k, v in out[0][1].items():
                    assert np.array_equal(f[k], v), (L, M, N, wt, shape, k)
    e = thermo.EnergyTG(pkg.channel_setup(16, 9, 6, STR2=STR2), t, (1, 1), cthecd, dhmin, dhmax, (0.1, -0.05))
    for bad in (dict(rhs_lds=2), dict(rhs_kc=0), dict(rhs_nxcd=0), dict(nope=1)):
        with pytest.raises(pkg.ChsError):
            e.set_tuning(**bad)
    e.close()
