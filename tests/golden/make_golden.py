"""Writes the golden fixtures in tests/golden/ from the compiled reference (oracle/_ref,
built by ``make -f oracle/ref.mk`` from the reference's Fortran sources).

Each ``tg_<L>x<M>x<N>.npz`` holds, for one channel mesh:
  setup arrays as dumped by the reference (prefix ``s_``),
  the seeded initial field (``q_init``, ``pr_init``; oracle.initial_field),
  the reference convection terms of that field (``conv``),
  the field after the NS=0 projection substep (``q_ns0``, ``pr_ns0``),
  the field after one RK3 step (NS=1,2,3) from there (``q_step1``, ``pr_step1``),
  a reference Poisson solve of a seeded compatible RHS (``poisson_rhs``, ``poisson_phi``).
Arrays are C-order views of the reference's Fortran layout (z, y, x).

Also ``props_2048.npz`` (property tables, make_props), ``tdma.npz`` (the reference's
batched Thomas solvers, make_tdma) and ``fishpack.npz`` (the main-domain FISHPACK Poisson
solve, make_fishpack), ``eng.npz`` (the energy substep, make_eng) and ``io.npz`` (the
main-domain momentum substep, make_io) and ``meanzx.npz`` (x-z mean statistics,
make_meanzx) and ``cfl_div_tg.npz`` (CFL_tg and DIVGCK_tg, make_cfl).

Run: python tests/golden/make_golden.py [tg] [props] [tdma] [fishpack] [eng] [io] [meanzx] [cfl]
(needs oracle/_ref)
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle, refrun, setup  # noqa: E402

CASES = [
    # L, M, N, STR2, REN, dt, seed
    (16, 16, 16, 1.8, 2800.0, 1.0e-3, 11),
    (20, 18, 30, 2.2, 6875.0, 5.0e-4, 12),
]


def compatible_rhs(mesh, seed):
    rng = np.random.Generator(np.random.Philox(key=seed))
    r = rng.standard_normal((mesh.N, mesh.M, mesh.L))
    w = (1.0 / mesh.DYFI)[None, :, None]
    r -= np.sum(r * w) / (np.sum(w) * mesh.L * mesh.N)
    return r


def make_case(L, M, N, STR2, REN, dt, seed):
    h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt)
    try:
        s = h.setup()
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        q, pr = oracle.initial_field(mesh, seed)
        h.write_fields("init.bin", q[0], q[1], q[2], pr)
        rhs = compatible_rhs(mesh, seed + 1000)
        h.write_cells("prhs.bin", rhs)
        h.run(["load init.bin", "conv c0", "substep 0 %r" % dt, "dump s0",
               "substep 1 %r" % dt, "substep 2 %r" % dt, "substep 3 %r" % dt, "dump s1",
               "poisson prhs.bin phi"])
        out = {f"s_{k}": np.asarray(v) for k, v in s.items()}
        out.update(q_init=q, pr_init=pr, conv=h.read_cells("c0", 3),
                   poisson_rhs=rhs, poisson_phi=h.read_cells("phi"),
                   params=np.array([L, M, N, STR2, REN, dt, seed], dtype=np.float64))
        u, v, w, p = h.read_q("s0")
        out.update(q_ns0=np.stack([u, v, w]), pr_ns0=p)
        u, v, w, p = h.read_q("s1")
        out.update(q_step1=np.stack([u, v, w]), pr_step1=p)
    finally:
        h.close()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"tg_{L}x{M}x{N}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def make_cfl():
    """``cfl_div_tg.npz``: CFLMM_tg (CFL_tg) and MAXDIVGV_tg (DIVGCK_tg) of the reference on
    the fields of the tg fixtures (the seeded start, which is not divergence-free, the NS = 0
    projection and one RK3 step), for each mesh: keys ``<L>x<M>x<N>_<field>_{cfl,div}``."""
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for (L, M, N, STR2, REN, dt, seed) in CASES:
        g = dict(np.load(os.path.join(here, f"tg_{L}x{M}x{N}.npz")))
        h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt)
        try:
            for tag, q, pr in (("init", g["q_init"], g["pr_init"]), ("ns0", g["q_ns0"], g["pr_ns0"]),
                               ("step1", g["q_step1"], g["pr_step1"])):
                h.write_fields("f.bin", q[0], q[1], q[2], pr)
                h.run(["load f.bin", "cfltg cfl.bin", "divcktg div.bin"])
                out[f"{L}x{M}x{N}_{tag}_cfl"] = np.fromfile(h.path("cfl.bin"), dtype=np.float64)
                out[f"{L}x{M}x{N}_{tag}_div"] = np.fromfile(h.path("div.bin"), dtype=np.float64)
        finally:
            h.close()
    path = os.path.join(here, "cfl_div_tg.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def make_props(T0=665.0, seed=21):
    """Property-table fixture: the reference's nondimensional table and splines for the
    synthetic 2048-point table (chapsim_amd.thermo.synthetic_table), and its
    THERM_PROP_UPDATE_FROM_DH outputs for table nodes, range ends, points just inside the
    1e-8 tolerance and seeded random values."""
    sys.path.insert(0, os.path.join(ROOT))
    import __graft_entry__
    __graft_entry__.load()
    from chapsim_amd import thermo
    rows = thermo.synthetic_table()
    h = refrun.Harness(8, 8, 8, iDomain=2, iThermoDynamics=1, T0=T0, Ti=T0,
                       iThermalWallType=(1, 1), thermalWallBC=(1.0e3, -1.0e3))
    try:
        thermo.write_table_file(h.path("NIST_WATER_23.5MP.DAT"), rows)
        h.run(["thermoinit", "dumptable table.bin"])
        t = h.read_table("table.bin")
        lo, hi = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
        rng = np.random.Generator(np.random.Philox(key=seed))
        dh = np.concatenate([t["DH"][::7], [lo, hi, lo - 1e-10, hi + 1e-10],
                             rng.uniform(lo, hi, 4000)])
        h.prop_lookup(dh)
        h.run(["thermoinit", "proplookup dh.bin props.bin"])
        props = np.fromfile(h.path("props.bin"), dtype=np.float64).reshape(7, len(dh))
    finally:
        h.close()
    out = {f"t_{k}": np.asarray(v) for k, v in t.items()}
    out.update(rows=rows, T0=np.array(T0), dh=dh, props=props)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "props_2048.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


TDMA_SIZES = [(7, 5), (16, 9), (3, 4), (4, 3), (33, 12)]


def make_tdma(seed=31):
    """Batched Thomas fixture: for each kind of the reference's TDMAIJI_CYC, TDMAIJJ_nonCYC,
    TDMAIJI_nonCYC, TDMAIJJ_CYC and each (isz, jsz) of TDMA_SIZES, seeded diagonally dominant
    coefficients, RHS and boundary values, and R after the reference's solve."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    out = {}
    h = refrun.Harness(8, 8, 8)
    try:
        cmds, cases = [], []
        for kind in range(4):
            for isz, jsz in TDMA_SIZES:
                A = rng.uniform(-1, 1, (jsz, isz))
                C_ = rng.uniform(-1, 1, (jsz, isz))
                B = 2.5 + rng.uniform(0, 1, (jsz, isz))
                R = rng.standard_normal((jsz, isz))
                nl = jsz if kind == 2 else isz
                bc = rng.standard_normal((2, nl)) if kind in (1, 2) else np.zeros((2, nl))
                tag = f"k{kind}_{isz}x{jsz}"
                with open(h.path(tag + ".in"), "wb") as fh:
                    fh.write(np.array([isz, jsz], dtype=np.int32).tobytes())
                    for x in (A, B, C_, R):
                        fh.write(np.ascontiguousarray(x).tobytes())   # C [jsz, isz] = Fortran (isz, jsz)
                    if kind in (1, 2):
                        fh.write(np.ascontiguousarray(bc).tobytes())  # C [2, nl] = Fortran (nl, 2)
                cmds.append(f"tdma {kind} {tag}.in {tag}.out")
                cases.append((tag, A, B, C_, R, bc))
        h.run(cmds)
        for tag, A, B, C_, R, bc in cases:
            isz = A.shape[1]
            out.update({f"{tag}_A": A, f"{tag}_B": B, f"{tag}_C": C_, f"{tag}_R": R, f"{tag}_bc": bc,
                        f"{tag}_X": np.fromfile(h.path(tag + ".out"), dtype=np.float64).reshape(-1, isz)})
    finally:
        h.close()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tdma.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


FISHPACK_MESHES = [(16, 12, 24, 2.2), (32, 20, 16, 2.2), (64, 8, 64, 2.0)]


def make_fishpack(seed=41):
    """Main-domain Poisson fixture: seeded compatible RHS (Delta-y-weighted mean removed) and
    the reference FISHPACK_POIS3D_SIMPLE solution (harness command poissonio) per mesh."""
    out = {}
    for L, M, N, STR2 in FISHPACK_MESHES:
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        rhs = compatible_rhs(mesh, seed + L)
        h = refrun.Harness(L, M, N, STR2=STR2, iDomain=2)
        try:
            h.write_cells("prhs.bin", rhs)
            h.run(["poissonio prhs.bin phi"])
            phi = h.read_cells("phi")
        finally:
            h.close()
        tag = f"m{L}x{M}x{N}"
        out.update({f"{tag}_rhs": rhs, f"{tag}_phi": phi,
                    f"{tag}_params": np.array([L, M, N, STR2], dtype=np.float64)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fishpack.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


ENG_CASES = [  # L, M, N, STR2, REN, dt, wall types, thermalWallBC_Dim, seed
    (16, 12, 8, 1.8, 2800.0, 1.0e-3, (1, 1), (2.0e4, 2.0e4), 51),
    (12, 10, 16, 2.0, 2800.0, 5.0e-4, (2, 2), (655.0, 675.0), 52),
]


def make_eng(T0=665.0):
    """Energy-equation fixture (SOLVERRK3_ENG_io on the main domain, table mode): per case the
    seeded DH and G, the reference's constants (dumpengpar) and the state after loading
    (loadeng: properties from DH, walls) and after each of the substeps NS = 1, 2, 3."""
    sys.path.insert(0, os.path.join(ROOT))
    import __graft_entry__
    __graft_entry__.load()
    from chapsim_amd import thermo
    rows = thermo.synthetic_table()
    t = thermo.table_setup(rows, T0)
    out = {}
    for L, M, N, STR2, REN, dt, wt, wbc, seed in ENG_CASES:
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        dh, g = oracle.eng_initial_field(mesh, t, seed)
        h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt, iDomain=2, iThermoDynamics=1, T0=T0, Ti=T0,
                           iThermalWallType=wt, thermalWallBC=wbc)
        try:
            thermo.write_table_file(h.path("NIST_WATER_23.5MP.DAT"), rows)
            h.write_eng("eng.bin", dh, g)
            h.run(["thermoinitall", "dumpengpar par.bin", "loadeng eng.bin", "dumpeng s0",
                   f"engsubstep 1 {dt!r}", "dumpeng s1", f"engsubstep 2 {dt!r}", "dumpeng s2",
                   f"engsubstep 3 {dt!r}", "dumpeng s3"])
            par = np.fromfile(h.path("par.bin"), dtype=np.float64)
            states = {s: h.read_eng(s) for s in ("s0", "s1", "s2", "s3")}
        finally:
            h.close()
        tag = f"e{L}x{M}x{N}"
        out[f"{tag}_params"] = np.array([L, M, N, STR2, REN, dt, wt[0], wt[1], wbc[0], wbc[1], T0, seed])
        out[f"{tag}_par"] = par
        out[f"{tag}_dh"] = dh
        out[f"{tag}_g"] = g
        for s, st in states.items():
            for k, v in st.items():
                out[f"{tag}_{s}_{k}"] = v
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eng.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


IO_CASES = [  # L, M, N, STR2, REN, dt, iGravity, iFlowDriven, seed
    (16, 12, 8, 1.8, 2800.0, 1.0e-3, 0, 1, 61),
    (12, 10, 16, 2.0, 4000.0, 5.0e-4, 1, 2, 62),
]


def make_io(T0=665.0):
    """Main-domain momentum fixture (SOLVERRK3_MOM_io, explicit viscous scheme, FFT99 Poisson
    solve): per case the seeded state, the reference's F_A / VOLM_io / CVISC, the staggered
    properties and the convection + viscous terms of stage 1, and the state after each of the
    substeps NS = 1, 2, 3 (each preceded by the staggered-property update SOLVE does)."""
    sys.path.insert(0, os.path.join(ROOT))
    import __graft_entry__
    __graft_entry__.load()
    from chapsim_amd import thermo
    rows = thermo.synthetic_table()
    out = {}
    for L, M, N, STR2, REN, dt, grav, fd, seed in IO_CASES:
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        st = oracle.io_initial_field(mesh, seed)
        h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt, iDomain=2, iThermoDynamics=1, T0=T0, Ti=T0,
                           iVisScheme=0, iGravity=grav, iFlowDriven=fd)
        try:
            thermo.write_table_file(h.path("NIST_WATER_23.5MP.DAT"), rows)
            refrun.write_io(h, "io.bin", st["G"], st["G0"], st["PR"], st["D0"], st["M0"], st["EXPLT0"])
            h.run(["thermoinitall", "dumpiopar par.bin", "loadio io.bin", "ioprops 1", "ioterms t1", "dumpio p1",
                   f"iosubstep 1 {dt!r}", "dumpio s1", "ioprops 2", f"iosubstep 2 {dt!r}", "dumpio s2",
                   "ioprops 3", f"iosubstep 3 {dt!r}", "dumpio s3"])
            par = np.fromfile(h.path("par.bin"), dtype=np.float64)
            p1 = refrun.read_io(h, "p1")
            t1 = refrun.read_ioterms(h, "t1")
            ss = {s: refrun.read_io(h, s) for s in ("s1", "s2", "s3")}
        finally:
            h.close()
        tag = f"io{L}x{M}x{N}"
        out[f"{tag}_params"] = np.array([L, M, N, STR2, REN, dt, grav, fd, seed, T0])
        out[f"{tag}_par"] = par
        for k, v in st.items():
            out[f"{tag}_init_{k}"] = v
        for k in ("DSTG", "MSTG", "DRHO"):
            out[f"{tag}_p1_{k}"] = p1[k]
        out[f"{tag}_t1_H"] = t1["H"]
        out[f"{tag}_t1_DIVU"] = t1["DIVU"]
        for s, r in ss.items():
            for k in ("G", "Q", "PR", "DPH", "DRHO", "RHSP", "EXPLT0"):
                out[f"{tag}_{s}_{k}"] = r[k]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "io.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def cn_viscosity(st, seed, walls_zero):
    """Current viscosity for the Crank-Nicolson fixtures: Viscousity0 with 10 % seeded noise;
    walls_zero: the ghost rows are the negatives of the wall cells, so the wall-row means the
    u, w y-solves use are zero."""
    rng = np.random.Generator(np.random.Philox(key=seed + 1000))
    m1 = st["M0"] * (1.0 + 0.1 * rng.standard_normal(st["M0"].shape))
    if walls_zero:
        M = m1.shape[1] - 2
        m1[:, 0] = -m1[:, 1]
        m1[:, M + 1] = -m1[:, M]
    return m1


def make_io_cn(T0=665.0):
    """Main-domain momentum fixture, Crank-Nicolson viscous scheme (iVisScheme = 1), from the
    IO_CASES states (io.npz holds them):
      cn_*   ``iocn`` of stage 1 with a noisy current viscosity: G after the three MOMFA_io
             calls, RHS_io as each receives it, EXPLT0, DRHOI_STG; CFLMM_io of CFL_io on
             the stage's starting velocity (``cflio``) and MAXDIVGV_io of DIVGCK_io on the
             starting G with the stage's DrhoDtP (``divckio``);
      s1..3  three substeps (``iosubstep``, FFT99 Poisson) with the wall-row viscosity means
             zero: G, PR, EXPLT0 after each, RHSP and DPH of the first.
    The reference runs on two ranks; the first inputs are compared on the first half of z
    only, the second anywhere (the tests say why)."""
    sys.path.insert(0, os.path.join(ROOT))
    import __graft_entry__
    __graft_entry__.load()
    from chapsim_amd import thermo
    rows = thermo.synthetic_table()
    out = {}
    for L, M, N, STR2, REN, dt, grav, fd, seed in IO_CASES:
        mesh = setup.Mesh(L, M, N, STR2=STR2)
        st = oracle.io_initial_field(mesh, seed)
        m1 = cn_viscosity(st, seed, False)
        m1w = cn_viscosity(st, seed, True)
        tag = f"io{L}x{M}x{N}"
        for run, visc, cmds in (
                ("cn", m1, ["ioprops 1", "cflio cfl.bin", "divckio dck.bin", f"iocn 1 {dt!r} c1"]),
                ("sub", m1w, ["ioprops 1", f"iosubstep 1 {dt!r}", "dumpio s1", "ioprops 2", f"iosubstep 2 {dt!r}",
                              "dumpio s2", "ioprops 3", f"iosubstep 3 {dt!r}", "dumpio s3"])):
            h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt, iDomain=2, iThermoDynamics=1, T0=T0, Ti=T0,
                               iVisScheme=1, iGravity=grav, iFlowDriven=fd)
            try:
                thermo.write_table_file(h.path("NIST_WATER_23.5MP.DAT"), rows)
                refrun.write_io(h, "io.bin", st["G"], st["G0"], st["PR"], st["D0"], st["M0"], st["EXPLT0"])
                visc.tofile(h.path("m1.bin"))
                h.run(["thermoinitall", "loadio io.bin", "loadvisc m1.bin"] + cmds)
                if run == "cn":
                    r = refrun.read_iocn(h, "c1")
                    for k in ("G", "RHS", "EXPLT0", "DRHOI"):
                        out[f"{tag}_cn_{k}"] = r[k]
                    out[f"{tag}_cn_CFLMM"] = np.fromfile(h.path("cfl.bin"), dtype=np.float64)
                    out[f"{tag}_cn_MAXDIV"] = np.fromfile(h.path("dck.bin"), dtype=np.float64)
                else:
                    for sname in ("s1", "s2", "s3"):
                        r = refrun.read_io(h, sname)
                        for k in ("G", "PR", "EXPLT0") + (("RHSP", "DPH") if sname == "s1" else ()):
                            out[f"{tag}_{sname}_{k}"] = r[k]
            finally:
                h.close()
        out[f"{tag}_params"] = np.array([L, M, N, STR2, REN, dt, grav, fd, seed, T0])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "io_cn.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def make_meanzx():
    """x-z mean statistics fixture (PP_MEAN_ZX_FLOW_TG) of the fields after one RK3 step held
    in the tg_*.npz fixtures: stats [M, 122] (or_mean_zx's order) and U1mean / U1maxx."""
    out = {}
    for name in ("tg_16x16x16.npz", "tg_20x18x30.npz"):
        g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), name))
        L, M, N, STR2, REN, dt = (g["params"][i] for i in range(6))
        L, M, N = int(L), int(M), int(N)
        h = refrun.Harness(L, M, N, STR2=STR2, REN=REN, DT=dt)
        try:
            q, p = g["q_step1"], g["pr_step1"]
            h.write_fields("f.bin", q[0], q[1], q[2], p)
            h.run(["load f.bin", "meanzx mz"])
            st, sc = refrun.read_meanzx(h, "mz")
        finally:
            h.close()
        tag = name[:-4]
        out[f"{tag}_stats"] = st
        out[f"{tag}_scal"] = sc
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meanzx.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    import sys as _sys
    which = _sys.argv[1:] or ["tg", "props", "tdma", "fishpack", "eng", "io", "io_cn", "meanzx", "cfl"]
    if "tg" in which:
        for case in CASES:
            make_case(*case)
    if "props" in which:
        make_props()
    if "tdma" in which:
        make_tdma()
    if "fishpack" in which:
        make_fishpack()
    if "eng" in which:
        make_eng()
    if "io" in which:
        make_io()
    if "io_cn" in which:
        make_io_cn()
    if "meanzx" in which:
        make_meanzx()
    if "cfl" in which:
        make_cfl()
