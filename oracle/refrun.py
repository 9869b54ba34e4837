"""Runner for the reference Fortran build in oracle/_ref (test infrastructure only).

Writes a ``readdata.ini`` in the reference's input format (READINI.f90:40-600: twelve
sections read in fixed key order), feeds raw input fields to ``oracle/_ref/ref_harness``
(oracle/ref_harness.f90) under ``mpirun -np 2`` and reassembles the per-rank dumps into
global arrays.  Only tests, golden-fixture scripts and the bench's CPU-baseline leg use it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
HARNESS = os.path.join(REF_DIR, "ref_harness")
PROGRAM = os.path.join(REF_DIR, "chapsim")
MPIRUN = "/opt/conda/bin/mpirun"


def available() -> bool:
    return os.path.exists(HARNESS) and os.path.exists(MPIRUN)


def ini_text(L, M, N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8, REN=2800.0, DT=1e-3,
             NTSTF=1, iIniField_tg=0, CFLGV=1.0e6, TSTOP=1.0e6, dtSave1=1.0e6,
             iterMonitor=1, iRandomType=2, VPER=0.6, SVPER=0.3, TimeReStart_tg=0.0,
             iDomain=1, iThermoDynamics=0, T0=680.15, Ti=680.15, L0=1.5e-3,
             iThermalWallType=(2, 2), thermalWallBC=(640.15, 660.15), iVisScheme=1,
             iGravity=0, iFlowDriven=1, timeThermoStart=1.0e6) -> str:
    """Channel as key lines: iDomain=1 puts the mesh on the turbulence-generator domain
    (NCL1_tg, HX_tg), iDomain=2 on the main (IO) domain (NCL1_io, HX_io).

    Output intervals are set far beyond the run so no instant/averaged files are written.
    """
    f = lambda x: repr(float(x))
    Ltg, Lio = (L, 0) if iDomain == 1 else (0, L)
    HXtg, HXio = (HX, 0.0) if iDomain == 1 else (0.0, HX)
    lines = [
        "[flowtype]", "iCase= 1", f"iDomain= {iDomain}", f"iThermoDynamics= {iThermoDynamics}",
        "iCHT= 0",
        "[geometry]", f"HX_tg= {f(HXtg)}", f"HX_io= {f(HXio)}", f"HZ= {f(HZ)}", "HRb= -1.0",
        "HRt= 1.0",
        "[mesh]", f"NCL1_tg= {Ltg}", f"NCL1_io= {Lio}", f"NCL3= {N}", f"NCL2= {M}",
        f"STR2= {f(STR2)}",
        "[boundary]", "BCX12_tg= 3, 3", "BCX12_io= 3, 3", "BCZ12= 3, 3", "BCY12= 1, 1",
        "[fluid]", f"REN= {f(REN)}", f"ReIni= {f(REN)}", "TLgRe= 0.0", f"iFlowDriven= {iFlowDriven}",
        "Cf_Given= 8.18E-3",
        "[thermohydraulics]", "iFluidMedia= 1", f"iGravity= {iGravity}", f"L0= {f(L0)}", f"T0= {f(T0)}",
        f"Ti= {f(Ti)}", f"iThermalWallType= {iThermalWallType[0]}, {iThermalWallType[1]}",
        f"thermalWallBC_Dim= {f(thermalWallBC[0])}, {f(thermalWallBC[1])}",
        "[conjugateHeatTransfer]", "HY_Solid= 0.1, 0.1", "HXs_Solid= 1.0, 1.0",
        "HXe_Solid= 3.0, 3.0", "NCL2_Solid= 8, 8", "Cp_Solid= 502.4, 502.4",
        "D_Soild= 8000.0, 8000.0", "K_Solid= 16.2, 16.2",
        "[initialisation]", f"iIniField_tg= {iIniField_tg}", "iIniField_io= 0",
        f"iRandomType= {iRandomType}", f"VPER= {f(VPER)}", f"SVPER= {f(SVPER)}",
        f"TimeReStart_tg= {f(TimeReStart_tg)}", "TimeReStart_io= 0.0", "iIniFieldType= 0",
        "iIniFieldTime= 0",
        "[numerics]", f"NTSTF= {NTSTF}", f"TSTOP= {f(TSTOP)}", "timeFlowStart= 0.0",
        f"timeThermoStart= {f(timeThermoStart)}", f"DT= {f(DT)}", "DtMin= 1.0E-9", f"CFLGV= {f(CFLGV)}",
        "[methods]", f"iVisScheme= {iVisScheme}", "iWeightedPre= 0",
        "[statistics]", f"dtSave1= {f(dtSave1)}", f"tRunAve= {f(dtSave1)}",
        "tRunAve_Reset= 0.0", f"dtAveView= {f(dtSave1)}", f"dtRawView= {f(dtSave1)}",
        f"iterMonitor= {iterMonitor}", "MGRID= 4", "JINI= 2",
        "[postprocess]", "iPostProcess= 0", "iPPInst= 0", "iPPSpectra= 0", "iPPQuadrants= 0",
        "iPPDimension= 0", "pp_instn_sz= 0",
    ]
    return "\n".join(lines) + "\n"


def _read_setup(path):
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a.copy()

    L, M, N = (int(v) for v in take(np.int32, 3))
    s = {"L": L, "M": M, "N": N}
    for k, n in (("IPV", L), ("IMV", L), ("KPV", N), ("KMV", N), ("KSYM", N)):
        s[k] = take(np.int32, n)
    s["JGPV"] = take(np.int32, M + 1)   # index 0..M
    s["JGMV"] = take(np.int32, M + 1)   # index 1..M+1
    sc = take(np.float64, 10)
    for i, k in enumerate(("DX", "DZ", "DXI", "DZI", "DXQI", "DZQI", "VOLM", "AREA",
                           "AMVR1", "APVRN")):
        s[k] = float(sc[i])
    rk = take(np.float64, 12)
    s["TGAM"], s["TROH"], s["TALP"] = rk[0:4], rk[4:8], rk[8:12]
    s["YND"] = take(np.float64, M + 1)
    s["YCC"] = take(np.float64, M)
    s["DYFI"] = take(np.float64, M)
    s["DYCI"] = take(np.float64, M + 1)
    s["WFF"] = take(np.float64, M + 1)
    s["WFB"] = take(np.float64, M + 1)
    for k in ("AMVR", "ACVR", "APVR"):
        s[k] = take(np.float64, 3 * (M + 1)).reshape(3, M + 1)   # [idr-1, j-1]
    for k in ("AMPH", "ACPH", "APPH"):
        s[k] = take(np.float64, M)
    s["AK1"] = take(np.float64, L)
    s["AK3"] = take(np.float64, N)
    s["VINI"] = take(np.float64, M)
    return s


def _gather(prefix, nranks, nvars, halo):
    """Reassemble per-rank dumps.  Returns [nvars, N, rows, L] C-order (z, y, x) arrays."""
    parts = []
    for r in range(nranks):
        with open(f"{prefix}.{r}", "rb") as fh:
            hdr = np.frombuffer(fh.read(24), dtype=np.int32)
            data = np.frombuffer(fh.read(), dtype=np.float64)
        myid, j0, n2, L, M, N = (int(v) for v in hdr)
        rows = n2 + 2 if halo else n2
        arr = data.reshape(nvars, N, rows, L)
        parts.append((j0, n2, arr))
    L, M, N = parts[0][2].shape[3], None, parts[0][2].shape[1]
    M = sum(p[1] for p in parts)
    rows = M + 2 if halo else M
    out = np.zeros((nvars, N, rows, L))
    for j0, n2, arr in parts:
        if halo:
            out[:, :, j0:j0 + n2, :] = arr[:, :, 1:n2 + 1, :]
            if j0 == 1:
                out[:, :, 0, :] = arr[:, :, 0, :]
            if j0 + n2 - 1 == M:
                out[:, :, M + 1, :] = arr[:, :, n2 + 1, :]
        else:
            out[:, :, j0 - 1:j0 - 1 + n2, :] = arr
    return out


class Harness:
    """One reference set-up; run() executes a list of harness commands in a fresh process."""

    def __init__(self, L, M, N, nranks=2, keep=False, **ini):
        if not available():
            raise RuntimeError("oracle/_ref/ref_harness is not built (make -f oracle/ref.mk)")
        self.L, self.M, self.N = L, M, N
        self.nranks = nranks
        self.ini = ini_text(L, M, N, **ini)
        self.dir = tempfile.mkdtemp(prefix="chs_ref_")
        self.keep = keep
        with open(os.path.join(self.dir, "readdata.ini"), "w") as fh:
            fh.write(self.ini)

    def close(self):
        if not self.keep:
            shutil.rmtree(self.dir, ignore_errors=True)

    def path(self, name):
        return os.path.join(self.dir, name)

    def write_fields(self, name, u, v, w, p):
        """u,v,w,p as [N, M+2, L] C-order arrays (z, y incl. halo rows 0..M+1, x)."""
        arr = np.stack([u, v, w, p]).astype(np.float64)
        arr.tofile(self.path(name))

    def write_cells(self, name, a):
        np.ascontiguousarray(a, dtype=np.float64).tofile(self.path(name))

    def run(self, commands, timeout=600):
        with open(self.path("harness.cmd"), "w") as fh:
            fh.write("\n".join(commands) + "\n")
        # READINI opens its monitor log with STATUS='new' (READINI.f90:102): start clean.
        shutil.rmtree(self.path("0_log_monitors"), ignore_errors=True)
        cmd = f"ulimit -s unlimited; {MPIRUN} -np {self.nranks} {HARNESS}"
        res = subprocess.run(["bash", "-c", cmd], cwd=self.dir, capture_output=True,
                             text=True, timeout=timeout)
        if res.returncode != 0:
            raise RuntimeError(f"ref_harness failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
        return res

    def setup(self):
        self.run(["dumpsetup setup.bin"])
        return _read_setup(self.path("setup.bin"))

    def read_q(self, name):
        """Returns u, v, w, p as [N, M+2, L] arrays."""
        a = _gather(self.path(name), self.nranks, 4, halo=True)
        return a[0], a[1], a[2], a[3]

    def read_eng(self, name):
        """Energy state written by ``dumpeng``: dict of DH, H, T, D, M, K, CP as [N, M+2, L]
        halo fields and RHS0 as [N, M, L] cells."""
        parts = []
        for r in range(self.nranks):
            with open(self.path(f"{name}.{r}"), "rb") as fh:
                hdr = np.frombuffer(fh.read(24), dtype=np.int32)
                data = np.frombuffer(fh.read(), dtype=np.float64)
            myid, j0, n2, L, M, N = (int(v) for v in hdr)
            nh = N * (n2 + 2) * L
            parts.append((j0, n2, data[:7 * nh].reshape(7, N, n2 + 2, L), data[7 * nh:].reshape(N, n2, L)))
        M = sum(p[1] for p in parts)
        N, L = parts[0][2].shape[1], parts[0][2].shape[3]
        h = np.zeros((7, N, M + 2, L))
        c = np.zeros((N, M, L))
        for j0, n2, hv, cv in parts:
            h[:, :, j0:j0 + n2, :] = hv[:, :, 1:n2 + 1, :]
            if j0 == 1:
                h[:, :, 0, :] = hv[:, :, 0, :]
            if j0 + n2 - 1 == M:
                h[:, :, M + 1, :] = hv[:, :, n2 + 1, :]
            c[:, j0 - 1:j0 - 1 + n2, :] = cv
        out = {k: h[i] for i, k in enumerate(("DH", "H", "T", "D", "M", "K", "CP"))}
        out["RHS0"] = c
        return out

    def write_eng(self, name, dh, g):
        """DH [N, M+2, L] and G [3, N, M+2, L] in the layout ``loadeng`` reads."""
        np.concatenate([np.asarray(dh, dtype=np.float64)[None], np.asarray(g, dtype=np.float64)]).tofile(
            self.path(name))

    def read_table(self, name):
        """Nondimensional property table + spline coefficients written by ``dumptable``."""
        with open(self.path(name), "rb") as fh:
            hdr = np.frombuffer(fh.read(20), dtype=np.int32)
            n = int(hdr[0])
            rest = np.frombuffer(fh.read(), dtype=np.float64)
        t = {"N": n, "IMIN_DH": int(hdr[1]), "IMAX_DH": int(hdr[2]), "IMIN_H": int(hdr[3]),
             "IMAX_H": int(hdr[4])}
        keys = ["DH", "H", "T", "D", "M", "K", "CP", "B"]
        for s in ("DHH", "HT", "HD", "HM", "HK", "HCp", "HB"):
            keys += [f"{s}_B", f"{s}_C", f"{s}_D"]
        for i, k in enumerate(keys):
            t[k] = rest[i * n:(i + 1) * n].copy()
        ref = rest[len(keys) * n:]
        for i, k in enumerate(("T0", "H0", "D0", "M0", "K0", "CP0", "B0")):
            t[k] = float(ref[i])
        return t

    def prop_lookup(self, dh, fin="dh.bin", fout="props.bin", table_file=None):
        """THERM_PROP_UPDATE_FROM_DH of the reference for each DH; returns [7, n]."""
        dh = np.ascontiguousarray(dh, dtype=np.float64)
        with open(self.path(fin), "wb") as fh:
            fh.write(np.array([len(dh)], dtype=np.int32).tobytes())
            fh.write(dh.tobytes())
        return fin, fout

    def read_cells(self, name, nvars=1):
        a = _gather(self.path(name), self.nranks, nvars, halo=False)
        return a if nvars > 1 else a[0]


def _split_dump(path, nranks, groups):
    """Per-rank dumps holding consecutive groups of arrays; groups = [(nvars, halo), ...].
    Returns one [nvars, N, rows, L] array per group, reassembled over the ranks."""
    ranks = []
    for r in range(nranks):
        with open(f"{path}.{r}", "rb") as fh:
            hdr = np.frombuffer(fh.read(24), dtype=np.int32)
            data = np.frombuffer(fh.read(), dtype=np.float64)
        myid, j0, n2, L, M, N = (int(v) for v in hdr)
        off, arrs = 0, []
        for nv, halo in groups:
            rows = n2 + 2 if halo else n2
            n = nv * N * rows * L
            arrs.append(data[off:off + n].reshape(nv, N, rows, L))
            off += n
        ranks.append((j0, n2, arrs))
    Mt = sum(r[1] for r in ranks)
    N, L = ranks[0][2][0].shape[1], ranks[0][2][0].shape[3]
    out = []
    for g, (nv, halo) in enumerate(groups):
        a = np.zeros((nv, N, Mt + 2 if halo else Mt, L))
        for j0, n2, arrs in ranks:
            x = arrs[g]
            if halo:
                a[:, :, j0:j0 + n2, :] = x[:, :, 1:n2 + 1, :]
                if j0 == 1:
                    a[:, :, 0, :] = x[:, :, 0, :]
                if j0 + n2 - 1 == Mt:
                    a[:, :, Mt + 1, :] = x[:, :, n2 + 1, :]
            else:
                a[:, :, j0 - 1:j0 - 1 + n2, :] = x
        out.append(a)
    return out


def write_io(h: "Harness", name, g, g0, pr, d0, m0, explt0):
    """Main-domain momentum state in the layout ``loadio`` reads: G, G0 [3, N, M+2, L], PR, D0,
    M0 [N, M+2, L], EXPLT0 [3, N, M, L] (written with zero halo rows)."""
    g = np.asarray(g, dtype=np.float64)
    N, Mh, L = g.shape[1:]
    ex = np.zeros((3, N, Mh, L))
    ex[:, :, 1:Mh - 1, :] = explt0
    np.concatenate([g, np.asarray(g0, dtype=np.float64), np.asarray(pr, dtype=np.float64)[None],
                    np.asarray(d0, dtype=np.float64)[None], np.asarray(m0, dtype=np.float64)[None],
                    ex]).tofile(h.path(name))


def read_io(h: "Harness", name):
    """State written by ``dumpio``: dict G, Q, DSTG, MSTG [3, N, M+2, L]; PR, DPH [N, M+2, L];
    EXPLT0 [3, N, M, L]; DRHO, RHSP [N, M, L]."""
    g, q, pr, ds, ms, dph, ex, dr, rp = _split_dump(
        h.path(name), h.nranks, [(3, True), (3, True), (1, True), (3, True), (3, True), (1, True),
                                 (3, False), (1, False), (1, False)])
    return dict(G=g, Q=q, PR=pr[0], DSTG=ds, MSTG=ms, DPH=dph[0], EXPLT0=ex, DRHO=dr[0], RHSP=rp[0])


def read_iocn(h: "Harness", name):
    """``iocn`` dump: G, DRHOI [3, N, M+2, L] after the three MOMFA_io calls; RHS [3, N, M, L]
    as each MOMFA_io receives it; EXPLT0 [3, N, M, L]."""
    g, dr, r, ex = _split_dump(h.path(name), h.nranks, [(3, True), (3, True), (3, False), (3, False)])
    return dict(G=g, DRHOI=dr, RHS=r, EXPLT0=ex)


def read_ioterms(h: "Harness", name):
    """``ioterms`` dump: H [3, N, M, L] (x, y, z momentum terms) and DIVU [N, M, L]."""
    a = h.read_cells(name, nvars=4)
    return dict(H=a[:3], DIVU=a[3])


def read_meanzx(h: "Harness", name):
    """``meanzx`` dump -> (stats [M, 122] in or_mean_zx's order, [U1mean, U1maxx])."""
    rows, tail = [], None
    for r in range(h.nranks):
        with open(h.path(f"{name}.{r}"), "rb") as fh:
            hdr = np.frombuffer(fh.read(24), dtype=np.int32)
            data = np.frombuffer(fh.read(), dtype=np.float64)
        myid, j0, n2, L, M, N = (int(v) for v in hdr)
        off = 0

        def take(n):
            nonlocal off
            a = data[off:off + n]
            off += n
            return a

        u1 = take(4 * n2).reshape(4, n2)
        up = take(3 * n2).reshape(3, n2)
        u2 = take(6 * n2).reshape(6, n2)
        u3 = take(10 * n2).reshape(10, n2)
        d1 = take(9 * n2).reshape(3, 3, n2)      # [N][M][J]
        dp = take(9 * n2).reshape(3, 3, n2)
        d2 = take(81 * n2).reshape(9, 9, n2)     # [L2][L1][J]
        tail = take(2)
        blk = np.concatenate([u1, up, u2, u3, d1.transpose(1, 0, 2).reshape(9, n2),
                              dp.transpose(1, 0, 2).reshape(9, n2), d2.transpose(1, 0, 2).reshape(81, n2)])
        rows.append((j0, blk.T))
    rows.sort(key=lambda x: x[0])
    return np.concatenate([b for _, b in rows]), np.array(tail)
