"""Oracle (test infrastructure only): numpy restatement of the reference's set-up routines
for the periodic plane channel on the turbulence-generator (TG) domain.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use anything under
oracle/, and only as the checker.  Index conventions: arrays keep the reference's 1-based
meaning through an explicit offset noted per array (e.g. ``dyci[j]`` is DYCI(J) for
J=0..M+1, entry 0 unused).

Pinned by: tests/test_oracle_setup.py against the reference's own set-up dump
(oracle/_ref/ref_harness ``dumpsetup``, fixture tests/golden/setup_16x16x16.npz) and against
closed forms (tanh mesh end points, telescoping sums, symmetric stencils).
"""
from __future__ import annotations

import numpy as np

# CONS_RKCOEF.f90:21-51 (NSST = 3): gamma, zeta and alpha = gamma + zeta, index 0..3.
TGAM = np.array([1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0])
TROH = np.array([0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0])
TALP = TGAM + TROH


class Mesh:
    """Channel mesh and coefficients, all FP64 like the reference (modules.f90:22, WP=8)."""

    def __init__(self, L, M, N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8, HYB=-1.0, HYT=1.0):
        if L % 2 or N % 2:
            raise ValueError("L and N must be even (FFT99 half-complex packing)")
        self.L, self.M, self.N = L, M, N
        self.HX, self.HZ, self.STR2 = float(HX), float(HZ), float(STR2)

        # INDEXSET.f90:26-31, :59-64 — periodic +/-1 maps (1-based values, 1-based index).
        ic = np.arange(1, L + 1)
        self.IPV = np.where(ic == L, 1, ic + 1).astype(np.int32)
        self.IMV = np.where(ic == 1, L, ic - 1).astype(np.int32)
        kc = np.arange(1, N + 1)
        self.KPV = np.where(kc == N, 1, kc + 1).astype(np.int32)
        self.KMV = np.where(kc == 1, N, kc - 1).astype(np.int32)
        # INDEXSET.f90:68-71 — KSYM(k) = k + N/2 wrapped.
        ks = kc + N // 2
        self.KSYM = np.where(ks > N, ks - N, ks).astype(np.int32)
        # INDEXSET.f90:86-95 (channel branch): JGPV(0..M), JGMV(1..M+1).
        self.JGPV = np.arange(0, M + 1, dtype=np.int32) + 1
        self.JGPV[0] = 1
        self.JGMV = np.arange(1, M + 2, dtype=np.int32) - 1
        self.JGMV[M] = M            # JGMV(NND2) = NCL2

        # CONSPARA.f90:44-73.
        self.DX = self.HX / L
        self.DXI = 1.0 / self.DX
        self.DXQI = self.DXI * self.DXI
        self.DZ = self.HZ / N
        self.DZI = 1.0 / self.DZ
        self.DZQI = self.DZI * self.DZI

        # COORDJR.f90:44-55 — tanh stretching on NND2 = M+1 nodes; YCC at cell centres.
        x2 = np.arange(M + 1, dtype=np.float64) / float(M)
        self.YND = np.tanh(self.STR2 * (x2 - 0.5)) / np.tanh(self.STR2 * 0.5)   # YND(1..M+1)
        self.YCC = (self.YND[:-1] + self.YND[1:]) * 0.5                          # YCC(1..M)
        # COORDJR.f90:121-145.
        self.DYFI = 1.0 / (self.YND[1:] - self.YND[:-1])                         # DYFI(1..M)
        dyci = np.zeros(M + 2)
        dyci[2:M + 1] = 1.0 / ((self.YND[2:] - self.YND[:-2]) * 0.5)
        dyci[1] = 2.0 / (self.YND[1] - self.YND[0])
        dyci[M + 1] = 2.0 / (self.YND[M] - self.YND[M - 1])
        self.DYCI = dyci                                                          # DYCI(0..M+1)
        # COORDJR.f90:425-442 — interpolation weights, 1-based J = 1..M+1 stored at [J].
        wff = np.zeros(M + 2)
        wfb = np.zeros(M + 2)
        for j in range(2, M + 1):
            jm = j - 1
            ycc_j, ycc_m, ynd_j = self.YCC[j - 1], self.YCC[jm - 1], self.YND[j - 1]
            wff[j] = (ynd_j - ycc_m) / (ycc_j - ycc_m)
            wfb[j] = (ycc_j - ynd_j) / (ycc_j - ycc_m)
        wfb[1], wff[1] = 1.0, 0.0
        wfb[M + 1], wff[M + 1] = 0.0, 1.0
        self.WFF, self.WFB = wff, wfb
        # COORDJR.f90:380-388 — VOLM_tg = sum over cells of 1/DYFI (no dx, dz factors).
        self.VOLM = float(L * N) * float(np.sum(1.0 / self.DYFI))
        # LAMPOISLPROF.f90:29-51 — Poiseuille 1.5 (1 - y^2) at cell centres.
        self.VINI = np.where((self.YCC < HYT) & (self.YCC > HYB),
                             (1.0 - self.YCC ** 2) * 1.5, 0.0)

        # TDMA_COEF.f90:44-65 — Poisson y-stencil with Neumann folding, J = 1..M at [J-1].
        d, c = self.DYFI, self.DYCI
        amph = d * c[1:M + 1]
        apph = d * c[2:M + 2]
        amph[0] = d[0] * c[1] * 0.5
        apph[M - 1] = d[M - 1] * c[M + 1] * 0.5
        acph = -(amph + apph)
        acph[0] = amph[0] + acph[0]
        amph[0] = 0.0
        acph[M - 1] = apph[M - 1] + acph[M - 1]
        apph[M - 1] = 0.0
        self.AMPH, self.ACPH, self.APPH = amph, acph, apph
        # TDMA_COEF.f90:67-104 — CN y-stencils [IDR-1, J] for J = 1..M+1 (index 0 unused).
        amvr = np.zeros((3, M + 2))
        acvr = np.zeros((3, M + 2))
        apvr = np.zeros((3, M + 2))
        for idr in (0, 2):
            for j in range(2, M):
                amvr[idr, j] = d[j - 1] * c[j]
                apvr[idr, j] = d[j - 1] * c[j + 1]
                acvr[idr, j] = -(amvr[idr, j] + apvr[idr, j])
            ucaj = 4.0 / (2.0 / c[2] + 1.0 / d[0])
            amvr[idr, 1] = 0.0
            apvr[idr, 1] = ucaj * c[2]
            acvr[idr, 1] = -ucaj * (c[2] + 2.0 * d[0])
            self.AMVR1 = ucaj * (2.0 * d[0])
            ucaj = 4.0 / (2.0 / c[M] + 1.0 / d[M - 1])
            amvr[idr, M] = ucaj * c[M]
            apvr[idr, M] = 0.0
            acvr[idr, M] = -ucaj * (2.0 * d[M - 1] + c[M])
            self.APVRN = ucaj * (2.0 * d[M - 1])
        for j in range(2, M + 1):
            amvr[1, j] = d[j - 2] * c[j]
            apvr[1, j] = d[j - 1] * c[j]
            acvr[1, j] = -(amvr[1, j] + apvr[1, j])
        amvr[1, 1] = apvr[1, 1] = 0.0
        acvr[1, 1] = 1.0
        amvr[1, M + 1] = apvr[1, M + 1] = 0.0
        acvr[1, M + 1] = 1.0
        self.AMVR, self.ACVR, self.APVR = amvr, acvr, apvr

        # possion3d_FFT99.f90:171-199 — modified wavenumbers (2 - 2cos(2 pi k / n)) / d^2.
        pi = 2.0 * np.arcsin(1.0)
        an = np.array([(k - 1) * 2.0 * pi if k <= N // 2 else -(N - k + 1) * 2.0 * pi
                       for k in range(1, N + 1)])
        self.AK3 = 2.0 * (1.0 - np.cos(an / N)) * self.DZQI
        ap = np.array([(i - 1) * 2.0 * pi if i <= L // 2 else -(L - i + 1) * 2.0 * pi
                       for i in range(1, L + 1)])
        self.AK1 = 2.0 * (1.0 - np.cos(ap / L)) * self.DXQI
