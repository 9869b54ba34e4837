"""Thermophysical property tables for the heated channel (table mode).

``synthetic_table`` writes the workload's 2048-point table in the reference's table-file
format (Property_Table_Nondimensionalization, thermo_properties_table_searching.f90:1247-1264:
a header token, the row count, then rows ``P H T D M K CP BETA``, SI units), and
``PropTable`` wraps the GPU lookup of THERM_PROP_UPDATE_FROM_DH (chs_prop_* in
include/chapsim_hip.h) and ``EnergyTG`` the energy substep SOLVERRK3_ENG_io (chs_eng_*).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .tg import _check, _coeff_struct, _dp, _p, lib, ChsError


def synthetic_table(n=2048, T_lo=600.0, T_hi=720.0, T_pc=660.0, cp_base=5.0e3, peak=8.0,
                    width_frac=0.02, rho_lo=700.0, mu_lo=8.0e-5, k_lo=0.5, k_hi=0.3,
                    drop=5.0, P=23.5e6, h_lo=1.5e6):
    """Rows [n, 8] = P, H, T, D, M, K, CP, BETA sorted by T.

    cp = cp_base (1 + (peak - 1) exp(-((T - T_pc) / w)^2)), w = width_frac (T_hi - T_lo);
    h = h_lo + integral of cp (closed form with erf); rho = rho_lo drop^(-(h - h_lo)/dh_tot)
    (5x decrease, steepest where h rises fastest, so rho*h stays monotone for a reference
    state in the upper half of the range); mu = mu_lo drop^(-(T - T_lo)/(T_hi - T_lo));
    k linear from k_lo to k_hi; BETA = -(1/rho) drho/dT.
    """
    T = np.linspace(T_lo, T_hi, n)
    w = width_frac * (T_hi - T_lo)
    a = peak - 1.0
    erf = np.vectorize(math.erf)
    h = h_lo + cp_base * (T - T_lo) + cp_base * a * w * math.sqrt(math.pi) / 2.0 * (
        erf((T - T_pc) / w) - math.erf((T_lo - T_pc) / w))
    cp = cp_base * (1.0 + a * np.exp(-((T - T_pc) / w) ** 2))
    dh_tot = h[-1] - h[0]
    lam = math.log(drop) / dh_tot
    rho = rho_lo * np.exp(-lam * (h - h_lo))
    mu = mu_lo * np.exp(-math.log(drop) * (T - T_lo) / (T_hi - T_lo))
    k = k_lo + (k_hi - k_lo) * (T - T_lo) / (T_hi - T_lo)
    beta = lam * cp
    return np.stack([np.full(n, P), h, T, rho, mu, k, cp, beta], axis=1)


def write_table_file(path, rows):
    """Table file as the reference reads it (list-directed; 17 significant digits so every
    value round-trips exactly)."""
    with open(path, "w") as fh:
        fh.write("#P(Pa),H(J/kg),T(K),D(kg/m3),M(Pa.s),K(W/m.K),CP(J/kg.K),BETA(1/K)\n")
        fh.write(f"{len(rows)}\n")
        for r in rows:
            fh.write(" ".join(f"{v:.17e}" for v in r) + "\n")


def read_table_file(path):
    with open(path) as fh:
        fh.readline()
        n = int(fh.readline().split()[0])
        rows = np.array([[float(x) for x in fh.readline().split()[:8]] for _ in range(n)])
    return rows


SPLINES = ("DHH", "HT", "HD", "HM", "HK", "HCp", "HB")


def table_setup(rows, T0):
    """chs_prop_table_setup: nondimensional table + splines (dict keyed like the reference's
    arrays: DH, H, ..., HT_B, HT_C, ...; IMIN_DH, ...; T0, H0, ...)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    n = rows.shape[0]
    lst = np.zeros(8 * n)
    spl = np.zeros(21 * n)
    idx = (C.c_int * 4)()
    ref = np.zeros(7)
    _check(lib().chs_prop_table_setup(n, _p(rows), float(T0), _p(lst), _p(spl), idx, _p(ref)))
    t = {"N": n, "IMIN_DH": idx[0], "IMAX_DH": idx[1], "IMIN_H": idx[2], "IMAX_H": idx[3]}
    for i, k in enumerate(("DH", "H", "T", "D", "M", "K", "CP", "B")):
        t[k] = lst[i * n:(i + 1) * n].copy()
    for s_i, s in enumerate(SPLINES):
        for c_i, c in enumerate("BCD"):
            o = (3 * s_i + c_i) * n
            t[f"{s}_{c}"] = spl[o:o + n].copy()
    for i, k in enumerate(("T0", "H0", "D0", "M0", "K0", "CP0", "B0")):
        t[k] = float(ref[i])
    return t


class _PropTable(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "imin_dh", "imax_dh", "imin_h", "imax_h")] + \
               [(k, _dp) for k in ("dh", "h", "t", "d", "m", "k", "cp", "b")] + \
               [(f"{s.lower()}_{c}", _dp) for s in SPLINES for c in "bcd"]


class PropTable:
    """A property table resident on the GPU; ``from_dh`` is THERM_PROP_UPDATE_FROM_DH."""

    def __init__(self, t: dict):
        n = int(t["N"])
        self._keep = {}

        def arr(key):
            a = np.ascontiguousarray(t[key], dtype=np.float64)
            if a.shape != (n,):
                raise ChsError(f"table array {key} must have {n} values")
            self._keep[key] = a
            return _p(a)

        fields = [n, int(t["IMIN_DH"]), int(t["IMAX_DH"]), int(t["IMIN_H"]), int(t["IMAX_H"])]
        fields += [arr(k) for k in ("DH", "H", "T", "D", "M", "K", "CP", "B")]
        fields += [arr(f"{s}_{c}") for s in SPLINES for c in "BCD"]
        self._t = _PropTable(*fields)
        h = C.c_void_p()
        rc = lib().chs_prop_create(C.byref(self._t), C.byref(h))
        if rc != 0:
            raise ChsError(f"chs_prop_create failed ({rc}): {lib().chs_prop_last_error(None).decode()}")
        self._h = h

    def from_dh(self, dh):
        """Returns [7, n]: H, T, D, M, K, Cp, B."""
        dh = np.ascontiguousarray(dh, dtype=np.float64).ravel()
        out = np.zeros(7 * len(dh))
        rc = lib().chs_prop_update_from_dh(self._h, len(dh), _p(dh), _p(out))
        if rc != 0:
            raise ChsError(f"chs_prop_update_from_dh failed ({rc}): "
                           f"{lib().chs_prop_last_error(self._h).decode()}")
        return out.reshape(7, len(dh))

    def close(self):
        if getattr(self, "_h", None):
            lib().chs_prop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _EngParams(C.Structure):
    _fields_ = [("wall_type", C.c_int * 2), ("cthecd", C.c_double), ("dh_min", C.c_double),
                ("dh_max", C.c_double), ("wallflux", _dp), ("wall_state", (C.c_double * 5) * 2),
                ("top_face", C.c_int)]


ENG_FIELDS = ("DH", "H", "T", "D", "M", "K", "CP")


class EnergyTG:
    """Thermal state of one heated periodic channel on the GPU; ``SOLVERRK3_ENG_io(ns, dt)`` is
    the reference's energy substep (SOLVERRK3_ENG_IO.f90:11-82) with table properties.

    coeffs: ``channel_setup`` dict; t: nondimensional table (``table_setup``); wall_type
    (bottom, top): 1 fixed heat flux, 2 fixed temperature (modules.f90:526-527); cthecd =
    1/(REN PRT0); dhmin/dhmax: DHmin/DHmax; wallflux: (bottom, top) scalars or a [2, N, L]
    array; wall_state: [2][T, H, D, M, K] of fixed-temperature walls; top_face: reference J
    of the top wall's ghost-row weights (M on one rank, the last rank's N2DO on several).
    Fields are C-order [N][M+2][L]."""

    def __init__(self, coeffs: dict, t: dict, wall_type, cthecd, dhmin, dhmax, wallflux,
                 wall_state=None, top_face=None):
        L, M, N = int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"])
        self.L, self.M, self.N = L, M, N
        self._keep = []
        self._c = _coeff_struct(coeffs, self._keep)
        self._tab = PropTable(t)
        wf = np.asarray(wallflux, dtype=np.float64)
        if wf.shape == (2,):
            wf = np.stack([np.full((N, L), wf[0]), np.full((N, L), wf[1])])
        wf = np.ascontiguousarray(wf.reshape(2, N, L))
        self._keep.append(wf)
        ws = ((C.c_double * 5) * 2)()
        if wall_state is not None:
            for w in range(2):
                for j in range(5):
                    ws[w][j] = float(wall_state[w][j])
        self._par = _EngParams((C.c_int * 2)(*[int(x) for x in wall_type]), float(cthecd), float(dhmin),
                               float(dhmax), _p(wf), ws, 0 if top_face is None else int(top_face))
        h = C.c_void_p()
        rc = lib().chs_eng_create(C.byref(self._c), C.byref(self._tab._t), C.byref(self._par), C.byref(h))
        if rc != 0:
            raise ChsError(f"chs_eng_create failed ({rc}): {lib().chs_eng_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            raise ChsError(f"{what} failed ({rc}): {lib().chs_eng_last_error(self._h).decode()}")

    def upload(self, dh, g, rhs0=None, init_props=True):
        """DH [N, M+2, L], G [3, N, M+2, L], RHS_ENERGY0 [N, M, L] (None: zero); init_props runs
        THERM_PROP_UPDATE on every row and BC_WALL_THERMAL, as the start-up does."""
        hs = (self.N, self.M + 2, self.L)
        dh = np.ascontiguousarray(dh, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        if dh.shape != hs or g.shape != (3,) + hs:
            raise ChsError(f"upload: DH must be {hs} and G (3,) + {hs}")
        r0 = None
        if rhs0 is not None:
            r0 = np.ascontiguousarray(rhs0, dtype=np.float64)
            if r0.shape != (self.N, self.M, self.L):
                raise ChsError("upload: RHS_ENERGY0 must be (N, M, L)")
        self._check(lib().chs_eng_upload(self._h, _p(dh), _p(g), None if r0 is None else _p(r0),
                                         1 if init_props else 0), "chs_eng_upload")

    def set_mass_flux(self, g_dev_ptr: int):
        self._check(lib().chs_eng_set_mass_flux(self._h, C.c_void_p(g_dev_ptr)), "chs_eng_set_mass_flux")

    def set_tuning(self, **kv):
        """Launch shapes (chs_eng_set_tuning), e.g. rhs_lds=0."""
        for k, v in kv.items():
            self._check(lib().chs_eng_set_tuning(self._h, k.encode(), int(v)), "chs_eng_set_tuning")

    def SOLVERRK3_ENG_io(self, ns, dt):
        self._check(lib().chs_substep_eng(self._h, int(ns), float(dt)), "chs_substep_eng")

    def download(self):
        """dict of DH, H, T, D, M, K, CP ([N, M+2, L]) and RHS0 ([N, M, L])."""
        hs = (self.N, self.M + 2, self.L)
        out = {k: np.zeros(hs) for k in ENG_FIELDS}
        out["RHS0"] = np.zeros((self.N, self.M, self.L))
        self._check(lib().chs_eng_download(self._h, *[_p(out[k]) for k in ENG_FIELDS + ("RHS0",)]),
                    "chs_eng_download")
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().chs_eng_destroy(self._h)
            self._h = None
        if getattr(self, "_tab", None) is not None:
            self._tab.close()
            self._tab = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
