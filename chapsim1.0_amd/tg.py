"""Host mirror of the reference's TG-domain momentum substep interface over the C-ABI of
libchapsim_hip.so (include/chapsim_hip.h).

The reference calls ``SOLVERRK3_MOM_tg(NS)`` with everything else in module globals
(SOLVERRK3_MOM_tg.f90:13, SOLVE.f90:114-134).  ``ChannelTG`` holds the same state on the
GPU and exposes the same steps: ``SOLVERRK3_MOM_tg(ns, dt, cvisc)`` for one substep
(ns = 0 is the projection of CALC_INITIALIZATION_tg), ``rk3_step`` for the stage loop of
SOLVE, ``cfl`` / ``max_div`` for CFL_tg / DIVGCK_tg.  Errors are raised as ``ChsError``
(the reference STOPs through ERRHDL, WRTHDL.f90:48-62).

Arrays use C-order numpy views of the reference's Fortran layout:
``q[c, k, J, i]`` = Q_tg(i+1, J, k+1, c+1) with J = 0..M+1, ``pr[k, J, i]``,
``convh0[c, k, J-1, i]``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchapsim_hip.so")

PLAIN_FP64 = 0
MIXED = 1

_dp = C.POINTER(C.c_double)
_lib = None

# Every entry point declared in include/chapsim_hip.h: name -> (restype, argtypes).
SIGNATURES = {
    "chs_version": (C.c_char_p, []),
    "chs_last_error": (C.c_char_p, [C.c_void_p]),
    "chs_channel_setup": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
                          + [_dp] * 16),
    "chs_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "chs_destroy": (None, [C.c_void_p]),
    "chs_upload_fields": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "chs_download_fields": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "chs_substep_tg": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
    "chs_step_tg": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "chs_run_tg": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "chs_set_step_state": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "chs_get_step_state": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_int)]),
    "chs_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "chs_set_refinement": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "chs_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "chs_tdma_lines": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "chs_tdma_last_error": (C.c_char_p, []),
    "chs_tdma_lines_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
    "chs_dev_malloc": (C.c_void_p, [C.c_size_t]),
    "chs_dev_free": (C.c_int, [C.c_void_p]),
    "chs_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "chs_poisson": (C.c_int, [C.c_void_p, _dp, _dp]),
    "chs_poisson_fishpack": (C.c_int, [C.c_void_p, _dp, _dp]),
    "chs_poisson_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp]),
    "chs_cfl_tg": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "chs_max_div": (C.c_int, [C.c_void_p, _dp]),
    "chs_sync": (C.c_int, [C.c_void_p]),
    "chs_device_fields": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp)]),
    "chs_profile_reset": (C.c_int, [C.c_void_p, C.c_char_p]),
    "chs_profile_get": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.POINTER(C.c_int)]),
    "chs_ens_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "chs_ens_destroy": (None, [C.c_void_p]),
    "chs_ens_channel": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "chs_ens_step": (C.c_int, [C.c_void_p, _dp, _dp]),
    "chs_ens_sync": (C.c_int, [C.c_void_p]),
    "chs_ens_set_batched": (C.c_int, [C.c_void_p, C.c_int]),
    "chs_prop_table_setup": (C.c_int, [C.c_int, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_int), _dp]),
    "chs_prop_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "chs_prop_destroy": (None, [C.c_void_p]),
    "chs_prop_last_error": (C.c_char_p, [C.c_void_p]),
    "chs_prop_update_from_dh": (C.c_int, [C.c_void_p, C.c_long, _dp, _dp]),
    "chs_eng_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "chs_eng_destroy": (None, [C.c_void_p]),
    "chs_eng_last_error": (C.c_char_p, [C.c_void_p]),
    "chs_eng_upload": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int]),
    "chs_eng_set_mass_flux": (C.c_int, [C.c_void_p, C.c_void_p]),
    "chs_eng_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "chs_substep_eng": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "chs_eng_download": (C.c_int, [C.c_void_p] + [_dp] * 8),
    "chs_mean_zx_tg": (C.c_int, [C.c_void_p, _dp, _dp]),
    "chs_restart_write_tg": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, _dp, _dp]),
    "chs_restart_read_tg": (C.c_int, [C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                      _dp, _dp, _dp]),
    "chs_restart_last_error": (C.c_char_p, []),
    "chs_interp_initial_tg": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                        C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "chs_random_field_tg": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_double, C.c_double, C.c_ulonglong, C.c_double, _dp, _dp]),
    "chs_io_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "chs_io_destroy": (None, [C.c_void_p]),
    "chs_io_last_error": (C.c_char_p, [C.c_void_p]),
    "chs_io_upload": (C.c_int, [C.c_void_p] + [_dp] * 6),
    "chs_io_set_properties": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "chs_io_props": (C.c_int, [C.c_void_p, C.c_int]),
    "chs_substep_io": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "chs_substep_io_momentum": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "chs_io_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "chs_io_set_mixed": (C.c_int, [C.c_void_p, C.c_int]),
    "chs_io_set_poisson": (C.c_int, [C.c_void_p, C.c_int]),
    "chs_io_upload_viscosity": (C.c_int, [C.c_void_p, _dp]),
    "chs_cfl_io": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double)]),
    "chs_max_div_io": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "chs_restart_write_io": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double] + [_dp] * 7),
    "chs_restart_read_io": (C.c_int, [C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), _dp] + [_dp] * 6),
    "chs_heated_cfl": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)]),
    "chs_io_download": (C.c_int, [C.c_void_p] + [_dp] * 9),
    "chs_io_sync": (C.c_int, [C.c_void_p]),
    "chs_heated_substep": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double]),
}


class ChsError(RuntimeError):
    pass


class _Coeffs(C.Structure):
    _fields_ = [("L", C.c_int), ("M", C.c_int), ("N", C.c_int),
                ("dxi", C.c_double), ("dzi", C.c_double), ("dxqi", C.c_double),
                ("dzqi", C.c_double), ("volm", C.c_double)] + \
               [(n, _dp) for n in ("dyfi", "dyci", "wff", "wfb", "amvr", "acvr", "apvr",
                                   "amph", "acph", "apph", "ak1", "ak3", "tgam", "troh", "talp")]


def lib():
    """Load libchapsim_hip.so; raises if it is not built (there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChsError(f"{LIB_PATH} is missing: build it (make -f chapsim1.0_amd/Makefile)")
        lb = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lb, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lb
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().chs_last_error(ctx)
        raise ChsError(f"chapsim_hip error {rc}: {msg.decode() if msg else ''}")


def channel_setup(L, M, N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8):
    """Mesh and coefficients of the reference's set-up, computed by chs_channel_setup."""
    scal = np.zeros(5)
    out = dict(dyfi=np.zeros(M), dyci=np.zeros(M + 1), wff=np.zeros(M + 1), wfb=np.zeros(M + 1),
               amvr=np.zeros(3 * (M + 1)), acvr=np.zeros(3 * (M + 1)), apvr=np.zeros(3 * (M + 1)),
               amph=np.zeros(M), acph=np.zeros(M), apph=np.zeros(M), ak1=np.zeros(L),
               ak3=np.zeros(N), tgam=np.zeros(4), troh=np.zeros(4), talp=np.zeros(4))
    keys = ("dyfi", "dyci", "wff", "wfb", "amvr", "acvr", "apvr", "amph", "acph", "apph",
            "ak1", "ak3", "tgam", "troh", "talp")
    _check(lib().chs_channel_setup(L, M, N, HX, HZ, STR2, _p(scal), *[_p(out[k]) for k in keys]))
    out.update(L=L, M=M, N=N, dxi=scal[0], dzi=scal[1], dxqi=scal[2], dzqi=scal[3], volm=scal[4])
    return out


def _coeff_struct(coeffs: dict, keep: list):
    arrays = {k: np.ascontiguousarray(v, dtype=np.float64)
              for k, v in coeffs.items() if isinstance(v, np.ndarray)}
    keep.append(arrays)
    return _Coeffs(int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"]), coeffs["dxi"],
                   coeffs["dzi"], coeffs["dxqi"], coeffs["dzqi"], coeffs["volm"],
                   *[_p(arrays[k]) for k in ("dyfi", "dyci", "wff", "wfb", "amvr", "acvr",
                                             "apvr", "amph", "acph", "apph", "ak1", "ak3",
                                             "tgam", "troh", "talp")])


class ChannelTG:
    """One periodic channel resident on the GPU (TG domain, flow only)."""

    def __init__(self, coeffs: dict, mode: int = PLAIN_FP64, handle=None, owner=None):
        self.L, self.M, self.N = int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"])
        self.mode = mode
        self._dyfi = np.array(coeffs["dyfi"], dtype=np.float64)
        self._keep = []
        self._owner = owner
        if handle is not None:          # a channel of an Ensemble (destroyed by it)
            self._h = handle
            return
        self._c = _coeff_struct(coeffs, self._keep)
        h = C.c_void_p()
        _check(lib().chs_create(C.byref(self._c), mode, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._owner is None:
            lib().chs_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -----------------------------------------------------------------------
    def upload(self, q, pr, convh0=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        L, M, N = self.L, self.M, self.N
        if q.shape != (3, N, M + 2, L) or pr.shape != (N, M + 2, L):
            raise ChsError("field shapes must be q[3,N,M+2,L], pr[N,M+2,L]")
        h0 = None
        if convh0 is not None:
            h0 = np.ascontiguousarray(convh0, dtype=np.float64)
            if h0.shape != (3, N, M, L):
                raise ChsError("convh0 shape must be [3,N,M,L]")
        _check(lib().chs_upload_fields(self._h, _p(q), _p(pr), _p(h0) if h0 is not None else None),
               self._h)

    def download(self, with_history=False):
        L, M, N = self.L, self.M, self.N
        q = np.zeros((3, N, M + 2, L))
        pr = np.zeros((N, M + 2, L))
        h0 = np.zeros((3, N, M, L)) if with_history else None
        _check(lib().chs_download_fields(self._h, _p(q), _p(pr), _p(h0) if with_history else None),
               self._h)
        return (q, pr, h0) if with_history else (q, pr)

    def PP_MEAN_ZX_FLOW_TG(self):
        """x-z mean statistics (chs_mean_zx_tg): stats [M, 122] and [U1mean, U1maxx]."""
        st = np.zeros((self.M, 122))
        sc = np.zeros(2)
        _check(lib().chs_mean_zx_tg(self._h, _p(st), _p(sc)), self._h)
        return st, sc

    def monitor(self, REN):
        """Instantaneous monitor quantities of PP_MONITOR_TG (processing/PP_SCN_MONITOR.f90:
        147-253) from the x-z means: Cf at the bottom and top walls, mean u, v, w and the
        normal stresses uu, uv, vv, ww at the centre row NCL2 / 2, U1mean, U1maxx."""
        st, sc = self.PP_MEAN_ZX_FLOW_TG()
        M = self.M
        dyfi = self._dyfi
        cf_b = (st[0, 0] - 0.0) * dyfi[0] * 2.0 * 2.0 / REN
        cf_t = (st[M - 1, 0] - 0.0) * dyfi[M - 1] * 2.0 * 2.0 / REN
        j = M // 2 - 1
        u1 = st[j, 0:3]
        uu = [st[j, 7] - u1[0] * u1[0], st[j, 8] - u1[0] * u1[1], st[j, 10] - u1[1] * u1[1],
              st[j, 12] - u1[2] * u1[2]]
        return dict(cf_bottom=cf_b, cf_top=cf_t, u_centre=u1.tolist(), uu_uv_vv_ww=uu, u1mean=sc[0],
                    u1maxx=sc[1])

    # -- the reference's call structure ----------------------------------------------
    def SOLVERRK3_MOM_tg(self, ns: int, dt: float, cvisc: float):
        _check(lib().chs_substep_tg(self._h, int(ns), float(dt), float(cvisc)), self._h)

    def rk3_step(self, dt: float, cvisc: float):
        _check(lib().chs_step_tg(self._h, float(dt), float(cvisc)), self._h)

    def run(self, nsteps: int, dt0: float, cflgv: float, dtmin: float, cvisc: float):
        """nsteps steps of SOLVE's loop with DT chosen on the device by BCAST_COMM_STEP's rule
        (chs_run_tg); set_step_state first."""
        _check(lib().chs_run_tg(self._h, int(nsteps), float(dt0), float(cflgv), float(dtmin), float(cvisc)),
               self._h)

    def set_step_state(self, dt: float, phytime: float = 0.0):
        _check(lib().chs_set_step_state(self._h, float(dt), float(phytime)), self._h)

    def step_state(self) -> dict:
        """DT, PhyTIME, CFLMM (NaN until formed) and the DT / DT0 < 0.01 stop flag."""
        dt, t, cf, st = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        _check(lib().chs_get_step_state(self._h, C.byref(dt), C.byref(t), C.byref(cf), C.byref(st)), self._h)
        return dict(dt=dt.value, phytime=t.value, cflmm=cf.value, stopped=bool(st.value))

    def set_graph(self, enable: bool):
        """Replay rk3_step from a hipGraph captured per (dt, cvisc)."""
        _check(lib().chs_set_graph(self._h, 1 if enable else 0), self._h)
        self.graph = bool(enable)

    def set_refinement(self, nref_step: int, nref_projection: int):
        _check(lib().chs_set_refinement(self._h, int(nref_step), int(nref_projection)), self._h)

    def set_tuning(self, **shapes):
        """Launch shapes, e.g. set_tuning(conv_by=8, conv_kc=16) (keys of chs_set_tuning)."""
        for k, v in shapes.items():
            _check(lib().chs_set_tuning(self._h, k.encode(), int(v)), self._h)

    def poisson(self, rhs):
        L, M, N = self.L, self.M, self.N
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if rhs.shape != (N, M, L):
            raise ChsError("rhs shape must be [N,M,L]")
        phi = np.zeros((N, M, L))
        _check(lib().chs_poisson(self._h, _p(rhs), _p(phi)), self._h)
        return phi

    def FISHPACK_POIS3D_SIMPLE(self, rhs):
        """Main-domain pressure solve with TRID0's gauge (chs_poisson_fishpack); rhs and the
        result are [N,M,L] cell arrays."""
        L, M, N = self.L, self.M, self.N
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if rhs.shape != (N, M, L):
            raise ChsError("rhs shape must be [N,M,L]")
        phi = np.zeros((N, M, L))
        _check(lib().chs_poisson_fishpack(self._h, _p(rhs), _p(phi)), self._h)
        return phi

    def poisson_device(self, rhs_ptr: int, phi_ptr: int, nref: int = 0, want_residual=True):
        """Solve on device arrays (cells, [N,M,L] float64) given as raw device pointers;
        returns max |rhs - L phi| (FP64) if want_residual."""
        res = C.c_double(-1.0)
        _check(lib().chs_poisson_device(self._h, C.c_void_p(rhs_ptr), C.c_void_p(phi_ptr), int(nref),
                                        C.byref(res) if want_residual else None), self._h)
        return res.value

    def cfl(self, dt: float) -> float:
        v = C.c_double()
        _check(lib().chs_cfl_tg(self._h, float(dt), C.byref(v)), self._h)
        return v.value

    def max_div(self) -> float:
        v = C.c_double()
        _check(lib().chs_max_div(self._h, C.byref(v)), self._h)
        return v.value

    def sync(self):
        _check(lib().chs_sync(self._h), self._h)

    def profile_reset(self, only: str | None = None):
        _check(lib().chs_profile_reset(self._h, only.encode() if only else None), self._h)

    def profile_get(self, kernel: str):
        avg = C.c_double()
        n = C.c_int()
        _check(lib().chs_profile_get(self._h, kernel.encode(), C.byref(avg), C.byref(n)), self._h)
        return avg.value, n.value


class Ensemble:
    """Independent channels of uneven size advanced together: one hipGraph replays one RK3
    step of every channel (chs_ens_step)."""

    def __init__(self, coeff_list, mode: int = PLAIN_FP64):
        self._keep = []
        structs = [_coeff_struct(c, self._keep) for c in coeff_list]
        self._structs = structs
        ptrs = (C.c_void_p * len(structs))(*[C.cast(C.pointer(s), C.c_void_p) for s in structs])
        h = C.c_void_p()
        _check(lib().chs_ens_create(len(structs), ptrs, mode, C.byref(h)))
        self._h = h
        self.channels = []
        for i, co in enumerate(coeff_list):
            ch = C.c_void_p()
            _check(lib().chs_ens_channel(self._h, i, C.byref(ch)))
            self.channels.append(ChannelTG(co, mode, handle=ch, owner=self))

    def step(self, dt, cvisc):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        cv = np.ascontiguousarray(cvisc, dtype=np.float64)
        _check(lib().chs_ens_step(self._h, _p(dt), _p(cv)))

    def sync(self):
        _check(lib().chs_ens_sync(self._h))

    def set_batched(self, batched: bool):
        """One launch per kernel kind over all channels (True) or each channel's own launches
        on parallel capture lanes (False, the default)."""
        _check(lib().chs_ens_set_batched(self._h, 1 if batched else 0))

    def close(self):
        if getattr(self, "_h", None):
            for ch in self.channels:
                ch._h = None
            lib().chs_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float64 device buffer from chs_dev_malloc (for device-pointer entry points)."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = lib().chs_dev_malloc(self.nbytes)
        if not self.ptr:
            raise ChsError("chs_dev_malloc failed")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        _check(lib().chs_memcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes, 1))
        return d

    def to_host(self):
        a = np.empty(self.shape)
        _check(lib().chs_memcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 0))
        return a

    def free(self):
        if self.ptr:
            lib().chs_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


TDMA_KINDS = {"IJI_CYC": 0, "IJJ_nonCYC": 1, "IJI_nonCYC": 2, "IJJ_CYC": 3}


def tdma_lines(kind, A, B, C_, R, bc=None, noalloc=False):
    """The reference's batched Thomas solvers (chs_tdma_lines) on host arrays given as C-order
    [jsz, isz] views of the Fortran A, B, C, R(isz, jsz); bc as [2, nlines] for the
    non-cyclic kinds.  kind: 0/"IJI_CYC" (TDMAIJI_CYC), 1/"IJJ_nonCYC", 2/"IJI_nonCYC",
    3/"IJJ_CYC".  Returns the solution; the arrays go through device memory.  noalloc: the
    entry point without allocation (chs_tdma_lines_noalloc)."""
    k = TDMA_KINDS[kind] if isinstance(kind, str) else int(kind)
    R = np.ascontiguousarray(R, dtype=np.float64)
    jsz, isz = R.shape
    dev = [DeviceArray.from_host(x) for x in (A, B, C_, R)]
    dbc = DeviceArray.from_host(bc) if k in (1, 2) else None
    try:
        fn = lib().chs_tdma_lines_noalloc if noalloc else lib().chs_tdma_lines
        rc = fn(k, isz, jsz, *[C.c_void_p(d.ptr) for d in dev], C.c_void_p(dbc.ptr) if dbc else None)
        if rc != 0:
            raise ChsError(f"chs_tdma_lines failed ({rc}): {lib().chs_tdma_last_error().decode()}")
        return dev[3].to_host()
    finally:
        for d in dev + ([dbc] if dbc else []):
            d.free()


class _IoParams(C.Structure):
    _fields_ = [("cvisc", C.c_double), ("gravity", C.c_int), ("f_a", C.c_double),
                ("flow_driven", C.c_int), ("cf_given", C.c_double), ("vis_scheme", C.c_int)]


class ChannelIO:
    """The main-domain (IO) flow of one periodic heated channel on the GPU, variable density
    and viscosity: ``ioprops(ns)`` is what SOLVE runs before the momentum substep
    (SOLVE.f90:122-127), ``SOLVERRK3_MOM_io(ns, dt)`` the substep (SOLVERRK3_MOM_IO.f90:11-120).
    gravity = |iGravity|, f_a = F_A, flow_driven = iFlowDriven, vis_scheme = iVisScheme (0
    explicit, 1 Crank-Nicolson).  Fields are C-order: G, G0 [3, N, M+2, L], PR, D0, M0, M1
    [N, M+2, L], EXPLT0 [3, N, M, L]."""

    def __init__(self, coeffs: dict, REN: float, gravity=0, f_a=0.0, flow_driven=1, cf_given=0.0, vis_scheme=0):
        self.L, self.M, self.N = int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"])
        self._keep = []
        self._c = _coeff_struct(coeffs, self._keep)
        self._par = _IoParams(1.0 / float(REN), int(gravity), float(f_a), int(flow_driven), float(cf_given),
                              int(vis_scheme))
        h = C.c_void_p()
        rc = lib().chs_io_create(C.byref(self._c), C.byref(self._par), C.byref(h))
        if rc != 0:
            raise ChsError(f"chs_io_create failed ({rc}): {lib().chs_io_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            raise ChsError(f"{what} failed ({rc}): {lib().chs_io_last_error(self._h).decode()}")

    def upload(self, G, PR, D0, M0, G0=None, EXPLT0=None, M1=None):
        """Host state; M1 (Crank-Nicolson only) is the current viscosity, default M0."""
        L, M, N = self.L, self.M, self.N
        hs = (N, M + 2, L)
        arrs = {}
        for k, a, shp in (("G", G, (3,) + hs), ("G0", G0, (3,) + hs), ("PR", PR, hs), ("D0", D0, hs),
                          ("M0", M0, hs), ("EXPLT0", EXPLT0, (3, N, M, L))):
            if a is None:
                arrs[k] = None
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != shp:
                raise ChsError(f"upload: {k} must have shape {shp}")
            arrs[k] = a
        P = lambda a: None if a is None else _p(a)
        self._check(lib().chs_io_upload(self._h, P(arrs["G"]), P(arrs["G0"]), P(arrs["PR"]), P(arrs["D0"]),
                                        P(arrs["M0"]), P(arrs["EXPLT0"])), "chs_io_upload")
        if M1 is not None:
            m1 = np.ascontiguousarray(M1, dtype=np.float64)
            if m1.shape != hs:
                raise ChsError(f"upload: M1 must have shape {hs}")
            self._check(lib().chs_io_upload_viscosity(self._h, _p(m1)), "chs_io_upload_viscosity")

    def set_properties(self, d0_dev: int, m0_dev: int):
        self._check(lib().chs_io_set_properties(self._h, C.c_void_p(d0_dev), C.c_void_p(m0_dev)),
                    "chs_io_set_properties")

    def ioprops(self, ns):
        self._check(lib().chs_io_props(self._h, int(ns)), "chs_io_props")

    def SOLVERRK3_MOM_io(self, ns, dt):
        self._check(lib().chs_substep_io(self._h, int(ns), float(dt)), "chs_substep_io")

    def momentum(self, ns, dt):
        """The momentum part of the substep alone: G before the projection."""
        self._check(lib().chs_substep_io_momentum(self._h, int(ns), float(dt)), "chs_substep_io_momentum")

    def set_poisson(self, fishpack: bool):
        """Pressure solver (chs_io_set_poisson): FFT99's gauge (False) or FISHPACK's (True)."""
        self._check(lib().chs_io_set_poisson(self._h, int(bool(fishpack))), "chs_io_set_poisson")

    def set_mixed(self, mixed: bool):
        """Mixed form (chs_io_set_mixed): EXPLT0 and the momentum increment in FP32."""
        self._check(lib().chs_io_set_mixed(self._h, int(bool(mixed))), "chs_io_set_mixed")

    def set_tuning(self, **kv):
        """Launch shapes (chs_io_set_tuning), e.g. mom_rows=2."""
        for k, v in kv.items():
            self._check(lib().chs_io_set_tuning(self._h, k.encode(), int(v)), "chs_io_set_tuning")

    def sync(self):
        self._check(lib().chs_io_sync(self._h), "chs_io_sync")

    def max_div(self) -> float:
        """DIVGCK_io (DIVGCK.f90:60-175): max |div G + DrhoDtP|."""
        v = C.c_double()
        self._check(lib().chs_max_div_io(self._h, C.byref(v)), "chs_max_div_io")
        return v.value

    def cfl(self, dt: float) -> float:
        """CFL_io (CFL.f90:59-107) x dt on the current velocity, with the context's DENSITY0 and
        viscosity (a flow-only run's DENSITY, Viscousity); cfl(1.0) is CFLMM_io."""
        v = C.c_double()
        self._check(lib().chs_cfl_io(self._h, float(dt), C.byref(v)), "chs_cfl_io")
        return v.value

    def download(self):
        """dict of G, Q, DSTG, MSTG [3, N, M+2, L]; PR, DPH [N, M+2, L]; EXPLT0 [3, N, M, L];
        DRHO, RHSP [N, M, L]."""
        L, M, N = self.L, self.M, self.N
        hs = (N, M + 2, L)
        o = dict(G=np.zeros((3,) + hs), Q=np.zeros((3,) + hs), PR=np.zeros(hs), DPH=np.zeros(hs),
                 EXPLT0=np.zeros((3, N, M, L)), DRHO=np.zeros((N, M, L)), RHSP=np.zeros((N, M, L)),
                 DSTG=np.zeros((3,) + hs), MSTG=np.zeros((3,) + hs))
        self._check(lib().chs_io_download(self._h, *[_p(o[k]) for k in ("G", "Q", "PR", "DPH", "EXPLT0", "DRHO",
                                                                         "RHSP", "DSTG", "MSTG")]),
                    "chs_io_download")
        return o

    def close(self):
        if getattr(self, "_h", None):
            lib().chs_io_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def heated_substep(eng, io: ChannelIO, ns: int, dt: float):
    """One stage of the heated channel as SOLVE runs it (SOLVE.f90:119-133): energy substep,
    staggered properties, SOLVERRK3_MOM_io (chs_heated_substep)."""
    rc = lib().chs_heated_substep(eng._h, io._h, int(ns), float(dt))
    if rc != 0:
        raise ChsError(f"chs_heated_substep failed ({rc}): {lib().chs_io_last_error(io._h).decode()}")


def heated_cfl(eng, io: ChannelIO, dt: float) -> float:
    """CFL_io x dt of a heated run: the momentum context's velocity with the energy context's
    current DENSITY and Viscousity (chs_heated_cfl); heated_cfl(eng, io, 1.0) is CFLMM_io."""
    v = C.c_double()
    rc = lib().chs_heated_cfl(eng._h, io._h, float(dt), C.byref(v))
    if rc != 0:
        raise ChsError(f"chs_heated_cfl failed ({rc}): {lib().chs_io_last_error(io._h).decode()}")
    return v.value


def next_dt(dt: float, dt0: float, cflgv: float, cflmm: float, dtmin: float) -> float:
    """The time-step rule of BCAST_COMM_STEP (SOLVE.f90:169-178): the given DT0, at most five
    times the last step and CFLGV / CFLMM (CFLMM = max of CFLMM_tg and CFLMM_io, SOLVE.f90:79-85,
    e.g. ChannelTG.cfl(1.0), ChannelIO.cfl(1.0)), not below DTMIN."""
    return max(min(dt0, 5.0 * dt, cflgv / cflmm), dtmin)


def restart_write_tg(directory, q, pr, iterg, phytime, REN, dt):
    """WRT_INSTANT_VARS_TG file set (chs_restart_write_tg) from halo arrays q [3, N, M+2, L], pr."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    _, N, Mh, L = q.shape
    rc = lib().chs_restart_write_tg(str(directory).encode(), L, Mh - 2, N, int(iterg), float(phytime), float(REN),
                                    float(dt), _p(q), _p(pr))
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())


def restart_read_tg(directory, tfm, L, M, N):
    """ReStart_INSTANT_VARS_TG(TFM) (chs_restart_read_tg): returns q, pr, ITERG, PhyTIME."""
    q = np.zeros((3, N, M + 2, L))
    pr = np.zeros((N, M + 2, L))
    it = C.c_int(0)
    t = np.zeros(1)
    rc = lib().chs_restart_read_tg(str(directory).encode(), float(tfm), L, M, N, C.byref(it), _p(t), _p(q), _p(pr))
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())
    return q, pr, it.value, float(t[0])


def restart_write_io(directory, q, pr, iterg, phytime, REN, dt, thermal=None):
    """WRT_INSTANT_VARS_IO file set (chs_restart_write_io) from halo arrays q [3, N, M+2, L]
    (Q_io), pr and, for a thermal run, thermal = dict T, D, H, M, K [N, M+2, L]."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    _, N, Mh, L = q.shape
    th = [None if thermal is None or thermal.get(k) is None else np.ascontiguousarray(thermal[k], dtype=np.float64)
          for k in "TDHMK"]
    rc = lib().chs_restart_write_io(str(directory).encode(), L, Mh - 2, N, int(iterg), float(phytime), float(REN),
                                    float(dt), _p(q), _p(pr), *[None if a is None else _p(a) for a in th])
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())


def restart_read_io(directory, tfm, L, M, N, ini_field_type=0, thermal=True):
    """ReStart_INSTANT_VARS_io(TFM), file part (chs_restart_read_io): dict FLOW [3, N, M+2, L]
    (G_io for ini_field_type 0, Q_io for 1), PR, ITERG, PhyTIME and, with thermal, T, D, H, DH."""
    hs = (N, M + 2, L)
    o = dict(FLOW=np.zeros((3,) + hs), PR=np.zeros(hs))
    if thermal:
        o.update({k: np.zeros(hs) for k in ("T", "D", "H", "DH")})
    it = C.c_int(0)
    t = np.zeros(1)
    th = [_p(o[k]) for k in ("T", "D", "H", "DH")] if thermal else [None] * 4
    rc = lib().chs_restart_read_io(str(directory).encode(), float(tfm), L, M, N, int(ini_field_type), C.byref(it),
                                   _p(t), _p(o["FLOW"]), _p(o["PR"]), *th)
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())
    o["ITERG"], o["PhyTIME"] = it.value, float(t[0])
    return o


def interp_initial_tg(qo, po, HXoo, HZoo, yndo, L, M, N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8):
    """INITIAL_INTERP_tg (chs_interp_initial_tg) from a coarse field given as halo arrays
    qo [3, No, Mo+2, Lo], po [No, Mo+2, Lo] and its y nodes yndo [Mo+1]; returns q, pr on the
    channel mesh (L, M, N, HX, HZ, STR2) with u scaled to unit bulk velocity."""
    qo = np.asarray(qo, dtype=np.float64)
    _, No, Mh, Lo = qo.shape
    Mo = Mh - 2
    parts = [np.ascontiguousarray(a[:, 1:Mo + 1, :]) for a in (qo[0], qo[1], qo[2], np.asarray(po, dtype=np.float64))]
    yndo = np.ascontiguousarray(yndo, dtype=np.float64)
    q = np.zeros((3, N, M + 2, L))
    pr = np.zeros((N, M + 2, L))
    rc = lib().chs_interp_initial_tg(Lo, Mo, No, float(HXoo), float(HZoo), _p(yndo), *[_p(a) for a in parts],
                                     L, M, N, float(HX), float(HZ), float(STR2), _p(q), _p(pr))
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())
    return q, pr


def random_field_tg(L, M, N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8, random_type=2, vper=0.6, svper=0.3, seed=0,
                    m_over_d=1.0):
    """random_FL_FLD_TG (chs_random_field_tg): the start field before the NS = 0 projection, as
    halo arrays q [3, N, M+2, L], pr [N, M+2, L].  random_type 3 is the reference's sine field;
    1 and 2 draw the perturbations from a seeded counter-based generator in place of the
    compiler's RANDOM_NUMBER (type 2 keeps the reference's dependence on I + K + J only).  The
    velocity halo rows hold the opposite wall's first row, as the reference's ring exchange
    leaves them for the projection substep."""
    q = np.zeros((3, N, M + 2, L))
    pr = np.zeros((N, M + 2, L))
    rc = lib().chs_random_field_tg(L, M, N, float(HX), float(HZ), float(STR2), int(random_type), float(vper),
                                   float(svper), int(seed), float(m_over_d), _p(q), _p(pr))
    if rc != 0:
        raise ChsError(lib().chs_restart_last_error().decode())
    return q, pr
