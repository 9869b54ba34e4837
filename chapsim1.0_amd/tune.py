"""Launch-shape search for one channel (the keys of chs_set_tuning).

Each key is swept over its candidates with the others fixed, timing only the kernels that the
key shapes (HIP events of chs_profile_*), and the fastest value is kept.  Fields are saved
before and restored after, so the channel's state is unchanged; the step graph is re-made
on the next step.  Shapes change no result except the summation order of the mass-flux
partials (conv_by, conv_kc, conv_lds, conv_dma) and the rounding of the chunked y-Thomas (ytdma_c).
"""
from __future__ import annotations

from .tg import ChsError

KERNELS = {
    "conv_by": ("conv_rhs",), "conv_kc": ("conv_rhs",), "conv_bx": ("conv_rhs",), "conv_fence": ("conv_rhs",),
    "conv_fx": ("conv_rhs", "cn_x"),
    "cnx_nlx": ("cn_x",), "cn_xz": ("cn_x", "cn_z", "cn_xz"), "cnxz_by": ("cn_xz",),
    "cnz_kz": ("cn_z", "cn_xz"), "cny_ky": ("cn_y",), "fftx_tk": ("divg_fftx", "fftx_inv"),
    "fftz_nl": ("fftz_fwd", "fftz_inv"), "nxcd": ("conv_rhs",),
    "fft_reg": ("divg_fftx", "fftz_fwd", "fftz_inv", "fftx_inv"),
    "ytdma_bs": ("ytdma",), "ytdma_rb": ("ytdma",), "ytdma_c": ("ytdma",), "conv_lds": ("conv_rhs", "cn_x"),
    "conv_dma": ("conv_rhs", "cn_x"), "cny3_ky": ("cn_y",), "divg_nxcd": ("divg_fftx",),
    "upd_kz": ("update",), "upd_by": ("update",), "upd_nxcd": ("update",),
    "upd_w2": ("update",), "cnzy": ("cn_z", "cn_y", "cn_zy"), "cnzy_ky": ("cn_zy",), "cnzy_nxcd": ("cn_zy",),
}


def candidates(L: int, M: int, N: int) -> dict[str, list[int]]:
    """Values of each key that satisfy its limits on an L x M x N mesh."""
    c = {
        "conv_by": [2, 4, 8, 16],
        "conv_kc": [k for k in (4, 8, 16, 32) if k <= N],
        "conv_bx": [1, 4] if L % 256 == 0 else [1],
        "conv_fence": [0, 1],
        "conv_fx": [f for f in (0, 1, 2) if f == 0 or (L % 64 == 0 and L * f <= 512 and M % f == 0)],
        "cn_xz": [0, 1],
        "cnxz_by": [1, 2],
        "cnx_nlx": [n for n in (1, 2, 4, 8, 16) if n * L * 8 <= 65536 and n <= M * N],
        "cnz_kz": [k for k in (8, 16, 32, 64) if k <= N],
        "cny_ky": [k for k in (16, 24, 32, 48, 64, 96) if k <= M],
        "fftx_tk": [t for t in (2, 4, 8, 16, 32) if N % t == 0 and (t // 2) * L <= 2048],
        "fftz_nl": [n for n in (1, 2, 4, 8, 16) if n * N <= 2048],
        "fft_reg": [0, 1],
        "nxcd": [1, 8],
        "ytdma_bs": [64, 128, 256],
        "ytdma_rb": [8, 16],
        "ytdma_c": [c for c in (0, 4, 8, 16) if c == 0 or (M <= 256 and -(-M // c) <= 16)],
        "conv_lds": [v for v in (0, 1, 2, 4) if v == 0 or (L % 64 == 0 and L <= 256 and M % v == 0 and L * v <= 1024)],
        "conv_dma": [v for v in (0, 2, 4) if v == 0 or (L % 64 == 0 and L <= 256 and M % v == 0 and L * v <= 1024)],
        "cny3_ky": [v for v in (-1, 0, 48, 64, 96) if v <= M],
        "divg_nxcd": [1, 8],
        "upd_kz": [k for k in (4, 8, 16) if k <= N],
        "upd_by": [1, 2, 4, 8, 16],
        "upd_nxcd": [1, 8],
        "upd_w2": [0, 1],
        "cnzy": [0, 1, 2],
        "cnzy_ky": [v for v in (0, 24, 48, 96) if v <= M],
        "cnzy_nxcd": [1, 8],
    }
    return {k: v for k, v in c.items() if v}


def _time(ch, names, dt, cvisc, steps):
    ch.profile_reset(None)
    for _ in range(steps):
        ch.rk3_step(dt, cvisc)
    tot = 0.0
    for n in names:
        avg, cnt = ch.profile_get(n)
        tot += avg * cnt
    return tot / steps


def autotune(ch, dt: float, cvisc: float, steps: int = 3, keys=None, verbose=False) -> dict:
    """Sweeps the launch shapes of channel ``ch`` (chapsim_amd.ChannelTG with fields loaded)
    at time step dt and viscosity cvisc; applies and returns the best value per key and the
    timings {key: {value: ms per RK3 step of the key's kernels}}."""
    saved = ch.download(with_history=True)
    graph = getattr(ch, "graph", False)
    ch.set_graph(False)
    best, table = {}, {}
    try:
        ch.rk3_step(dt, cvisc)                       # CN tables, warm caches
        for key, vals in candidates(ch.L, ch.M, ch.N).items():
            if keys is not None and key not in keys:
                continue
            table[key] = {}
            for v in vals:
                try:
                    ch.set_tuning(**{key: v})
                except ChsError:              # excluded by the current value of another key
                    continue
                table[key][v] = _time(ch, KERNELS[key], dt, cvisc, steps)
            best[key] = min(table[key], key=table[key].get)
            ch.set_tuning(**{key: best[key]})
            if verbose:
                print(f"tune {key}: " + ", ".join(f"{v}={t:.4f}" for v, t in table[key].items())
                      + f" -> {best[key]}", flush=True)
    finally:
        ch.profile_reset("none")
        ch.upload(saved[0], saved[1], saved[2])
        ch.set_graph(graph)
    return {"best": best, "ms": table}
