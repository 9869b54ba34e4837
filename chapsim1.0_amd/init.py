"""Seeded initial channel fields.

The reference builds its initial TG field from the laminar profile (LAMPOISLPROF.f90:29-51,
1.5 (1 - y^2) at cell centres) plus random perturbations from the compiler's intrinsic RNG
(IniField_FLOW_tg, INITIAL_FL_FLD_TG.f90:165-213), which no other compiler reproduces.  This
generator keeps the same structure with a counter-based RNG (numpy Philox) so a field is a
pure function of (seed, mesh): u = Poiseuille + U(-a, a), v = U(-a, a) off the walls,
w = U(-a, a), wall ghost rows as BC_WALL_Q_tg.  The field is not divergence-free; the NS = 0
substep projects it, as CALC_INITIALIZATION_tg does (INITIAL_FL_FLD_TG.f90:121).
"""
from __future__ import annotations

import numpy as np


def poiseuille_perturbed(coeffs: dict, seed: int, amp: float = 0.1):
    L, M, N = int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"])
    dyfi = np.asarray(coeffs["dyfi"])
    ynd = np.concatenate([[-1.0], -1.0 + np.cumsum(1.0 / dyfi)])
    ycc = 0.5 * (ynd[:-1] + ynd[1:])
    vini = np.where(np.abs(ycc) < 1.0, (1.0 - ycc ** 2) * 1.5, 0.0)
    rng = np.random.Generator(np.random.Philox(key=seed))
    q = np.zeros((3, N, M + 2, L))
    pert = rng.uniform(-amp, amp, size=(3, N, M, L))
    q[0, :, 1:M + 1, :] = vini[None, :, None] + pert[0]
    q[1, :, 2:M + 1, :] = pert[1, :, 1:, :]
    q[2, :, 1:M + 1, :] = pert[2]
    return q, np.zeros((N, M + 2, L))


def heated_channel_initial(coeffs: dict, table: dict, seed: int, amp: float = 0.05, t_amp: float = 0.01):
    """Seeded start of the heated main-domain channel: T / T0 = 1 + t_amp (smooth + noise)
    mapped to rho h through the nondimensional table (monotone in T; the ghost rows take the
    first / last cell's value), G = rho u with the laminar profile plus amp noise (v and w
    noise only, zero on the walls) and zero pressure.  Returns DH, PR [N, M+2, L] and
    G [3, N, M+2, L]; densities follow from DH in the energy context."""
    L, M, N = int(coeffs["L"]), int(coeffs["M"]), int(coeffs["N"])
    dyfi = np.asarray(coeffs["dyfi"])
    ynd = np.concatenate([[-1.0], -1.0 + np.cumsum(1.0 / dyfi)])
    ycc = 0.5 * (ynd[:-1] + ynd[1:])
    rng = np.random.Generator(np.random.Philox(key=seed))
    x = np.arange(L) / L
    z = np.arange(N) / N
    T = 1.0 + t_amp * (np.sin(2 * np.pi * x)[None, None, :] * np.cos(2 * np.pi * z)[:, None, None]
                       + 0.2 * rng.standard_normal((N, M, L)))
    dh = np.zeros((N, M + 2, L))
    dh[:, 1:M + 1, :] = np.interp(T, table["T"], table["DH"])
    dh[:, 0, :] = dh[:, 1, :]
    dh[:, M + 1, :] = dh[:, M, :]
    g = np.zeros((3, N, M + 2, L))
    g[0, :, 1:M + 1, :] = 1.5 * (1.0 - ycc ** 2)[None, :, None] + rng.uniform(-amp, amp, (N, M, L))
    g[1, :, 2:M + 1, :] = rng.uniform(-amp, amp, (N, M - 1, L))
    g[2, :, 1:M + 1, :] = rng.uniform(-amp, amp, (N, M, L))
    return dh, g, np.zeros((N, M + 2, L))


def channel_ynd(M: int, STR2: float):
    """y nodes YND(1..M+1) of the tanh-stretched channel (COORDJR.f90:47-51)."""
    x2 = np.arange(M + 1) / M
    return np.tanh(STR2 * (x2 - 0.5)) / np.tanh(STR2 * 0.5)


def add_perturbation(q, seed: int, amp: float):
    """Seeded uniform perturbation U(-amp, amp) of u, w (rows 1..M) and v (rows 2..M), in place."""
    _, N, Mh, L = q.shape
    M = Mh - 2
    rng = np.random.Generator(np.random.Philox(key=seed))
    pert = rng.uniform(-amp, amp, size=(3, N, M, L))
    q[0, :, 1:M + 1, :] += pert[0]
    q[1, :, 2:M + 1, :] += pert[1, :, 1:, :]
    q[2, :, 1:M + 1, :] += pert[2]
    return q


def ren_from_retau(re_tau: float) -> float:
    """Bulk Reynolds number on the half height from Re_tau by Dean's correlation
    Re_tau = 0.09 (2 REN)^0.88 (REN = 2800 <-> Re_tau ~ 180, as the reference's templates)."""
    return (re_tau / 0.09) ** (1.0 / 0.88) / 2.0


def _smooth(v: int) -> bool:
    for p in (2, 3, 5):
        while v % p == 0:
            v //= p
    return v == 1


def ensemble_cases(n: int = 32, seed: int = 4):
    """Seeded channel sizes of the batched ensemble workload: L, N in [64, 192] and M in
    [64, 128], L and N even with no prime factor above 5 (FFT99's condition,
    lib/fft99/FFTFAX.f90:56-91), Re_tau uniform in [180, 395] (REN by Dean's correlation),
    dt = 5e-4, box 4 pi x 2 x 2 pi, tanh stretching 1.8 .. 2.0 with Re_tau."""
    ok = [v for v in range(64, 193, 2) if _smooth(v)]
    rng = np.random.Generator(np.random.Philox(key=seed))
    cases = []
    for _ in range(n):
        L = int(rng.choice(ok))
        N = int(rng.choice(ok))
        M = int(rng.integers(64, 129))
        rt = float(rng.uniform(180.0, 395.0))
        cases.append(dict(L=L, M=M, N=N, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8 + 0.2 * (rt - 180) / 215,
                          REN=ren_from_retau(rt), dt=5.0e-4, re_tau=rt))
    return cases
