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
