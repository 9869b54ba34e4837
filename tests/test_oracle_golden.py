"""C oracle (oracle/tg_oracle.c) against golden vectors written by the compiled reference."""
import numpy as np
import pytest

from oracle import oracle, setup

CASES = ["tg_16x16x16.npz", "tg_20x18x30.npz"]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def make(g):
    L, M, N, STR2, REN, dt, seed = (g["params"][i] for i in range(7))
    m = setup.Mesh(int(L), int(M), int(N), STR2=STR2)
    o = oracle.OracleTG(m, REN)
    o.set_fields(g["q_init"], g["pr_init"])
    return o, dt


@pytest.mark.parametrize("name", CASES)
def test_convection(golden, name):
    g = golden(name)
    o, _ = make(g)
    cx, cy, cz = o.convection()
    ref = g["conv"]
    assert rel(cx, ref[0]) < 1e-13
    assert rel(cy[:, 1:], ref[1][:, 1:]) < 1e-13   # v rows J >= 2 only (CONVECTION_Y.f90:34-35)
    assert rel(cz, ref[2]) < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_projection_and_step(golden, name):
    g = golden(name)
    o, dt = make(g)
    o.substep(0, dt)
    for c in range(3):
        assert rel(o.q[c], g["q_ns0"][c]) < 1e-12, c
    assert rel(o.pr, g["pr_ns0"]) < 1e-10, rel(o.pr, g["pr_ns0"])
    o.step(dt)
    for c in range(3):
        assert rel(o.q[c], g["q_step1"][c]) < 1e-12, c
    assert rel(o.pr, g["pr_step1"]) < 1e-10, rel(o.pr, g["pr_step1"])


@pytest.mark.parametrize("name", CASES)
def test_poisson(golden, name):
    g = golden(name)
    o, _ = make(g)
    phi = o.poisson(g["poisson_rhs"])
    assert rel(phi, g["poisson_phi"]) < 1e-10
