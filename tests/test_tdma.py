"""The reference's batched Thomas solvers (TDMAIJI_CYC, TDMAIJJ_nonCYC, TDMAIJI_nonCYC,
TDMAIJJ_CYC).  The oracle restatement (or_tdma_lines) is pinned bit for bit against vectors the
compiled reference wrote (tests/golden/tdma.npz, make_golden.py make_tdma) and against a dense
solve; the GPU entry point chs_tdma_lines is compared with both, bit for bit."""
import numpy as np
import pytest

from oracle import oracle

KINDS = range(4)


def _cases(golden):
    g = golden("tdma.npz")
    tags = sorted({k.rsplit("_", 1)[0] for k in g})
    return g, tags


def test_oracle_matches_reference_bitwise(golden):
    g, tags = _cases(golden)
    assert len(tags) == 20
    for t in tags:
        X = oracle.tdma_lines(int(t[1]), g[t + "_A"], g[t + "_B"], g[t + "_C"], g[t + "_R"], g[t + "_bc"])
        assert np.array_equal(X, g[t + "_X"]), t


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_solves_the_system(kind):
    """Dense check of the restatement: the residual of the assembled (cyclic) tridiagonal
    system, boundary values moved to the right-hand side."""
    rng = np.random.default_rng(kind)
    isz, jsz = 9, 6
    A, C_ = rng.uniform(-1, 1, (2, jsz, isz))
    B = 2.5 + rng.uniform(0, 1, (jsz, isz))
    R = rng.standard_normal((jsz, isz))
    alongI, cyc = kind in (0, 2), kind in (0, 3)
    nl = jsz if alongI else isz
    bc = rng.standard_normal((2, nl)) if not cyc else np.zeros((2, nl))
    X = oracle.tdma_lines(kind, A, B, C_, R, bc)
    for line in range(nl):
        sl = (line, slice(None)) if alongI else (slice(None), line)
        a, b, c, r, x = A[sl], B[sl], C_[sl], R[sl].copy(), X[sl]
        Mx = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        if cyc:
            Mx[0, -1], Mx[-1, 0] = a[0], c[-1]
        else:
            r[0] -= a[0] * bc[0, line]
            r[-1] -= c[-1] * bc[1, line]
        assert np.abs(Mx @ x - r).max() < 1e-14


@pytest.mark.gpu
def test_gpu_matches_reference_bitwise(pkg, golden):
    g, tags = _cases(golden)
    for t in tags:
        print(f"tdma {t}", flush=True)
        X = pkg.tdma_lines(int(t[1]), g[t + "_A"], g[t + "_B"], g[t + "_C"], g[t + "_R"], g[t + "_bc"])
        assert np.array_equal(X, g[t + "_X"]), (t, np.abs(X - g[t + "_X"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("kind", KINDS)
def test_gpu_matches_oracle_many_lines(pkg, kind):
    """Wide batches (more lines than a block) and long lines, bit for bit against the oracle."""
    rng = np.random.default_rng(10 + kind)
    isz, jsz = (300, 517) if kind in (0, 2) else (517, 300)
    A, C_ = rng.uniform(-1, 1, (2, jsz, isz))
    B = 2.5 + rng.uniform(0, 1, (jsz, isz))
    R = rng.standard_normal((jsz, isz))
    nl = jsz if kind in (0, 2) else isz
    bc = rng.standard_normal((2, nl))
    X = pkg.tdma_lines(kind, A, B, C_, R, bc)
    assert np.array_equal(X, oracle.tdma_lines(kind, A, B, C_, R, bc))


@pytest.mark.gpu
def test_gpu_rejects_bad_arguments(pkg):
    z = np.zeros((4, 2))
    with pytest.raises(pkg.ChsError):
        pkg.tdma_lines(0, z, z + 1, z, z)          # cyclic line of 2 points
    with pytest.raises(pkg.ChsError):
        pkg.tdma_lines(7, z, z + 1, z, z)
