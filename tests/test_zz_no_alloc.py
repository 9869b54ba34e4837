"""chs_tdma_lines_noalloc (the batched Thomas solvers without allocation, work buffer in a
static per-GPU array) against chs_tdma_lines on the same inputs, bit for bit, on two calls in
a row, the second smaller, for every kind; and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def system(kind, isz, jsz, seed):
    rng = np.random.default_rng(seed)
    A, C_ = rng.uniform(-1, 1, (2, jsz, isz))
    B = 2.5 + rng.uniform(0, 1, (jsz, isz))
    R = rng.standard_normal((jsz, isz))
    nl = jsz if kind in (0, 2) else isz
    return A, B, C_, R, rng.standard_normal((2, nl))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_noalloc_equals_allocating_entry_point(pkg, kind):
    for (isz, jsz), seed in (((300, 517), 1), ((40, 23), 2)):   # the second call is smaller
        if kind in (1, 3):
            isz, jsz = jsz, isz
        print(f"kind {kind} {isz}x{jsz}", flush=True)
        A, B, C_, R, bc = system(kind, isz, jsz, 100 * kind + seed)
        ref = pkg.tdma_lines(kind, A, B, C_, R, bc)
        got = pkg.tdma_lines(kind, A, B, C_, R, bc, noalloc=True)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_sizes_beyond_its_buffer(pkg):
    n, nl = 1024, 700   # 3 n nl > 2^21 doubles
    z = np.zeros((nl, n))
    with pytest.raises(pkg.ChsError):
        pkg.tdma_lines(0, z, z + 3.0, z, z, noalloc=True)
