"""The oracle's per-tile POINV operations (the LAPACK++/BLAS++ calls of
examples/potrf/trtri_L.h and lauum.h) against what fixes them: exact
integer products for the multiplications (numpy's product is then exact),
and L X = B, L Linv = I to rounding for the triangular solves and the
inverse.  A transposed operand, a wrong triangle or a dropped alpha fails
one of these."""
import numpy as np
import pytest

import oracle


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


def rint(rng, *shape):
    return F(rng.integers(-8, 9, size=shape))


def lower_spd_factor(rng, n):
    """A well-conditioned lower-triangular tile with junk in its strict upper part."""
    L = np.tril(rng.uniform(-1, 1, (n, n))) / np.sqrt(n) + np.diag(rng.uniform(1, 2, n))
    junk = np.triu(rng.uniform(-9, 9, (n, n)), 1)
    return F(L + junk), F(L)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 5, 64), (64, 64, 64), (96, 40, 7)])
def test_gemm_xn_exact_on_integers(m, n, k):
    rng = np.random.default_rng(m * 100 + n)
    for ta in (0, 1):
        A = rint(rng, k, m) if ta else rint(rng, m, k)
        B, C = rint(rng, k, n), rint(rng, m, n)
        want = 2.0 * ((A.T if ta else A) @ B) - 3.0 * C
        oracle.tile_gemm_xn(ta, A, B, C, alpha=2.0, beta=-3.0)
        assert np.array_equal(C, want)


@pytest.mark.parametrize("n,k", [(1, 3), (33, 17), (64, 64)])
def test_syrk_lt_exact_lower_only(n, k):
    rng = np.random.default_rng(n)
    A, C = rint(rng, k, n), rint(rng, n, n)
    up = np.triu(C, 1).copy()
    want = np.tril(-1.0 * (A.T @ A) + 2.0 * C)
    oracle.tile_syrk_lt(A, C, alpha=-1.0, beta=2.0)
    assert np.array_equal(np.tril(C), want)
    assert np.array_equal(np.triu(C, 1), up)  # the strict upper triangle is untouched


@pytest.mark.parametrize("m,n", [(1, 4), (50, 50), (64, 9)])
def test_trmm_llt_exact_and_ignores_upper(m, n):
    rng = np.random.default_rng(m + 7 * n)
    Lfull = rint(rng, m, m)
    B = rint(rng, m, n)
    want = np.tril(Lfull).T @ B
    oracle.tile_trmm_llt(Lfull, B)
    assert np.array_equal(B, want)


@pytest.mark.parametrize("n", [1, 31, 64, 100])
def test_lauum_exact_lower_only(n):
    rng = np.random.default_rng(n)
    A = rint(rng, n, n)
    L = np.tril(A)
    up = np.triu(A, 1).copy()
    oracle.tile_lauum(A)
    assert np.array_equal(np.tril(A), np.tril(L.T @ L))
    assert np.array_equal(np.triu(A, 1), up)


@pytest.mark.parametrize("m,n", [(1, 1), (40, 13), (64, 64), (128, 3)])
def test_triangular_solves_residual(m, n):
    rng = np.random.default_rng(m * 3 + n)
    # left: L m x m, B m x n
    Lj, L = lower_spd_factor(rng, m)
    B0 = F(rng.uniform(-1, 1, (m, n)))
    X = B0.copy(order="F")
    oracle.tile_trsm_lln(Lj, X)
    assert np.abs(L @ X - B0).max() <= 1e-13 * np.abs(B0).max() * m
    # right: L n x n, B m x n, alpha = -1: X L = -B
    Lj, L = lower_spd_factor(rng, n)
    X = B0.copy(order="F")
    oracle.tile_trsm_rln(Lj, X, alpha=-1.0)
    assert np.abs(X @ L + B0).max() <= 1e-13 * np.abs(B0).max() * n


@pytest.mark.parametrize("n", [1, 16, 64, 129])
def test_trtri_inverse_and_info(n):
    rng = np.random.default_rng(n)
    Aj, L = lower_spd_factor(rng, n)
    up = np.triu(Aj, 1).copy()
    assert oracle.tile_trtri(Aj) == 0
    Linv = np.tril(Aj)
    assert np.abs(L @ Linv - np.eye(n)).max() <= 1e-13 * n
    assert np.array_equal(np.triu(Aj, 1), up)
    # an exact zero on the diagonal: LAPACK info = its 1-based index
    if n > 1:
        Z = lower_spd_factor(rng, n)[0]
        Z[n // 2, n // 2] = 0.0
        assert oracle.tile_trtri(Z) == n // 2 + 1
