"""Pins of the oracle's checker functions by constructions whose answer the
mathematics fixes: exactly representable factors with planted symmetric
perturbations (check_dpotrf, testing_dpotrf.cc:277-335), planted errors on
exact integer solves (the solve backward error) and the mirrored dense norm
(||A||_inf of a symmetric matrix held by its lower tiles).  No GPU."""
import numpy as np
import pytest

EPS = np.finfo(np.float64).eps


def _stored(N, NB, S):
    """What getLAPACKMatrix (pmw.h:122-139) returns for the symmetric S held
    in lower tiles: lower tiles as stored, diagonal tiles whole, zeros above."""
    D = np.zeros_like(S)
    for i in range(N):
        for jb in range(0, i // NB + 1):
            j0, j1 = jb * NB, min(N, (jb + 1) * NB)
            D[i, j0:j1] = S[i, j0:j1]
    return D


def _int_factor(N, rng, heavy_col0=False):
    """Lower triangular L with small integer entries: L L^T is exact in FP64."""
    L = np.tril(rng.integers(-2, 3, (N, N)).astype(np.float64), -1)
    L[np.diag_indices(N)] = 8.0 + rng.integers(0, 3, N)
    if heavy_col0:
        L[1:, 0] = 6.0  # row 0 of L L^T carries the largest symmetric row sum
    return L


def _planted(N, rng, scale=2.0 ** -30, density=0.2):
    """Symmetric E of dyadic entries, exactly representable next to L L^T."""
    E = rng.integers(-3, 4, (N, N)).astype(np.float64) * scale
    E[rng.random((N, N)) > density] = 0.0
    return np.tril(E) + np.tril(E, -1).T


@pytest.mark.parametrize("N,NB,heavy", [(96, 32, False), (100, 32, True), (130, 64, False), (77, 77, True),
                                        (300, 64, True)])
def test_check_dpotrf_planted_perturbation_closed_form(oracle, N, NB, heavy):
    """check_dpotrf(L, A0) with A0 = L L^T - E returns ||E||_inf / (||A0||_inf
    N eps), ||A0||_inf of A0 as getLAPACKMatrix stores it (diagonal tiles
    whole), for the dense and the tiled forms alike."""
    rng = np.random.default_rng(N + NB)
    L = _int_factor(N, rng, heavy)
    E = _planted(N, rng)
    S = L @ L.T - E  # exact: integers plus dyadic E
    A0 = _stored(N, NB, S)
    # the factor as it leaves POTRF: L below, the input's upper part untouched
    Lst = _stored(N, NB, np.tril(L, -1) + np.diag(np.diag(L)) + np.triu(S, 1))
    want = np.abs(E).sum(1).max() / (np.abs(A0).sum(1).max() * N * EPS)
    got_dense = oracle.check_dpotrf_dense(Lst, A0)
    got_tiled = oracle.check_dpotrf_tiled(N, NB, oracle.dense_to_tiles(N, NB, Lst), oracle.dense_to_tiles(N, NB, A0))
    assert got_dense == pytest.approx(want, rel=1e-13)
    assert got_tiled == pytest.approx(want, rel=1e-13)
    # with more than one tile row the denominator is not the symmetric norm
    # when the largest row is near the top (a mirrored ||A0|| would differ)
    if heavy and N > NB:
        sym = np.abs(S).sum(1).max()
        assert abs(sym - np.abs(A0).sum(1).max()) > 1.0


def test_check_dpotrf_residual_is_mirrored(oracle):
    """A perturbation in one strictly lower entry counts in two rows of the
    mirrored residual: a planted E with a single off-diagonal pair whose
    row-sum maximum needs the mirror."""
    N, NB = 64, 32
    rng = np.random.default_rng(5)
    L = _int_factor(N, rng)
    E = np.zeros((N, N))
    E[40, 3] = E[3, 40] = 2.0 ** -20
    E[3, 7] = E[7, 3] = 2.0 ** -20  # row 3 holds two entries: one from the mirror
    S = L @ L.T - E
    A0 = _stored(N, NB, S)
    Lst = _stored(N, NB, np.tril(L) + np.triu(S, 1))
    want = 2.0 ** -19 / (np.abs(A0).sum(1).max() * N * EPS)
    assert oracle.check_dpotrf_dense(Lst, A0) == pytest.approx(want, rel=1e-13)
    assert oracle.check_dpotrf_tiled(N, NB, oracle.dense_to_tiles(N, NB, Lst),
                                     oracle.dense_to_tiles(N, NB, A0)) == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("N,NB", [(160, 32), (300, 64)])
def test_check_dpotrf_dense_and_tiled_on_a_real_factor(oracle, N, NB):
    """On a computed factor the numerator ||L L^T - A0|| is rounding noise of
    the product L L^T, which the dense and tiled forms sum in different
    orders, so they agree only in size; their exact agreement is pinned on
    the exactly representable planted cases above."""
    A0 = oracle.plgsy_tiles(N, NB, N, 3872)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    d = oracle.check_dpotrf_dense(oracle.tiles_to_dense(N, NB, L), oracle.tiles_to_dense(N, NB, A0))
    t = oracle.check_dpotrf_tiled(N, NB, L, A0)
    assert 0 < d <= 60 and 0 < t <= 60
    assert 0.5 < t / d < 2.0


def test_norm_inf_sym_tiled_exact_on_integers(oracle):
    """||A||_inf of the mirrored dense matrix, exactly, on integer data whose
    largest row sum needs the upper (mirrored) part; the diagonal tiles' upper
    part holds garbage that must not be read."""
    for (N, NB) in [(97, 32), (128, 32), (50, 64)]:
        rng = np.random.default_rng(N)
        S = rng.integers(-9, 10, (N, N)).astype(np.float64)
        S = np.tril(S) + np.tril(S, -1).T
        S[1:, 0] = 50.0
        S[0, 1:] = 50.0
        st = _stored(N, NB, np.tril(S) + np.triu(np.full((N, N), 1e6), 1))
        t = oracle.dense_to_tiles(N, NB, st)
        assert oracle.norm_inf_sym_tiled(N, NB, t) == np.abs(S).sum(1).max()


@pytest.mark.parametrize("gen", ["plgsy", "spd", "uniform"])
def test_norm_inf_sym_tiled_generators(oracle, gen):
    N, NB = 300, 64  # ragged edge tiles
    t = {"plgsy": lambda: oracle.plgsy_tiles(N, NB, N, 3872), "spd": lambda: oracle.spd_tiles(N, NB, 0.5, 9),
         "uniform": lambda: oracle.spd_uniform_tiles(N, NB, 0.5, 9)}[gen]()
    D = oracle.tiles_to_dense(N, NB, t)
    S = np.tril(D) + np.tril(D, -1).T
    assert oracle.norm_inf_sym_tiled(N, NB, t) == pytest.approx(np.linalg.norm(S, np.inf), rel=1e-15)


def test_solve_backward_error_planted_closed_form(oracle):
    """B = A X exactly (integers); X' = X + d e_j in one column gives the
    residual -d A e_j, so the backward error of that column is
    d max_i |A_ij| / (||A||_inf ||x'||_inf + ||b||_inf) and the maximum over
    the columns picks it."""
    N, NB, nrhs = 90, 32, 4
    rng = np.random.default_rng(17)
    S = rng.integers(-20, 21, (N, N)).astype(np.float64)
    S = np.tril(S) + np.tril(S, -1).T
    S[60, :] = S[:, 60] = 30.0  # a heavy row only the mirror completes in lower tiles
    X = rng.integers(-5, 6, (N, nrhs)).astype(np.float64)
    B = S @ X
    t = oracle.dense_to_tiles(N, NB, np.tril(S) + np.triu(np.full((N, N), 7e5), 1))
    assert oracle.solve_backward_error(N, NB, t, X, B) == 0.0
    d, j, col = 2.0 ** -10, 13, 2
    Xp = X.copy()
    Xp[j, col] += d
    an = np.abs(S).sum(1).max()
    want_col = d * np.abs(S[:, j]).max() / (an * np.abs(Xp[:, col]).max() + np.abs(B[:, col]).max())
    assert oracle.solve_backward_error(N, NB, t, Xp, B) == pytest.approx(want_col, rel=1e-14)
    # a larger error in another column takes over the maximum
    Xq = Xp.copy()
    Xq[60, 0] += 4 * d
    want0 = 4 * d * np.abs(S[:, 60]).max() / (an * np.abs(Xq[:, 0]).max() + np.abs(B[:, 0]).max())
    assert oracle.solve_backward_error(N, NB, t, Xq, B) == pytest.approx(max(want0, want_col), rel=1e-14)
    R = oracle.residual_tiled(N, NB, t, Xq, B)
    assert np.array_equal(R, B - S @ Xq)


def test_spd_uniform_values_and_product(oracle):
    """B(i,j) = 2 (ran >> 11) 2^-53 - 1 from the LCG at stream position
    i + j N (checked against Python integer stepping), and A = B B^T / N +
    alpha I against numpy."""
    a, c, mask = 6364136223846793005, 1, (1 << 64) - 1
    N, NB, alpha, seed = 40, 16, 0.75, 2024
    x = seed
    vals = []
    for _ in range(N * N):
        vals.append(2.0 * ((x >> 11) * 2.0 ** -53) - 1.0)
        x = (a * x + c) & mask
    Bw = np.array(vals).reshape(N, N).T  # position i + j N -> B[i, j]
    Bo = np.array([[oracle.lib().oracle_spd_uval(i, j, N, seed) for j in range(N)] for i in range(N)])
    assert np.array_equal(Bo, Bw)
    assert Bw.min() >= -1.0 and Bw.max() < 1.0
    t = oracle.spd_uniform_tiles(N, NB, alpha, seed)
    D = oracle.tiles_to_dense(N, NB, t)
    want = Bw @ Bw.T / N + alpha * np.eye(N)
    np.testing.assert_allclose(np.tril(D), np.tril(want), rtol=0, atol=1e-15)
    # first and second moments of a larger draw: uniform on [-1, 1)
    big = np.array([oracle.lib().oracle_spd_uval(i, 0, 20000, 7) for i in range(20000)])
    assert abs(big.mean()) < 0.02 and abs(big.var() - 1.0 / 3.0) < 0.01
