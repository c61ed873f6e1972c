"""The CPU oracle pinned against the reference's own generator, LAPACK
fixtures, closed forms and invariants (no GPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def golden_plgsy():
    return np.load(os.path.join(GOLDEN, "plgsy_ref.npz"))


@pytest.fixture(scope="module")
def lapack():
    return np.load(os.path.join(GOLDEN, "lapack_tiles.npz"))


def test_flops_closed_form(oracle):
    for n in (1, 2, 32, 160, 4096, 65536):
        assert oracle.lib().oracle_flops_dpotrf(n) == pytest.approx(n ** 3 / 3 + n ** 2 / 2 + n / 6, rel=1e-15)


def test_rnd64_jump_brute_force(oracle):
    a, c, mask = 6364136223846793005, 1, (1 << 64) - 1
    for seed in (0, 3872, 2 ** 63 + 11):
        x = seed
        for n in range(300):
            assert oracle.lib().oracle_rnd64_jump(n, seed) == x
            x = (a * x + c) & mask


def test_rnd64_jump_golden(oracle, golden_plgsy):
    for n, want in zip(golden_plgsy["jump_n"], golden_plgsy["jump_seed3872"]):
        assert oracle.lib().oracle_rnd64_jump(int(n), 3872) == int(want)


@pytest.mark.parametrize("N,NB", [(160, 32), (100, 32)])
def test_plgsy_bit_exact_vs_reference_golden(oracle, golden_plgsy, N, NB):
    t = oracle.plgsy_tiles(N, NB, N, 3872)
    ref = golden_plgsy[f"plgsy_{N}_{NB}"]
    # compare only the entries the generator defines (edge tiles are partial)
    nt = oracle.nt_of(N, NB)
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            a = oracle.tile_view(t, N, NB, m, n)[:mb, :nb]
            b = oracle.tile_view(ref, N, NB, m, n)[:mb, :nb]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_plgsy_bit_exact_vs_reference_build(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes

    for (bump, m, n, gM, m0, n0) in [(7.0, 13, 9, 50, 26, 13), (0.0, 9, 13, 50, 13, 26), (50.0, 11, 11, 50, 22, 22)]:
        A = np.zeros((m, n), order="F")
        B = np.zeros((m, n), order="F")
        oracle.lib().oracle_core_plgsy(bump, m, n, oracle.ptr(A), m, gM, m0, n0, 99)
        R.ref_core_plgsy(bump, m, n, ctypes.c_void_p(B.ctypes.data), m, gM, m0, n0, 99)
        assert np.array_equal(A.view(np.uint64), B.view(np.uint64))


def test_plgsy_symmetric_and_bumped(oracle):
    N, NB = 96, 32
    t = oracle.plgsy_tiles(N, NB, N, 3872)
    D = oracle.tiles_to_dense(N, NB, t)
    full = np.tril(D) + np.tril(D, -1).T
    # diagonal tiles hold the mirrored upper part
    for k in range(3):
        tile = oracle.tile_view(t, N, NB, k, k)
        assert np.array_equal(tile, tile.T)
    assert np.all(np.abs(np.diag(full)) > N - 1)  # |0.5 - u| <= 0.5 plus bump N
    assert np.all(np.abs(full - np.diag(np.diag(full))) <= 0.5)


def test_tile_kernels_vs_lapack(oracle, lapack):
    O = oracle
    for n in (1, 7, 32, 45, 64):
        A = np.asfortranarray(lapack[f"potrf_A_{n}"].copy())
        assert O.lib().oracle_dpotrf_tile(n, O.ptr(A), n) == 0
        np.testing.assert_allclose(np.tril(A), lapack[f"potrf_L_{n}"], rtol=0, atol=1e-13)
    for (m, n) in [(1, 1), (5, 8), (64, 32), (37, 45)]:
        L = np.asfortranarray(lapack[f"trsm_L_{m}_{n}"])
        B = np.asfortranarray(lapack[f"trsm_B_{m}_{n}"].copy())
        O.lib().oracle_dtrsm_rlt(m, n, O.ptr(L), n, O.ptr(B), m)
        np.testing.assert_allclose(B, lapack[f"trsm_X_{m}_{n}"], rtol=0, atol=1e-12)
    for (n, k) in [(1, 1), (8, 5), (32, 64), (45, 37)]:
        A = np.asfortranarray(lapack[f"syrk_A_{n}_{k}"])
        C = np.asfortranarray(lapack[f"syrk_C_{n}_{k}"].copy())
        O.lib().oracle_dsyrk_ln(n, k, -1.0, O.ptr(A), n, 1.0, O.ptr(C), n)
        want = lapack[f"syrk_out_{n}_{k}"]
        np.testing.assert_allclose(np.tril(C), np.tril(want), rtol=0, atol=1e-12)
        # strict upper triangle is not referenced
        np.testing.assert_array_equal(np.triu(C, 1), np.triu(lapack[f"syrk_C_{n}_{k}"], 1))
    for (m, n, k) in [(1, 1, 1), (8, 5, 3), (64, 32, 48), (37, 45, 29)]:
        A = np.asfortranarray(lapack[f"gemm_A_{m}_{n}_{k}"])
        B = np.asfortranarray(lapack[f"gemm_B_{m}_{n}_{k}"])
        C = np.asfortranarray(lapack[f"gemm_C_{m}_{n}_{k}"].copy())
        O.lib().oracle_dgemm_nt(m, n, k, -1.0, O.ptr(A), m, O.ptr(B), n, 1.0, O.ptr(C), m)
        np.testing.assert_allclose(C, lapack[f"gemm_out_{m}_{n}_{k}"], rtol=0, atol=1e-12)


def test_potrf_tile_info_nonspd(oracle):
    n = 20
    A = np.asfortranarray(np.eye(n))
    A[7, 7] = -1.0
    assert oracle.lib().oracle_dpotrf_tile(n, oracle.ptr(A), n) == 8
    A = np.asfortranarray(np.eye(n))
    A[0, 0] = np.nan
    assert oracle.lib().oracle_dpotrf_tile(n, oracle.ptr(A), n) == 1


@pytest.mark.parametrize("N,NB", [(160, 32), (257, 64), (512, 128)])
def test_tiled_dag_bitwise_equals_sequential(oracle, N, NB):
    t = oracle.plgsy_tiles(N, NB, N, 3872)
    a, b = t.copy(), t.copy()
    assert oracle.potrf_tiled(N, NB, a, 4) == 0
    assert oracle.potrf_tiled_seq(N, NB, b) == 0
    assert np.array_equal(a, b)  # schedule-independent (graph fixes the per-tile order)
    c = t.copy()
    assert oracle.potrf_tiled(N, NB, c, 1) == 0
    assert np.array_equal(a, c)


@pytest.mark.parametrize("N,NB", [(160, 32), (300, 64)])
def test_tiled_matches_numpy_cholesky(oracle, N, NB):
    t = oracle.spd_tiles(N, NB, 1.0, 7)
    D0 = oracle.tiles_to_dense(N, NB, t)
    L = t.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    Ln = np.linalg.cholesky(D0 + np.tril(D0, -1).T)
    Lo = np.tril(oracle.tiles_to_dense(N, NB, L))
    assert np.abs(Lo - Ln).max() / np.abs(Ln).max() < 1e-13


def test_task_counts_closed_form(oracle):
    for N, NB in [(160, 32), (512, 32)]:
        nt = N // NB
        t = oracle.spd_tiles(N, NB, 1.0, 1)
        assert oracle.potrf_tiled(N, NB, t, 2) == 0
        c = oracle.last_task_counts()
        assert list(c) == [nt, nt * (nt - 1) // 2, nt * (nt - 1) // 2, nt * (nt - 1) * (nt - 2) // 6]


def test_priorities_closed_form(oracle):
    nt = 16
    for K in range(nt):
        assert oracle.priority(oracle.POTRF, nt, K) == (nt - K) ** 3
        for M in range(K + 1, nt):
            assert oracle.priority(oracle.TRSM, nt, M, K) == (nt - M) ** 3 + 3 * (2 * nt - K - M - 1) * (M - K)
            assert oracle.priority(oracle.SYRK, nt, K, M) == (nt - K) ** 3 + 3 * (K - M)
            for Nn in range(K + 1, M):
                assert oracle.priority(oracle.GEMM, nt, M, Nn, K) == (
                    (nt - M) ** 3 + 3 * (2 * nt - M - Nn - 3) * (M - Nn) + 6 * (M - K))


def test_successor_keys(oracle):
    nt = 6
    s = oracle.successors(oracle.POTRF, nt, 2)
    assert s == [(-1, 2, 2, 0)] + [(oracle.TRSM, m, 2, 0) for m in range(3, nt)]
    s = oracle.successors(oracle.TRSM, nt, 4, 1)
    assert s == ([(-1, 4, 1, 0), (oracle.SYRK, 1, 4, 0)] + [(oracle.GEMM, 4, n, 1) for n in range(2, 4)] +
                 [(oracle.GEMM, m, 4, 1) for m in range(5, nt)])
    assert oracle.successors(oracle.SYRK, nt, 1, 2) == [(oracle.POTRF, 2, 0, 0)]
    assert oracle.successors(oracle.SYRK, nt, 1, 4) == [(oracle.SYRK, 2, 4, 0)]
    assert oracle.successors(oracle.GEMM, nt, 5, 2, 1) == [(oracle.TRSM, 5, 2, 0)]
    assert oracle.successors(oracle.GEMM, nt, 5, 3, 1) == [(oracle.GEMM, 5, 3, 2)]


def test_check_dpotrf_reference_case(oracle):
    """The reference's own ctest: N=160, NB=32, PLGSY(bump=N, seed=3872),
    check_dpotrf <= 60 (examples/CMakeLists.txt:17, testing_dpotrf.cc:277)."""
    N, NB = 160, 32
    A0 = oracle.plgsy_tiles(N, NB, N, 3872)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    r_dense = oracle.check_dpotrf_dense(oracle.tiles_to_dense(N, NB, L), oracle.tiles_to_dense(N, NB, A0))
    r_tiled = oracle.check_dpotrf_tiled(N, NB, L, A0)
    assert r_dense <= 60 and r_tiled <= 60
    # the two forms sum the rounding noise of L L^T in different orders; their
    # exact agreement is pinned on planted exact data (test_oracle_checks.py)
    assert r_tiled == pytest.approx(r_dense, rel=0.5, abs=0.5)


def test_check_dpotrf_catches_planted_error(oracle):
    N, NB = 160, 32
    A0 = oracle.plgsy_tiles(N, NB, N, 3872)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    oracle.tile_view(L, N, NB, 3, 1)[5, 7] += 1e-9  # one wrong entry in the factor
    assert oracle.check_dpotrf_tiled(N, NB, L, A0) > 60
    assert oracle.check_dpotrf_dense(oracle.tiles_to_dense(N, NB, L), oracle.tiles_to_dense(N, NB, A0)) > 60


def test_tiled_info_nonspd(oracle):
    N, NB = 128, 32
    t = oracle.spd_tiles(N, NB, 1.0, 3)
    oracle.tile_view(t, N, NB, 2, 2)[4, 4] = -1e3  # global diagonal 68 -> info 69
    assert oracle.potrf_tiled(N, NB, t.copy()) == 2 * NB + 5
    assert oracle.potrf_tiled_seq(N, NB, t.copy()) == 2 * NB + 5


def test_spd_generator_exact_integer(oracle):
    N, NB, alpha, seed = 48, 16, 0.25, 11
    t = oracle.spd_tiles(N, NB, alpha, seed)
    Bi = np.array([[oracle.lib().oracle_spd_bint(i, j, N, seed) for j in range(N)] for i in range(N)],
                  dtype=np.int64)
    assert Bi.min() >= -127 and Bi.max() <= 127
    S = Bi @ Bi.T
    want = S.astype(np.float64) * (1.0 / (127.0 * 127.0 * N))
    want[np.diag_indices(N)] += alpha
    D = oracle.tiles_to_dense(N, NB, t)
    assert np.array_equal(np.tril(D), np.tril(want))


def test_potrs_and_residual_vs_lapack(oracle, lapack):
    for (n, nrhs) in [(40, 1), (96, 3)]:
        A = lapack[f"potrs_A_{n}_{nrhs}"]
        NB = 32
        t = oracle.dense_to_tiles(n, NB, np.tril(A))
        L = t.copy()
        assert oracle.potrf_tiled(n, NB, L) == 0
        X = oracle.potrs_tiled(n, NB, L, lapack[f"potrs_B_{n}_{nrhs}"])
        np.testing.assert_allclose(X, lapack[f"potrs_X_{n}_{nrhs}"], rtol=0, atol=1e-12)
        R = oracle.residual_tiled(n, NB, t, X, lapack[f"potrs_B_{n}_{nrhs}"])
        np.testing.assert_allclose(R, lapack[f"potrs_B_{n}_{nrhs}"] - A @ X, rtol=0, atol=1e-13)
        assert oracle.solve_backward_error(n, NB, t, X, lapack[f"potrs_B_{n}_{nrhs}"]) < 1e-15
