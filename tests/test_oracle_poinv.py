"""The CPU oracle's POINV family (TRTRI, LAUUM, POTRI of examples/potrf/
trtri_L.h, lauum.h, potri.h) pinned against LAPACK/BLAS fixtures, closed
forms and invariants (no GPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "lapack_poinv.npz"))


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_flops_closed_forms(oracle):
    L = oracle.lib()
    for n in (1, 2, 32, 160, 4096, 65536):
        assert L.oracle_flops_dtrtri(n) == pytest.approx(n ** 3 / 3 + 2 * n / 3, rel=1e-14)
        assert L.oracle_flops_dpotri(n) == pytest.approx(2 * n ** 3 / 3 + n ** 2 / 2 + 5 * n / 6, rel=1e-14)
        assert L.oracle_flops_dlauum(n) == pytest.approx(n ** 3 / 3 + n ** 2 / 2 + n / 6, rel=1e-14)


def test_tile_ops_vs_blas(oracle, fx):
    L, p = oracle.lib(), oracle.ptr
    for (m, n) in [(1, 1), (5, 8), (37, 45)]:
        X = F(fx[f"trsm_rln_B_{m}_{n}"])
        L.oracle_dtrsm_rln(m, n, -1.0, p(F(fx[f"trsm_rln_L_{m}_{n}"])), n, p(X), m)
        assert rel(X, fx[f"trsm_rln_X_{m}_{n}"]) < 1e-13
        Lm = F(fx[f"trsm_lln_L_{m}_{n}"])
        Y = F(fx[f"trsm_lln_B_{m}_{n}"])
        L.oracle_dtrsm_lln(m, n, p(Lm), m, p(Y), m)
        assert rel(Y, fx[f"trsm_lln_X_{m}_{n}"]) < 1e-13
        Z = F(fx[f"trsm_lln_B_{m}_{n}"])
        L.oracle_dtrmm_llt(m, n, p(Lm), m, p(Z), m)
        assert rel(Z, fx[f"trmm_llt_X_{m}_{n}"]) < 1e-13
    for (n, k) in [(1, 1), (8, 5), (45, 37)]:
        C = F(fx[f"syrk_lt_C_{n}_{k}"])
        L.oracle_dsyrk_lt(n, k, 1.0, p(F(fx[f"syrk_lt_A_{n}_{k}"])), k, 1.0, p(C), n)
        assert rel(np.tril(C), fx[f"syrk_lt_out_{n}_{k}"]) < 1e-13
    for (m, n, k) in [(1, 1, 1), (8, 5, 3), (37, 45, 29)]:
        B = F(fx[f"gemm_xn_B_{m}_{n}_{k}"])
        C = F(fx[f"gemm_xn_C_{m}_{n}_{k}"])
        L.oracle_dgemm_xn(0, m, n, k, 1.0, p(F(fx[f"gemm_xn_A_{m}_{n}_{k}"])), m, p(B), k, 1.0, p(C), m)
        assert rel(C, fx[f"gemm_nn_out_{m}_{n}_{k}"]) < 1e-13
        C = F(fx[f"gemm_xn_C_{m}_{n}_{k}"])
        L.oracle_dgemm_xn(1, m, n, k, 1.0, p(F(fx[f"gemm_xn_At_{m}_{n}_{k}"])), k, p(B), k, 1.0, p(C), m)
        assert rel(C, fx[f"gemm_tn_out_{m}_{n}_{k}"]) < 1e-13


@pytest.mark.parametrize("n", [1, 7, 45])
def test_tile_trtri_lauum_vs_lapack(oracle, fx, n):
    A = F(fx[f"L_{n}"])
    assert oracle.lib().oracle_dtrtri_tile(n, oracle.ptr(A), n) == 0
    assert rel(np.tril(A), fx[f"trtri_{n}"]) < 1e-13
    B = F(fx[f"L_{n}"])
    oracle.lib().oracle_dlauum_tile(n, oracle.ptr(B), n)
    assert rel(np.tril(B), fx[f"lauum_{n}"]) < 1e-13


@pytest.mark.parametrize("N,NB", [(160, 32), (100, 32), (45, 16), (7, 4), (100, 100)])
def test_tiled_vs_lapack(oracle, fx, N, NB):
    L = fx[f"L_{N}"]
    t = oracle.dense_to_tiles(N, NB, L)
    assert oracle.trtri_tiled(N, NB, t) == 0
    assert rel(np.tril(oracle.tiles_to_dense(N, NB, t)), fx[f"trtri_{N}"]) < 1e-13
    t = oracle.dense_to_tiles(N, NB, L)
    oracle.lauum_tiled(N, NB, t)
    assert rel(np.tril(oracle.tiles_to_dense(N, NB, t)), fx[f"lauum_{N}"]) < 1e-13
    t = oracle.dense_to_tiles(N, NB, L)
    assert oracle.potri_tiled(N, NB, t) == 0
    assert rel(np.tril(oracle.tiles_to_dense(N, NB, t)), fx[f"potri_{N}"]) < 1e-13


def test_invariants(oracle, fx):
    """L Linv = I, lauum = L^T L, potri(L) A = I: closed forms independent of
    the task order."""
    N, NB = 160, 32
    L = fx[f"L_{N}"]
    t = oracle.dense_to_tiles(N, NB, L)
    oracle.trtri_tiled(N, NB, t)
    Li = np.tril(oracle.tiles_to_dense(N, NB, t))
    assert np.abs(L @ Li - np.eye(N)).max() < 1e-13
    t = oracle.dense_to_tiles(N, NB, L)
    oracle.lauum_tiled(N, NB, t)
    assert rel(np.tril(oracle.tiles_to_dense(N, NB, t)), np.tril(L.T @ L)) < 1e-14
    t = oracle.dense_to_tiles(N, NB, L)
    oracle.potri_tiled(N, NB, t)
    P = np.tril(oracle.tiles_to_dense(N, NB, t))
    P = P + np.tril(P, -1).T
    assert np.abs(P @ (L @ L.T) - np.eye(N)).max() < 1e-13


def test_check_dtrtri_and_planted_error(oracle, fx):
    N, NB = 160, 32
    L = fx[f"L_{N}"]
    t = oracle.dense_to_tiles(N, NB, L)
    oracle.trtri_tiled(N, NB, t)
    Li = np.tril(oracle.tiles_to_dense(N, NB, t))
    assert oracle.check_dtrtri(L, Li) <= 10.0
    bad = Li.copy()
    bad[90, 17] += 1e-9 * np.abs(Li).max()
    assert oracle.check_dtrtri(L, bad) > 10.0


def test_trtri_singular_info(oracle, fx):
    N, NB = 100, 32
    L = fx[f"L_{N}"].copy()
    L[70, 70] = 0.0
    t = oracle.dense_to_tiles(N, NB, L)
    before = t.copy()
    assert oracle.trtri_tiled(N, NB, t) == 71
    assert np.array_equal(t, before)
    assert oracle.potri_tiled(N, NB, t) == 71
