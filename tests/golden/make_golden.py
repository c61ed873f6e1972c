"""Writes the golden fixtures of tests/golden/ (run in the development
container, where the reference tree and scipy are present):

  plgsy_ref.npz  -- PLGSY tiles and LCG skip-ahead values produced by the
                    reference's own generator headers (examples/potrf/
                    core_plgsy.h, random.h) compiled into oracle/_ref.
  lapack_tiles.npz -- FP64 tile-kernel results from LAPACK/BLAS through
                    scipy: dpotrf (lower), dtrsm (R,L,T,N), dsyrk (L,N),
                    dgemm (N,T), dpotrs, on small seeded inputs.
  lapack_poinv.npz -- the POINV family from LAPACK/BLAS through scipy:
                    dtrtri, dlauum, dpotri (lower, non-unit) of seeded
                    Cholesky factors, and the tile operations of the TRTRI
                    and LAUUM graphs: dtrsm (R,L,N,N), dtrsm (L,L,N,N),
                    dtrmm (L,L,T,N), dsyrk (L,T), dgemm (N,N) and (T,N).

Usage: python tests/golden/make_golden.py   (after `make -C oracle all ref`)
       python tests/golden/make_golden.py poinv   (lapack_poinv.npz only)
"""
import ctypes
import os
import sys

import numpy as np
import scipy.linalg.blas as sb
import scipy.linalg.lapack as sl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def plgsy_from_reference(N, NB, bump, seed):
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref_plgsy.so is not built (make -C oracle ref)")
    nt = O.nt_of(N, NB)
    tiles = O.alloc_tiles(N, NB)
    for n in range(nt):
        for m in range(n, nt):
            idx = O.lib().oracle_tile_index(m, n, nt)
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            R.ref_core_plgsy(float(bump), mb, nb, ctypes.c_void_p(tiles.ctypes.data + idx * NB * NB * 8), NB, N,
                             m * NB, n * NB, seed)
    return tiles


def main():
    R = O.ref_lib()
    out = {}
    # the reference ctest case (N=160, NB=32, bump=N, seed=3872) and a ragged one
    for (N, NB) in [(160, 32), (100, 32)]:
        out[f"plgsy_{N}_{NB}"] = plgsy_from_reference(N, NB, N, 3872)
    jumps = np.array([0, 1, 2, 3, 63, 64, 1000, 123456789, 2**40 + 17, 2**63 + 5], dtype=np.uint64)
    out["jump_n"] = jumps
    out["jump_seed3872"] = np.array([R.ref_rnd64_jump(int(n), 3872) for n in jumps], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "plgsy_ref.npz"), **out)

    rng = np.random.default_rng(20241203)
    lt = {}
    for n in (1, 7, 32, 45, 64):
        B = rng.uniform(-1, 1, (n, n))
        A = B @ B.T / n + np.eye(n)
        L, info = sl.dpotrf(A, lower=1, clean=1)
        assert info == 0
        lt[f"potrf_A_{n}"] = A
        lt[f"potrf_L_{n}"] = np.tril(L)
    for (m, n) in [(1, 1), (5, 8), (64, 32), (37, 45)]:
        B = rng.uniform(-1, 1, (n, n))
        A = B @ B.T / n + np.eye(n)
        L = np.tril(sl.dpotrf(A, lower=1, clean=1)[0])
        X = rng.uniform(-1, 1, (m, n))
        Y = sb.dtrsm(1.0, L, X, side=1, lower=1, trans_a=1, diag=0)
        lt[f"trsm_L_{m}_{n}"] = L
        lt[f"trsm_B_{m}_{n}"] = X
        lt[f"trsm_X_{m}_{n}"] = Y
    for (n, k) in [(1, 1), (8, 5), (32, 64), (45, 37)]:
        A = rng.uniform(-1, 1, (n, k))
        C = rng.uniform(-1, 1, (n, n))
        Cout = sb.dsyrk(-1.0, A, beta=1.0, c=C, lower=1)
        lt[f"syrk_A_{n}_{k}"] = A
        lt[f"syrk_C_{n}_{k}"] = C
        lt[f"syrk_out_{n}_{k}"] = Cout
    for (m, n, k) in [(1, 1, 1), (8, 5, 3), (64, 32, 48), (37, 45, 29)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (n, k))
        C = rng.uniform(-1, 1, (m, n))
        Cout = sb.dgemm(-1.0, A, B, beta=1.0, c=C, trans_b=1)
        lt[f"gemm_A_{m}_{n}_{k}"] = A
        lt[f"gemm_B_{m}_{n}_{k}"] = B
        lt[f"gemm_C_{m}_{n}_{k}"] = C
        lt[f"gemm_out_{m}_{n}_{k}"] = Cout
    for (n, nrhs) in [(40, 1), (96, 3)]:
        B = rng.uniform(-1, 1, (n, n))
        A = B @ B.T / n + np.eye(n)
        L = sl.dpotrf(A, lower=1, clean=1)[0]
        rhs = rng.uniform(-1, 1, (n, nrhs))
        X, info = sl.dpotrs(L, rhs, lower=1)
        assert info == 0
        lt[f"potrs_A_{n}_{nrhs}"] = A
        lt[f"potrs_B_{n}_{nrhs}"] = rhs
        lt[f"potrs_X_{n}_{nrhs}"] = X
    np.savez_compressed(os.path.join(HERE, "lapack_tiles.npz"), **lt)
    print("wrote", os.listdir(HERE))


def poinv():
    rng = np.random.default_rng(20261203)
    out = {}

    def factor(n):
        B = rng.uniform(-1, 1, (n, n))
        A = B @ B.T / n + np.eye(n)
        L, info = sl.dpotrf(A, lower=1, clean=1)
        assert info == 0
        return np.tril(L)

    for n in (1, 7, 45, 100, 160):
        L = factor(n)
        Li, info = sl.dtrtri(L, lower=1, unitdiag=0)
        assert info == 0
        U, info = sl.dlauum(L, lower=1)
        assert info == 0
        P, info = sl.dpotri(L, lower=1)
        assert info == 0
        out[f"L_{n}"] = L
        out[f"trtri_{n}"] = np.tril(Li)
        out[f"lauum_{n}"] = np.tril(U)
        out[f"potri_{n}"] = np.tril(P)
    for (m, n) in [(1, 1), (5, 8), (37, 45)]:
        L = factor(n)
        X = rng.uniform(-1, 1, (m, n))
        out[f"trsm_rln_L_{m}_{n}"] = L
        out[f"trsm_rln_B_{m}_{n}"] = X
        out[f"trsm_rln_X_{m}_{n}"] = sb.dtrsm(-1.0, L, X, side=1, lower=1, trans_a=0, diag=0)
        L2 = factor(m)
        Y = rng.uniform(-1, 1, (m, n))
        out[f"trsm_lln_L_{m}_{n}"] = L2
        out[f"trsm_lln_B_{m}_{n}"] = Y
        out[f"trsm_lln_X_{m}_{n}"] = sb.dtrsm(1.0, L2, Y, side=0, lower=1, trans_a=0, diag=0)
        out[f"trmm_llt_X_{m}_{n}"] = sb.dtrmm(1.0, L2, Y, side=0, lower=1, trans_a=1, diag=0)
    for (n, k) in [(1, 1), (8, 5), (45, 37)]:
        A = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (n, n))
        out[f"syrk_lt_A_{n}_{k}"] = A
        out[f"syrk_lt_C_{n}_{k}"] = C
        out[f"syrk_lt_out_{n}_{k}"] = np.tril(sb.dsyrk(1.0, A, beta=1.0, c=C, lower=1, trans=1))
    for (m, n, k) in [(1, 1, 1), (8, 5, 3), (37, 45, 29)]:
        A = rng.uniform(-1, 1, (m, k))
        At = rng.uniform(-1, 1, (k, m))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        out[f"gemm_xn_A_{m}_{n}_{k}"] = A
        out[f"gemm_xn_At_{m}_{n}_{k}"] = At
        out[f"gemm_xn_B_{m}_{n}_{k}"] = B
        out[f"gemm_xn_C_{m}_{n}_{k}"] = C
        out[f"gemm_nn_out_{m}_{n}_{k}"] = sb.dgemm(1.0, A, B, beta=1.0, c=C)
        out[f"gemm_tn_out_{m}_{n}_{k}"] = sb.dgemm(1.0, At, B, beta=1.0, c=C, trans_a=1)
    np.savez_compressed(os.path.join(HERE, "lapack_poinv.npz"), **out)
    print("wrote lapack_poinv.npz")


if __name__ == "__main__":
    if sys.argv[1:] == ["poinv"]:
        poinv()
    else:
        main()
        poinv()
