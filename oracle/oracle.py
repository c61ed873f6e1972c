"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only the test suite, the smoke check and the benchmark's CPU-baseline leg use
this module.  It never backs the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_ull = ctypes.c_ulonglong
_p = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        sig = {
            "oracle_flops_dpotrf": (_d, [_d]),
            "oracle_rnd64_jump": (_ull, [_ull, _ull]),
            "oracle_core_plgsy": (None, [_d, _i, _i, _p, _i, _i, _i, _i, _ull]),
            "oracle_tile_index": (_l, [_i, _i, _i]),
            "oracle_num_tiles": (_i, [_i, _i]),
            "oracle_plgsy_tiled": (None, [_i, _i, _d, _ull, _p]),
            "oracle_tiles_to_dense": (None, [_i, _i, _p, _p]),
            "oracle_dense_to_tiles": (None, [_i, _i, _p, _p]),
            "oracle_spd_bint": (_i, [_l, _l, _l, _ull]),
            "oracle_spd_tiled": (None, [_i, _i, _d, _ull, _p]),
            "oracle_spd_uval": (_d, [_l, _l, _l, _ull]),
            "oracle_spd_uniform_tiled": (None, [_i, _i, _d, _ull, _p]),
            "oracle_dpotrf_tile": (_i, [_i, _p, _i]),
            "oracle_dtrsm_rlt": (None, [_i, _i, _p, _i, _p, _i]),
            "oracle_dsyrk_ln": (None, [_i, _i, _d, _p, _i, _d, _p, _i]),
            "oracle_dgemm_nt": (None, [_i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
            "oracle_priority": (_i, [_i, _i, _i, _i, _i]),
            "oracle_successors": (_i, [_i, _i, _i, _i, _i, _p, _i]),
            "oracle_dpotrf_tiled": (_i, [_i, _i, _p, _i]),
            "oracle_dpotrf_tiled_seq": (_i, [_i, _i, _p]),
            "oracle_last_task_counts": (None, [_p]),
            "oracle_check_dpotrf": (_d, [_p, _p, _i]),
            "oracle_check_dpotrf_tiled": (_d, [_i, _i, _p, _p, _i]),
            "oracle_dpotrs_tiled": (None, [_i, _i, _p, _i, _p, _i]),
            "oracle_dresidual_tiled": (None, [_i, _i, _p, _i, _p, _i, _p, _i, _p, _i]),
            "oracle_solve_backward_error": (_d, [_i, _i, _p, _i, _p, _i, _p, _i]),
            "oracle_norm_inf_sym_tiled": (_d, [_i, _i, _p]),
            "oracle_dtrsm_rln": (None, [_i, _i, _d, _p, _i, _p, _i]),
            "oracle_dtrsm_lln": (None, [_i, _i, _p, _i, _p, _i]),
            "oracle_dtrmm_llt": (None, [_i, _i, _p, _i, _p, _i]),
            "oracle_dsyrk_lt": (None, [_i, _i, _d, _p, _i, _d, _p, _i]),
            "oracle_dgemm_xn": (None, [_i, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
            "oracle_dtrtri_tile": (_i, [_i, _p, _i]),
            "oracle_dlauum_tile": (None, [_i, _p, _i]),
            "oracle_dtrtri_tiled": (_i, [_i, _i, _p]),
            "oracle_dlauum_tiled": (None, [_i, _i, _p]),
            "oracle_dpotri_tiled": (_i, [_i, _i, _p]),
            "oracle_check_dtrtri": (_d, [_p, _p, _i]),
            "oracle_flops_dtrtri": (_d, [_d]),
            "oracle_flops_dlauum": (_d, [_d]),
            "oracle_flops_dpotri": (_d, [_d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ref_lib():
    """oracle/_ref/libref_plgsy.so built from the reference's own headers, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_plgsy.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_core_plgsy.restype = None
        R.ref_core_plgsy.argtypes = [_d, _i, _i, _p, _i, _i, _i, _i, _ull]
        R.ref_rnd64_jump.restype = _ull
        R.ref_rnd64_jump.argtypes = [_ull, _ull]
        R.ref_check_norm.restype = _i
        R.ref_check_norm.argtypes = [_d, _d]
        _REF = R
    return _REF


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nt_of(N, NB):
    return (N + NB - 1) // NB


def alloc_tiles(N, NB, dtype=np.float64):
    nt = nt_of(N, NB)
    return np.zeros(nt * (nt + 1) // 2 * NB * NB, dtype=dtype)


def tile_view(tiles, N, NB, m, n):
    """Column-major NB x NB view of tile (m, n) (as a Fortran-ordered array)."""
    idx = lib().oracle_tile_index(m, n, nt_of(N, NB))
    return tiles[idx * NB * NB:(idx + 1) * NB * NB].reshape(NB, NB).T


def plgsy_tiles(N, NB, bump, seed):
    t = alloc_tiles(N, NB)
    lib().oracle_plgsy_tiled(N, NB, float(bump), seed, ptr(t))
    return t


def spd_tiles(N, NB, alpha, seed):
    t = alloc_tiles(N, NB)
    lib().oracle_spd_tiled(N, NB, float(alpha), seed, ptr(t))
    return t


def spd_uniform_tiles(N, NB, alpha, seed):
    t = alloc_tiles(N, NB)
    lib().oracle_spd_uniform_tiled(N, NB, float(alpha), seed, ptr(t))
    return t


def norm_inf_sym_tiled(N, NB, tiles):
    return lib().oracle_norm_inf_sym_tiled(N, NB, ptr(tiles))


def tiles_to_dense(N, NB, tiles):
    A = np.zeros((N, N), dtype=np.float64, order="F")
    lib().oracle_tiles_to_dense(N, NB, ptr(tiles), ptr(A))
    return A


def dense_to_tiles(N, NB, A):
    A = np.asfortranarray(A, dtype=np.float64)
    t = alloc_tiles(N, NB)
    lib().oracle_dense_to_tiles(N, NB, ptr(A), ptr(t))
    return t


def potrf_tiled(N, NB, tiles, nthreads=0):
    return lib().oracle_dpotrf_tiled(N, NB, ptr(tiles), nthreads)


def potrf_tiled_seq(N, NB, tiles):
    return lib().oracle_dpotrf_tiled_seq(N, NB, ptr(tiles))


def last_task_counts():
    c = np.zeros(4, dtype=np.int64)
    lib().oracle_last_task_counts(ptr(c))
    return c


def check_dpotrf_dense(Lfull, A0):
    Lfull = np.asfortranarray(Lfull, dtype=np.float64)
    A0 = np.asfortranarray(A0, dtype=np.float64)
    return lib().oracle_check_dpotrf(ptr(Lfull), ptr(A0), Lfull.shape[0])


def check_dpotrf_tiled(N, NB, Ltiles, A0tiles, nthreads=0):
    return lib().oracle_check_dpotrf_tiled(N, NB, ptr(Ltiles), ptr(A0tiles), nthreads)


def potrs_tiled(N, NB, Ltiles, B):
    B = np.array(B, dtype=np.float64, order="F", copy=True)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    lib().oracle_dpotrs_tiled(N, NB, ptr(Ltiles), B.shape[1], ptr(B), N)
    return B


def residual_tiled(N, NB, Atiles, X, B):
    X = np.asfortranarray(X, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    R = np.zeros_like(B, order="F")
    lib().oracle_dresidual_tiled(N, NB, ptr(Atiles), B.shape[1], ptr(X), N, ptr(B), N, ptr(R), N)
    return R


def solve_backward_error(N, NB, Atiles, X, B):
    X = np.asfortranarray(X, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    return lib().oracle_solve_backward_error(N, NB, ptr(Atiles), B.shape[1], ptr(X), N, ptr(B), N)


def successors(kind, nt, k0, k1=0, k2=0):
    out = np.zeros(4 * (2 * nt + 4), dtype=np.int32)
    n = lib().oracle_successors(kind, nt, k0, k1, k2, ptr(out), 2 * nt + 4)
    return [tuple(int(v) for v in out[4 * i:4 * i + 4]) for i in range(n)]


def priority(kind, nt, k0, k1=0, k2=0):
    return lib().oracle_priority(kind, nt, k0, k1, k2)


POTRF, TRSM, SYRK, GEMM = 0, 1, 2, 3


# ---- per-tile POINV operations (column-major numpy tiles, in place)


def _fc(a):
    if not (a.dtype == np.float64 and a.flags.f_contiguous):
        raise TypeError("tiles must be Fortran-ordered float64 arrays")
    return a


def tile_trsm_rln(L, B, alpha=-1.0):
    """B := alpha B L^{-1} (trtri_L.h:109-116)."""
    lib().oracle_dtrsm_rln(B.shape[0], B.shape[1], alpha, ptr(_fc(L)), L.shape[0], ptr(_fc(B)), B.shape[0])


def tile_trsm_lln(L, B):
    """B := L^{-1} B (trtri_L.h:64-71, alpha = 1)."""
    lib().oracle_dtrsm_lln(B.shape[0], B.shape[1], ptr(_fc(L)), L.shape[0], ptr(_fc(B)), B.shape[0])


def tile_trmm_llt(L, B):
    """B := L^T B (lauum.h:71-77, alpha = 1)."""
    lib().oracle_dtrmm_llt(B.shape[0], B.shape[1], ptr(_fc(L)), L.shape[0], ptr(_fc(B)), B.shape[0])


def tile_syrk_lt(A, C, alpha=1.0, beta=1.0):
    """lower(C) := alpha A^T A + beta C, A k x n (lauum.h:116-121)."""
    lib().oracle_dsyrk_lt(C.shape[0], A.shape[0], alpha, ptr(_fc(A)), A.shape[0], beta, ptr(_fc(C)), C.shape[0])


def tile_gemm_xn(ta, A, B, C, alpha=1.0, beta=1.0):
    """C := alpha op(A) B + beta C, op(A) = A^T if ta (trtri_L.h:179-185, lauum.h:160-166)."""
    k = A.shape[0] if ta else A.shape[1]
    lib().oracle_dgemm_xn(int(ta), C.shape[0], C.shape[1], k, alpha, ptr(_fc(A)), A.shape[0], ptr(_fc(B)),
                          B.shape[0], beta, ptr(_fc(C)), C.shape[0])


def tile_trtri(A):
    """lower(A) := lower(A)^{-1}; returns LAPACK info (trtri_L.h:32)."""
    return lib().oracle_dtrtri_tile(A.shape[0], ptr(_fc(A)), A.shape[0])


def tile_lauum(A):
    """lower(A) := L^T L (lauum.h:36)."""
    lib().oracle_dlauum_tile(A.shape[0], ptr(_fc(A)), A.shape[0])


def trtri_tiled(N, NB, tiles):
    """In place L := L^{-1} over lower tiles (examples/potrf/trtri_L.h order)."""
    return lib().oracle_dtrtri_tiled(N, NB, ptr(tiles))


def lauum_tiled(N, NB, tiles):
    """In place A := L^T L over lower tiles (examples/potrf/lauum.h order)."""
    lib().oracle_dlauum_tiled(N, NB, ptr(tiles))


def potri_tiled(N, NB, tiles):
    """In place A := (L L^T)^{-1} from the factor (examples/potrf/potri.h)."""
    return lib().oracle_dpotri_tiled(N, NB, ptr(tiles))


def check_dtrtri(L, Linv):
    """check_dtrtri of examples/potrf/testing_dtrtri.cc:269-325 (dense, lower)."""
    L = np.asfortranarray(L, dtype=np.float64)
    Linv = np.asfortranarray(Linv, dtype=np.float64)
    return lib().oracle_check_dtrtri(ptr(L), ptr(Linv), L.shape[0])
