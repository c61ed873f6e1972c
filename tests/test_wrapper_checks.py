"""The Python wrappers reject tensors whose dtype, size, layout or device
would let a kernel read or write past an allocation (the C ABI takes raw
pointers and cannot see sizes).  The checks run before the device check, so
CPU tensors exercise them here; no GPU is needed."""
import numpy as np
import pytest
import torch

import ttg_amd.chol as C


def tiles(N, NB, dtype=torch.float64, short=0):
    return torch.zeros(C.num_tiles(N, NB) * NB * NB - short, dtype=dtype)


def cols(N, nrhs, dtype=torch.float64):
    return torch.zeros(nrhs, N, dtype=dtype).t()  # column-major N x nrhs


def raises(fn, *words):
    with pytest.raises(C.CholError) as e:
        fn()
    msg = str(e.value)
    for w in words:
        assert w in msg, msg
    return msg


def test_tiled_drivers_reject_short_or_mistyped_tiles():
    N, NB = 300, 64
    raises(lambda: C.dpotrf_tiled(tiles(N, NB, short=1), N, NB), "needs")
    raises(lambda: C.dpotrf_tiled(tiles(N, NB, torch.float32), N, NB), "float64")
    raises(lambda: C.dtrtri_tiled(tiles(N, NB, short=NB), N, NB), "needs")
    raises(lambda: C.dlauum_tiled(tiles(N, NB, torch.float32), N, NB), "float64")
    raises(lambda: C.dpotri_tiled(tiles(N, NB, short=7), N, NB), "needs")
    raises(lambda: C.plgsy_tiled(tiles(N, NB, short=1), N, NB, N, 3872), "needs")
    raises(lambda: C.dpotrf_tiled(tiles(N, NB)[::2], N // 2, NB), "contiguous")
    raises(lambda: C.dpotrf_tiled(tiles(N, NB), -1, NB), "bad sizes")
    # well-formed host tensors reach the device check
    raises(lambda: C.dpotrf_tiled(tiles(N, NB), N, NB), "device tensors only")


def test_info_must_be_int32():
    N, NB = 128, 64
    raises(lambda: C.dpotrf_tiled(tiles(N, NB), N, NB, torch.zeros(1, dtype=torch.int64)), "int32")
    raises(lambda: C.dpotrf_tiled(tiles(N, NB), N, NB, torch.zeros(0, dtype=torch.int32)), "at least 1")


def test_solves_reject_wrong_right_hand_sides():
    N, NB = 200, 64
    t = tiles(N, NB)
    raises(lambda: C.dpotrs_tiled(t, N, NB, cols(N - 1, 2)), "N=200 rows")
    raises(lambda: C.dpotrs_tiled(t, N, NB, cols(N, 2, torch.float32)), "float64")
    raises(lambda: C.dpotrs_tiled(t, N, NB, torch.zeros(N, 2, dtype=torch.float64)), "column-major")
    raises(lambda: C.dresidual_tiled(t, N, NB, cols(N, 3), cols(N, 2), cols(N, 2)), "3 columns")
    raises(lambda: C.dresidual_tiled(t, N, NB, cols(N, 2), cols(N, 2), cols(N, 1)), "1 columns")
    raises(lambda: C.dposv_mixed(t, N, NB, cols(N, 2), cols(N, 1)), "1 columns")
    raises(lambda: C.dposv_mixed(tiles(N, NB, short=1), N, NB, cols(N, 2), cols(N, 2)), "needs")
    # columns that overlap (leading dimension below N)
    X = torch.as_strided(torch.zeros(2 * N, dtype=torch.float64), (N, 2), (1, N - 8))
    raises(lambda: C.dpotrs_tiled(t, N, NB, X), "leading dimension")


def test_per_tile_calls_check_shapes():
    def cm(r, c, dtype=torch.float64):
        return torch.zeros(c, r, dtype=dtype).t()  # column-major r x c

    A = cm(64, 64)
    raises(lambda: C.dpotrf_tile(cm(64, 32)), "expected (64, 64)")
    raises(lambda: C.dpotrf_tile(cm(64, 64, torch.float32)), "float64")
    raises(lambda: C.dpotrf_tile(torch.zeros(64, 64, dtype=torch.float64)), "column-major")
    raises(lambda: C.dtrsm_rlt(cm(32, 32), A), "expected (64, 64)")
    raises(lambda: C.dgemm_nt(A, cm(64, 32), A), "expected (None, 64)")
    raises(lambda: C.dgemm_nt(A, A, cm(64, 32)), "expected (64, 64)")
    raises(lambda: C.dsyrk_ln(A, cm(32, 32)), "expected (64, 64)")
    raises(lambda: C.dpotrf_tile(A), "device tensors only")
    # a norm takes any dense tile, row- or column-major, but no strided gaps
    raises(lambda: C.dnrm2(A, torch.zeros(1, dtype=torch.float64)), "device tensors only")
    raises(lambda: C.dnrm2(torch.zeros(64, 64, dtype=torch.float64)[:, :8], torch.zeros(1, dtype=torch.float64)),
           "dense")


def test_generators_and_checks_need_full_buffers():
    N, NB = 256, 64
    raises(lambda: C.spd_tiled(tiles(N, NB), N, NB, 1.0, 1, torch.zeros(N * N - 1, dtype=torch.float64)), "N*N")
    raises(lambda: C.spd_uniform_tiled(tiles(N, NB, short=1), N, NB, 1.0, 1,
                                       torch.zeros(N * N, dtype=torch.float64)), "needs")
    raises(lambda: C.dcheck_dpotrf_tiled(tiles(N, NB), tiles(N, NB, short=1), N, NB), "needs")
    raises(lambda: C.dposv_mixed_factor(tiles(N, NB, torch.float32), N, NB, 1, C.PREC_FP32), "float64")
    raises(lambda: C.selftest_mfma_layout(2, torch.zeros(64 * 64 * 4, dtype=torch.float64)), "needs")


def test_batched_checks_dtypes_and_info():
    A = torch.zeros(64 * 64 + 32 * 32, dtype=torch.float32)
    b = torch.zeros(96, dtype=torch.float32)
    raises(lambda: C.Batch.packed([64, 32], A.double(), b), "float32")
    raises(lambda: C.Batch.packed([64, 33], A, b), "larger than its buffers")
    raises(lambda: C.Batch.packed([64, 0], A, b), "n >= 1")
    raises(lambda: C.Batch([torch.zeros(8, 4)], [torch.zeros(8)]), "square")
    raises(lambda: C.Batch([torch.zeros(8, 8)], [torch.zeros(7)]), "8 contiguous")
    batch = C.Batch.packed([64, 32], A, b)
    raises(lambda: C.sposv_batched(batch, torch.zeros(1, dtype=torch.int32)), "at least 2")
    raises(lambda: C.sposv_batched(batch, None), "info")


def test_graph_binding_refuses_short_host_tiles():
    import ttg_amd.graph as G

    N, NB = 300, 64
    nt = (N + NB - 1) // NB
    short = np.zeros(nt * (nt + 1) // 2 * NB * NB - 1)
    with pytest.raises(ValueError, match="needs"):
        G.dpotrf(N, NB, short)
    with pytest.raises(ValueError, match="needs"):
        G.dplgsy(N, NB, float(N), 3872, short)
    with pytest.raises(TypeError):
        G.dpotrf(N, NB, short.astype(np.float32))
