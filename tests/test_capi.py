"""The C-ABI library loads, exports every symbol include/ttg_chol.h declares,
and rejects bad arguments before touching the GPU (no GPU needed)."""
import ctypes
import re

import pytest

import ttg_amd.chol as C


def declared_symbols():
    text = open(C.HEADER).read()
    return sorted(set(re.findall(r"\b(ttg_chol_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers each one
    assert set(names) <= set(C.SIGNATURES), set(names) - set(C.SIGNATURES)


def test_version_and_error_strings():
    L = C.lib()
    assert b"gfx950" in L.ttg_chol_version()
    assert L.ttg_chol_last_error() is not None


@pytest.mark.parametrize("call", [
    lambda L: L.ttg_chol_dpotrf_tile(None, -1, None, 1, None, 0),
    lambda L: L.ttg_chol_dpotrf_tile(None, 600, ctypes.c_void_p(16), 600, None, 0),
    lambda L: L.ttg_chol_dpotrf_tile(None, 8, ctypes.c_void_p(16), 4, None, 0),
    lambda L: L.ttg_chol_dtrsm_rlt(None, 8, 8, None, 8, None, 8),
    lambda L: L.ttg_chol_dtrsm_rlt(None, 8, 8, ctypes.c_void_p(16), 7, ctypes.c_void_p(16), 8),
    lambda L: L.ttg_chol_dgemm_nt(None, 8, 8, 8, -1.0, ctypes.c_void_p(16), 4, ctypes.c_void_p(16), 8, 1.0,
                                  ctypes.c_void_p(16), 8),
    lambda L: L.ttg_chol_dsyrk_ln(None, -2, 8, -1.0, None, 8, 1.0, None, 8),
    lambda L: L.ttg_chol_dpotrf_tiled(None, 100, 0, None, None, 0),
    lambda L: L.ttg_chol_dpotrf_tiled(None, 100, 48, ctypes.c_void_p(16), None, 0),
    lambda L: L.ttg_chol_dtrtri_tiled(None, 100, 0, None, None),
    lambda L: L.ttg_chol_dtrtri_tiled(None, 100, 48, ctypes.c_void_p(16), None),
    lambda L: L.ttg_chol_dlauum_tiled(None, 100, 1024, ctypes.c_void_p(16)),
    lambda L: L.ttg_chol_dpotri_tiled(None, 100, 64, None, None),
])
def test_bad_arguments_rejected_without_launch(call):
    L = C.lib()
    rc = call(L)
    assert rc in (C.EINVAL, C.ENOTSUP)
    assert L.ttg_chol_last_error().decode() != ""


def test_group_end_without_begin_is_state_error():
    L = C.lib()
    assert L.ttg_chol_group_end(None) == C.ESTATE


def test_no_cpu_fallback():
    import torch

    with pytest.raises(C.CholError):
        C.dpotrf_tile(torch.eye(4, dtype=torch.float64))


def test_graph_library_exports_every_declared_symbol():
    import ttg_amd.graph as G

    text = open(G.HEADER).read()
    names = sorted(set(re.findall(r"\b(ttg_graph_\w+)\s*\(", text)))
    assert names == sorted(G.SIGNATURES)
    L = G.lib()
    assert all(hasattr(L, n) for n in names)


def test_graph_rejects_bad_arguments_without_gpu():
    import numpy as np

    import ttg_amd.graph as G

    t = np.zeros(4, dtype=np.float64)
    assert G.lib().ttg_graph_dpotrf(0, 32, t.ctypes.data, None, None) == -1
    assert G.lib().ttg_graph_dpotrf(100, 1024, t.ctypes.data, None, None) == -1
    assert G.lib().ttg_graph_dplgsy(100, 0, 1.0, 1, t.ctypes.data) == -1


def test_batched_rejects_bad_sizes_without_launch():
    import ctypes

    L = C.lib()
    n = (ctypes.c_int * 2)(64, 2000)
    lda = (ctypes.c_int * 2)(64, 2000)
    ptrs = (ctypes.c_void_p * 2)(16, 16)
    assert L.ttg_chol_sposv_batched(None, 2, n, ptrs, lda, ptrs, ctypes.c_void_p(16), -1) == C.EINVAL
    n2 = (ctypes.c_int * 1)(64)
    lda2 = (ctypes.c_int * 1)(32)
    p1 = (ctypes.c_void_p * 1)(16)
    assert L.ttg_chol_sposv_batched(None, 1, n2, p1, lda2, p1, ctypes.c_void_p(16), -1) == C.EINVAL
    assert L.ttg_chol_sposv_batched(None, 1, n2, p1, n2, p1, ctypes.c_void_p(16), 99) == C.EINVAL
    assert L.ttg_chol_sposv_batched(None, 0, None, None, None, None, None, -1) == C.OK
    assert L.ttg_chol_batched_num_configs() >= 2
