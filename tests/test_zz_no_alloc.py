"""The allocation-free per-tile TRSM (ttg_chol_dtrsm_rlt_noalloc) through
the C ABI: bitwise the result of ttg_chol_dtrsm_rlt (the same kernels; only
the block-inverse scratch differs), on two calls in a row, the second with a
smaller tile, and within the LAPACK bar of the exact solve; sizes beyond its
fixed scratch and an open launch group are refused."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_util import dev_colmajor, host_colmajor

pytestmark = pytest.mark.gpu


def _lower(n, rng):
    M = rng.uniform(-1, 1, (n, n))
    return np.linalg.cholesky(M @ M.T / n + np.eye(n))


def _call(L, name, stream, m, n, dL, dB):
    fn = getattr(L, name)
    return fn(stream, m, n, ctypes.c_void_p(dL.data_ptr()), dL.stride(1), ctypes.c_void_p(dB.data_ptr()),
              dB.stride(1))


def test_noalloc_trsm_matches_allocating_entry(dev):
    import ttg_amd.chol as C

    L = C.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(41)
    for (m, n) in [(512, 512), (200, 96)]:  # the second call is smaller
        print(f"dtrsm_rlt_noalloc m={m} n={n}", flush=True)
        Lh = _lower(n, rng)
        Bh = rng.uniform(-1, 1, (m, n))
        dL = dev_colmajor(Lh, dev)
        dB1 = dev_colmajor(Bh, dev)
        dB2 = dev_colmajor(Bh, dev)
        assert _call(L, "ttg_chol_dtrsm_rlt", stream, m, n, dL, dB1) == C.OK
        assert _call(L, "ttg_chol_dtrsm_rlt_noalloc", stream, m, n, dL, dB2) == C.OK, \
            L.ttg_chol_last_error().decode()
        got, ref = host_colmajor(dB2), host_colmajor(dB1)
        assert np.array_equal(got, ref)
        want = np.linalg.solve(Lh, Bh.T).T  # B L^{-T}
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_noalloc_trsm_refuses_what_it_cannot_serve(dev):
    import ttg_amd.chol as C

    L = C.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dL = torch.zeros(513, 513, dtype=torch.float64, device=dev).t()  # column-major
    dB = torch.zeros(513, 513, dtype=torch.float64, device=dev).t()
    assert _call(L, "ttg_chol_dtrsm_rlt_noalloc", stream, 64, 513, dL, dB) == C.EINVAL
    assert "scratch" in L.ttg_chol_last_error().decode()
    assert L.ttg_chol_group_begin(stream) == C.OK
    try:
        assert _call(L, "ttg_chol_dtrsm_rlt_noalloc", stream, 64, 64, dL, dB) == C.ESTATE
    finally:
        assert L.ttg_chol_group_end(stream) == C.OK
