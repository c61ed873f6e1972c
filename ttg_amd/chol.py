"""ctypes binding of libttg_chol.so (the C ABI of include/ttg_chol.h).

Device buffers are torch CUDA tensors (HIP on ROCm); streams are torch
streams.  Every call goes to the HIP library; nothing here computes on the
CPU.  A missing library or a failed launch raises :class:`CholError`.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttg_chol.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ttg_chol.h")

OK, EINVAL, EHIP, ENOTSUP, ENOMEM, ESTATE, ENOCONV = 0, -1, -2, -3, -4, -5, -6
PREC_FP64, PREC_FP32, PREC_F16X2, PREC_F16, PREC_BF16 = 0, 1, 2, 3, 4
F_LOOKAHEAD, F_GRAPH, F_PROFILE = 0x1, 0x2, 0x4
PROF_UPDATE, PROF_TRSM, PROF_POTRF = 0, 1, 2
F_DEFAULT = F_LOOKAHEAD | F_GRAPH
MAX_TILE = 512

_d, _i, _u, _ull, _p = ctypes.c_double, ctypes.c_int, ctypes.c_uint, ctypes.c_ulonglong, ctypes.c_void_p

# name -> (restype, argtypes); must match include/ttg_chol.h
SIGNATURES = {
    "ttg_chol_version": (ctypes.c_char_p, []),
    "ttg_chol_last_error": (ctypes.c_char_p, []),
    "ttg_chol_num_devices": (_i, []),
    "ttg_chol_dpotrf_tile": (_i, [_p, _i, _p, _i, _p, _i]),
    "ttg_chol_dtrsm_rlt": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_dsyrk_ln": (_i, [_p, _i, _i, _d, _p, _i, _d, _p, _i]),
    "ttg_chol_dgemm_nt": (_i, [_p, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "ttg_chol_dnrm2": (_i, [_p, _i, _p, _p]),
    "ttg_chol_group_begin": (_i, [_p]),
    "ttg_chol_group_end": (_i, [_p]),
    "ttg_chol_group_last_launches": (_i, []),
    "ttg_chol_dpotrf_tiled": (_i, [_p, _i, _i, _p, _p, _u]),
    "ttg_chol_dpotrs_tiled": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_dresidual_tiled": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _i, _p, _i]),
    "ttg_chol_dtrtri_tiled": (_i, [_p, _i, _i, _p, _p]),
    "ttg_chol_release_plans": (_i, []),
    "ttg_chol_dlauum_tiled": (_i, [_p, _i, _i, _p]),
    "ttg_chol_dpotri_tiled": (_i, [_p, _i, _i, _p, _p]),
    "ttg_chol_plgsy_tiled": (_i, [_p, _i, _i, _d, _ull, _p]),
    "ttg_chol_dplgsy_tile": (_i, [_p, _d, _i, _i, _p, _i, _i, _i, _i, _ull]),
    "ttg_chol_spd_tiled": (_i, [_p, _i, _i, _d, _ull, _p, _p]),
    "ttg_chol_selftest_mfma_layout": (_i, [_p, _i, _p]),
    "ttg_chol_profile_read": (_i, [_i, _p, _p, _p]),
    "ttg_chol_profile_enable": (_i, [_i]),
    "ttg_chol_dposv_mixed": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _i, _i, _d, _i, _p, _p, _p]),
    "ttg_chol_sposv_batched": (_i, [_p, _i, _p, _p, _p, _p, _p, _i]),
    "ttg_chol_batched_num_configs": (_i, []),
    "ttg_chol_batched_config": (_i, [_i, _p, _p]),
}


class CholError(RuntimeError):
    pass


_LIB = None


def lib():
    """Load libttg_chol.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CholError(f"{LIB_PATH} is not built; run `python -m ttg_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != OK:
        msg = lib().ttg_chol_last_error().decode()
        raise CholError(f"{what} failed ({rc}): {msg}")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _require_cuda(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise CholError("ttg_amd.chol works on device tensors only (no CPU path)")


def _ld(t):
    """Leading dimension of a column-major 2-D tensor (a transposed row-major one)."""
    if t.dim() == 1:
        return max(1, t.shape[0])
    if t.shape[0] > 1 and t.stride(0) != 1:
        raise CholError("tensor must be column-major (stride(0) == 1)")
    if t.shape[1] == 1:
        return max(1, t.shape[0])
    return max(1, t.stride(1))


def colmajor(t):
    """Column-major copy of a 2-D tensor (a transposed view of a row-major tensor)."""
    return t.t().contiguous().t()


# ---------------------------------------------------------------- per tile


def dpotrf_tile(A, info=None, info_offset=0, stream=None):
    _require_cuda(A, info)
    n = A.shape[0]
    _check(lib().ttg_chol_dpotrf_tile(_stream(stream), n, _ptr(A), _ld(A), _ptr(info), info_offset),
           "dpotrf_tile")


def dtrsm_rlt(L, B, stream=None):
    _require_cuda(L, B)
    m, n = B.shape
    _check(lib().ttg_chol_dtrsm_rlt(_stream(stream), m, n, _ptr(L), _ld(L), _ptr(B), _ld(B)), "dtrsm_rlt")


def dsyrk_ln(A, C, alpha=-1.0, beta=1.0, stream=None):
    _require_cuda(A, C)
    n, k = A.shape
    _check(lib().ttg_chol_dsyrk_ln(_stream(stream), n, k, alpha, _ptr(A), _ld(A), beta, _ptr(C), _ld(C)),
           "dsyrk_ln")


def dgemm_nt(A, B, C, alpha=-1.0, beta=1.0, stream=None):
    _require_cuda(A, B, C)
    m, k = A.shape
    n = B.shape[0]
    _check(lib().ttg_chol_dgemm_nt(_stream(stream), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                   _ptr(C), _ld(C)), "dgemm_nt")


def dnrm2(x, out, stream=None):
    _require_cuda(x, out)
    _check(lib().ttg_chol_dnrm2(_stream(stream), x.numel(), _ptr(x), _ptr(out)), "dnrm2")


class group:
    """Context manager: per-tile calls on `stream` inside are launched as
    batched kernels when the block exits (ttg_chol_group_begin/end)."""

    def __init__(self, stream=None):
        self.stream = stream

    def __enter__(self):
        _check(lib().ttg_chol_group_begin(_stream(self.stream)), "group_begin")
        return self

    def __exit__(self, *exc):
        _check(lib().ttg_chol_group_end(_stream(self.stream)), "group_end")
        return False


def group_last_launches():
    return lib().ttg_chol_group_last_launches()


# ---------------------------------------------------------------- tiled


def nt_of(N, NB):
    return (N + NB - 1) // NB


def num_tiles(N, NB):
    nt = nt_of(N, NB)
    return nt * (nt + 1) // 2


def tile_index(m, n, nt):
    return n * nt - n * (n - 1) // 2 + (m - n)


def alloc_tiles(N, NB, dtype=None, device="cuda"):
    import torch

    return torch.zeros(num_tiles(N, NB) * NB * NB, dtype=dtype or torch.float64, device=device)


def dpotrf_tiled(tiles, N, NB, info=None, flags=F_DEFAULT, stream=None):
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dpotrf_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info), flags), "dpotrf_tiled")


def dpotrs_tiled(tiles, N, NB, B, stream=None):
    _require_cuda(tiles, B)
    nrhs = 1 if B.dim() == 1 else B.shape[1]
    _check(lib().ttg_chol_dpotrs_tiled(_stream(stream), N, NB, _ptr(tiles), nrhs, _ptr(B), _ld(B)), "dpotrs_tiled")


def dresidual_tiled(tiles, N, NB, X, B, R, stream=None):
    _require_cuda(tiles, X, B, R)
    nrhs = 1 if B.dim() == 1 else B.shape[1]
    _check(lib().ttg_chol_dresidual_tiled(_stream(stream), N, NB, _ptr(tiles), nrhs, _ptr(X), _ld(X), _ptr(B),
                                          _ld(B), _ptr(R), _ld(R)), "dresidual_tiled")


def release_plans():
    """Free every cached plan and workspace of the library (synchronises the
    device); the next call on a matrix builds its plan again."""
    _check(lib().ttg_chol_release_plans(), "release_plans")


def dtrtri_tiled(tiles, N, NB, info=None, stream=None):
    """tiles := L^{-1} (lower tiles, in place); the TRTRI graph of
    examples/potrf/trtri_L.h.  info (device int32 tensor) receives 0 or the
    1-based index of the first zero diagonal entry."""
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dtrtri_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info)), "dtrtri_tiled")


def dlauum_tiled(tiles, N, NB, stream=None):
    """tiles := L^T L (lower part, in place); the LAUUM graph of
    examples/potrf/lauum.h."""
    _require_cuda(tiles)
    _check(lib().ttg_chol_dlauum_tiled(_stream(stream), N, NB, _ptr(tiles)), "dlauum_tiled")


def dpotri_tiled(tiles, N, NB, info=None, stream=None):
    """tiles := A^{-1} (lower part, in place) from the Cholesky factor of A;
    the POTRI graph of examples/potrf/potri.h (TRTRI then LAUUM)."""
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dpotri_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info)), "dpotri_tiled")


def selftest_mfma_layout(kind, out, stream=None):
    _require_cuda(out)
    _check(lib().ttg_chol_selftest_mfma_layout(_stream(stream), kind, _ptr(out)), "selftest_mfma_layout")


def profile_read(kind):
    """(total ms, launches, algorithmic flops) of the profiled launches of `kind`."""
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    fl = ctypes.c_double()
    _check(lib().ttg_chol_profile_read(kind, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl)), "profile_read")
    return ms.value, n.value, fl.value


def dposv_mixed(A, N, NB, B, X, prec=PREC_FP32, tol=1e-12, max_iter=30, info=None, stream=None, check=True):
    """Mixed-precision solve with FP64 refinement; returns (iterations, backward error)."""
    import torch

    _require_cuda(A, B, X)
    if info is None:
        info = torch.zeros(1, dtype=torch.int32, device=A.device)
    nrhs = 1 if B.dim() == 1 else B.shape[1]
    it = ctypes.c_int()
    be = ctypes.c_double()
    rc = lib().ttg_chol_dposv_mixed(_stream(stream), N, NB, _ptr(A), nrhs, _ptr(B), _ld(B), _ptr(X), _ld(X), prec,
                                    tol, max_iter, _ptr(info), ctypes.byref(it), ctypes.byref(be))
    if rc == ENOCONV and not check:
        return it.value, be.value
    _check(rc, "dposv_mixed")
    return it.value, be.value


def profile_enable(on):
    _check(lib().ttg_chol_profile_enable(1 if on else 0), "profile_enable")


def dplgsy_tile(A, bump, gM, m0, n0, seed, stream=None):
    """One CORE_dplgsy tile (column-major device tensor A) at global offset (m0, n0)."""
    _require_cuda(A)
    m, n = A.shape
    _check(lib().ttg_chol_dplgsy_tile(_stream(stream), float(bump), m, n, _ptr(A), _ld(A), gM, m0, n0, seed),
           "dplgsy_tile")


def plgsy_tiled(tiles, N, NB, bump, seed, stream=None):
    _require_cuda(tiles)
    _check(lib().ttg_chol_plgsy_tiled(_stream(stream), N, NB, float(bump), seed, _ptr(tiles)), "plgsy_tiled")


def spd_tiled(tiles, N, NB, alpha, seed, work, stream=None):
    _require_cuda(tiles, work)
    _check(lib().ttg_chol_spd_tiled(_stream(stream), N, NB, float(alpha), seed, _ptr(tiles), _ptr(work)),
           "spd_tiled")


class Batch:
    """A batch of FP32 SPD systems for sposv_batched: host arrays of sizes,
    leading dimensions and device pointers, built once and reusable."""

    def __init__(self, mats, rhs):
        if len(mats) != len(rhs):
            raise ValueError("one right-hand side per matrix")
        self.count = len(mats)
        n = [int(m.shape[0]) for m in mats]
        self.n = (ctypes.c_int * self.count)(*n)
        self.lda = (ctypes.c_int * self.count)(*n)
        self.A = (ctypes.c_void_p * self.count)(*[m.data_ptr() for m in mats])
        self.b = (ctypes.c_void_p * self.count)(*[v.data_ptr() for v in rhs])

    @classmethod
    def packed(cls, sizes, A, b):
        """Matrices packed back to back in one float32 tensor A (each n_i^2,
        column-major, lda = n_i) and right-hand sides in b (each n_i)."""
        self = cls.__new__(cls)
        self.count = len(sizes)
        self.n = (ctypes.c_int * self.count)(*sizes)
        self.lda = (ctypes.c_int * self.count)(*sizes)
        a0, b0 = A.data_ptr(), b.data_ptr()
        offa, offb, pa, pb = 0, 0, [], []
        for s in sizes:
            pa.append(a0 + 4 * offa)
            pb.append(b0 + 4 * offb)
            offa += s * s
            offb += s
        if offa > A.numel() or offb > b.numel():
            raise ValueError("packed batch larger than its buffers")
        self.A = (ctypes.c_void_p * self.count)(*pa)
        self.b = (ctypes.c_void_p * self.count)(*pb)
        return self


def sposv_batched(batch, info, cfg=-1, stream=None):
    """Factor + solve every system of `batch` in one launch (info: int32 device
    tensor of batch.count entries)."""
    _require_cuda(info)
    _check(lib().ttg_chol_sposv_batched(_stream(stream), batch.count, batch.n, batch.A, batch.lda, batch.b,
                                        _ptr(info), cfg), "sposv_batched")


def batched_configs():
    """[(waves per workgroup, workgroups per CU)] of the batched launch configurations."""
    out = []
    for c in range(lib().ttg_chol_batched_num_configs()):
        w, g = ctypes.c_int(), ctypes.c_int()
        _check(lib().ttg_chol_batched_config(c, ctypes.byref(w), ctypes.byref(g)), "batched_config")
        out.append((w.value, g.value))
    return out
