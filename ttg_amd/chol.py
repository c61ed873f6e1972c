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
    "ttg_chol_dtrsm_rlt_noalloc": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_dsyrk_ln": (_i, [_p, _i, _i, _d, _p, _i, _d, _p, _i]),
    "ttg_chol_dgemm_nt": (_i, [_p, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "ttg_chol_dnrm2": (_i, [_p, _i, _p, _p]),
    "ttg_chol_group_begin": (_i, [_p]),
    "ttg_chol_group_end": (_i, [_p]),
    "ttg_chol_group_last_launches": (_i, []),
    "ttg_chol_group_last_splits": (_i, []),
    "ttg_chol_dpotrf_tiled": (_i, [_p, _i, _i, _p, _p, _u]),
    "ttg_chol_dpotrs_tiled": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_dresidual_tiled": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _i, _p, _i]),
    "ttg_chol_dtrtri_tiled": (_i, [_p, _i, _i, _p, _p]),
    "ttg_chol_release_plans": (_i, []),
    "ttg_chol_set_knob": (_i, [_i, _i]),
    "ttg_chol_spotrf_tile": (_i, [_p, _i, _p, _i, _p, _i]),
    "ttg_chol_strsm_rlt": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_ssyrk_ln": (_i, [_p, _i, _i, ctypes.c_float, _p, _i, ctypes.c_float, _p, _i]),
    "ttg_chol_sgemm_nt": (_i, [_p, _i, _i, _i, ctypes.c_float, _p, _i, _p, _i, ctypes.c_float, _p, _i]),
    "ttg_chol_spotrs_tiled": (_i, [_p, _i, _i, _p, _i, _p, _i]),
    "ttg_chol_dtrtri_tile": (_i, [_p, _i, _p, _i, _p, _i]),
    "ttg_chol_dtrsm_lln": (_i, [_p, _i, _i, _d, _p, _i, _p, _i]),
    "ttg_chol_dtrsm_rln": (_i, [_p, _i, _i, _d, _p, _i, _p, _i]),
    "ttg_chol_dgemm_nn": (_i, [_p, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "ttg_chol_dlauum_tile": (_i, [_p, _i, _p, _i]),
    "ttg_chol_dtrmm_llt": (_i, [_p, _i, _i, _d, _p, _i, _p, _i]),
    "ttg_chol_dsyrk_lt": (_i, [_p, _i, _i, _d, _p, _i, _d, _p, _i]),
    "ttg_chol_dgemm_tn": (_i, [_p, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]),
    "ttg_chol_get_knob": (_i, [_i]),
    "ttg_chol_dlauum_tiled": (_i, [_p, _i, _i, _p]),
    "ttg_chol_dpotri_tiled": (_i, [_p, _i, _i, _p, _p]),
    "ttg_chol_plgsy_tiled": (_i, [_p, _i, _i, _d, _ull, _p]),
    "ttg_chol_dplgsy_tile": (_i, [_p, _d, _i, _i, _p, _i, _i, _i, _i, _ull]),
    "ttg_chol_spd_tiled": (_i, [_p, _i, _i, _d, _ull, _p, _p]),
    "ttg_chol_spd_uniform_tiled": (_i, [_p, _i, _i, _d, _ull, _p, _p]),
    "ttg_chol_dposv_mixed_factor": (_i, [_p, _i, _i, _i, _i, _p]),
    "ttg_chol_dcheck_dpotrf_tiled": (_i, [_p, _i, _i, _p, _p, _p]),
    "ttg_chol_selftest_mfma_layout": (_i, [_p, _i, _p]),
    "ttg_chol_profile_read": (_i, [_i, _p, _p, _p]),
    "ttg_chol_profile_enable": (_i, [_i]),
    "ttg_chol_dposv_mixed": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _i, _i, _d, _i, _p, _p, _p]),
    "ttg_chol_dposv_krylov": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _i, _i, _d, _d, _i, _p, _p, _p]),
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
    """Every tensor on the GPU, and on the one device the call runs on."""
    import torch

    devs = set()
    for t in ts:
        if t is None:
            continue
        if not t.is_cuda:
            raise CholError("ttg_amd.chol works on device tensors only (no CPU path)")
        devs.add(t.device.index)
    if len(devs) > 1:
        raise CholError(f"tensors on different devices {sorted(devs)}")
    if devs and devs.pop() != torch.cuda.current_device():
        raise CholError("tensors are not on the current device (the stream's device)")


# ---- argument checks: the C ABI takes raw pointers and cannot see sizes, so
# a short or mistyped tensor is rejected here, before any kernel could read
# or write past its allocation.  They run before the device check, so they
# hold for any tensor.


def _dtype_name(t):
    return str(t.dtype).replace("torch.", "")


def _need_dtype(t, want, name):
    import torch

    want_t = {"f64": torch.float64, "f32": torch.float32, "i32": torch.int32}[want]
    if t.dtype != want_t:
        raise CholError(f"{name} must be {str(want_t).replace('torch.', '')}, got {_dtype_name(t)}")


def _span(t):
    """Elements from the first to one past the last element a strided view reaches."""
    if t.numel() == 0:
        return 0
    return 1 + sum((d - 1) * st for d, st in zip(t.shape, t.stride()))


def _need_storage(t, elems, name):
    """t's storage holds `elems` elements from t's first element."""
    avail = t.untyped_storage().nbytes() // t.element_size() - t.storage_offset()
    if avail < elems:
        raise CholError(f"{name} holds {avail} elements from its first one, needs {elems}")


def _check_tiles(t, N, NB, name="tiles", dtype="f64"):
    if not isinstance(N, int) or not isinstance(NB, int) or N < 0 or NB <= 0:
        raise CholError(f"bad sizes N={N!r}, NB={NB!r}")
    _need_dtype(t, dtype, name)
    if not t.is_contiguous():
        raise CholError(f"{name} must be contiguous (lower tile-packed layout)")
    need = num_tiles(N, NB) * NB * NB
    if t.numel() < need:
        raise CholError(f"{name} has {t.numel()} elements, the lower tile-packed layout of N={N}, NB={NB} "
                        f"needs {need}")


def _check_cols(X, N, name, nrhs=None, dtype="f64"):
    """A column-major N x nrhs matrix (or a vector of N); returns nrhs."""
    _need_dtype(X, dtype, name)
    if X.dim() == 1:
        cols = 1
        if X.shape[0] != N or (N > 1 and X.stride(0) != 1):
            raise CholError(f"{name} must be a contiguous vector of {N} elements, got shape {tuple(X.shape)}")
    elif X.dim() == 2:
        cols = X.shape[1]
        if X.shape[0] != N:
            raise CholError(f"{name} must have N={N} rows, got shape {tuple(X.shape)}")
        if N > 1 and X.stride(0) != 1:
            raise CholError(f"{name} must be column-major (stride(0) == 1)")
        if cols > 1 and X.stride(1) < max(1, N):
            raise CholError(f"{name}: leading dimension {X.stride(1)} < N={N}")
    else:
        raise CholError(f"{name} must be 1-D or 2-D, got {X.dim()}-D")
    if nrhs is not None and cols != nrhs:
        raise CholError(f"{name} has {cols} columns, expected {nrhs}")
    _need_storage(X, _span(X), name)
    return cols


def _check_tile(A, name, rows=None, cols=None):
    """A column-major 2-D float64 tile."""
    _need_dtype(A, "f64", name)
    if A.dim() != 2:
        raise CholError(f"{name} must be 2-D, got {A.dim()}-D")
    if rows is not None and A.shape[0] != rows or cols is not None and A.shape[1] != cols:
        raise CholError(f"{name} has shape {tuple(A.shape)}, expected ({rows}, {cols})")
    if A.shape[0] > 1 and A.stride(0) != 1:
        raise CholError(f"{name} must be column-major (stride(0) == 1)")
    if A.shape[1] > 1 and A.stride(1) < max(1, A.shape[0]):
        raise CholError(f"{name}: leading dimension {A.stride(1)} < rows {A.shape[0]}")
    _need_storage(A, _span(A), name)


def _check_info(info, count=1, name="info"):
    if info is None:
        return
    _need_dtype(info, "i32", name)
    if info.numel() < count:
        raise CholError(f"{name} must hold at least {count} int32 entries, got {info.numel()}")


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
    n = A.shape[0] if A.dim() == 2 else -1
    _check_tile(A, "A", n, n)
    _check_info(info)
    _require_cuda(A, info)
    _check(lib().ttg_chol_dpotrf_tile(_stream(stream), n, _ptr(A), _ld(A), _ptr(info), info_offset),
           "dpotrf_tile")


def dtrsm_rlt(L, B, stream=None):
    _check_tile(B, "B")
    m, n = B.shape
    _check_tile(L, "L", n, n)
    _require_cuda(L, B)
    _check(lib().ttg_chol_dtrsm_rlt(_stream(stream), m, n, _ptr(L), _ld(L), _ptr(B), _ld(B)), "dtrsm_rlt")


def dsyrk_ln(A, C, alpha=-1.0, beta=1.0, stream=None):
    _check_tile(A, "A")
    n, k = A.shape
    _check_tile(C, "C", n, n)
    _require_cuda(A, C)
    _check(lib().ttg_chol_dsyrk_ln(_stream(stream), n, k, alpha, _ptr(A), _ld(A), beta, _ptr(C), _ld(C)),
           "dsyrk_ln")


def dgemm_nt(A, B, C, alpha=-1.0, beta=1.0, stream=None):
    _check_tile(A, "A")
    m, k = A.shape
    _check_tile(B, "B", None, k)
    n = B.shape[0]
    _check_tile(C, "C", m, n)
    _require_cuda(A, B, C)
    _check(lib().ttg_chol_dgemm_nt(_stream(stream), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                   _ptr(C), _ld(C)), "dgemm_nt")


# ---- tile operations of the POINV graphs (examples/potrf/trtri_L.h, lauum.h)


def dtrtri_tile(A, info=None, info_offset=0, stream=None):
    """A := A^{-1} on its lower triangle (non-unit)."""
    n = A.shape[0] if A.dim() == 2 else -1
    _check_tile(A, "A", n, n)
    _check_info(info)
    _require_cuda(A, info)
    _check(lib().ttg_chol_dtrtri_tile(_stream(stream), n, _ptr(A), _ld(A), _ptr(info), info_offset),
           "dtrtri_tile")


def _trsm_like(fn, name, L, B, alpha, left, stream):
    _check_tile(B, "B")
    m, n = B.shape
    order = m if left else n
    _check_tile(L, "L", order, order)
    _require_cuda(L, B)
    _check(getattr(lib(), fn)(_stream(stream), m, n, alpha, _ptr(L), _ld(L), _ptr(B), _ld(B)), name)


def dtrsm_lln(L, B, alpha=1.0, stream=None):
    """B := alpha L^{-1} B (L lower, non-unit)."""
    _trsm_like("ttg_chol_dtrsm_lln", "dtrsm_lln", L, B, alpha, True, stream)


def dtrsm_rln(L, B, alpha=-1.0, stream=None):
    """B := alpha B L^{-1} (L lower, non-unit)."""
    _trsm_like("ttg_chol_dtrsm_rln", "dtrsm_rln", L, B, alpha, False, stream)


def dtrmm_llt(L, B, alpha=1.0, stream=None):
    """B := alpha L^T B (L lower, non-unit)."""
    _trsm_like("ttg_chol_dtrmm_llt", "dtrmm_llt", L, B, alpha, True, stream)


def dlauum_tile(A, stream=None):
    """lower(A) := L^T L, L = lower(A)."""
    n = A.shape[0] if A.dim() == 2 else -1
    _check_tile(A, "A", n, n)
    _require_cuda(A)
    _check(lib().ttg_chol_dlauum_tile(_stream(stream), n, _ptr(A), _ld(A)), "dlauum_tile")


def dgemm_nn(A, B, C, alpha=1.0, beta=1.0, stream=None):
    """C := alpha A B + beta C."""
    _check_tile(A, "A")
    m, k = A.shape
    _check_tile(B, "B", k, None)
    n = B.shape[1]
    _check_tile(C, "C", m, n)
    _require_cuda(A, B, C)
    _check(lib().ttg_chol_dgemm_nn(_stream(stream), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                   _ptr(C), _ld(C)), "dgemm_nn")


def dgemm_tn(A, B, C, alpha=1.0, beta=1.0, stream=None):
    """C := alpha A^T B + beta C (A k x m)."""
    _check_tile(A, "A")
    k, m = A.shape
    _check_tile(B, "B", k, None)
    n = B.shape[1]
    _check_tile(C, "C", m, n)
    _require_cuda(A, B, C)
    _check(lib().ttg_chol_dgemm_tn(_stream(stream), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta,
                                   _ptr(C), _ld(C)), "dgemm_tn")


def dsyrk_lt(A, C, alpha=1.0, beta=1.0, stream=None):
    """C := alpha A^T A + beta C on the lower triangle (A k x n)."""
    _check_tile(A, "A")
    k, n = A.shape
    _check_tile(C, "C", n, n)
    _require_cuda(A, C)
    _check(lib().ttg_chol_dsyrk_lt(_stream(stream), n, k, alpha, _ptr(A), _ld(A), beta, _ptr(C), _ld(C)),
           "dsyrk_lt")


def dnrm2(x, out, stream=None):
    _need_dtype(x, "f64", "x")
    _need_dtype(out, "f64", "out")
    dense = x.is_contiguous() or (x.dim() == 2 and x.t().is_contiguous())  # row- or column-major, no gaps
    if not dense or out.numel() < 1:
        raise CholError("dnrm2: x must be dense (contiguous row- or column-major) and out hold one double")
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


def group_last_splits():
    """How often the last group launched early for a call that depended on
    an earlier call of the group."""
    return lib().ttg_chol_group_last_splits()


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
    _check_tiles(tiles, N, NB)
    _check_info(info)
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dpotrf_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info), flags), "dpotrf_tiled")


def dpotrs_tiled(tiles, N, NB, B, stream=None):
    _check_tiles(tiles, N, NB)
    nrhs = _check_cols(B, N, "B")
    _require_cuda(tiles, B)
    _check(lib().ttg_chol_dpotrs_tiled(_stream(stream), N, NB, _ptr(tiles), nrhs, _ptr(B), _ld(B)), "dpotrs_tiled")


def spotrs_tiled(tiles, N, NB, B, stream=None):
    """B := (L L^T)^{-1} B with an FP32 factor (lower float tiles), FP64 B."""
    _check_tiles(tiles, N, NB, dtype="f32")
    nrhs = _check_cols(B, N, "B")
    _require_cuda(tiles, B)
    _check(lib().ttg_chol_spotrs_tiled(_stream(stream), N, NB, _ptr(tiles), nrhs, _ptr(B), _ld(B)), "spotrs_tiled")


def dresidual_tiled(tiles, N, NB, X, B, R, stream=None):
    _check_tiles(tiles, N, NB)
    nrhs = _check_cols(B, N, "B")
    _check_cols(X, N, "X", nrhs)
    _check_cols(R, N, "R", nrhs)
    _require_cuda(tiles, X, B, R)
    _check(lib().ttg_chol_dresidual_tiled(_stream(stream), N, NB, _ptr(tiles), nrhs, _ptr(X), _ld(X), _ptr(B),
                                          _ld(B), _ptr(R), _ld(R)), "dresidual_tiled")


def backward_error(tiles, N, NB, X, B, stream=None):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) of a
    solve, A symmetric from its lower FP64 tiles, on the device: the residual
    by ttg_chol_dresidual_tiled, ||A||_inf as the largest row sum of |A| (the
    same residual kernel applied to |A| and a vector of ones)."""
    import torch

    nrhs = _check_cols(B, N, "B")
    _check_cols(X, N, "X", nrhs)
    _check_tiles(tiles, N, NB)
    _require_cuda(tiles, X, B)
    dev = tiles.device
    ones = torch.ones(N, 1, dtype=torch.float64, device=dev)
    zeros = torch.zeros(N, 1, dtype=torch.float64, device=dev)
    rs = torch.empty(N, 1, dtype=torch.float64, device=dev)
    dresidual_tiled(tiles.abs(), N, NB, ones, zeros, rs, stream)  # rs = -|A| 1
    anorm = (-rs).max()
    X2 = X.reshape(N, 1) if X.dim() == 1 else X
    B2 = B.reshape(N, 1) if B.dim() == 1 else B
    R = torch.empty(nrhs, N, dtype=torch.float64, device=dev).t()
    dresidual_tiled(tiles, N, NB, X2, B2, R, stream)
    rn = R.abs().amax(dim=0)
    e = rn / (anorm * X2.abs().amax(dim=0) + B2.abs().amax(dim=0))
    if torch.isnan(rn).any():
        return float("nan")
    return float(e.max().item())


def dcheck_dpotrf_tiled(A0, L, N, NB, stream=None):
    """The factorization check of testing_dpotrf.cc:277-335 on the device:
    (||A0||_inf, ||L L^T - A0||_inf, ||L L^T - A0||_inf / (||A0||_inf N eps))."""
    import torch

    _check_tiles(A0, N, NB, "A0")
    _check_tiles(L, N, NB, "L")
    _require_cuda(A0, L)
    out = torch.empty(3, dtype=torch.float64, device=A0.device)
    _check(lib().ttg_chol_dcheck_dpotrf_tiled(_stream(stream), N, NB, _ptr(A0), _ptr(L), _ptr(out)),
           "dcheck_dpotrf_tiled")
    a, r, q = (float(v) for v in out.cpu())
    return a, r, q


def dposv_mixed_factor(L, N, NB, nrhs, prec, stream=None):
    """The working factor of the last dposv_mixed call with these arguments,
    widened to FP64, into the tiles L."""
    _check_tiles(L, N, NB, "L")
    _require_cuda(L)
    _check(lib().ttg_chol_dposv_mixed_factor(_stream(stream), N, NB, nrhs, prec, _ptr(L)), "dposv_mixed_factor")


def release_plans():
    """Free every cached plan and workspace of the library (synchronises the
    device); the next call on a matrix builds its plan again."""
    _check(lib().ttg_chol_release_plans(), "release_plans")


# schedule knobs of the planned factorizations (include/ttg_chol.h)
KNOBS = {"group": 0, "potrf_cus": 1, "tile_group": 2, "paired": 3, "sweep": 4}


def set_knob(name, value):
    """Set a schedule knob (None clears it); plans built afterwards use it,
    so the caller releases the cached plans (release_plans) after a change."""
    _check(lib().ttg_chol_set_knob(KNOBS[name], -1 if value is None else int(value)), f"set_knob({name})")


def get_knob(name):
    """The value of a schedule knob that a plan built now would use."""
    v = lib().ttg_chol_get_knob(KNOBS[name])
    if v < 0:
        _check(v, f"get_knob({name})")
    return v


def knobs():
    return {k: get_knob(k) for k in KNOBS}


def dtrtri_tiled(tiles, N, NB, info=None, stream=None):
    """tiles := L^{-1} (lower tiles, in place); the TRTRI graph of
    examples/potrf/trtri_L.h.  info (device int32 tensor) receives 0 or the
    1-based index of the first zero diagonal entry."""
    _check_tiles(tiles, N, NB)
    _check_info(info)
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dtrtri_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info)), "dtrtri_tiled")


def dlauum_tiled(tiles, N, NB, stream=None):
    """tiles := L^T L (lower part, in place); the LAUUM graph of
    examples/potrf/lauum.h."""
    _check_tiles(tiles, N, NB)
    _require_cuda(tiles)
    _check(lib().ttg_chol_dlauum_tiled(_stream(stream), N, NB, _ptr(tiles)), "dlauum_tiled")


def dpotri_tiled(tiles, N, NB, info=None, stream=None):
    """tiles := A^{-1} (lower part, in place) from the Cholesky factor of A;
    the POTRI graph of examples/potrf/potri.h (TRTRI then LAUUM)."""
    _check_tiles(tiles, N, NB)
    _check_info(info)
    _require_cuda(tiles, info)
    _check(lib().ttg_chol_dpotri_tiled(_stream(stream), N, NB, _ptr(tiles), _ptr(info)), "dpotri_tiled")


def selftest_mfma_layout(kind, out, stream=None):
    need = {0: 64 * 64 * 4, 1: 64 * 64 * 4, 2: 512 * 64 * 4, 3: 512 * 64 * 16}.get(kind)
    _need_dtype(out, "f64", "out")
    if need is None or out.numel() < need or not out.is_contiguous():
        raise CholError(f"selftest_mfma_layout: kind {kind} needs a contiguous float64 out of {need} elements")
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

    _check_tiles(A, N, NB, "A")
    nrhs = _check_cols(B, N, "B")
    _check_cols(X, N, "X", nrhs)
    _check_info(info)
    _require_cuda(A, B, X, info)
    if info is None:
        info = torch.zeros(1, dtype=torch.int32, device=A.device)
    it = ctypes.c_int()
    be = ctypes.c_double()
    rc = lib().ttg_chol_dposv_mixed(_stream(stream), N, NB, _ptr(A), nrhs, _ptr(B), _ld(B), _ptr(X), _ld(X), prec,
                                    tol, max_iter, _ptr(info), ctypes.byref(it), ctypes.byref(be))
    if rc == ENOCONV and not check:
        return it.value, be.value
    _check(rc, "dposv_mixed")
    return it.value, be.value


def dposv_krylov(A, N, NB, B, X, prec=PREC_F16, shift=0.0, tol=1e-12, max_iter=100, info=None, stream=None,
                 check=True):
    """A X = B by FP64 preconditioned CG with the low-precision factor of A
    (diagonal scaled by 1 + shift) as preconditioner; returns (CG
    iterations, backward error)."""
    import torch

    _check_tiles(A, N, NB, "A")
    nrhs = _check_cols(B, N, "B")
    _check_cols(X, N, "X", nrhs)
    _check_info(info)
    _require_cuda(A, B, X, info)
    if info is None:
        info = torch.zeros(1, dtype=torch.int32, device=A.device)
    it = ctypes.c_int()
    be = ctypes.c_double()
    rc = lib().ttg_chol_dposv_krylov(_stream(stream), N, NB, _ptr(A), nrhs, _ptr(B), _ld(B), _ptr(X), _ld(X), prec,
                                     float(shift), tol, max_iter, _ptr(info), ctypes.byref(it), ctypes.byref(be))
    if rc == ENOCONV and not check:
        return it.value, be.value
    _check(rc, "dposv_krylov")
    return it.value, be.value


def profile_enable(on):
    _check(lib().ttg_chol_profile_enable(1 if on else 0), "profile_enable")


def dplgsy_tile(A, bump, gM, m0, n0, seed, stream=None):
    """One CORE_dplgsy tile (column-major device tensor A) at global offset (m0, n0)."""
    _check_tile(A, "A")
    _require_cuda(A)
    m, n = A.shape
    _check(lib().ttg_chol_dplgsy_tile(_stream(stream), float(bump), m, n, _ptr(A), _ld(A), gM, m0, n0, seed),
           "dplgsy_tile")


def plgsy_tiled(tiles, N, NB, bump, seed, stream=None):
    _check_tiles(tiles, N, NB)
    _require_cuda(tiles)
    _check(lib().ttg_chol_plgsy_tiled(_stream(stream), N, NB, float(bump), seed, _ptr(tiles)), "plgsy_tiled")


def _check_work(work, N):
    _need_dtype(work, "f64", "work")
    if not work.is_contiguous() or work.numel() < N * N:
        raise CholError(f"work must be a contiguous float64 tensor of at least N*N = {N * N} elements")


def spd_tiled(tiles, N, NB, alpha, seed, work, stream=None):
    _check_tiles(tiles, N, NB)
    _check_work(work, N)
    _require_cuda(tiles, work)
    _check(lib().ttg_chol_spd_tiled(_stream(stream), N, NB, float(alpha), seed, _ptr(tiles), _ptr(work)),
           "spd_tiled")


def spd_uniform_tiled(tiles, N, NB, alpha, seed, work, stream=None):
    """A = B B^T / N + alpha I with B FP64 uniform on [-1, 1) (seeded LCG)."""
    _check_tiles(tiles, N, NB)
    _check_work(work, N)
    _require_cuda(tiles, work)
    _check(lib().ttg_chol_spd_uniform_tiled(_stream(stream), N, NB, float(alpha), seed, _ptr(tiles), _ptr(work)),
           "spd_uniform_tiled")


class Batch:
    """A batch of FP32 SPD systems for sposv_batched: host arrays of sizes,
    leading dimensions and device pointers, built once and reusable."""

    def __init__(self, mats, rhs):
        if len(mats) != len(rhs):
            raise ValueError("one right-hand side per matrix")
        for i, (m, v) in enumerate(zip(mats, rhs)):
            _need_dtype(m, "f32", f"matrix {i}")
            _need_dtype(v, "f32", f"right-hand side {i}")
            if m.dim() != 2 or m.shape[0] != m.shape[1] or not m.is_contiguous() or m.numel() == 0:
                raise CholError(f"matrix {i} must be a non-empty contiguous square tensor, got {tuple(m.shape)}")
            if v.numel() != m.shape[0] or not v.is_contiguous():
                raise CholError(f"right-hand side {i} must be {m.shape[0]} contiguous floats")
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
        _need_dtype(A, "f32", "A")
        _need_dtype(b, "f32", "b")
        if not A.is_contiguous() or not b.is_contiguous():
            raise CholError("packed batch buffers must be contiguous")
        if any(int(s) < 1 for s in sizes):
            raise CholError("every matrix of a batch needs n >= 1")
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
            raise CholError("packed batch larger than its buffers")
        self.A = (ctypes.c_void_p * self.count)(*pa)
        self.b = (ctypes.c_void_p * self.count)(*pb)
        return self


def sposv_batched(batch, info, cfg=-1, stream=None):
    """Factor + solve every system of `batch` in one launch (info: int32 device
    tensor of batch.count entries)."""
    _check_info(info, batch.count)
    if info is None:
        raise CholError("sposv_batched needs an info tensor of batch.count int32 entries")
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
