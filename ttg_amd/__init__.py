"""ttg_amd: MI355X-native tiled Cholesky (the POTRF/TRSM/SYRK/GEMM graph of
TTG's examples/potrf) on hand-written gfx950 HIP kernels.

The compute path is the C-ABI library ``ttg_amd/libttg_chol.so`` (declared in
``include/ttg_chol.h``); :mod:`ttg_amd.chol` binds it with ctypes for Python
callers.  There is no CPU fallback: calls fail loudly if the library or a GPU
is missing.
"""
from . import chol  # noqa: F401

__all__ = ["chol"]
