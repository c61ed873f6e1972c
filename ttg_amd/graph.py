"""ctypes binding of libttg_graph.so (include/ttg_graph.h): the examples/potrf
task graph on the single-GPU TTG runtime.  Host numpy tiles in, host tiles
out; there is no CPU path."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libttg_graph.so")
DRIVER = os.path.join(HERE, "bin", "testing_dpotrf")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ttg_graph.h")

SIGNATURES = {
    "ttg_graph_dpotrf": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "ttg_graph_dtrtri": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "ttg_graph_dlauum": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ttg_graph_dpotri": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "ttg_graph_dposv_mixed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "ttg_graph_dplgsy": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulong,
                                        ctypes.c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run python -m ttg_amd.build")
        L = ctypes.CDLL(LIB)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _tiles(t, N, NB):
    """The host tile array of the lower tile-packed layout of (N, NB): the
    library writes every tile, so a short array is refused here."""
    if not (isinstance(t, np.ndarray) and t.dtype == np.float64 and t.flags.c_contiguous):
        raise TypeError("tiles must be a contiguous float64 numpy array")
    if not isinstance(N, int) or not isinstance(NB, int) or N < 0 or NB <= 0:
        raise ValueError(f"bad sizes N={N!r}, NB={NB!r}")
    nt = (N + NB - 1) // NB
    need = nt * (nt + 1) // 2 * NB * NB
    if t.size < need:
        raise ValueError(f"tiles has {t.size} elements, the lower tile-packed layout of N={N}, NB={NB} needs {need}")
    return t.ctypes.data


def dpotrf(N, NB, tiles):
    """Factor host tiles in place with the task graph; returns the runtime's
    statistics (tasks, rounds, kernel launches, largest round) and LAPACK's
    info (0, or the 1-based column of the first non-positive pivot, at which
    the graph stopped)."""
    stats = (ctypes.c_long * 4)()
    info = ctypes.c_int(-1)
    rc = lib().ttg_graph_dpotrf(N, NB, _tiles(tiles, N, NB), ctypes.addressof(stats), ctypes.addressof(info))
    if rc != 0:
        raise ValueError(f"ttg_graph_dpotrf(N={N}, NB={NB}) rejected its arguments")
    return {"tasks": stats[0], "rounds": stats[1], "launches": stats[2], "max_round": stats[3], "info": info.value}


def _stats(stats, info=None):
    d = {"tasks": stats[0], "rounds": stats[1], "launches": stats[2], "max_round": stats[3]}
    if info is not None:
        d["info"] = info.value
    return d


def dtrtri(N, NB, tiles):
    """L := L^{-1} on host lower tiles with the TRTRI task graph."""
    stats, info = (ctypes.c_long * 4)(), ctypes.c_int(-1)
    if lib().ttg_graph_dtrtri(N, NB, _tiles(tiles, N, NB), ctypes.addressof(stats), ctypes.addressof(info)) != 0:
        raise ValueError(f"ttg_graph_dtrtri(N={N}, NB={NB}) rejected its arguments")
    return _stats(stats, info)


def dlauum(N, NB, tiles):
    """A := L^T L on host lower tiles with the LAUUM task graph."""
    stats = (ctypes.c_long * 4)()
    if lib().ttg_graph_dlauum(N, NB, _tiles(tiles, N, NB), ctypes.addressof(stats)) != 0:
        raise ValueError(f"ttg_graph_dlauum(N={N}, NB={NB}) rejected its arguments")
    return _stats(stats)


def dpotri(N, NB, tiles):
    """A := (L L^T)^{-1} from the factor in host lower tiles with the POTRI task graph."""
    stats, info = (ctypes.c_long * 4)(), ctypes.c_int(-1)
    if lib().ttg_graph_dpotri(N, NB, _tiles(tiles, N, NB), ctypes.addressof(stats), ctypes.addressof(info)) != 0:
        raise ValueError(f"ttg_graph_dpotri(N={N}, NB={NB}) rejected its arguments")
    return _stats(stats, info)


def dposv_mixed(N, NB, tiles, B, tol=1e-12, max_iter=30):
    """Solve A X = B (A: host FP64 lower tiles, B: N x nrhs) with the FP32
    factorization run by the task graph and FP64 refinement; returns
    (X, {"iters", "berr", "info"})."""
    B = np.asfortranarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    if B.shape[0] != N:
        raise ValueError(f"B has {B.shape[0]} rows, expected N={N}")
    X = np.zeros_like(B, order="F")
    it, be, info = ctypes.c_int(-1), ctypes.c_double(float("nan")), ctypes.c_int(-1)
    rc = lib().ttg_graph_dposv_mixed(N, NB, _tiles(tiles, N, NB), B.shape[1], B.ctypes.data, N, X.ctypes.data, N,
                                     float(tol), int(max_iter), ctypes.addressof(it), ctypes.addressof(be),
                                     ctypes.addressof(info))
    if rc == -1:
        raise ValueError(f"ttg_graph_dposv_mixed(N={N}, NB={NB}) rejected its arguments")
    if rc != 0:
        raise RuntimeError(f"ttg_graph_dposv_mixed failed on the device ({rc})")
    return X, {"iters": it.value, "berr": be.value, "info": info.value}


def dplgsy(N, NB, bump, seed, tiles):
    rc = lib().ttg_graph_dplgsy(N, NB, float(bump), seed, _tiles(tiles, N, NB))
    if rc != 0:
        raise ValueError(f"ttg_graph_dplgsy(N={N}, NB={NB}) rejected its arguments")
