"""The POINV task graphs (ttg_amd/examples/potrf/trtri_L.h, lauum.h,
potri.h: the reference's TRTRI, LAUUM and POTRI graphs with device bodies)
against the oracle's tiled TRTRI/LAUUM/POTRI, which follow the same task
order: bitwise on small-integer data for LAUUM (every product and sum is
exact), within 1e-12 relative otherwise (FP64), with the task counts the
graphs' successor rules fix."""
import numpy as np
import pytest

from gpu_util import lower_tiles_rel_err

pytestmark = pytest.mark.gpu

CASES = [(160, 32), (700, 64), (1000, 128), (2048, 512)]


def tiles_count(nt):
    return nt * (nt + 1) // 2


def graph_tasks(nt):
    # dispatch + result per tile; one diagonal task per tile column; two
    # tasks per strictly-lower tile; one GEMM per (K, M, N) ordered triple
    return 2 * tiles_count(nt) + nt + nt * (nt - 1) + nt * (nt - 1) * (nt - 2) // 6


def factor(oracle, N, NB, seed):
    A = oracle.spd_tiles(N, NB, 1.0, seed)
    assert oracle.potrf_tiled(N, NB, A) == 0
    return A


def int_lower(oracle, N, NB, seed):
    """Lower tiles holding small integers (upper parts of the diagonal tiles too)."""
    rng = np.random.default_rng(seed)
    t = oracle.alloc_tiles(N, NB)
    t[:] = rng.integers(-4, 5, size=t.size)
    return t


@pytest.mark.parametrize("N,NB", CASES)
def test_graph_trtri_vs_oracle(oracle, N, NB):
    import ttg_amd.graph as G

    L = factor(oracle, N, NB, 3)
    want = L.copy()
    assert oracle.trtri_tiled(N, NB, want) == 0
    got = L.copy()
    print(f"graph dtrtri N={N} NB={NB}", flush=True)
    st = G.dtrtri(N, NB, got)
    assert st["info"] == 0
    assert st["tasks"] == graph_tasks(oracle.nt_of(N, NB)), st
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= 1e-12


@pytest.mark.parametrize("N,NB", CASES)
def test_graph_lauum_vs_oracle(oracle, N, NB):
    import ttg_amd.graph as G

    L = int_lower(oracle, N, NB, 5)
    want = L.copy()
    oracle.lauum_tiled(N, NB, want)
    got = L.copy()
    print(f"graph dlauum N={N} NB={NB}", flush=True)
    st = G.dlauum(N, NB, got)
    assert st["tasks"] == graph_tasks(oracle.nt_of(N, NB)), st
    assert lower_tiles_rel_err(oracle, N, NB, got, want) == 0.0
    # real data
    L = factor(oracle, N, NB, 6)
    want = L.copy()
    oracle.lauum_tiled(N, NB, want)
    got = L.copy()
    G.dlauum(N, NB, got)
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= 1e-12


@pytest.mark.parametrize("N,NB", CASES)
def test_graph_potri_vs_oracle(oracle, N, NB):
    import ttg_amd.graph as G

    L = factor(oracle, N, NB, 7)
    want = L.copy()
    assert oracle.potri_tiled(N, NB, want) == 0
    got = L.copy()
    print(f"graph dpotri N={N} NB={NB}", flush=True)
    st = G.dpotri(N, NB, got)
    assert st["info"] == 0
    nt = oracle.nt_of(N, NB)
    # both graphs' tasks, less the TRTRI graph's result writers (its tiles go
    # to the LAUUM dispatcher instead)
    assert st["tasks"] == 2 * graph_tasks(nt) - tiles_count(nt), st
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= 1e-12


@pytest.mark.parametrize("col", [1, 70, 700])
def test_graph_trtri_stops_with_info(oracle, col):
    """An exact zero on the diagonal: the graph stops with LAPACK's info, as
    the oracle (and LAPACK dtrtri) reports it."""
    import ttg_amd.graph as G

    N, NB = 700, 64
    L = factor(oracle, N, NB, 9)
    t, i = (col - 1) // NB, (col - 1) % NB
    oracle.tile_view(L, N, NB, t, t)[i, i] = 0.0
    assert oracle.trtri_tiled(N, NB, L.copy()) == col
    st = G.dtrtri(N, NB, L)
    assert st["info"] == col, st
