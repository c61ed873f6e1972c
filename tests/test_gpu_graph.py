"""The examples/potrf task graph on the single-GPU TTG runtime
(ttg_amd/include/ttg.h + ttg_amd/examples/potrf): the reference driver's
check (testing_dpotrf -N 160 -t 32, the reference's ctest case
examples/CMakeLists.txt potrf), PLGSY through device tasks bit-exact against
the oracle, and the graph's factor against the oracle's tiled factor
(FP64 bar: max |graph - oracle| / max |oracle| <= 1e-12 over the lower
tiles) with the task count the graph's successor rules fix."""
import subprocess

import numpy as np
import pytest

from gpu_util import lower_tiles_rel_err

pytestmark = pytest.mark.gpu


def expected_tasks(nt):
    tiles = nt * (nt + 1) // 2
    # dispatch + result per tile; POTRF per column; TRSM and SYRK per
    # strictly-lower tile; GEMM per (m > n > k)
    return 2 * tiles + nt + 2 * (nt * (nt - 1) // 2) + nt * (nt - 1) * (nt - 2) // 6


def run_driver(args, exe=None):
    import ttg_amd.graph as G

    exe = exe or G.DRIVER
    print(exe, *args, flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def test_driver_check_reference_case():
    r = run_driver(["-N", "160", "-t", "32"])
    assert r.returncode == 0
    assert "Factorization is CORRECT" in r.stdout


@pytest.mark.parametrize("args", [["-N", "160", "-t", "32"], ["-N", "1000", "-t", "128"]])
def test_driver_tile_norm_checks(args):
    """The DEBUG_TILES_VALUES build: every task checks the 2-norms of its
    input tiles against the norms their producers set and the result sink
    checks the final tiles, so a run that ends CORRECT received every tile
    its producer sent (a mismatch aborts with the task and tile named)."""
    from ttg_amd.build import DEBUG_DRIVER

    r = run_driver(args, DEBUG_DRIVER)
    assert r.returncode == 0
    assert "Factorization is CORRECT" in r.stdout
    assert "producer sent" not in r.stderr


def test_driver_timed_mode():
    r = run_driver(["-N", "2048", "-t", "256", "-x", "-n", "2"])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("TTG Execution Time (milliseconds) : ")]
    assert len(lines) == 2
    gflops = float(lines[-1].split()[-2])
    assert gflops > 0


@pytest.mark.parametrize("N,NB", [(160, 32), (700, 64), (1536, 512)])
def test_graph_plgsy_bit_exact(oracle, N, NB):
    import ttg_amd.graph as G

    want = oracle.plgsy_tiles(N, NB, N, 3872)
    got = np.full_like(want, np.nan)
    G.dplgsy(N, NB, N, 3872, got)
    nt = oracle.nt_of(N, NB)
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            a = oracle.tile_view(got, N, NB, m, n)[:mb, :nb]
            b = oracle.tile_view(want, N, NB, m, n)[:mb, :nb]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (m, n)


@pytest.mark.parametrize("N,NB,gen", [(160, 32, "plgsy"), (700, 64, "plgsy"), (1000, 128, "spd"),
                                      (2048, 256, "spd"), (4096, 256, "uniform")])
def test_graph_dpotrf_vs_oracle(oracle, N, NB, gen):
    """(4096, 256, uniform) is cfg1: A = B B^T/N + I with B FP64 uniform."""
    import ttg_amd.graph as G

    A0 = {"plgsy": lambda: oracle.plgsy_tiles(N, NB, N, 3872), "spd": lambda: oracle.spd_tiles(N, NB, 1.0, 5),
          "uniform": lambda: oracle.spd_uniform_tiles(N, NB, 1.0, 2024)}[gen]()
    ref = A0.copy()
    assert oracle.potrf_tiled(N, NB, ref) == 0
    got = A0.copy()
    print(f"graph dpotrf N={N} NB={NB} {gen}", flush=True)
    st = G.dpotrf(N, NB, got)
    nt = oracle.nt_of(N, NB)
    assert st["tasks"] == expected_tasks(nt), st
    assert st["rounds"] >= 3 * nt - 2, st  # POTRF -> TRSM -> SYRK chain per column
    assert st["launches"] < st["tasks"], st  # rounds batch their tile calls
    assert lower_tiles_rel_err(oracle, N, NB, got, ref) <= 1e-12


@pytest.mark.parametrize("N,NB,col", [(700, 64, 300), (1000, 128, 1), (1000, 128, 129), (1000, 128, 1000)])
def test_graph_stops_with_lapack_info(oracle, N, NB, col):
    """A matrix whose leading minor of order `col` is not positive definite
    (a_{col,col} planted negative, 1-based): the graph stops cleanly with
    LAPACK's info = col, the oracle's tiled factor reports the same, and the
    next factorization through the same runtime runs in full."""
    import ttg_amd.graph as G

    A0 = oracle.spd_tiles(N, NB, 1.0, 5)
    t, i = (col - 1) // NB, (col - 1) % NB
    oracle.tile_view(A0, N, NB, t, t)[i, i] = -1.0
    ref = A0.copy()
    want = oracle.potrf_tiled(N, NB, ref)
    assert want == col
    got = A0.copy()
    print(f"graph dpotrf N={N} NB={NB}, planted pivot at {col}", flush=True)
    st = G.dpotrf(N, NB, got)
    assert st["info"] == want, st
    assert st["tasks"] < expected_tasks(oracle.nt_of(N, NB)), st
    # the runtime is ready for the next graph
    good = oracle.spd_tiles(N, NB, 1.0, 6)
    ref = good.copy()
    assert oracle.potrf_tiled(N, NB, ref) == 0
    st = G.dpotrf(N, NB, good)
    assert st["info"] == 0 and st["tasks"] == expected_tasks(oracle.nt_of(N, NB)), st
    assert lower_tiles_rel_err(oracle, N, NB, good, ref) <= 1e-12


@pytest.mark.parametrize("flows", [0, 1, 2, 4, 8])
def test_chain_benchmark(flows):
    """The device-task chain benchmark (task-benchmarks/chain-ttg-dev.cc):
    n + 1 dependent device tasks that forward F values (or, for F = 0,
    only their keys) run to the end and report the reference's lines."""
    from ttg_amd.build import CHAIN_BENCH

    r = run_driver([str(flows), "-n", "500"], CHAIN_BENCH)
    assert r.returncode == 0
    assert f"Graph {flows} is connected." in r.stdout
    assert "# of tasks = 500" in r.stdout
    assert "# of tasks run = 501" in r.stdout  # keys 0 .. 500: the chain ran to its end
    avg = float(r.stdout.strip().splitlines()[-1].split("avg")[-1])
    assert 0.0 < avg < 10000.0


CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import ttg_amd.graph as G
A = np.load(sys.argv[2])
st = G.dpotrf(int(sys.argv[3]), int(sys.argv[4]), A)
np.save(sys.argv[5], A)
print("rounds", st["rounds"], "max_round", st["max_round"], flush=True)
"""


def test_graph_factor_independent_of_round_sizes(oracle, tmp_path):
    """Round sizes only regroup tasks that are ready together, and the graph
    fixes every tile's update order, so the factor is bitwise the same with
    one task per round (TTG_ROUND_MIN = TTG_ROUND_MAX = 1, set before the
    runtime starts, hence in a child process) as with the default sizes."""
    import os
    import sys

    import ttg_amd.graph as G

    N, NB = 1000, 128
    A0 = oracle.spd_tiles(N, NB, 1.0, 8)
    np.save(tmp_path / "a0.npy", A0)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, TTG_ROUND_MIN="1", TTG_ROUND_MAX="1")
    print("child: graph dpotrf N=1000 NB=128, one task per round", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, root, str(tmp_path / "a0.npy"), str(N), str(NB),
                        str(tmp_path / "one.npy")], env=env, capture_output=True, text=True, timeout=100)
    print(r.stdout[-500:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    got = A0.copy()
    st = G.dpotrf(N, NB, got)
    one = np.load(tmp_path / "one.npy")
    nt = oracle.nt_of(N, NB)
    assert st["max_round"] > 1 and "max_round 1" in r.stdout
    assert int(r.stdout.split()[1]) == expected_tasks(nt) - 2 * (nt * (nt + 1) // 2)  # device tasks, one per round
    for n in range(nt):
        for m in range(n, nt):
            a = oracle.tile_view(got, N, NB, m, n)
            b = oracle.tile_view(one, N, NB, m, n)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (m, n)


@pytest.mark.parametrize("N,NB,kappa", [(2048, 256, 1e4), (1000, 128, 1e3), (4096, 512, 1e5)])
def test_graph_mixed_solve(oracle, N, NB, kappa):
    """Mixed precision through the task graph: the potrf graph over FP32
    tiles (device tasks on the S tile kernels) and FP64 refinement reach the
    FP64 backward error bound 1e-12, checked by the oracle; ragged N
    exercises the padding of the FP32 factor."""
    import ttg_amd.graph as G

    sigma2 = sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2
    A0 = oracle.spd_tiles(N, NB, 4 * sigma2 / (kappa - 1), 31)
    B = np.random.default_rng(3).uniform(-1, 1, (N, 4))
    print(f"graph mixed solve N={N} NB={NB} kappa={kappa:g}", flush=True)
    X, st = G.dposv_mixed(N, NB, A0.copy(), B, tol=1e-12, max_iter=30)
    assert st["info"] == 0, st
    assert 1 <= st["iters"] <= 30 and st["berr"] <= 1e-12, st
    assert oracle.solve_backward_error(N, NB, A0, X, B) <= 1e-12


def test_graph_mixed_solve_reports_info(oracle):
    """A matrix that is not positive definite: the FP32 graph stops with
    LAPACK's info and nothing is solved."""
    import ttg_amd.graph as G

    N, NB, col = 1000, 128, 300
    A0 = oracle.spd_tiles(N, NB, 1.0, 5)
    t, i = (col - 1) // NB, (col - 1) % NB
    oracle.tile_view(A0, N, NB, t, t)[i, i] = -1.0
    _, st = G.dposv_mixed(N, NB, A0, np.ones((N, 1)))
    assert st["info"] == col, st
