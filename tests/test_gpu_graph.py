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
                                      (2048, 256, "spd")])
def test_graph_dpotrf_vs_oracle(oracle, N, NB, gen):
    import ttg_amd.graph as G

    A0 = oracle.plgsy_tiles(N, NB, N, 3872) if gen == "plgsy" else oracle.spd_tiles(N, NB, 1.0, 5)
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
    avg = float(r.stdout.strip().splitlines()[-1].split("avg")[-1])
    assert 0.0 < avg < 10000.0
