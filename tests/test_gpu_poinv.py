"""POINV family on the GPU (TRTRI, LAUUM, POTRI over the lower tiles; the
graphs of examples/potrf/trtri_L.h, lauum.h, potri.h) against the oracle.
FP64 bar: max |gpu - oracle| / max |oracle| <= 1e-12 over the lower tiles,
plus the closed forms L L^{-1} = I and A A^{-1} = I and the reference's
check_dtrtri bound at sizes where the oracle is too slow."""
import numpy as np
import pytest
import torch

from gpu_util import lower_tiles_rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12
SIZES = [(160, 32), (100, 32), (1000, 128), (1536, 512), (2048, 256)]


def factor(oracle, N, NB, seed=5):
    A0 = oracle.spd_tiles(N, NB, 1.0, seed)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    return A0, L


def strict_upper_of_diag(oracle, N, NB, t):
    nt = oracle.nt_of(N, NB)
    out = []
    for k in range(nt):
        d = min(NB, N - k * NB)
        out.append(np.triu(oracle.tile_view(t, N, NB, k, k)[:d, :d], 1).copy())
    return out


@pytest.mark.parametrize("N,NB", SIZES)
def test_dtrtri_tiled_vs_oracle(dev, oracle, N, NB):
    import ttg_amd.chol as C

    _, L = factor(oracle, N, NB)
    # garbage in the unreferenced strict upper part of the diagonal tiles
    nt = oracle.nt_of(N, NB)
    for k in range(nt):
        v = oracle.tile_view(L, N, NB, k, k)
        d = min(NB, N - k * NB)
        v[:d, :d] += np.triu(np.full((d, d), 7.0), 1)
    want = L.copy()
    assert oracle.trtri_tiled(N, NB, want) == 0
    t = to_dev(L, dev)
    info = torch.full((1,), -1, dtype=torch.int32, device=dev)
    C.dtrtri_tiled(t, N, NB, info)
    got = to_host(t)
    assert int(info.item()) == 0
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= TOL
    for a, b in zip(strict_upper_of_diag(oracle, N, NB, got), strict_upper_of_diag(oracle, N, NB, L)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("N,NB", SIZES)
def test_dlauum_tiled_vs_oracle(dev, oracle, N, NB):
    import ttg_amd.chol as C

    _, L = factor(oracle, N, NB, seed=9)
    want = L.copy()
    oracle.lauum_tiled(N, NB, want)
    t = to_dev(L, dev)
    C.dlauum_tiled(t, N, NB)
    got = to_host(t)
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= TOL
    # a second call on fresh data reuses the plan and its work copy
    t2 = to_dev(L, dev)
    C.dlauum_tiled(t2, N, NB)
    assert np.array_equal(to_host(t2), got)


@pytest.mark.parametrize("N,NB", [(160, 32), (1000, 128), (2048, 512)])
def test_dpotri_tiled_vs_oracle(dev, oracle, N, NB):
    import ttg_amd.chol as C

    A0, L = factor(oracle, N, NB, seed=11)
    want = L.copy()
    assert oracle.potri_tiled(N, NB, want) == 0
    t = to_dev(L, dev)
    info = torch.full((1,), -1, dtype=torch.int32, device=dev)
    C.dpotri_tiled(t, N, NB, info)
    got = to_host(t)
    assert int(info.item()) == 0
    assert lower_tiles_rel_err(oracle, N, NB, got, want) <= TOL
    X = np.tril(oracle.tiles_to_dense(N, NB, got))
    X = X + np.tril(X, -1).T
    A = np.tril(oracle.tiles_to_dense(N, NB, A0))
    A = A + np.tril(A, -1).T
    assert np.abs(X @ A - np.eye(N)).max() < 1e-12


def test_poinv_after_gpu_potrf_large(dev, oracle):
    """potrf -> trtri and potrf -> potri on the GPU at N=4096, NB=512: the
    closed forms on the GPU (FP64 products) and check_dtrtri at N=2048."""
    import ttg_amd.chol as C

    N, NB = 4096, 512
    A0 = oracle.spd_tiles(N, NB, 1.0, 21)
    t = to_dev(A0, dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    L = to_host(t)
    assert int(info.item()) == 0
    Ld = torch.from_numpy(np.tril(oracle.tiles_to_dense(N, NB, L))).to(dev)
    t2 = to_dev(L, dev)
    C.dtrtri_tiled(t2, N, NB, info)
    Li = torch.from_numpy(np.tril(oracle.tiles_to_dense(N, NB, to_host(t2)))).to(dev)
    assert int(info.item()) == 0
    eye = torch.eye(N, dtype=torch.float64, device=dev)
    assert (Ld @ Li - eye).abs().max().item() < 1e-12
    t3 = to_dev(L, dev)
    C.dpotri_tiled(t3, N, NB, info)
    X = np.tril(oracle.tiles_to_dense(N, NB, to_host(t3)))
    X = torch.from_numpy(X + np.tril(X, -1).T).to(dev)
    A = np.tril(oracle.tiles_to_dense(N, NB, A0))
    A = torch.from_numpy(A + np.tril(A, -1).T).to(dev)
    assert (X @ A - eye).abs().max().item() < 1e-11
    # the reference's TRTRI check at N=2048
    N2, NB2 = 2048, 256
    _, L2 = factor(oracle, N2, NB2, seed=3)
    t4 = to_dev(L2, dev)
    C.dtrtri_tiled(t4, N2, NB2, info)
    Li2 = np.tril(oracle.tiles_to_dense(N2, NB2, to_host(t4)))
    assert oracle.check_dtrtri(np.tril(oracle.tiles_to_dense(N2, NB2, L2)), Li2) <= 10.0


def test_dtrtri_singular_info(dev, oracle):
    import ttg_amd.chol as C

    N, NB = 300, 128
    _, L = factor(oracle, N, NB, seed=2)
    v = oracle.tile_view(L, N, NB, 2, 2)
    v[13, 13] = 0.0  # global index 2*128 + 13
    t = to_dev(L, dev)
    info = torch.full((1,), -1, dtype=torch.int32, device=dev)
    C.dtrtri_tiled(t, N, NB, info)
    assert int(info.item()) == 2 * NB + 13 + 1
    got = to_host(t)
    assert lower_tiles_rel_err(oracle, N, NB, got, L) == 0.0
    C.dpotri_tiled(t, N, NB, info)
    assert int(info.item()) == 2 * NB + 13 + 1
    assert lower_tiles_rel_err(oracle, N, NB, to_host(t), L) == 0.0


def test_release_plans_then_rebuild(dev, oracle):
    """Plans freed by ttg_chol_release_plans are rebuilt on the next call and
    give the same bits (POTRI and the mixed solve)."""
    import ttg_amd.chol as C

    N, NB = 1000, 128
    _, L = factor(oracle, N, NB, seed=13)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    t = to_dev(L, dev)
    C.dpotri_tiled(t, N, NB, info)
    first = to_host(t)
    A = torch.from_numpy(oracle.spd_tiles(1024, 128, 1.0, 3)).to(dev)
    B = (torch.arange(2 * 1024, dtype=torch.float64, device=dev).reshape(2, 1024) % 7 - 3).t()
    X1 = torch.empty_like(B.t()).t()
    C.dposv_mixed(A, 1024, 128, B, X1, prec=C.PREC_FP32, tol=1e-12, check=False)
    C.release_plans()
    t = to_dev(L, dev)
    C.dpotri_tiled(t, N, NB, info)
    assert np.array_equal(to_host(t), first)
    X2 = torch.empty_like(B.t()).t()
    C.dposv_mixed(A, 1024, 128, B, X2, prec=C.PREC_FP32, tol=1e-12, check=False)
    assert torch.equal(X1, X2)


def _run_poinv_driver(args):
    import os
    import subprocess

    import ttg_amd.graph as G

    exe = os.path.join(os.path.dirname(G.DRIVER), "testing_dpoinv")
    print("testing_dpoinv", *args, flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def test_poinv_driver_check_and_lines():
    """bin/testing_dpoinv: PLGSY(bump=N, seed=3872), POTRF + POTRI on the
    device, the inversion check, and the reference driver's output lines."""
    r = _run_poinv_driver(["-N", "160", "-t", "32", "-nruns", "1"])
    assert r.returncode == 0
    assert "Inversion is CORRECT" in r.stdout
    lines = [l for l in r.stdout.splitlines() if l.startswith("POINV (POTRF+POTRI)")]
    assert len(lines) == 3 and all("TTG Execution Time (milliseconds) : " in l for l in lines)
    r = _run_poinv_driver(["-N", "1536", "-t", "512", "-seq", "1", "-nruns", "1"])
    assert r.returncode == 0
    assert "Inversion is CORRECT" in r.stdout
    kinds = [l.split("nthreads= ")[1].split()[1] for l in r.stdout.splitlines()
             if l.startswith("POINV (POTRF+TRTRI+LAUUM)")]
    assert kinds == ["POTRF", "TRTRI", "LAUUM", "TTG"]


@pytest.mark.parametrize("seq,kind,n_lines", [(0, "POTRF+POTRI", 3), (1, "POTRF+TRTRI+LAUUM", 4), (2, "POINV", 1)])
def test_poinv_driver_task_graphs(seq, kind, n_lines):
    """bin/testing_dpoinv -graph: the reference's POTRF, POTRI, TRTRI and
    LAUUM task graphs on the TTG runtime (device tasks over a host matrix),
    the inversion check and the reference driver's lines, for each -seq."""
    r = _run_poinv_driver(["-N", "700", "-t", "64", "-seq", str(seq), "-nruns", "1", "-graph"])
    assert r.returncode == 0
    assert "Inversion is CORRECT" in r.stdout
    lines = [l for l in r.stdout.splitlines() if l.startswith(f"POINV ({kind})")]
    assert len(lines) == n_lines and all("TTG Execution Time (milliseconds) : " in l for l in lines)
