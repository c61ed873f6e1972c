"""Per-tile HIP kernels (the C-ABI replacements of hipsolverDpotrf,
hipblasDtrsm/Dsyrk/Dgemm of examples/potrf/potrf.h) against the oracle and
the LAPACK fixtures.  FP64 bar: max |gpu - oracle| / max |oracle| <= 1e-12."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from gpu_util import dev_colmajor, host_colmajor, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def spd(n, rng):
    B = rng.uniform(-1, 1, (n, n))
    return B @ B.T / max(n, 1) + np.eye(n)


@pytest.mark.parametrize("n,lda", [(1, 1), (7, 9), (32, 32), (45, 50), (64, 64), (100, 128), (256, 256),
                                   (500, 512), (512, 512)])
def test_dpotrf_tile(dev, oracle, n, lda):
    import ttg_amd.chol as C

    rng = np.random.default_rng(n)
    A = np.full((lda, n), np.nan)
    A[:n, :n] = spd(n, rng)
    A[:n, :n][np.triu_indices(n, 1)] = 7.0  # strict upper: must stay untouched
    ref = np.asfortranarray(A[:n, :n].copy())
    assert oracle.lib().oracle_dpotrf_tile(n, oracle.ptr(ref), n) == 0
    dA = dev_colmajor(A, dev)
    info = torch.full((1,), -7, dtype=torch.int32, device=dev)
    C.dpotrf_tile(dA[:n, :], info)
    got = host_colmajor(dA)
    assert int(info.item()) == 0
    assert rel_err(np.tril(got[:n, :n]), np.tril(ref)) <= TOL
    assert np.all(got[:n, :n][np.triu_indices(n, 1)] == 7.0)
    assert np.all(np.isnan(got[n:, :]))  # padding rows beyond n untouched


@pytest.mark.parametrize("n,bad,want", [(20, 7, 8), (100, 40, 41), (512, 300, 301)])
def test_dpotrf_tile_info(dev, n, bad, want):
    import ttg_amd.chol as C

    A = np.eye(n)
    A[bad, bad] = -1.0
    dA = dev_colmajor(A, dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    C.dpotrf_tile(dA, info, info_offset=1000)
    assert int(info.item()) == 1000 + want
    A = np.eye(n)
    A[bad, bad] = np.nan
    dA = dev_colmajor(A, dev)
    C.dpotrf_tile(dA, info)
    assert int(info.item()) == want


@pytest.mark.parametrize("m,n", [(1, 1), (64, 32), (100, 45), (333, 512), (512, 512), (512, 96), (7, 500)])
def test_dtrsm_rlt(dev, oracle, m, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(m * 1000 + n)
    L = np.asfortranarray(np.linalg.cholesky(spd(n, rng)))
    Lfull = L + np.triu(np.full((n, n), 3.0), 1)  # junk above the diagonal is not referenced
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    ref = B.copy(order="F")
    oracle.lib().oracle_dtrsm_rlt(m, n, oracle.ptr(L), n, oracle.ptr(ref), m)
    dL, dB = dev_colmajor(Lfull, dev), dev_colmajor(B, dev)
    C.dtrsm_rlt(dL, dB)
    assert rel_err(host_colmajor(dB), ref) <= TOL


@pytest.mark.parametrize("n,k", [(1, 1), (32, 64), (45, 37), (128, 512), (512, 512), (300, 17)])
def test_dsyrk_ln(dev, oracle, n, k):
    import ttg_amd.chol as C

    rng = np.random.default_rng(n + 7 * k)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
    Cm = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    ref = Cm.copy(order="F")
    oracle.lib().oracle_dsyrk_ln(n, k, -1.0, oracle.ptr(A), n, 1.0, oracle.ptr(ref), n)
    dA, dC = dev_colmajor(A, dev), dev_colmajor(Cm, dev)
    C.dsyrk_ln(dA, dC)
    got = host_colmajor(dC)
    assert rel_err(np.tril(got), np.tril(ref)) <= TOL
    assert np.array_equal(np.triu(got, 1), np.triu(Cm, 1))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (8, 5, 3), (128, 128, 16), (512, 512, 512), (300, 200, 100),
                                   (16, 512, 512), (37, 45, 29)])
def test_dgemm_nt(dev, oracle, m, n, k):
    import ttg_amd.chol as C

    rng = np.random.default_rng(m + 3 * n + 11 * k)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
    Cm = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    ref = Cm.copy(order="F")
    oracle.lib().oracle_dgemm_nt(m, n, k, -1.0, oracle.ptr(A), m, oracle.ptr(B), n, 1.0, oracle.ptr(ref), m)
    dA, dB, dC = dev_colmajor(A, dev), dev_colmajor(B, dev), dev_colmajor(Cm, dev)
    C.dgemm_nt(dA, dB, dC)
    assert rel_err(host_colmajor(dC), ref) <= TOL


def test_dgemm_nt_alpha_beta(dev, oracle):
    import ttg_amd.chol as C

    rng = np.random.default_rng(5)
    m, n, k = 256, 256, 64
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
    Cm = np.full((m, n), np.nan)
    dA, dB, dC = dev_colmajor(A, dev), dev_colmajor(B, dev), dev_colmajor(Cm, dev)
    C.dgemm_nt(dA, dB, dC, alpha=0.5, beta=0.0)  # beta = 0: C is not read
    assert rel_err(host_colmajor(dC), 0.5 * A @ B.T) <= TOL
    Cm = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    dC = dev_colmajor(Cm, dev)
    C.dgemm_nt(dA, dB, dC, alpha=2.0, beta=-3.0)
    assert rel_err(host_colmajor(dC), 2.0 * A @ B.T - 3.0 * Cm) <= TOL


def test_lapack_fixtures_on_gpu(dev):
    import ttg_amd.chol as C

    lt = np.load(os.path.join(GOLDEN, "lapack_tiles.npz"))
    for n in (1, 7, 32, 45, 64):
        dA = dev_colmajor(lt[f"potrf_A_{n}"], dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        C.dpotrf_tile(dA, info)
        assert int(info.item()) == 0
        assert np.abs(np.tril(host_colmajor(dA)) - lt[f"potrf_L_{n}"]).max() <= 1e-13
    for (m, n) in [(1, 1), (5, 8), (64, 32), (37, 45)]:
        dB = dev_colmajor(lt[f"trsm_B_{m}_{n}"], dev)
        C.dtrsm_rlt(dev_colmajor(lt[f"trsm_L_{m}_{n}"], dev), dB)
        assert np.abs(host_colmajor(dB) - lt[f"trsm_X_{m}_{n}"]).max() <= 1e-12
    for (n, k) in [(1, 1), (8, 5), (32, 64), (45, 37)]:
        dC = dev_colmajor(lt[f"syrk_C_{n}_{k}"], dev)
        C.dsyrk_ln(dev_colmajor(lt[f"syrk_A_{n}_{k}"], dev), dC)
        assert np.abs(np.tril(host_colmajor(dC)) - np.tril(lt[f"syrk_out_{n}_{k}"])).max() <= 1e-12
    for (m, n, k) in [(1, 1, 1), (8, 5, 3), (64, 32, 48), (37, 45, 29)]:
        dC = dev_colmajor(lt[f"gemm_C_{m}_{n}_{k}"], dev)
        C.dgemm_nt(dev_colmajor(lt[f"gemm_A_{m}_{n}_{k}"], dev), dev_colmajor(lt[f"gemm_B_{m}_{n}_{k}"], dev), dC)
        assert np.abs(host_colmajor(dC) - lt[f"gemm_out_{m}_{n}_{k}"]).max() <= 1e-12


def test_group_launches_match_single_launches(dev):
    """Independent tile ops recorded in one group run as batched kernels and
    give bitwise the same results as one launch per op."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(9)
    n = 256
    spds = [dev_colmajor(spd(n, rng), dev) for _ in range(3)]
    Ls = [dev_colmajor(np.linalg.cholesky(spd(n, rng)), dev) for _ in range(2)]
    Bs = [dev_colmajor(rng.uniform(-1, 1, (n, n)), dev) for _ in range(4)]
    Cs = [dev_colmajor(rng.uniform(-1, 1, (n, n)), dev) for _ in range(4)]

    def run(grouped):
        a = [t.clone() for t in spds]
        b = [t.clone() for t in Bs]
        c = [t.clone() for t in Cs]
        a = [x.t().contiguous().t() for x in a]
        b = [x.t().contiguous().t() for x in b]
        c = [x.t().contiguous().t() for x in c]
        info = torch.zeros(3, dtype=torch.int32, device=dev)

        def body():
            for i in range(3):
                C.dpotrf_tile(a[i], info[i:i + 1])
            C.dtrsm_rlt(Ls[0], b[0])
            C.dtrsm_rlt(Ls[1], b[1])
            C.dsyrk_ln(b[2], c[0])
            C.dgemm_nt(b[2], b[3], c[1])
            C.dgemm_nt(b[3], b[2], c[2])
            C.dgemm_nt(b[2], b[2], c[3], alpha=0.5, beta=2.0)

        if grouped:
            with C.group():
                body()
            assert C.group_last_launches() == 5  # potrf, dinv+trsm, two (alpha, beta) update batches
            assert C.group_last_splits() == 0  # independent calls: one part
        else:
            body()
        torch.cuda.synchronize()
        return [x.cpu() for x in a + b + c] + [info.cpu()]

    for x, y in zip(run(False), run(True)):
        assert torch.equal(x, y)


def test_group_keeps_dependent_calls_in_order(dev):
    """A task body that issues dependent calls inside one group -- SYRK into a
    tile, then POTRF of it, then TRSM of a panel by the new factor, then GEMM
    from that panel -- gets them in program order: the group launches what it
    holds whenever a call reads or writes what an earlier one writes (or
    writes what one reads), so the result is bitwise the one of separate
    launches, and each dependent step split the group once."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(21)
    n = 256
    A0 = spd(n, rng)
    W0 = rng.uniform(-1, 1, (n, n)) * 0.05
    P0 = rng.uniform(-1, 1, (n, n))
    G0 = rng.uniform(-1, 1, (n, n))

    def run(grouped):
        a, w, p, g = (dev_colmajor(x, dev) for x in (A0, W0, P0, G0))
        info = torch.zeros(1, dtype=torch.int32, device=dev)

        def body():
            C.dsyrk_ln(w, a)          # a -= w w^T (lower)
            C.dpotrf_tile(a, info)    # factor the updated tile
            C.dtrsm_rlt(a, p)         # p := p a^{-T}
            C.dgemm_nt(p, p, g)       # g -= p p^T

        if grouped:
            with C.group():
                body()
            assert C.group_last_splits() == 3
        else:
            body()
        torch.cuda.synchronize()
        assert int(info.item()) == 0
        return [x.cpu() for x in (a, p, g)]

    for x, y in zip(run(False), run(True)):
        assert torch.equal(x, y)


def test_dnrm2(dev):
    import ttg_amd.chol as C

    x = torch.randn(512 * 512, dtype=torch.float64, device=dev)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    C.dnrm2(x, out)
    assert abs(out.item() - torch.linalg.norm(x).item()) <= 1e-12 * out.item()


def test_dnrm2_in_group_sees_written_values(dev):
    """Inside a launch group a norm of a tile that an operation recorded
    earlier in the group writes is taken after the group's launches (it sees
    the written tile); the norm of the same tile before its operation, and of
    a read-only input, is taken at once (the per-task norms of the
    tile-norm debug check)."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(12)
    n = 256
    A = dev_colmajor(spd(n, rng), dev)
    L = dev_colmajor(np.linalg.cholesky(spd(n, rng)), dev)
    B = dev_colmajor(rng.uniform(-1, 1, (n, n)), dev)
    A0, B0 = A.cpu().numpy().copy(), B.cpu().numpy().copy()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    norms = torch.zeros(5, dtype=torch.float64, device=dev)
    with C.group():
        C.dnrm2(A, norms[0:1])
        C.dpotrf_tile(A, info)
        C.dnrm2(A, norms[1:2])
        C.dnrm2(B, norms[2:3])
        C.dtrsm_rlt(L, B)
        C.dnrm2(B, norms[3:4])
        C.dnrm2(L, norms[4:5])
    torch.cuda.synchronize()
    got = norms.cpu().numpy()
    A1, B1 = A.cpu().numpy(), B.cpu().numpy()
    assert int(info.item()) == 0
    assert not np.array_equal(A0, A1) and not np.array_equal(B0, B1)
    want = [np.linalg.norm(A0), np.linalg.norm(A1), np.linalg.norm(B0), np.linalg.norm(B1),
            np.linalg.norm(L.cpu().numpy())]
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-12 * w, (got, want)


def test_plgsy_tiles_in_group_match_single_launches(dev, oracle):
    """PLGSY tiles recorded in a launch group (one batched launch) are
    bitwise the tiles of one launch each, and those of the oracle's
    generator, ragged edge tiles and diagonal bumps included."""
    import ttg_amd.chol as C

    N, NB, bump, seed = 1000, 256, 1000.0, 3872
    nt = oracle.nt_of(N, NB)
    want = oracle.plgsy_tiles(N, NB, bump, seed)
    keys = [(m, n) for n in range(nt) for m in range(n, nt)]

    def gen(grouped):
        tiles = {k: torch.full((NB, NB), float("nan"), dtype=torch.float64, device=dev).t() for k in keys}

        def body():
            for (m, n), t in tiles.items():
                mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
                C.dplgsy_tile(t[:mb, :nb], bump, N, m * NB, n * NB, seed)

        if grouped:
            with C.group():
                body()
            assert C.group_last_launches() == 1
        else:
            body()
        torch.cuda.synchronize()
        return {k: t.cpu().numpy() for k, t in tiles.items()}

    single, batched = gen(False), gen(True)
    for (m, n) in keys:
        mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
        a, b = single[(m, n)][:mb, :nb], batched[(m, n)][:mb, :nb]
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (m, n)
        ref = oracle.tile_view(want, N, NB, m, n)[:mb, :nb]
        assert np.array_equal(a.view(np.uint64), ref.view(np.uint64)), (m, n)


def test_float_tile_ops_against_fp64_oracle(dev, oracle):
    """The FP32 tile operations of the graph over float tiles: GEMM and SYRK
    bitwise on small integers (exact in FP32), POTRF and TRSM within FP32
    rounding of the FP64 oracle (8 n eps_32 relative)."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(4)
    n, k = 256, 200

    def cm32(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32).T)).to(dev).t()

    def host(t):
        torch.cuda.synchronize()
        return t.t().cpu().numpy().T.astype(np.float64)

    lib = C.lib()
    A = rng.integers(-8, 9, size=(n, k)).astype(np.float64)
    Bm = rng.integers(-8, 9, size=(n, k)).astype(np.float64)
    C0 = rng.integers(-8, 9, size=(n, n)).astype(np.float64)
    dA, dB, dC = cm32(A), cm32(Bm), cm32(C0)
    assert lib.ttg_chol_sgemm_nt(None, n, n, k, -1.0, dA.data_ptr(), n, dB.data_ptr(), n, 1.0, dC.data_ptr(), n) == 0
    assert np.array_equal(host(dC), C0 - A @ Bm.T)
    dC = cm32(C0)
    assert lib.ttg_chol_ssyrk_ln(None, n, k, -1.0, dA.data_ptr(), n, 1.0, dC.data_ptr(), n) == 0
    got = host(dC)
    assert np.array_equal(np.tril(got), np.tril(C0 - A @ A.T))
    assert np.array_equal(np.triu(got, 1), np.triu(C0, 1))
    # POTRF and TRSM
    M = rng.uniform(-1, 1, (n, n))
    S = M @ M.T / n + np.eye(n)
    S32 = S.astype(np.float32).astype(np.float64)
    want = np.asfortranarray(S32.copy())
    assert oracle.lib().oracle_dpotrf_tile(n, oracle.ptr(want), n) == 0
    L64 = np.tril(want)
    dS = cm32(S32)
    info = torch.full((1,), -3, dtype=torch.int32, device=dev)
    assert lib.ttg_chol_spotrf_tile(None, n, dS.data_ptr(), n, info.data_ptr(), 0) == 0
    L32 = np.tril(host(dS))
    assert int(info.item()) == 0
    eps32 = np.finfo(np.float32).eps
    assert np.abs(L32 - L64).max() <= 8 * n * eps32 * np.abs(L64).max()
    Bt = rng.uniform(-1, 1, (100, n))
    dBt = cm32(Bt)
    dL = cm32(L32)
    assert lib.ttg_chol_strsm_rlt(None, 100, n, dL.data_ptr(), n, dBt.data_ptr(), 100) == 0
    X = host(dBt)
    assert np.abs(X @ L32.T - Bt.astype(np.float32)).max() <= 8 * n * eps32 * np.abs(Bt).max() * np.abs(L32).max()
