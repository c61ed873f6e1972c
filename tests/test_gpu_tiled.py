"""Tiled drivers on the GPU: the FP64 factorization of the examples/potrf
graph with grouped launches, the generators, solves and residuals, against
the oracle.  FP64 bar: max |gpu - oracle| / max |oracle| <= 1e-12 over the
lower tiles; bit-exact for the generators."""
import numpy as np
import pytest
import torch

from gpu_util import lower_tiles_rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("N,NB", [(160, 32), (1000, 128), (2048, 512)])
def test_plgsy_bit_exact(dev, oracle, N, NB):
    import ttg_amd.chol as C

    want = oracle.plgsy_tiles(N, NB, N, 3872)
    t = torch.full((want.size,), np.nan, dtype=torch.float64, device=dev)
    C.plgsy_tiled(t, N, NB, N, 3872)
    got = to_host(t)
    nt = oracle.nt_of(N, NB)
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            a = oracle.tile_view(got, N, NB, m, n)[:mb, :nb]
            b = oracle.tile_view(want, N, NB, m, n)[:mb, :nb]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (m, n)


@pytest.mark.parametrize("N,NB,alpha", [(256, 64, 1.0), (300, 128, 1e-3), (1024, 256, 1.0)])
def test_spd_bit_exact(dev, oracle, N, NB, alpha):
    import ttg_amd.chol as C

    want = oracle.spd_tiles(N, NB, alpha, 42)
    t = torch.zeros(want.size, dtype=torch.float64, device=dev)
    work = torch.empty(N * N, dtype=torch.float64, device=dev)
    C.spd_tiled(t, N, NB, alpha, 42, work)
    got = to_host(t)
    nt = oracle.nt_of(N, NB)
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            a = oracle.tile_view(got, N, NB, m, n)[:mb, :nb]
            b = oracle.tile_view(want, N, NB, m, n)[:mb, :nb]
            # diagonal tiles are written whole (both triangles), exactly
            assert np.array_equal(a, b), (m, n)


@pytest.mark.parametrize("N,NB,alpha", [(256, 64, 1.0), (300, 128, 1e-3), (2048, 512, 1e-4)])
def test_spd_uniform_vs_oracle(dev, oracle, N, NB, alpha):
    """B FP64 uniform: B is bit-identical, so the product differs from the
    oracle's only by summation order (a few ulps of the largest entry)."""
    import ttg_amd.chol as C

    want = oracle.spd_uniform_tiles(N, NB, alpha, 2024)
    t = torch.zeros(want.size, dtype=torch.float64, device=dev)
    work = torch.empty(N * N, dtype=torch.float64, device=dev)
    C.spd_uniform_tiled(t, N, NB, alpha, 2024, work)
    got = to_host(t)
    Bh = work.cpu().numpy().reshape(N, N).T  # column-major B
    assert np.array_equal(Bh[:5, :3], np.array([[oracle.lib().oracle_spd_uval(i, j, N, 2024) for j in range(3)]
                                                for i in range(5)]))
    nt = oracle.nt_of(N, NB)
    scale = 0.0
    err = 0.0
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            a = oracle.tile_view(got, N, NB, m, n)[:mb, :nb]
            b = oracle.tile_view(want, N, NB, m, n)[:mb, :nb]
            err = max(err, np.abs(a - b).max())
            scale = max(scale, np.abs(b).max())
    # N products of magnitude <= 1 summed in another order, over N: rounding
    # of order sqrt(N) eps of the largest entry (a wrong operand, scale or
    # shift is off by far more)
    assert err <= 16 * np.finfo(np.float64).eps * scale * np.sqrt(N), (err, scale)


@pytest.mark.parametrize("N,NB,gen", [(160, 32, "plgsy"), (1000, 128, "spd"), (2048, 256, "spd"),
                                      (4096, 512, "spd"), (1536, 512, "plgsy"), (4096, 256, "uniform")])
def test_dpotrf_tiled_vs_oracle(dev, oracle, N, NB, gen):
    """(4096, 256, uniform) is cfg1: A = B B^T/N + I with B FP64 uniform."""
    import ttg_amd.chol as C

    A0 = {"plgsy": lambda: oracle.plgsy_tiles(N, NB, N, 3872), "spd": lambda: oracle.spd_tiles(N, NB, 1.0, 5),
          "uniform": lambda: oracle.spd_uniform_tiles(N, NB, 1.0, 2024)}[gen]()
    ref = A0.copy()
    assert oracle.potrf_tiled(N, NB, ref) == 0
    t = to_dev(A0, dev)
    info = torch.full((1,), -3, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    got = to_host(t)
    assert int(info.item()) == 0
    assert lower_tiles_rel_err(oracle, N, NB, got, ref) <= TOL
    assert oracle.check_dpotrf_tiled(N, NB, got, A0) <= 60  # testing_dpotrf.cc:315-324


def test_dpotrf_tiled_flags_bitwise_identical(dev, oracle):
    """Lookahead and graph capture change only the launch schedule: every tile
    sees the same kernels in the same order, so the factor is bitwise equal."""
    import ttg_amd.chol as C

    N, NB = 3072, 512
    A0 = to_dev(oracle.spd_tiles(N, NB, 1.0, 8), dev)
    outs = []
    for flags in (0, C.F_LOOKAHEAD, C.F_GRAPH, C.F_DEFAULT):
        for rep in range(2):  # a captured graph is replayed on the second call
            t = A0.clone()
            info = torch.full((1,), 5, dtype=torch.int32, device=dev)
            C.dpotrf_tiled(t, N, NB, info, flags=flags)
            torch.cuda.synchronize()
            assert int(info.item()) == 0
            outs.append(t.cpu())
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


def test_dpotrf_tiled_replay_same_buffer(dev, oracle):
    import ttg_amd.chol as C

    N, NB = 2048, 512
    A0 = oracle.spd_tiles(N, NB, 1.0, 9)
    t = to_dev(A0, dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    first = None
    for _ in range(3):
        t.copy_(to_dev(A0, dev))
        C.dpotrf_tiled(t, N, NB, info)
        got = to_host(t)
        if first is None:
            first = got
        assert np.array_equal(got, first)


@pytest.mark.parametrize("N,NB,k,local", [(1024, 256, 2, 17), (2048, 512, 3, 500)])
def test_dpotrf_tiled_info_nonspd(dev, oracle, N, NB, k, local):
    import ttg_amd.chol as C

    A0 = oracle.spd_tiles(N, NB, 1.0, 3)
    oracle.tile_view(A0, N, NB, k, k)[local, local] = -1e6
    want = oracle.potrf_tiled(N, NB, A0.copy())
    assert want == k * NB + local + 1
    t = to_dev(A0, dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    assert int(info.item()) == want
    # a later good factorization on the same plan resets info
    t.copy_(to_dev(oracle.spd_tiles(N, NB, 1.0, 4), dev))
    C.dpotrf_tiled(t, N, NB, info)
    assert int(info.item()) == 0


@pytest.mark.parametrize("N,NB,nrhs", [(160, 32, 1), (1000, 128, 3), (2048, 512, 16), (1536, 256, 40),
                                       (2600, 512, 20), (1024, 512, 1)])
def test_dpotrs_and_residual_vs_oracle(dev, oracle, N, NB, nrhs):
    import ttg_amd.chol as C

    A0 = oracle.spd_tiles(N, NB, 0.5, 21)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    rng = np.random.default_rng(N + nrhs)
    B = np.asfortranarray(rng.uniform(-1, 1, (N, nrhs)))
    Xref = oracle.potrs_tiled(N, NB, L, B)
    dL = to_dev(L, dev)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev).t()  # column-major N x nrhs
    X = dB.clone().t().contiguous().t()
    C.dpotrs_tiled(dL, N, NB, X)
    Xg = X.t().cpu().numpy().T
    assert np.abs(Xg - Xref).max() / np.abs(Xref).max() <= TOL
    dA = to_dev(A0, dev)
    R = torch.empty_like(X)
    C.dresidual_tiled(dA, N, NB, X, dB, R)
    Rg = R.t().cpu().numpy().T
    Rref = oracle.residual_tiled(N, NB, A0, Xg, B)
    assert np.abs(Rg - Rref).max() <= 1e-13
    assert oracle.solve_backward_error(N, NB, A0, Xg, B) <= 1e-14


def _solve_with_sweep(C, fn, L, N, NB, B, dev, sweep):
    """fn (dpotrs_tiled / spotrs_tiled) on a fresh copy of B with the SWEEP
    knob set (plans rebuilt), as a host array."""
    C.set_knob("sweep", sweep)
    C.release_plans()
    try:
        X = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev).t()
        fn(L, N, NB, X)
        return X.t().cpu().numpy().T
    finally:
        C.set_knob("sweep", None)
        C.release_plans()


@pytest.mark.parametrize("N,NB,nrhs", [(160, 32, 1), (1000, 128, 3), (2048, 512, 16), (1600, 512, 5),
                                       (3072, 256, 16), (4096, 512, 33)])
def test_sweep_forms_agree(dev, oracle, N, NB, nrhs):
    """Both triangular sweeps as one persistent launch (SWEEP = 1, 2 and 4
    workgroups per CU: the same summation order, so bitwise one result) and
    as per-step launches (SWEEP = 0): FP64 within the FP64 bar of the oracle's
    solve; FP32 (the factor rounded to float) within 1e-5 of the oracle's FP64
    solve with that same rounded factor (A = B B^T/N + I/2, condition ~4)."""
    import ttg_amd.chol as C

    A0 = oracle.spd_tiles(N, NB, 0.5, 5 + N)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    rng = np.random.default_rng(N + 7 * nrhs)
    B = np.asfortranarray(rng.uniform(-1, 1, (N, nrhs)))
    Xref = oracle.potrs_tiled(N, NB, L, B)
    dL = to_dev(L, dev)
    got = {sw: _solve_with_sweep(C, C.dpotrs_tiled, dL, N, NB, B, dev, sw) for sw in (0, 1, 2, 4)}
    for sw, X in got.items():
        assert np.abs(X - Xref).max() / np.abs(Xref).max() <= TOL, f"FP64 sweep={sw}"
    assert np.array_equal(got[1], got[2]) and np.array_equal(got[2], got[4])
    L32 = L.astype(np.float32)
    Xref32 = oracle.potrs_tiled(N, NB, L32.astype(np.float64), B)
    dL32 = to_dev(L32, dev)
    got32 = {sw: _solve_with_sweep(C, C.spotrs_tiled, dL32, N, NB, B, dev, sw) for sw in (0, 1, 2, 4)}
    for sw, X in got32.items():
        assert np.abs(X - Xref32).max() / np.abs(Xref32).max() <= 1e-5, f"FP32 sweep={sw}"
    assert np.array_equal(got32[1], got32[2]) and np.array_equal(got32[2], got32[4])


def test_large_factor_solve_backward_error(dev):
    """Size-independent property at a realistic size: factor N=16384, solve,
    and the FP64 residual gives a backward error at the level of eps."""
    import ttg_amd.chol as C

    N, NB, nrhs = 16384, 512, 4
    t = C.alloc_tiles(N, NB)
    work = torch.empty(N * N, dtype=torch.float64, device=dev)
    C.spd_tiled(t, N, NB, 1.0, 77, work)
    del work
    A0 = t.clone()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    assert int(info.item()) == 0
    g = torch.Generator(device="cpu").manual_seed(1)
    B = (torch.rand(nrhs, N, generator=g, dtype=torch.float64) * 2 - 1).to(dev).t()
    X = B.clone().t().contiguous().t()
    C.dpotrs_tiled(t, N, NB, X)
    R = torch.empty_like(X)
    C.dresidual_tiled(A0, N, NB, X, B, R)
    torch.cuda.synchronize()
    # ||A||_inf <= N * max|a_ij| ~ N/3 + 1 for this construction; use the exact bound via row sums later
    rn = R.abs().max(dim=0).values
    xn = X.abs().max(dim=0).values
    bn = B.abs().max(dim=0).values
    anorm = 1.0 + N * (1.0 / 3.0 + 0.05)  # upper bound of ||A||_inf for B B^T/N + I with |b| <= 1
    berr = (rn / (anorm * xn + bn)).max().item()
    assert berr <= 1e-15


def test_dpotrf_tiled_full_size_sampled_residual(dev):
    """The FP64 factorization at its full single-GPU size (N=32768, NB=512):
    entries of L L^T - A on 48 sampled rows (every column up to the row),
    held to the reference's check_dpotrf bound entry by entry,
    |(L L^T - A)_ij| <= 60 N eps max|A| (testing_dpotrf.cc:315-324), with the
    products on the GPU.  The oracle cannot factor this size in test time."""
    import ttg_amd.chol as C

    N, NB = 32768, 512
    nt = N // NB
    t = C.alloc_tiles(N, NB)
    work = torch.empty(N * N, dtype=torch.float64, device=dev)
    C.spd_tiled(t, N, NB, 1.0, 17, work)
    del work
    A0 = t.clone()
    info = torch.full((1,), -3, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0

    def tidx(m, n):
        return n * nt - n * (n - 1) // 2 + (m - n)

    def row(tiles, i, upto):
        """Row i of the lower triangle, columns 0..upto (tile-packed, column-major tiles)."""
        m, r = divmod(i, NB)
        parts = [tiles[tidx(m, n) * NB * NB:(tidx(m, n) + 1) * NB * NB].view(NB, NB)[:, r] for n in range(m + 1)]
        return torch.cat(parts)[:upto + 1]

    g = torch.Generator().manual_seed(5)
    rows = sorted(set(int(v) for v in torch.randint(0, N, (48,), generator=g)))
    amax = max(A0[tidx(k, k) * NB * NB:(tidx(k, k) + 1) * NB * NB].abs().max().item() for k in range(nt))
    bound = 60 * N * np.finfo(np.float64).eps * amax
    R = torch.zeros(len(rows), N, dtype=torch.float64, device=dev)
    for a, i in enumerate(rows):
        R[a, :i + 1] = row(t, i, i)
    P = R @ R.t()  # (L L^T) on the sampled rows and columns
    worst = 0.0
    for a, i in enumerate(rows):
        for b, j in enumerate(rows[:a + 1]):
            m, n = i // NB, j // NB
            aij = A0[tidx(m, n) * NB * NB + (i % NB) + (j % NB) * NB].item()
            worst = max(worst, abs(P[a, b].item() - aij))
    assert worst <= bound, (worst, bound)
