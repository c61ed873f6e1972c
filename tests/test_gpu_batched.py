"""Batched variable-size FP32 factor + solve (one persistent launch) against
the oracle's FP64 tile factorization (pinned to LAPACK): per matrix, the
factor agrees with the oracle's to the FP32 bar max|L - L_oracle| / max|L_oracle|
<= 16 n u32, and the solve reaches backward error <= 16 n u32, where
u32 = 2^-24.  Sizes cover 1, the block edges (31/32/33, 63/64/65, 511-513)
and the maximum 1024; a non-SPD matrix reports its pivot and leaves its
right-hand side alone without disturbing the others; every launch
configuration gives the same bits."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24

SIZES = [1, 2, 5, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 511, 512, 513, 700, 1000, 1023, 1024]


def make_batch(sizes, seed):
    rng = np.random.default_rng(seed)
    mats, rhs = [], []
    for n in sizes:
        B = rng.uniform(-1, 1, (n, n))
        mats.append((B @ B.T / n + np.eye(n)).astype(np.float32))
        rhs.append(rng.uniform(-1, 1, n).astype(np.float32))
    return mats, rhs


def to_dev(mats, rhs, dev):
    # a symmetric matrix is its own transpose: row-major storage is column-major A
    return ([torch.from_numpy(m.copy()).to(dev) for m in mats], [torch.from_numpy(v.copy()).to(dev) for v in rhs])


def oracle_factor(oracle, a):
    n = a.shape[0]
    L = np.asfortranarray(a.astype(np.float64))
    assert oracle.lib().oracle_dpotrf_tile(n, oracle.ptr(L), n) == 0
    return np.tril(L)


def run(mats_d, rhs_d, dev, cfg=-1):
    import ttg_amd.chol as C

    info = torch.full((len(mats_d),), -7, dtype=torch.int32, device=dev)
    batch = C.Batch(mats_d, rhs_d)
    print(f"sposv_batched count={batch.count} cfg={cfg}", flush=True)
    C.sposv_batched(batch, info, cfg=cfg)
    torch.cuda.synchronize()
    return info.cpu().numpy()


def test_batched_vs_oracle(dev, oracle):
    sizes = SIZES + list(np.random.default_rng(3).integers(32, 1025, 40))
    mats, rhs = make_batch(sizes, 11)
    mats_d, rhs_d = to_dev(mats, rhs, dev)
    info = run(mats_d, rhs_d, dev)
    assert (info == 0).all(), info
    for i, n in enumerate(sizes):
        Lg = np.tril(mats_d[i].cpu().numpy().T.astype(np.float64))  # column-major lower
        Lo = oracle_factor(oracle, mats[i])
        assert np.abs(Lg - Lo).max() / np.abs(Lo).max() <= 16 * n * U32, (i, n)
        A = mats[i].astype(np.float64)
        x = rhs_d[i].cpu().numpy().astype(np.float64)
        b = rhs[i].astype(np.float64)
        berr = np.abs(b - A @ x).max() / (np.abs(A).sum(1).max() * np.abs(x).max() + np.abs(b).max())
        assert berr <= 16 * n * U32, (i, n, berr)


def test_batched_nonspd_reports_pivot(dev):
    sizes = [64, 300, 40, 1024]
    mats, rhs = make_batch(sizes, 5)
    bad = np.eye(300, dtype=np.float32)
    bad[137, 137] = -1.0  # first non-positive pivot at column 138
    mats[1] = bad
    mats_d, rhs_d = to_dev(mats, rhs, dev)
    info = run(mats_d, rhs_d, dev)
    assert list(info) == [0, 138, 0, 0]
    assert np.array_equal(rhs_d[1].cpu().numpy(), rhs[1])
    for i in (0, 2, 3):
        A = mats[i].astype(np.float64)
        x = rhs_d[i].cpu().numpy().astype(np.float64)
        assert np.abs(A @ x - rhs[i]).max() <= 1e-3


def test_batched_configs_bitwise_equal(dev):
    import ttg_amd.chol as C

    sizes = [1024, 33, 500, 32, 257, 1, 900]
    mats, rhs = make_batch(sizes, 7)
    outs = []
    for cfg in range(len(C.batched_configs())):
        mats_d, rhs_d = to_dev(mats, rhs, dev)
        info = run(mats_d, rhs_d, dev, cfg)
        assert (info == 0).all()
        outs.append([t.cpu().numpy() for t in mats_d + rhs_d])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_batched_many_small_and_empty(dev):
    import ttg_amd.chol as C

    sizes = list(np.random.default_rng(9).integers(1, 48, 3000))
    mats, rhs = make_batch(sizes, 13)
    mats_d, rhs_d = to_dev(mats, rhs, dev)
    info = run(mats_d, rhs_d, dev)
    assert (info == 0).all()
    for i in range(0, 3000, 97):
        A = mats[i].astype(np.float64)
        x = rhs_d[i].cpu().numpy().astype(np.float64)
        assert np.abs(A @ x - rhs[i]).max() <= 1e-4 * max(1.0, np.abs(x).max())
    empty = torch.zeros(1, dtype=torch.int32, device=dev)
    C.sposv_batched(C.Batch([], []), empty)
