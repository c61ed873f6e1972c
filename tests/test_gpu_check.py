"""The device factor check (testing_dpotrf.cc:277-335 in tiles, L L^T on
FP64 MFMA) against planted exact cases and the oracle's check, the exported
mixed-precision factor, and the device backward error the benchmark reports
against the oracle's."""
import numpy as np
import pytest
import torch

from gpu_util import dev_colmajor, to_dev

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _stored(N, NB, S):
    D = np.zeros_like(S)
    for i in range(N):
        D[i, :min(N, (i // NB + 1) * NB)] = S[i, :min(N, (i // NB + 1) * NB)]
    return D


@pytest.mark.parametrize("N,NB", [(256, 64), (300, 128), (1024, 512)])
def test_device_check_planted_closed_form(dev, oracle, N, NB):
    """Integer L, dyadic symmetric E, A0 = L L^T - E exactly: the device
    returns ||E||_inf / (||A0||_inf N eps), ||A0||_inf over the stored lower
    tiles with the diagonal tiles whole (getLAPACKMatrix)."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(N)
    L = np.tril(rng.integers(-2, 3, (N, N)).astype(np.float64), -1)
    L[np.diag_indices(N)] = 8.0 + rng.integers(0, 3, N)
    L[1:, 0] = 6.0
    E = rng.integers(-3, 4, (N, N)).astype(np.float64) * 2.0 ** -30
    E[rng.random((N, N)) > 0.2] = 0.0
    E = np.tril(E) + np.tril(E, -1).T
    S = L @ L.T - E
    A0 = _stored(N, NB, S)
    Lst = _stored(N, NB, np.tril(L) + np.triu(S, 1))
    want_a = np.abs(A0).sum(1).max()
    want_r = np.abs(E).sum(1).max()
    a, r, q = C.dcheck_dpotrf_tiled(to_dev(oracle.dense_to_tiles(N, NB, A0), dev),
                                    to_dev(oracle.dense_to_tiles(N, NB, Lst), dev), N, NB)
    assert a == want_a
    assert r == pytest.approx(want_r, rel=1e-13)
    assert q == pytest.approx(want_r / (want_a * N * EPS), rel=1e-13)


@pytest.mark.parametrize("N,NB,gen", [(160, 32, "plgsy"), (1000, 128, "spd"), (2048, 512, "uniform")])
def test_device_check_matches_oracle_on_a_factor(dev, oracle, N, NB, gen):
    """On a computed factor the residual is rounding noise summed in another
    order, so the two checks agree in the denominator to rounding and in the
    quotient in size (both within the reference's bound of 60)."""
    import ttg_amd.chol as C

    A0 = {"plgsy": lambda: oracle.plgsy_tiles(N, NB, N, 3872), "spd": lambda: oracle.spd_tiles(N, NB, 1.0, 5),
          "uniform": lambda: oracle.spd_uniform_tiles(N, NB, 1.0, 2024)}[gen]()
    t = to_dev(A0, dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    C.dpotrf_tiled(t, N, NB, info)
    assert int(info.item()) == 0
    a, r, q = C.dcheck_dpotrf_tiled(to_dev(A0, dev), t, N, NB)
    want = oracle.check_dpotrf_tiled(N, NB, t.cpu().numpy(), A0)
    D = oracle.tiles_to_dense(N, NB, A0)
    assert a == pytest.approx(np.abs(D).sum(1).max(), rel=1e-14)
    assert 0 < q <= 60 and 0 < want <= 60
    assert 0.33 < q / want < 3.0
    # a planted error in the factor is caught by both
    bad = t.clone()
    oracle_bad = bad.cpu().numpy()
    oracle.tile_view(oracle_bad, N, NB, 1, 0)[3, 5] += 1e-8
    bad.copy_(torch.from_numpy(oracle_bad))
    _, _, qb = C.dcheck_dpotrf_tiled(to_dev(A0, dev), bad, N, NB)
    assert qb > 60 and oracle.check_dpotrf_tiled(N, NB, oracle_bad, A0) > 60
    assert qb == pytest.approx(oracle.check_dpotrf_tiled(N, NB, oracle_bad, A0), rel=1e-3)


def test_device_check_nan_propagates(dev, oracle):
    import ttg_amd.chol as C

    N, NB = 256, 64
    A0 = oracle.spd_tiles(N, NB, 1.0, 3)
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    oracle.tile_view(L, N, NB, 2, 1)[0, 0] = np.nan
    _, r, q = C.dcheck_dpotrf_tiled(to_dev(A0, dev), to_dev(L, dev), N, NB)
    assert np.isnan(r) and np.isnan(q)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_mixed_factor_export(dev, oracle, prec):
    """The exported working factor: FP64 plan equal to the oracle's factor at
    the FP64 bar, FP32 / split-FP16 within the FP32 factor's distance, and
    the device check of each within its precision's bound."""
    import ttg_amd.chol as C

    N, NB, nrhs = 1024, 256, 2
    A0 = oracle.spd_tiles(N, NB, 1.0, 13)
    ref = A0.copy()
    assert oracle.potrf_tiled(N, NB, ref) == 0
    dA = to_dev(A0, dev)
    B = dev_colmajor(np.random.default_rng(0).uniform(-1, 1, (N, nrhs)), dev)
    X = torch.zeros(nrhs, N, dtype=torch.float64, device=dev).t()
    C.dposv_mixed(dA, N, NB, B, X, prec=prec)
    L = C.alloc_tiles(N, NB)
    C.dposv_mixed_factor(L, N, NB, nrhs, prec)
    from gpu_util import lower_tiles_rel_err

    err = lower_tiles_rel_err(oracle, N, NB, L.cpu().numpy(), ref)
    assert err <= (1e-12 if prec == 0 else 1e-5), err
    _, _, q = C.dcheck_dpotrf_tiled(dA, L, N, NB)
    if prec == 0:
        assert q <= 60
    else:  # an FP32 factor: ||A - L L^T|| / ||A|| near the FP32 unit roundoff
        assert q * N * EPS <= 1e-5


def test_mixed_factor_export_needs_a_solve(dev):
    import ttg_amd.chol as C

    L = C.alloc_tiles(384, 128)
    with pytest.raises(C.CholError, match="no dposv_mixed call"):
        C.dposv_mixed_factor(L, 384, 128, 5, C.PREC_FP32)


@pytest.mark.parametrize("gen", ["spd", "uniform", "plgsy"])
def test_device_backward_error_matches_oracle(dev, oracle, gen):
    """C.backward_error (what the benchmark reports) against the oracle's
    solve backward error on the same solution, N=2048."""
    import ttg_amd.chol as C

    N, NB, nrhs = 2048, 256, 4
    A0 = {"plgsy": lambda: oracle.plgsy_tiles(N, NB, N, 3872), "spd": lambda: oracle.spd_tiles(N, NB, 1e-3, 5),
          "uniform": lambda: oracle.spd_uniform_tiles(N, NB, 1e-3, 2024)}[gen]()
    rng = np.random.default_rng(7)
    Bh = rng.uniform(-1, 1, (N, nrhs))
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    Xh = oracle.potrs_tiled(N, NB, L, Bh)
    Xh[17, 2] += 1e-9  # a planted error so the quantity is not pure rounding
    dA = to_dev(A0, dev)
    got = C.backward_error(dA, N, NB, dev_colmajor(Xh, dev), dev_colmajor(Bh, dev))
    want = oracle.solve_backward_error(N, NB, A0, Xh, Bh)
    assert got == pytest.approx(want, rel=1e-6)
