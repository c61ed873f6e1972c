"""Mixed-precision solve with FP64 iterative refinement (no reference
counterpart): pinned by the FP64 backward error recomputed by the oracle on
the CPU, and by agreement with the oracle's FP64 solution within the
forward-error bound kappa * tol."""
import numpy as np
import pytest
import torch

from gpu_util import to_dev

pytestmark = pytest.mark.gpu


def sigma2():
    return sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2


@pytest.mark.parametrize("N,NB,nrhs,kappa,prec", [(1024, 256, 4, 1e4, 1), (2048, 512, 16, 1e4, 1),
                                                  (1000, 128, 3, 1e3, 1), (1536, 512, 2, 1e4, 0),
                                                  (1024, 256, 4, 1e4, 2), (2048, 512, 16, 1e4, 2),
                                                  (1000, 128, 3, 1e3, 2), (3072, 512, 8, 1e5, 2),
                                                  (4096, 512, 8, 1e6, 2), (1024, 256, 4, 10, 3),
                                                  (2048, 512, 2, 30, 3), (1024, 256, 4, 10, 4)])
def test_dposv_mixed_reaches_fp64_backward_error(dev, oracle, N, NB, nrhs, kappa, prec):
    import ttg_amd.chol as C

    alpha = 4 * sigma2() / (kappa - 1)
    A0 = oracle.spd_tiles(N, NB, alpha, 13)
    rng = np.random.default_rng(N)
    Bh = rng.uniform(-1, 1, (N, nrhs))
    dA = to_dev(A0, dev)
    dB = torch.from_numpy(np.ascontiguousarray(Bh.T)).to(dev).t()
    X = torch.zeros(nrhs, N, dtype=torch.float64, device=dev).t()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    its, berr = C.dposv_mixed(dA, N, NB, dB, X, prec=prec, tol=1e-12, max_iter=30, info=info)
    assert int(info.item()) == 0
    assert berr <= 1e-12
    Xh = X.t().cpu().numpy().T
    # the oracle recomputes the backward error from scratch
    assert oracle.solve_backward_error(N, NB, A0, Xh, Bh) <= 1e-12
    # and the solution agrees with the FP64 oracle solve within kappa * 1e-12
    L = A0.copy()
    assert oracle.potrf_tiled(N, NB, L) == 0
    Xref = oracle.potrs_tiled(N, NB, L, Bh)
    assert np.abs(Xh - Xref).max() / np.abs(Xref).max() <= kappa * 1e-11
    # A is left unchanged
    assert np.array_equal(dA.cpu().numpy(), A0)
    if prec != 0:
        assert its >= 1  # a low-precision factor needs refinement to reach 1e-12


def test_dposv_mixed_nonspd_reports_info(dev, oracle):
    import ttg_amd.chol as C

    N, NB = 1024, 256
    A0 = oracle.spd_tiles(N, NB, 1.0, 3)
    oracle.tile_view(A0, N, NB, 1, 1)[10, 10] = -5.0
    dA = to_dev(A0, dev)
    dB = torch.ones(N, 1, dtype=torch.float64, device=dev)
    X = torch.zeros(N, 1, dtype=torch.float64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    its, berr = C.dposv_mixed(dA, N, NB, dB, X, prec=1, info=info)
    assert int(info.item()) == NB + 11


def test_refinement_stops_when_stagnating(dev, oracle):
    """An unreachable tolerance (1e-300) with 30 allowed steps: the FP32
    refinement reaches the FP64 floor within a few steps, and the loop then
    stops after two steps that do not halve the backward error (ENOCONV)
    instead of running all 30 FP64 residuals and solves."""
    import ttg_amd.chol as C

    N, NB, r = 2048, 512, 2
    A0 = oracle.spd_tiles(N, NB, 0.5, 13)
    Bh = np.random.default_rng(2).uniform(-1, 1, (N, r))
    dA = to_dev(A0, dev)
    dB = torch.from_numpy(np.ascontiguousarray(Bh.T)).to(dev).t()
    X = torch.zeros(r, N, dtype=torch.float64, device=dev).t()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    with pytest.raises(C.CholError):
        C.dposv_mixed(dA, N, NB, dB, X, prec=C.PREC_FP32, tol=1e-300, max_iter=30, info=info)
    its, berr = C.dposv_mixed(dA, N, NB, dB, X, prec=C.PREC_FP32, tol=1e-300, max_iter=30, info=info, check=False)
    assert int(info.item()) == 0
    assert 2 <= its <= 8, its
    Xh = X.t().cpu().numpy().T
    assert oracle.solve_backward_error(N, NB, A0, Xh, Bh) <= 1e-15
    assert berr <= 1e-15


@pytest.mark.parametrize("prec,kappa,shift", [(3, 1e4, 0.0), (4, 1e4, 0.0), (3, 1e6, 1e-3), (1, 1e4, 0.0)])
def test_krylov_solve_reaches_tolerance(dev, oracle, prec, kappa, shift):
    """Preconditioned CG with the low-precision factor (FP16 or BF16 single
    products, FP32) reaches the FP64 backward error 1e-12, checked by the
    oracle; at kappa 1e6 the FP16 factor of A itself meets a non-positive
    pivot, and the factor of A with its diagonal scaled by 1 + 1e-3 serves."""
    import ttg_amd.chol as C

    N, NB, r = 4096, 512, 4
    sigma2 = sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2
    A0 = oracle.spd_tiles(N, NB, 4 * sigma2 / (kappa - 1), 41)
    Bh = np.random.default_rng(8).uniform(-1, 1, (N, r))
    dA = torch.from_numpy(A0).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(Bh.T)).to(dev).t()
    X = torch.zeros(r, N, dtype=torch.float64, device=dev).t()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    print(f"krylov prec={prec} kappa={kappa:g} shift={shift:g}", flush=True)
    its, be = C.dposv_krylov(dA, N, NB, dB, X, prec=prec, shift=shift, tol=1e-12, max_iter=300, info=info)
    assert int(info.item()) == 0 and be <= 1e-12 and its >= 1
    Xh = X.t().cpu().numpy().T
    assert oracle.solve_backward_error(N, NB, A0, Xh, Bh) <= 1e-12


def test_krylov_refuses_bad_shift(dev):
    import ttg_amd.chol as C

    A = torch.zeros(C.num_tiles(256, 128) * 128 * 128, dtype=torch.float64, device=dev)
    B = torch.zeros(1, 256, dtype=torch.float64, device=dev).t()
    with pytest.raises(C.CholError):
        C.dposv_krylov(A, 256, 128, B, B.clone(), shift=-1.0)
