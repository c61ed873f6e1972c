"""The schedules of the planned solve against each other: the diagonal POTRF
on its own CUs beside the bulk update (TTG_CHOL_POTRF_CUS), and the paired
split-FP16 schedule whose bulk updates
apply two panels in one pass over C (TTG_CHOL_PAIRED).  All of them only
reorder independent work or fuse two FP32 updates of a tile in their own
order, so the solution must be bitwise the one of the plain two-stream /
hipGraph schedule (a race between the streams would show as a difference).

Each schedule is chosen when its plan is built, and plans are cached per
shape and right-hand-side count, so the two runs use different (otherwise
unused) RHS counts; a right-hand side's solution does not depend on how many
others are solved with it.  A fixed number of refinement steps (tol 1e-300) keeps
the two runs on the same arithmetic."""
import numpy as np
import pytest
import torch

from gpu_util import to_dev

pytestmark = pytest.mark.gpu


def _solve(C, dA, N, NB, Bh, nrhs, prec, dev):
    dB = torch.from_numpy(np.ascontiguousarray(Bh[:, :nrhs].T)).to(dev).t()
    X = torch.zeros(nrhs, N, dtype=torch.float64, device=dev).t()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    its, _ = C.dposv_mixed(dA, N, NB, dB, X, prec=prec, tol=1e-300, max_iter=2, info=info, check=False)
    assert int(info.item()) == 0 and its == 2
    return X.t().cpu().numpy().T


@pytest.mark.parametrize("prec,r1,r2", [(1, 6, 7), (0, 9, 11), (2, 3, 5)])
def test_split_schedules_match_plain(dev, oracle, monkeypatch, prec, r1, r2):
    import ttg_amd.chol as C

    N, NB = 4096, 512
    sigma2 = sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2
    A0 = oracle.spd_tiles(N, NB, 4 * sigma2 / (1e4 - 1), 21)
    Bh = np.random.default_rng(5).uniform(-1, 1, (N, r2))
    dA = to_dev(A0, dev)
    monkeypatch.setenv("TTG_CHOL_POTRF_CUS", "8")
    Xs = _solve(C, dA, N, NB, Bh, r1, prec, dev)
    monkeypatch.setenv("TTG_CHOL_POTRF_CUS", "0")
    Xp = _solve(C, dA, N, NB, Bh, r2, prec, dev)
    assert np.array_equal(Xs, Xp[:, :r1])
    assert oracle.solve_backward_error(N, NB, A0, Xs, Bh[:, :r1]) <= 1e-12


@pytest.mark.parametrize("G,r1", [(2, 12), (3, 14), (4, 16)])
def test_grouped_schedule_matches_unpaired(dev, oracle, monkeypatch, G, r1):
    """Split-FP16 solve: the grouped schedule (G-panel bulk updates, eager
    updates inside a group) gives bitwise the solution of the one-step
    schedule.  Each right-hand side is solved independently, so the first
    r1 columns of an (r1 + 1)-column solve (a plan of its own) must match."""
    import ttg_amd.chol as C

    N, NB = 5120, 512  # NT = 10: several groups, and a remainder group for G = 3, 4
    sigma2 = sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2
    A0 = oracle.spd_tiles(N, NB, 4 * sigma2 / (1e4 - 1), 23)
    Bh = np.random.default_rng(6).uniform(-1, 1, (N, r1 + 1))
    dA = to_dev(A0, dev)
    monkeypatch.setenv("TTG_CHOL_PAIRED", "1")
    monkeypatch.setenv("TTG_CHOL_GROUP", str(G))
    Xp = _solve(C, dA, N, NB, Bh, r1, C.PREC_F16X2, dev)
    monkeypatch.setenv("TTG_CHOL_PAIRED", "0")
    Xu = _solve(C, dA, N, NB, Bh, r1 + 1, C.PREC_F16X2, dev)
    assert np.array_equal(Xp, Xu[:, :r1])
    assert oracle.solve_backward_error(N, NB, A0, Xp, Bh[:, :r1]) <= 1e-12


def test_knobs_set_get_and_refuse(dev):
    import ttg_amd.chol as C

    try:
        C.set_knob("group", 4)
        assert C.get_knob("group") == 4
        C.set_knob("group", None)
        assert C.get_knob("group") == 3
        for name, bad in (("group", 1), ("group", 9), ("tile_group", 0), ("paired", 2), ("potrf_cus", 65), ("sweep", 9)):
            with pytest.raises(C.CholError):
                C.set_knob(name, bad)
    finally:
        for k in C.KNOBS:
            C.set_knob(k, None)
        C.release_plans()


def test_tuned_schedule_is_bitwise_the_default(dev, oracle, tmp_path, monkeypatch):
    """The autotuner's search (every candidate runs the same arithmetic in
    another order of independent work) ends with a choice whose solution is
    bitwise the default schedule's; the trials record every candidate."""
    import ttg_amd.autotune as T
    import ttg_amd.chol as C

    monkeypatch.setenv("TTG_AMD_TUNING", str(tmp_path / "tuning.json"))
    N, NB, r = 5120, 512, 4
    sigma2 = sum(k * k for k in range(-127, 128)) / 255.0 / 127.0 ** 2
    A0 = oracle.spd_tiles(N, NB, 4 * sigma2 / (1e4 - 1), 29)
    Bh = np.random.default_rng(9).uniform(-1, 1, (N, r))
    dA = to_dev(A0, dev)
    try:
        T.apply_schedule({k: None for k in C.KNOBS})
        Xd = _solve(C, dA, N, NB, Bh, r, C.PREC_F16X2, dev)
        dB = torch.from_numpy(np.ascontiguousarray(Bh.T)).to(dev).t()
        X = torch.zeros(r, N, dtype=torch.float64, device=dev).t()

        def run():
            C.dposv_mixed(dA, N, NB, dB, X, prec=C.PREC_F16X2, tol=1e-300, max_iter=2, check=False)

        choice, trials = T.tune_schedule(run, N, NB, C.PREC_F16X2, r, reps=1, warmup=1, use_cache=False)
        print("choice", choice, "trials", trials, flush=True)
        assert set(choice) == set(C.KNOBS) and len(trials) >= 5
        assert C.knobs() == choice
        Xt = _solve(C, dA, N, NB, Bh, r, C.PREC_F16X2, dev)
        assert np.array_equal(Xt, Xd)
    finally:
        T.apply_schedule({k: None for k in C.KNOBS})
