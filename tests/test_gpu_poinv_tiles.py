"""The per-tile POINV operations of libttg_chol (the device replacements of
the LAPACK++/BLAS++ calls in examples/potrf/trtri_L.h and lauum.h) against
the oracle's tile operations: bitwise on small-integer inputs for the
multiplications (every product and sum is exact), and within 1e-12 relative
(FP64, a different summation order) for the triangular solves and the
inverse.  Each op also runs inside a launch group beside other ops of other
sizes, which must give bitwise the eager result."""
import numpy as np
import pytest
import torch

from gpu_util import dev_colmajor, host_colmajor

pytestmark = pytest.mark.gpu

SIZES = [(512, 512), (200, 64), (37, 5), (1, 1), (96, 384)]


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


def ints(rng, *shape):
    return F(rng.integers(-8, 9, size=shape))


def lower_factor(rng, n):
    """Well-conditioned lower tile with junk in the strict upper part."""
    L = np.tril(rng.uniform(-1, 1, (n, n))) / np.sqrt(n) + np.diag(rng.uniform(1, 2, n))
    return F(L + np.triu(rng.uniform(-9, 9, (n, n)), 1))


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("m,n", SIZES)
def test_gemm_nn_tn_exact(dev, oracle, m, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(m + n)
    k = max(1, (m + n) // 3)
    for ta in (0, 1):
        A = ints(rng, k, m) if ta else ints(rng, m, k)
        B, C0 = ints(rng, k, n), ints(rng, m, n)
        want = C0.copy(order="F")
        oracle.tile_gemm_xn(ta, A, B, want, alpha=-2.0, beta=3.0)
        dC = dev_colmajor(C0, dev)
        (C.dgemm_tn if ta else C.dgemm_nn)(dev_colmajor(A, dev), dev_colmajor(B, dev), dC, alpha=-2.0, beta=3.0)
        assert np.array_equal(host_colmajor(dC), want), ta


@pytest.mark.parametrize("n,k", [(512, 512), (130, 77), (32, 3), (1, 1)])
def test_syrk_lt_exact(dev, oracle, n, k):
    import ttg_amd.chol as C

    rng = np.random.default_rng(n * k)
    A, C0 = ints(rng, k, n), ints(rng, n, n)
    want = C0.copy(order="F")
    oracle.tile_syrk_lt(A, want, alpha=-1.0, beta=1.0)
    dC = dev_colmajor(C0, dev)
    C.dsyrk_lt(dev_colmajor(A, dev), dC, alpha=-1.0, beta=1.0)
    assert np.array_equal(host_colmajor(dC), want)  # the strict upper part included: untouched


@pytest.mark.parametrize("m,n", SIZES)
def test_trmm_llt_exact(dev, oracle, m, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(3 * m + n)
    L, B0 = ints(rng, m, m), ints(rng, m, n)
    want = B0.copy(order="F")
    oracle.tile_trmm_llt(L, want)
    dB = dev_colmajor(B0, dev)
    C.dtrmm_llt(dev_colmajor(L, dev), dB)
    assert np.array_equal(host_colmajor(dB), want)


@pytest.mark.parametrize("n", [512, 200, 33, 1])
def test_lauum_exact(dev, oracle, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(n)
    A0 = ints(rng, n, n)
    want = A0.copy(order="F")
    oracle.tile_lauum(want)
    dA = dev_colmajor(A0, dev)
    C.dlauum_tile(dA)
    assert np.array_equal(host_colmajor(dA), want)


@pytest.mark.parametrize("m,n", SIZES)
def test_triangular_solves(dev, oracle, m, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(7 * m + n)
    B0 = F(rng.uniform(-1, 1, (m, n)))
    L = lower_factor(rng, m)
    want = B0.copy(order="F")
    oracle.tile_trsm_lln(L, want)
    dB = dev_colmajor(B0, dev)
    C.dtrsm_lln(dev_colmajor(L, dev), dB, alpha=1.0)
    assert rel(host_colmajor(dB), want) <= 1e-12
    L = lower_factor(rng, n)
    want = B0.copy(order="F")
    oracle.tile_trsm_rln(L, want, alpha=-1.0)
    dB = dev_colmajor(B0, dev)
    C.dtrsm_rln(dev_colmajor(L, dev), dB, alpha=-1.0)
    assert rel(host_colmajor(dB), want) <= 1e-12


@pytest.mark.parametrize("n", [512, 300, 64, 1])
def test_trtri_tile(dev, oracle, n):
    import ttg_amd.chol as C

    rng = np.random.default_rng(n + 11)
    A0 = lower_factor(rng, n)
    want = A0.copy(order="F")
    assert oracle.tile_trtri(want) == 0
    dA = dev_colmajor(A0, dev)
    info = torch.full((1,), -7, dtype=torch.int32, device=dev)
    C.dtrtri_tile(dA, info)
    got = host_colmajor(dA)
    assert int(info.item()) == 0
    assert rel(np.tril(got), np.tril(want)) <= 1e-12
    assert np.array_equal(np.triu(got, 1), np.triu(A0, 1))  # strict upper untouched
    if n > 1:  # an exact zero on the diagonal: LAPACK's info, shifted by info_offset
        A0[n // 3, n // 3] = 0.0
        C.dtrtri_tile(dev_colmajor(A0, dev), info, info_offset=1000)
        assert int(info.item()) == 1000 + n // 3 + 1


def test_grouped_poinv_ops_match_eager(dev):
    """Ops of every kind and of several sizes in one launch group (one
    batched launch per pass) give bitwise the results of eager calls."""
    import ttg_amd.chol as C

    rng = np.random.default_rng(5)
    cases = []
    for (m, n) in [(512, 512), (200, 64), (37, 5)]:
        cases.append(("lln", lower_factor(rng, m), F(rng.uniform(-1, 1, (m, n)))))
        cases.append(("rln", lower_factor(rng, n), F(rng.uniform(-1, 1, (m, n)))))
        cases.append(("trmm", lower_factor(rng, m), F(rng.uniform(-1, 1, (m, n)))))
        cases.append(("trtri", lower_factor(rng, m), None))
        cases.append(("lauum", lower_factor(rng, n), None))
        cases.append(("nn", F(rng.uniform(-1, 1, (m, n))), F(rng.uniform(-1, 1, (n, m)))))
        cases.append(("tn", F(rng.uniform(-1, 1, (n, m))), F(rng.uniform(-1, 1, (n, m)))))
        cases.append(("syrk", F(rng.uniform(-1, 1, (m, n))), None))

    def run(grouped):
        outs = []
        calls = []
        for kind, a, b in cases:
            da = dev_colmajor(a, dev)
            if kind in ("lln", "rln", "trmm"):
                db = dev_colmajor(b, dev)
                f = {"lln": C.dtrsm_lln, "rln": C.dtrsm_rln, "trmm": C.dtrmm_llt}[kind]
                calls.append(lambda f=f, da=da, db=db: f(da, db, alpha=-0.5))
                outs.append(db)
            elif kind == "trtri":
                calls.append(lambda da=da: C.dtrtri_tile(da))
                outs.append(da)
            elif kind == "lauum":
                calls.append(lambda da=da: C.dlauum_tile(da))
                outs.append(da)
            elif kind == "syrk":
                dc = dev_colmajor(F(np.eye(a.shape[1])), dev)
                calls.append(lambda da=da, dc=dc: C.dsyrk_lt(da, dc, alpha=1.5, beta=0.5))
                outs.append(dc)
            else:
                db = dev_colmajor(b, dev)
                mm = a.shape[1] if kind == "tn" else a.shape[0]
                dc = dev_colmajor(F(np.ones((mm, b.shape[1]))), dev)
                f = C.dgemm_tn if kind == "tn" else C.dgemm_nn
                calls.append(lambda f=f, da=da, db=db, dc=dc: f(da, db, dc, alpha=0.75, beta=-1.0))
                outs.append(dc)
        if grouped:
            with C.group():
                for c in calls:
                    c()
        else:
            for c in calls:
                c()
        return [host_colmajor(o) for o in outs]

    eager, grouped = run(False), run(True)
    for (kind, _, _), e, g in zip(cases, eager, grouped):
        assert np.array_equal(e, g), kind
