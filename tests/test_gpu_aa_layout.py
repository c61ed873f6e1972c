"""Lane layout of the MFMA builtins the kernels are written against
(collected first: every other GPU test depends on it)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def drow(kind, g, r):
    # D row of element r held by lane group g, as the kernels assume
    return g + 4 * r if kind == 0 else 4 * g + r


@pytest.mark.parametrize("kind", [0, 1])
def test_mfma_16x16x4_layout(dev, kind):
    import ttg_amd.chol as C

    out = torch.full((64 * 64 * 4,), -1.0, dtype=torch.float64, device=dev)
    C.selftest_mfma_layout(kind, out)
    got = out.cpu().numpy().reshape(64, 64, 4)
    want = np.zeros((64, 64, 4))
    for m in range(64):  # B one-hot at (k = m // 16, j = m % 16); A lane l = (i = l % 16, k = l // 16)
        for l in range(64):
            if l % 16 == m % 16:
                for r in range(4):
                    i = drow(kind, l // 16, r)
                    want[m, l, r] = (m // 16) * 16 + i + 1
    assert np.array_equal(got, want)


def test_mfma_f16_16x16x32_layout(dev):
    # A lane l = (i = l % 16, k = 8 (l // 16) + e); B lane l = (k = 8 (l // 16) + e, j = l % 16);
    # D lane l element r = (i = 4 (l // 16) + r, j = l % 16)
    import ttg_amd.chol as C

    out = torch.full((512 * 64 * 4,), -1.0, dtype=torch.float64, device=dev)
    C.selftest_mfma_layout(2, out)
    got = out.cpu().numpy().reshape(512, 64, 4)
    want = np.zeros((512, 64, 4))
    for q in range(512):
        L, e = q // 8, q % 8
        k0, j0 = 8 * (L // 16) + e, L % 16
        for l in range(64):
            if l % 16 == j0:
                for r in range(4):
                    i = 4 * (l // 16) + r
                    lp = i + 16 * (k0 // 8)
                    want[q, l, r] = lp * 8 + k0 % 8 + 1
    assert np.array_equal(got, want)


def test_mfma_f16_32x32x16_layout(dev):
    # A lane l = (i = l % 32, k = 8 (l // 32) + e); B lane l = (k = 8 (l // 32) + e, j = l % 32);
    # D lane l element r = (i = 8 (r // 4) + 4 (l // 32) + r % 4, j = l % 32)
    import ttg_amd.chol as C

    out = torch.full((512 * 64 * 16,), -1.0, dtype=torch.float64, device=dev)
    C.selftest_mfma_layout(3, out)
    got = out.cpu().numpy().reshape(512, 64, 16)
    want = np.zeros((512, 64, 16))
    for q in range(512):
        L, e = q // 8, q % 8
        k0, j0 = 8 * (L // 32) + e, L % 32
        for l in range(64):
            if l % 32 == j0:
                for r in range(16):
                    i = 8 * (r // 4) + 4 * (l // 32) + r % 4
                    lp = i + 32 * (k0 // 8)
                    want[q, l, r] = lp * 8 + k0 % 8 + 1
    assert np.array_equal(got, want)
