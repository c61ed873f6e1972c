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
