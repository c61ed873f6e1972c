"""Helpers shared by the GPU tests: move oracle tile arrays to the device and
back, and compare factors."""
import numpy as np
import torch


def to_dev(a, dev="cuda"):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def dev_colmajor(a, dev="cuda"):
    """Column-major device tensor (a transposed view) holding the numpy array a."""
    a = np.asarray(a, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).t()


def host_colmajor(t):
    torch.cuda.synchronize()
    return t.t().cpu().numpy().T


def rel_err(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = max(np.abs(np.asarray(b)).max(), 1e-300)
    return d / s


def lower_tiles_rel_err(O, N, NB, got, want):
    """max |got - want| / max |want| over the defined lower entries of the tiles."""
    nt = O.nt_of(N, NB)
    num, den = 0.0, 0.0
    for n in range(nt):
        for m in range(n, nt):
            mb, nb = min(NB, N - m * NB), min(NB, N - n * NB)
            g = O.tile_view(got, N, NB, m, n)[:mb, :nb]
            w = O.tile_view(want, N, NB, m, n)[:mb, :nb]
            if m == n:
                g, w = np.tril(g), np.tril(w)
            num = max(num, np.abs(g - w).max())
            den = max(den, np.abs(w).max())
    return num / den
