"""Synthetic input fields for the workloads (inputs, not part of the solver path)."""
from __future__ import annotations

import numpy as np


def gaussian_rff(pos: np.ndarray, ell: float = 0.1, modes: int = 1024, seed: int = 0) -> np.ndarray:
    """Unit-variance Gaussian random field with squared-exponential correlation of length
    ell, drawn as random Fourier features: g(x) = sqrt(2/M) sum_m cos(w_m . x + p_m) with
    w_m ~ N(0, ell^-2 I), p_m ~ U(0, 2 pi) from numpy's PCG64 generator seeded with `seed`
    (w drawn first, as an (M, 3) array, then p).  pos: (N, 3).  The sum is evaluated in
    FP64 on the host CPU threads (torch), in chunks of 65536 points."""
    import torch
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.normal(0.0, 1.0 / ell, size=(modes, 3)))
    ph = torch.from_numpy(rng.uniform(0.0, 2.0 * np.pi, size=modes))
    scale = np.sqrt(2.0 / modes)
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.empty(pos.shape[0])
    chunk = 1 << 16
    for s in range(0, pos.shape[0], chunk):
        x = torch.from_numpy(pos[s:s + chunk])
        out[s:s + chunk] = (scale * torch.cos(x @ w.T + ph).sum(dim=1)).numpy()
    return out


def lognormal_conductivity(pos: np.ndarray, ell: float = 0.1, modes: int = 1024, seed: int = 0) -> np.ndarray:
    """k = exp(g) with g = gaussian_rff(...): lognormal, mu = 0, sigma = 1."""
    return np.exp(gaussian_rff(pos, ell, modes, seed))
