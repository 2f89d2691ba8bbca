"""Seeded synthetic Gaussian-mixture workloads (numpy PCG64 streams).

Named workloads (points x dims):
  mix10_100k_d12    100,000 x 12: 10 isotropic Gaussians, centres U[-10,10]^12,
                    std 1, equal weights, plus 5% uniform noise in [-15,15]^12
  mix30_1m_d12      1,000,000 x 12: 30 Gaussians, centres U[-20,20]^12, std
                    U[0.5,2], Dirichlet(1) weights, plus 2% uniform noise over
                    the bounding box of the cluster points
  mix100_10m_d12    10,000,000 x 12: 100 Gaussians, centres U[-50,50]^12, std
                    U[0.5,3], Zipf(1.1) sizes, plus 1% noise
  aniso50_2m_d128   2,000,000 x 128: 50 anisotropic Gaussians (random
                    orthogonal rotation, axis std log-uniform [0.1,2]),
                    centres N(0,10^2)
  aniso200_5m_d32   5,000,000 x 32: 200 anisotropic Gaussians, centres
                    N(0,15^2), plus 2% uniform noise
All points are shuffled.  `n` may be overridden to get a smaller sample of
the same mixture.
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    "mix10_100k_d12": dict(n=100_000, d=12, kind="iso", k=10, centre=("uniform", 10.0), std=(1.0, 1.0),
                           weights="equal", noise=0.05, noise_box=15.0),
    "mix30_1m_d12": dict(n=1_000_000, d=12, kind="iso", k=30, centre=("uniform", 20.0), std=(0.5, 2.0),
                         weights="dirichlet", noise=0.02, noise_box=None),
    "mix100_10m_d12": dict(n=10_000_000, d=12, kind="iso", k=100, centre=("uniform", 50.0), std=(0.5, 3.0),
                           weights="zipf", noise=0.01, noise_box=None),
    "aniso50_2m_d128": dict(n=2_000_000, d=128, kind="aniso", k=50, centre=("normal", 10.0), std=(0.1, 2.0),
                            weights="equal", noise=0.0, noise_box=None),
    "aniso200_5m_d32": dict(n=5_000_000, d=32, kind="aniso", k=200, centre=("normal", 15.0), std=(0.1, 2.0),
                            weights="equal", noise=0.02, noise_box=None),
}


def _sizes(rng, total, k, weights):
    if weights == "equal":
        base = np.full(k, total // k, np.int64)
        base[: total - base.sum()] += 1
        return base
    if weights == "dirichlet":
        w = rng.dirichlet(np.ones(k))
    elif weights == "zipf":
        w = 1.0 / np.arange(1, k + 1, dtype=np.float64) ** 1.1
        w /= w.sum()
    else:
        raise ValueError(weights)
    return rng.multinomial(total, w)


def _rotation(rng, d):
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))[None, :]


def make(name: str, n: int | None = None, seed: int = 0) -> np.ndarray:
    cfg = WORKLOADS[name]
    n = int(n or cfg["n"])
    d = cfg["d"]
    rng = np.random.default_rng(seed)
    n_noise = int(round(cfg["noise"] * n))
    n_clu = n - n_noise
    k = cfg["k"]
    kind, scale = cfg["centre"]
    centres = rng.uniform(-scale, scale, (k, d)) if kind == "uniform" else rng.normal(0.0, scale, (k, d))
    sizes = _sizes(rng, n_clu, k, cfg["weights"])
    out = np.empty((n, d), np.float32)
    pos = 0
    lo, hi = cfg["std"]
    for c in range(k):
        m = int(sizes[c])
        if cfg["kind"] == "iso":
            std = rng.uniform(lo, hi)
            pts = rng.standard_normal((m, d)) * std + centres[c]
        else:
            axis = np.exp(rng.uniform(np.log(lo), np.log(hi), d))
            rot = _rotation(rng, d)
            pts = (rng.standard_normal((m, d)) * axis) @ rot.T + centres[c]
        out[pos:pos + m] = pts
        pos += m
    if n_noise:
        if cfg["noise_box"] is not None:
            b = cfg["noise_box"]
            lo_b, hi_b = np.full(d, -b), np.full(d, b)
        else:
            lo_b, hi_b = out[:pos].min(axis=0).astype(np.float64), out[:pos].max(axis=0).astype(np.float64)
        out[pos:] = rng.uniform(lo_b, hi_b, (n_noise, d))
    perm = rng.permutation(n)
    return out[perm]


def tiny_blobs(n: int, d: int, centres: int, seed: int = 0, spread: float = 8.0) -> np.ndarray:
    """Small seeded test mixture (isotropic, unit std)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (centres, d))
    lab = rng.integers(0, centres, n)
    return (c[lab] + rng.standard_normal((n, d))).astype(np.float32)
