"""Seeded synthetic workloads of the five benchmark configurations.

Every generator returns float32 row-major (data, queries).  Seeds: data
0xC0FFEE00 + cfg, queries 0xC0FFEF00 + cfg, cluster centres 0xC0FFF000 + cfg
(numpy PCG64).  The same buffers are fed to the GPU path and to the oracle.
"""
from __future__ import annotations

import numpy as np

DATA_SEED = 0xC0FFEE00
QUERY_SEED = 0xC0FFEF00
CENTRE_SEED = 0xC0FFF000


def _normalize(x: np.ndarray) -> np.ndarray:
    n = np.sqrt(np.einsum("ij,ij->i", x.astype(np.float64), x.astype(np.float64)))
    n[n == 0] = 1.0
    return (x / n[:, None]).astype(np.float32)


def gaussian_unit(n: int, d: int, seed: int) -> np.ndarray:
    """iid N(0,1) coordinates, rows scaled to unit L2 norm."""
    rng = np.random.default_rng(seed)
    return _normalize(rng.standard_normal((n, d), dtype=np.float32))


def clustered_unit(n: int, d: int, clusters: int, sigma: float, seed: int, centre_seed: int,
                   chunk: int = 1 << 20) -> np.ndarray:
    """Mixture of `clusters` Gaussian clusters: centre ~ N(0, I), point =
    centre + sigma * N(0, I), then scaled to unit L2 norm."""
    centres = np.random.default_rng(centre_seed).standard_normal((clusters, d), dtype=np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lab = rng.integers(0, clusters, size=e - s)
        x = centres[lab] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        out[s:e] = _normalize(x)
    return out


def sparse_exponential(n: int, d: int, zero_frac: float, seed: int) -> np.ndarray:
    """Nonnegative vectors: each coordinate 0 with probability zero_frac,
    else Exponential(1)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_exponential((n, d), dtype=np.float32)
    x[rng.random((n, d), dtype=np.float32) < zero_frac] = 0.0
    return x


def config(cfg: int, n: int | None = None, q: int | None = None):
    """(data, queries) of benchmark configuration cfg (sizes overridable)."""
    if cfg == 0:
        n = n or 100_000
        q = q or 1_000
        return gaussian_unit(n, 100, DATA_SEED + 0), gaussian_unit(q, 100, QUERY_SEED + 0)
    if cfg == 1:
        n = n or 1_200_000
        q = q or 10_000
        return (clustered_unit(n, 100, 1000, 0.3, DATA_SEED + 1, CENTRE_SEED + 1),
                clustered_unit(q, 100, 1000, 0.3, QUERY_SEED + 1, CENTRE_SEED + 1))
    if cfg == 2:
        n = n or 1_000_000
        q = q or 5_000
        return sparse_exponential(n, 128, 0.7, DATA_SEED + 2), sparse_exponential(q, 128, 0.7, QUERY_SEED + 2)
    if cfg == 3:
        n = n or 4_000_000
        q = q or 50_000
        return (clustered_unit(n, 128, 2000, 0.25, DATA_SEED + 3, CENTRE_SEED + 3),
                clustered_unit(q, 128, 2000, 0.25, QUERY_SEED + 3, CENTRE_SEED + 3))
    if cfg == 4:
        n = n or 10_000_000
        q = q or 2_048
        return (clustered_unit(n, 100, 5000, 0.3, DATA_SEED + 4, CENTRE_SEED + 4),
                clustered_unit(q, 100, 5000, 0.3, QUERY_SEED + 4, CENTRE_SEED + 4))
    raise ValueError(f"unknown configuration {cfg}")
