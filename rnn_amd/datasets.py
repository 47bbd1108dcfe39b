"""Seeded synthetic inputs for the exact-kNN workloads.

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` in a fixed
order, in float64, and rounds to float32 last where the workload is FP32, so
the same seed gives the same bits on every machine running this numpy.

Naming follows the GSKNN C API (ChenhanYu/rnn include/gsknn.h:90-116):
``amap`` selects the m *candidate* rows of table XA, ``bmap`` selects the n
*query* rows of table XB, and heap row j belongs to query ``bmap[j]``
(frame/gsknn.c:1167-1182, test/test_dgsknn.c:133-143).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def sqnorm(X: np.ndarray) -> np.ndarray:
    """Row squared norms in the table's own precision, summed in p order."""
    acc = np.zeros(X.shape[0], dtype=X.dtype)
    for p in range(X.shape[1]):
        col = X[:, p]
        acc = acc + col * col
    return acc


def uniform_table(N: int, d: int, seed: int, dtype=np.float64) -> np.ndarray:
    return rng(seed).random((N, d)).astype(dtype)


def gaussian_mixture(N: int, d: int, clusters: int, seed: int, sigma=0.1,
                     sigma_range=None, dtype=np.float64) -> np.ndarray:
    """Centres U[-1,1]^d, equal weights, isotropic sigma per cluster.

    ``sigma_range=(lo, hi)`` draws each cluster's sigma log-uniform in [lo, hi].
    """
    g = rng(seed)
    centres = g.uniform(-1.0, 1.0, size=(clusters, d))
    if sigma_range is not None:
        lo, hi = sigma_range
        sig = np.exp(g.uniform(np.log(lo), np.log(hi), size=clusters))
    else:
        sig = np.full(clusters, float(sigma))
    comp = g.integers(0, clusters, size=N)
    X = centres[comp] + sig[comp, None] * g.standard_normal((N, d))
    return X.astype(dtype)


def low_rank(N: int, d: int, latent: int, seed: int, noise=0.01,
             dtype=np.float32, W=None) -> np.ndarray:
    """x = z W + noise * eps, z ~ N(0, I_latent), W ~ N(0,1)^{latent x d}."""
    g = rng(seed)
    if W is None:
        W = g.standard_normal((latent, d))
    out = np.empty((N, d), dtype=dtype)
    step = 1 << 16
    for s in range(0, N, step):
        e = min(N, s + step)
        z = g.standard_normal((e - s, latent))
        out[s:e] = (z @ W + noise * g.standard_normal((e - s, d))).astype(dtype)
    return out


@dataclass
class Problem:
    """One gsknn call: candidates (XA, amap), queries (XB, bmap), heap k."""
    XA: np.ndarray
    XA2: np.ndarray
    amap: np.ndarray
    XB: np.ndarray
    XB2: np.ndarray
    bmap: np.ndarray
    k: int
    extra_amaps: list = field(default_factory=list)  # further calls, same heaps
    prefill_amap: np.ndarray | None = None            # call made before the timed one

    @property
    def m(self):
        return int(self.amap.size)

    @property
    def n(self):
        return int(self.bmap.size)

    @property
    def d(self):
        return int(self.XA.shape[1])


def empty_heap(n: int, k: int, dtype=np.float64):
    D = np.full((n, k), np.inf, dtype=dtype)
    I = np.full((n, k), -1, dtype=np.int32)
    return D, I


def config1(scale: int = 1) -> Problem:
    """Plain dgsknn FP64: N=8192, d=16, U[0,1), seed 0; pi = permutation(8192).

    amap = pi[0:2048] (candidates), bmap = pi[6144:8192] (queries), k = 16; the
    second call merges the disjoint candidates pi[2048:4096] into the same
    heaps.  ``scale`` > 1 divides N, m, n for quick checks.
    """
    N, m = 8192 // scale, 2048 // scale
    X = uniform_table(N, 16, seed=0)
    pi = rng(0).permutation(N).astype(np.int32)
    X2 = sqnorm(X)
    return Problem(X, X2, pi[0:m].copy(), X, X2, pi[N - m:N].copy(), 16,
                   extra_amaps=[pi[m:2 * m].copy()])


def config2(scale: int = 1) -> Problem:
    """Drop-in dgsknn FP64 at scale: N=131072, d=64, 32-cluster mixture
    (sigma 0.1), seed 1; amap, bmap = 65536 uniform ints with repeats; the
    heaps are pre-filled by one call over 8192 distinct random candidates;
    k = 32."""
    N, m = 131072 // scale, 65536 // scale
    X = gaussian_mixture(N, 64, 32, seed=1, sigma=0.1)
    g = rng(101)
    amap = g.integers(0, N, size=m).astype(np.int32)
    bmap = g.integers(0, N, size=m).astype(np.int32)
    pre = g.choice(N, size=max(1, 8192 // scale), replace=False).astype(np.int32)
    X2 = sqnorm(X)
    return Problem(X, X2, amap, X, X2, bmap, 32, prefill_amap=pre)


def config3(scale: int = 1) -> Problem:
    """Precision form: FP32 points x = z W + 0.01 eps, latent 10, d = 256,
    N = 262144, seed 2; amap = pi[0:131072], bmap = pi[131072:262144]; k = 64."""
    N = 262144 // scale
    X = low_rank(N, 256, 10, seed=2)
    pi = rng(102).permutation(N).astype(np.int32)
    h = N // 2
    X2 = sqnorm(X)
    return Problem(X, X2, pi[:h].copy(), X, X2, pi[h:].copy(), 64)


def config4_points(N: int = 1 << 20, d: int = 32, clusters: int = 256, seed: int = 3):
    """All-NN sweep points: 256-cluster mixture, sigma log-uniform [0.02, 0.2]."""
    return gaussian_mixture(N, d, clusters, seed=seed, sigma_range=(0.02, 0.2),
                            dtype=np.float32)


def config5_reference(N: int = 2_000_000, d: int = 128, seed: int = 4):
    """Serving reference set: low-rank generator with latent dimension 16."""
    g = rng(seed)
    W = g.standard_normal((16, d))
    return low_rank(N, d, 16, seed=seed + 1000, W=W), W


def config5_queries(q: int, W: np.ndarray, seed: int) -> np.ndarray:
    return low_rank(q, W.shape[1], W.shape[0], seed=seed, W=W)
