"""Python 3 counterpart of the reference's python/example.py.

    python -m rnn_amd.example [--double]

Builds a tiny table (N=10 points, d=3), asks for the k=3 nearest of the
first m=4 points for each of the first n=3 points, and prints the heap rows
after `sgsknn`/`dgsknn` (rnn_amd.gsknn, same argument order as the
reference's binding: d, k, Xq, Xq2, q, Xr, Xr2, r, D, I).  The reference
example also calls the `*_ref` oracle; here the rows are checked against a
brute-force numpy computation instead.  Needs the built libgsknn.so and a
GPU.
"""
from __future__ import annotations

import argparse

import numpy as np

from . import gsknn


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--double", action="store_true", help="dgsknn instead of sgsknn")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)

    N, m, n, d, k = 10, 4, 3, 3, 3
    dt = np.float64 if args.double else np.float32
    rng = np.random.default_rng(args.seed)
    X = rng.random((N, d)).astype(dt)
    X2 = np.zeros(N, dtype=dt)
    for i in range(N):  # squared norms in coordinate order
        for p in range(d):
            X2[i] += X[i, p] * X[i, p]
    q = np.arange(n, dtype=np.int32)   # query points
    r = np.arange(m, dtype=np.int32)   # reference (candidate) points
    D = np.full((n, k), 999999.9, dtype=dt)
    I = np.full((n, k), -1, dtype=np.int32)

    fn = gsknn.dgsknn if args.double else gsknn.sgsknn
    fn(d, k, X, X2, q, X, X2, r, D, I)
    print("gsknn:")
    print(D)
    print(I)

    # Brute force: each row is a max-heap of the k nearest by (distance, index).
    ok = True
    for j in range(n):
        dist = ((X[r].astype(np.float64) - X[q[j]].astype(np.float64)) ** 2).sum(1)
        want = sorted(zip(dist.tolist(), r.tolist()))[:k]
        got = sorted(zip(D[j].astype(np.float64).tolist(), I[j].tolist()))
        ok &= [w[1] for w in want] == [g[1] for g in got]
        ok &= np.allclose([w[0] for w in want], [g[0] for g in got], rtol=1e-5, atol=1e-6)
    print("matches brute force:", ok)
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())
