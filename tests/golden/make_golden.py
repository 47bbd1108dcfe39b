"""Write the golden fixtures in this directory from the reference's own oracle.

    make -C oracle ref && python tests/golden/make_golden.py [l2|l1]

The reference ships no golden vectors (its drivers print instead of assert,
test/test_dgsknn.c:13-104), so the fixtures are made here: seeded inputs, run
through ``dgsknn_ref``/``sgsknn_ref`` compiled from the reference's sources
(oracle/_ref/libgsknn_ref.so, built by oracle/Makefile), outputs stored
beside the inputs.  Only data is written; nothing of the reference is copied.

The stored D/I are the raw heap arrays after each call, so the fixtures pin
the heap layout (insertion order under the reference's replace-on-<= rule) as
well as the neighbour sets.

``l1`` writes gsknn_ref_l1_fixtures.npz from the reference's optimized
library (oracle/_ref/libgsknn_opt.so) run through its KNN_1NORM kernel-table
entries (oracle.ref_opt_l1): the L1 norm has no oracle form in the
reference, so its own kernels are the source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402


def norms(X):
    acc = np.zeros(X.shape[0], dtype=X.dtype)
    for p in range(X.shape[1]):
        acc = acc + X[:, p] * X[:, p]
    return acc


def run_ref(X, amap, bmap, D, I, calls=1, extra=()):
    X2 = norms(X)
    outs = []
    for am in (amap,) + tuple(np.ascontiguousarray(e, dtype=np.int32) for e in extra):
        oracle.ref_gsknn(X, X2, am, X, X2, bmap, D, I)
        outs.append((D.copy(), I.copy()))
    return X2, outs


def case(name, X, amap, bmap, k, ro, extra=(), D0=None, I0=None):
    X = np.ascontiguousarray(X)
    n = bmap.size
    if D0 is None:
        D = np.full((n, k), ro, dtype=X.dtype)
        I = np.full((n, k), -1, dtype=np.int32)
    else:
        D, I = D0.copy(), I0.copy()
    Din, Iin = D.copy(), I.copy()
    X2, outs = run_ref(X, amap.astype(np.int32), bmap.astype(np.int32), D, I, extra=extra)
    rec = {f"{name}__X": X, f"{name}__X2": X2, f"{name}__amap": amap.astype(np.int32),
           f"{name}__bmap": bmap.astype(np.int32), f"{name}__D0": Din, f"{name}__I0": Iin}
    for c, am in enumerate(extra):
        rec[f"{name}__amap{c + 2}"] = am.astype(np.int32)
    for c, (Dc, Ic) in enumerate(outs):
        rec[f"{name}__D{c + 1}"] = Dc
        rec[f"{name}__I{c + 1}"] = Ic
    return rec


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libgsknn_ref.so missing: run `make -C oracle ref`")
    g = np.random.Generator(np.random.PCG64(2024))
    recs = {}

    # Reduced config-1 shape: N = 1024 U[0,1) points, d = 16, 256 candidates,
    # 256 queries, k = 16, then a second call with disjoint candidates.
    X = np.random.Generator(np.random.PCG64(0)).random((1024, 16))
    pi = np.random.Generator(np.random.PCG64(0)).permutation(1024)
    recs.update(case("c1small", X, pi[:256], pi[768:], 16, 1.79e308, extra=(pi[256:512],)))

    # k > m: rows stay partially filled with (ro, -1).
    X = g.random((20, 3))
    recs.update(case("partial", X, np.arange(5), np.arange(7, 14), 8, 1.79e308))

    # Repeated candidate indices (duplicates are kept).
    X = g.random((50, 4))
    recs.update(case("repeats", X, g.integers(0, 50, 40), g.integers(0, 50, 10), 6, 1.79e308))

    # Integer lattice: exact distances with many exact ties.
    pts = np.array([(x, y) for x in range(8) for y in range(8)], dtype=np.float64)
    recs.update(case("lattice", pts, g.permutation(64), np.arange(0, 64, 4), 10, 1.79e308))

    # d = 1 and an odd shape.
    X = g.random((64, 1))
    recs.update(case("d1", X, np.arange(0, 64, 2), np.arange(1, 64, 2), 5, 1.79e308))
    X = g.random((120, 5))
    recs.update(case("odd", X, g.permutation(120)[:37], g.permutation(120)[:23], 7, 1.79e308))

    # Tight initial radius: candidates beyond ro never enter.
    X = g.random((200, 6))
    recs.update(case("radius", X, g.permutation(200)[:150], g.permutation(200)[:40], 9, 0.4))

    # FP32 (sgsknn_ref).
    X = g.random((512, 8)).astype(np.float32)
    recs.update(case("f32", X, g.permutation(512)[:200], g.permutation(512)[:100], 12,
                     np.float32(1.79e30)))

    out = os.path.join(HERE, "gsknn_ref_fixtures.npz")
    np.savez_compressed(out, **recs)
    print(out, os.path.getsize(out), "bytes,", len(recs), "arrays")


def l1_case(name, X, amap, bmap, k):
    X = np.ascontiguousarray(X, dtype=np.float64)
    amap = np.ascontiguousarray(amap, dtype=np.int32)
    bmap = np.ascontiguousarray(bmap, dtype=np.int32)
    D = np.full((bmap.size, k), 1.79e308)
    I = np.full((bmap.size, k), -1, dtype=np.int32)
    oracle.ref_opt_l1(X, amap, X, bmap, D, I)
    return {f"{name}__X": X, f"{name}__amap": amap, f"{name}__bmap": bmap,
            f"{name}__D": D, f"{name}__I": I}


def main_l1():
    if oracle.ref_opt_lib() is None:
        raise SystemExit("oracle/_ref/libgsknn_opt.so missing: run `make -C oracle ref`")
    g = np.random.Generator(np.random.PCG64(77))
    recs = {}
    # d = 64 (one dimension block) and d = 600 (three blocks of 256, 256, 88).
    X = g.standard_normal((160, 64))
    recs.update(l1_case("l1_d64", X, g.permutation(160)[:140], g.permutation(160)[:40], 16))
    X = g.standard_normal((48, 600))
    recs.update(l1_case("l1_d600", X, g.permutation(48)[:36], g.permutation(48)[:12], 8))
    # d = 1, and more neighbours than a 4-wide candidate group.
    X = g.standard_normal((94, 1))
    recs.update(l1_case("l1_d1", X, g.permutation(94)[:64], g.permutation(94)[:30], 5))
    X = g.standard_normal((600, 5))
    recs.update(l1_case("l1_k40", X, g.permutation(600)[:560], g.permutation(600)[:50], 40))
    # Integer lattice: exact L1 distances with many exact ties.
    pts = np.array([(x, y, z) for x in range(5) for y in range(5) for z in range(5)],
                   dtype=np.float64)
    recs.update(l1_case("l1_lattice", pts, g.permutation(125), np.arange(0, 125, 5), 10))
    out = os.path.join(HERE, "gsknn_ref_l1_fixtures.npz")
    np.savez_compressed(out, **recs)
    print(out, os.path.getsize(out), "bytes,", len(recs), "arrays")


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("all", "l2"):
        main()
    if which in ("all", "l1"):
        main_l1()
