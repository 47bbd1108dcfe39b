"""Pins of the CPU oracle (oracle/gsknn_oracle.c) -- CPU only.

The oracle restates frame/gsknn_ref.c:63-162 + frame/dgsknn_ref_impl.h:41-103
+ frame/gsknn_heap.c:52-162.  It is checked against
  * the golden fixtures made from the reference's own oracle sources
    (tests/golden/make_golden.py), bit for bit, heap layout included;
  * the reference oracle compiled live from its sources, when present;
  * brute force in long double, exact lattice ties, chunking and permutation
    invariants, the d = 1 special case, and planted mistakes that the pins
    must catch.
"""
import numpy as np
import pytest

from oracle import oracle


def _run_compat(c, calls=None):
    X, X2 = c["X"], c["X2"]
    D, I = c["D0"].copy(), c["I0"].copy()
    outs = []
    amaps = [c["amap"]] + [c[f"amap{i}"] for i in range(2, 5) if f"amap{i}" in c]
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, c["bmap"], D, I, canonical=False)
        outs.append((D.copy(), I.copy()))
    return outs


@pytest.mark.parametrize("name", ["c1small", "partial", "repeats", "lattice", "d1", "odd",
                                  "radius", "f32"])
def test_oracle_matches_reference_fixture_bitwise(fixtures, name):
    c = fixtures[name]
    outs = _run_compat(c)
    for i, (D, I) in enumerate(outs, start=1):
        np.testing.assert_array_equal(I, c[f"I{i}"])
        assert np.array_equal(D.view(np.uint8), c[f"D{i}"].view(np.uint8)), "distance bits differ"


def test_fixture_norms_match_oracle_norms(fixtures):
    for name, c in fixtures.items():
        np.testing.assert_array_equal(oracle.sqnorm(c["X"]), c["X2"])


@pytest.mark.skipif(oracle.ref_lib() is None, reason="reference oracle not built here")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(300, 7, 11), (513, 33, 17), (129, 64, 32)])
def test_oracle_matches_live_reference(dtype, shape):
    m, d, k = shape
    g = np.random.Generator(np.random.PCG64(m * d + k))
    X = g.standard_normal((700, d)).astype(dtype)
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, 700, m).astype(np.int32)
    bmap = g.integers(0, 700, 97).astype(np.int32)
    ro = 1.79e308 if dtype == np.float64 else 1.79e30
    D1 = np.full((97, k), ro, dtype=dtype)
    I1 = np.full((97, k), -1, dtype=np.int32)
    D2, I2 = D1.copy(), I1.copy()
    oracle.gsknn(X, X2, amap, X, X2, bmap, D1, I1)
    oracle.ref_gsknn(X, X2, amap, X, X2, bmap, D2, I2)
    np.testing.assert_array_equal(I1, I2)
    assert np.array_equal(D1.view(np.uint8), D2.view(np.uint8))


def _canon(X, amap, bmap, k, dtype=np.float64):
    X = X.astype(dtype)
    X2 = oracle.sqnorm(X)
    D = np.full((bmap.size, k), np.inf, dtype=dtype)
    I = np.full((bmap.size, k), -1, dtype=np.int32)
    oracle.gsknn(X, X2, amap, X, X2, bmap, D, I, canonical=True)
    return oracle.sort_rows(D, I)


def test_canonical_matches_brute_force():
    g = np.random.Generator(np.random.PCG64(7))
    X = g.random((400, 9))
    amap = g.permutation(400)[:300].astype(np.int32)
    bmap = g.permutation(400)[:50].astype(np.int32)
    Ds, Is = _canon(X, amap, bmap, 12)
    Db, Ib = oracle.brute(X, amap, X, bmap, 12)
    np.testing.assert_array_equal(Is, Ib)
    scale = np.add.outer(oracle.sqnorm(X)[bmap], np.zeros(12)) + 9.0
    assert np.all(np.abs(Ds - Db) <= 1e-12 * scale)


def test_lattice_ties_go_to_lower_index():
    # Integer points: every distance is exact, so ties are real ties.
    pts = np.array([(x, y, z) for x in range(5) for y in range(5) for z in range(4)],
                   dtype=np.float64)
    g = np.random.Generator(np.random.PCG64(3))
    amap = g.permutation(pts.shape[0]).astype(np.int32)
    bmap = np.arange(0, pts.shape[0], 3).astype(np.int32)
    Ds, Is = _canon(pts, amap, bmap, 11)
    Db, Ib = oracle.brute(pts, amap, pts, bmap, 11)
    np.testing.assert_array_equal(Is, Ib)
    np.testing.assert_array_equal(Ds, Db)


def test_canonical_is_chunking_and_permutation_invariant():
    g = np.random.Generator(np.random.PCG64(11))
    X = np.floor(g.random((300, 4)) * 6)  # many exact ties
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, 300, 280).astype(np.int32)  # repeats too
    bmap = g.integers(0, 300, 40).astype(np.int32)
    k = 9
    ref = _canon(X, amap, bmap, k)
    D = np.full((40, k), np.inf)
    I = np.full((40, k), -1, dtype=np.int32)
    for lo, hi in [(0, 13), (13, 200), (200, 280)]:
        oracle.gsknn(X, X2, amap[lo:hi].copy(), X, X2, bmap, D, I, canonical=True)
    got = oracle.sort_rows(D, I)
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[0], ref[0])
    perm = _canon(X, g.permutation(amap).astype(np.int32), bmap, k)
    np.testing.assert_array_equal(perm[1], ref[1])


def test_d1_reduces_to_sorting_abs_difference():
    g = np.random.Generator(np.random.PCG64(5))
    x = np.round(g.random(200) * 1000) / 8.0  # exact binary fractions
    X = x[:, None]
    amap = np.arange(0, 200, 2, dtype=np.int32)
    bmap = np.arange(1, 200, 2, dtype=np.int32)
    Ds, Is = _canon(X, amap, bmap, 6)
    for j, b in enumerate(bmap):
        key = sorted(((x[a] - x[b]) ** 2, a) for a in amap)[:6]
        assert [i for _, i in key] == list(Is[j])
        assert [dd for dd, _ in key] == list(Ds[j])


def test_rows_are_max_heaps_and_partial_fill(fixtures):
    c = fixtures["partial"]
    D, I = c["D1"], c["I1"]
    assert oracle.rows_are_heaps(D, I)
    # 5 candidates, k = 8: three (ro, -1) entries remain in every row.
    assert np.all((I == -1).sum(axis=1) == 3)
    assert np.all(D[I == -1] == 1.79e308)


def test_self_query_is_nearest():
    g = np.random.Generator(np.random.PCG64(9))
    X = g.random((100, 5))
    idx = np.arange(100, dtype=np.int32)
    Ds, Is = _canon(X, idx, idx, 3)
    np.testing.assert_array_equal(Is[:, 0], idx)


def _py_variant(c, variant):
    """Python restatements with one plausible mistake planted."""
    X, X2, amap, bmap = c["X"], c["X2"], c["amap"], c["bmap"]
    k = c["D1"].shape[1]
    A, B = X[amap], X[bmap]
    dot = B @ A.T
    if variant == "sign":
        dist = 2.0 * dot + X2[amap][None, :] + X2[bmap][:, None]
    elif variant == "transpose":
        dist = -2.0 * dot + X2[amap][None, :] + X2[bmap][:, None]
        dist = dist.T[: bmap.size, : amap.size] if dist.shape[0] == dist.shape[1] else dist[:, ::-1]
    elif variant == "dropped_term":
        dist = -2.0 * dot + X2[bmap][:, None]
    else:
        dist = -2.0 * dot + X2[amap][None, :] + X2[bmap][:, None]
    order = np.lexsort((np.broadcast_to(amap, dist.shape), dist), axis=1)[:, :k]
    return np.take_along_axis(np.broadcast_to(amap, dist.shape), order, axis=1)


@pytest.mark.parametrize("variant", ["sign", "transpose", "dropped_term"])
def test_planted_mistakes_are_caught_by_fixture(fixtures, variant):
    c = fixtures["c1small"]
    _, Iref = oracle.sort_rows(c["D1"], c["I1"])
    ok = _py_variant(c, "none")
    np.testing.assert_array_equal(ok, Iref)  # the faithful restatement agrees
    bad = _py_variant(c, variant)
    assert not np.array_equal(bad, Iref)


@pytest.mark.skipif(oracle.ref_opt_lib() is None, reason="reference library not built here")
@pytest.mark.parametrize("d", [16, 64, 300])
def test_oracle_matches_reference_optimized_path(d):
    """The reference's production dgsknn_var1 (AVX micro-kernels, blocked in
    256-dimension slices above d = 256; oracle/_ref/libgsknn_opt.so) returns
    the oracle's neighbour sets on tie-free data, distances within its
    different summation order's rounding, with 1 and with 4 threads."""
    g = np.random.Generator(np.random.PCG64(70 + d))
    X = g.standard_normal((3000, d))
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, 3000, 2500).astype(np.int32)
    bmap = g.integers(0, 3000, 300).astype(np.int32)
    k = 24
    Do = np.full((300, k), np.inf)
    Io = np.full((300, k), -1, dtype=np.int32)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Do, Io)
    Dos, Ios = oracle.sort_rows(Do, Io)
    for threads in (1, 4):
        Dr = np.full((300, k), 1.79e308)
        Ir = np.full((300, k), -1, dtype=np.int32)
        oracle.ref_opt_gsknn(X, X2, amap, X, X2, bmap, Dr, Ir, threads=threads)
        Drs, Irs = oracle.sort_rows(Dr, Ir)
        np.testing.assert_array_equal(Irs, Ios)
        scale = X2[bmap][:, None] + X2.max()
        assert np.all(np.abs(Drs - Dos) <= 1e-12 * scale)


def test_direct_knn_f32in_matches_long_double_brute_force():
    """The FP64 direct-difference checker of the BF16 form equals long-double
    brute force on FP32 points: same neighbour sets in (distance, index)
    order, distances within one rounding of the FP64 sum (the long-double
    sum rounded to FP64 is the reference); lattice points give exact ties
    that the lower index must win."""
    g = np.random.Generator(np.random.PCG64(91))
    for d, lattice in ((1, False), (7, False), (33, False), (5, True)):
        if lattice:
            X = g.integers(-3, 4, (300, d)).astype(np.float32)
        else:
            X = g.standard_normal((300, d)).astype(np.float32)
        amap = g.permutation(300)[:250].astype(np.int32)
        bmap = g.integers(0, 300, 40).astype(np.int32)
        k = 12
        D = np.full((40, k), np.inf)
        I = np.full((40, k), -1, dtype=np.int32)
        oracle.direct_knn_f32in(X, amap, X, bmap, D, I)
        Ds, Is = oracle.sort_rows(D, I)
        Db, Ib = oracle.brute(X.astype(np.float64), amap, X.astype(np.float64), bmap, k)
        np.testing.assert_array_equal(Is, Ib)
        np.testing.assert_allclose(Ds, Db, rtol=4 * d * 2.0 ** -53, atol=0)
        if lattice:
            assert np.array_equal(Ds, Db)  # integer coordinates: every sum is exact


def test_rows_are_heaps_rejects_a_planted_violation():
    """rows_are_heaps is a real check: a valid max-heap row passes, and the
    same row with a child swapped above its parent, or (canonically) an
    equal-distance child of larger index above its parent, fails."""
    D = np.array([[9.0, 7.0, 8.0, 1.0, 2.0, 3.0]])
    I = np.array([[5, 4, 3, 2, 1, 0]], dtype=np.int32)
    assert oracle.rows_are_heaps(D, I)
    Dv = D.copy()
    Dv[0, [0, 1]] = Dv[0, [1, 0]]  # root 7 below its child 9
    Iv = I.copy()
    Iv[0, [0, 1]] = Iv[0, [1, 0]]
    assert not oracle.rows_are_heaps(Dv, Iv)
    Dv = D.copy()
    Dv[0, 4] = 7.5  # a grandchild above its parent 7
    assert not oracle.rows_are_heaps(Dv, I)
    # Canonical order: parent (7, 4) with an equal-distance child (7, 9).
    Dt = np.array([[9.0, 7.0, 8.0, 7.0]])
    It = np.array([[0, 4, 3, 9]], dtype=np.int32)
    assert oracle.rows_are_heaps(Dt, It)  # distance order only
    assert not oracle.rows_are_heaps(Dt, It, canonical=True)
    It[0, 3] = 1  # equal distance, smaller index below: a canonical heap
    assert oracle.rows_are_heaps(Dt, It, canonical=True)
