"""Pins of the L1 (KNN_1NORM) oracle -- CPU only.

The oracle (oracle/gsknn_oracle.c, oracle_dgsknn_l1) restates the reference's
L1 kernels, which its FP64 kernel tables register at index KNN_1NORM
(kernels/x86_64/sandybridge/gsknn_kernel.h:125-163) and its driver never
selects (frame/gsknn.c:265, :409): |a_p - b_p| summed in p order in blocks of
256 dimensions, block sums added in order (frame/gsknn.c:996-1086,
rnn_rank_k_abs_int_d8x4.c, rnn_r_1norm_int_d8x4_row.c).  It is checked
against
  * fixtures made from those kernels (tests/golden/make_golden.py l1): raw
    heap arrays bit for bit on tie-free inputs, distance multisets on an
    integer lattice whose ties the reference breaks by its 4-wide prefilter;
  * the reference's optimized library compiled live from its sources, when
    present, at several d on both sides of the 256-dimension block;
  * brute force in long double, exact integer lattices (canonical ties), the
    d = 1 special case, and a planted mistake (an unblocked sum) that the
    fixture must catch.
"""
import numpy as np
import pytest

from oracle import oracle


def _compat(c):
    n, k = c["D"].shape
    D = np.full((n, k), 1.79e308)
    I = np.full((n, k), -1, dtype=np.int32)
    oracle.gsknn_l1(c["X"], c["amap"], c["X"], c["bmap"], D, I)
    return D, I


@pytest.mark.parametrize("name", ["l1_d64", "l1_d600", "l1_d1", "l1_k40"])
def test_l1_oracle_matches_reference_kernels_bitwise(l1_fixtures, name):
    c = l1_fixtures[name]
    D, I = _compat(c)
    np.testing.assert_array_equal(I, c["I"])
    assert np.array_equal(D.view(np.uint8), c["D"].view(np.uint8)), "distance bits differ"


def test_l1_lattice_distances_match_reference(l1_fixtures):
    c = l1_fixtures["l1_lattice"]
    D, I = _compat(c)
    np.testing.assert_array_equal(np.sort(D, axis=1), np.sort(c["D"], axis=1))
    # every returned index carries its own exact L1 distance
    X = c["X"]
    for j, b in enumerate(c["bmap"]):
        for Dr, Ir in ((D, I), (c["D"], c["I"])):
            np.testing.assert_array_equal(np.abs(X[Ir[j]] - X[b]).sum(axis=1), Dr[j])


@pytest.mark.skipif(oracle.ref_opt_lib() is None, reason="reference library not built here")
@pytest.mark.parametrize("d", [1, 3, 17, 64, 255, 256, 257, 513])
def test_l1_oracle_matches_live_reference(d):
    g = np.random.Generator(np.random.PCG64(1000 + d))
    X = g.standard_normal((400, d))
    amap = g.permutation(400)[:300].astype(np.int32)
    bmap = g.permutation(400)[:37].astype(np.int32)
    D1 = np.full((37, 12), 1.79e308)
    I1 = np.full((37, 12), -1, dtype=np.int32)
    D2, I2 = D1.copy(), I1.copy()
    oracle.gsknn_l1(X, amap, X, bmap, D1, I1)
    oracle.ref_opt_l1(X, amap, X, bmap, D2, I2)
    np.testing.assert_array_equal(I1, I2)
    assert np.array_equal(D1.view(np.uint8), D2.view(np.uint8))


def _canon(X, amap, bmap, k, dtype=np.float64):
    X = np.ascontiguousarray(X, dtype=dtype)
    D = np.full((bmap.size, k), np.inf, dtype=dtype)
    I = np.full((bmap.size, k), -1, dtype=np.int32)
    oracle.gsknn_l1(X, amap, X, bmap, D, I, canonical=True)
    return oracle.sort_rows(D, I)


def test_l1_canonical_matches_brute_force():
    g = np.random.Generator(np.random.PCG64(11))
    X = g.standard_normal((300, 9))
    amap = g.permutation(300)[:200].astype(np.int32)
    bmap = g.permutation(300)[:50].astype(np.int32)
    Ds, Is = _canon(X, amap, bmap, 10)
    Db, Ib = oracle.brute_l1(X, amap, X, bmap, 10)
    np.testing.assert_array_equal(Is, Ib)
    np.testing.assert_allclose(Ds, Db, rtol=1e-14, atol=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_l1_lattice_ties_go_to_lower_index(dtype):
    pts = np.array([(x, y, z) for x in range(6) for y in range(6) for z in range(6)])
    g = np.random.Generator(np.random.PCG64(3))
    amap = g.permutation(216).astype(np.int32)
    bmap = np.arange(0, 216, 7, dtype=np.int32)
    Ds, Is = _canon(pts, amap, bmap, 12, dtype)
    for j, b in enumerate(bmap):
        key = sorted((int(np.abs(pts[a] - pts[b]).sum()), int(a)) for a in amap)[:12]
        assert [i for _, i in key] == list(Is[j])
        assert [dd for dd, _ in key] == [int(v) for v in Ds[j]]


def test_l1_d1_reduces_to_sorting_abs_difference():
    g = np.random.Generator(np.random.PCG64(6))
    x = np.round(g.random(150) * 1000) / 8.0
    X = x[:, None]
    amap = np.arange(0, 150, 2, dtype=np.int32)
    bmap = np.arange(1, 150, 2, dtype=np.int32)
    Ds, Is = _canon(X, amap, bmap, 7)
    for j, b in enumerate(bmap):
        key = sorted((abs(x[a] - x[b]), a) for a in amap)[:7]
        assert [i for _, i in key] == list(Is[j])
        assert [dd for dd, _ in key] == list(Ds[j])


def test_l1_f32_matches_brute_force_sets():
    g = np.random.Generator(np.random.PCG64(12))
    X = g.standard_normal((400, 24)).astype(np.float32)
    amap = g.permutation(400)[:300].astype(np.int32)
    bmap = g.permutation(400)[:40].astype(np.int32)
    Ds, Is = _canon(X, amap, bmap, 8, np.float32)
    Db, Ib = oracle.brute_l1(X.astype(np.float64), amap, X.astype(np.float64), bmap, 8)
    np.testing.assert_array_equal(Is, Ib)
    np.testing.assert_allclose(Ds, Db, rtol=2e-6, atol=0)


def test_l1_unblocked_sum_is_caught_by_fixture(l1_fixtures):
    """A restatement that sums all 600 dimensions in one block (the planted
    mistake) disagrees with the reference's bits; the blocked one agrees."""
    c = l1_fixtures["l1_d600"]
    X, I, D = c["X"], c["I"], c["D"]

    def dist(q, a, block):
        s = 0.0
        for p0 in range(0, q.size, block):
            t = 0.0
            for p in range(p0, min(q.size, p0 + block)):
                t += abs(q[p] - a[p])
            s = t if p0 == 0 else s + t
        return s

    good = bad = 0
    for j, b in enumerate(c["bmap"][:6]):
        for t in range(I.shape[1]):
            good += dist(X[b], X[I[j, t]], 256) == D[j, t]
            bad += dist(X[b], X[I[j, t]], 600) == D[j, t]
    assert good == 6 * I.shape[1]
    assert bad < good
