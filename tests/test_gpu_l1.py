"""GPU parity of the L1 (KNN_1NORM) path of libgsknn.so, through the C-ABI.

Bar: neighbour indices and distances bit-identical to the L1 oracle
(oracle_dgsknn_l1 / oracle_sgsknn_l1), which is itself bit-identical to the
reference's own L1 kernels run through its kernel tables (tests/
test_oracle_l1.py).  The fixtures made from those kernels are also checked
against the GPU directly.  Inputs without exact ties are compared after
sorting rows by (distance, index); integer lattices (exact ties) under the
canonical lower-index rule.  Cases cover the register paths (d = 8, 16, 32,
64), the LDS path (other d, partial chunks, d > 256 with several summation
blocks), candidate splits, partial groups, pre-filled heaps, heap row maps
and the host-pointer entry points.
"""
import numpy as np
import pytest

from oracle import oracle
from rnn_amd import gsknn

pytestmark = pytest.mark.gpu


def _sorted(D, I):
    return oracle.sort_rows(np.ascontiguousarray(D), np.ascontiguousarray(I))


def assert_bitwise(D, I, Dref, Iref):
    Ds, Is = _sorted(D, I)
    Dr, Ir = _sorted(Dref, Iref)
    np.testing.assert_array_equal(Is, Ir)
    assert np.array_equal(Ds.view(np.uint8), Dr.view(np.uint8)), \
        f"max |dD| = {np.nanmax(np.abs(Ds - Dr))}"


def oracle_l1(X, amap, bmap, D, I):
    oracle.gsknn_l1(X, amap, X, bmap, D, I, canonical=True)


def gpu_l1_dev(X, amap, bmap, D, I, hrow=None):
    import torch
    dev = torch.device("cuda")
    tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    tX = torch.from_numpy(X).to(dev)
    gsknn.knn1_dev(tX, torch.from_numpy(amap).to(dev), tX, torch.from_numpy(bmap).to(dev), tD,
                   tI, hrow=None if hrow is None else torch.from_numpy(hrow).to(dev))
    torch.cuda.synchronize()
    return tD.cpu().numpy(), tI.cpu().numpy()


def empty(n, k, dtype):
    return np.full((n, k), np.inf, dtype=dtype), np.full((n, k), -1, dtype=np.int32)


@pytest.mark.parametrize("name", ["l1_d64", "l1_d600", "l1_d1", "l1_k40"])
def test_l1_reference_fixtures(l1_fixtures, name):
    """GPU rows equal the reference's own L1 kernel output (tie-free data)."""
    c = l1_fixtures[name]
    print(f"[gpu] l1 fixture {name}", flush=True)
    X = c["X"]
    n, k = c["D"].shape
    D, I = empty(n, k, np.float64)
    gsknn.dgsknn_1norm_heap(X, c["amap"], X, c["bmap"], D, I)
    assert_bitwise(D, I, c["D"], c["I"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d", [1, 5, 8, 16, 20, 32, 64, 65, 100, 300])
def test_l1_matches_oracle(dtype, d):
    print(f"[gpu] l1 dtype={np.dtype(dtype).name} d={d}", flush=True)
    g = np.random.Generator(np.random.PCG64(100 + d))
    X = g.standard_normal((1500, d)).astype(dtype)
    amap = g.permutation(1500)[:1203].astype(np.int32)  # a partial last group
    bmap = g.permutation(1500)[:333].astype(np.int32)
    k = 16
    D, I = empty(333, k, dtype)
    Dr, Ir = D.copy(), I.copy()
    Dg, Ig = gpu_l1_dev(X, amap, bmap, D, I)
    oracle_l1(X, amap, bmap, Dr, Ir)
    assert_bitwise(Dg, Ig, Dr, Ir)


@pytest.mark.parametrize("d", [16, 64, 40])
def test_l1_candidate_splits_and_prefill(d):
    """Few queries, many candidates (split launch + merge), on rows already
    holding a first call's result."""
    print(f"[gpu] l1 splits d={d}", flush=True)
    g = np.random.Generator(np.random.PCG64(7 + d))
    X = g.standard_normal((40000, d))
    first = g.permutation(40000)[:3000].astype(np.int32)
    amap = g.permutation(40000)[:36001].astype(np.int32)
    bmap = g.permutation(40000)[:70].astype(np.int32)
    k = 24
    D, I = empty(70, k, np.float64)
    Dr, Ir = D.copy(), I.copy()
    D1, I1 = gpu_l1_dev(X, first, bmap, D, I)
    Dg, Ig = gpu_l1_dev(X, amap, bmap, D1.copy(), I1.copy())
    oracle_l1(X, first, bmap, Dr, Ir)
    oracle_l1(X, amap, bmap, Dr, Ir)
    assert_bitwise(Dg, Ig, Dr, Ir)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_l1_lattice_ties_lower_index(dtype):
    pts = np.array([(x, y, z, w) for x in range(7) for y in range(7) for z in range(7)
                    for w in range(3)], dtype=dtype)
    g = np.random.Generator(np.random.PCG64(21))
    amap = g.permutation(len(pts)).astype(np.int32)
    bmap = np.arange(0, len(pts), 5, dtype=np.int32)
    k = 20
    D, I = empty(bmap.size, k, dtype)
    Dr, Ir = D.copy(), I.copy()
    Dg, Ig = gpu_l1_dev(pts, amap, bmap, D, I)
    oracle_l1(pts, amap, bmap, Dr, Ir)
    assert_bitwise(Dg, Ig, Dr, Ir)


def test_l1_heap_row_map_and_partial_fill():
    g = np.random.Generator(np.random.PCG64(31))
    X = g.standard_normal((600, 16))
    amap = g.permutation(600)[:9].astype(np.int32)  # k > m: rows stay partly empty
    bmap = g.permutation(600)[:200].astype(np.int32)
    hrow = g.permutation(300)[:200].astype(np.int32)
    k = 12
    pool_D, pool_I = empty(300, k, np.float64)
    Dg, Ig = gpu_l1_dev(X, amap, bmap, pool_D, pool_I, hrow=hrow)
    Dr, Ir = empty(200, k, np.float64)
    oracle_l1(X, amap, bmap, Dr, Ir)
    assert_bitwise(Dg[hrow], Ig[hrow], Dr, Ir)
    assert np.all((Ig[hrow] == -1).sum(axis=1) == k - 9)
    untouched = np.setdiff1d(np.arange(300), hrow)
    assert np.all(Ig[untouched] == -1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_l1_host_entry_points(dtype):
    g = np.random.Generator(np.random.PCG64(41))
    X = g.standard_normal((900, 24)).astype(dtype)
    amap = g.permutation(900)[:700].astype(np.int32)
    bmap = g.permutation(900)[:150].astype(np.int32)
    k = 10
    D, I = empty(150, k, dtype)
    Dr, Ir = D.copy(), I.copy()
    if dtype == np.float64:
        gsknn.dgsknn_1norm_heap(X, amap, X, bmap, D, I)
    else:
        gsknn.sgsknn_1norm_heap(X, amap, X, bmap, D, I)
    oracle_l1(X, amap, bmap, Dr, Ir)
    assert_bitwise(D, I, Dr, Ir)


def test_l1_full_size_sampled_rows():
    """65536 x 65536 pairs at d = 64, k = 32 (the FP64 benchmark shape); 48
    sampled rows against the oracle, every row filled and sorted descending."""
    import torch
    from rnn_amd import datasets as ds
    print("[gpu] l1 full size", flush=True)
    p = ds.config2()
    dev = torch.device("cuda")
    tX = torch.from_numpy(p.XA).to(dev)
    D = torch.full((p.n, p.k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((p.n, p.k), -1, dtype=torch.int32, device=dev)
    gsknn.knn1_dev(tX, torch.from_numpy(p.amap).to(dev), tX, torch.from_numpy(p.bmap).to(dev),
                   D, I)
    torch.cuda.synchronize()
    Dg, Ig = D.cpu().numpy(), I.cpu().numpy()
    rows = np.linspace(0, p.n - 1, 48).astype(np.int64)
    Dr, Ir = empty(rows.size, p.k, np.float64)
    oracle.gsknn_l1(p.XA, p.amap, p.XB, p.bmap[rows].copy(), Dr, Ir, canonical=True)
    assert_bitwise(Dg[rows], Ig[rows], Dr, Ir)
    assert np.all(Ig >= 0)
    assert np.all(np.diff(Dg, axis=1) <= 0)  # rows come back sorted descending


@pytest.mark.parametrize("dtype,k", [(np.float64, 300), (np.float32, 520)])
def test_l1_large_k(dtype, k):
    """Heaps too large for 64 queries per workgroup (32 or 16 instead)."""
    print(f"[gpu] l1 large k={k} {np.dtype(dtype).name}", flush=True)
    g = np.random.Generator(np.random.PCG64(k))
    X = g.standard_normal((3000, 16)).astype(dtype)
    amap = g.permutation(3000)[:2600].astype(np.int32)
    bmap = g.permutation(3000)[:150].astype(np.int32)
    D, I = empty(150, k, dtype)
    Dr, Ir = D.copy(), I.copy()
    Dg, Ig = gpu_l1_dev(X, amap, bmap, D, I)
    oracle_l1(X, amap, bmap, Dr, Ir)
    assert_bitwise(Dg, Ig, Dr, Ir)
