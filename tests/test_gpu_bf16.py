"""GPU parity of the BF16 candidate pass + FP64 re-rank in its register form
(k <= 128: upper-bound sets in registers, one list segment per lane
group, candidates gathered into a panel) against the FP64 oracle.

Bar: neighbour indices exact; distances within
|D - D_ref| <= 1e-12 * max(|D_ref|, |q|^2 + |r|^2) of the oracle's (the
re-rank sums direct differences, the oracle the reference's expanded form),
and the distance multisets of tie-free rows equal to the FP64
direct-difference oracle's within the same bar.
"""
import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import gsknn

from test_gpu_gsknn import TOL64, _sorted, assert_rows

pytestmark = pytest.mark.gpu


def _run(X, amaps, bmap, D, I, hrow=None, pool=None):
    """Calls in sequence into the same rows; returns the rows (or the pool)."""
    import torch
    dev = torch.device("cuda")
    tX = torch.from_numpy(X).to(dev)
    prep = gsknn.bf16_prep_dev(tX)
    if pool is not None:
        tD, tI = torch.from_numpy(pool[0]).to(dev), torch.from_numpy(pool[1]).to(dev)
    else:
        tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    th = torch.from_numpy(hrow).to(dev) if hrow is not None else None
    over = torch.zeros(1, dtype=torch.int32, device=dev)
    tb = torch.from_numpy(bmap).to(dev)
    for am in amaps:
        gsknn.bf16_knn_dev(tX, prep, torch.from_numpy(am).to(dev), tX, prep, tb, tD, tI,
                           hrow=th, overflow=over)
    torch.cuda.synchronize()
    return tD.cpu().numpy(), tI.cpu().numpy(), int(over.item())


def _oracle(X, amaps, bmap, k, canonical=False):
    X64 = X.astype(np.float64)
    X2 = oracle.sqnorm(X64)
    Dr, Ir = ds.empty_heap(bmap.size, k)
    for am in amaps:
        oracle.gsknn(X64, X2, am, X64, X2, bmap, Dr, Ir, canonical=canonical)
    return Dr, Ir, X2


@pytest.mark.parametrize("d,k", [(256, 48), (256, 64), (97, 50), (50, 64), (8, 56), (200, 63),
                                 (64, 1), (100, 3), (256, 7), (33, 17), (128, 32), (256, 10),
                                 (256, 65), (128, 100), (70, 128)])
def test_register_form_two_calls(d, k):
    # Low-rank points, two calls of disjoint candidates into the same rows.
    g = np.random.Generator(np.random.PCG64(1000 + d + k))
    N = 9000
    X = ds.low_rank(N, d, 10, seed=d + 3)
    perm = g.permutation(N).astype(np.int32)
    amaps = [perm[:3500].copy(), perm[3500:6000].copy()]
    bmap = perm[6000:].copy()
    print(f"[gpu] bf16 register form d={d} k={k}", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, over = _run(X, amaps, bmap, D, I)
    assert over == 0
    Dr, Ir, X2 = _oracle(X, amaps, bmap, k)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
    assert oracle.rows_are_heaps(Dg, Ig)
    # FP64 direct differences (the re-rank's definition) on the union.
    Dd, Id = ds.empty_heap(bmap.size, k)
    oracle.direct_knn_f32in(X, np.concatenate(amaps), X, bmap, Dd, Id)
    assert_rows(Dg, Ig, Dd, Id, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)


def test_register_form_lattice_ties():
    # Integer points (exact in BF16): many exact ties at the k-th distance;
    # canonical (distance, lower index) result, two calls.
    g = np.random.Generator(np.random.PCG64(78))
    X = np.floor(g.random((4000, 24)) * 4).astype(np.float32)
    amaps = [g.integers(0, 4000, 1800).astype(np.int32),
             g.integers(0, 4000, 1500).astype(np.int32)]
    bmap = g.permutation(4000)[:700].astype(np.int32)
    k = 60
    print("[gpu] bf16 register form lattice", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, _ = _run(X, amaps, bmap, D, I)
    Dr, Ir, _ = _oracle(X, amaps, bmap, k, canonical=True)
    assert_rows(Dg, Ig, Dr, Ir)


def test_register_form_candidate_splits():
    # Few queries against many candidates: the filter splits the candidates
    # over workgroups (per-split lists, four segments each).
    g = np.random.Generator(np.random.PCG64(79))
    X = ds.low_rank(70000, 64, 8, seed=6)
    amap = np.arange(70000, dtype=np.int32)
    bmap = g.permutation(70000)[:45].astype(np.int32)
    k = 52
    print("[gpu] bf16 register form splits n=45 m=70000", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, over = _run(X, [amap], bmap, D, I)
    assert over == 0
    Dr, Ir, X2 = _oracle(X, [amap], bmap, k)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)


def test_register_form_heap_row_map():
    # Rows scattered in a larger pool through hrow; the other rows untouched.
    g = np.random.Generator(np.random.PCG64(80))
    X = ds.low_rank(8000, 128, 10, seed=7)
    amap = g.permutation(8000)[:5000].astype(np.int32)
    bmap = g.permutation(8000)[:900].astype(np.int32)
    k = 64
    pool = 1300
    hrow = g.permutation(pool)[:bmap.size].astype(np.int32)
    Dp = np.full((pool, k), 7.5, dtype=np.float64)
    Ip = np.full((pool, k), -3, dtype=np.int32)
    Dp[hrow], Ip[hrow] = np.inf, -1
    print("[gpu] bf16 register form heap-row map", flush=True)
    Dg, Ig, _ = _run(X, [amap], bmap, None, None, hrow=hrow, pool=(Dp, Ip))
    Dr, Ir, X2 = _oracle(X, [amap], bmap, k)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg[hrow], Ig[hrow], Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
    untouched = np.setdiff1d(np.arange(pool), hrow)
    assert np.all(Dg[untouched] == 7.5) and np.all(Ig[untouched] == -3)


def test_register_form_candidate_counts():
    # Each query keeps at least k candidates in its lists (empty rows), and
    # the count API reports nsplit x 4 lists per query.
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(81))
    X = ds.low_rank(6000, 256, 10, seed=8)
    amap = g.permutation(6000)[:4000].astype(np.int32)
    bmap = g.permutation(6000)[:1000].astype(np.int32)
    k = 64
    tX = torch.from_numpy(X).to(dev)
    prep = gsknn.bf16_prep_dev(tX)
    work = torch.empty(gsknn.bf16_work_bytes(amap.size, bmap.size, 256, k), dtype=torch.uint8,
                       device=dev)
    D = torch.full((bmap.size, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((bmap.size, k), -1, dtype=torch.int32, device=dev)
    gsknn.bf16_knn_dev(tX, prep, torch.from_numpy(amap).to(dev), tX, prep,
                       torch.from_numpy(bmap).to(dev), D, I, work=work)
    cnt = gsknn.bf16_candidate_counts(work, amap.size, bmap.size, 256, k).cpu().numpy()
    assert cnt.shape == (bmap.size,)
    assert np.all(cnt >= k) and np.all(cnt <= amap.size)


@pytest.mark.parametrize("d,k", [(80, 32), (150, 10), (200, 100), (70, 64)])
def test_rows_padded_to_filter_k_steps(d, k):
    # ceil(d / 32) not a power of two: prepared rows are padded to the
    # filter's 32 KS dims (KS = 4 or 8 here); both filter forms.
    g = np.random.Generator(np.random.PCG64(2000 + d + k))
    N = 7000
    X = ds.low_rank(N, d, 10, seed=d + 11)
    perm = g.permutation(N).astype(np.int32)
    amaps = [perm[:4500].copy()]
    bmap = perm[4500:].copy()
    print(f"[gpu] bf16 padded rows d={d} k={k}", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, _ = _run(X, amaps, bmap, D, I)
    Dr, Ir, X2 = _oracle(X, amaps, bmap, k)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)


@pytest.mark.parametrize("warm", [0, 1, 5, 18, 40])
def test_register_form_warm_up_pass(warm, monkeypatch):
    # The register form's warm-up pass over the first tiles of each split
    # (upper bounds from each lane's smallest value per tile), then the main
    # pass over every tile; 0 = none, 18 = the whole split (7 splits of 18
    # tiles here), 40 = more than the split holds.  Carried-over rows.
    g = np.random.Generator(np.random.PCG64(3000 + warm))
    N = 12000
    X = ds.low_rank(N, 256, 10, seed=12)
    perm = g.permutation(N).astype(np.int32)
    amaps = [perm[:700].copy(), perm[700:8700].copy()]
    bmap = perm[9000:].copy()
    k = 64
    monkeypatch.setenv("GSKNN_BF16_WARM", str(warm))
    print(f"[gpu] bf16 register form warm-up {warm}", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, over = _run(X, amaps, bmap, D, I)
    assert over == 0
    Dr, Ir, X2 = _oracle(X, amaps, bmap, k)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
