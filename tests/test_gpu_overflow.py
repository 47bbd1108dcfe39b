"""The re-scan fallback of the two-pass forms: when a query's candidate list
(in one candidate split) outgrows its capacity (the register forms with
per-lane upper-bound sets -- FP32 at d <= 256 and k <= 128, BF16 at the k
it covers -- list into four lane-group segments of clamp(32 k, 1024, 8192)
/ 4 entries; the FP32 LDS-heap form holds clamp(16 k, 512, 4096)), the
re-rank re-scans every candidate of the split exactly.
Thousands of candidates at exactly the same distance from every query
(duplicated rows with distinct indices) all reach the k-th boundary; with
16384 queries a launch splits the candidates at most a few ways, so every
split's list overflows.  The rows must still equal the oracle's (ties to
the lower index; checked on sampled rows), and the overflow count must say
that the fallback ran for every query."""
import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import gsknn

pytestmark = pytest.mark.gpu

NQ = 16384


def _tied_table(d, dtype, copies=3000):
    g = np.random.Generator(np.random.PCG64(606 + d))
    base = g.standard_normal(d)
    m = copies + 200
    X = np.empty((m, d))
    X[:copies] = base                     # copies: one distance per query
    X[copies:] = base + 4.0 + g.standard_normal((200, d))  # farther points
    q = g.standard_normal((NQ, d)) * 0.1 + base + 1.0
    X = np.concatenate([X, q]).astype(dtype)
    amap = g.permutation(m).astype(np.int32)
    bmap = np.arange(m, m + NQ, dtype=np.int32)
    rows = np.sort(g.choice(NQ, 96, replace=False))
    return X, amap, bmap, rows


@pytest.mark.parametrize("d", [32, 128])
def test_fp32_list_overflow_rescans_exactly(d):
    import torch
    print(f"[gpu] fp32 overflow d={d}", flush=True)
    X, amap, bmap, rows = _tied_table(d, np.float32)
    X2 = oracle.sqnorm(X)
    k = 10
    m, n = amap.size, bmap.size
    off = gsknn.lib().gsknn_sknn_overflow_offset(m, n, d, k, gsknn.MODE_AUTO)
    assert off != 2 ** 64 - 1, "the call takes the FP32 MFMA form"
    dev = torch.device("cuda")
    work = torch.zeros(gsknn.work_bytes_s(m, n, d, k), dtype=torch.uint8, device=dev)
    D = torch.full((n, k), float("inf"), dtype=torch.float32, device=dev)
    I = torch.full((n, k), -1, dtype=torch.int32, device=dev)
    t = {name: torch.from_numpy(v).to(dev) for name, v in dict(X=X, X2=X2, am=amap, bm=bmap).items()}
    gsknn.sknn_dev(t["X"], t["X2"], t["am"], t["X"], t["X2"], t["bm"], D, I, work=work)
    torch.cuda.synchronize()
    nover = int(work[off:off + 4].view(torch.int32).item())
    bsub = np.ascontiguousarray(bmap[rows])
    Dr, Ir = ds.empty_heap(rows.size, k, np.float32)
    oracle.gsknn(X, X2, amap, X, X2, bsub, Dr, Ir, canonical=True)
    Ds, Is = oracle.sort_rows(D.cpu().numpy()[rows], I.cpu().numpy()[rows])
    Drs, Irs = oracle.sort_rows(Dr, Ir)
    np.testing.assert_array_equal(Is, Irs)
    assert np.array_equal(Ds.view(np.uint8), Drs.view(np.uint8))
    assert nover >= n, f"{nover} (split, query) lists overflowed for {n} queries"


@pytest.mark.parametrize("d,k,copies", [(32, 10, 12000), (128, 10, 12000), (64, 64, 12000)])
def test_bf16_list_overflow_rescans_exactly(d, k, copies):
    # The register form (k <= 64) lists in four lane-group segments of
    # capacity clamp(32 k, 1024, 8192) / 4: 12000 tied copies give each of
    # the few candidate splits about 3000, so every segment overflows.
    import torch
    print(f"[gpu] bf16 overflow d={d} k={k}", flush=True)
    X, amap, bmap, rows = _tied_table(d, np.float32, copies)
    dev = torch.device("cuda")
    tX = torch.from_numpy(X).to(dev)
    prep = gsknn.bf16_prep_dev(tX)
    over = torch.zeros(1, dtype=torch.int32, device=dev)
    D = torch.full((bmap.size, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((bmap.size, k), -1, dtype=torch.int32, device=dev)
    gsknn.bf16_knn_dev(tX, prep, torch.from_numpy(amap).to(dev), tX, prep,
                       torch.from_numpy(bmap).to(dev), D, I, overflow=over)
    torch.cuda.synchronize()
    X64 = X.astype(np.float64)
    Db, Ib = oracle.brute(X64, amap, X64, np.ascontiguousarray(bmap[rows]), k)
    Dg, Ig = oracle.sort_rows(D.cpu().numpy()[rows], I.cpu().numpy()[rows])
    np.testing.assert_array_equal(Ig, Ib)
    assert np.all(np.abs(Dg - Db) <= 1e-12 * np.maximum(Db, 1.0))
    assert int(over.item()) >= bmap.size
