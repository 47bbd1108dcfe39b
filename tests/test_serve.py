"""Serving loop: each request's result is the same alone and in the mix, and
every request's rows match the oracle (FP32 mode: bitwise; BF16 mode, FP64
direct differences: indices exact, distances within 1e-12 at the norm
scale)."""
import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import serve


def test_trace_shape():
    W = np.random.Generator(np.random.PCG64(0)).standard_normal((4, 16))
    reqs = serve.make_trace(W, 50, rate=200.0, seed=5, d=16, max_q=64)
    assert len(reqs) == 50
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr) and arr[0] == 0.0
    sizes = [r.queries.shape[0] for r in reqs]
    assert min(sizes) >= 1 and max(sizes) <= 64
    assert {r.k for r in reqs} <= {10, 32, 100}
    assert all(r.queries.dtype == np.float32 and r.queries.shape[1] == 16 for r in reqs)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["bf16", "fp32"])
def test_alone_equals_in_mix(mode):
    import torch
    dev = torch.device("cuda")
    N, d = 40000, 128
    g = np.random.Generator(np.random.PCG64(4))
    W = g.standard_normal((16, d))
    Rn = ds.low_rank(N, d, 16, seed=11, W=W)
    R = torch.from_numpy(Rn).to(dev)
    srv = serve.KnnServer(R, mode=mode, hbm_budget=2 << 30, max_batch=2048, streams=3)
    reqs = serve.make_trace(W, 24, rate=1000.0, seed=6, d=d, max_q=600)
    print(f"[gpu] serve {mode}: {sum(r.queries.shape[0] for r in reqs)} queries", flush=True)
    _, mixed = srv.run(reqs, realtime=False)
    assert srv.batches_launched < len(reqs)  # requests were batched together
    got = {r.rid: (r.D.numpy().copy(), r.I.numpy().copy()) for r in mixed}
    for r in reqs:
        alone = serve.Request(r.rid, 0.0, r.queries, r.k)
        srv.run([alone], realtime=False)
        Da, Ia = alone.D.numpy(), alone.I.numpy()
        Dm, Im = got[r.rid]
        assert np.array_equal(Ia, Im) and np.array_equal(Da.view(np.uint8), Dm.view(np.uint8))
    # every request against the oracle
    amap = np.arange(N, dtype=np.int32)
    for r in reqs:
        q = r.queries.shape[0]
        bmap = np.arange(q, dtype=np.int32)
        Dm, Im = got[r.rid]
        if mode == "bf16":
            R64, Q64 = Rn.astype(np.float64), r.queries.astype(np.float64)
            R2, Q2 = oracle.sqnorm(R64), oracle.sqnorm(Q64)
            Dr = np.full((q, r.k), np.inf)
            Ir = np.full((q, r.k), -1, dtype=np.int32)
            oracle.gsknn(R64, R2, amap, Q64, Q2, bmap, Dr, Ir, canonical=True)
            a, b = oracle.sort_rows(Dm, Im), oracle.sort_rows(Dr, Ir)
            np.testing.assert_array_equal(a[1], b[1])
            scale = np.maximum(np.abs(b[0]), Q2[:, None] + R2[b[1]])
            assert np.all(np.abs(a[0] - b[0]) <= 1e-12 * scale)
        else:
            R2, Q2 = oracle.sqnorm(Rn), oracle.sqnorm(r.queries)
            Dr = np.full((q, r.k), np.inf, dtype=np.float32)
            Ir = np.full((q, r.k), -1, dtype=np.int32)
            oracle.gsknn(Rn, R2, amap, np.ascontiguousarray(r.queries), Q2, bmap, Dr, Ir,
                         canonical=True)
            a, b = oracle.sort_rows(Dm, Im), oracle.sort_rows(Dr, Ir)
            np.testing.assert_array_equal(a[1], b[1])
            assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
