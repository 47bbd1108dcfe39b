"""GPU parity of libgsknn.so against the oracle, through the C-ABI.

Bars (neighbour indices exact; distances per precision):
  * every FP64 and FP32 path (exact VALU kernels, FP64 MFMA filter + oracle
    distances, FP32 MFMA filter + oracle-arithmetic re-rank): distances
    bit-identical to the oracle, which is itself bit-identical to the
    reference's dgsknn_ref;
  * BF16 candidates + FP64 direct-difference re-rank:
    |D - D_ref| <= 1e-12 * max(|D_ref|, |q|^2 + |r|^2).
Inputs without exact ties are compared with the reference's own tie rule
(the oracle in compat mode); inputs with exact ties (integer lattices) with
the canonical (distance, lower index) rule.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import gsknn

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _sorted(D, I):
    return oracle.sort_rows(np.ascontiguousarray(D), np.ascontiguousarray(I))


def assert_rows(D, I, Dref, Iref, X2q=None, X2c_rows=None, tol=None):
    """Indices exact after sorting by (dist, idx); distances bitwise or within tol."""
    Ds, Is = _sorted(D, I)
    Dr, Ir = _sorted(Dref, Iref)
    np.testing.assert_array_equal(Is, Ir)
    if tol is None:
        assert np.array_equal(Ds.view(np.uint8), Dr.view(np.uint8)), \
            f"max |dD| = {np.nanmax(np.abs(Ds - Dr))}"
    else:
        # Equal entries (placeholders at +inf included) pass; any other pair
        # must be finite and within the norm-scaled tolerance.
        same = Ds == Dr
        with np.errstate(invalid="ignore"):
            scale = np.where(np.isfinite(Dr), np.maximum(np.abs(Dr), X2q[:, None] + X2c_rows), 0.0)
            ok = same | (np.isfinite(Ds) & np.isfinite(Dr) & (np.abs(Ds - Dr) <= tol * scale))
        assert ok.all(), f"{(~ok).sum()} distances beyond tolerance"


def run_gpu(X, X2, amaps, bmap, D, I):
    for am in amaps:
        if X.dtype == np.float64:
            gsknn.dgsknn_heap(X, X2, am, X, X2, bmap, D, I)
        else:
            gsknn.sgsknn_heap(X, X2, am, X, X2, bmap, D, I)


def run_oracle(X, X2, amaps, bmap, D, I, canonical=False):
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, bmap, D, I, canonical=canonical)


@pytest.mark.parametrize("name", ["c1small", "partial", "repeats", "odd", "d1", "radius", "f32"])
def test_reference_fixtures(fixtures, name):
    c = fixtures[name]
    print(f"[gpu] fixture {name}", flush=True)
    amaps = [c["amap"]] + [c[f"amap{i}"] for i in range(2, 5) if f"amap{i}" in c]
    D, I = c["D0"].copy(), c["I0"].copy()
    for i, am in enumerate(amaps, start=1):
        run_gpu(c["X"], c["X2"], [am], c["bmap"], D, I)
        assert_rows(D, I, c[f"D{i}"], c[f"I{i}"])
        assert oracle.rows_are_heaps(D, I)


def test_lattice_ties_lower_index_wins(fixtures):
    c = fixtures["lattice"]
    X, X2, am, bm = c["X"], c["X2"], c["amap"], c["bmap"]
    for mode in ("exact", "mfma"):
        print(f"[gpu] lattice {mode}", flush=True)
        D, I = c["D0"].copy(), c["I0"].copy()
        Dr, Ir = c["D0"].copy(), c["I0"].copy()
        import torch
        dev = torch.device("cuda")
        t = {k: torch.from_numpy(v).to(dev) for k, v in
             dict(X=X, X2=X2, am=am, bm=bm, D=D, I=I).items()}
        gsknn.dknn_dev(t["X"], t["X2"], t["am"], t["X"], t["X2"], t["bm"], t["D"], t["I"],
                       mode=gsknn.MODE_EXACT if mode == "exact" else gsknn.MODE_MFMA)
        torch.cuda.synchronize()
        run_oracle(X, X2, [am], bm, Dr, Ir, canonical=True)
        assert_rows(t["D"].cpu().numpy(), t["I"].cpu().numpy(), Dr, Ir)


def test_config1_full_both_calls():
    p = ds.config1()
    print("[gpu] config1 N=8192 d=16 m=n=2048 k=16", flush=True)
    D, I = ds.empty_heap(p.n, p.k)
    Dr, Ir = D.copy(), I.copy()
    for am in [p.amap] + p.extra_amaps:
        gsknn.dgsknn_heap(p.XA, p.XA2, am, p.XB, p.XB2, p.bmap, D, I)
        oracle.gsknn(p.XA, p.XA2, am, p.XB, p.XB2, p.bmap, Dr, Ir)
        assert_rows(D, I, Dr, Ir)


def test_set_device_binds_host_api_to_that_device():
    # gsknn_set_device rejects devices that do not exist and leaves the
    # selection unchanged; the host-pointer API then stages and launches on
    # the selected device's own stream and buffers.
    import torch
    L = gsknn.lib()
    count = torch.cuda.device_count()
    assert L.gsknn_set_device(-1) == 1
    assert L.gsknn_set_device(count) == 1
    assert "gsknn_set_device" in gsknn.last_error()
    p = ds.config1()
    for dev in range(count):
        assert L.gsknn_set_device(dev) == 0
        print(f"[gpu] set_device {dev} config1 m=n=2048 d=16 k=16", flush=True)
        D, I = ds.empty_heap(p.n, p.k)
        Dr, Ir = D.copy(), I.copy()
        gsknn.dgsknn_heap(p.XA, p.XA2, p.amap, p.XB, p.XB2, p.bmap, D, I)
        oracle.gsknn(p.XA, p.XA2, p.amap, p.XB, p.XB2, p.bmap, Dr, Ir)
        assert_rows(D, I, Dr, Ir)
    assert L.gsknn_set_device(0) == 0


def test_config1_literal_second_call():
    # The other reading of the second call: a disjoint query set into the
    # same heap rows.
    p = ds.config1()
    pi = np.random.Generator(np.random.PCG64(0)).permutation(8192).astype(np.int32)
    D, I = ds.empty_heap(p.n, p.k)
    Dr, Ir = D.copy(), I.copy()
    for bm in (p.bmap, pi[2048:4096].copy()):
        gsknn.dgsknn_heap(p.XA, p.XA2, p.amap, p.XB, p.XB2, bm, D, I)
        oracle.gsknn(p.XA, p.XA2, p.amap, p.XB, p.XB2, bm, Dr, Ir)
        assert_rows(D, I, Dr, Ir)


@pytest.mark.parametrize("mode", ["auto", "exact"])
def test_config2_scaled(mode, monkeypatch):
    # Config-2 generator at 1/8 scale (N=16384, m=n=8192, d=64, k=32) with the
    # 1024-candidate prefill call.
    p = ds.config2(scale=8)
    print(f"[gpu] config2/8 mode={mode}", flush=True)
    import torch
    dev = torch.device("cuda")
    X = torch.from_numpy(p.XA).to(dev)
    X2 = torch.from_numpy(p.XA2).to(dev)
    D, I = ds.empty_heap(p.n, p.k)
    Dr, Ir = D.copy(), I.copy()
    tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    bm = torch.from_numpy(p.bmap).to(dev)
    m = gsknn.MODE_AUTO if mode == "auto" else gsknn.MODE_EXACT
    for am in (p.prefill_amap, p.amap):
        gsknn.dknn_dev(X, X2, torch.from_numpy(am).to(dev), X, X2, bm, tD, tI, mode=m)
        torch.cuda.synchronize()
        oracle.gsknn(p.XA, p.XA2, am, p.XB, p.XB2, p.bmap, Dr, Ir)
        Dg, Ig = tD.cpu().numpy(), tI.cpu().numpy()
        assert_rows(Dg, Ig, Dr, Ir)


@pytest.mark.parametrize("d", [16, 64, 100])
def test_candidate_split_path(d):
    # Few queries, many candidates: the launch splits the candidates over
    # workgroups and merges partial lists.
    g = np.random.Generator(np.random.PCG64(d))
    X = g.standard_normal((30000, d))
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, 30000, 25000).astype(np.int32)
    bmap = g.integers(0, 30000, 70).astype(np.int32)
    D, I = ds.empty_heap(70, 20)
    Dr, Ir = D.copy(), I.copy()
    print(f"[gpu] split d={d}", flush=True)
    gsknn.dgsknn_heap(X, X2, amap, X, X2, bmap, D, I)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir)
    assert_rows(D, I, Dr, Ir)


def test_sgsknn_fp32_bitwise():
    g = np.random.Generator(np.random.PCG64(42))
    X = g.standard_normal((6000, 24)).astype(np.float32)
    X2 = oracle.sqnorm(X)
    amap = g.permutation(6000)[:4000].astype(np.int32)
    bmap = g.permutation(6000)[:900].astype(np.int32)
    D, I = ds.empty_heap(900, 17, np.float32)
    Dr, Ir = D.copy(), I.copy()
    print("[gpu] sgsknn fp32", flush=True)
    gsknn.sgsknn_heap(X, X2, amap, X, X2, bmap, D, I)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir)
    assert_rows(D, I, Dr, Ir)


def test_python_binding_mirror():
    # python/gsknn.py argument order: (d, k, Xq, Xq2, q, Xr, Xr2, r, D, I).
    g = np.random.Generator(np.random.PCG64(1))
    Xq = g.random((300, 6))
    Xr = g.random((500, 6))
    Xq2, Xr2 = oracle.sqnorm(Xq), oracle.sqnorm(Xr)
    q = np.arange(0, 300, 3, dtype=np.int32)
    r = np.arange(0, 500, 2, dtype=np.int32)
    D = np.full((q.size, 5), 1.79e308)
    I = np.full((q.size, 5), -1, dtype=np.int32)
    gsknn.dgsknn(6, 5, Xq, Xq2, q, Xr, Xr2, r, D, I)
    Dr, Ir = np.full_like(D, 1.79e308), np.full_like(I, -1)
    oracle.gsknn(Xr, Xr2, r, Xq, Xq2, q, Dr, Ir)
    assert_rows(D, I, Dr, Ir)
    Ds = D.astype(np.float32)
    Xq32, Xr32 = Xq.astype(np.float32), Xr.astype(np.float32)
    Xq232, Xr232 = oracle.sqnorm(Xq32), oracle.sqnorm(Xr32)
    Ds = np.full((q.size, 5), 1.79e30, dtype=np.float32)
    Is = np.full((q.size, 5), -1, dtype=np.int32)
    gsknn.sgsknn(6, 5, Xq32, Xq232, q, Xr32, Xr232, r, Ds, Is)
    Dr, Ir = np.full_like(Ds, 1.79e30), np.full_like(Is, -1)
    oracle.gsknn(Xr32, Xr232, r, Xq32, Xq232, q, Dr, Ir)
    assert_rows(Ds, Is, Dr, Ir)


def test_device_api_with_heap_row_map():
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(8))
    X = g.random((2000, 40))
    X2 = oracle.sqnorm(X)
    amap = g.permutation(2000)[:1500].astype(np.int32)
    bmap = g.permutation(2000)[:300].astype(np.int32)
    hrow = g.permutation(500)[:300].astype(np.int32)  # rows scattered in a 500-row pool
    k = 12
    pool_D = np.full((500, k), np.inf)
    pool_I = np.full((500, k), -1, dtype=np.int32)
    tD, tI = torch.from_numpy(pool_D).to(dev), torch.from_numpy(pool_I).to(dev)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    gsknn.dknn_dev(tX, tX2, torch.from_numpy(amap).to(dev), tX, tX2,
                   torch.from_numpy(bmap).to(dev), tD, tI,
                   hrow=torch.from_numpy(hrow).to(dev), mode=gsknn.MODE_EXACT)
    torch.cuda.synchronize()
    Dg, Ig = tD.cpu().numpy(), tI.cpu().numpy()
    Dr, Ir = ds.empty_heap(300, k)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir)
    assert_rows(Dg[hrow], Ig[hrow], Dr, Ir)
    untouched = np.setdiff1d(np.arange(500), hrow)
    assert np.all(Ig[untouched] == -1)


def test_large_k_header_layout_var3():
    # r > 512 goes through dgsknn_var3 on heapCreate_d rows with the 3-slot
    # header and a 4-ary heap (frame/gsknn_heap.c:257-308).
    g = np.random.Generator(np.random.PCG64(13))
    X = g.random((3000, 8))
    X2 = oracle.sqnorm(X)
    L = gsknn.lib()
    k, n = 600, 40
    bmap = g.permutation(3000)[:n].astype(np.int32)
    h = L.heapCreate_d(n, k, 1.79e308)
    ldk = h.contents.ldk
    Dr, Ir = ds.empty_heap(n, k)
    Dr[:] = 1.79e308
    for m in (250, 1200):  # first call leaves rows partially filled
        amap = g.permutation(3000)[:m].astype(np.int32)
        print(f"[gpu] var3 k={k} m={m}", flush=True)
        L.dgsknn(m, n, 8, k, gsknn._p(X), gsknn._p(X2), gsknn._p(amap), gsknn._p(X), gsknn._p(X2),
                 gsknn._p(bmap), ctypes.cast(h, ctypes.c_void_p))
        assert gsknn.last_error() == ""
        oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir, canonical=True)
        Dh = np.ctypeslib.as_array(h.contents.D, shape=(n * ldk,)).reshape(n, ldk)
        Ih = np.ctypeslib.as_array(h.contents.I, shape=(n * ldk,)).reshape(n, ldk)
        fill = Ih[:, 1]
        assert np.all(Ih[:, 0] == k) and np.all(Ih[:, 2] == ldk)
        Ds, Is = _sorted(Dh[:, 3:3 + k], Ih[:, 3:3 + k])
        Drs, Irs = _sorted(Dr, Ir)
        np.testing.assert_array_equal(Is, Irs)
        assert np.array_equal(Ds, Drs)
        assert np.all(fill == min(k, 250 if m == 250 else k))
    L.heapFree_d(ctypes.cast(h, ctypes.c_void_p))


def test_sq2nrm_matrix():
    g = np.random.Generator(np.random.PCG64(21))
    X = g.random((400, 9))
    X2 = oracle.sqnorm(X)
    amap = g.permutation(400)[:70].astype(np.int32)
    bmap = g.permutation(400)[:50].astype(np.int32)
    ldc = 72
    C = np.full(ldc * 50, -7.0)
    L = gsknn.lib()
    L.dgssq2nrm(70, 50, 9, gsknn._p(X), gsknn._p(X2), gsknn._p(amap), gsknn._p(X), gsknn._p(X2),
                gsknn._p(bmap), gsknn._p(C), ldc)
    C = C.reshape(50, ldc)
    A, B = X[amap], X[bmap]
    s = np.zeros((50, 70))
    for p in range(9):
        s = s + B[:, p:p + 1] * A[:, p][None, :]
    ref = (s * -2.0 + X2[amap][None, :]) + X2[bmap][:, None]
    assert np.array_equal(C[:, :70], ref)
    assert np.all(C[:, 70:] == -7.0)


def test_deterministic_and_empty():
    p = ds.config1(scale=4)
    D1, I1 = ds.empty_heap(p.n, p.k)
    D2, I2 = D1.copy(), I1.copy()
    gsknn.dgsknn_heap(p.XA, p.XA2, p.amap, p.XB, p.XB2, p.bmap, D1, I1)
    gsknn.dgsknn_heap(p.XA, p.XA2, p.amap, p.XB, p.XB2, p.bmap, D2, I2)
    assert np.array_equal(D1, D2) and np.array_equal(I1, I2)
    D3, I3 = D1.copy(), I1.copy()
    gsknn.dgsknn_heap(p.XA, p.XA2, p.amap[:0].copy(), p.XB, p.XB2, p.bmap, D3, I3)
    assert np.array_equal(D1, D3) and np.array_equal(I1, I3)


def _bf16_run(XA32, amap, XB32, bmap, D, I, hrow=None):
    import torch
    dev = torch.device("cuda")
    tA = torch.from_numpy(XA32).to(dev)
    tB = tA if XB32 is XA32 else torch.from_numpy(XB32).to(dev)
    pA = gsknn.bf16_prep_dev(tA)
    pB = pA if XB32 is XA32 else gsknn.bf16_prep_dev(tB)
    tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    over = torch.zeros(1, dtype=torch.int32, device=dev)
    gsknn.bf16_knn_dev(tA, pA, torch.from_numpy(amap).to(dev), tB, pB,
                       torch.from_numpy(bmap).to(dev), tD, tI, overflow=over)
    torch.cuda.synchronize()
    return tD.cpu().numpy(), tI.cpu().numpy(), int(over.item())


# k = 600 and 700 take the 2-wave filter shape, k = 1100 the 1-wave shape
# (heaps too large for 4 waves of queries in LDS).
@pytest.mark.parametrize("d,k", [(256, 64), (128, 32), (32, 16), (50, 10), (256, 600),
                                 (256, 1100), (64, 700)])
def test_bf16_rerank_matches_fp64_oracle(d, k):
    # Low-rank generator of the precision workload (latent 10, noise 0.01).
    g = np.random.Generator(np.random.PCG64(d + k))
    N = 6000
    X = ds.low_rank(N, d, 10, seed=d)
    perm = g.permutation(N).astype(np.int32)
    amap, bmap = perm[:4000].copy(), perm[4000:].copy()
    D, I = ds.empty_heap(bmap.size, k)
    print(f"[gpu] bf16 d={d} k={k}", flush=True)
    Dg, Ig, over = _bf16_run(X, amap, X, bmap, D.copy(), I.copy())
    # Oracle: dgsknn_ref arithmetic on the FP64-upcast points.
    X64 = X.astype(np.float64)
    X2 = oracle.sqnorm(X64)
    Dr, Ir = D.copy(), I.copy()
    oracle.gsknn(X64, X2, amap, X64, X2, bmap, Dr, Ir)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
    # and the FP64 direct-difference definition of the re-rank
    Dd, Id = D.copy(), I.copy()
    oracle.direct_knn_f32in(X, amap, X, bmap, Dd, Id)
    assert_rows(Dg, Ig, Dd, Id, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
    print(f"[gpu] bf16 overflowed queries: {over}", flush=True)


def test_bf16_carry_over_and_ties():
    # Integer points (exact in BF16) with many exact ties, two calls into the
    # same heap rows: canonical (distance, lower index) result.
    g = np.random.Generator(np.random.PCG64(77))
    X = np.floor(g.random((3000, 40)) * 4).astype(np.float32)
    amap1 = g.integers(0, 3000, 1500).astype(np.int32)
    amap2 = g.integers(0, 3000, 1700).astype(np.int32)
    bmap = g.permutation(3000)[:500].astype(np.int32)
    D, I = ds.empty_heap(500, 20)
    Dg, Ig, _ = _bf16_run(X, amap1, X, bmap, D, I)
    Dg, Ig, _ = _bf16_run(X, amap2, X, bmap, Dg.copy(), Ig.copy())
    X64 = X.astype(np.float64)
    X2 = oracle.sqnorm(X64)
    Dr, Ir = ds.empty_heap(500, 20)
    for am in (amap1, amap2):
        oracle.gsknn(X64, X2, am, X64, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(Dg, Ig, Dr, Ir)


# k = 700 takes the 2-wave filter shape and k = 1300 the 1-wave shape at
# d = 24 (heaps too large for 4 waves of queries in LDS).
@pytest.mark.parametrize("d,k,lattice", [(24, 17, False), (100, 32, False), (256, 64, False),
                                         (12, 10, True), (24, 700, False), (24, 1300, False),
                                         (128, 1, False), (64, 65, False), (128, 100, False),
                                         (33, 128, False), (16, 120, True), (128, 200, False),
                                         (200, 100, False), (129, 10, False)])
def test_sgsknn_fp32_mfma_rerank_bitwise(d, k, lattice):
    # n >= 4096 takes the FP32 MFMA candidate pass + re-rank in the oracle's
    # FP32 arithmetic; distances must still be the oracle's bits.
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(d * 7 + k))
    N = 12000
    if lattice:
        X = np.floor(g.random((N, d)) * 5).astype(np.float32)
    else:
        X = g.standard_normal((N, d)).astype(np.float32)
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, N, 6000).astype(np.int32)
    bmap = g.permutation(N)[:4500].astype(np.int32)
    D, I = ds.empty_heap(bmap.size, k, np.float32)
    print(f"[gpu] sgsknn mfma d={d} k={k} lattice={lattice}", flush=True)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    gsknn.sknn_dev(tX, tX2, torch.from_numpy(amap).to(dev), tX, tX2,
                   torch.from_numpy(bmap).to(dev), tD, tI)
    torch.cuda.synchronize()
    Dg, Ig = tD.cpu().numpy(), tI.cpu().numpy()
    # FP32 distances tie exactly now and then (ulp ~ 6e-5 at |D| ~ 500): the
    # canonical oracle fixes the index choice; the reference's own rule must
    # give the same distance multiset.
    Dr, Ir = D.copy(), I.copy()
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(Dg, Ig, Dr, Ir)
    Dc, Ic = D.copy(), I.copy()
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dc, Ic, canonical=False)
    assert np.array_equal(_sorted(Dg, Ig)[0], _sorted(Dc, Ic)[0])


def test_bf16_and_fp32_few_queries_many_candidates():
    # Few queries against many candidates: the filter kernels split the
    # candidates over workgroups (per-split bounds and lists).
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(99))
    X = ds.low_rank(60000, 64, 8, seed=5)
    amap = np.arange(60000, dtype=np.int32)
    bmap = g.permutation(60000)[:37].astype(np.int32)
    k = 24
    print("[gpu] split bf16/fp32 n=37 m=60000", flush=True)
    D, I = ds.empty_heap(bmap.size, k)
    Dg, Ig, over = _bf16_run(X, amap, X, bmap, D.copy(), I.copy())
    X64 = X.astype(np.float64)
    X2 = oracle.sqnorm(X64)
    Dr, Ir = D.copy(), I.copy()
    oracle.gsknn(X64, X2, amap, X64, X2, bmap, Dr, Ir, canonical=True)
    Is = _sorted(Dr, Ir)[1]
    assert_rows(Dg, Ig, Dr, Ir, X2q=X2[bmap], X2c_rows=X2[Is], tol=TOL64)
    X2f = oracle.sqnorm(X)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2f).to(dev)
    Df, If = ds.empty_heap(bmap.size, k, np.float32)
    tD, tI = torch.from_numpy(Df).to(dev), torch.from_numpy(If).to(dev)
    gsknn.sknn_dev(tX, tX2, torch.from_numpy(amap).to(dev), tX, tX2,
                   torch.from_numpy(bmap).to(dev), tD, tI)
    torch.cuda.synchronize()
    Dr, Ir = Df.copy(), If.copy()
    oracle.gsknn(X, X2f, amap, X, X2f, bmap, Dr, Ir, canonical=True)
    assert_rows(tD.cpu().numpy(), tI.cpu().numpy(), Dr, Ir)


@pytest.mark.parametrize("k,ldk,dtype", [(1, 1, np.float64), (7, 9, np.float32), (32, 32, np.float64),
                                         (64, 70, np.float32), (600, 600, np.float64),
                                         (4096, 4096, np.float32)])
def test_rows_ascending_matches_oracle_sort(k, ldk, dtype):
    # Sorted output on the device == the oracle's (distance, index) sort, on
    # rows with duplicates, exact ties broken by index, and (+inf, -1) slots.
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(k * 7 + ldk))
    n = 257
    print(f"[gpu] rows_ascending k={k} ldk={ldk} {np.dtype(dtype).name}", flush=True)
    D = np.round(g.random((n, ldk)) * 50).astype(dtype)       # many exact ties
    I = g.integers(0, 1000, size=(n, ldk)).astype(np.int32)
    D[::5, : k // 3] = np.inf                                  # unfilled slots
    I[::5, : k // 3] = -1
    if ldk > 1:
        D[1::7, 0] = D[1::7, 1]                                # duplicate entries
        I[1::7, 0] = I[1::7, 1]
    tD, tI = torch.from_numpy(D.copy()).to(dev), torch.from_numpy(I.copy()).to(dev)
    gsknn.rows_ascending_dev(tD[:, :k], tI[:, :k])
    torch.cuda.synchronize()
    Dg, Ig = tD.cpu().numpy(), tI.cpu().numpy()
    Dr, Ir = oracle.sort_rows(D[:, :k].copy(), I[:, :k].copy())
    assert np.array_equal(Dg[:, :k].view(np.uint8), Dr.view(np.uint8))
    np.testing.assert_array_equal(Ig[:, :k], Ir)
    np.testing.assert_array_equal(Dg[:, k:], D[:, k:])        # padding columns untouched
    np.testing.assert_array_equal(Ig[:, k:], I[:, k:])


def test_rows_ascending_after_knn_and_example():
    # gsknn rows -> ascending order equals the oracle's sorted rows; the
    # Python 3 port of the reference example runs and checks itself.
    import torch
    from rnn_amd import example
    dev = torch.device("cuda")
    p = ds.config1(scale=4)
    D, I = ds.empty_heap(p.n, p.k)
    tX, tX2 = torch.from_numpy(p.XA).to(dev), torch.from_numpy(p.XA2).to(dev)
    tD, tI = torch.from_numpy(D.copy()).to(dev), torch.from_numpy(I.copy()).to(dev)
    gsknn.dknn_dev(tX, tX2, torch.from_numpy(p.amap).to(dev), tX, tX2,
                   torch.from_numpy(p.bmap).to(dev), tD, tI)
    gsknn.rows_ascending_dev(tD, tI)
    torch.cuda.synchronize()
    Dr, Ir = D.copy(), I.copy()
    oracle.gsknn(p.XA, p.XA2, p.amap, p.XB, p.XB2, p.bmap, Dr, Ir, canonical=True)
    Dr, Ir = oracle.sort_rows(Dr, Ir)
    np.testing.assert_array_equal(tI.cpu().numpy(), Ir)
    assert np.array_equal(tD.cpu().numpy().view(np.uint8), Dr.view(np.uint8))
    assert example.main([]) == 0
    assert example.main(["--double"]) == 0


@pytest.mark.parametrize("dtype,d", [(np.float64, 16), (np.float32, 32), (np.float64, 64),
                                     (np.float32, 8)])
def test_exact_scalar_candidate_form(dtype, d):
    """Exact mode at >= 2^26 pairs with d in {8, 16, 32, 64} takes the
    scalar-candidate block routine (candidates as SGPR operands): 48 sampled
    rows bitwise equal to the oracle, partial last group and candidate splits
    included."""
    import torch
    print(f"[gpu] exact scalar form {np.dtype(dtype).name} d={d}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(300 + d))
    N = 12000
    X = g.standard_normal((N, d)).astype(dtype)
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, N, 8195).astype(np.int32)
    bmap = g.integers(0, N, 8192).astype(np.int32)
    k = 20
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((8192, k), float("inf"), dtype=tX.dtype, device=dev)
    I = torch.full((8192, k), -1, dtype=torch.int32, device=dev)
    call = gsknn.dknn_dev if dtype == np.float64 else gsknn.sknn_dev
    call(tX, tX2, torch.from_numpy(amap).to(dev), tX, tX2, torch.from_numpy(bmap).to(dev), D, I,
         mode=gsknn.MODE_EXACT)
    torch.cuda.synchronize()
    rows = np.linspace(0, 8191, 48).astype(np.int64)
    Dr, Ir = ds.empty_heap(rows.size, k, dtype)
    oracle.gsknn(X, X2, amap, X, X2, bmap[rows].copy(), Dr, Ir, canonical=True)
    assert_rows(D.cpu().numpy()[rows], I.cpu().numpy()[rows], Dr, Ir)


@pytest.mark.parametrize("k,nq", [(100, 300), (12, 5000), (300, 40)])
def test_split_merge_paths_exact(k, nq):
    """Exact mode with candidate splits and the two merge kernels: one wave
    per query when the lists fit (k (1 + splits) <= 1024 entries), one thread
    per query otherwise; rows pre-filled by a first call so the resident
    maximum filters the lists."""
    import torch
    print(f"[gpu] split merge k={k} nq={nq}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(k + nq))
    X = g.standard_normal((9000, 16))
    X2 = oracle.sqnorm(X)
    first = g.integers(0, 9000, 700).astype(np.int32)
    amap = g.integers(0, 9000, 8001).astype(np.int32)
    bmap = g.integers(0, 9000, nq).astype(np.int32)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((nq, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((nq, k), -1, dtype=torch.int32, device=dev)
    tb = torch.from_numpy(bmap).to(dev)
    for am in (first, amap):
        gsknn.dknn_dev(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2, tb, D, I,
                       mode=gsknn.MODE_EXACT)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(nq, k)
    oracle.gsknn(X, X2, first, X, X2, bmap, Dr, Ir, canonical=True)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(D.cpu().numpy(), I.cpu().numpy(), Dr, Ir)


@pytest.mark.parametrize("d,k", [(64, 500), (40, 420), (100, 700)])
def test_mfma_large_k_single_wave_shape(d, k):
    """FP64 MFMA, LDS-tile kernel, with heaps so large that only one wave of
    16 queries fits a workgroup's LDS (fewer waves than candidate blocks per
    tile)."""
    import torch
    print(f"[gpu] mfma large k d={d} k={k}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(d * 1000 + k))
    X = g.standard_normal((5000, d))
    X2 = oracle.sqnorm(X)
    amap = g.integers(0, 5000, 3001).astype(np.int32)
    bmap = g.integers(0, 5000, 200).astype(np.int32)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((200, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((200, k), -1, dtype=torch.int32, device=dev)
    gsknn.dknn_dev(tX, tX2, torch.from_numpy(amap).to(dev), tX, tX2,
                   torch.from_numpy(bmap).to(dev), D, I, mode=gsknn.MODE_MFMA_TILE)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(200, k)
    oracle.gsknn(X, X2, amap, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(D.cpu().numpy(), I.cpu().numpy(), Dr, Ir)


def _mfma_case(X, amaps, bmap, k, mode):
    import torch
    dev = torch.device("cuda")
    X2 = oracle.sqnorm(X)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    nq = bmap.size
    D = torch.full((nq, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((nq, k), -1, dtype=torch.int32, device=dev)
    tb = torch.from_numpy(bmap).to(dev)
    for am in amaps:
        gsknn.dknn_dev(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2, tb, D, I, mode=mode)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(nq, k)
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    return D.cpu().numpy(), I.cpu().numpy(), Dr, Ir, X2


@pytest.mark.parametrize("form", ["direct", "tile"])
@pytest.mark.parametrize("d,k,nq", [(64, 32, 1000), (64, 1, 700), (63, 200, 500), (50, 400, 300),
                                    (33, 17, 37), (32, 20, 800), (17, 32, 900), (16, 16, 600),
                                    (5, 9, 400), (3, 8, 300), (64, 64, 20), (128, 32, 700),
                                    (100, 24, 500), (65, 40, 60)])
def test_mfma_direct_and_tile_forms(form, d, k, nq):
    """FP64 MFMA in both forms: the direct form (each wave streams candidate
    operands into registers) and the LDS-tile kernel.  Partial last tile (3001
    candidates), pre-filled rows, candidate splits for few queries; the
    direct form's line deal (d = 16, 32, 64, and 128 with one query block and
    one candidate block per wave), its contiguous runs moved back to stay
    inside the row (d = 5, 17, 33, 50, 63, 65, 100), and d = 3, below its run
    length (falls back to the tile kernel)."""
    print(f"[gpu] mfma form={form} d={d} k={k} nq={nq}", flush=True)
    g = np.random.Generator(np.random.PCG64(d * 7919 + k * 31 + nq))
    X = g.standard_normal((6000, d))
    first = g.integers(0, 6000, 500).astype(np.int32)
    amap = g.integers(0, 6000, 3001).astype(np.int32)
    bmap = g.integers(0, 6000, nq).astype(np.int32)
    mode = gsknn.MODE_MFMA if form == "direct" else gsknn.MODE_MFMA_TILE
    Dg, Ig, Dr, Ir, X2 = _mfma_case(X, (first, amap), bmap, k, mode)
    assert_rows(Dg, Ig, Dr, Ir)


@pytest.mark.parametrize("form", ["direct", "tile"])
def test_mfma_forms_lattice_ties(form):
    """Integer lattice at d = 40: every MFMA dot product is exact, so rows
    hold exact ties and the lower index must win, bit-identical distances."""
    print(f"[gpu] mfma lattice form={form}", flush=True)
    g = np.random.Generator(np.random.PCG64(4040))
    X = g.integers(-3, 4, (5000, 40)).astype(np.float64)
    amap = g.integers(0, 5000, 4000).astype(np.int32)
    bmap = g.integers(0, 5000, 600).astype(np.int32)
    mode = gsknn.MODE_MFMA if form == "direct" else gsknn.MODE_MFMA_TILE
    Dg, Ig, Dr, Ir, _ = _mfma_case(X, (amap,), bmap, 24, mode)
    assert_rows(Dg, Ig, Dr, Ir)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d,m,k,nq", [(16, 700, 16, 300), (7, 2048, 64, 100), (64, 1500, 33, 50),
                                      (33, 130, 5, 900), (1, 1999, 40, 64), (24, 900, 16, 200),
                                      (32, 2000, 64, 60)])
def test_small_m_one_wave_per_query(dtype, d, m, k, nq):
    """Exact calls too small to fill the chip (few queries, m <= 2048,
    k <= 64): one wave per query, candidates staged through LDS, k rounds of
    a wave arg-min.  Bit-identical to the oracle on carried-over rows (a first
    call, then a second over other candidates) written through a heap-row
    map; lane slots left partly empty (m not a multiple of 64); d from 1 to
    64 (chunks of 512 rows for d <= 16, 256 for d 17..32 at d = 24 and 32,
    128 above)."""
    import torch
    print(f"[gpu] small-m {np.dtype(dtype).name} d={d} m={m} k={k} nq={nq}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(d * 131 + m + k))
    X = g.standard_normal((4000, d)).astype(dtype)
    X2 = oracle.sqnorm(X).astype(dtype)
    first = g.integers(0, 4000, 300).astype(np.int32)
    amap = g.integers(0, 4000, m).astype(np.int32)
    bmap = g.integers(0, 4000, nq).astype(np.int32)
    hrow = g.permutation(nq + 17)[:nq].astype(np.int32)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((nq + 17, k), float("inf"), dtype=torch.from_numpy(X).dtype, device=dev)
    I = torch.full((nq + 17, k), -1, dtype=torch.int32, device=dev)
    th = torch.from_numpy(hrow).to(dev)
    run = gsknn.dknn_dev if dtype == np.float64 else gsknn.sknn_dev
    for am in (first, amap):
        run(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2, torch.from_numpy(bmap).to(dev), D, I,
            hrow=th, mode=gsknn.MODE_EXACT)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(nq, k, dtype)
    for am in (first, amap):
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(D.cpu().numpy()[hrow], I.cpu().numpy()[hrow], Dr, Ir)
    untouched = np.setdiff1d(np.arange(nq + 17), hrow)
    assert np.all(I.cpu().numpy()[untouched] == -1)


def test_small_m_lattice_ties():
    """Integer lattice (exact ties) through the one-wave-per-query kernel:
    the lower index wins, bit-identical distances, duplicates kept."""
    print("[gpu] small-m lattice", flush=True)
    pts = np.array([(x, y, z) for x in range(9) for y in range(9) for z in range(9)],
                   dtype=np.float64)
    X2 = oracle.sqnorm(pts)
    g = np.random.Generator(np.random.PCG64(77))
    amap = np.concatenate([g.permutation(len(pts))[:900], g.integers(0, len(pts), 300)]).astype(np.int32)
    bmap = np.arange(0, len(pts), 3, dtype=np.int32)
    D, I = ds.empty_heap(bmap.size, 24)
    Dr, Ir = D.copy(), I.copy()
    gsknn.dgsknn_heap(pts, X2, amap, pts, X2, bmap, D, I)
    oracle.gsknn(pts, X2, amap, pts, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(D, I, Dr, Ir)


@pytest.mark.parametrize("form", ["direct", "tile"])
@pytest.mark.parametrize("d,nq", [(64, 300), (48, 2500)])
def test_mfma_forms_heap_row_map(form, d, nq):
    """FP64 MFMA in both forms writing through a heap-row map into a larger
    pool (rows outside the map untouched), with a pre-filling call, at a
    size that splits the candidates (300 queries) and one that does not."""
    import torch
    print(f"[gpu] mfma hrow form={form} d={d} nq={nq}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(d * 17 + nq))
    X = g.standard_normal((8000, d))
    X2 = oracle.sqnorm(X)
    first = g.integers(0, 8000, 700).astype(np.int32)
    amap = g.integers(0, 8000, 6001).astype(np.int32)
    bmap = g.integers(0, 8000, nq).astype(np.int32)
    pool = nq + 123
    hrow = g.permutation(pool)[:nq].astype(np.int32)
    k = 20
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((pool, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((pool, k), -1, dtype=torch.int32, device=dev)
    mode = gsknn.MODE_MFMA if form == "direct" else gsknn.MODE_MFMA_TILE
    for am in (first, amap):
        gsknn.dknn_dev(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2,
                       torch.from_numpy(bmap).to(dev), D, I,
                       hrow=torch.from_numpy(hrow).to(dev), mode=mode)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(nq, k)
    for am in (first, amap):
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    Dg, Ig = D.cpu().numpy(), I.cpu().numpy()
    assert_rows(Dg[hrow], Ig[hrow], Dr, Ir)
    untouched = np.setdiff1d(np.arange(pool), hrow)
    assert np.all(Ig[untouched] == -1)


@pytest.mark.parametrize("form", ["direct", "tile"])
@pytest.mark.parametrize("m,k", [(5, 12), (31, 31), (33, 40)])
def test_mfma_forms_tiny_candidate_sets(form, m, k):
    """Fewer candidates than one tile and than k: partly empty tile slots
    never enter a row, and rows keep (inf, -1) placeholders beyond m."""
    print(f"[gpu] mfma tiny form={form} m={m} k={k}", flush=True)
    g = np.random.Generator(np.random.PCG64(m * 100 + k))
    X = g.standard_normal((500, 64))
    amap = g.permutation(500)[:m].astype(np.int32)
    bmap = g.integers(0, 500, 70).astype(np.int32)
    mode = gsknn.MODE_MFMA if form == "direct" else gsknn.MODE_MFMA_TILE
    Dg, Ig, Dr, Ir, X2 = _mfma_case(X, (amap,), bmap, k, mode)
    assert_rows(Dg, Ig, Dr, Ir)
    assert np.all((Ig >= 0).sum(axis=1) == min(m, k))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("k,spread", [(8, 3), (16, 2), (40, 3), (16, 0)])
def test_small_m_selection_ties(dtype, k, spread):
    """One-wave-per-query kernel on tie-heavy inputs: coordinates drawn from
    {-spread .. spread} (spread 0: every point the same, so every distance
    ties and the selection takes the lowest indices), repeated candidate
    ids, two calls into carried-over rows.  The sort-based selection and its
    fallback (more than 64 entries at or below the threshold) both meet the
    canonical (distance, index) order bit for bit."""
    import torch
    print(f"[gpu] small-m ties {np.dtype(dtype).name} k={k} spread={spread}", flush=True)
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(500 + k + spread))
    X = g.integers(-spread, spread + 1, (3000, 6)).astype(dtype)
    X2 = oracle.sqnorm(X).astype(dtype)
    amaps = [g.integers(0, 3000, 700).astype(np.int32), g.integers(0, 3000, 1500).astype(np.int32)]
    bmap = g.integers(0, 3000, 150).astype(np.int32)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D = torch.full((bmap.size, k), float("inf"), dtype=torch.from_numpy(X).dtype, device=dev)
    I = torch.full((bmap.size, k), -1, dtype=torch.int32, device=dev)
    run = gsknn.dknn_dev if dtype == np.float64 else gsknn.sknn_dev
    for am in amaps:
        run(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2, torch.from_numpy(bmap).to(dev), D, I,
            mode=gsknn.MODE_EXACT)
    torch.cuda.synchronize()
    Dr, Ir = ds.empty_heap(bmap.size, k, dtype)
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(D.cpu().numpy(), I.cpu().numpy(), Dr, Ir)


@pytest.mark.parametrize("warm", [0, 3, 1000])
@pytest.mark.parametrize("k", [32, 100])
def test_sgsknn_register_form_warm_up_and_carried_rows(warm, k, monkeypatch):
    """The FP32 register-form filter (d <= 256, k <= 128) with warm-up passes
    of 0 tiles, a few, and more than a split holds, on rows carried over from
    a first call (resident distances seed the upper-bound sets): bitwise the
    oracle's FP32 rows."""
    import torch
    monkeypatch.setenv("GSKNN_F32_WARM", str(warm))
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(4000 + warm + k))
    N, d = 14000, 96
    X = ds.low_rank(N, d, 12, seed=21).astype(np.float32)
    X2 = oracle.sqnorm(X)
    amaps = [g.integers(0, N, 900).astype(np.int32), g.integers(0, N, 7000).astype(np.int32)]
    bmap = g.permutation(N)[:5000].astype(np.int32)
    print(f"[gpu] sgsknn register form warm={warm} k={k}", flush=True)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    D, I = ds.empty_heap(bmap.size, k, np.float32)
    tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
    for am in amaps:
        gsknn.sknn_dev(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2,
                       torch.from_numpy(bmap).to(dev), tD, tI)
    torch.cuda.synchronize()
    Dr, Ir = D.copy(), I.copy()
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(tD.cpu().numpy(), tI.cpu().numpy(), Dr, Ir)


@pytest.mark.parametrize("d,k", [(96, 1), (96, 64), (128, 100), (40, 200), (40, 256)])
def test_sgsknn_select_and_heap_rerank_agree(d, k, monkeypatch):
    """The FP32 re-rank's selection form (k <= 256: a pool of recomputed
    candidates, rank selection) and its heap form (GSKNN_F32_SELECT=0) on
    rows carried over from a first call whose candidates repeat in the
    second (identical (distance, index) pairs in the row and the lists):
    the same bytes in the same order, and bitwise the oracle's FP32 rows."""
    import torch
    dev = torch.device("cuda")
    g = np.random.Generator(np.random.PCG64(5100 + d + k))
    N = 12000
    X = ds.low_rank(N, d, 10, seed=23).astype(np.float32)
    X2 = oracle.sqnorm(X)
    a0 = g.integers(0, N, 3000).astype(np.int32)
    amaps = [a0, np.concatenate([a0[:1500], g.integers(0, N, 6000).astype(np.int32)])]
    bmap = g.permutation(N)[:4000].astype(np.int32)
    tX, tX2 = torch.from_numpy(X).to(dev), torch.from_numpy(X2).to(dev)
    rows = {}
    for sel in ("1", "0"):
        monkeypatch.setenv("GSKNN_F32_SELECT", sel)
        print(f"[gpu] sgsknn re-rank select={sel} d={d} k={k}", flush=True)
        D, I = ds.empty_heap(bmap.size, k, np.float32)
        tD, tI = torch.from_numpy(D).to(dev), torch.from_numpy(I).to(dev)
        for am in amaps:
            gsknn.sknn_dev(tX, tX2, torch.from_numpy(am).to(dev), tX, tX2,
                           torch.from_numpy(bmap).to(dev), tD, tI)
        torch.cuda.synchronize()
        rows[sel] = (tD.cpu().numpy(), tI.cpu().numpy())
    assert np.array_equal(rows["1"][0].view(np.uint32), rows["0"][0].view(np.uint32))
    np.testing.assert_array_equal(rows["1"][1], rows["0"][1])
    Dr, Ir = ds.empty_heap(bmap.size, k, np.float32)
    for am in amaps:
        oracle.gsknn(X, X2, am, X, X2, bmap, Dr, Ir, canonical=True)
    assert_rows(rows["1"][0], rows["1"][1], Dr, Ir)
