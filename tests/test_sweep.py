"""RP-tree all-NN sweep: tree construction (CPU) and GPU parity with the
composition of per-leaf oracle calls."""
import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import sweep


def _cpu_points(N=4096, d=8, seed=3):
    import torch
    return torch.from_numpy(ds.config4_points(N, d, 16, seed))


def test_tree_partitions_points_into_bounded_leaves():
    X = _cpu_points()
    perm, leaves = sweep.build_tree(X, 0, seed=3, leaf_lo=64, leaf_hi=512)
    p = perm.numpy()
    assert np.array_equal(np.sort(p), np.arange(X.shape[0]))
    starts = [s for s, _ in leaves]
    sizes = [n for _, n in leaves]
    assert starts[0] == 0 and sum(sizes) == X.shape[0]
    assert all(s1 + n1 == s2 for (s1, n1), (s2, _) in zip(leaves, leaves[1:]))
    assert max(sizes) <= 512 and min(sizes) >= 32
    assert len(set(sizes)) > 1  # uneven leaves


def test_tree_is_seeded_and_differs_across_iterations():
    X = _cpu_points()
    p0, l0 = sweep.build_tree(X, 0, seed=3, leaf_lo=64, leaf_hi=512)
    p0b, l0b = sweep.build_tree(X, 0, seed=3, leaf_lo=64, leaf_hi=512)
    p1, _ = sweep.build_tree(X, 1, seed=3, leaf_lo=64, leaf_hi=512)
    assert np.array_equal(p0.numpy(), p0b.numpy()) and l0 == l0b
    assert not np.array_equal(p0.numpy(), p1.numpy())


def test_median_split_of_root():
    # With one split forced (leaf_hi below N, leaf_lo above N/2), the two
    # leaves are the lower and upper halves of the projection.
    import torch
    g = np.random.Generator(np.random.PCG64(1))
    X = torch.from_numpy(g.standard_normal((200, 3)).astype(np.float32))
    perm, leaves = sweep.build_tree(X, 0, seed=9, leaf_lo=150, leaf_hi=199)
    assert leaves == [(0, 100), (100, 100)]
    gg = np.random.Generator(np.random.PCG64([9, 0]))
    gg.uniform(np.log(150), np.log(199), size=1)
    direc = gg.standard_normal((1, 3))
    direc /= np.linalg.norm(direc)
    proj = (X.numpy() * direc.astype(np.float32)).sum(1)
    left = set(perm.numpy()[:100])
    assert max(proj[list(left)]) <= min(proj[list(set(range(200)) - left)])


def test_items_cover_each_leaf_once():
    leaves = [(0, 100), (100, 37), (137, 64)]
    it = sweep.make_items(leaves, 32)
    assert it[0][1] == 100  # largest first
    cover = {}
    for s, n, q0, _ in it:
        cover.setdefault((s, n), []).append(q0)
    assert sorted(cover[(0, 100)]) == [0, 32, 64, 96]
    assert sorted(cover[(100, 37)]) == [0, 32]


@pytest.mark.gpu
def test_sweep_matches_per_leaf_oracle(tmp_path, monkeypatch):
    # Iteration 0 runs the exact form directly on the empty rows; iteration 1
    # the MFMA form directly (then captured), iterations 2-4 replay it; the
    # fifth iteration's tree comes from the second batch of four, built on
    # the side stream.  The heap rows are checked after every iteration
    # against the composition of per-leaf oracle calls.
    import torch
    monkeypatch.setenv("GSKNN_TUNE_CACHE", str(tmp_path / "tune.json"))
    dev = torch.device("cuda")
    N, d, k = 1 << 14, 32, 16
    Xn = ds.config4_points(N, d, 64, seed=3)
    X = torch.from_numpy(Xn).to(dev)
    sw = sweep.RPTreeSweep(X, k, seed=3, leaf_lo=64, leaf_hi=1024)
    X2 = oracle.sqnorm(Xn)
    assert np.array_equal(sw.X2.cpu().numpy(), X2)
    Dr = np.full((N, k), np.inf, dtype=np.float32)
    Ir = np.full((N, k), -1, dtype=np.int32)
    forms = []
    for _ in range(5):
        perm, leaves, info = sw.step()
        forms.append((info["form"], info["graph"]))
        print(f"[gpu] sweep iteration {info['iteration']}: {info['sweep_ms']:.3f} ms", flush=True)
        perm = perm.cpu().numpy()
        for s, n in leaves:
            ids = np.ascontiguousarray(perm[s:s + n])
            Ds, Is = np.ascontiguousarray(Dr[ids]), np.ascontiguousarray(Ir[ids])
            oracle.gsknn(Xn, X2, ids, Xn, X2, ids, Ds, Is, canonical=True)
            Dr[ids], Ir[ids] = Ds, Is
        a = oracle.sort_rows(sw.D.cpu().numpy(), sw.I.cpu().numpy())
        b = oracle.sort_rows(Dr, Ir)
        np.testing.assert_array_equal(a[1], b[1])
        assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    assert sw.graph is not None
    assert forms == [("exact", False), ("mfma", False), ("mfma", True), ("mfma", True),
                     ("mfma", True)]


@pytest.mark.gpu
def test_batched_trees_equal_single_trees():
    # Several trees sorted together per level (one key space numbering the
    # nodes of all trees) give each tree's single build, permutation and
    # leaves alike.
    import torch
    dev = torch.device("cuda")
    X = torch.from_numpy(ds.config4_points(20000, 16, 32, seed=5)).to(dev)
    print("[gpu] batched trees N=20000 d=16", flush=True)
    perms, leaves = sweep.build_trees(X, [3, 4, 5, 6, 7], seed=11, leaf_lo=100, leaf_hi=900)
    for t, it in enumerate([3, 4, 5, 6, 7]):
        p1, l1 = sweep.build_tree(X, it, seed=11, leaf_lo=100, leaf_hi=900)
        assert l1 == leaves[t]
        assert torch.equal(p1, perms[t])


@pytest.mark.gpu
@pytest.mark.parametrize("pretune", [False, True])
def test_sweep_config4_parameters(tmp_path, monkeypatch, pretune):
    """The config-4 sweep parameters (k = 32, leaves of 256..4096 points,
    d = 32, 256-cluster mixture) at N = 2^16: exact median splits of a power
    of two give leaves of 256, 512, 1024 and 2048 points (a 4096-point node
    always lies above its stop threshold, drawn below 4096).  Iterations (the
    exact form, then the MFMA form launched directly and graph-replayed; with
    pretune both forms are tuned and the MFMA graph captured before the
    first step, so every MFMA iteration replays), each equal bitwise to the
    composition of per-leaf oracle calls."""
    import torch
    monkeypatch.setenv("GSKNN_TUNE_CACHE", str(tmp_path / "tune.json"))
    dev = torch.device("cuda")
    N, d, k = 1 << 16, 32, 32
    Xn = ds.config4_points(N, d, 256, seed=3)
    X = torch.from_numpy(Xn).to(dev)
    sw = sweep.RPTreeSweep(X, k, seed=3, leaf_lo=256, leaf_hi=4096)
    if pretune:
        sw.tune()
        assert sw.config is not None and sw.config_mfma is not None and sw.graph is not None
    X2 = oracle.sqnorm(Xn)
    Dr = np.full((N, k), np.inf, dtype=np.float32)
    Ir = np.full((N, k), -1, dtype=np.int32)
    biggest = 0
    for _ in range(3 if pretune else 2):
        perm, leaves, info = sw.step()
        if pretune:
            assert info["graph"] == (info["iteration"] > 0)
        print(f"[gpu] config-4 sweep iteration {info['iteration']}: {len(leaves)} leaves", flush=True)
        sizes = [n for _, n in leaves]
        assert min(sizes) >= 128 and max(sizes) <= 4096
        biggest = max(biggest, max(sizes))
        perm = perm.cpu().numpy()
        for s, n in leaves:
            ids = np.ascontiguousarray(perm[s:s + n])
            Ds, Is = np.ascontiguousarray(Dr[ids]), np.ascontiguousarray(Ir[ids])
            oracle.gsknn(Xn, X2, ids, Xn, X2, ids, Ds, Is, canonical=True)
            Dr[ids], Ir[ids] = Ds, Is
        a = oracle.sort_rows(sw.D.cpu().numpy(), sw.I.cpu().numpy())
        b = oracle.sort_rows(Dr, Ir)
        np.testing.assert_array_equal(a[1], b[1])
        assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    assert biggest == 2048
    assert sw.graph is not None


@pytest.mark.gpu
def test_device_tree_partition_and_median():
    # The device build (fused per-level key kernel) partitions the points
    # into bounded leaves, and its root split is the median of the
    # projection summed in p order in FP32, ties by point id.
    import torch
    dev = torch.device("cuda")
    X = _cpu_points(N=6000, d=8)
    print("[gpu] device tree N=6000 d=8", flush=True)
    perm, leaves = sweep.build_tree(X.to(dev), 0, seed=3, leaf_lo=64, leaf_hi=512)
    p = perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(X.shape[0]))
    sizes = [n for _, n in leaves]
    assert sum(sizes) == X.shape[0] and max(sizes) <= 512 and min(sizes) >= 32
    assert all(s1 + n1 == s2 for (s1, n1), (s2, _) in zip(leaves, leaves[1:]))
    g = np.random.Generator(np.random.PCG64(1))
    Y = g.standard_normal((200, 3)).astype(np.float32)
    Y[50:60] = Y[40]  # exact ties in the projection
    perm, leaves = sweep.build_tree(torch.from_numpy(Y).to(dev), 0, seed=9, leaf_lo=150,
                                    leaf_hi=199)
    assert leaves == [(0, 100), (100, 100)]
    gg = np.random.Generator(np.random.PCG64([9, 0]))
    gg.uniform(np.log(150), np.log(199), size=1)
    direc = gg.standard_normal((1, 3))
    direc = (direc / np.linalg.norm(direc)).astype(np.float32)[0]
    proj = np.zeros(200, dtype=np.float32)
    for q in range(3):
        proj = (proj + Y[:, q] * direc[q]).astype(np.float32)
    order = sorted(range(200), key=lambda i: (proj[i], i))
    assert list(perm.cpu().numpy()[:100]) == sorted(order[:100], key=lambda i: (proj[i], i))


@pytest.mark.gpu
@pytest.mark.parametrize("N,d,k,lo,hi,mfma,exact_first",
                         [(10000, 16, 20, 40, 700, "1", True), (9000, 32, 7, 100, 900, "1", False),
                          (7000, 32, 64, 80, 600, "1", False), (6000, 32, 32, 64, 512, "0", True)])
def test_sweep_odd_leaves_both_forms(tmp_path, monkeypatch, N, d, k, lo, hi, mfma, exact_first):
    """Leaf sizes that are not multiples of the tile or query block (N not a
    power of two), odd k; the FP32 MFMA form (from the empty rows on, with
    exact_first=False) and (GSKNN_SWEEP_MFMA=0) the exact VALU form, three
    iterations each, bitwise against the composition of per-leaf oracle
    calls."""
    import torch
    monkeypatch.setenv("GSKNN_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setenv("GSKNN_SWEEP_MFMA", mfma)
    dev = torch.device("cuda")
    Xn = ds.config4_points(N, d, 40, seed=5)
    X = torch.from_numpy(Xn).to(dev)
    sw = sweep.RPTreeSweep(X, k, seed=11, leaf_lo=lo, leaf_hi=hi, exact_first=exact_first)
    assert (sw.mfma is not None) == (mfma == "1")
    X2 = oracle.sqnorm(Xn)
    Dr = np.full((N, k), np.inf, dtype=np.float32)
    Ir = np.full((N, k), -1, dtype=np.int32)
    for _ in range(3):
        perm, leaves, info = sw.step()
        print(f"[gpu] sweep N={N} d={d} k={k} mfma={mfma} iteration {info['iteration']}", flush=True)
        assert any(n % 32 for _, n in leaves)
        perm = perm.cpu().numpy()
        for s, n in leaves:
            ids = np.ascontiguousarray(perm[s:s + n])
            Ds, Is = np.ascontiguousarray(Dr[ids]), np.ascontiguousarray(Ir[ids])
            oracle.gsknn(Xn, X2, ids, Xn, X2, ids, Ds, Is, canonical=True)
            Dr[ids], Ir[ids] = Ds, Is
        a = oracle.sort_rows(sw.D.cpu().numpy(), sw.I.cpu().numpy())
        b = oracle.sort_rows(Dr, Ir)
        np.testing.assert_array_equal(a[1], b[1])
        assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
