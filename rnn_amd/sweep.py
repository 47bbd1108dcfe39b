"""Randomized-projection-tree all-nearest-neighbour sweep on one GPU.

Each tree iteration splits the point set recursively: a node draws a stop
threshold log-uniform in [leaf_lo, leaf_hi]; a node larger than its threshold
projects its points on a seeded random unit direction and splits at the exact
median (ties by point id), so leaf sizes are uneven.  Every leaf then runs
gsknn with amap = bmap = the leaf's point ids, merging into global heap rows
(row = point id) -- the composition of per-leaf reference gsknn calls
(frame/gsknn.c:1138-1184).  Heaps carry over between iterations; a neighbour
found again in a later iteration is kept twice, as the reference's heap would.

On the device each tree level is one fused kernel of libgsknn.so
(gsknn_tree_level_keys_dev: node lookup, projection and packed sort key per
point) and one sort of the keys.

The leaves of one iteration run in ONE persistent launch of the leaf-sweep
kernel (libgsknn.so: gsknn_sweep_leaves_[sd]dev) over a device work list;
that launch is captured once in a hipGraph and replayed for every iteration
after the tree build rewrites the work list in place.  The launch shape
(queries per workgroup, candidates per LDS tile, persistent grid) comes from
``autotune``, which times the candidates on the device and caches the winner.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import gsknn

QPB_CHOICES = (64, 32)
CT_CHOICES = (16, 32, 64)
GRID_CHOICES = (512, 1024, 2048, 4096, 8192, 16384)
# Dimensions with the scalar-candidate block routine: no LDS candidate tile,
# so the tile size is not a tuning knob there.
SCALAR_DIMS = (8, 16, 32, 64)


def _orderable(f):
    """float32 tensor -> int64 keys with the same order."""
    import torch
    b = f.contiguous().view(torch.int32)
    b = torch.where(b < 0, b ^ 0x7FFFFFFF, b)
    return b.to(torch.int64) + (1 << 31)


def build_tree(X, iteration, seed, leaf_lo=256, leaf_hi=4096):
    """One tree over the rows of device tensor X (N, d), float32.

    Returns (perm, leaves): perm an int32 device tensor of point ids grouped
    by leaf, leaves a list of (start, size) in perm order.  Node bookkeeping
    is vectorized (numpy on the host, one pass of device work per level).
    """
    import torch
    dev = X.device
    N, d = X.shape
    g = np.random.Generator(np.random.PCG64([seed, iteration]))
    perm = torch.arange(N, device=dev, dtype=torch.int64)
    n_start = np.zeros(1, dtype=np.int64)
    n_size = np.full(1, N, dtype=np.int64)
    leaf_parts = []
    while n_start.size:
        thr = np.exp(g.uniform(np.log(leaf_lo), np.log(leaf_hi), size=n_start.size))
        split = (n_size > thr) & (n_size >= 2)
        leaf_parts.append(np.stack([n_start[~split], n_size[~split]], axis=1))
        if not split.any():
            break
        starts, sizes = n_start[split], n_size[split]
        dirs = g.standard_normal((starts.size, d))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        dirs_t = torch.from_numpy(dirs.astype(np.float32)).to(dev)
        starts_t = torch.from_numpy(starts).to(dev)
        total = int(sizes.sum())
        # order inside each node by (projection, point id): one sort of the
        # packed key (node, projection, id) when it fits in 63 bits, else a
        # sort by id followed by a stable sort by (node, projection)
        id_bits = max(1, int(N - 1).bit_length())
        node_bits = max(1, int(starts.size - 1).bit_length())
        fits = node_bits + 32 + id_bits <= 63
        if fits and X.is_cuda:
            # positions and keys from one fused kernel (libgsknn.so)
            offs_t = torch.from_numpy(np.cumsum(sizes) - sizes).to(dev)
            pos = torch.empty(total, dtype=torch.int64, device=dev)
            key = torch.empty(total, dtype=torch.int64, device=dev)
            gsknn.tree_level_keys_dev(X, perm, starts_t, offs_t, dirs_t, total, id_bits, pos, key)
            perm[pos] = torch.sort(key).values & ((1 << id_bits) - 1)
            half = sizes // 2
            n_start = np.stack([starts, starts + half], axis=1).reshape(-1)
            n_size = np.stack([half, sizes - half], axis=1).reshape(-1)
            continue
        sizes_t = torch.from_numpy(sizes).to(dev)
        node = torch.repeat_interleave(torch.arange(starts.size, device=dev), sizes_t,
                                       output_size=total)
        offs = torch.cumsum(sizes_t, 0) - sizes_t
        pos = starts_t[node] + (torch.arange(node.numel(), device=dev) - offs[node])
        pts = perm[pos]
        proj = (X[pts] * dirs_t[node]).sum(dim=1)
        if fits:
            key = (node << (32 + id_bits)) | (_orderable(proj) << id_bits) | pts
            perm[pos] = pts[torch.argsort(key)]
        else:
            o1 = torch.argsort(node * N + pts)
            node1, pts1, proj1 = node[o1], pts[o1], proj[o1]
            o2 = torch.sort(node1 * (1 << 33) + _orderable(proj1), stable=True).indices
            perm[pos] = pts1[o2]
        half = sizes // 2
        n_start = np.stack([starts, starts + half], axis=1).reshape(-1)
        n_size = np.stack([half, sizes - half], axis=1).reshape(-1)
    leaves = np.concatenate(leaf_parts) if leaf_parts else np.zeros((0, 2), np.int64)
    leaves = leaves[np.argsort(leaves[:, 0], kind="stable")]
    return perm.to(torch.int32), [(int(s_), int(n_)) for s_, n_ in leaves]


def make_items(leaves, qpb):
    """Work list [nitems][4] = (start, size, first query, 0), largest leaves
    first (ties in leaf order, query blocks ascending)."""
    lv = np.asarray(leaves, dtype=np.int64).reshape(-1, 2)
    nblk = (lv[:, 1] + qpb - 1) // qpb
    leaf = np.repeat(np.arange(lv.shape[0]), nblk)
    first = np.cumsum(nblk) - nblk
    q0 = (np.arange(leaf.size) - first[leaf]) * qpb
    items = np.stack([lv[leaf, 0], lv[leaf, 1], q0, np.zeros_like(q0)], axis=1)
    items = items[np.argsort(-items[:, 1], kind="stable")]
    return items.astype(np.int32).reshape(-1, 4)


def _cache_path():
    return os.environ.get("GSKNN_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gsknn_tune.json"))


def _launch(dtype, d, k, X, X2, perm, items, nitems, D, I, qpb, ct, grid, stream):
    L = gsknn.lib()
    fn = L.gsknn_sweep_leaves_ddev if dtype == "f64" else L.gsknn_sweep_leaves_sdev
    rc = fn(d, k, gsknn._dp(X), gsknn._dp(X2), gsknn._dp(perm), gsknn._dp(items),
            gsknn._dp(nitems), gsknn._dp(D), gsknn._dp(I), D.stride(0), qpb, ct, grid,
            gsknn._stream_ptr(stream))
    if rc:
        raise gsknn.GsknnError(gsknn.last_error())


def autotune(X, X2, k, perm, leaves, dtype="f32", cache=True, log=None):
    """Pick (qpb, ct, grid) for the leaf sweep by timing each candidate on one
    iteration's work list against scratch heaps.  Cached per shape."""
    import torch
    dev = X.device
    N, d = X.shape
    sizes = np.array([n for _, n in leaves])
    key = f"{dtype}-d{d}-k{k}-N{N}-leaves{int(sizes.min())}-{int(sizes.max())}"
    path = _cache_path()
    table = {}
    if cache and os.path.exists(path):
        try:
            table = json.load(open(path))
        except (OSError, ValueError):
            table = {}
        if key in table:
            return tuple(table[key]["config"]), table[key]
    tdt = torch.float64 if dtype == "f64" else torch.float32
    D = torch.full((N, k), float("inf"), dtype=tdt, device=dev)
    I = torch.full((N, k), -1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    elem = 8 if dtype == "f64" else 4
    results = []
    for qpb in QPB_CHOICES:
        items = torch.from_numpy(make_items(leaves, qpb)).to(dev)
        nitems = torch.tensor([items.shape[0]], dtype=torch.int32, device=dev)
        for ct in (CT_CHOICES[:1] if d in SCALAR_DIMS else CT_CHOICES):
            if int(gsknn.lib().gsknn_sweep_lds_bytes(elem, d, k, qpb, ct)) > 160 * 1024:
                continue
            for grid in GRID_CHOICES:
                D.fill_(float("inf"))
                I.fill_(-1)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                _launch(dtype, d, k, X, X2, perm, items, nitems, D, I, qpb, ct, grid, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1)
                results.append((ms, qpb, ct, grid))
                if log:
                    log(f"autotune qpb={qpb} ct={ct} grid={grid}: {ms:.3f} ms")
    results.sort()
    best = results[0]
    entry = {"config": [best[1], best[2], best[3]], "ms": best[0],
             "tried": [[r[1], r[2], r[3], r[0]] for r in results]}
    if cache:
        table[key] = entry
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            json.dump(table, open(path, "w"), indent=1)
        except OSError:
            pass
    return (best[1], best[2], best[3]), entry


class RPTreeSweep:
    """All-NN sweep state: points, global heap rows, the captured launch."""

    def __init__(self, X, k, seed=3, leaf_lo=256, leaf_hi=4096, dtype="f32", config=None,
                 use_graph=True, log=None):
        import torch
        self.X = X
        self.k = k
        self.seed = seed
        self.leaf_lo, self.leaf_hi = leaf_lo, leaf_hi
        self.dtype = dtype
        self.log = log
        dev = X.device
        N, d = X.shape
        tdt = torch.float64 if dtype == "f64" else torch.float32
        self.X2 = torch.empty(N, dtype=tdt, device=dev)
        gsknn.sqnorm_dev(X, self.X2)
        self.D = torch.full((N, k), float("inf"), dtype=tdt, device=dev)
        self.I = torch.full((N, k), -1, dtype=torch.int32, device=dev)
        self.perm = torch.empty(N, dtype=torch.int32, device=dev)
        self.config = config
        self.use_graph = use_graph
        self.graph = None
        self.tune_info = None
        self.iteration = 0
        self.timings = []

    def _ensure_config(self, perm, leaves):
        if self.config is None:
            self.config, self.tune_info = autotune(self.X, self.X2, self.k, perm, leaves,
                                                   self.dtype, log=self.log)
        qpb = self.config[0]
        import torch
        if not hasattr(self, "items"):
            N = self.X.shape[0]
            max_items = N // qpb + N // max(1, min(self.leaf_lo, qpb)) + 16
            self.items = torch.zeros((max_items, 4), dtype=torch.int32, device=self.X.device)
            self.nitems = torch.zeros(1, dtype=torch.int32, device=self.X.device)

    def _launch_captured(self):
        import torch
        qpb, ct, grid = self.config
        stream = torch.cuda.current_stream()
        _launch(self.dtype, self.X.shape[1], self.k, self.X, self.X2, self.perm, self.items,
                self.nitems, self.D, self.I, qpb, ct, grid, stream)

    def step(self):
        """One tree iteration: build the tree, rewrite the work list, launch."""
        import torch
        t0 = time.perf_counter()
        perm, leaves = build_tree(self.X, self.iteration, self.seed, self.leaf_lo, self.leaf_hi)
        self._ensure_config(perm, leaves)
        qpb = self.config[0]
        items = make_items(leaves, qpb)
        self.perm.copy_(perm)
        self.items[: items.shape[0]].copy_(torch.from_numpy(items))
        self.nitems.fill_(items.shape[0])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if self.use_graph and self.graph is not None:
            e0.record(stream)
            self.graph.replay()
            e1.record(stream)
        else:
            # First iteration (or no graph): launch directly; the direct
            # launch also sets the kernel's attributes before any capture.
            e0.record(stream)
            self._launch_captured()
            e1.record(stream)
        torch.cuda.synchronize()
        sweep_ms = e0.elapsed_time(e1)
        if self.use_graph and self.graph is None:
            # Capture the launch once; later iterations rewrite perm/items/
            # nitems in place and replay it.  Capturing does not execute.
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._launch_captured()
            torch.cuda.synchronize()
        pairs = float(sum(n * n for _, n in leaves))
        info = {"iteration": self.iteration, "tree_ms": (t1 - t0) * 1e3, "sweep_ms": sweep_ms,
                "leaves": len(leaves), "pairs": pairs,
                "leaf_sizes": sorted({n for _, n in leaves})}
        self.timings.append(info)
        self.iteration += 1
        return perm, leaves, info
