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


def plan_tree(N, d, iteration, seed, leaf_lo=256, leaf_hi=4096):
    """The node structure of one tree, on the host alone.

    Exact median splits make every node's size a function of its parent's
    (halves), so the whole tree follows from N and the seeded draws: per
    level, a stop threshold log-uniform in [leaf_lo, leaf_hi] for each node,
    then a unit direction for each node that splits.  Returns (levels,
    leaves): levels a list of (starts, sizes, dirs) for the nodes split at
    each level (int64, int64, float64 [nodes, d]), leaves a list of (start,
    size) in perm order.
    """
    g = np.random.Generator(np.random.PCG64([seed, iteration]))
    n_start = np.zeros(1, dtype=np.int64)
    n_size = np.full(1, N, dtype=np.int64)
    leaf_parts = []
    levels = []
    while n_start.size:
        thr = np.exp(g.uniform(np.log(leaf_lo), np.log(leaf_hi), size=n_start.size))
        split = (n_size > thr) & (n_size >= 2)
        leaf_parts.append(np.stack([n_start[~split], n_size[~split]], axis=1))
        if not split.any():
            break
        starts, sizes = n_start[split], n_size[split]
        dirs = g.standard_normal((starts.size, d))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        levels.append((starts, sizes, dirs))
        half = sizes // 2
        n_start = np.stack([starts, starts + half], axis=1).reshape(-1)
        n_size = np.stack([half, sizes - half], axis=1).reshape(-1)
    leaves = np.concatenate(leaf_parts) if leaf_parts else np.zeros((0, 2), np.int64)
    leaves = leaves[np.argsort(leaves[:, 0], kind="stable")]
    return levels, [(int(s_), int(n_)) for s_, n_ in leaves]


def build_tree(X, iteration, seed, leaf_lo=256, leaf_hi=4096, perm_out=None, plan=None):
    """One tree over the rows of device tensor X (N, d), float32.

    Returns (perm, leaves): perm an int32 device tensor of point ids grouped
    by leaf (written into perm_out when given), leaves a list of (start,
    size) in perm order.  The node structure comes from plan_tree on the
    host (or `plan`, its result); the per-level node arrays go to the device
    in one copy, and each level is one fused key kernel (libgsknn.so) and one
    sort of the keys, on the current stream with no host synchronisation.
    """
    import torch
    dev = X.device
    N, d = X.shape
    levels, leaves = plan if plan is not None else plan_tree(N, d, iteration, seed, leaf_lo, leaf_hi)
    perm = torch.arange(N, device=dev, dtype=torch.int64)
    id_bits = max(1, int(N - 1).bit_length())
    fused = X.is_cuda and all(max(1, int(st.size - 1).bit_length()) + 32 + id_bits <= 63
                              for st, _, _ in levels)
    if levels and fused:
        # One upload: per level, starts and element offsets (int64), then
        # the directions (float32).
        nodes = np.array([st.size for st, _, _ in levels], dtype=np.int64)
        node_off = np.concatenate([[0], np.cumsum(nodes)])
        ints = np.concatenate([np.concatenate([st, np.cumsum(sz) - sz]) for st, sz, _ in levels])
        dirs = np.concatenate([dr for _, _, dr in levels]).astype(np.float32)
        ints_t = torch.from_numpy(ints).pin_memory().to(dev, non_blocking=True)
        dirs_t = torch.from_numpy(dirs).pin_memory().to(dev, non_blocking=True)
        int_off = 0
        for li, (starts, sizes, _) in enumerate(levels):
            nn = starts.size
            starts_t = ints_t[int_off:int_off + nn]
            offs_t = ints_t[int_off + nn:int_off + 2 * nn]
            int_off += 2 * nn
            total = int(sizes.sum())
            pos = torch.empty(total, dtype=torch.int64, device=dev)
            key = torch.empty(total, dtype=torch.int64, device=dev)
            gsknn.tree_level_keys_dev(X, perm, starts_t, offs_t,
                                      dirs_t[node_off[li]:node_off[li + 1]], total, id_bits,
                                      pos, key)
            perm[pos] = torch.sort(key).values & ((1 << id_bits) - 1)
    else:
        for starts, sizes, dirs in levels:
            _split_level_torch(X, perm, starts, sizes, dirs, id_bits)
    if perm_out is not None:
        perm_out.copy_(perm)
        return perm_out, leaves
    return perm.to(torch.int32), leaves


def _split_level_torch(X, perm, starts, sizes, dirs, id_bits):
    """One level with torch operations (CPU tensors, or node counts too
    large for the packed key): order each split node by (projection, id)."""
    import torch
    dev = X.device
    N = X.shape[0]
    dirs_t = torch.from_numpy(dirs.astype(np.float32)).to(dev)
    starts_t = torch.from_numpy(starts).to(dev)
    total = int(sizes.sum())
    sizes_t = torch.from_numpy(sizes).to(dev)
    node = torch.repeat_interleave(torch.arange(starts.size, device=dev), sizes_t,
                                   output_size=total)
    offs = torch.cumsum(sizes_t, 0) - sizes_t
    pos = starts_t[node] + (torch.arange(node.numel(), device=dev) - offs[node])
    pts = perm[pos]
    proj = (X[pts] * dirs_t[node]).sum(dim=1)
    node_bits = max(1, int(starts.size - 1).bit_length())
    if node_bits + 32 + id_bits <= 63:
        key = (node << (32 + id_bits)) | (_orderable(proj) << id_bits) | pts
        perm[pos] = pts[torch.argsort(key)]
    else:
        o1 = torch.argsort(node * N + pts)
        node1, pts1, proj1 = node[o1], pts[o1], proj[o1]
        o2 = torch.sort(node1 * (1 << 33) + _orderable(proj1), stable=True).indices
        perm[pos] = pts1[o2]


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
    """All-NN sweep state: points, global heap rows, the captured launch.

    step() runs one iteration.  With prefetch (the default on a GPU) the
    next iteration's tree is planned on the host and built on a second
    stream while the current sweep runs, into one of two spare permutation
    buffers, so an iteration costs about the larger of the two instead of
    their sum.
    """

    def __init__(self, X, k, seed=3, leaf_lo=256, leaf_hi=4096, dtype="f32", config=None,
                 use_graph=True, log=None, prefetch=True):
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
        self.prefetch = prefetch and X.is_cuda
        self._bufs = [torch.empty(N, dtype=torch.int32, device=dev) for _ in range(2)]
        self._side = torch.cuda.Stream(device=dev) if self.prefetch else None
        self._pending = None
        self._last_end = None

    def _build(self, it, after=None):
        """Build tree `it` into spare buffer it % 2; on the side stream with
        prefetch (after event `after`: the buffer's last reader is done).
        Once the launch shape is known, the tree's work list is also made
        here (on the host, while the current sweep runs) in pinned memory.
        Returns (perm buffer, leaves, start event, done event, work list)."""
        import torch
        buf = self._bufs[it % 2]
        stream = self._side if self.prefetch else torch.cuda.current_stream()
        t0 = time.perf_counter()
        plan = plan_tree(self.X.shape[0], self.X.shape[1], it, self.seed, self.leaf_lo,
                         self.leaf_hi)
        with torch.cuda.stream(stream):
            if after is not None:
                stream.wait_event(after)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            _, leaves = build_tree(self.X, it, self.seed, self.leaf_lo, self.leaf_hi, perm_out=buf,
                                   plan=plan)
            e1.record(stream)
        items = None
        if self.config is not None:
            items = torch.from_numpy(make_items(leaves, self.config[0])).pin_memory()
        self._plan_ms = (time.perf_counter() - t0) * 1e3
        return buf, leaves, e0, e1, items

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
        """One tree iteration: (wait for / build) its tree, rewrite the work
        list, launch the sweep; with prefetch, start the next tree."""
        import torch
        main = torch.cuda.current_stream()
        t0 = time.perf_counter()
        if self._pending is None:
            self._pending = self._build(self.iteration)
        buf, leaves, et0, et1, items = self._pending
        self._pending = None
        if self.config is None:
            torch.cuda.synchronize()  # the autotuner times launches on this tree
        self._ensure_config(buf, leaves)
        qpb = self.config[0]
        if items is None:
            items = torch.from_numpy(make_items(leaves, qpb)).pin_memory()
        main.wait_event(et1)
        self.perm.copy_(buf)
        copied = torch.cuda.Event()
        copied.record(main)
        self.items[: items.shape[0]].copy_(items, non_blocking=True)
        self.nitems.fill_(items.shape[0])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(main)
        if self.use_graph and self.graph is not None:
            self.graph.replay()
        else:
            # First iteration (or no graph): launch directly; the direct
            # launch also sets the kernel's attributes before any capture.
            self._launch_captured()
        e1.record(main)
        if self.prefetch:
            # The next tree goes to the other buffer (read by the previous
            # iteration's copy, long finished); it waits for this copy only
            # so that the two streams never touch one buffer at once.  Its
            # build may still run when this step returns: the next step's
            # sweep waits for it on the device.
            self._pending = self._build(self.iteration + 1, after=copied)
        e1.synchronize()
        t1 = time.perf_counter()
        sweep_ms = e0.elapsed_time(e1)
        tree_ms = et0.elapsed_time(et1)
        if self.use_graph and self.graph is None:
            # Capture the launch once; later iterations rewrite perm/items/
            # nitems in place and replay it.  Capturing does not execute.
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._launch_captured()
            torch.cuda.synchronize()
        pairs = float(sum(n * n for _, n in leaves))
        info = {"iteration": self.iteration, "tree_ms": tree_ms, "sweep_ms": sweep_ms,
                "tree_plan_ms": self._plan_ms,
                "iteration_ms": (t1 - t0) * 1e3, "leaves": len(leaves), "pairs": pairs,
                "leaf_sizes": sorted({n for _, n in leaves})}
        self.timings.append(info)
        self.iteration += 1
        return buf, leaves, info
