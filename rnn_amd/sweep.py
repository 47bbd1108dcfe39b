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
kernel (libgsknn.so: gsknn_sweep_leaves_mfma_sdev, the FP32 MFMA form, for
d in {16, 32} and k <= 64; gsknn_sweep_leaves_[sd]dev, the exact VALU form,
otherwise) over a device work list;
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
    host (or `plan`, its result).  See build_trees.
    """
    outs = None if perm_out is None else [perm_out]
    perms, leaves = build_trees(X, [iteration], seed, leaf_lo, leaf_hi, perm_outs=outs,
                                plans=None if plan is None else [plan])
    return perms[0], leaves[0]


def build_trees(X, iterations, seed, leaf_lo=256, leaf_hi=4096, perm_outs=None, plans=None):
    """The trees of several iterations at once (rows of X, (N, d) float32).

    Returns (perms, leaves): one int32 tensor of point ids grouped by leaf per
    tree (written into perm_outs[t] when given) and one leaf list of (start,
    size) per tree.  On the device the trees advance together level by
    level: each level is one fused key kernel over the split nodes of every
    tree (libgsknn.so: node lookup, projection and packed sort key per point)
    and one sort of the keys, the key's node field numbering the nodes of all
    trees in order; a level whose node count does not fit the 63-bit key
    beside the projection and the point id is sorted tree by tree.  The per-
    level node arrays go to the device in one copy; no host synchronisation.
    """
    import torch
    dev = X.device
    N, d = X.shape
    T = len(iterations)
    if plans is None:
        plans = [plan_tree(N, d, it, seed, leaf_lo, leaf_hi) for it in iterations]
    id_bits = max(1, int(N - 1).bit_length())
    perm = torch.arange(N, device=dev, dtype=torch.int64).repeat(T)  # tree t at t N
    nlev = max((len(pl[0]) for pl in plans), default=0)
    if X.is_cuda and nlev:
        # Per level and group of trees sorted together: node starts (with
        # the tree's offset), element offsets and directions.
        groups = []  # (level, [trees], fused)
        for li in range(nlev):
            trees = [t for t in range(T) if li < len(plans[t][0])]
            nodes = sum(plans[t][0][li][0].size for t in trees)
            if max(1, int(nodes - 1).bit_length()) + 32 + id_bits <= 63:
                groups.append((li, trees, True))
                continue
            for t in trees:
                nodes = plans[t][0][li][0].size
                groups.append((li, [t], max(1, int(nodes - 1).bit_length()) + 32 + id_bits <= 63))
        ints, dirs, meta = [np.zeros(0, np.int64)], [np.zeros(0, np.float32)], []
        ioff = doff = 0
        for li, trees, fused in groups:
            if not fused:
                meta.append(None)
                continue
            st = np.concatenate([plans[t][0][li][0] + t * N for t in trees])
            sz = np.concatenate([plans[t][0][li][1] for t in trees])
            dr = np.concatenate([plans[t][0][li][2] for t in trees]).astype(np.float32)
            ints.append(np.concatenate([st, np.cumsum(sz) - sz]))
            dirs.append(dr.reshape(-1))
            meta.append((st.size, int(sz.sum()), ioff, doff))
            ioff += 2 * st.size
            doff += dr.size
        ints_t = torch.from_numpy(np.concatenate(ints)).pin_memory().to(dev, non_blocking=True)
        dirs_t = torch.from_numpy(np.concatenate(dirs)).pin_memory().to(dev, non_blocking=True)
        for (li, trees, fused), m in zip(groups, meta):
            if not fused:
                # Too many nodes for the packed key: this tree's level with
                # torch operations (two sorts).
                t = trees[0]
                _split_level_torch(X, perm[t * N:(t + 1) * N], *plans[t][0][li], id_bits)
                continue
            nn, total, io, do = m
            starts_t = ints_t[io:io + nn]
            offs_t = ints_t[io + nn:io + 2 * nn]
            pos = torch.empty(total, dtype=torch.int64, device=dev)
            key = torch.empty(total, dtype=torch.int64, device=dev)
            gsknn.tree_level_keys_dev(X, perm, starts_t, offs_t, dirs_t[do:do + nn * d].view(nn, d),
                                      total, id_bits, pos, key)
            perm[pos] = torch.sort(key).values & ((1 << id_bits) - 1)
    else:
        for t in range(T):
            pt = perm[t * N:(t + 1) * N]
            for starts, sizes, dirs_ in plans[t][0]:
                _split_level_torch(X, pt, starts, sizes, dirs_, id_bits)
    perms = []
    for t in range(T):
        pt = perm[t * N:(t + 1) * N]
        if perm_outs is not None:
            perm_outs[t].copy_(pt)
            perms.append(perm_outs[t])
        else:
            perms.append(pt.to(torch.int32))
    return perms, [pl[1] for pl in plans]


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
    # Ordering the leaves (stable) and then expanding them orders the items.
    lv = lv[np.argsort(-lv[:, 1], kind="stable")]
    nblk = (lv[:, 1] + qpb - 1) // qpb
    leaf = np.repeat(np.arange(lv.shape[0]), nblk)
    first = np.cumsum(nblk) - nblk
    items = np.zeros((leaf.size, 4), dtype=np.int32)
    items[:, 0] = lv[leaf, 0]
    items[:, 1] = lv[leaf, 1]
    items[:, 2] = (np.arange(leaf.size) - first[leaf]) * qpb
    return items


def _cache_path():
    return os.environ.get("GSKNN_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gsknn_tune.json"))


def _launch(dtype, d, k, X, X2, perm, items, nitems, D, I, qpb, ct, grid, stream, mfma=None):
    L = gsknn.lib()
    if mfma is not None:
        xn, mx = mfma
        rc = L.gsknn_sweep_leaves_mfma_sdev(d, k, gsknn._dp(X), gsknn._dp(X2), gsknn._dp(xn),
                                            gsknn._dp(mx), gsknn._dp(perm), gsknn._dp(items),
                                            gsknn._dp(nitems), gsknn._dp(D), gsknn._dp(I),
                                            D.stride(0), qpb, grid, gsknn._stream_ptr(stream))
        if rc:
            raise gsknn.GsknnError(gsknn.last_error())
        return
    fn = L.gsknn_sweep_leaves_ddev if dtype == "f64" else L.gsknn_sweep_leaves_sdev
    rc = fn(d, k, gsknn._dp(X), gsknn._dp(X2), gsknn._dp(perm), gsknn._dp(items),
            gsknn._dp(nitems), gsknn._dp(D), gsknn._dp(I), D.stride(0), qpb, ct, grid,
            gsknn._stream_ptr(stream))
    if rc:
        raise gsknn.GsknnError(gsknn.last_error())


def autotune(X, X2, k, perm, leaves, dtype="f32", cache=True, log=None, mfma=None):
    """Pick (qpb, ct, grid) for the leaf sweep by timing each candidate on one
    iteration's work list against scratch heaps.  Cached per shape.  With
    mfma = (xn, mx) the FP32 MFMA form is tuned (qpb 32, ct unused: 0)."""
    import torch
    dev = X.device
    N, d = X.shape
    sizes = np.array([n for _, n in leaves])
    key = f"{dtype}{'-mfma' if mfma is not None else ''}-d{d}-k{k}-N{N}-leaves{int(sizes.min())}-{int(sizes.max())}"
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
    for qpb in ((32,) if mfma is not None else QPB_CHOICES):
        items = torch.from_numpy(make_items(leaves, qpb)).to(dev)
        nitems = torch.tensor([items.shape[0]], dtype=torch.int32, device=dev)
        for ct in ((0,) if mfma is not None else CT_CHOICES[:1] if d in SCALAR_DIMS else CT_CHOICES):
            if mfma is None and int(gsknn.lib().gsknn_sweep_lds_bytes(elem, d, k, qpb, ct)) > 160 * 1024:
                continue
            for grid in GRID_CHOICES:
                D.fill_(float("inf"))
                I.fill_(-1)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                _launch(dtype, d, k, X, X2, perm, items, nitems, D, I, qpb, ct, grid, stream, mfma)
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

    step() runs one iteration.  Trees are built `batch` iterations at a time
    (build_trees: one sort per level for all of them).  With prefetch (the
    default on a GPU) a worker thread plans the next batch on the host (node
    structure and work lists) and the batch is built on a second stream while
    the current batch's sweeps run, into the other half of a double set of
    permutation buffers; `iterations`, when given, stops building past that
    many.  The first iteration, on empty heap rows, runs the exact VALU form
    of the sweep (nearly every pair of a cold row needs its exact distance);
    later iterations run the FP32 MFMA form where it applies (d in {16, 32},
    k <= 64; exact_first=False runs it from iteration 0), launched once
    directly and then replayed from a captured hipGraph.  Both forms give the
    oracle's rows bit for bit.
    """

    def __init__(self, X, k, seed=3, leaf_lo=256, leaf_hi=4096, dtype="f32", config=None,
                 use_graph=True, log=None, prefetch=True, batch=4, exact_first=True,
                 iterations=None):
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
        # FP32 MFMA form where it applies (GSKNN_SWEEP_MFMA=0 keeps the exact
        # VALU form, for comparison): both give the oracle's rows bit for bit.
        self.mfma = None
        if (dtype == "f32" and X.is_cuda and os.environ.get("GSKNN_SWEEP_MFMA", "1") != "0"
                and gsknn.sweep_mfma_supported(d, k, 32)):
            self.mfma = gsknn.sweep_norms_dev(X, self.X2)
        self.D = torch.full((N, k), float("inf"), dtype=tdt, device=dev)
        self.I = torch.full((N, k), -1, dtype=torch.int32, device=dev)
        self.perm = torch.empty(N, dtype=torch.int32, device=dev)
        self.config = config          # exact VALU form: (qpb, ct, grid)
        self.config_mfma = None       # MFMA form: (32, 0, grid)
        self.use_graph = use_graph
        self.graph = None
        self._graph_mfma = None
        self.tune_info = None
        self.tune_info_mfma = None
        self.iteration = 0
        self.exact_first = exact_first
        self.iterations = iterations
        self.timings = []
        self.prefetch = prefetch and X.is_cuda
        self.batch = max(1, int(batch))
        self._bufs = [[torch.empty(N, dtype=torch.int32, device=dev) for _ in range(self.batch)]
                      for _ in range(2)]
        self._side = torch.cuda.Stream(device=dev) if self.prefetch else None
        self._pool = None
        if self.prefetch:
            from concurrent.futures import ThreadPoolExecutor
            self._pool = ThreadPoolExecutor(max_workers=1)
        self._plans = {}       # first iteration of a batch -> future of its plan
        self._trees = {}       # iteration -> (buffer, leaves, batch record, items, qpb)
        self._set_free = [None, None]  # event: the buffer set's last reader is done

    def _uses_mfma(self, it):
        return self.mfma is not None and (it > 0 or not self.exact_first)

    def _batch_its(self, it0):
        end = it0 + self.batch
        if self.iterations is not None:
            end = min(end, self.iterations)
        return list(range(it0, end))

    def _plan_job(self, it0):
        """Host planning of a batch: node structure, and the work list of
        each tree whose launch shape is known already."""
        its = self._batch_its(it0)
        qpbs = [32 if self._uses_mfma(it) else (self.config[0] if self.config else None)
                for it in its]
        N, d = self.X.shape

        def work():
            plans = [plan_tree(N, d, it, self.seed, self.leaf_lo, self.leaf_hi) for it in its]
            items = [make_items(pl[1], q) if q else None for pl, q in zip(plans, qpbs)]
            return plans, items, qpbs
        return self._pool.submit(work) if self._pool is not None else work

    def _build_batch(self, it0):
        """Build the trees of a batch into buffer set (it0 // batch) % 2, on
        the side stream with prefetch (after that set's last copy); with
        prefetch the next batch's planning starts on the worker thread."""
        import torch
        its = self._batch_its(it0)
        if not its:
            return
        bset = (it0 // self.batch) % 2
        stream = self._side if self.prefetch else torch.cuda.current_stream()
        t0 = time.perf_counter()
        job = self._plans.pop(it0, None) or self._plan_job(it0)
        plans, items, qpbs = job.result() if hasattr(job, "result") else job()
        wait_ms = (time.perf_counter() - t0) * 1e3
        with torch.cuda.stream(stream):
            if self._set_free[bset] is not None:
                stream.wait_event(self._set_free[bset])
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            _, leaves = build_trees(self.X, its, self.seed, self.leaf_lo, self.leaf_hi,
                                    perm_outs=self._bufs[bset][:len(its)], plans=plans)
            e1.record(stream)
        rec = {"e0": e0, "e1": e1, "n": len(its), "plan_wait_ms": wait_ms}
        for t, it in enumerate(its):
            self._trees[it] = (self._bufs[bset][t], leaves[t], rec, items[t], qpbs[t])
        if self.prefetch:
            nxt = it0 + self.batch
            if self._batch_its(nxt):
                self._plans[nxt] = self._plan_job(nxt)

    def _ensure_items(self):
        import torch
        if not hasattr(self, "items"):
            N = self.X.shape[0]
            qmin = min(q for q in QPB_CHOICES + (32,))
            max_items = N // qmin + N // max(1, min(self.leaf_lo, qmin)) + 16
            self.items = torch.zeros((max_items, 4), dtype=torch.int32, device=self.X.device)
            self.nitems = torch.zeros(1, dtype=torch.int32, device=self.X.device)

    def tune(self):
        """Autotune both forms' launch shapes on the first tree and capture
        the steady form's graph, before the first step (so that no step
        includes them).  Optional: step() tunes a form on first use."""
        import torch
        if self.iteration != 0:
            return
        if 0 not in self._trees:
            self._build_batch(0)
        buf, leaves, rec = self._trees[0][:3]
        rec["e1"].synchronize()
        if self.config is None and (self.exact_first or self.mfma is None):
            self.config, self.tune_info = autotune(self.X, self.X2, self.k, buf, leaves,
                                                   self.dtype, log=self.log)
        if self.mfma is not None and self.config_mfma is None:
            self.config_mfma, self.tune_info_mfma = autotune(
                self.X, self.X2, self.k, buf, leaves, self.dtype, log=self.log, mfma=self.mfma)
        self._ensure_items()
        steady = self.mfma is not None
        if self.use_graph and self.graph is None:
            # The steady form's kernel has run (its tuning), so its attributes
            # are set; capturing does not execute.
            self.graph = torch.cuda.CUDAGraph()
            self._graph_mfma = steady
            with torch.cuda.graph(self.graph):
                self._launch(steady)
            torch.cuda.synchronize()

    def _launch(self, use_mfma):
        import torch
        qpb, ct, grid = self.config_mfma if use_mfma else self.config
        _launch(self.dtype, self.X.shape[1], self.k, self.X, self.X2, self.perm, self.items,
                self.nitems, self.D, self.I, qpb, ct, grid, torch.cuda.current_stream(),
                self.mfma if use_mfma else None)

    def step(self):
        """One tree iteration: (wait for / build) its tree, rewrite the work
        list, launch the sweep; with prefetch, start the next tree batch."""
        import torch
        main = torch.cuda.current_stream()
        t0 = time.perf_counter()
        it = self.iteration
        if self.iterations is not None and it >= self.iterations:
            raise ValueError(f"RPTreeSweep: {self.iterations} iterations were announced")
        if it not in self._trees:
            self._build_batch(it - it % self.batch)
        buf, leaves, rec, items_np, items_qpb = self._trees.pop(it)
        use_mfma = self._uses_mfma(it)
        if (use_mfma and self.config_mfma is None) or (not use_mfma and self.config is None):
            torch.cuda.synchronize()  # the autotuner times launches on this tree
            if use_mfma:
                self.config_mfma, self.tune_info_mfma = autotune(
                    self.X, self.X2, self.k, buf, leaves, self.dtype, log=self.log,
                    mfma=self.mfma)
            else:
                self.config, self.tune_info = autotune(self.X, self.X2, self.k, buf, leaves,
                                                       self.dtype, log=self.log)
        self._ensure_items()
        qpb = (self.config_mfma if use_mfma else self.config)[0]
        if items_np is None or items_qpb != qpb:
            items_np = make_items(leaves, qpb)
        items = torch.from_numpy(items_np).pin_memory()
        main.wait_event(rec["e1"])
        self.perm.copy_(buf)
        copied = torch.cuda.Event()
        copied.record(main)
        self.items[: items.shape[0]].copy_(items, non_blocking=True)
        self.nitems.fill_(items.shape[0])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(main)
        replayed = False
        if self.use_graph and self.graph is not None and use_mfma == self._graph_mfma:
            self.graph.replay()
            replayed = True
        else:
            # A form's first launch goes directly (it also sets the kernel's
            # attributes before any capture).
            self._launch(use_mfma)
        e1.record(main)
        if it % self.batch == self.batch - 1 or it + 1 == self.iterations:
            self._set_free[(it // self.batch) % 2] = copied
        if self.prefetch and it % self.batch == 0:
            # The next batch goes to the other buffer set, whose last copy
            # (the previous batch's last step) is already enqueued; its build
            # runs beside this batch's sweeps.
            self._build_batch(it + self.batch)
        e1.synchronize()
        t1 = time.perf_counter()
        sweep_ms = e0.elapsed_time(e1)
        rec["e1"].synchronize()
        tree_ms = rec["e0"].elapsed_time(rec["e1"]) / rec["n"]
        if self.use_graph and not replayed and (self.graph is None or use_mfma != self._graph_mfma) \
                and (use_mfma or self.mfma is None):
            # Capture the steady form's launch once; later iterations rewrite
            # perm/items/nitems in place and replay it.  Capturing does not
            # execute.
            self.graph = torch.cuda.CUDAGraph()
            self._graph_mfma = use_mfma
            with torch.cuda.graph(self.graph):
                self._launch(use_mfma)
            torch.cuda.synchronize()
        pairs = float(sum(n * n for _, n in leaves))
        info = {"iteration": it, "tree_ms": tree_ms, "sweep_ms": sweep_ms,
                "plan_wait_ms": rec["plan_wait_ms"], "form": "mfma" if use_mfma else "exact",
                "graph": replayed, "iteration_ms": (t1 - t0) * 1e3, "leaves": len(leaves),
                "pairs": pairs, "leaf_sizes": sorted({n for _, n in leaves})}
        self.timings.append(info)
        self.iteration += 1
        return buf, leaves, info

    def close(self):
        """Wait for the worker thread (planning ahead) to finish."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
