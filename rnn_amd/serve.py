"""Continuous-batching exact-kNN server over a device-resident reference set.

Requests (a block of q query points and a neighbour count k) arrive over
time.  The dispatcher admits arrived requests while the HBM budget allows,
forms batches from admitted requests of one k (oldest first, up to
``max_batch`` queries), and launches each batch on the first free stream of a
small pool, so batches that arrive while others run execute concurrently on
the device.  Each batch is one exact kNN call of libgsknn.so against the whole
reference set (amap = all references), into fresh heap rows:

  mode "bf16": BF16 MFMA candidates + FP64 direct-difference re-rank
               (gsknn_bf16_knn_dev), FP64 distances, exact k-sets;
  mode "fp32": the sgsknn path (gsknn_sknn_dev), the reference oracle's FP32
               distances.

A pair's distance does not depend on the batch it is computed in and the
selection is by the total order (distance, index), so a request's result is
the same alone and in any mix.  Latency = completion time - arrival time.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import gsknn


@dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the trace
    queries: np.ndarray     # (q, d) float32
    k: int
    pinned: object = None   # pinned host copy of the queries
    D: object = None        # results (host, pinned), filled on completion
    I: object = None
    done: float = -1.0
    admitted: float = -1.0
    started: float = -1.0


@dataclass
class _Batch:
    reqs: list
    stream: object
    event: object
    nbytes: int
    tensors: list = field(default_factory=list)


class KnnServer:
    def __init__(self, R, mode="bf16", hbm_budget=8 << 30, max_batch=16384, streams=3):
        import torch
        assert mode in ("bf16", "fp32")
        self.torch = torch
        self.R = R
        self.N, self.d = R.shape
        self.mode = mode
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        dev = R.device
        self.dev = dev
        self.amap = torch.arange(self.N, dtype=torch.int32, device=dev)
        if mode == "bf16":
            self.prep = gsknn.bf16_prep_dev(R)
        else:
            self.R2 = torch.empty(self.N, dtype=torch.float32, device=dev)
            gsknn.sqnorm_dev(R, self.R2)
        torch.cuda.synchronize()
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(streams)]
        self.busy = {}
        self.reserved = 0
        self._fixed = {}
        self.batches_launched = 0
        self.batch_sizes = []

    # -- sizing -------------------------------------------------------------
    def _batch_bytes(self, nq, k):
        d, N = self.d, self.N
        b = nq * d * 4 + nq * k * 12 + nq * 8
        if self.mode == "bf16":
            b += int(gsknn.lib().gsknn_bf16_prep_bytes(nq, d))
            b += gsknn.bf16_work_bytes(N, nq, d, k)
        else:
            b += gsknn.work_bytes_s(N, nq, d, k)
        return b

    def _fixed_bytes(self, k):
        """The part of a batch's workspace that does not grow with its
        queries (the candidate panel of the register-form filters: the whole
        reference set in position order), taken as the workspace of one
        query."""
        if k not in self._fixed:
            d, N = self.d, self.N
            w = gsknn.bf16_work_bytes(N, 1, d, k) if self.mode == "bf16" else gsknn.work_bytes_s(N, 1, d, k)
            self._fixed[k] = int(w)
        return self._fixed[k]

    def _share_bytes(self, nq, k):
        """A request's share of a batch: its batch bytes less the fixed part,
        which the budget holds once per stream instead."""
        return max(0, self._batch_bytes(nq, k) - self._fixed_bytes(k))

    # -- one batch -----------------------------------------------------------
    def _launch(self, reqs, stream):
        torch = self.torch
        k = reqs[0].k
        nq = sum(r.queries.shape[0] for r in reqs)
        nbytes = self._batch_bytes(nq, k)
        with torch.cuda.stream(stream):
            Q = torch.empty((nq, self.d), dtype=torch.float32, device=self.dev)
            off = 0
            for r in reqs:
                q = r.queries.shape[0]
                Q[off:off + q].copy_(r.pinned, non_blocking=True)
                off += q
            bmap = torch.arange(nq, dtype=torch.int32, device=self.dev)
            if self.mode == "bf16":
                D = torch.full((nq, k), float("inf"), dtype=torch.float64, device=self.dev)
                I = torch.full((nq, k), -1, dtype=torch.int32, device=self.dev)
                prepQ = gsknn.bf16_prep_dev(Q, stream=stream)
                work = torch.empty(max(16, gsknn.bf16_work_bytes(self.N, nq, self.d, k)),
                                   dtype=torch.uint8, device=self.dev)
                gsknn.bf16_knn_dev(self.R, self.prep, self.amap, Q, prepQ, bmap, D, I,
                                   work=work, stream=stream)
                keep = [Q, bmap, D, I, prepQ, work]
            else:
                D = torch.full((nq, k), float("inf"), dtype=torch.float32, device=self.dev)
                I = torch.full((nq, k), -1, dtype=torch.int32, device=self.dev)
                Q2 = torch.empty(nq, dtype=torch.float32, device=self.dev)
                gsknn.sqnorm_dev(Q, Q2, stream=stream)
                work = torch.empty(max(16, gsknn.work_bytes_s(self.N, nq, self.d, k)),
                                   dtype=torch.uint8, device=self.dev)
                gsknn.sknn_dev(self.R, self.R2, self.amap, Q, Q2, bmap, D, I, work=work,
                               stream=stream)
                keep = [Q, bmap, D, I, Q2, work]
            off = 0
            for r in reqs:
                q = r.queries.shape[0]
                r.D = torch.empty((q, k), dtype=D.dtype, pin_memory=True)
                r.I = torch.empty((q, k), dtype=torch.int32, pin_memory=True)
                r.D.copy_(D[off:off + q], non_blocking=True)
                r.I.copy_(I[off:off + q], non_blocking=True)
                off += q
            ev = torch.cuda.Event()
            ev.record(stream)
        self.batches_launched += 1
        self.batch_sizes.append(nq)
        return _Batch(reqs, stream, ev, nbytes, keep)

    # -- trace driver --------------------------------------------------------
    def run(self, requests, realtime=True, log=None):
        """Serve `requests` (sorted by arrival).  With realtime=True arrivals
        follow the wall clock; otherwise every request is available at once.
        Returns (wall seconds, requests with results and times)."""
        torch = self.torch
        for r in requests:
            if r.pinned is None:
                r.pinned = torch.from_numpy(np.ascontiguousarray(r.queries)).pin_memory()
        future = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()            # arrived, not yet admitted
        ready = collections.OrderedDict()        # k -> deque of admitted requests
        inflight = []
        done = 0
        total = len(requests)
        t0 = time.perf_counter()
        while done < total:
            now = time.perf_counter() - t0
            while future and (not realtime or future[0].arrival <= now):
                waiting.append(future.popleft())
            # admission by HBM budget: each request's share of a batch, plus
            # the largest fixed part once per stream (a batch in flight per
            # stream at most)
            while waiting:
                r = waiting[0]
                need = self._share_bytes(r.queries.shape[0], r.k)
                fixed = len(self.streams) * max(self._fixed_bytes(kk) for kk in {r.k, *self._fixed})
                if self.reserved + need + fixed > self.budget and (self.reserved > 0 or inflight):
                    break
                waiting.popleft()
                r.admitted = now
                self.reserved += need
                r._need = need
                ready.setdefault(r.k, collections.deque()).append(r)
            # launch on free streams: oldest admitted request's k first
            for s in self.streams:
                if self.busy.get(s):
                    continue
                cands = [(q[0].admitted, k) for k, q in ready.items() if q]
                if not cands:
                    break
                _, k = min(cands)
                q = ready[k]
                batch, nq = [], 0
                while q and (not batch or nq + q[0].queries.shape[0] <= self.max_batch):
                    r = q.popleft()
                    r.started = now
                    batch.append(r)
                    nq += r.queries.shape[0]
                b = self._launch(batch, s)
                self.busy[s] = True
                inflight.append(b)
            # completions
            still = []
            for b in inflight:
                if b.event.query():
                    t = time.perf_counter() - t0
                    for r in b.reqs:
                        r.done = t
                        self.reserved -= r._need
                    done += len(b.reqs)
                    self.busy[b.stream] = False
                else:
                    still.append(b)
            progressed = len(still) != len(inflight)
            inflight = still
            if not progressed:
                if inflight and all(self.busy.get(s) for s in self.streams):
                    inflight[0].event.synchronize()
                elif not waiting and not any(ready.values()) and future:
                    time.sleep(max(0.0, min(0.002, future[0].arrival - (time.perf_counter() - t0))))
                else:
                    time.sleep(0.0002)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, requests


def make_trace(R_gen_W, n_requests, rate, seed, d, max_q=8192, ks=(10, 32, 100)):
    """Poisson arrivals at `rate` req/s; sizes floor(exp(U ln(max_q+1))) in
    [1, max_q]; k uniform in `ks`; query points from the reference generator."""
    from . import datasets as ds
    g = np.random.Generator(np.random.PCG64(seed))
    gaps = g.exponential(1.0 / rate, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.clip(np.floor(np.exp(g.uniform(0, np.log(max_q + 1), size=n_requests))), 1,
                    max_q).astype(int)
    kk = g.choice(np.asarray(ks), size=n_requests)
    total = int(sizes.sum())
    Qall = ds.low_rank(total, d, R_gen_W.shape[0], seed=seed + 1, W=R_gen_W)
    reqs, off = [], 0
    for i in range(n_requests):
        q = int(sizes[i])
        reqs.append(Request(i, float(arrivals[i]), Qall[off:off + q], int(kk[i])))
        off += q
    return reqs


def latency_stats(requests):
    lat = np.array([r.done - r.arrival for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
