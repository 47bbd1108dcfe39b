"""Python binding of libgsknn.so, mirroring the reference's python/gsknn.py.

Same function names and argument order as ChenhanYu/rnn python/gsknn.py:
``dgsknn(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I)`` and ``sgsknn(...)`` call
``[ds]gsknn_var1`` with the query set first (python/gsknn.py:145-183), where
``d`` is the dimension, ``k`` the neighbour count, ``q``/``r`` the query and
reference (candidate) index arrays, and D/I the (n, k) result rows, updated in
place.  The reference's size assert ``D.size == n * d`` (python/gsknn.py:157)
confuses d and k; here the rows are checked against k.

The reference also binds ``dgsknn_ref``/``sgsknn_ref``, its CPU oracle; those
are not part of this library (the CPU checker lives in oracle/).

The library is found next to this file (rnn_amd/lib/libgsknn.so) or under
``$GSKNN_DIR/lib`` as in the reference (python/gsknn.py:32).  There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgsknn.so")

MODE_AUTO, MODE_EXACT, MODE_MFMA, MODE_MFMA_TILE = 0, 1, 2, 3

_lib = None


class GsknnError(RuntimeError):
    pass


def lib_path() -> str:
    if os.path.exists(LIB_PATH):
        return LIB_PATH
    env = os.environ.get("GSKNN_DIR")
    if env and os.path.exists(os.path.join(env, "lib", "libgsknn.so")):
        return os.path.join(env, "lib", "libgsknn.so")
    raise GsknnError(
        f"libgsknn.so not found at {LIB_PATH}: build it with `python -m rnn_amd.build`")


def lib() -> ctypes.CDLL:
    """Load libgsknn.so.

    PyTorch ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    libamdhip64.so.7).  Importing torch first makes the dynamic linker bind
    this library's libamdhip64.so.7 dependency to that same copy, so device
    pointers and streams from torch and from this library belong to one
    runtime; loading the system runtime first would leave two in the process.
    """
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(lib_path())
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    L.gsknn_last_error.restype = ctypes.c_char_p
    L.gsknn_set_device.argtypes = [ctypes.c_int]
    L.gsknn_set_device.restype = ctypes.c_int
    L.gsknn_dknn_dev.argtypes = [i32] * 4 + [vp] * 8 + [i32, vp, i32, vp, sz, vp]
    L.gsknn_dknn_dev.restype = i32
    L.gsknn_dknn_noalloc_dev.argtypes = [i32] * 4 + [vp] * 8 + [i32, vp, i32, vp]
    L.gsknn_dknn_noalloc_dev.restype = i32
    L.gsknn_sknn_dev.argtypes = [i32] * 4 + [vp] * 8 + [i32, vp, i32, vp, sz, vp]
    L.gsknn_sknn_dev.restype = i32
    L.gsknn_dknn_work_bytes.argtypes = [i32] * 5
    L.gsknn_dknn_work_bytes.restype = sz
    L.gsknn_sknn_overflow_offset.argtypes = [i32] * 5
    L.gsknn_sknn_overflow_offset.restype = sz
    L.gsknn_dknn_fallback_offset.argtypes = [i32] * 5
    L.gsknn_dknn_fallback_offset.restype = sz
    L.gsknn_sknn_work_bytes.argtypes = [i32] * 5
    L.gsknn_sknn_work_bytes.restype = sz
    L.gsknn_bf16_prep_bytes.argtypes = [i32, i32]
    L.gsknn_bf16_prep_bytes.restype = sz
    L.gsknn_bf16_prep_dev.argtypes = [i32, i32, vp, vp, vp]
    L.gsknn_bf16_work_bytes.argtypes = [i32] * 4
    L.gsknn_bf16_work_bytes.restype = sz
    L.gsknn_bf16_knn_dev.argtypes = [i32] * 5 + [vp] * 3 + [i32] + [vp] * 5 + [i32, vp, vp, sz, vp, vp]
    L.gsknn_bf16_count_offset.argtypes = [i32] * 4 + [ctypes.POINTER(i32)]
    L.gsknn_bf16_count_offset.restype = sz
    L.gsknn_dsqnorm_dev.argtypes = [i32, i32, vp, vp, vp]
    L.gsknn_ssqnorm_dev.argtypes = [i32, i32, vp, vp, vp]
    L.gsknn_tree_level_keys_dev.argtypes = [ctypes.c_longlong, i32, i32, i32] + [vp] * 8
    L.gsknn_sweep_norms_sdev.argtypes = [i32, i32] + [vp] * 5
    L.gsknn_sweep_mfma_supported.argtypes = [i32] * 3
    L.gsknn_sweep_leaves_mfma_sdev.argtypes = [i32] * 2 + [vp] * 9 + [i32] * 3 + [vp]
    L.gsknn_rows_ascending_ddev.argtypes = [i32, i32, vp, vp, i32, vp]
    L.gsknn_rows_ascending_sdev.argtypes = [i32, i32, vp, vp, i32, vp]
    for name in ("dgsknn_var1", "sgsknn_var1"):
        getattr(L, name).argtypes = [i32] * 4 + [vp] * 8
        getattr(L, name).restype = None
    for name in ("dgsknn", "sgsknn", "dgsknn_var3"):
        getattr(L, name).argtypes = [i32] * 4 + [vp] * 7
        getattr(L, name).restype = None
    L.dgssq2nrm.argtypes = [i32] * 3 + [vp] * 7 + [i32]
    for name in ("dgsknn_1norm", "sgsknn_1norm"):
        getattr(L, name).argtypes = [i32] * 4 + [vp] * 5
        getattr(L, name).restype = None
    for name in ("gsknn_dknn1_dev", "gsknn_sknn1_dev"):
        getattr(L, name).argtypes = [i32] * 4 + [vp] * 6 + [i32, vp, vp, sz, vp]
        getattr(L, name).restype = i32
    for name in ("gsknn_dknn1_work_bytes", "gsknn_sknn1_work_bytes"):
        getattr(L, name).argtypes = [i32] * 4
        getattr(L, name).restype = sz
    L.heapCreate_d.argtypes = [i32, i32, ctypes.c_double]
    L.heapCreate_d.restype = ctypes.POINTER(HeapT)
    L.heapCreate_s.argtypes = [i32, i32, ctypes.c_float]
    L.heapCreate_s.restype = ctypes.POINTER(HeapT)
    L.heapAttach_d.argtypes = [i32, i32, vp, vp]
    L.heapAttach_d.restype = ctypes.POINTER(HeapT)
    L.heapAttach_s.argtypes = [i32, i32, vp, vp]
    L.heapAttach_s.restype = ctypes.POINTER(HeapT)
    L.heapFree_d.argtypes = [vp]
    L.heapFree_s.argtypes = [vp]
    L.heapSelect_d.argtypes = [i32, i32, vp, vp, vp, vp]
    L.heapSelect_s.argtypes = [i32, i32, vp, vp, vp, vp]
    L.gsknn_malloc_aligned.argtypes = [i32, i32, i32]
    L.gsknn_malloc_aligned.restype = vp


class HeapT(ctypes.Structure):
    """heap_t, byte-identical to the reference (include/gsknn.h:75-88)."""
    _fields_ = [("m", ctypes.c_int), ("k", ctypes.c_int), ("ldk", ctypes.c_int),
                ("d", ctypes.c_int), ("ro", ctypes.c_double),
                ("D", ctypes.POINTER(ctypes.c_double)), ("ro_s", ctypes.c_float),
                ("D_s", ctypes.POINTER(ctypes.c_float)), ("I", ctypes.POINTER(ctypes.c_int)),
                ("type", ctypes.c_int), ("prec", ctypes.c_int)]


def last_error() -> str:
    return lib().gsknn_last_error().decode()


def _raise_if_error():
    msg = last_error()
    if msg:
        raise GsknnError(msg)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I, dtype):
    # python/gsknn.py:60-80, with the row check taken against k.
    M, N = Xr.shape[0], Xq.shape[0]
    m, n = r.size, q.size
    assert Xr.shape[1] == Xq.shape[1] == d
    assert M >= m and N >= n
    assert Xq2.size == N and Xr2.size == M
    assert D.size == n * k and I.size == n * k
    for a in (Xq, Xr, Xq2, Xr2, D):
        assert a.dtype == dtype and a.flags.c_contiguous
    for a in (q, r, I):
        assert a.dtype == np.int32 and a.flags.c_contiguous
    return m, n


def dgsknn(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I):
    """python/gsknn.py:145-183: FP64 kNN of queries Xq[q] among Xr[r]."""
    m, n = _check(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I, np.float64)
    lib().dgsknn_var1(n, m, d, k, _p(Xq), _p(Xq2), _p(q), _p(Xr), _p(Xr2), _p(r), _p(D), _p(I))
    _raise_if_error()


def sgsknn(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I):
    """python/gsknn.py:55-95: FP32 kNN of queries Xq[q] among Xr[r]."""
    m, n = _check(d, k, Xq, Xq2, q, Xr, Xr2, r, D, I, np.float32)
    lib().sgsknn_var1(n, m, d, k, _p(Xq), _p(Xq2), _p(q), _p(Xr), _p(Xr2), _p(r), _p(D), _p(I))
    _raise_if_error()


# ---------------------------------------------------------------------------
# C-API style calls on numpy arrays (dgsknn/sgsknn argument order).
# ---------------------------------------------------------------------------
def dgsknn_heap(XA, XA2, amap, XB, XB2, bmap, D, I):
    """dgsknn(m, n, d, k, XA, XA2, amap, XB, XB2, bmap, heap) on (n, k) rows
    attached as a heap_t (frame/gsknn.c:1138-1184)."""
    n, k = D.shape
    h = lib().heapAttach_d(n, k, _p(D), _p(I))
    try:
        lib().dgsknn(amap.size, bmap.size, XA.shape[1], k, _p(XA), _p(XA2), _p(amap), _p(XB),
                     _p(XB2), _p(bmap), ctypes.cast(h, ctypes.c_void_p))
        _raise_if_error()
    finally:
        _free_shell(h)


def sgsknn_heap(XA, XA2, amap, XB, XB2, bmap, D, I):
    n, k = D.shape
    h = lib().heapAttach_s(n, k, _p(D), _p(I))
    try:
        lib().sgsknn(amap.size, bmap.size, XA.shape[1], k, _p(XA), _p(XA2), _p(amap), _p(XB),
                     _p(XB2), _p(bmap), ctypes.cast(h, ctypes.c_void_p))
        _raise_if_error()
    finally:
        _free_shell(h)


def dgsknn_1norm_heap(XA, amap, XB, bmap, D, I):
    """L1 kNN (KNN_1NORM) through dgsknn_1norm on (n, k) rows attached as a
    heap_t: m = amap.size candidates XA[amap], n = bmap.size queries."""
    n, k = D.shape
    h = lib().heapAttach_d(n, k, _p(D), _p(I))
    try:
        lib().dgsknn_1norm(amap.size, bmap.size, XA.shape[1], k, _p(XA), _p(amap), _p(XB),
                           _p(bmap), ctypes.cast(h, ctypes.c_void_p))
        _raise_if_error()
    finally:
        _free_shell(h)


def sgsknn_1norm_heap(XA, amap, XB, bmap, D, I):
    n, k = D.shape
    h = lib().heapAttach_s(n, k, _p(D), _p(I))
    try:
        lib().sgsknn_1norm(amap.size, bmap.size, XA.shape[1], k, _p(XA), _p(amap), _p(XB),
                           _p(bmap), ctypes.cast(h, ctypes.c_void_p))
        _raise_if_error()
    finally:
        _free_shell(h)


def _free_shell(h):
    # heapFree_[sd] frees D/I too (frame/gsknn_heap.c:321-337); the arrays here
    # belong to numpy, so only the heap_t shell is released.
    ctypes.CDLL(None).free(ctypes.cast(h, ctypes.c_void_p))


# ---------------------------------------------------------------------------
# Device-array calls (torch tensors on the GPU; no host staging).
# ---------------------------------------------------------------------------
def _dp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def dknn_dev(XA, XA2, amap, XB, XB2, bmap, D, I, hrow=None, mode=MODE_AUTO, work=None,
             stream=None):
    """FP64 kNN on device tensors, enqueued on ``stream`` (torch's current
    stream by default).  D/I: (rows, ldk) device tensors."""
    k = D.shape[1] if hrow is None else D.shape[1]
    m, n, d = amap.numel(), bmap.numel(), XA.shape[1]
    wb = work.numel() * work.element_size() if work is not None else 0
    rc = lib().gsknn_dknn_dev(m, n, d, k, _dp(XA), _dp(XA2), _dp(amap), _dp(XB), _dp(XB2),
                              _dp(bmap), _dp(D), _dp(I), D.stride(0), _dp(hrow), mode,
                              _dp(work), wb, _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


NOALLOC_SCRATCH_BYTES = 64 << 20  # GSKNN_NOALLOC_SCRATCH_BYTES (include/gsknn_hip.h)


def dknn_noalloc_dev(XA, XA2, amap, XB, XB2, bmap, D, I, hrow=None, mode=MODE_AUTO, stream=None):
    """gsknn_dknn_noalloc_dev: dknn_dev without any allocation (workspace
    from the library's fixed per-GPU scratch array); returns the C return
    code (0 on success) instead of raising."""
    k = D.shape[1]
    m, n, d = amap.numel(), bmap.numel(), XA.shape[1]
    return lib().gsknn_dknn_noalloc_dev(m, n, d, k, _dp(XA), _dp(XA2), _dp(amap), _dp(XB),
                                        _dp(XB2), _dp(bmap), _dp(D), _dp(I), D.stride(0),
                                        _dp(hrow), mode, _stream_ptr(stream))


def sknn_dev(XA, XA2, amap, XB, XB2, bmap, D, I, hrow=None, mode=MODE_AUTO, work=None,
             stream=None):
    """FP32 kNN on device tensors (sgsknn semantics); distances are the
    oracle's FP32 bits in every mode."""
    k = D.shape[1]
    m, n, d = amap.numel(), bmap.numel(), XA.shape[1]
    wb = work.numel() * work.element_size() if work is not None else 0
    rc = lib().gsknn_sknn_dev(m, n, d, k, _dp(XA), _dp(XA2), _dp(amap), _dp(XB), _dp(XB2),
                              _dp(bmap), _dp(D), _dp(I), D.stride(0), _dp(hrow), mode, _dp(work),
                              wb, _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


def knn1_dev(XA, amap, XB, bmap, D, I, hrow=None, work=None, stream=None):
    """L1 kNN (KNN_1NORM) on device tensors, FP64 or FP32 by XA's dtype."""
    k = D.shape[1]
    m, n, d = amap.numel(), bmap.numel(), XA.shape[1]
    wb = work.numel() * work.element_size() if work is not None else 0
    fn = lib().gsknn_dknn1_dev if XA.element_size() == 8 else lib().gsknn_sknn1_dev
    rc = fn(m, n, d, k, _dp(XA), _dp(amap), _dp(XB), _dp(bmap), _dp(D), _dp(I), D.stride(0),
            _dp(hrow), _dp(work), wb, _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


def work_bytes_l1(m, n, d, k, elem_bytes=8) -> int:
    fn = lib().gsknn_dknn1_work_bytes if elem_bytes == 8 else lib().gsknn_sknn1_work_bytes
    return int(fn(m, n, d, k))


def work_bytes_d(m, n, d, k, mode=MODE_AUTO) -> int:
    return int(lib().gsknn_dknn_work_bytes(m, n, d, k, mode))


def work_bytes_s(m, n, d, k, mode=MODE_AUTO) -> int:
    return int(lib().gsknn_sknn_work_bytes(m, n, d, k, mode))


def sqnorm_dev(X, X2, stream=None):
    fn = lib().gsknn_dsqnorm_dev if X.element_size() == 8 else lib().gsknn_ssqnorm_dev
    rc = fn(X.shape[0], X.shape[1], _dp(X), _dp(X2), _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


def sweep_norms_dev(X, X2, stream=None):
    """Norms the FP32 MFMA leaf sweep's error bound needs: xn (N,) with
    xn[i] = |X[i]| rounded up, and mx (2,) = {max xn, max |X2|}."""
    import torch
    xn = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
    mx = torch.zeros(2, dtype=torch.float32, device=X.device)
    rc = lib().gsknn_sweep_norms_sdev(X.shape[0], X.shape[1], _dp(X), _dp(X2), _dp(xn), _dp(mx),
                                      _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())
    return xn, mx


def sweep_mfma_supported(d, k, qpb=32) -> bool:
    return bool(lib().gsknn_sweep_mfma_supported(d, k, qpb))


def tree_level_keys_dev(X, perm, starts, offs, dirs, total, id_bits, pos, keys, stream=None):
    """One RP-tree level on the device (see gsknn_tree_level_keys_dev)."""
    rc = lib().gsknn_tree_level_keys_dev(int(total), starts.numel(), X.shape[1], id_bits, _dp(X),
                                         _dp(perm), _dp(starts), _dp(offs), _dp(dirs), _dp(pos),
                                         _dp(keys), _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


def rows_ascending_dev(D, I, stream=None):
    """Rewrite heap rows (n, k) of device tensors D/I in place in ascending
    (distance, index) order -- the sorted-output step tree codes apply to
    gsknn heaps (frame/old/dgsrnn_directKQuery.cpp:110-123)."""
    assert D.dim() == 2 and D.shape == I.shape and D.stride(1) == 1 and I.stride(1) == 1
    assert D.stride(0) == I.stride(0)
    fn = lib().gsknn_rows_ascending_ddev if D.element_size() == 8 else lib().gsknn_rows_ascending_sdev
    rc = fn(D.shape[0], D.shape[1], _dp(D), _dp(I), D.stride(0), _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())


# ---------------------------------------------------------------------------
# BF16 candidate pass + FP64 re-rank (device tensors).
# ---------------------------------------------------------------------------
def bf16_prep_dev(X, stream=None):
    """Prepare an FP32 device table (N, d) for the BF16 path; returns the
    prepared buffer (uint8 device tensor)."""
    import torch
    N, d = X.shape
    nb = int(lib().gsknn_bf16_prep_bytes(N, d))
    prep = torch.empty(max(nb, 16), dtype=torch.uint8, device=X.device)
    rc = lib().gsknn_bf16_prep_dev(N, d, _dp(X), _dp(prep), _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())
    return prep


def bf16_work_bytes(m, n, d, k) -> int:
    return int(lib().gsknn_bf16_work_bytes(m, n, d, k))


def bf16_candidate_counts(work, m, n, d, k):
    """Per-query candidate counts (summed over candidate splits) of the last
    bf16_knn_dev call that used the device byte tensor ``work``."""
    ns = ctypes.c_int(0)
    off = int(lib().gsknn_bf16_count_offset(m, n, d, k, ctypes.byref(ns)))
    cnt = work[off:off + 4 * n * ns.value].view(dtype=__import__("torch").int32)
    return cnt.view(ns.value, n).sum(dim=0)


def bf16_knn_dev(XA, prepA, amap, XB, prepB, bmap, D, I, hrow=None, work=None, overflow=None,
                 stream=None):
    """Exact kNN (FP64 distances) of queries XB[bmap] among XA[amap] from a
    BF16 candidate pass; D (float64) / I (int32) rows updated in place."""
    m, n, d, k = amap.numel(), bmap.numel(), XA.shape[1], D.shape[1]
    wb = work.numel() * work.element_size() if work is not None else 0
    rc = lib().gsknn_bf16_knn_dev(m, n, d, k, XA.shape[0], _dp(XA), _dp(prepA), _dp(amap),
                                  XB.shape[0], _dp(XB), _dp(prepB), _dp(bmap), _dp(D), _dp(I),
                                  D.stride(0), _dp(hrow), _dp(work), wb, _dp(overflow),
                                  _stream_ptr(stream))
    if rc:
        raise GsknnError(last_error())
