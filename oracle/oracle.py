"""ctypes front end of the CPU checker (test infrastructure only).

Loads ``oracle/liboracle.so`` (this repository's C restatement of the
reference's ``dgsknn_ref``/``sgsknn_ref``, frame/gsknn_ref.c:63-162) and, when
built, ``oracle/_ref/libgsknn_ref.so`` (the reference's own oracle sources
compiled by ``oracle/Makefile``) and ``oracle/_ref/libgsknn_opt.so`` (its
optimized library, whose kernel tables reach the L1 kernels).  Only tests, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` import this module; the product library
never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgsknn_ref.so")
REF_OPT_PATH = os.path.join(HERE, "_ref", "libgsknn_opt.so")

_lib = None
_ref = None
_ref_opt = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def ref_lib():
    """The reference oracle compiled from its own sources, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def ref_opt_lib():
    """The reference's optimized library (frame/gsknn.c and its x86-64 AVX
    kernels) compiled from its own sources, or None."""
    global _ref_opt
    if _ref_opt is None and os.path.exists(REF_OPT_PATH):
        _ref_opt = ctypes.CDLL(REF_OPT_PATH)
    return _ref_opt


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(X, X2, idx, dtype):
    assert X.dtype == dtype and X2.dtype == dtype and idx.dtype == np.int32
    assert X.flags.c_contiguous and X2.flags.c_contiguous and idx.flags.c_contiguous


def gsknn(XA, XA2, amap, XB, XB2, bmap, D, I, canonical=False, threads=None):
    """Blocked oracle: m = amap.size candidates, n = bmap.size heap rows.

    ``D``/``I`` are (n, ldk) arrays updated in place.  ``canonical`` selects the
    (distance, index) tie order instead of the reference's replace-on-<= rule.
    """
    dt = XA.dtype
    _check(XA, XA2, amap, dt)
    _check(XB, XB2, bmap, dt)
    assert D.dtype == dt and I.dtype == np.int32 and D.shape == I.shape
    assert D.flags.c_contiguous and I.flags.c_contiguous
    m, n = amap.size, bmap.size
    d = XA.shape[1]
    assert XB.shape[1] == d and D.shape[0] == n
    ldk = D.shape[1]
    k = ldk
    fn = lib().oracle_dgsknn if dt == np.float64 else lib().oracle_sgsknn
    _with_threads(threads, lambda: fn(
        ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(k),
        _p(XA), _p(XA2), _p(amap), _p(XB), _p(XB2), _p(bmap), _p(D), _p(I),
        ctypes.c_int(ldk), ctypes.c_int(1 if canonical else 0)))


def gsknn_l1(XA, amap, XB, bmap, D, I, canonical=False, threads=None):
    """L1 (KNN_1NORM) oracle, same roles and rows as ``gsknn``, no norms."""
    dt = XA.dtype
    assert XB.dtype == dt and D.dtype == dt and I.dtype == np.int32 and D.shape == I.shape
    for a in (XA, XB, amap, bmap, D, I):
        assert a.flags.c_contiguous
    assert amap.dtype == np.int32 and bmap.dtype == np.int32
    m, n, d = amap.size, bmap.size, XA.shape[1]
    assert XB.shape[1] == d and D.shape[0] == n
    ldk = D.shape[1]
    fn = lib().oracle_dgsknn_l1 if dt == np.float64 else lib().oracle_sgsknn_l1
    _with_threads(threads, lambda: fn(
        ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(ldk),
        _p(XA), _p(amap), _p(XB), _p(bmap), _p(D), _p(I), ctypes.c_int(ldk),
        ctypes.c_int(1 if canonical else 0)))


def ref_opt_l1(XA, amap, XB, bmap, D, I):
    """The reference's own FP64 L1 path: its dgsknn_var1 (frame/gsknn.c:
    883-1103) run with kernel-table slot 0 pointed at the KNN_1NORM entries
    (rankk_d[1], micro_d[1]; kernels/x86_64/sandybridge/gsknn_kernel.h:
    125-163), restored afterwards.  m = amap.size candidates, n = bmap.size
    heap rows of stride k = D.shape[1] (<= 512)."""
    r = ref_opt_lib()
    if r is None:
        raise FileNotFoundError(REF_OPT_PATH)
    m, n, d, k = amap.size, bmap.size, XA.shape[1], D.shape[1]
    assert XA.dtype == np.float64 and XB.dtype == np.float64 and k <= 512
    micro = (ctypes.c_void_p * 2).in_dll(r, "micro_d")
    rankk = (ctypes.c_void_p * 2).in_dll(r, "rankk_d")
    saved = (micro[0], rankk[0])
    zeros_a = np.zeros(XA.shape[0])
    zeros_b = np.zeros(XB.shape[0])
    micro[0], rankk[0] = micro[1], rankk[1]
    try:
        # var1 roles: queries first (heap rows), candidates second.
        r.dgsknn_var1(ctypes.c_int(n), ctypes.c_int(m), ctypes.c_int(d), ctypes.c_int(k),
                      _p(XB), _p(zeros_b), _p(bmap), _p(XA), _p(zeros_a), _p(amap), _p(D), _p(I))
    finally:
        micro[0], rankk[0] = saved


def ref_opt_gsknn(XA, XA2, amap, XB, XB2, bmap, D, I, threads=1):
    """The reference's own optimized FP64 kNN (its dgsknn_var1 with the x86-64
    AVX micro-kernels, frame/gsknn.c:883-1103): m = amap.size candidates, n =
    bmap.size heap rows of stride k = D.shape[1] (<= 512).  `threads` is the
    reference's GSKNN_IC_NT (read from the environment at each call)."""
    r = ref_opt_lib()
    if r is None:
        raise FileNotFoundError(REF_OPT_PATH)
    m, n, d, k = amap.size, bmap.size, XA.shape[1], D.shape[1]
    assert XA.dtype == np.float64 and XB.dtype == np.float64 and k <= 512
    prev = os.environ.get("GSKNN_IC_NT")
    os.environ["GSKNN_IC_NT"] = str(int(threads))
    try:
        r.dgsknn_var1(ctypes.c_int(n), ctypes.c_int(m), ctypes.c_int(d), ctypes.c_int(k),
                      _p(XB), _p(XB2), _p(bmap), _p(XA), _p(XA2), _p(amap), _p(D), _p(I))
    finally:
        if prev is None:
            os.environ.pop("GSKNN_IC_NT", None)
        else:
            os.environ["GSKNN_IC_NT"] = prev


def brute_l1(XA, amap, XB, bmap, k):
    """Long-double L1 brute force; rows ascending by (dist, idx)."""
    XA = np.ascontiguousarray(XA, dtype=np.float64)
    XB = np.ascontiguousarray(XB, dtype=np.float64)
    m, n, d = amap.size, bmap.size, XA.shape[1]
    kk = min(k, m)
    D = np.full((n, kk), np.inf)
    I = np.full((n, kk), -1, dtype=np.int32)
    lib().oracle_brute_l1_d(ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(kk),
                            _p(XA), _p(amap), _p(XB), _p(bmap), _p(D), _p(I), ctypes.c_int(kk))
    return D, I


def ref_gsknn(XA, XA2, amap, XB, XB2, bmap, D, I):
    """The reference's dgsknn_ref/sgsknn_ref compiled from its sources."""
    r = ref_lib()
    if r is None:
        raise FileNotFoundError(REF_PATH)
    dt = XA.dtype
    m, n, d, k = amap.size, bmap.size, XA.shape[1], D.shape[1]
    fn = r.dgsknn_ref if dt == np.float64 else r.sgsknn_ref
    fn(ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(k),
       _p(XA), _p(XA2), _p(amap), _p(XB), _p(XB2), _p(bmap), _p(D), _p(I))


def direct_knn_f32in(XA, amap, XB, bmap, D, I):
    """FP64 direct-difference distances on FP32 points, canonical selection."""
    assert XA.dtype == np.float32 and XB.dtype == np.float32 and D.dtype == np.float64
    m, n, d, k = amap.size, bmap.size, XA.shape[1], D.shape[1]
    lib().oracle_direct_knn_f32in(
        ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(k),
        _p(XA), _p(amap), _p(XB), _p(bmap), _p(D), _p(I), ctypes.c_int(k))


def brute(XA, amap, XB, bmap, k):
    """Long-double direct-difference brute force; rows ascending by (dist, idx)."""
    XA = np.ascontiguousarray(XA, dtype=np.float64)
    XB = np.ascontiguousarray(XB, dtype=np.float64)
    m, n, d = amap.size, bmap.size, XA.shape[1]
    kk = min(k, m)
    D = np.full((n, kk), np.inf)
    I = np.full((n, kk), -1, dtype=np.int32)
    lib().oracle_brute_d(ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(kk),
                         _p(XA), _p(amap), _p(XB), _p(bmap), _p(D), _p(I), ctypes.c_int(kk))
    return D, I


def heap_select(x, alpha, D, I, canonical=False):
    """One row of heapSelect_[sd] (frame/gsknn_heap.c:104-162)."""
    fn = lib().oracle_heap_select_d if D.dtype == np.float64 else lib().oracle_heap_select_s
    fn(ctypes.c_int(x.size), ctypes.c_int(D.size), _p(x), _p(alpha), _p(D), _p(I),
       ctypes.c_int(1 if canonical else 0))


def sort_rows(D, I):
    """Copies of D, I with each row ascending by (distance, index)."""
    D = np.ascontiguousarray(D).copy()
    I = np.ascontiguousarray(I).copy()
    n, k = D.shape
    fn = lib().oracle_sort_rows_d if D.dtype == np.float64 else lib().oracle_sort_rows_s
    fn(ctypes.c_int(n), ctypes.c_int(k), _p(D), _p(I), ctypes.c_int(k))
    return D, I


def rows_are_heaps(D, I, canonical=False):
    n, k = D.shape
    fn = lib().oracle_rows_are_heaps_d if D.dtype == np.float64 else lib().oracle_rows_are_heaps_s
    return bool(fn(ctypes.c_int(n), ctypes.c_int(k), _p(np.ascontiguousarray(D)),
                   _p(np.ascontiguousarray(I)), ctypes.c_int(k), ctypes.c_int(1 if canonical else 0)))


def sqnorm(X):
    X = np.ascontiguousarray(X)
    out = np.empty(X.shape[0], dtype=X.dtype)
    fn = lib().oracle_sqnorm_d if X.dtype == np.float64 else lib().oracle_sqnorm_s
    fn(ctypes.c_int(X.shape[0]), ctypes.c_int(X.shape[1]), _p(X), _p(out))
    return out


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def _with_threads(threads, fn):
    # OpenMP reads OMP_NUM_THREADS once per process; per-call control goes
    # through omp_set_num_threads in libgomp.
    if threads is None:
        return fn()
    gomp = ctypes.CDLL("libgomp.so.1")
    prev = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(ctypes.c_int(int(threads)))
    try:
        return fn()
    finally:
        gomp.omp_set_num_threads(ctypes.c_int(prev))
