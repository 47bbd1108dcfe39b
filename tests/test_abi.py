"""C-ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, and its host-side heap utilities behave like the
reference's (frame/gsknn_heap.c, frame/gsknn_util.c)."""
import ctypes
import os
import re

import numpy as np

from oracle import oracle
from rnn_amd import gsknn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("gsknn.h", "gsknn_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#.*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name not in ("if", "sizeof", "defined"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    L = gsknn.lib()
    names = declared_functions()
    assert {"dgsknn", "sgsknn", "dgsknn_var1", "heapCreate_d", "gsknn_dknn_dev"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_heap_t_layout_matches_reference():
    # include/gsknn.h:75-88 on LP64: 4 ints, double, 2 pointers around a
    # float, int pointer, two enums.
    H = gsknn.HeapT
    assert ctypes.sizeof(H) == 64
    offs = {f: getattr(H, f).offset for f, _ in H._fields_}
    assert offs == {"m": 0, "k": 4, "ldk": 8, "d": 12, "ro": 16, "D": 24, "ro_s": 32,
                    "D_s": 40, "I": 48, "type": 56, "prec": 60}


def test_heap_create_layouts():
    L = gsknn.lib()
    h = L.heapCreate_d(3, 7, 2.5)
    c = h.contents
    assert (c.m, c.k, c.ldk, c.d) == (3, 7, 7, 2)
    D = np.ctypeslib.as_array(c.D, shape=(21,))
    I = np.ctypeslib.as_array(c.I, shape=(21,))
    assert np.all(D == 2.5) and np.all(I == -1)
    L.heapFree_d(ctypes.cast(h, ctypes.c_void_p))
    # k > 512: ldk = ((k + 2) / 4 + 1) * 4 and the 3-slot header
    h = L.heapCreate_d(2, 600, 9.0)
    c = h.contents
    assert c.ldk == ((600 + 3 - 1) // 4 + 1) * 4 and c.d == 4
    D = np.ctypeslib.as_array(c.D, shape=(2 * c.ldk,)).reshape(2, c.ldk)
    I = np.ctypeslib.as_array(c.I, shape=(2 * c.ldk,)).reshape(2, c.ldk)
    assert np.all(I[:, 0] == 600) and np.all(I[:, 1] == 0) and np.all(I[:, 2] == c.ldk)
    assert np.all(D[:, 0] == 9.0) and np.all(D[:, 3:603] == 9.0) and np.all(I[:, 3:603] == -1)
    L.heapFree_d(ctypes.cast(h, ctypes.c_void_p))
    h = L.heapCreate_s(4, 5, 1.5)
    c = h.contents
    assert c.prec == 1 and c.ldk == 5
    assert np.all(np.ctypeslib.as_array(c.D_s, shape=(20,)) == np.float32(1.5))
    L.heapFree_s(ctypes.cast(h, ctypes.c_void_p))


def test_host_heap_select_matches_oracle_compat():
    L = gsknn.lib()
    g = np.random.Generator(np.random.PCG64(4))
    for dt, fn in ((np.float64, L.heapSelect_d), (np.float32, L.heapSelect_s)):
        x = np.floor(g.random(500) * 50).astype(dt)  # ties
        alpha = np.arange(500, dtype=np.int32)
        D1 = np.full(16, 1e30, dtype=dt)
        I1 = np.full(16, -1, dtype=np.int32)
        D2, I2 = D1.copy(), I1.copy()
        fn(500, 16, gsknn._p(x), gsknn._p(alpha), gsknn._p(D1), gsknn._p(I1))
        oracle.heap_select(x, alpha, D2, I2, canonical=False)
        np.testing.assert_array_equal(D1, D2)
        np.testing.assert_array_equal(I1, I2)


def test_bubble_sort_and_swaps():
    L = gsknn.lib()
    g = np.random.Generator(np.random.PCG64(5))
    D = np.floor(g.random(40) * 9)
    I = np.arange(40, dtype=np.int32)
    L.bubbleSort_d(40, gsknn._p(D), gsknn._p(I))
    order = np.argsort(np.floor(np.random.Generator(np.random.PCG64(5)).random(40) * 9),
                       kind="stable")
    np.testing.assert_array_equal(I, order)
    x = np.array([1.0, 2.0, 3.0])
    L.swap_double(gsknn._p(x), 0, 2)
    assert list(x) == [3.0, 2.0, 1.0]


def test_malloc_aligned():
    L = gsknn.lib()
    p = L.gsknn_malloc_aligned(10, 10, 8)
    assert p and p % 32 == 0
    ctypes.CDLL(None).free(ctypes.c_void_p(p))


def test_work_bytes_planning_is_host_only():
    # Split planning: few queries, many candidates -> partial-list workspace.
    assert gsknn.work_bytes_d(2_000_000, 10, 128, 32) > 2_000_000 * 8
    # An unsplit FP64 MFMA launch of the direct form (d <= 64, one stage)
    # needs only the candidates' shifted norms (one double per candidate
    # position).
    n, k2 = 65536, 40
    assert gsknn.work_bytes_d(65536, n, 64, 32) == 65536 * 8
    # The operand-ring form (d > 64) runs in two stages: the shifted norms,
    # then key heaps of k2 = 40 entries per query (8-byte key, 4-byte id),
    # the fallback count (16 bytes) and the two fallback lists (4 bytes per
    # query each).
    assert gsknn.work_bytes_d(65536, n, 96, 32) == 65536 * 8 + n * k2 * 12 + 16 + 2 * n * 4
    assert gsknn.work_bytes_d(0, 10, 8, 4) == 0


def test_l1_large_r_rejected_before_any_gpu_work():
    # The reference's large-r header layout has no L1 form.
    X = np.zeros((4, 3))
    idx = np.arange(4, dtype=np.int32)
    D = np.zeros((4, 600))
    I = np.zeros((4, 600), dtype=np.int32)
    try:
        gsknn.dgsknn_1norm_heap(X, idx, X, idx, D, I)
    except gsknn.GsknnError as e:
        assert "r=600" in str(e)
    else:
        raise AssertionError("dgsknn_1norm accepted r = 600")


def test_l1_work_bytes_planning_is_host_only():
    assert gsknn.work_bytes_l1(2_000_000, 10, 64, 32, 8) > 0
    assert gsknn.work_bytes_l1(2_000_000, 10, 100, 32, 4) > 0
    assert gsknn.work_bytes_l1(0, 10, 8, 4, 8) == 0
