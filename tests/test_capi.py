"""A compiled C caller of the boundary (tests/c/capi_driver.c): it includes
include/gsknn.h, static-asserts the heap_t layout of the reference
(include/gsknn.h:75-88: 64 bytes, field offsets 0, 4, 8, 12, 16, 24, 32, 40,
48, 56, 60), links libgsknn.so and runs config 1's two calls through
heapCreate_d -> dgsknn -> heapFree_d, as the reference's drivers do
(test/test_dgsknn.c:133-191)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import gsknn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "capi_driver.c")
INC = os.path.join(ROOT, "include")


def _cc():
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    return cc


def test_c_driver_compiles_against_the_header(tmp_path):
    """The header compiles in plain C (C11) and heap_t has the reference's
    layout (the static assertions hold); compile only, no GPU."""
    obj = tmp_path / "capi_driver.o"
    subprocess.run([_cc(), "-std=c11", "-Wall", "-Werror", "-I", INC, "-c", SRC, "-o", str(obj)],
                   check=True)
    assert obj.exists()


@pytest.mark.gpu
def test_c_driver_config1_two_calls(tmp_path):
    libdir = os.path.dirname(gsknn.lib_path())
    exe = tmp_path / "capi_driver"
    subprocess.run([_cc(), "-std=c11", "-O2", "-I", INC, SRC, "-o", str(exe), "-L", libdir,
                    "-lgsknn", f"-Wl,-rpath,{libdir}"], check=True)
    p = ds.config1()
    X = np.ascontiguousarray(p.XA)
    assert p.XB is p.XA or np.array_equal(p.XB, p.XA)
    calls = [p.amap] + p.extra_amaps
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([X.shape[0], X.shape[1], p.n, p.k, len(calls)], dtype=np.int32).tofile(f)
        X.tofile(f)
        p.XA2.astype(np.float64).tofile(f)
        p.bmap.astype(np.int32).tofile(f)
        for am in calls:
            np.array([am.size], dtype=np.int32).tofile(f)
            am.astype(np.int32).tofile(f)
    out = tmp_path / "out.bin"
    print(f"[gpu] C driver config1 N={X.shape[0]} d={X.shape[1]} n={p.n} k={p.k}", flush=True)
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = np.dtype([("D", np.float64, (p.k,)), ("I", np.int32, (p.k,))])
    got = np.fromfile(out, dtype=rec).reshape(len(calls), p.n)
    Dr = np.full((p.n, p.k), 1.79e308)
    Ir = np.full((p.n, p.k), -1, dtype=np.int32)
    for c, am in enumerate(calls):
        oracle.gsknn(X, p.XA2, am, X, p.XA2, p.bmap, Dr, Ir)
        Ds, Is = oracle.sort_rows(np.ascontiguousarray(got[c]["D"]), np.ascontiguousarray(got[c]["I"]))
        Drs, Irs = oracle.sort_rows(Dr, Ir)
        np.testing.assert_array_equal(Is, Irs)
        assert np.array_equal(Ds, Drs)
