"""FP64 MFMA two-stage form through the C-ABI.

Unsplit per-wave launches keep, per query, the k2 = k + slack smallest
candidates under their MFMA key (no row read while streaming); a finalize
kernel forms the oracle distances of those entries, merges them with the
resident row and checks that no candidate left out can reach the k-th
distance; queries it cannot settle are re-run by the one-stage kernel.
Rows must be bit-identical to the oracle's in every case:
  * random data, cold and carried-over rows, every per-wave form (direct,
    operand ring, streamed-query ring); few or no fallbacks;
  * integer lattices (exact ties at the k-th distance): fallbacks counted
    and required, rows against the canonical-order oracle;
  * a heap-row map into a larger pool with ldk > k;
  * a caller workspace too small for the key heaps (the call runs in one
    stage) gives the same rows.
The sizes keep the launches unsplit (enough queries to fill the chip) and
the candidate sets small, so the oracle checks every row in seconds.  The
direct form (d <= 64) takes the two-stage form only under GSKNN_MFMA_TWO=2,
which the library reads once per process: its cases run in a child process.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle
from rnn_amd import datasets as ds
from rnn_amd import gsknn

pytestmark = pytest.mark.gpu


def _al16(x):
    return (x + 15) // 16 * 16


def _call(t, D, I, k, m, n, d, work_bytes=None, hrow=None):
    """gsknn_dknn_dev (AUTO) on a caller workspace; returns the two-stage
    fallback count (-1 when the call did not take the two-stage form)."""
    import torch
    L = gsknn.lib()
    need = L.gsknn_dknn_work_bytes(m, n, d, k, gsknn.MODE_AUTO)
    wb = need if work_bytes is None else work_bytes
    work = torch.zeros(max(wb, 16), dtype=torch.uint8, device="cuda")
    gsknn.dknn_dev(t["X"], t["X2"], t["am"], t["X"], t["X2"], t["bm"], D, I, hrow=hrow,
                   mode=gsknn.MODE_AUTO, work=work[:wb] if wb else None)
    torch.cuda.synchronize()
    off = L.gsknn_dknn_fallback_offset(m, n, d, k, gsknn.MODE_AUTO)
    if off == (1 << 64) - 1 or wb < need:
        return -1
    return int(work[off:off + 4].view(torch.int32).item())


def _tensors(X, am, bm):
    import torch
    dev = torch.device("cuda")
    X2 = (X * X).sum(1)
    return dict(X=torch.from_numpy(X).to(dev), X2=torch.from_numpy(X2).to(dev),
                am=torch.from_numpy(am).to(dev), bm=torch.from_numpy(bm).to(dev),
                Xh=X, X2h=X2, amh=am, bmh=bm)


def _assert_bitwise(D, I, Dr, Ir):
    Ds, Is = oracle.sort_rows(D, I)
    Drs, Irs = oracle.sort_rows(Dr, Ir)
    np.testing.assert_array_equal(Is, Irs)
    assert np.array_equal(Ds.view(np.uint8), Drs.view(np.uint8)), \
        f"max |dD| = {np.nanmax(np.abs(Ds - Drs))}"


def _random_case(d, n, warm):
    import torch
    g = np.random.Generator(np.random.PCG64(1000 + d))
    N, m, k = 80000, 2048, 32
    X = ds.gaussian_mixture(N, d, 32, seed=d)
    am = g.integers(0, N, m).astype(np.int32)
    bm = g.integers(0, N, n).astype(np.int32)
    t = _tensors(X, am, bm)
    Dr = np.full((n, k), np.inf)
    Ir = np.full((n, k), -1, dtype=np.int32)
    if warm:
        # Rows carried over from a call on 1024 other candidates.
        pre = g.integers(0, N, 1024).astype(np.int32)
        oracle.gsknn(X, t["X2h"], pre, X, t["X2h"], bm, Dr, Ir)
    D = torch.from_numpy(Dr.copy()).cuda()
    I = torch.from_numpy(Ir.copy()).cuda()
    print(f"[gpu] two-stage d={d} n={n} m={m} k={k} warm={warm}", flush=True)
    nfb = _call(t, D, I, k, m, n, d)
    print(f"[gpu]   fallbacks {nfb}", flush=True)
    assert nfb >= 0, "the call did not take the two-stage form"
    assert nfb <= n // 1000
    oracle.gsknn(X, t["X2h"], am, X, t["X2h"], bm, Dr, Ir)
    _assert_bitwise(D.cpu().numpy(), I.cpu().numpy(), Dr, Ir)


def _lattice_case(d):
    # Coordinates in {0, 1, 2}: hundreds of candidates share each distance,
    # so many k-th distances have more than the slack of exact ties within
    # the MFMA margin and their queries must go through the fallback.
    import torch
    g = np.random.Generator(np.random.PCG64(77 + d))
    N, m, k = 6000, 2048, 32
    n = 65536 if d == 64 else 32768
    X = g.integers(0, 3, (N, d)).astype(np.float64)
    am = g.integers(0, N, m).astype(np.int32)
    bm = g.integers(0, N, n).astype(np.int32)
    t = _tensors(X, am, bm)
    D = torch.full((n, k), float("inf"), dtype=torch.float64, device="cuda")
    I = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
    print(f"[gpu] two-stage lattice d={d} n={n} m={m} k={k}", flush=True)
    nfb = _call(t, D, I, k, m, n, d)
    print(f"[gpu]   fallbacks {nfb}", flush=True)
    assert nfb > 0
    Dr = np.full((n, k), np.inf)
    Ir = np.full((n, k), -1, dtype=np.int32)
    oracle.gsknn(X, t["X2h"], am, X, t["X2h"], bm, Dr, Ir, canonical=True)
    _assert_bitwise(D.cpu().numpy(), I.cpu().numpy(), Dr, Ir)


@pytest.mark.parametrize("d", [96, 128, 256, 258, 283, 300, 320])
@pytest.mark.parametrize("warm", [False, True])
def test_two_stage_random_rows_match_oracle(d, warm):
    _random_case(d, 32768, warm)


def test_two_stage_lattice_ties_fall_back():
    _lattice_case(96)


def test_direct_form_two_stage_in_child_process():
    env = dict(os.environ, GSKNN_MFMA_TWO="2")
    r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "child"], env=env,
                       cwd=ROOT, capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "child ok" in r.stdout


def test_two_stage_heap_row_map_and_ldk():
    import torch
    g = np.random.Generator(np.random.PCG64(5))
    N, m, n, d, k, ldk, rows = 80000, 2048, 32768, 128, 24, 40, 50000
    X = ds.gaussian_mixture(N, d, 32, seed=3)
    am = g.integers(0, N, m).astype(np.int32)
    bm = g.integers(0, N, n).astype(np.int32)
    hr = g.permutation(rows)[:n].astype(np.int32)
    t = _tensors(X, am, bm)
    pool_D = np.full((rows, ldk), 7.0)
    pool_I = np.full((rows, ldk), -5, dtype=np.int32)
    pool_D[:, :k] = np.inf
    pool_I[:, :k] = -1
    D = torch.from_numpy(pool_D.copy()).cuda()
    I = torch.from_numpy(pool_I.copy()).cuda()
    print(f"[gpu] two-stage hrow d={d} n={n} k={k} ldk={ldk}", flush=True)
    nfb = _call(t, D[:, :k], I[:, :k], k, m, n, d, hrow=torch.from_numpy(hr).cuda())
    assert nfb >= 0
    Dg, Ig = D.cpu().numpy(), I.cpu().numpy()
    # Untouched: rows outside the map and the columns past k.
    other = np.setdiff1d(np.arange(rows), hr)
    assert np.array_equal(Dg[other], pool_D[other]) and np.array_equal(Ig[other], pool_I[other])
    assert np.all(Dg[:, k:] == 7.0) and np.all(Ig[:, k:] == -5)
    Dr = np.full((n, k), np.inf)
    Ir = np.full((n, k), -1, dtype=np.int32)
    oracle.gsknn(X, t["X2h"], am, X, t["X2h"], bm, Dr, Ir)
    _assert_bitwise(np.ascontiguousarray(Dg[hr, :k]), np.ascontiguousarray(Ig[hr, :k]), Dr, Ir)


def test_small_workspace_runs_one_stage_with_same_rows():
    import torch
    g = np.random.Generator(np.random.PCG64(9))
    N, m, n, d, k = 80000, 2048, 32768, 96, 32
    X = ds.gaussian_mixture(N, d, 32, seed=4)
    am = g.integers(0, N, m).astype(np.int32)
    bm = g.integers(0, N, n).astype(np.int32)
    t = _tensors(X, am, bm)
    out = []
    for wb in (None, _al16(8 * m)):  # full workspace, then only the shifted norms
        D = torch.full((n, k), float("inf"), dtype=torch.float64, device="cuda")
        I = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
        print(f"[gpu] two-stage workspace {wb} d={d} n={n}", flush=True)
        nfb = _call(t, D, I, k, m, n, d, work_bytes=wb)
        assert (nfb >= 0) == (wb is None)
        out.append((D.cpu().numpy(), I.cpu().numpy()))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0].view(np.uint8), out[1][0].view(np.uint8))
    Dr = np.full((n, k), np.inf)
    Ir = np.full((n, k), -1, dtype=np.int32)
    oracle.gsknn(X, t["X2h"], am, X, t["X2h"], bm, Dr, Ir)
    _assert_bitwise(out[0][0], out[0][1], Dr, Ir)


if __name__ == "__main__" and sys.argv[1:] == ["child"]:
    # Direct form (d = 64), two-stage under GSKNN_MFMA_TWO=2.
    for warm in (False, True):
        _random_case(64, 65536, warm)
    _lattice_case(64)
    print("child ok", flush=True)
