"""gsknn_dknn_noalloc_dev (the entry point that allocates nothing) through
the C-ABI: its rows equal gsknn_dknn_dev's bit for bit on two calls in a
row that use the fixed scratch array (candidate-split launches, and the
two-stage form's key heaps), the second smaller than the first; a call
whose key heaps exceed the array runs in one stage; a call needing more
scratch than the array holds even then is refused with GSKNN_ELIMIT before
any work is enqueued."""
import numpy as np
import pytest

from rnn_amd import gsknn

pytestmark = pytest.mark.gpu

GSKNN_ELIMIT = 4


def _case(g, N, d, m, n, k):
    import torch
    dev = torch.device("cuda")
    X = g.standard_normal((N, d))
    X2 = (X * X).sum(1)
    t = dict(X=torch.from_numpy(X).to(dev), X2=torch.from_numpy(X2).to(dev),
             am=torch.from_numpy(g.integers(0, N, m).astype(np.int32)).to(dev),
             bm=torch.from_numpy(g.integers(0, N, n).astype(np.int32)).to(dev))
    return t


def _rows(t, k, n, fn, **kw):
    import torch
    D = torch.full((n, k), float("inf"), dtype=torch.float64, device="cuda")
    I = torch.full((n, k), -1, dtype=torch.int32, device="cuda")
    rc = fn(t["X"], t["X2"], t["am"], t["X"], t["X2"], t["bm"], D, I, **kw)
    torch.cuda.synchronize()
    return rc, D.cpu().numpy(), I.cpu().numpy()


@pytest.mark.parametrize("mode", [gsknn.MODE_AUTO, gsknn.MODE_EXACT])
def test_noalloc_matches_dknn_dev_two_calls(mode):
    g = np.random.Generator(np.random.PCG64(7 + mode))
    # Two calls in a row, the second smaller; both split their candidates,
    # so both run through the scratch array.
    for (N, d, m, n, k) in [(40000, 64, 60000, 300, 32), (9000, 48, 20000, 70, 12)]:
        need = gsknn.lib().gsknn_dknn_work_bytes(m, n, d, k, mode)
        print(f"[gpu] noalloc mode={mode} m={m} n={n} d={d} k={k} scratch={need}", flush=True)
        assert 0 < need <= gsknn.NOALLOC_SCRATCH_BYTES
        t = _case(g, N, d, m, n, k)
        rc, Dn, In = _rows(t, k, n, gsknn.dknn_noalloc_dev, mode=mode)
        assert rc == 0, gsknn.last_error()
        _, Dr, Ir = _rows(t, k, n, lambda *a, **kw: gsknn.dknn_dev(*a, **kw), mode=mode)
        np.testing.assert_array_equal(In, Ir)
        assert np.array_equal(Dn.view(np.uint8), Dr.view(np.uint8))


def test_noalloc_two_stage_form():
    # Unsplit operand-ring launches (d > 64, enough queries to fill the chip)
    # take the two-stage form, whose key heaps and fallback lists live in the
    # scratch array too: two calls in a row, the second smaller.  The third
    # call's key heaps (200000 queries x 40 entries x 12 bytes) exceed the
    # array, so it runs in one stage there; its rows still equal those of
    # gsknn_dknn_dev, which takes the two-stage form.
    g = np.random.Generator(np.random.PCG64(11))
    L = gsknn.lib()
    for (N, d, m, n, k) in [(60000, 96, 2048, 32768, 32), (50000, 128, 1024, 32768, 24),
                            (250000, 96, 1024, 200000, 32)]:
        two = L.gsknn_dknn_fallback_offset(m, n, d, k, gsknn.MODE_AUTO) != (1 << 64) - 1
        need = L.gsknn_dknn_work_bytes(m, n, d, k, gsknn.MODE_AUTO)
        print(f"[gpu] noalloc two-stage m={m} n={n} d={d} k={k} two={two} need={need}", flush=True)
        assert two
        assert (need <= gsknn.NOALLOC_SCRATCH_BYTES) == (n < 100000)
        t = _case(g, N, d, m, n, k)
        rc, Dn, In = _rows(t, k, n, gsknn.dknn_noalloc_dev)
        assert rc == 0, gsknn.last_error()
        _, Dr, Ir = _rows(t, k, n, lambda *a, **kw: gsknn.dknn_dev(*a, **kw))
        np.testing.assert_array_equal(In, Ir)
        assert np.array_equal(Dn.view(np.uint8), Dr.view(np.uint8))


def test_noalloc_refuses_beyond_scratch():
    import torch
    g = np.random.Generator(np.random.PCG64(3))
    N, d, m, n, k = 4096, 8, 3_000_000, 1000, 200
    need = gsknn.lib().gsknn_dknn_work_bytes(m, n, d, k, gsknn.MODE_AUTO)
    print(f"[gpu] noalloc limit: needs {need} bytes", flush=True)
    assert need > gsknn.NOALLOC_SCRATCH_BYTES
    dev = torch.device("cuda")
    X = torch.from_numpy(g.standard_normal((N, d))).to(dev)
    X2 = (X * X).sum(1)
    am = torch.from_numpy(g.integers(0, N, m).astype(np.int32)).to(dev)
    bm = torch.from_numpy(g.integers(0, N, n).astype(np.int32)).to(dev)
    D = torch.full((n, k), float("inf"), dtype=torch.float64, device=dev)
    I = torch.full((n, k), -1, dtype=torch.int32, device=dev)
    rc = gsknn.dknn_noalloc_dev(X, X2, am, X, X2, bm, D, I)
    torch.cuda.synchronize()
    assert rc == GSKNN_ELIMIT and "scratch" in gsknn.last_error()
    assert torch.all(I == -1)
