"""lg_diffuse_noalloc: the diffusion entry point that allocates nothing (its
scratch is a fixed per-GPU array inside the library) gives the bytes of
lg_diffuse, called through the C ABI, on two calls in a row with the second
smaller than the first, and refuses sizes beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph(rng, n):
    """A symmetric optical-depth matrix with +INF for unconnected pairs (ld = n
    rounded up to the tile), zero diagonal, and a light vector."""
    ld = ((n + 127) // 128) * 128
    D = np.full((ld, ld), np.inf, np.float32)
    m = rng.uniform(0.05, 4.0, size=(n, n)).astype(np.float32)
    m = np.minimum(m, m.T)
    m[rng.uniform(size=(n, n)) < 0.3] = np.inf
    m = np.minimum(m, m.T)
    np.fill_diagonal(m, 0.0)
    D[:n, :n] = m
    L = rng.uniform(0.0, 2.0, size=(n, 3)).astype(np.float32)
    return D, L, ld


def _call(lga, ctx, name, n, db, ld, sigma, lb, ob):
    sig = (C.c_double * 3)(*sigma)
    rc = getattr(lga.lib(), name)(ctx.handle, n, db.ptr, ld, sig, lb.ptr, ob.ptr)
    return rc


def test_noalloc_equals_diffuse_twice(lga, ctx):
    rng = np.random.default_rng(17)
    sigma = (1.1, 1.0, 0.9)
    for n in (3000, 700):  # second call smaller: stale scratch beyond n must not matter
        D, L, ld = _graph(rng, n)
        db, lb = ctx.upload(D), ctx.upload(L)
        o1, o2 = ctx.alloc(n * 12), ctx.alloc(n * 12)
        assert _call(lga, ctx, "lg_diffuse", n, db, ld, sigma, lb, o1) == 0
        before = ctx.bytes_in_use()
        assert _call(lga, ctx, "lg_diffuse_noalloc", n, db, ld, sigma, lb, o2) == 0
        ctx.sync()
        assert ctx.bytes_in_use() == before
        want = o1.read(np.float32, (n, 3))
        got = o2.read(np.float32, (n, 3))
        np.testing.assert_array_equal(got, want)
        assert np.all(np.isfinite(got)) and got.max() > 0
        for b in (db, lb, o1, o2):
            b.free()


def test_noalloc_rejects_sizes_beyond_scratch(lga, ctx):
    n = 65537
    dummy = ctx.alloc(256)
    sig = (C.c_double * 3)(1.0, 1.0, 1.0)
    rc = lga.lib().lg_diffuse_noalloc(ctx.handle, n, dummy.ptr, n, sig, dummy.ptr, dummy.ptr)
    assert rc == lga._abi.LG_ERR_INVALID
    assert b"scratch limit" in lga.lib().lg_last_error()
    dummy.free()
