"""The allocation-free entry point piafs_tridiag_host_noalloc against the
allocating piafs_tridiag_host on the same systems, bit for bit, on two calls
in a row (the second smaller, so that it runs on scratch the first one left
behind), the zero-pivot status, and the size limit of its scratch array."""
import ctypes

import numpy as np
import pytest

import piafs_amd

pytestmark = pytest.mark.gpu


def _systems(rng, n, ns):
    a = rng.uniform(-1, 1, (n, ns))
    c = rng.uniform(-1, 1, (n, ns))
    b = 3.0 + rng.uniform(0, 1, (n, ns))
    x = rng.uniform(-1, 1, (n, ns))
    a[0] = 0.0
    c[-1] = 0.0
    return a, b, c, x


def _call(fn, a, b, c, x):
    a, b, c, x = (np.ascontiguousarray(t, dtype=np.float64).copy() for t in (a, b, c, x))
    n, ns = x.shape
    p = lambda t: t.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = fn(p(a), p(b), p(c), p(x), n, ns)
    return rc, x


def test_noalloc_matches_allocating_entry_point():
    L = piafs_amd.lib()
    rng = np.random.default_rng(2027)
    for n, ns in ((1025, 700), (200, 33)):   # second call smaller than the first
        a, b, c, x = _systems(rng, n, ns)
        rc0, ref = _call(L.piafs_tridiag_host, a, b, c, x)
        rc1, got = _call(L.piafs_tridiag_host_noalloc, a, b, c, x)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got, ref), (n, ns)
    # a zero pivot is reported (-1) as by the allocating entry point; the next
    # call starts from a cleared flag
    a, b, c, x = _systems(rng, 64, 5)
    b[0, 2] = 0.0
    assert _call(L.piafs_tridiag_host_noalloc, a, b, c, x)[0] == -1
    assert _call(L.piafs_tridiag_host, a, b, c, x)[0] == -1
    b[0, 2] = 3.0
    assert _call(L.piafs_tridiag_host_noalloc, a, b, c, x)[0] == 0
    # beyond the scratch array: refused with 2, nothing launched
    cap = piafs_amd.NOALLOC_SCRATCH_DOUBLES
    n = 4096
    ns = (cap - 1) // 4 // n + 1
    z = np.zeros((n, ns))
    assert _call(L.piafs_tridiag_host_noalloc, z, z + 1.0, z, z)[0] == 2
