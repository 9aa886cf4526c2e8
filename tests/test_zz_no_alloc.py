"""rtdbs_bruteforce_noalloc (the validator without any allocation) against
rtdbs_bruteforce, through the C ABI: bit-exact labels, ids and neighbour
counts on two calls in a row, the second on a smaller input, so the shared
per-GPU scratch array holds the first call's contents when the second runs.
Sizes beyond the scratch array are refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _noalloc(R, pts, eps, mp, want_nn):
    pts = np.ascontiguousarray(pts, np.float32)
    n, dim = pts.shape
    d_pts = R.DeviceBuffer(pts.nbytes)
    d_pts.upload(pts)
    d_l, d_c, d_n = R.DeviceBuffer(n * 4), R.DeviceBuffer(n * 4), R.DeviceBuffer(n * 4)
    R._check(R.lib().rtdbs_bruteforce_noalloc(dim, eps, mp, ctypes.c_void_p(d_pts.addr()), n,
                                              ctypes.c_void_p(d_l.addr()), ctypes.c_void_p(d_c.addr()),
                                              ctypes.c_void_p(d_n.addr()) if want_nn else None))
    out = [b.download(np.empty(n, np.int32)) for b in (d_l, d_c)]
    out.append(d_n.download(np.empty(n, np.int32)) if want_nn else None)
    return out


@pytest.mark.parametrize("want_nn", [True, False])
def test_noalloc_equals_allocating_validator(engine, want_nn):
    R = engine
    cases = [
        (WL.walks3d(6000, seed=41, n_walks=50, box=(20.0, 20.0, 5.0)), 0.7, 10),  # first call
        (WL.blobs2d(1500, seed=42), 0.8, 5),                                        # then smaller, other dim
        (WL.walks3d(700, seed=43, n_walks=20, box=(10.0, 10.0, 3.0)), 0.9, 4),
    ]
    for pts, eps, mp in cases:
        l0, c0, n0 = R.bruteforce(pts, eps, mp)
        l1, c1, n1 = _noalloc(R, pts, eps, mp, want_nn)
        assert np.array_equal(l1, l0) and np.array_equal(c1, c0)
        if want_nn:
            assert np.array_equal(n1, n0)


def test_noalloc_refuses_beyond_scratch(engine):
    R = engine
    rc = R.lib().rtdbs_bruteforce_noalloc(3, 1.0, 5, ctypes.c_void_p(16), (1 << 20) + 1, ctypes.c_void_p(16),
                                          ctypes.c_void_p(16), None)
    assert rc == R.EINVAL
    assert b"scratch" in R.lib().rtdbs_last_error()
