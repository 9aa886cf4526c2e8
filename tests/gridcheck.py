"""Independent numpy restatement of the engine's grid quantities, for pinning
the work counters (rtdbs_stats) to values computed without the engine.

The engine bins each FP32 coordinate as floor((double(x) - lo) * inv_side),
negative or NaN to 0, clamped to n - 1 (the grid the stream reports through
rtdbs_stream_get_info). A point's stencil is the (2k+1)^d box of cells around
its own, clipped to the grid: every candidate the range count, the merge and
the output kernel may test for it.
"""
import itertools

import numpy as np


def cell_coords(points, info):
    p = np.asarray(points, np.float32)
    if p.ndim == 1:
        p = p[:, None]
    d = p.shape[1]
    inv = 1.0 / float(info["cell_side"])
    lo = np.asarray(info["lo"][:d], np.float64)
    n = np.asarray(info["grid"][:d], np.int64)
    ix = np.floor((p.astype(np.float64) - lo[None, :]) * inv)
    ix = np.where(ix >= 0.0, ix, 0.0)  # negative or NaN
    ix = np.minimum(ix, (n - 1)[None, :].astype(np.float64))
    return ix.astype(np.int64), n


def _linear(ix, n):
    key = np.zeros(len(ix), np.int64)
    for a in range(ix.shape[1]):
        key = key * n[a] + ix[:, a]
    return key


def cells_and_stencils(points, info):
    """(number of non-empty cells, per-point stencil candidate counts)."""
    ix, n = cell_coords(points, info)
    d = ix.shape[1]
    k = int(info["cell_div"])
    key = _linear(ix, n)
    uk, first, inv, cnt = np.unique(key, return_index=True, return_inverse=True, return_counts=True)
    uc = ix[first]
    total = np.zeros(len(uk), np.int64)
    for off in itertools.product(range(-k, k + 1), repeat=d):
        c2 = uc + np.asarray(off, np.int64)[None, :]
        ok = np.all((c2 >= 0) & (c2 < n[None, :]), axis=1)
        idx = np.flatnonzero(ok)
        k2 = _linear(c2[idx], n)
        pos = np.minimum(np.searchsorted(uk, k2), len(uk) - 1)
        hit = uk[pos] == k2
        np.add.at(total, idx[hit], cnt[pos[hit]])
    return len(uk), total[inv.reshape(-1)]
