"""Algorithmic byte counts for the roofline of the range-count kernel.

B_rc = 4*d * sum_{c in Cq} sum_{c' in N(c)} n_c'  +  (4*d + 4) * |Q|

over an accounting grid of side exactly Eps (N(c) = the 3^d cells around c,
n_c = window points in c): each query cell stages its stencil once, reads
its queries' coordinates and writes one flag per query. Q is the set of points
whose core status the kernel must compute: every window point except those in
cells the engine classifies as dense (>= MinPts points in an interior cell of
its dense-safe grid, all core by construction). Cq = accounting cells holding
at least one point of Q. The figure depends only on the window contents, not
on the engine's tile shapes.
"""
import numpy as np


def _cell_ids(p, lo, side, n):
    ix = np.floor((p.astype(np.float64) - lo[None, :]) / side).astype(np.int64)
    ix = np.clip(ix, 0, np.asarray(n)[None, :] - 1)
    return ix


def dense_mask(points, eps, min_pts, info):
    """Points the engine treats as core without counting (same grid)."""
    p = np.asarray(points, np.float32)
    d = p.shape[1]
    if not info.get("dense_shortcut"):
        return np.zeros(len(p), bool)
    side = info["cell_side"]
    lo = np.asarray(info["lo"][:d], np.float64)
    n = np.asarray(info["grid"][:d], np.int64)
    inv = 1.0 / side
    ix = np.floor((p.astype(np.float64) - lo[None, :]) * inv)
    ix = np.clip(np.where(ix >= 0, ix, 0), 0, n[None, :] - 1).astype(np.int64)
    key = np.zeros(len(p), np.int64)
    for a in range(d):
        key = key * n[a] + ix[:, a]
    uk, inv_idx, cnt = np.unique(key, return_inverse=True, return_counts=True)
    interior = np.all((ix > 0) & (ix < n[None, :] - 1), axis=1)
    return (cnt[inv_idx] >= min_pts) & interior


def range_count_bytes(points, eps, q_mask):
    """B_rc for one window (bytes)."""
    p = np.asarray(points, np.float32).astype(np.float64)
    d = p.shape[1]
    lo = p.min(axis=0)
    ix = np.floor((p - lo[None, :]) / float(eps)).astype(np.int64)
    n = ix.max(axis=0) + 3
    ix = ix + 1  # room for the -1 neighbour
    key = np.zeros(len(p), np.int64)
    for a in range(d):
        key = key * n[a] + ix[:, a]
    uk, cnt = np.unique(key, return_counts=True)
    qcells = np.unique(key[q_mask])
    # sum over query cells of the points in their 3^d stencil
    offs = np.array(np.meshgrid(*[[-1, 0, 1]] * d, indexing="ij")).reshape(d, -1).T
    mult = np.ones(d, np.int64)
    for a in range(d - 2, -1, -1):
        mult[a] = mult[a + 1] * n[a + 1]
    total = 0
    for o in offs:
        nk = qcells + int(np.dot(o, mult))
        pos = np.searchsorted(uk, nk)
        pos = np.minimum(pos, len(uk) - 1)
        hit = uk[pos] == nk
        total += int(cnt[pos[hit]].sum())
    nq = int(q_mask.sum())
    return 4 * d * total + (4 * d + 4) * nq
