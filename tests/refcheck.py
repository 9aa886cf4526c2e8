"""Shared checking helpers for the tests.

`sklearn_reference` derives the expected labels with scikit-learn's DBSCAN
(an independent implementation) on the exact FP32 squared distances the
reference computes: D[i, j] = fl32(((dx*dx + dy*dy) + dz*dz)) and eps_sk =
the largest double below fl32(eps*eps), so sklearn's `D <= eps_sk` equals the
reference's strict `d2 < eps2` (util.cpp:314). sklearn counts the point itself
(min_samples includes it, as util.cpp:308 does), seeds clusters at ascending
core index and gives a border point to the first cluster that reaches it -
the rule of util.cpp:325-351.
"""
import numpy as np


def sqdist_matrix_fp32(pts):
    p = np.asarray(pts, dtype=np.float32)
    if p.ndim == 1:
        p = p[:, None]
    d = None
    for a in range(p.shape[1]):
        da = p[:, None, a] - p[None, :, a]  # float32
        t = da * da
        d = t if d is None else d + t
    return d


def eps2_fp32(eps):
    e = np.float32(eps)
    return np.float32(e * e)


def sklearn_reference(pts, eps, min_pts):
    from sklearn.cluster import DBSCAN

    D = sqdist_matrix_fp32(pts).astype(np.float64)
    e2 = float(eps2_fp32(eps))
    eps_sk = np.nextafter(e2, -np.inf)
    m = DBSCAN(eps=eps_sk, min_samples=int(min_pts), metric="precomputed").fit(D)
    n = len(D)
    core = np.zeros(n, bool)
    core[m.core_sample_indices_] = True
    lab = m.labels_
    label = np.full(n, 2, np.int32)
    cid = np.arange(n, dtype=np.int32)
    canon = {}
    for i in np.flatnonzero(core):
        canon.setdefault(lab[i], i)  # ascending i: first core = minimum core index
    for i in range(n):
        if lab[i] >= 0:
            label[i] = 0 if core[i] else 1
            cid[i] = canon[lab[i]]
    return label, cid


def reference_check(pts, eps, label, cid, ref_label, ref_cid):
    """The reference's own acceptance rules (util.cpp:399-438): labels equal,
    core cids equal, every border has an eps-neighbour j != i with the same cid.
    Returns a list of violations (empty = accepted)."""
    bad = []
    if not np.array_equal(label, ref_label):
        i = int(np.flatnonzero(label != ref_label)[0])
        bad.append("label[%d] = %d, expected %d" % (i, label[i], ref_label[i]))
        return bad
    core = label == 0
    if not np.array_equal(cid[core], ref_cid[core]):
        i = int(np.flatnonzero(core & (cid != ref_cid))[0])
        bad.append("core cid[%d] = %d, expected %d" % (i, cid[i], ref_cid[i]))
    e2 = eps2_fp32(eps)
    p = np.asarray(pts, np.float32)
    if p.ndim == 1:
        p = p[:, None]
    for i in np.flatnonzero(label == 1):
        d = None
        for a in range(p.shape[1]):
            da = p[:, a] - p[i, a]
            d = da * da if d is None else d + da * da
        nb = (d < e2)
        nb[i] = False
        if not np.any(cid[nb] == cid[i]):
            bad.append("border %d has no neighbour in cluster %d" % (i, cid[i]))
            break
    return bad


def window_at(data, W, S, t):
    """Slot-indexed window after slide t (0-based) of the reference ring
    (dbscan.cpp:2289: stride t overwrites slots [u*S, u*S+S), u = t mod W/S)."""
    win = np.array(data[:W], copy=True)
    units = W // S
    for k in range(t + 1):
        u = k % units
        win[u * S:(u + 1) * S] = data[W + k * S: W + (k + 1) * S]
    return win
