"""Pins the CPU oracle (oracle/rtdbs_oracle.c) before it is trusted.

The reference ships no tests, fixtures or golden vectors for this path (its
repository has no test directory), and its code needs CUDA/OptiX, so it
cannot be built here. The oracle is pinned instead by:
  * scikit-learn's DBSCAN on the exact FP32 squared distances (independent
    implementation; see tests/refcheck.py for why it is exact),
  * closed forms on lattices with exact ties at d2 == eps2,
  * connected components of the core graph (scipy) for the canonical ids,
  * the brute-force and grid oracles agreeing bit for bit,
  * the ring semantics of the reference's slide loop.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from tests.refcheck import eps2_fp32, reference_check, sklearn_reference, sqdist_matrix_fp32, window_at


def _data(kind, n, dim, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(0, 10, (n, dim)).astype(np.float32)
    if kind == "blobs":
        c = rng.uniform(0, 10, (5, dim))
        p = c[rng.integers(0, 5, n)] + rng.normal(0, 0.4, (n, dim))
        return p.astype(np.float32)
    if kind == "lattice":  # many exact ties at distance 0.5 multiples
        g = rng.integers(0, 12, (n, dim)).astype(np.float32) * np.float32(0.5)
        return g
    raise ValueError(kind)


CASES = [
    (kind, n, dim, eps, mp, seed)
    for seed, (kind, n, dim, eps, mp) in enumerate([
        ("uniform", 600, 2, 0.6, 4),
        ("uniform", 600, 3, 1.2, 5),
        ("uniform", 500, 1, 0.05, 3),
        ("blobs", 700, 2, 0.3, 5),
        ("blobs", 700, 3, 0.5, 10),
        ("blobs", 400, 1, 0.02, 4),
        ("lattice", 500, 2, 0.5, 3),   # eps2 == exact tie distance: ties are not neighbours
        ("lattice", 500, 3, 0.5, 2),
        ("lattice", 500, 2, 0.7072, 4),  # diagonal 0.5*sqrt(2)=0.70711 just inside
        ("blobs", 300, 2, 0.3, 1),     # min_pts = 1: every point core
        ("blobs", 300, 3, 0.5, 400),   # min_pts > n: all noise
    ])
]


@pytest.mark.parametrize("kind,n,dim,eps,mp,seed", CASES)
def test_bruteforce_matches_sklearn(oracle, kind, n, dim, eps, mp, seed):
    pts = _data(kind, n, dim, seed)
    label, cid = oracle.bruteforce(pts, eps, mp)
    ref_label, ref_cid = sklearn_reference(pts, eps, mp)
    assert np.array_equal(label, ref_label)
    assert np.array_equal(cid, ref_cid)
    assert reference_check(pts, eps, label, cid, ref_label, ref_cid) == []


@pytest.mark.parametrize("kind,n,dim,eps,mp,seed", CASES)
def test_grid_matches_bruteforce(oracle, kind, n, dim, eps, mp, seed):
    pts = _data(kind, n, dim, seed)
    a = oracle.bruteforce(pts, eps, mp)
    b = oracle.grid(pts, eps, mp)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_grid_matches_bruteforce_larger(oracle):
    rng = np.random.default_rng(7)
    for dim, eps, mp in [(2, 0.25, 6), (3, 0.6, 8)]:
        c = rng.uniform(0, 20, (30, dim))
        p = (c[rng.integers(0, 30, 4000)] + rng.normal(0, 0.5, (4000, dim))).astype(np.float32)
        a = oracle.bruteforce(p, eps, mp)
        b = oracle.grid(p, eps, mp, threads=4)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_neighbor_counts_match_fp32_matrix(oracle):
    pts = _data("lattice", 300, 2, 3)
    nn = oracle.neighbor_counts(pts, 0.5)
    D = sqdist_matrix_fp32(pts)
    assert np.array_equal(nn, (D < eps2_fp32(0.5)).sum(axis=1))  # self included (d = 0)


def test_eps2_is_fp32_product(oracle):
    for e in [0.8, 0.5, 1.0, 0.035, 1.3, 0.1]:
        assert np.float32(oracle.eps2(e)) == eps2_fp32(e)


def _lattice(m):
    ij = np.array([(i, j) for i in range(m) for j in range(m)], np.float32)
    return ij


def test_lattice_ties_are_not_neighbours(oracle):
    # unit lattice, eps = 1: every nearest pair sits at d2 == eps2 exactly
    pts = _lattice(6)
    l1, c1 = oracle.bruteforce(pts, 1.0, 1)
    assert np.all(l1 == 0) and np.array_equal(c1, np.arange(36))  # each point its own cluster
    l2, c2 = oracle.bruteforce(pts, 1.0, 2)
    assert np.all(l2 == 2) and np.array_equal(c2, np.arange(36))


def test_lattice_closed_form(oracle):
    # eps just above 1: 4-neighbours count, diagonals (d2 = 2) do not.
    m = 7
    pts = _lattice(m)
    eps = float(np.nextafter(np.float32(1.0), np.float32(2.0)))
    label, cid = oracle.bruteforce(pts, eps, 5)
    idx = lambda i, j: i * m + j
    exp_label = np.full(m * m, 2, np.int32)
    exp_cid = np.arange(m * m, dtype=np.int32)
    interior = [(i, j) for i in range(1, m - 1) for j in range(1, m - 1)]
    root = min(idx(i, j) for i, j in interior)
    for i, j in interior:
        exp_label[idx(i, j)] = 0
        exp_cid[idx(i, j)] = root
    for i in range(m):
        for j in range(m):
            on_edge = i in (0, m - 1) or j in (0, m - 1)
            corner = i in (0, m - 1) and j in (0, m - 1)
            if on_edge and not corner:
                exp_label[idx(i, j)] = 1
                exp_cid[idx(i, j)] = root
    assert np.array_equal(label, exp_label)
    assert np.array_equal(cid, exp_cid)


def test_shared_border_takes_minimum_cluster_id(oracle):
    # 1-D: cluster R (slots 0-3) and cluster L (slots 4-7), border b (slot 8)
    # within eps of a core of each. b must take R's id 0, not L's 4.
    pts = np.array([1.0, 1.1, 1.2, 1.3, 0.0, 0.1, 0.2, 0.3, 0.65], np.float32)
    label, cid = oracle.bruteforce(pts, 0.36, 4)
    assert list(label) == [0, 0, 0, 0, 0, 0, 0, 0, 1]
    assert list(cid) == [0, 0, 0, 0, 4, 4, 4, 4, 0]
    # the same with the slots of the two clusters swapped
    pts2 = np.concatenate([pts[4:8], pts[0:4], pts[8:]])
    label2, cid2 = oracle.bruteforce(pts2, 0.36, 4)
    assert list(cid2) == [0, 0, 0, 0, 4, 4, 4, 4, 0]
    g = oracle.grid(pts, 0.36, 4)
    assert np.array_equal(g[0], label) and np.array_equal(g[1], cid)


def test_canonical_ids_are_min_core_slot_of_component(oracle):
    rng = np.random.default_rng(11)
    pts = (rng.uniform(0, 8, (900, 2))).astype(np.float32)
    eps, mp = 0.45, 4
    label, cid = oracle.grid(pts, eps, mp)
    core = label == 0
    D = sqdist_matrix_fp32(pts) < eps2_fp32(eps)
    ci = np.flatnonzero(core)
    A = csr_matrix(D[np.ix_(ci, ci)])
    ncomp, comp = connected_components(A, directed=False)
    for k in range(ncomp):
        members = ci[comp == k]
        assert np.all(cid[members] == members.min())
    # noise keeps its own slot; borders point at a cluster adjacent to them
    noise = np.flatnonzero(label == 2)
    assert np.array_equal(cid[noise], noise)
    for i in np.flatnonzero(label == 1):
        adj = ci[D[i, ci]]
        assert cid[i] == cid[adj].min()


def test_strictness_is_observable(oracle):
    # A planted '<=' (sklearn fed eps2 itself) must disagree on tie data,
    # so the tie cases above really test the strict comparison.
    from sklearn.cluster import DBSCAN

    pts = _lattice(6)
    D = sqdist_matrix_fp32(pts).astype(np.float64)
    m = DBSCAN(eps=float(eps2_fp32(1.0)), min_samples=2, metric="precomputed").fit(D)
    assert np.all(m.labels_ >= 0)  # with <= every lattice point is core
    label, _ = oracle.bruteforce(pts, 1.0, 2)
    assert np.all(label == 2)


def test_stream_ring_semantics(oracle):
    rng = np.random.default_rng(3)
    W, S, slides = 120, 30, 9
    data = (rng.uniform(0, 3, (W + slides * S, 2))).astype(np.float32)
    labels, cids = oracle.stream(data, W, S, 0.3, 4, slides, method="bruteforce")
    for t in range(slides):
        win = window_at(data, W, S, t)
        l, c = oracle.bruteforce(win, 0.3, 4)
        assert np.array_equal(labels[t], l) and np.array_equal(cids[t], c)
    lg, cg = oracle.stream(data, W, S, 0.3, 4, slides, method="grid")
    assert np.array_equal(lg, labels) and np.array_equal(cg, cids)


def test_stream_rejects_bad_ring(oracle):
    data = np.zeros((100, 2), np.float32)
    with pytest.raises(RuntimeError):
        oracle.stream(data, 50, 7, 0.3, 4, 2)  # S does not divide W
    with pytest.raises(RuntimeError):
        oracle.stream(data, 50, 10, 0.3, 4, 10)  # not enough points
