"""Writes the small golden fixtures of the slide step (tests/golden/*.npz).

Each fixture is a seeded point stream plus the expected per-slide `label` and
`cid` arrays from the oracle's brute-force restatement. Before a fixture is
saved, every slide is cross-checked against scikit-learn's DBSCAN on the exact
FP32 squared distances (tests/refcheck.py), so the vectors do not rest on the
oracle alone. Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.refcheck import sklearn_reference, window_at  # noqa: E402

CASES = [
    # name, dim, W, S, slides, eps, min_pts, kind, seed
    ("blobs_w64_d2", 2, 64, 16, 4, 0.4, 3, "blobs", 1),
    ("blobs_w512_d3", 3, 512, 128, 4, 0.6, 5, "blobs", 2),
    ("uniform_w512_d1", 1, 512, 64, 3, 0.02, 3, "uniform", 3),
    ("lattice_w512_d2", 2, 512, 128, 3, 0.5, 3, "lattice", 4),
    ("blobs_w2048_d2", 2, 2048, 256, 3, 0.25, 6, "blobs", 5),
    ("lattice_w2048_d3", 3, 2048, 512, 3, 0.5, 4, "lattice", 6),
]


def points(kind, n, dim, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(0, 8, (n, dim)).astype(np.float32)
    if kind == "lattice":  # exact ties at d2 == eps2
        return (rng.integers(0, 10, (n, dim)).astype(np.float32) * np.float32(0.5))
    c = rng.uniform(0, 6, (5, dim))
    return (c[rng.integers(0, 5, n)] + rng.normal(0, 0.35, (n, dim))).astype(np.float32)


def main():
    for name, dim, W, S, slides, eps, mp, kind, seed in CASES:
        data = points(kind, W + slides * S, dim, seed)
        labels, cids = O.stream(data, W, S, eps, mp, slides, method="bruteforce")
        for t in range(slides):
            sl, sc = sklearn_reference(window_at(data, W, S, t), eps, mp)
            assert np.array_equal(sl, labels[t]) and np.array_equal(sc, cids[t]), (name, t)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), data=data, labels=labels, cids=cids,
                            params=np.array([dim, W, S, slides, mp], np.int32), eps=np.float32(eps))
        print("wrote", name, np.bincount(labels.ravel(), minlength=3))


if __name__ == "__main__":
    main()
