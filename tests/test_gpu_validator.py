"""The on-device brute-force validator (rtdbs_bruteforce): first pinned to
the oracle and the golden fixtures, then used as an independent check of the
slide engine at full window sizes, where the CPU oracle is too slow for every
slide.

Bar: bit-exact label, cid and neighbour counts (integer outputs).
"""
import glob
import os

import numpy as np
import pytest

from rtdbstream_amd import workloads as WL
from tests.refcheck import window_at

pytestmark = pytest.mark.gpu


def _mixed(n, dim, seed, spread=6.0, sigma=0.25, noise=0.3):
    rng = np.random.default_rng(seed)
    nb = int(n * (1 - noise))
    c = rng.uniform(0, spread, (6, dim))
    p = c[rng.integers(0, 6, nb)] + rng.normal(0, sigma, (nb, dim))
    q = rng.uniform(0, spread, (n - nb, dim))
    allp = np.concatenate([p, q]).astype(np.float32)
    return allp[rng.permutation(n)]


CASES = [
    # name, points, eps, min_pts
    ("mixed3d", lambda: _mixed(3000, 3, 1), 0.3, 6),
    ("mixed2d", lambda: _mixed(2500, 2, 2), 0.2, 5),
    ("mixed1d", lambda: _mixed(1500, 1, 3), 0.01, 4),
    ("lattice", lambda: (np.random.default_rng(4).integers(0, 30, (2000, 2)) * 0.5).astype(np.float32), 0.5, 3),
    ("duplicates", lambda: np.random.default_rng(5).uniform(0, 2, (50, 3)).astype(np.float32)[
        np.random.default_rng(6).integers(0, 50, 1200)], 0.3, 8),
    ("min_pts_1", lambda: _mixed(800, 3, 7), 0.2, 1),
    ("all_noise", lambda: _mixed(800, 3, 8), 0.05, 500),
    ("one_point", lambda: np.zeros((1, 3), np.float32), 1.0, 1),
]


@pytest.mark.parametrize("name,make,eps,mp", CASES, ids=[c[0] for c in CASES])
def test_bruteforce_matches_oracle(engine, oracle, name, make, eps, mp):
    pts = make()
    label, cid, nn = engine.bruteforce(pts, eps, mp)
    ol, oc = oracle.bruteforce(pts, eps, mp)
    assert np.array_equal(nn, oracle.neighbor_counts(pts, eps))
    assert np.array_equal(label, ol)
    assert np.array_equal(cid, oc)


FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p)[:-4] for p in FIX])
def test_bruteforce_reproduces_fixture_windows(engine, path):
    z = np.load(path)
    dim, W, S, slides, mp = (int(v) for v in z["params"])
    for t in range(slides):
        win = window_at(z["data"], W, S, t)
        label, cid, _ = engine.bruteforce(win, float(z["eps"]), mp)
        assert np.array_equal(label, z["labels"][t]) and np.array_equal(cid, z["cids"][t]), t


def test_bruteforce_rejects_bad_arguments(engine):
    pts = np.zeros((4, 3), np.float32)
    for dim, eps, mp in ((0, 1.0, 1), (4, 1.0, 1), (3, 0.0, 1), (3, float("nan"), 1), (3, 1.0, 0)):
        with pytest.raises(engine.RtdbsError):
            engine.bruteforce(pts.reshape(-1)[: 4 * max(dim, 1)], eps, mp, dim=dim)


def test_engine_matches_bruteforce_walks3d_every_slide(engine):
    """walks3d at full size (W 100,000, S 10,000): all 50 slides."""
    cfg = WL.WORKLOADS["walks3d"]
    W, S, n = cfg["window"], cfg["stride"], cfg["slides"]
    data = WL.generate("walks3d")
    s = engine.Stream(3, W, S, cfg["eps"], cfg["min_pts"])
    s.init(data[:W])
    for t in range(n):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        bl, bc, _ = engine.bruteforce(window_at(data, W, S, t), cfg["eps"], cfg["min_pts"])
        assert np.array_equal(l, bl), "slide %d labels" % t
        assert np.array_equal(c, bc), "slide %d cids" % t
    s.close()


def test_engine_matches_bruteforce_skewed3d_later_slides(engine):
    """skewed3d at full size (W 1,000,000): slides 2-4, beyond the two
    slides the CPU oracle checks."""
    cfg = WL.WORKLOADS["skewed3d"]
    W, S = cfg["window"], cfg["stride"]
    data = WL.generate("skewed3d", n=W + 5 * S)
    s = engine.Stream(3, W, S, cfg["eps"], cfg["min_pts"])
    s.init(data[:W])
    for t in range(5):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        if t >= 2:
            bl, bc, _ = engine.bruteforce(window_at(data, W, S, t), cfg["eps"], cfg["min_pts"])
            assert np.array_equal(l, bl), "slide %d labels" % t
            assert np.array_equal(c, bc), "slide %d cids" % t
    s.close()


def test_engine_matches_bruteforce_batched_streams(engine):
    """Eight of the 64 batch streams (2-D and 3-D, varied W/Eps/MinPts),
    slid together in one batch for 5 slides."""
    specs = WL.multi_stream_specs()[:8]
    datas = [WL.stream_points(dict(sp, slides=5)) for sp in specs]
    streams = []
    for sp, d in zip(specs, datas):
        st = engine.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        st.init(d[:sp["window"]])
        streams.append(st)
    b = engine.Batch(streams, use_graph=True)
    for t in range(5):
        strides = [d[sp["window"] + t * sp["stride"]: sp["window"] + (t + 1) * sp["stride"]]
                   for sp, d in zip(specs, datas)]
        labels, cids = b.slide(strides)
        for i, (sp, d) in enumerate(zip(specs, datas)):
            bl, bc, _ = engine.bruteforce(window_at(d, sp["window"], sp["stride"], t), sp["eps"], sp["min_pts"],
                                          dim=sp["dim"])
            assert np.array_equal(labels[i], bl) and np.array_equal(cids[i], bc), (i, t)
    b.close()
    for st in streams:
        st.close()
