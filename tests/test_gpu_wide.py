"""Wide and drifting streams: the grid has no per-axis cap, so a window
spanning more than 10^5 Eps on one axis, or a stream drifting away from its
first window, keeps every point in its own cell (nothing is clamped into the
border cells). Labels stay bit-exact against the all-pairs validator, and the
range count tests no candidate beyond the points' stencils (the work bound a
clamped grid breaks: far points piled into one border cell)."""
import numpy as np
import pytest

from tests.gridcheck import cells_and_stencils
from tests.refcheck import window_at

pytestmark = pytest.mark.gpu


def _blobs_along_x(n, seed, length, n_blobs=400, sigma=0.4):
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.0, length, n_blobs)
    which = rng.integers(0, n_blobs, n)
    p = np.empty((n, 3))
    p[:, 0] = cx[which] + rng.normal(0, sigma, n)
    p[:, 1] = rng.normal(0, sigma, n)
    p[:, 2] = rng.normal(0, sigma, n)
    noise = rng.random(n) < 0.1
    p[noise, 0] = rng.uniform(0.0, length, noise.sum())
    return p.astype(np.float32)


def _check_stream(engine, data, W, S, eps, mp, slides):
    s = engine.Stream(3, W, S, eps, mp)
    s.init(data[:W])
    s.set_stats(True)
    for t in range(slides):
        label, cid = s.slide(data[W + t * S: W + (t + 1) * S])
        win = window_at(data, W, S, t)
        bl, bc, _ = engine.bruteforce(win, eps, mp)
        assert np.array_equal(label, bl), ("labels", t)
        assert np.array_equal(cid, bc), ("cids", t)
    info = s.info()
    st = s.stats()
    s.close()
    return info, st


def _stencil_bound(data, W, S, slides, info, eps, mp):
    """Sum over the non-dense points of every window of their stencil sizes:
    the most candidates the range count may test."""
    from rtdbstream_amd import roofline as RF

    hi = 0
    for t in range(slides):
        win = window_at(data, W, S, t)
        dense = RF.dense_mask(win, eps, mp, info)
        _, sten = cells_and_stencils(win, info)
        hi += int(sten[~dense].sum())
    return hi


def test_extent_of_1e5_eps_on_one_axis(engine):
    eps, mp, W, S, slides = 0.5, 10, 20000, 2000, 3
    length = 1.2e5 * eps
    data = _blobs_along_x(W + slides * S, seed=61, length=length)
    info, st = _check_stream(engine, data, W, S, eps, mp, slides)
    # every point binned without clamping: the grid covers the data on x
    span = info["grid"][0] * info["cell_side"]
    assert info["lo"][0] <= float(data[:, 0].min()) and info["lo"][0] + span >= float(data[:, 0].max())
    assert info["grid"][0] * info["cell_side"] >= 1e5 * eps
    # the count stays within the points' own stencils
    assert st["count_tests"] <= _stencil_bound(data, W, S, slides, info, eps, mp)


def test_drifting_stream_stays_unclamped(engine):
    # blobs drifting along +x by 1.5 units per slide: after 12 slides the
    # window sits well beyond the first window's bounds
    eps, mp, W, S, slides = 0.6, 8, 8000, 800, 12
    n = W + slides * S
    t = np.arange(n) / S  # slide index of each point
    base = _blobs_along_x(n, seed=63, length=40.0, n_blobs=30)
    base[:, 0] += 1.5 * t
    info, st = _check_stream(engine, base, W, S, eps, mp, slides)
    span = info["grid"][0] * info["cell_side"]
    assert info["lo"][0] + span >= float(base[:, 0].max())
    assert st["count_tests"] <= _stencil_bound(base, W, S, slides, info, eps, mp)
