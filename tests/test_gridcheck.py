"""The numpy grid helper that pins the work counters, checked by brute force
on small inputs: a point's stencil holds every point whose cell differs from
its own by at most k on each axis, and the cell count is the number of
distinct cells."""
import numpy as np
import pytest

from tests.gridcheck import cell_coords, cells_and_stencils


@pytest.mark.parametrize("dim,k", [(1, 1), (2, 2), (3, 2), (3, 3), (2, 1)])
def test_stencils_by_brute_force(dim, k):
    rng = np.random.default_rng(dim * 10 + k)
    pts = rng.uniform(-1.0, 4.0, (400, dim)).astype(np.float32)
    pts[:5] = np.float32(100.0)  # far outside the grid: clamped into the last cells
    pts[5:8] = np.float32(-50.0)
    info = {"cell_side": 0.3, "lo": [0.0, 0.0, 0.0], "grid": [10, 11, 9][:dim] + [1] * (3 - dim), "cell_div": k}
    nc, sten = cells_and_stencils(pts, info)
    ix, n = cell_coords(pts, info)
    assert np.all(ix >= 0) and np.all(ix < n[None, :])
    diff = np.abs(ix[:, None, :] - ix[None, :, :]).max(axis=2)
    assert np.array_equal(sten, (diff <= k).sum(axis=1))
    assert nc == len({tuple(r) for r in ix})


def test_binning_matches_engine_rule():
    # floor((x - lo) * (1 / side)) in double, negative and NaN to 0, clamped to n - 1
    info = {"cell_side": 0.5, "lo": [1.0], "grid": [4, 1, 1], "cell_div": 1}
    pts = np.array([[1.0], [1.49], [1.5], [0.99], [np.nan], [2.99], [3.0], [1e9]], np.float32)
    ix, _ = cell_coords(pts, info)
    assert ix[:, 0].tolist() == [0, 0, 1, 0, 0, 3, 3, 3]
