"""Golden fixtures (tests/golden/*.npz, written and sklearn-checked by
tests/golden/make_fixtures.py): the oracle must reproduce them on the CPU and
the HIP engine on the GPU, bit for bit."""
import glob
import os

import numpy as np
import pytest

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    z = np.load(path)
    dim, W, S, slides, mp = (int(v) for v in z["params"])
    return z["data"], z["labels"], z["cids"], dim, W, S, slides, mp, float(z["eps"])


def test_fixtures_present():
    assert len(FIX) >= 6


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p)[:-4] for p in FIX])
def test_oracle_reproduces_fixture(oracle, path):
    data, labels, cids, dim, W, S, slides, mp, eps = _load(path)
    for method in ("bruteforce", "grid"):
        l, c = oracle.stream(data, W, S, eps, mp, slides, method=method)
        assert np.array_equal(l, labels) and np.array_equal(c, cids), method


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p)[:-4] for p in FIX])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_engine_reproduces_fixture(engine, path, k):
    data, labels, cids, dim, W, S, slides, mp, eps = _load(path)
    s = engine.Stream(dim, W, S, eps, mp, cell_div=k)
    s.init(data[:W])
    for t in range(slides):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        assert np.array_equal(l, labels[t]), (t, int(np.flatnonzero(l != labels[t])[0]))
        assert np.array_equal(c, cids[t]), t
    s.close()
