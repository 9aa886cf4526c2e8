"""Work counters (rtdbs_stats), pinned to numbers computed without the engine
(tests/gridcheck.py, rtdbstream_amd/roofline.py on the stream's own grid):

- cells: non-empty grid cells of every window;
- dense_points: points of interior cells holding >= MinPts points;
- border_tests: every non-core point tests its whole stencil (no early exit),
  so the total is exactly the sum of the non-core points' stencil sizes;
- count_tests: non-core queries test their whole stencil (those whose
  stencil holds fewer than MinPts points may test nothing), core queries stop
  at MinPts or later, no query tests beyond its stencil;
- and counting must not change any label.
"""
import numpy as np
import pytest

from rtdbstream_amd import roofline as RF
from rtdbstream_amd import workloads as WL
from tests.gridcheck import cells_and_stencils
from tests.refcheck import window_at

pytestmark = pytest.mark.gpu


def expected_counts(windows, labels, eps, mp, info):
    """Exact and bounding counter values over a run's windows."""
    e = dict(cells=0, dense_points=0, border_tests=0, count_lo=0, count_hi=0, queries=0)
    for win, lab in zip(windows, labels):
        nc, sten = cells_and_stencils(win, info)
        dense = RF.dense_mask(win, eps, mp, info)
        core = lab == 0
        assert not np.any(dense & ~core)  # dense points are core
        e["cells"] += nc
        e["dense_points"] += int(dense.sum())
        e["queries"] += int((~dense).sum())
        e["border_tests"] += int(sten[~core].sum())
        # a non-core query whose stencil holds fewer than MinPts points may be
        # settled without a test (nothing can make it core)
        e["count_lo"] += int(sten[~core & (sten >= mp)].sum()) + mp * int((core & ~dense).sum())
        e["count_hi"] += int(sten[~dense].sum())
    return e


def _run(engine, data, W, S, eps, mp, slides, stats):
    s = engine.Stream(3, W, S, eps, mp)
    s.init(data[:W])
    if stats:
        s.set_stats(True)
    out, cells = [], 0
    for t in range(slides):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        out.append((l.copy(), c.copy()))
        cells += s.info()["cells"]
    st = s.stats() if stats else None
    info = s.info()
    s.close()
    return out, st, cells, info


@pytest.mark.parametrize("mode", ["1", "2"])
def test_counters_pinned(engine, monkeypatch, mode):
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    W, S, slides, eps, mp = 20000, 2000, 4, 0.5, 10
    data = WL.walks3d(W + slides * S, seed=9)
    plain, _, _, _ = _run(engine, data, W, S, eps, mp, slides, False)
    counted, st, cells, info = _run(engine, data, W, S, eps, mp, slides, True)
    for (l1, c1), (l2, c2) in zip(plain, counted):
        assert np.array_equal(l1, l2) and np.array_equal(c1, c2)
    wins = [window_at(data, W, S, t) for t in range(slides)]
    e = expected_counts(wins, [l for l, _ in counted], eps, mp, info)
    assert st["slides"] == slides
    assert st["cells"] == cells == e["cells"]
    assert st["dense_points"] == e["dense_points"] > 0
    assert st["count_queries"] == e["queries"] == W * slides - e["dense_points"]
    assert st["border_tests"] == e["border_tests"]
    assert e["count_lo"] <= st["count_tests"] <= e["count_hi"]
    assert st["cell_parallel"] == (1 if mode == "1" else 0)
    # both paths merge with the cell-parallel kernel
    assert st["merge_tests"] > 0 and 0 < st["unites"] <= st["merge_pairs"]


def test_batch_counters_per_member(engine):
    specs = WL.multi_stream_specs()[:3]
    datas = [WL.stream_points(dict(sp, slides=3)) for sp in specs]
    streams = []
    for sp, d in zip(specs, datas):
        st = engine.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        st.init(d[:sp["window"]])
        streams.append(st)
    b = engine.Batch(streams, use_graph=True)
    b.set_stats(True)
    labels = [[] for _ in specs]
    for t in range(3):
        ls, _ = b.slide([d[sp["window"] + t * sp["stride"]: sp["window"] + (t + 1) * sp["stride"]]
                         for sp, d in zip(specs, datas)])
        for i in range(len(specs)):
            labels[i].append(ls[i].copy())
    for i, (sp, d) in enumerate(zip(specs, datas)):
        st = b.stats(i)
        W, S = sp["window"], sp["stride"]
        e = expected_counts([window_at(d, W, S, t) for t in range(3)], labels[i], sp["eps"], sp["min_pts"],
                            streams[i].info())
        assert st["slides"] == 3
        assert st["cells"] == e["cells"]
        assert st["dense_points"] == e["dense_points"]
        assert st["count_queries"] + st["dense_points"] == 3 * W
        assert st["border_tests"] == e["border_tests"]
        assert e["count_lo"] <= st["count_tests"] <= e["count_hi"]
    b.set_stats(False)
    b.slide([d[sp["window"] + 3 * sp["stride"] - sp["stride"]: sp["window"] + 3 * sp["stride"]]
             for sp, d in zip(specs, datas)])
    assert b.stats(0)["slides"] == 0
    with pytest.raises(engine.RtdbsError):
        b.stats(3)
    b.close()
    for st in streams:
        st.close()
