"""Work counters (rtdbs_stats): invariants on both neighbourhood-kernel forms,
and counting must not change any label."""
import numpy as np
import pytest

from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu


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
    s.close()
    return out, st, cells


@pytest.mark.parametrize("mode", ["1", "2"])
def test_counter_invariants(engine, monkeypatch, mode):
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    W, S, slides = 20000, 2000, 4
    data = WL.walks3d(W + slides * S, seed=9)
    plain, _, _ = _run(engine, data, W, S, 0.5, 10, slides, False)
    counted, st, cells = _run(engine, data, W, S, 0.5, 10, slides, True)
    for (l1, c1), (l2, c2) in zip(plain, counted):
        assert np.array_equal(l1, l2) and np.array_equal(c1, c2)
    assert st["slides"] == slides
    assert st["cells"] == cells
    # every point is either in a dense cell or counted
    assert st["count_queries"] + st["dense_points"] == W * slides
    assert st["dense_points"] > 0
    # a counted point tests at least itself; with early exit at most its stencil
    assert st["count_tests"] >= st["count_queries"]
    assert st["border_tests"] > 0
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
    for t in range(3):
        b.slide([d[sp["window"] + t * sp["stride"]: sp["window"] + (t + 1) * sp["stride"]]
                 for sp, d in zip(specs, datas)])
    for i, sp in enumerate(specs):
        st = b.stats(i)
        assert st["slides"] == 3
        assert st["count_queries"] + st["dense_points"] == 3 * sp["window"]
    b.set_stats(False)
    b.slide([d[sp["window"] + 3 * sp["stride"] - sp["stride"]: sp["window"] + 3 * sp["stride"]]
             for sp, d in zip(specs, datas)])
    assert b.stats(0)["slides"] == 0
    with pytest.raises(engine.RtdbsError):
        b.stats(3)
    b.close()
    for st in streams:
        st.close()
