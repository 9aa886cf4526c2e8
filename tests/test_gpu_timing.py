"""Per-phase timing (rtdbs_*_set_timing / _get_timing): every recorded slide
is counted exactly once, also past the engine's ring of event sets, and
timing never changes a label."""
import numpy as np
import pytest

from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu

PHASES = ("sort_expire", "rank_compact", "merge_final", "cells", "rows", "count", "union", "border")


def _stream(engine, W=4000, S=400, seed=5):
    data = WL.walks3d(W + 100 * S, seed=seed)
    s = engine.Stream(3, W, S, 0.5, 10)
    s.init(data[:W])
    return s, data


def _slides(s, data, t0, n, W=4000, S=400):
    out = []
    for t in range(t0, t0 + n):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        out.append((l.copy(), c.copy()))
    return out


def test_every_phase_counted_past_the_ring(engine):
    s, data = _stream(engine)
    s.set_timing(1)
    _slides(s, data, 0, 70)  # more slides than the ring of event sets holds
    tm = s.timing()
    assert set(PHASES) <= set(tm)
    for ph in PHASES:
        ms, n = tm[ph]
        assert n == 70 and ms > 0.0, (ph, ms, n)
    # reading does not count a slide twice
    assert s.timing()["count"][1] == 70
    s.close()


def test_count_only_and_off(engine):
    s, data = _stream(engine)
    s.set_timing(2)
    _slides(s, data, 0, 5)
    tm = s.timing()
    assert tm["count"][1] == 5 and tm["count"][0] > 0.0
    assert all(tm[ph][1] == 0 for ph in PHASES if ph != "count")
    s.set_timing(0)
    _slides(s, data, 5, 3)
    assert s.timing()["count"][1] == 0  # switching resets the totals
    s.close()


def test_timing_does_not_change_labels(engine):
    a, data = _stream(engine, seed=6)
    b, _ = _stream(engine, seed=6)
    b.set_timing(1)
    for (l1, c1), (l2, c2) in zip(_slides(a, data, 0, 6), _slides(b, data, 0, 6)):
        assert np.array_equal(l1, l2) and np.array_equal(c1, c2)
    a.close()
    b.close()


def test_batch_timing(engine):
    specs = WL.multi_stream_specs()[:3]
    datas = [WL.stream_points(dict(sp, slides=4)) for sp in specs]
    streams = []
    for sp, d in zip(specs, datas):
        st = engine.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        st.init(d[:sp["window"]])
        streams.append(st)
    b = engine.Batch(streams, use_graph=False)
    b.set_timing(1)
    for t in range(4):
        b.slide([d[sp["window"] + t * sp["stride"]: sp["window"] + (t + 1) * sp["stride"]]
                 for sp, d in zip(specs, datas)])
    tm = b.timing()
    assert all(tm[ph][1] == 4 for ph in PHASES)
    b.close()
    for st in streams:
        st.close()
