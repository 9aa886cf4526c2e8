"""Batched and served slides: each result must equal the same stream's plain
(alone) result bit for bit, and the oracle's; failure paths are forced from
the host (tiny budget, allocation hook, cancel) so every kernel exits normally."""
import threading

import numpy as np
import pytest

from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _specs():
    # uneven windows, 1-D/2-D/3-D, different eps / min_pts / stencil reach
    return [
        dict(dim=2, W=3000, S=300, eps=0.8, mp=5, k=0, gen="blobs2d", seed=11),
        dict(dim=3, W=4000, S=1000, eps=0.9, mp=10, k=0, gen="walks3d", seed=12),
        dict(dim=2, W=1200, S=100, eps=0.6, mp=4, k=3, gen="blobs2d", seed=13),
        dict(dim=3, W=6000, S=600, eps=1.1, mp=20, k=1, gen="walks3d", seed=14),
        dict(dim=1, W=2000, S=500, eps=0.05, mp=3, k=2, gen="line", seed=15),
        dict(dim=3, W=2500, S=2500, eps=1.5, mp=8, k=2, gen="walks3d", seed=16),
    ]


def _data(sp, slides):
    n = sp["W"] + slides * sp["S"]
    if sp["gen"] == "line":
        rng = np.random.default_rng(sp["seed"])
        return np.sort(rng.uniform(0, 20, (n, 1)), axis=0).astype(np.float32)[rng.permutation(n)]
    if sp["gen"] == "blobs2d":
        return WL.blobs2d(n, seed=sp["seed"])
    return WL.walks3d(n, seed=sp["seed"], n_walks=40, box=(20.0, 20.0, 5.0))


def _stream(R, sp, data):
    s = R.Stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], cell_div=sp["k"])
    s.init(data[:sp["W"]])
    return s


def _stride(sp, data, t):
    return data[sp["W"] + t * sp["S"]: sp["W"] + (t + 1) * sp["S"]]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_plain_and_oracle(engine, oracle, use_graph):
    R = engine
    slides = 5
    specs = _specs()
    datas = [_data(sp, slides) for sp in specs]
    plain = []
    for sp, d in zip(specs, datas):
        s = _stream(R, sp, d)
        plain.append([tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)])
        s.close()
    streams = [_stream(R, sp, d) for sp, d in zip(specs, datas)]
    b = R.Batch(streams, use_graph=use_graph)
    for t in range(slides):
        labels, cids = b.slide([_stride(sp, d, t) for sp, d in zip(specs, datas)])
        for i in range(len(specs)):
            assert np.array_equal(labels[i], plain[i][t][0]), (i, t)
            assert np.array_equal(cids[i], plain[i][t][1]), (i, t)
    b.close()
    for s in streams:
        s.close()
    for i, (sp, d) in enumerate(zip(specs, datas)):
        rl, rc = oracle.stream(d, sp["W"], sp["S"], sp["eps"], sp["mp"], slides, method="grid")
        for t in range(slides):
            assert np.array_equal(plain[i][t][0], rl[t]) and np.array_equal(plain[i][t][1], rc[t]), (i, t)


def test_graph_replays_with_eager_work_between(engine):
    """Several graph replays, with the program's own eager slides of other
    streams and synchronisations in between; every replay checked."""
    R = engine
    slides = 6
    specs = _specs()[:3]
    datas = [_data(sp, slides) for sp in specs]
    ref = []
    for sp, d in zip(specs, datas):
        s = _stream(R, sp, d)
        ref.append([tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)])
        s.close()
    other_sp = _specs()[3]
    other_d = _data(other_sp, slides)
    other = _stream(R, other_sp, other_d)
    streams = [_stream(R, sp, d) for sp, d in zip(specs, datas)]
    b = R.Batch(streams, use_graph=True)
    for t in range(slides):
        labels, cids = b.slide([_stride(sp, d, t) for sp, d in zip(specs, datas)])
        other.slide(_stride(other_sp, other_d, t))  # eager work between replays
        R.lib().rtdbs_device_synchronize()
        for i in range(len(specs)):
            assert np.array_equal(labels[i], ref[i][t][0]) and np.array_equal(cids[i], ref[i][t][1]), (i, t)
    b.close()
    other.close()
    for s in streams:
        s.close()


def test_batch_device_buffers(engine):
    R = engine
    slides = 3
    specs = _specs()[:4]
    datas = [_data(sp, slides) for sp in specs]
    ref = []
    for sp, d in zip(specs, datas):
        s = _stream(R, sp, d)
        ref.append([tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)])
        s.close()
    bufs = []
    for sp, d in zip(specs, datas):
        db = R.DeviceBuffer(d.nbytes)
        db.upload(d)
        bufs.append((db, R.DeviceBuffer(sp["W"] * 4), R.DeviceBuffer(sp["W"] * 4)))
    streams = [_stream(R, sp, d) for sp, d in zip(specs, datas)]
    b = R.Batch(streams, use_graph=True)
    for t in range(slides):
        b.slide_async([bufs[i][0].addr((sp["W"] + t * sp["S"]) * sp["dim"] * 4) for i, sp in enumerate(specs)],
                      [x[1].addr() for x in bufs], [x[2].addr() for x in bufs])
        b.sync()
        for i, sp in enumerate(specs):
            l = np.empty(sp["W"], np.int32)
            c = np.empty(sp["W"], np.int32)
            bufs[i][1].download(l)
            bufs[i][2].download(c)
            assert np.array_equal(l, ref[i][t][0]) and np.array_equal(c, ref[i][t][1]), (i, t)
    b.close()
    for s in streams:
        s.close()


def test_window_roundtrip_and_restore(engine):
    """A stream is its window and cursor: restoring them reproduces every
    later slide exactly (checkpoint / resume)."""
    R = engine
    sp = _specs()[1]
    slides = 7
    d = _data(sp, slides)
    s = _stream(R, sp, d)
    for t in range(3):
        s.slide(_stride(sp, d, t))
    win = s.window()
    from tests.refcheck import window_at
    assert np.array_equal(win, window_at(d, sp["W"], sp["S"], 2))
    rest = [tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(3, slides)]
    s2 = R.Stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], cell_div=sp["k"])
    s2.restore(win, 3)
    for t in range(3, slides):
        l, c = s2.slide(_stride(sp, d, t))
        assert np.array_equal(l, rest[t - 3][0]) and np.array_equal(c, rest[t - 3][1])
    s.close()
    s2.close()


def test_server_mix_equals_alone(engine):
    """Requests from several client threads arrive while others are on the
    device; each result equals the stream's plain result."""
    R = engine
    slides = 5
    specs = _specs()
    datas = [_data(sp, slides) for sp in specs]
    alone = []
    for sp, d in zip(specs, datas):
        s = _stream(R, sp, d)
        alone.append([tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)])
        s.close()
    srv = R.Server(hbm_budget_bytes=8 << 30, max_batch=4, lanes=2)
    sids = [srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]], cell_div=sp["k"])
            for sp, d in zip(specs, datas)]
    results = {}
    errors = []

    def client(idx):
        try:
            sp, d = specs[idx], datas[idx]
            tickets = [srv.submit(sids[idx], _stride(sp, d, t)) for t in range(slides)]
            results[idx] = [srv.wait(tk, 30000) for tk in tickets]
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    th = [threading.Thread(target=client, args=(i,)) for i in range(len(specs))]
    for t in th:
        t.start()
    for t in th:
        t.join(60)
    assert not errors, errors
    for i in range(len(specs)):
        for t in range(slides):
            assert np.array_equal(results[i][t][0], alone[i][t][0]), (i, t)
            assert np.array_equal(results[i][t][1], alone[i][t][1]), (i, t)
    st = srv.stats()
    assert st["completed"] == len(specs) * slides and st["failed"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_server_admission_budget_and_reuse(engine):
    R = engine
    sp = _specs()[0]
    d = _data(sp, 2)
    need = R.Stream.device_bytes(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"])
    srv = R.Server(hbm_budget_bytes=int(need * 1.5), max_batch=2, lanes=1)
    a = srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]])
    with pytest.raises(R.RtdbsError) as e:
        srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]])
    assert e.value.code == R.EBUDGET
    big = R.make_params(3, 1_000_000, 1000, 0.5, 5)
    with pytest.raises(R.RtdbsError) as e:
        srv.open_stream(3, 1_000_000, 1000, 0.5, 5, np.zeros((1_000_000, 3), np.float32))
    assert e.value.code == R.EBUDGET
    srv.close_stream(a)
    b = srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]])
    l, c = srv.wait(srv.submit(b, _stride(sp, d, 0)), 30000)
    s = _stream(R, sp, d)
    l2, c2 = s.slide(_stride(sp, d, 0))
    assert np.array_equal(l, l2) and np.array_equal(c, c2)
    s.close()
    srv.close()
    del big


def test_server_allocation_failure_isolated(engine):
    R = engine
    specs = _specs()[:2]
    datas = [_data(sp, 2) for sp in specs]
    srv = R.Server(hbm_budget_bytes=8 << 30, max_batch=4, lanes=1)
    a = srv.open_stream(specs[0]["dim"], specs[0]["W"], specs[0]["S"], specs[0]["eps"], specs[0]["mp"],
                        datas[0][:specs[0]["W"]])
    R.lib().rtdbs_test_fail_allocations(1)
    with pytest.raises(R.RtdbsError) as e:
        srv.open_stream(specs[1]["dim"], specs[1]["W"], specs[1]["S"], specs[1]["eps"], specs[1]["mp"],
                        datas[1][:specs[1]["W"]])
    assert e.value.code == R.ENOMEM
    assert srv.stats()["streams"] == 1
    l, c = srv.wait(srv.submit(a, _stride(specs[0], datas[0], 0)), 30000)
    s = _stream(R, specs[0], datas[0])
    l2, c2 = s.slide(_stride(specs[0], datas[0], 0))
    assert np.array_equal(l, l2) and np.array_equal(c, c2)
    s.close()
    srv.close()


def test_server_evict_readmit_and_cancel(engine):
    R = engine
    sp = _specs()[1]
    slides = 6
    d = _data(sp, slides)
    s = _stream(R, sp, d)
    alone = [tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)]
    s.close()
    srv = R.Server(hbm_budget_bytes=8 << 30, max_batch=2, lanes=1)
    sid = srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]])
    for t in range(2):
        l, c = srv.wait(srv.submit(sid, _stride(sp, d, t)), 30000)
        assert np.array_equal(l, alone[t][0]) and np.array_equal(c, alone[t][1])
    used = srv.stats()["bytes_in_use"]
    srv.evict(sid)
    assert srv.stats()["bytes_in_use"] == used - R.Stream.device_bytes(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"])
    tk = srv.submit(sid, _stride(sp, d, 2))  # waits while evicted
    assert not srv.poll(tk)
    srv.readmit(sid)
    l, c = srv.wait(tk, 30000)
    assert np.array_equal(l, alone[2][0]) and np.array_equal(c, alone[2][1])
    # cancel a queued request: evict so it cannot start, then cancel
    srv.evict(sid)
    tk2 = srv.submit(sid, _stride(sp, d, 3))
    srv.cancel(tk2)
    with pytest.raises(R.RtdbsError) as e:
        srv.wait(tk2, 1000)
    assert e.value.code == R.ECANCELLED
    srv.readmit(sid)
    # the cancelled stride was never applied: slide 3 proceeds normally
    l, c = srv.wait(srv.submit(sid, _stride(sp, d, 3)), 30000)
    assert np.array_equal(l, alone[3][0]) and np.array_equal(c, alone[3][1])
    srv.close()


def test_server_failed_swap_in_reports_its_cause(engine):
    """A swap-in whose device allocation fails fails only the request that
    asked for it, and the error names the call that failed (not only the
    status); the stream's next request swaps it in and slides as if the
    failed request had never been made."""
    R = engine
    specs = _specs()[:2]
    datas = [_data(sp, 2) for sp in specs]
    sizes = [R.Stream.device_bytes(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], cell_div=sp["k"])
             for sp in specs]
    srv = R.Server(hbm_budget_bytes=max(sizes), max_batch=4, lanes=1, swap=True)  # one stream at a time
    sids = [srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]], cell_div=sp["k"])
            for sp, d in zip(specs, datas)]
    srv.wait(srv.submit(sids[0], _stride(specs[0], datas[0], 0)), 30000)  # stream 0 on the device
    R.lib().rtdbs_test_fail_allocations(1)
    with pytest.raises(R.RtdbsError) as e:
        srv.wait(srv.submit(sids[1], _stride(specs[1], datas[1], 0)), 30000)
    assert e.value.code == R.ENOMEM
    assert "allocation hook" in str(e.value)
    assert srv.stats()["failed"] == 1
    l, c = srv.wait(srv.submit(sids[1], _stride(specs[1], datas[1], 0)), 30000)
    s = _stream(R, specs[1], datas[1])
    l2, c2 = s.slide(_stride(specs[1], datas[1], 0))
    assert np.array_equal(l, l2) and np.array_equal(c, c2)
    s.close()
    srv.close()


def test_server_swaps_under_a_binding_budget(engine):
    """A budget that holds about half the streams: with swapping on, every
    stream still opens, requests of streams not on the device swap them in
    (least recently used idle streams out), device bytes never exceed the
    budget, and every result equals the stream's plain result."""
    R = engine
    slides = 4
    specs = _specs()
    datas = [_data(sp, slides) for sp in specs]
    plain = []
    for sp, d in zip(specs, datas):
        s = _stream(R, sp, d)
        plain.append([tuple(x.copy() for x in s.slide(_stride(sp, d, t))) for t in range(slides)])
        s.close()
    sizes = [R.Stream.device_bytes(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], cell_div=sp["k"])
             for sp in specs]
    budget = max(sizes) + sum(sorted(sizes)[:2])  # the largest stream fits, never all of them
    assert budget < sum(sizes)
    srv = R.Server(hbm_budget_bytes=budget, max_batch=4, lanes=2, swap=True)
    sids = [srv.open_stream(sp["dim"], sp["W"], sp["S"], sp["eps"], sp["mp"], d[:sp["W"]], cell_div=sp["k"])
            for sp, d in zip(specs, datas)]
    assert srv.stats()["resident"] < len(specs)
    for t in range(slides):
        order = np.random.default_rng(t).permutation(len(specs))
        tickets = [(int(i), srv.submit(sids[i], _stride(specs[i], datas[i], t))) for i in order]
        for i, tk in tickets:
            l, c = srv.wait(tk, 60000)
            assert np.array_equal(l, plain[i][t][0]) and np.array_equal(c, plain[i][t][1]), (i, t)
        st = srv.stats()
        assert st["bytes_in_use"] <= budget
    st = srv.stats()
    assert st["swap_ins"] > 0 and st["swap_outs"] > 0 and st["failed"] == 0
    with pytest.raises(R.RtdbsError) as e:  # larger than the whole budget
        srv.open_stream(3, 200000, 20000, 0.5, 10, np.zeros((200000, 3), np.float32))
    assert e.value.code == R.EBUDGET
    for sid in sids:
        srv.close_stream(sid)
    assert srv.stats()["bytes_in_use"] == 0
    srv.close()
