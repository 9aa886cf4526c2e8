"""The two largest workloads at their real shape, checked stream by stream.

- 64 batched streams (W log-uniform in [5,000, 200,000], 3.21 M window
  points, 2-D blobs and 3-D walks): one graph batch, so one launch covers all
  of them and the large-launch kernel set runs (point-parallel count and
  output kernels, first cores and elements, the merge on the large-batch
  grid). Every stream on every slide equals the on-device brute-force
  validator; slide 0 of every stream equals the CPU oracle.
- 256 served streams (W log-uniform in [2,000, 500,000], 23 M window points,
  all three generators) through the request server, 2 slides each, all
  submitted at once: every result equals the validator.

This is the reference's per-stream acceptance (util.cpp:383-442) at the
strongest form of its rules: labels and ids bit for bit.
"""
import numpy as np
import pytest

from rtdbstream_amd import workloads as WL
from tests.refcheck import window_at

pytestmark = pytest.mark.gpu


def _stride(sp, d, t):
    W, S = sp["window"], sp["stride"]
    return d[W + t * S: W + (t + 1) * S]


@pytest.mark.timeout(900)
def test_64_streams_one_batch_vs_validator_and_oracle(engine, oracle):
    slides = 3
    specs = WL.multi_stream_specs(slides=slides + 1)
    assert sum(sp["window"] for sp in specs) > 1_000_000  # the large-launch kernel set
    datas = [WL.stream_points(sp) for sp in specs]
    streams = []
    for sp, d in zip(specs, datas):
        s = engine.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        s.init(d[:sp["window"]])
        streams.append(s)
    b = engine.Batch(streams, use_graph=True)
    for t in range(slides):
        labels, cids = b.slide([_stride(sp, d, t) for sp, d in zip(specs, datas)])
        for i, (sp, d) in enumerate(zip(specs, datas)):
            win = window_at(d, sp["window"], sp["stride"], t)
            bl, bc, _ = engine.bruteforce(win, sp["eps"], sp["min_pts"], dim=sp["dim"])
            assert np.array_equal(labels[i], bl), ("labels", i, t)
            assert np.array_equal(cids[i], bc), ("cids", i, t)
            if t == 0:
                ol, oc = oracle.stream(d, sp["window"], sp["stride"], sp["eps"], sp["min_pts"], 1)
                assert np.array_equal(labels[i], ol[0]) and np.array_equal(cids[i], oc[0]), ("oracle", i)
    # one more slide with the counters on: the launch is the point-parallel
    # (large-batch) form, and the counting instantiations give the same labels
    b.set_stats(True)
    t = slides
    labels, cids = b.slide([_stride(sp, d, t) for sp, d in zip(specs, datas)])
    for i, (sp, d) in enumerate(zip(specs, datas)):
        assert b.stats(i)["cell_parallel"] == 0
        bl, bc, _ = engine.bruteforce(window_at(d, sp["window"], sp["stride"], t), sp["eps"], sp["min_pts"],
                                      dim=sp["dim"])
        assert np.array_equal(labels[i], bl) and np.array_equal(cids[i], bc), ("counted", i)
    b.close()
    for s in streams:
        s.close()


@pytest.mark.timeout(900)
def test_256_served_streams_vs_validator(engine):
    slides = 2
    specs = WL.served_stream_specs(slides=slides)
    datas = [WL.stream_points(sp) for sp in specs]
    srv = engine.Server(hbm_budget_bytes=64 << 30, max_batch=64, lanes=3)
    sids = [srv.open_stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], d[:sp["window"]])
            for sp, d in zip(specs, datas)]
    # every request at once, interleaved across streams: the lanes form
    # batches of whatever is ready
    tickets = []
    for t in range(slides):
        for i in np.random.default_rng(3 + t).permutation(len(specs)):
            tickets.append((int(i), t, srv.submit(sids[i], _stride(specs[i], datas[i], t))))
    results = {}
    for i, t, tk in tickets:
        results[(i, t)] = srv.wait(tk, 300000)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["mean_batch"] > 1.0  # requests were batched
    for (i, t), (l, c) in sorted(results.items()):
        sp, d = specs[i], datas[i]
        bl, bc, _ = engine.bruteforce(window_at(d, sp["window"], sp["stride"], t), sp["eps"], sp["min_pts"],
                                      dim=sp["dim"])
        assert np.array_equal(l, bl), ("labels", i, t)
        assert np.array_equal(c, bc), ("cids", i, t)
