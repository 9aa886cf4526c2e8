"""Host-buffer slides (the reference's per-slide H2D of the stride and D2H of
label/cid, dbscan.cpp:2289,2344-2345, into pinned buffers, :54-55) and the
server's timeout and lifecycle races.

Every form of the host path must give the same labels as the device path:
pageable buffers (through the engine's pinned staging), engine-pinned
buffers (DMA directly), and pinned buffers laid out so that a batch's copies
merge into one each way.
"""
import threading

import numpy as np
import pytest

from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _walks(n, seed):
    return WL.walks3d(n, seed=seed, n_walks=40, box=(20.0, 20.0, 5.0))


def test_stream_host_forms_agree(engine):
    R = engine
    W, S, slides, eps, mp = 6000, 600, 4, 0.9, 10
    data = _walks(W + slides * S, 21)
    mk = lambda: R.Stream(3, W, S, eps, mp)
    forms = [mk() for _ in range(4)]
    for s in forms:
        s.init(data[:W])
    pin = R.HostBuffer(S * 3 * 4 + 2 * W * 4 + 2 * W * 4 + 64)
    stride_pin = pin.array(np.float32, S * 3)
    out_adj = pin.array(np.int32, 2 * W, offset=S * 3 * 4)           # label then cid, adjacent
    lab_far = pin.array(np.int32, W, offset=S * 3 * 4 + 2 * W * 4)  # label and cid apart
    cid_far = pin.array(np.int32, W, offset=S * 3 * 4 + 3 * W * 4 + 64)
    d_in, d_l, d_c = R.DeviceBuffer(S * 12), R.DeviceBuffer(W * 4), R.DeviceBuffer(W * 4)
    for t in range(slides):
        stride = data[W + t * S: W + (t + 1) * S]
        l0, c0 = forms[0].slide(stride)  # pageable
        stride_pin[:] = stride.reshape(-1)
        l1, c1 = forms[1].slide(stride_pin, out_adj[:W], out_adj[W:])  # pinned, merged copies
        l2, c2 = forms[2].slide(stride_pin, lab_far, cid_far)  # pinned, separate copies
        d_in.upload(stride)
        forms[3].slide_async(d_in.addr(), d_l.addr(), d_c.addr())
        forms[3].sync()
        l3, c3 = d_l.download(np.empty(W, np.int32)), d_c.download(np.empty(W, np.int32))
        for l, c in ((l1, c1), (l2, c2), (l3, c3)):
            assert np.array_equal(l, l0) and np.array_equal(c, c0), t
    for s in forms:
        s.close()
    pin.free()


def test_batch_pinned_contiguous_equals_pageable_and_plain(engine):
    R = engine
    specs = [(2, 3000, 300, 0.8, 5), (3, 4000, 400, 0.9, 10), (2, 1200, 100, 0.6, 4), (3, 5000, 1000, 1.1, 20)]
    slides = 4
    datas = []
    for i, (dim, W, S, eps, mp) in enumerate(specs):
        n = W + slides * S
        datas.append(WL.blobs2d(n, seed=30 + i) if dim == 2 else _walks(n, 30 + i))

    def streams():
        out = []
        for (dim, W, S, eps, mp), d in zip(specs, datas):
            s = R.Stream(dim, W, S, eps, mp)
            s.init(d[:W])
            out.append(s)
        return out

    a, b_, plain = streams(), streams(), streams()
    ba, bb = R.Batch(a, use_graph=True), R.Batch(b_, use_graph=False)
    in_bytes = sum(S * dim * 4 for dim, W, S, _, _ in specs)
    out_bytes = sum(2 * W * 4 for _, W, _, _, _ in specs)
    hin, hout = R.HostBuffer(in_bytes), R.HostBuffer(out_bytes)
    # strides one after the other; each member's label then cid, members in order
    ins, labs, cids = [], [], []
    oi = oo = 0
    for dim, W, S, _, _ in specs:
        ins.append(hin.array(np.float32, S * dim, offset=oi))
        oi += S * dim * 4
        labs.append(hout.array(np.int32, W, offset=oo))
        cids.append(hout.array(np.int32, W, offset=oo + W * 4))
        oo += 2 * W * 4
    for t in range(slides):
        strides = [d[W + t * S: W + (t + 1) * S] for (dim, W, S, _, _), d in zip(specs, datas)]
        for x, s in zip(ins, strides):
            x[:] = s.reshape(-1)
        la, ca = ba.slide(ins, labs, cids)
        lb, cb = bb.slide(strides)
        for i, s in enumerate(plain):
            lp, cp = s.slide(strides[i])
            assert np.array_equal(la[i], lp) and np.array_equal(ca[i], cp), (i, t)
            assert np.array_equal(lb[i], lp) and np.array_equal(cb[i], cp), (i, t)
    ba.close()
    bb.close()
    for s in a + b_ + plain:
        s.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_pipelined_slides_equal_plain(engine, use_graph):
    # rtdbs_batch_slide_pipelined: each call returns once the previous call's
    # outputs are in place; outputs alternate between two pinned sets so that
    # a call's buffers stay untouched until the next call returns
    R = engine
    specs = [(2, 3000, 300, 0.8, 5), (3, 4000, 400, 0.9, 10), (2, 1200, 100, 0.6, 4), (3, 5000, 1000, 1.1, 20)]
    slides = 6
    datas = []
    for i, (dim, W, S, eps, mp) in enumerate(specs):
        n = W + (slides + 2) * S
        datas.append(WL.blobs2d(n, seed=40 + i) if dim == 2 else _walks(n, 40 + i))

    def streams():
        out = []
        for (dim, W, S, eps, mp), d in zip(specs, datas):
            s = R.Stream(dim, W, S, eps, mp)
            s.init(d[:W])
            out.append(s)
        return out

    pipe, plain = streams(), streams()
    bp = R.Batch(pipe, use_graph=use_graph)
    in_bytes = sum(S * dim * 4 for dim, W, S, _, _ in specs)
    out_bytes = sum(2 * W * 4 for _, W, _, _, _ in specs)
    hin = R.HostBuffer(in_bytes * slides)
    houts = [R.HostBuffer(out_bytes) for _ in range(2)]
    ins = []
    oi = 0
    for t in range(slides):
        row = []
        for (dim, W, S, _, _), d in zip(specs, datas):
            a = hin.array(np.float32, S * dim, offset=oi)
            a[:] = d[W + t * S: W + (t + 1) * S].reshape(-1)
            row.append(a)
            oi += S * dim * 4
        ins.append(row)
    outs = []
    for h in houts:
        labs, cids, oo = [], [], 0
        for _, W, _, _, _ in specs:
            labs.append(h.array(np.int32, W, offset=oo))
            cids.append(h.array(np.int32, W, offset=oo + W * 4))
            oo += 2 * W * 4
        outs.append((labs, cids))
    expect = []
    for t in range(slides):
        expect.append([tuple(x.copy() for x in s.slide(d[W + t * S: W + (t + 1) * S]))
                       for s, (dim, W, S, _, _), d in zip(plain, specs, datas)])

    def check(t):
        labs, cids = outs[t % 2]
        for i in range(len(specs)):
            assert np.array_equal(labs[i], expect[t][i][0]) and np.array_equal(cids[i], expect[t][i][1]), (t, i)

    for t in range(slides):
        labs, cids = outs[t % 2]
        bp.slide_pipelined(ins[t], labs, cids)
        if t >= 1:
            check(t - 1)  # delivered when call t returned
    bp.drain()
    check(slides - 1)
    # a synchronous slide after pipelined ones continues the same windows
    t = slides
    strides = [d[W + t * S: W + (t + 1) * S] for (dim, W, S, _, _), d in zip(specs, datas)]
    lb, cb = bp.slide(strides)
    for i, s in enumerate(plain):
        lp, cp = s.slide(strides[i])
        assert np.array_equal(lb[i], lp) and np.array_equal(cb[i], cp), i
    # buffers outside engine-pinned memory are refused
    with pytest.raises(R.RtdbsError):
        bp.slide_pipelined([np.array(x) for x in ins[0]], *outs[0])
    bp.close()
    for s in pipe + plain:
        s.close()
    hin.free()
    for h in houts:
        h.free()


def test_output_arrays_are_checked(engine):
    s = engine.Stream(2, 1000, 100, 0.8, 5)
    s.init(WL.blobs2d(1000, seed=1))
    stride = WL.blobs2d(100, seed=2)
    for bad in (np.empty(999, np.int32), np.empty(1000, np.int64), np.empty(2000, np.int32)[::2],
                np.empty(1000, np.float32)):
        with pytest.raises(ValueError):
            s.slide(stride, label=bad)
    ro = np.empty(1000, np.int32)
    ro.flags.writeable = False
    with pytest.raises(ValueError):
        s.slide(stride, cid=ro)
    s.close()


def test_server_wait_times_out_then_completes(engine):
    R = engine
    W, S = 2000, 200
    data = WL.blobs2d(W + 2 * S, seed=5)
    ref = R.Stream(2, W, S, 0.8, 5)
    ref.init(data[:W])
    want = ref.slide(data[W:W + S])
    ref.close()
    srv = R.Server(hbm_budget_bytes=1 << 30, max_batch=4, lanes=1)
    sid = srv.open_stream(2, W, S, 0.8, 5, data[:W])
    srv.evict(sid)  # its requests wait until it is readmitted
    tk = srv.submit(sid, data[W:W + S])
    with pytest.raises(R.RtdbsError) as e:
        srv.wait(tk, 50)
    assert e.value.code == R.ETIMEOUT
    assert srv.poll(tk) is False
    srv.readmit(sid)
    l, c = srv.wait(tk, 60000)
    assert np.array_equal(l, want[0]) and np.array_equal(c, want[1])
    srv.close()


def _run_together(*fns, limit=60.0):
    errs = []

    def wrap(f):
        try:
            f()
        except Exception as ex:  # noqa: BLE001 - collected and checked below
            errs.append(ex)

    ts = [threading.Thread(target=wrap, args=(f,), daemon=True) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(limit)
    assert not any(t.is_alive() for t in ts), "a server call did not return"
    return errs


def test_server_close_races_evict_and_readmit(engine):
    R = engine
    W, S = 2000, 200
    data = WL.blobs2d(W, seed=6)
    srv = R.Server(hbm_budget_bytes=1 << 30, max_batch=4, lanes=2)
    for _ in range(5):
        # close and evict at once on an idle stream: both return
        sid = srv.open_stream(2, W, S, 0.8, 5, data)
        errs = _run_together(lambda: srv.close_stream(sid), lambda: srv.evict(sid))
        assert all(isinstance(x, R.RtdbsError) and x.code == R.ENOTFOUND for x in errs)
        # readmit and close at once on an evicted stream: no leak, no crash
        sid = srv.open_stream(2, W, S, 0.8, 5, data)
        srv.evict(sid)
        errs = _run_together(lambda: srv.readmit(sid), lambda: srv.close_stream(sid))
        assert all(isinstance(x, R.RtdbsError) and x.code == R.ENOTFOUND for x in errs)
        # two readmits at once: one stream, its bytes counted once
        sid = srv.open_stream(2, W, S, 0.8, 5, data)
        one = srv.stats()["bytes_in_use"]
        srv.evict(sid)
        errs = _run_together(lambda: srv.readmit(sid), lambda: srv.readmit(sid))
        assert not errs
        assert srv.stats()["bytes_in_use"] == one
        srv.close_stream(sid)
        assert srv.stats()["bytes_in_use"] == 0 and srv.stats()["streams"] == 0
    srv.close()
