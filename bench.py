"""Benchmark of the MI355X streaming DBSCAN slide step.

Headline (`value`): the 64-stream batch - 64 independent 2-D/3-D streams,
W log-uniform in [5,000, 200,000], S = W/10, Eps in [0.3, 1.2], MinPts in
{5, 10, 20} (rtdbstream_amd/workloads.py multi_stream_specs) - all sliding
together, one HIP-graph replay per slide. A step is one slide of every
stream: insert each stride, expire the oldest, recluster every window,
labels and ids written. Inputs (all strides) are resident in HBM when the
timed region starts and outputs stay in HBM; `value` is stride points
ingested per second over the whole batch. `vs_plain` compares it with the same
streams slid one at a time through the plain path on the same GPU.

Also on the same JSON line:
  transfer_inclusive - the same two legs timed the reference's way
                  (dbscan.cpp:2289 stride H2D through :2344-2345 label/cid
                  D2H, into pinned host buffers, :54-55): strides from pinned
                  host memory, labels and ids back in pinned host memory;
                  and the batch once more through pipelined slides, whose
                  transfers overlap the next slide's kernels
  plain         - single-stream legs: walks3d (W 100,000) and skewed3d
                  (W 1,000,000), device-resident and transfer-inclusive
  roofline      - the range-count kernel of the 64-stream launch: algorithmic
                  bytes per launch (rtdbstream_amd/roofline.py) / its average
                  duration from HIP events recorded by its own dispatch on the
                  batch's stream; `traffic` is the HBM bytes per launch from
                  the rocprofv3 counter passes summarised under profiles/
  cpu_baseline  - the CPU port (oracle grid restatement, OpenMP on all host
                  threads) on a bounded sample of the 64-stream workload,
                  labels compared; plus the single-thread brute force of the
                  reference's CPU path on blobs2d (W 10,000)
  served        - the request server (256 streams, Poisson arrivals at a
                  swept rate): throughput and p50/p99 latency, host buffers
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "points/sec & slides/sec + p99 slide latency, batched/served vs plain, exact labels"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def window_at(data, W, S, t):
    win = np.array(data[:W], copy=True)
    units = W // S
    for k in range(t + 1):
        u = k % units
        win[u * S:(u + 1) * S] = data[W + k * S: W + (k + 1) * S]
    return win


def pct(v, q):
    return round(float(np.percentile(v, q)), 4)


def tune(R, dim, W, S, eps, mp, data, enabled):
    """Grid side and kernel form from the project's autotuner (trial slides
    on the stream's first W + 9S points); the defaults when disabled."""
    if not enabled:
        return {"cell_div": 0, "lanes_per_query": 0, "neighbor_mode": 0, "cellpar_per_cu": 0, "rows_per_cu": 0,
                "from_cache": False, "trials": 0}
    return R.autotune(dim, W, S, eps, mp, data[:W + 9 * S], slides=8)


def tuned_stream(R, sp, t, data):
    s = R.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], cell_div=t["cell_div"],
                 lanes_per_query=t["lanes_per_query"], neighbor_mode=t["neighbor_mode"],
                 cellpar_per_cu=t["cellpar_per_cu"], rows_per_cu=t["rows_per_cu"])
    s.init(data[:sp["window"]])
    return s


def measured_traffic(name):
    """HBM bytes per launch of a kernel on a workload, from the two rocprofv3
    counter passes (FETCH_SIZE, WRITE_SIZE, each scaled by what it reads for a
    copy kernel of known size) summarised under profiles/<name>_traffic.md;
    None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.md" % name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def range_count_bytes(specs, datas, infos, ts):
    """Algorithmic bytes of one range-count launch over the given streams,
    averaged over the windows of slides ts."""
    from rtdbstream_amd import roofline as RF

    total = 0.0
    for sp, d, info in zip(specs, datas, infos):
        W, S = sp["window"], sp["stride"]
        per = []
        for t in ts:
            win = window_at(d, W, S, t)
            q = ~RF.dense_mask(win, sp["eps"], sp["min_pts"], info)
            per.append(RF.range_count_bytes(win, sp["eps"], q))
        total += float(np.mean(per))
    return total


def roofline_obj(B, avg_ms, kernel, traffic_name):
    achieved = B / (avg_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(traffic_name),
            "kernel": kernel, "algorithmic_bytes_per_launch": int(B), "avg_launch_us": round(avg_ms * 1e3, 2)}


class PinnedSlides:
    """Pinned host buffers of the transfer-inclusive legs: per slide, every
    stream's stride one after the other (the layout a batched receive would
    hand over), and per stream its label then cid, streams in order - the
    layout in which a batch's copies merge into one each way."""

    def __init__(self, R, specs, datas, t0, nslides):
        self.specs = specs
        in_b = sum(sp["stride"] * sp["dim"] * 4 for sp in specs)
        out_b = sum(2 * sp["window"] * 4 for sp in specs)
        self.hin = R.HostBuffer(in_b * nslides)
        self.hout = R.HostBuffer(out_b)
        self.ins = []
        o = 0
        for t in range(t0, t0 + nslides):
            row = []
            for sp, d in zip(specs, datas):
                W, S, dim = sp["window"], sp["stride"], sp["dim"]
                a = self.hin.array(np.float32, S * dim, offset=o)
                a[:] = d[W + t * S: W + (t + 1) * S].reshape(-1)
                row.append(a)
                o += S * dim * 4
            self.ins.append(row)
        self.labels, self.cids = [], []
        o = 0
        for sp in specs:
            W = sp["window"]
            self.labels.append(self.hout.array(np.int32, W, offset=o))
            self.cids.append(self.hout.array(np.int32, W, offset=o + W * 4))
            o += 2 * W * 4

    def free(self):
        self.hin.free()
        self.hout.free()


def batched_only(R, WL, n_streams, steps, warmup):
    """Only the batched graph replays of the 64-stream workload (default
    grid and kernel forms, no trial slides, no plain legs): a clean launch
    sequence for kernel traces and counter passes."""
    specs = WL.multi_stream_specs(n_streams=n_streams, slides=warmup + steps)
    datas = [WL.stream_points(sp) for sp in specs]
    bufs = []
    for sp, d in zip(specs, datas):
        db = R.DeviceBuffer(d.nbytes)
        db.upload(d)
        bufs.append((db, R.DeviceBuffer(sp["window"] * 4), R.DeviceBuffer(sp["window"] * 4)))
    streams = []
    for sp, d in zip(specs, datas):
        s = R.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        s.init(d[:sp["window"]])
        streams.append(s)
    b = R.Batch(streams, use_graph=True)
    a0 = None
    for t in range(warmup + steps):
        if t == warmup:
            R.lib().rtdbs_device_synchronize()
            a0 = time.perf_counter()
        b.slide_async([bufs[i][0].addr((sp["window"] + t * sp["stride"]) * sp["dim"] * 4)
                       for i, sp in enumerate(specs)], [x[1].addr() for x in bufs], [x[2].addr() for x in bufs])
        b.sync()
    el = time.perf_counter() - a0
    b.close()
    for s in streams:
        s.close()
    return {"streams": len(specs), "steps": steps, "ms_per_slide": round(el / steps * 1e3, 4),
            "points_per_s": round(sum(sp["stride"] for sp in specs) * steps / el, 1)}


def batched_leg(R, WL, n_streams, steps, warmup, autotune=True):
    """The 64-stream workload: plain (one stream at a time) and batched (one
    graph replay per slide), each device-resident and transfer-inclusive."""
    extra = 3
    total = warmup + 2 * steps + extra
    specs = WL.multi_stream_specs(n_streams=n_streams, slides=total + 2)
    datas = [WL.stream_points(sp) for sp in specs]
    bufs = []
    for sp, d in zip(specs, datas):
        db = R.DeviceBuffer(d.nbytes)
        db.upload(d)
        W = sp["window"]
        bufs.append((db, R.DeviceBuffer(W * 4), R.DeviceBuffer(W * 4), R.DeviceBuffer(W * 4), R.DeviceBuffer(W * 4)))
    t0 = time.perf_counter()
    tuned = [tune(R, sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], d, autotune)
             for sp, d in zip(specs, datas)]
    tune_s = time.perf_counter() - t0

    def in_addr(i, t):
        sp = specs[i]
        return bufs[i][0].addr((sp["window"] + t * sp["stride"]) * sp["dim"] * 4)

    pts_per_slide = sum(sp["stride"] for sp in specs)
    out = {"streams": len(specs), "window_points": sum(sp["window"] for sp in specs),
           "window_range": [min(sp["window"] for sp in specs), max(sp["window"] for sp in specs)],
           "stride_points_per_slide": pts_per_slide}

    # --- device-resident, plain: slides 0 .. warmup+steps-1
    streams = [tuned_stream(R, sp, tu, d) for sp, d, tu in zip(specs, datas, tuned)]
    for t in range(warmup):
        for i, s in enumerate(streams):
            s.slide_async(in_addr(i, t), bufs[i][1].addr(), bufs[i][2].addr())
            s.sync()
    R.lib().rtdbs_device_synchronize()
    lat_plain = []
    a0 = time.perf_counter()
    for t in range(warmup, warmup + steps):
        for i, s in enumerate(streams):
            a = time.perf_counter()
            s.slide_async(in_addr(i, t), bufs[i][1].addr(), bufs[i][2].addr())
            s.sync()
            lat_plain.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    t_plain = time.perf_counter() - a0
    # --- transfer-inclusive, plain: the next `steps` slides from pinned host buffers
    tx = PinnedSlides(R, specs, datas, warmup + steps, steps)
    lat_plain_x = []
    R.lib().rtdbs_device_synchronize()
    a0 = time.perf_counter()
    for k in range(steps):
        for i, s in enumerate(streams):
            a = time.perf_counter()
            s.slide(tx.ins[k][i], tx.labels[i], tx.cids[i])
            lat_plain_x.append((time.perf_counter() - a) * 1e3)
    t_plain_x = time.perf_counter() - a0
    plain_last = [(l.copy(), c.copy()) for l, c in zip(tx.labels, tx.cids)]
    infos = [s.info() for s in streams]
    for s in streams:
        s.close()

    # --- batched, same streams from the same first windows
    streams = [tuned_stream(R, sp, tu, d) for sp, d, tu in zip(specs, datas, tuned)]
    b = R.Batch(streams, use_graph=True)

    def bslide(t):
        b.slide_async([in_addr(i, t) for i in range(len(streams))], [x[3].addr() for x in bufs],
                      [x[4].addr() for x in bufs])
        b.sync()

    for t in range(warmup):
        bslide(t)
    R.lib().rtdbs_device_synchronize()
    lat_b = []
    a0 = time.perf_counter()
    for t in range(warmup, warmup + steps):
        a = time.perf_counter()
        bslide(t)
        lat_b.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    t_batch = time.perf_counter() - a0
    same_dev = True
    for i, sp in enumerate(specs):
        W = sp["window"]
        l1, c1, l2, c2 = (np.empty(W, np.int32) for _ in range(4))
        bufs[i][1].download(l1)
        bufs[i][2].download(c1)
        bufs[i][3].download(l2)
        bufs[i][4].download(c2)
        same_dev = same_dev and np.array_equal(l1, l2) and np.array_equal(c1, c2)
    # --- transfer-inclusive, batched: the same pinned buffers (two copies per slide)
    lat_bx = []
    R.lib().rtdbs_device_synchronize()
    a0 = time.perf_counter()
    for k in range(steps):
        a = time.perf_counter()
        b.slide(tx.ins[k], tx.labels, tx.cids)
        lat_bx.append((time.perf_counter() - a) * 1e3)
    t_batch_x = time.perf_counter() - a0
    same_x = all(np.array_equal(l, pl) and np.array_equal(c, pc)
                 for l, c, (pl, pc) in zip(tx.labels, tx.cids, plain_last))
    # --- transfer-inclusive, pipelined (rtdbs_batch_slide_pipelined): fresh
    # streams brought to the same windows by device slides, then the same
    # pinned strides, outputs alternating between two pinned sets; one
    # slide's transfers overlap the next slide's kernels
    ps = [tuned_stream(R, sp, tu, d) for sp, d, tu in zip(specs, datas, tuned)]
    bp = R.Batch(ps, use_graph=True)
    for t in range(warmup + steps):
        bp.slide_async([in_addr(i, t) for i in range(len(ps))], [x[3].addr() for x in bufs],
                       [x[4].addr() for x in bufs])
    bp.sync()
    hout2 = R.HostBuffer(tx.hout.nbytes)
    outsets = [(tx.labels, tx.cids), ([], [])]
    o = 0
    for sp in specs:
        W = sp["window"]
        outsets[1][0].append(hout2.array(np.int32, W, offset=o))
        outsets[1][1].append(hout2.array(np.int32, W, offset=o + W * 4))
        o += 2 * W * 4
    pipe_warm = 2  # the first calls set up the double-buffered areas and capture their graphs
    for k in range(steps):
        if k == pipe_warm:
            a0 = time.perf_counter()
        bp.slide_pipelined(tx.ins[k], *outsets[k % 2])
    bp.drain()
    t_pipe_x = time.perf_counter() - a0
    n_pipe = steps - pipe_warm
    last = outsets[(steps - 1) % 2]
    same_p = all(np.array_equal(l, pl) and np.array_equal(c, pc)
                 for l, c, (pl, pc) in zip(last[0], last[1], plain_last))
    bp.close()
    for s in ps:
        s.close()
    hout2.free()
    tx.free()
    # --- range-count kernel of the batched launch: HIP events on the batch's
    # own stream, a few more slides without the graph
    bt = R.Batch(streams, use_graph=False)
    bt.set_timing(2)
    t_extra = list(range(warmup + 2 * steps, warmup + 2 * steps + extra))
    for t in t_extra:
        bt.slide_async([in_addr(i, t) for i in range(len(streams))], [x[3].addr() for x in bufs],
                       [x[4].addr() for x in bufs])
        bt.sync()
    cnt_ms, cnt_n = bt.timing()["count"]
    bt.set_timing(1)  # per-phase breakdown (events between phases: outside the timed region)
    t_ph = warmup + 2 * steps + extra
    bt.slide_async([in_addr(i, t_ph) for i in range(len(streams))], [x[3].addr() for x in bufs],
                   [x[4].addr() for x in bufs])
    bt.sync()
    phase_ms = {k: round(v[0] / max(1, v[1]), 4) for k, v in bt.timing().items() if v[1]}
    # the merge's work in one more slide, counted (the counting kernels are not
    # the timed ones): unite calls, touching element pairs, distance tests
    bt.set_timing(0)
    bt.set_stats(True)
    bt.slide_async([in_addr(i, t_ph + 1) for i in range(len(streams))], [x[3].addr() for x in bufs],
                   [x[4].addr() for x in bufs])
    bt.sync()
    tot = {"unites": 0, "merge_pairs": 0, "merge_tests": 0, "count_tests": 0, "count_queries": 0}
    for i in range(len(streams)):
        st = bt.stats(i)
        for k in tot:
            tot[k] += st[k]
    bt.set_stats(False)
    union_s = phase_ms.get("union", 0.0) / 1e3
    merge = dict(tot, union_ms=phase_ms.get("union"),
                 unites_per_s=round(tot["unites"] / union_s, 1) if union_s else None,
                 pairs_per_s=round(tot["merge_pairs"] / union_s, 1) if union_s else None)
    b.close()
    bt.close()
    for s in streams:
        s.close()
    forms = {}
    for tu in tuned:
        f = "k%d/lpq%d/mode%d" % (tu["cell_div"], tu["lanes_per_query"], tu["neighbor_mode"])
        forms[f] = forms.get(f, 0) + 1
    pts = pts_per_slide * steps
    nsl = len(specs) * steps
    out.update({
        "autotune": {"forms": forms, "from_cache": sum(bool(tu["from_cache"]) for tu in tuned),
                     "tune_s": round(tune_s, 2)},
        "t_batch": t_batch, "steps": steps,
        "plain_points_per_s": round(pts / t_plain, 1), "plain_slides_per_s": round(nsl / t_plain, 1),
        "batched_points_per_s": round(pts / t_batch, 1), "batched_slides_per_s": round(nsl / t_batch, 1),
        "speedup": round(t_plain / t_batch, 3),
        "plain_p50_ms": pct(lat_plain, 50), "plain_p99_ms": pct(lat_plain, 99),
        "batched_step_p50_ms": pct(lat_b, 50), "batched_step_p99_ms": pct(lat_b, 99),
        "labels_equal_plain": bool(same_dev),
        "phase_ms": phase_ms,
        "merge_work": merge,
        "transfer_inclusive": {
            "plain_points_per_s": round(pts / t_plain_x, 1), "plain_slides_per_s": round(nsl / t_plain_x, 1),
            "batched_points_per_s": round(pts / t_batch_x, 1), "batched_slides_per_s": round(nsl / t_batch_x, 1),
            "speedup": round(t_plain_x / t_batch_x, 3),
            "plain_p50_ms": pct(lat_plain_x, 50), "plain_p99_ms": pct(lat_plain_x, 99),
            "batched_step_p50_ms": pct(lat_bx, 50), "batched_step_p99_ms": pct(lat_bx, 99),
            "labels_equal_plain": bool(same_x),
            "pipelined_points_per_s": round(pts_per_slide * n_pipe / t_pipe_x, 1),
            "pipelined_slides_per_s": round(len(specs) * n_pipe / t_pipe_x, 1),
            "pipelined_speedup": round((t_plain_x / nsl) / (t_pipe_x / (len(specs) * n_pipe)), 3),
            "pipelined_labels_equal_plain": bool(same_p),
            "host_buffers": "pinned (rtdbs_host_alloc); per slide one stride H2D and one label+cid D2H "
                            "for the batch, three copies per plain slide; pipelined: the same copies on two "
                            "copy streams, overlapping the next slide's kernels"},
        "_count": (cnt_ms / max(1, cnt_n), t_extra, specs, datas, infos),
    })
    return out


def plain_leg(R, WL, name, steps, warmup, autotune=True, roof_sample=4):
    """One stream of a named workload: device-resident slides (timed, with
    the range-count kernel's own events), then the same number of slides
    with host buffers in pinned memory."""
    cfg = WL.WORKLOADS[name]
    sp = dict(cfg, index=0, gen=name)
    W, S, dim = cfg["window"], cfg["stride"], cfg["dim"]
    extra = 3
    total = warmup + 2 * steps + extra
    data = WL.generate(name, n=W + total * S)
    d_data = R.DeviceBuffer(data.nbytes)
    d_data.upload(data)
    d_label, d_cid = R.DeviceBuffer(W * 4), R.DeviceBuffer(W * 4)
    t0 = time.perf_counter()
    tu = tune(R, dim, W, S, cfg["eps"], cfg["min_pts"], data, autotune)
    tune_s = time.perf_counter() - t0
    s = tuned_stream(R, sp, tu, data)

    def slide(t):
        s.slide_async(d_data.addr((W + t * S) * dim * 4), d_label.addr(), d_cid.addr())
        s.sync()

    for t in range(warmup):
        slide(t)
    s.set_timing(2)  # events around the range-count kernel only
    lat = []
    R.lib().rtdbs_device_synchronize()
    a0 = time.perf_counter()
    for t in range(warmup, warmup + steps):
        a = time.perf_counter()
        slide(t)
        lat.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    el = time.perf_counter() - a0
    timing = s.timing()
    s.set_timing(0)
    info = s.info()
    # transfer-inclusive: pinned stride in, pinned label + cid out
    tx = PinnedSlides(R, [sp], [data], warmup + steps, steps)
    latx = []
    a0 = time.perf_counter()
    for k in range(steps):
        a = time.perf_counter()
        s.slide(tx.ins[k][0], tx.labels[0], tx.cids[0])
        latx.append((time.perf_counter() - a) * 1e3)
    elx = time.perf_counter() - a0
    tx.free()
    s.set_timing(1)
    for t in range(warmup + 2 * steps, warmup + 2 * steps + extra):
        slide(t)
    phase_ms = {k: round(v[0] / max(1, v[1]), 4) for k, v in s.timing().items() if v[1]}
    s.close()
    ts = list(range(warmup, warmup + steps))
    ts = ts[:: max(1, len(ts) // roof_sample)][:roof_sample]
    ms, n = timing["count"]
    B = range_count_bytes([sp], [data], [info], ts)
    return {
        "workload": name, "dim": dim, "window": W, "stride": S, "eps": cfg["eps"], "min_pts": cfg["min_pts"],
        "steps": steps,
        "points_per_s": round(steps * S / el, 1), "slides_per_s": round(steps / el, 2),
        "ms_per_slide": round(el / steps * 1e3, 4), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "transfer_inclusive": {"points_per_s": round(steps * S / elx, 1), "ms_per_slide": round(elx / steps * 1e3, 4),
                               "p50_ms": pct(latx, 50), "p99_ms": pct(latx, 99)},
        "phase_ms": phase_ms,
        "autotune": {k: tu[k] for k in ("cell_div", "lanes_per_query", "neighbor_mode", "cellpar_per_cu",
                                        "rows_per_cu", "from_cache", "trials")},
        "tune_s": round(tune_s, 2),
        "cells_last": info["cells"],
        "roofline": roofline_obj(B, ms / max(1, n),
                                 "k_count" if tu["lanes_per_query"] == 0 else "k_count_points<%d>" % tu["lanes_per_query"],
                                 "k_count_%s" % name),
    }


def cpu_baseline(R, WL, seconds, bf_seconds):
    """CPU port (oracle grid restatement, OpenMP) on a bounded sample of the
    64-stream workload: slides of the streams in turn, each stream's GPU
    labels compared bit for bit. Then the single-thread brute force of the
    reference's CPU path (util.cpp:299-354) on blobs2d."""
    from oracle import oracle as O

    threads = O.max_threads()
    per = 10  # slides per stream at most
    specs = WL.multi_stream_specs(slides=per)
    order = np.random.default_rng(0).permutation(len(specs))
    n_done, n_streams, pts, cpu_time, exact = 0, 0, 0, 0.0, True
    for i in order:
        if cpu_time >= seconds:
            break
        sp = specs[i]
        d = WL.stream_points(sp)
        W, S = sp["window"], sp["stride"]
        s = R.Stream(sp["dim"], W, S, sp["eps"], sp["min_pts"])
        s.init(d[:W])
        n_streams += 1
        for t in range(per):
            if cpu_time >= seconds and t > 0:
                break
            label, cid = s.slide(d[W + t * S: W + (t + 1) * S])
            win = window_at(d, W, S, t)
            a = time.perf_counter()
            rl, rc = O.grid(win, sp["eps"], sp["min_pts"], threads=threads)
            cpu_time += time.perf_counter() - a
            exact = exact and np.array_equal(label, rl) and np.array_equal(cid, rc)
            n_done += 1
            pts += S
        s.close()
    out = {"value": round(pts / cpu_time, 1), "unit": "points/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(), "host_threads_visible": os.cpu_count(),
           "sample": "%d slides of the 64-stream workload (%d streams in a seeded order, up to %d slides each): "
                     "oracle grid restatement, OpenMP on %d threads, %.1f s" % (n_done, n_streams, per, threads,
                                                                               cpu_time),
           "slides_per_s": round(n_done / cpu_time, 3), "labels_equal_gpu": bool(exact)}
    # single-thread brute force (O(W^2)), blobs2d W 10,000
    cfg = WL.WORKLOADS["blobs2d"]
    W, S = cfg["window"], cfg["stride"]
    data = WL.generate("blobs2d")
    s = R.Stream(2, W, S, cfg["eps"], cfg["min_pts"])
    s.init(data[:W])
    nb, bt, ex2 = 0, 0.0, True
    while (bt < bf_seconds or nb < 1) and nb < cfg["slides"]:
        label, cid = s.slide(data[W + nb * S: W + (nb + 1) * S])
        win = window_at(data, W, S, nb)
        a = time.perf_counter()
        rl, rc = O.bruteforce(win, cfg["eps"], cfg["min_pts"])
        bt += time.perf_counter() - a
        ex2 = ex2 and np.array_equal(label, rl) and np.array_equal(cid, rc)
        nb += 1
    s.close()
    out["bruteforce_blobs2d"] = {"value": round(nb * S / bt, 1), "unit": "points/s", "cores": 1,
                                 "ms_per_slide": round(bt / nb * 1e3, 2), "slides": nb,
                                 "labels_equal_gpu": bool(ex2),
                                 "sample": "%d slides of blobs2d (W %d, S %d): literal restatement of the "
                                           "reference's CPU DBSCAN, one thread, %.1f s" % (nb, W, S, bt)}
    return out


def served_leg(R, WL, n_streams, slides, factors, max_batch, lanes, budget_gb, seed=7, data_slides=10,
               swap_frac=0.5, swap_factor=1.0):
    """Request server under Poisson arrivals at `factors` x the serial plain
    capacity: throughput and p50/p99 slide latency (submit -> labels and ids
    in the client's pinned host buffers). Each stream holds W + data_slides*S
    points; slide t inserts stride t mod data_slides (a full slide of work;
    the windows repeat with that period). The sweep stops one rate after the
    server falls behind the offered rate (its capacity). Every stream's last
    result is compared with its plain result."""
    specs = WL.served_stream_specs(n_streams=n_streams, slides=data_slides)
    datas = [WL.stream_points(sp) for sp in specs]
    # client buffers in pinned memory: every stream's strides, and one label
    # + cid pair per stream that its requests (run in order) write in turn
    in_b = sum(sp["stride"] * sp["dim"] * 4 * data_slides for sp in specs)
    out_b = sum(2 * sp["window"] * 4 for sp in specs)
    hin, hout = R.HostBuffer(in_b), R.HostBuffer(out_b)
    strides, labs, cids = [], [], []
    oi = oo = 0
    for sp, d in zip(specs, datas):
        W, S, dim = sp["window"], sp["stride"], sp["dim"]
        row = []
        for t in range(data_slides):
            a = hin.array(np.float32, S * dim, offset=oi)
            a[:] = d[W + t * S: W + (t + 1) * S].reshape(-1)
            row.append(a)
            oi += S * dim * 4
        strides.append(row)
        labs.append(hout.array(np.int32, W, offset=oo))
        cids.append(hout.array(np.int32, W, offset=oo + W * 4))
        oo += 2 * W * 4

    def stride(i, t):
        return strides[i][t % data_slides]

    # serial plain slides on the same GPU, same buffers
    plain_ms, last = [], []
    for i, sp in enumerate(specs):
        s = R.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        s.init(datas[i][:sp["window"]])
        for t in range(slides):
            a = time.perf_counter()
            s.slide(stride(i, t), labs[i], cids[i])
            plain_ms.append((time.perf_counter() - a) * 1e3)
        last.append((labs[i].copy(), cids[i].copy()))
        s.close()
    plain_total_s = sum(plain_ms) / 1e3
    capacity = len(plain_ms) / plain_total_s
    stream_bytes = sum(R.Stream.device_bytes(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
                       for sp in specs)
    out = {"streams": n_streams, "slides_per_stream": slides,
           "window_range": [min(sp["window"] for sp in specs), max(sp["window"] for sp in specs)],
           "window_points": sum(sp["window"] for sp in specs), "stream_device_gb": round(stream_bytes / 2**30, 2),
           "hbm_budget_gb": budget_gb, "max_batch": max_batch, "lanes": lanes,
           "client_buffers": "pinned (rtdbs_host_alloc)",
           "plain_serial_slides_per_s": round(capacity, 1),
           "plain_serial_points_per_s": round(sum(sp["stride"] for sp in specs) * slides / plain_total_s, 1),
           "plain_p50_ms": pct(plain_ms, 50), "plain_p99_ms": pct(plain_ms, 99), "sweep": []}
    rng = np.random.default_rng(seed)
    exact = True
    behind = 0
    # the sweep at the full budget, then one run whose budget holds only
    # swap_frac of the streams' device bytes (streams swap through host memory)
    runs = [(f, int(budget_gb * (1 << 30)), False) for f in factors] + [(swap_factor, int(swap_frac * stream_bytes),
                                                                          True)]
    for f, budget, swap in runs:
        if not swap and behind >= 2:
            continue
        order = np.repeat(np.arange(n_streams), slides)
        rng.shuffle(order)
        gaps = rng.exponential(1.0 / (f * capacity), size=len(order))
        srv = R.Server(hbm_budget_bytes=budget, max_batch=max_batch, lanes=lanes, swap=swap)
        failed = []
        try:
            sids = [srv.open_stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"],
                                    datas[i][:sp["window"]]) for i, sp in enumerate(specs)]
            nxt = [0] * n_streams
            tickets = []
            t0 = time.perf_counter()
            due = t0
            for k, i in enumerate(order):
                due += gaps[k]
                while True:
                    now = time.perf_counter()
                    if now >= due:
                        break
                    if due - now > 2e-4:
                        time.sleep((due - now) * 0.5)
                t = nxt[i]
                nxt[i] += 1
                tickets.append(srv.submit(sids[i], stride(i, t), labs[i], cids[i]))
            for tk in tickets:
                try:
                    srv.wait(tk, 120000)
                except R.RtdbsError as e:  # a failed request is reported, not fatal
                    failed.append(str(e))
            wall = time.perf_counter() - t0
            st = srv.stats()
        except R.RtdbsError as e:  # opening or submitting failed: the run is reported
            failed.append(str(e))
        finally:
            srv.close()
        if failed:
            row = {"rate_factor": f, "requests": len(order), "failed": len(failed), "error": failed[0]}
            log("served f=%.2f%s: %d of %d requests failed, first: %s" %
                (f, " (swapping)" if swap else "", len(failed), len(order), failed[0]))
            if swap:
                out["swap_run"] = row
            else:
                out["sweep"].append(row)
                behind = 2  # no higher rates after a failure
            continue
        for i in range(n_streams):
            exact = exact and np.array_equal(labs[i], last[i][0]) and np.array_equal(cids[i], last[i][1])
        nreq = len(order)
        rate = nreq / wall
        row = {"rate_factor": f, "offered_slides_per_s": round(f * capacity, 1), "slides_per_s": round(rate, 1),
               "points_per_s": round(sum(specs[i]["stride"] for i in order) / wall, 1),
               "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
               "mean_batch": round(st["mean_batch"], 2), "batches": st["batches"]}
        if swap:
            row.update({"budget_gb": round(budget / 2**30, 3), "swap_ins": st["swap_ins"],
                        "swap_outs": st["swap_outs"], "failed": st["failed"]})
            out["swap_run"] = row
            log("served with swapping (budget %.2f of the streams' bytes): %s" % (swap_frac, row))
            continue
        out["sweep"].append(row)
        log("served f=%.2f: %s" % (f, row))
        if rate < 0.9 * f * capacity:
            behind += 1
    ok = [r for r in out["sweep"] if "error" not in r]
    sat = max(ok, key=lambda r: r["slides_per_s"]) if ok else None
    out["capacity_slides_per_s"] = sat["slides_per_s"] if sat else None
    out["capacity_points_per_s"] = sat["points_per_s"] if sat else None
    out["capacity_vs_plain"] = round(sat["slides_per_s"] / capacity, 3) if sat else None
    out["labels_equal_plain"] = bool(exact)
    hin.free()
    hout.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed slides of the 64-stream batch")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch-streams", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=30, help="timed slides of each single-stream leg")
    ap.add_argument("--plain", default="walks3d,skewed3d", help="single-stream legs ('' for none)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--bf-seconds", type=float, default=5.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-streams", type=int, default=256)
    ap.add_argument("--serve-slides", type=int, default=20, help="requests per served stream")
    ap.add_argument("--serve-rates", default="0.5,1,2,4,8,16,32",
                    help="offered rates as multiples of the serial plain capacity (stops past saturation)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a plain device copy of this size (profiler byte-counter calibration)")
    ap.add_argument("--only", default="",
                    help="run only this leg (batched, plain, cpu, served; batched-kernels: the 64-stream graph "
                         "replays alone, for profiling)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default grid side and kernel form instead of the autotuner's choice")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs every stream on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import rtdbstream_amd as R
    from rtdbstream_amd import workloads as WL

    R.lib()
    if R.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        src, dst = R.DeviceBuffer(nb), R.DeviceBuffer(nb)
        R._check(R.lib().rtdbs_device_copy(R.ctypes.c_void_p(dst.addr()), R.ctypes.c_void_p(src.addr()), nb))
        log("calibration copy: %d MiB read + %d MiB written by k_copy16" % (args.calib_copy_mb, args.calib_copy_mb))
        del src, dst
    only = args.only
    autotune = not args.no_autotune
    out = {"metric": METRIC}
    if only == "batched-kernels":
        out["batched_only"] = batched_only(R, WL, args.batch_streams, args.steps, args.warmup)
        print(json.dumps(out), flush=True)
        return
    if not only or only == "batched":
        log("batched leg: %d streams, %d warmup + %d timed slides" % (args.batch_streams, args.warmup, args.steps))
        bres = batched_leg(R, WL, args.batch_streams, args.steps, args.warmup, autotune)
        avg_ms, ts, specs, datas, infos = bres.pop("_count")
        t_batch = bres.pop("t_batch")
        B = range_count_bytes(specs, datas, infos, ts)
        out.update({
            "value": bres["batched_points_per_s"], "unit": "points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(t_batch / args.steps * 1e3, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32",
            "data": "synthetic: 64 seeded streams (rtdbstream_amd/workloads.py multi_stream_specs), "
                    "strides resident in HBM, labels and ids written to HBM",
            "config": {"workload": "64 batched streams: 2-D blobs / 3-D walks, W %d-%d (%d window points), "
                                   "S = W/10, Eps 0.3-1.2, MinPts 5/10/20, one graph replay per slide"
                                   % (bres["window_range"][0], bres["window_range"][1], bres["window_points"]),
                       "streams": bres["streams"], "window_points": bres["window_points"],
                       "stride_points_per_slide": bres["stride_points_per_slide"], "path": "batched"},
            "vs_plain": bres["speedup"],
            "slides_per_s": bres["batched_slides_per_s"],
            "p50_ms": bres["batched_step_p50_ms"], "p99_ms": bres["batched_step_p99_ms"],
            "exact_labels": {"batched_vs_plain": bres["labels_equal_plain"],
                             "transfer_inclusive_vs_plain": bres["transfer_inclusive"]["labels_equal_plain"],
                             "pipelined_vs_plain": bres["transfer_inclusive"]["pipelined_labels_equal_plain"]},
            "transfer_inclusive": dict(bres["transfer_inclusive"]),
            "roofline": roofline_obj(B, avg_ms, "k_count_points (64-stream launch)", "k_count_points_64streams"),
            "batched": bres,
        })
        log("batched: %s" % {k: v for k, v in bres.items() if k not in ("phase_ms",)})
        log("roofline: %s" % out["roofline"])
    if (not only or only == "plain") and args.plain:
        out["plain"] = {}
        for name in args.plain.split(","):
            log("plain leg: %s, %d timed slides" % (name, args.plain_steps))
            steps = args.plain_steps if name != "skewed3d" else max(5, args.plain_steps // 3)
            out["plain"][name] = plain_leg(R, WL, name, steps, args.warmup + 1, autotune)
            log("plain %s: %s" % (name, out["plain"][name]))
    if (not only or only == "cpu") and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(R, WL, args.cpu_seconds, args.bf_seconds)
        log("cpu baseline: %s" % out["cpu_baseline"])
    if (not only or only == "served") and not args.no_serve:
        log("served leg: %d streams x %d slides" % (args.serve_streams, args.serve_slides))
        try:
            out["served"] = served_leg(R, WL, args.serve_streams, args.serve_slides,
                                       [float(x) for x in args.serve_rates.split(",")], 64, 3, 64)
        except Exception as e:  # reported in the line; the other legs' figures stand
            out["served"] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("served leg failed: %s" % out["served"]["error"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
