"""Benchmark of the MI355X streaming DBSCAN slide step.

Default workload (the configuration the headline metric is quoted on): the
3-D trajectory-like stream "walks3d" - W 100,000, S 10,000, Eps 0.5, MinPts 10,
one stream, plain per-slide GPU path. A step is one slide: insert the stride,
expire the oldest stride, recluster the window, labels available on the device.
All inputs are resident in HBM before the timed region.

Also reported on the same JSON line:
  roofline      - the range-count kernel: algorithmic bytes per launch
                  (rtdbstream_amd/roofline.py) / its average duration from HIP
                  events recorded by its own dispatch; `traffic` is the HBM
                  bytes per launch from the rocprofv3 counter passes summarised
                  in profiles/k_count_<workload>_traffic.md
  cpu_baseline  - the CPU port (oracle, OpenMP) on a bounded sample of the same
                  slides, whose labels are also compared with the GPU's
  batched       - the 64-stream batch: one launch per phase, HIP graph, vs the
                  same streams slid one at a time (plain)
  served        - the request server under Poisson arrivals: throughput and
                  p50/p99 latency
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "points/sec & slides/sec + p99 slide latency, batched/served vs plain, exact labels"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def window_at(data, W, S, t):
    win = np.array(data[:W], copy=True)
    units = W // S
    for k in range(t + 1):
        u = k % units
        win[u * S:(u + 1) * S] = data[W + k * S: W + (k + 1) * S]
    return win


def tune(R, dim, W, S, eps, mp, data, enabled):
    """Grid side and kernel form from the project's autotuner (trial slides
    on the stream's first W + 9S points); the defaults when disabled."""
    if not enabled:
        return {"cell_div": 0, "lanes_per_query": 0, "neighbor_mode": 0, "from_cache": False, "trials": 0}
    return R.autotune(dim, W, S, eps, mp, data[:W + 9 * S], slides=8)


def tuned_stream(R, dim, W, S, eps, mp, t):
    return R.Stream(dim, W, S, eps, mp, cell_div=t["cell_div"], lanes_per_query=t["lanes_per_query"],
                    neighbor_mode=t["neighbor_mode"])


def plain_leg(R, WL, name, steps, warmup, autotune=True):
    cfg = WL.WORKLOADS[name]
    W, S, dim, eps, mp = cfg["window"], cfg["stride"], cfg["dim"], cfg["eps"], cfg["min_pts"]
    total_slides = warmup + steps
    extra = 10
    data = WL.generate(name, n=W + (total_slides + extra) * S)
    d_data = R.DeviceBuffer(data.nbytes)
    d_data.upload(data)
    d_label = R.DeviceBuffer(W * 4)
    d_cid = R.DeviceBuffer(W * 4)
    t0 = time.perf_counter()
    tuned = tune(R, dim, W, S, eps, mp, data, autotune)
    tune_s = time.perf_counter() - t0
    s = tuned_stream(R, dim, W, S, eps, mp, tuned)
    s.init(data[:W])

    def slide(t):
        s.slide_async(d_data.addr((W + t * S) * dim * 4), d_label.addr(), d_cid.addr())
        s.sync()

    for t in range(warmup):
        slide(t)
    s.set_timing(2)  # events around the range-count kernel only
    lat = []
    R.lib().rtdbs_device_synchronize()
    t0 = time.perf_counter()
    for t in range(warmup, total_slides):
        a = time.perf_counter()
        slide(t)
        lat.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    elapsed = time.perf_counter() - t0
    timing = s.timing()
    info = s.info()
    # per-phase breakdown from a separate pass (events between every phase
    # add launch gaps, so they stay out of the timed region)
    s.set_timing(1)
    for t in range(total_slides, total_slides + min(steps, extra)):
        slide(t)
    phase_ms = {k: round(v[0] / max(1, v[1]), 4) for k, v in s.timing().items()}
    s.close()
    return dict(data=data, cfg=cfg, cfg_name=name, elapsed=elapsed, lat=lat, timing=timing, info=info, W=W, S=S, dim=dim,
                phase_ms=phase_ms, tuned={k: tuned[k] for k in ("cell_div", "lanes_per_query", "neighbor_mode",
                                                                 "from_cache", "trials")},
                tune_s=round(tune_s, 2),
                eps=eps, min_pts=mp, warmup=warmup, steps=steps)


def measured_traffic(workload):
    """HBM bytes per launch of the range-count kernel on this workload, from
    the two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE, each scaled by
    what it reads for a copy kernel of known size) summarised under profiles/.
    None when no such summary exists for the workload."""
    path = os.path.join(ROOT, "profiles", "k_count_%s_traffic.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def roofline(WL, res, sample=8):
    from rtdbstream_amd import roofline as RF

    data, W, S, eps, mp = res["data"], res["W"], res["S"], res["eps"], res["min_pts"]
    info = res["info"]
    t_first = res["warmup"]
    ts = list(range(t_first, t_first + res["steps"]))
    if len(ts) > sample:
        ts = ts[:: max(1, len(ts) // sample)][:sample]
    bytes_per = []
    for t in ts:
        win = window_at(data, W, S, t)
        q = ~RF.dense_mask(win, eps, mp, info)
        bytes_per.append(RF.range_count_bytes(win, eps, q))
    B = float(np.mean(bytes_per))
    ms, n = res["timing"]["count"]
    avg_s = ms / max(1, n) / 1e3
    achieved = B / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(res["cfg_name"]),
            "kernel": "k_count", "algorithmic_bytes_per_launch": int(B),
            "avg_launch_us": round(avg_s * 1e6, 2)}


def batched_roofline(WL, bres):
    """Range-count roofline of the 64-stream launch: algorithmic bytes summed
    over the streams' windows / the launch's HIP-event duration."""
    from rtdbstream_amd import roofline as RF

    avg_ms, ts, specs, datas, infos = bres.pop("_count_kernel")
    total = 0.0
    for sp, d, info in zip(specs, datas, infos):
        W, S = sp["window"], sp["stride"]
        per = []
        for t in ts:
            win = window_at(d, W, S, t)
            q = ~RF.dense_mask(win, sp["eps"], sp["min_pts"], info)
            per.append(RF.range_count_bytes(win, sp["eps"], q))
        total += float(np.mean(per))
    achieved = total / (avg_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_count (64-stream launch)",
            "algorithmic_bytes_per_launch": int(total), "avg_launch_us": round(avg_ms * 1e3, 2)}


def cpu_baseline(R, WL, res, seconds, max_slides=400):
    """CPU port (oracle grid, OpenMP) on a bounded sample of the same workload
    (about `seconds` of CPU work); the GPU labels of the same slides are
    computed alongside and compared bit for bit."""
    from oracle import oracle as O

    W, S, dim, eps, mp = res["W"], res["S"], res["dim"], res["eps"], res["min_pts"]
    name = res["cfg_name"]
    data = WL.generate(name, n=W + max_slides * S)
    threads = O.max_threads()
    s = R.Stream(dim, W, S, eps, mp)
    s.init(data[:W])
    win = np.array(data[:W], copy=True)
    units = W // S
    n_done, cpu_time, exact = 0, 0.0, True
    t = 0
    while (cpu_time < seconds or n_done < 2) and t < max_slides:
        stride = data[W + t * S: W + (t + 1) * S]
        label, cid = s.slide(stride)
        u = t % units
        win[u * S:(u + 1) * S] = stride
        a = time.perf_counter()
        rl, rc = O.grid(win, eps, mp, threads=threads)
        cpu_time += time.perf_counter() - a
        exact = exact and np.array_equal(label, rl) and np.array_equal(cid, rc)
        n_done += 1
        t += 1
    s.close()
    return {"value": round(n_done * S / cpu_time, 1), "unit": "points/s", "cores": threads, "kind": "port",
            "sample": "%d slides of %s (W=%d, S=%d): oracle grid restatement, OpenMP on %d threads, %.1f s"
                      % (n_done, name, W, S, threads, cpu_time),
            "slides_per_s": round(n_done / cpu_time, 3)}, exact, n_done


def batched_leg(R, WL, n_streams, slides, warmup, autotune=True):
    specs = WL.multi_stream_specs(n_streams=n_streams, slides=slides + warmup + 2)
    streams, bufs, datas = [], [], []
    for sp in specs:
        d = WL.stream_points(sp)
        datas.append(d)
        db = R.DeviceBuffer(d.nbytes)
        db.upload(d)
        bufs.append((db, R.DeviceBuffer(sp["window"] * 4), R.DeviceBuffer(sp["window"] * 4),
                     R.DeviceBuffer(sp["window"] * 4), R.DeviceBuffer(sp["window"] * 4)))
    # every stream's grid side and kernel form from the autotuner (both legs)
    t0 = time.perf_counter()
    tuned = [tune(R, sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], d, autotune)
             for sp, d in zip(specs, datas)]
    tune_s = time.perf_counter() - t0
    # plain: one stream at a time
    for sp, d, tu in zip(specs, datas, tuned):
        s = tuned_stream(R, sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], tu)
        s.init(d[:sp["window"]])
        streams.append(s)

    def in_addr(i, t):
        sp = specs[i]
        return bufs[i][0].addr((sp["window"] + t * sp["stride"]) * sp["dim"] * 4)

    for t in range(warmup):
        for i, s in enumerate(streams):
            s.slide_async(in_addr(i, t), bufs[i][1].addr(), bufs[i][2].addr())
            s.sync()
    R.lib().rtdbs_device_synchronize()
    t0 = time.perf_counter()
    lat_plain = []
    for t in range(warmup, warmup + slides):
        for i, s in enumerate(streams):
            a = time.perf_counter()
            s.slide_async(in_addr(i, t), bufs[i][1].addr(), bufs[i][2].addr())
            s.sync()
            lat_plain.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    t_plain = time.perf_counter() - t0
    for s in streams:
        s.close()
    # batched: same streams from the same first windows, one graph replay per slide
    streams = []
    for sp, d, tu in zip(specs, datas, tuned):
        s = tuned_stream(R, sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"], tu)
        s.init(d[:sp["window"]])
        streams.append(s)
    b = R.Batch(streams, use_graph=True)
    bt = R.Batch(streams, use_graph=False)  # same streams, events around the range count

    def bslide(t):
        b.slide_async([in_addr(i, t) for i in range(len(streams))], [x[3].addr() for x in bufs],
                      [x[4].addr() for x in bufs])
        b.sync()

    for t in range(warmup):
        bslide(t)
    R.lib().rtdbs_device_synchronize()
    t0 = time.perf_counter()
    lat_b = []
    for t in range(warmup, warmup + slides):
        a = time.perf_counter()
        bslide(t)
        lat_b.append((time.perf_counter() - a) * 1e3)
    R.lib().rtdbs_device_synchronize()
    t_batch = time.perf_counter() - t0
    # exact: batched labels of the last slide equal the plain ones, per stream
    # (checked before the timing pass below moves the windows on)
    same = True
    for i, sp in enumerate(specs):
        W = sp["window"]
        l1, c1, l2, c2 = (np.empty(W, np.int32) for _ in range(4))
        bufs[i][1].download(l1)
        bufs[i][2].download(c1)
        bufs[i][3].download(l2)
        bufs[i][4].download(c2)
        same = same and np.array_equal(l1, l2) and np.array_equal(c1, c2)
    # range-count kernel of the batched launch: HIP events on the batch's own
    # stream, two more slides without the graph
    bt.set_timing(2)
    t_extra = []
    for t in range(warmup + slides, warmup + slides + 2):
        bt.slide_async([in_addr(i, t) for i in range(len(streams))], [x[3].addr() for x in bufs],
                       [x[4].addr() for x in bufs])
        bt.sync()
        t_extra.append(t)
    cnt_ms, cnt_n = bt.timing()["count"]
    infos = [st.info() for st in streams]
    b.close()
    bt.close()
    for s in streams:
        s.close()
    pts = sum(sp["stride"] for sp in specs) * slides
    nsl = len(specs) * slides
    forms = {}
    for tu in tuned:
        f = "k%d/lpq%d/mode%d" % (tu["cell_div"], tu["lanes_per_query"], tu["neighbor_mode"])
        forms[f] = forms.get(f, 0) + 1
    return {
        "streams": len(specs), "slides_per_stream": slides,
        "autotune": {"forms": forms, "from_cache": sum(bool(tu["from_cache"]) for tu in tuned),
                     "tune_s": round(tune_s, 2)},
        "window_range": [min(sp["window"] for sp in specs), max(sp["window"] for sp in specs)],
        "plain_points_per_s": round(pts / t_plain, 1), "plain_slides_per_s": round(nsl / t_plain, 1),
        "batched_points_per_s": round(pts / t_batch, 1), "batched_slides_per_s": round(nsl / t_batch, 1),
        "speedup": round(t_plain / t_batch, 3),
        "plain_p50_ms": round(float(np.percentile(lat_plain, 50)), 3),
        "plain_p99_ms": round(float(np.percentile(lat_plain, 99)), 3),
        "batched_step_p50_ms": round(float(np.percentile(lat_b, 50)), 3),
        "batched_step_p99_ms": round(float(np.percentile(lat_b, 99)), 3),
        "labels_equal_plain": bool(same),
        "_count_kernel": (cnt_ms / max(1, cnt_n), t_extra, specs, datas, infos),
    }


def served_leg(R, WL, n_streams, slides, factors, max_batch, lanes, budget_gb, seed=7):
    """Request server under Poisson arrivals at `factors` x the serial plain
    capacity: throughput and p50/p99 slide latency (submit -> labels in host
    memory). Each stream's last result is compared with its plain result."""
    specs = WL.served_stream_specs(n_streams=n_streams, slides=slides)
    datas = [WL.stream_points(sp) for sp in specs]

    def stride(i, t):
        sp = specs[i]
        return datas[i][sp["window"] + t * sp["stride"]: sp["window"] + (t + 1) * sp["stride"]]

    # serial plain slides on the same GPU (host buffers, like the server)
    plain_ms, last = [], []
    for i, sp in enumerate(specs):
        s = R.Stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"])
        s.init(datas[i][:sp["window"]])
        for t in range(slides):
            a = time.perf_counter()
            l, c = s.slide(stride(i, t))
            plain_ms.append((time.perf_counter() - a) * 1e3)
        last.append((l.copy(), c.copy()))
        s.close()
    plain_total_s = sum(plain_ms) / 1e3
    capacity = len(plain_ms) / plain_total_s
    out = {"streams": n_streams, "slides_per_stream": slides,
           "window_range": [min(sp["window"] for sp in specs), max(sp["window"] for sp in specs)],
           "hbm_budget_gb": budget_gb, "max_batch": max_batch, "lanes": lanes,
           "plain_serial_slides_per_s": round(capacity, 1),
           "plain_serial_points_per_s": round(sum(sp["stride"] for sp in specs) * slides / plain_total_s, 1),
           "plain_p50_ms": round(float(np.percentile(plain_ms, 50)), 3),
           "plain_p99_ms": round(float(np.percentile(plain_ms, 99)), 3), "sweep": []}
    rng = np.random.default_rng(seed)
    exact = True
    for f in factors:
        srv = R.Server(hbm_budget_bytes=int(budget_gb * (1 << 30)), max_batch=max_batch, lanes=lanes)
        sids = [srv.open_stream(sp["dim"], sp["window"], sp["stride"], sp["eps"], sp["min_pts"],
                                datas[i][:sp["window"]]) for i, sp in enumerate(specs)]
        order = np.repeat(np.arange(n_streams), slides)
        rng.shuffle(order)
        gaps = rng.exponential(1.0 / (f * capacity), size=len(order))
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
            tickets.append((i, t, srv.submit(sids[i], stride(i, t))))
        res_last = {}
        for i, t, tk in tickets:
            l, c = srv.wait(tk, 120000)
            if t == slides - 1:
                res_last[i] = (l, c)
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        for i in range(n_streams):
            exact = exact and np.array_equal(res_last[i][0], last[i][0]) and np.array_equal(res_last[i][1], last[i][1])
        nreq = len(order)
        out["sweep"].append({"rate_factor": f, "offered_slides_per_s": round(f * capacity, 1),
                             "slides_per_s": round(nreq / wall, 1),
                             "points_per_s": round(sum(specs[i]["stride"] for i in order) / wall, 1),
                             "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                             "mean_batch": round(st["mean_batch"], 2), "batches": st["batches"]})
        log("served f=%.2f: %s" % (f, out["sweep"][-1]))
    out["labels_equal_plain"] = bool(exact)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="walks3d")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--batch-streams", type=int, default=64)
    ap.add_argument("--batch-slides", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-streams", type=int, default=256)
    ap.add_argument("--serve-slides", type=int, default=3)
    ap.add_argument("--serve-rates", default="0.5,1,2,4")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a plain device copy of this size (profiler byte-counter calibration)")
    ap.add_argument("--no-cpu", action="store_true")
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
    log("plain leg: %s, %d warmup + %d timed slides" % (args.workload, args.warmup, args.steps))
    res = plain_leg(R, WL, args.workload, args.steps, args.warmup, autotune=not args.no_autotune)
    W, S = res["W"], res["S"]
    ms_per = res["elapsed"] / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": round(args.steps * S / res["elapsed"], 1),
        "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded %s generator (rtdbstream_amd/workloads.py), inputs resident in HBM" % args.workload,
        "config": {"workload": args.workload, "dim": res["dim"], "window": W, "stride": S, "eps": res["eps"],
                   "min_pts": res["min_pts"], "streams": 1, "path": "plain"},
        "slides_per_s": round(args.steps / res["elapsed"], 2),
        "window_points_per_s": round(args.steps * W / res["elapsed"], 1),
        "p50_ms": round(float(np.percentile(res["lat"], 50)), 4),
        "p99_ms": round(float(np.percentile(res["lat"], 99)), 4),
        "phase_ms": res["phase_ms"],
        "grid": {"cell_div": res["info"]["cell_div"], "cells_last": res["info"]["cells"]},
        "autotune": dict(res["tuned"], tune_s=res["tune_s"]),
    }
    log("plain: %.1f points/s, %.3f ms/slide" % (out["value"], ms_per))
    out["roofline"] = roofline(WL, res)
    log("roofline: %s" % out["roofline"])
    if not args.no_cpu:
        cb, exact, nchk = cpu_baseline(R, WL, res, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["exact_labels"] = {"slides_checked": nchk, "equal": bool(exact)}
        log("cpu baseline: %s exact=%s" % (cb, exact))
    if not args.no_batch:
        log("batched leg: %d streams x %d slides" % (args.batch_streams, args.batch_slides))
        out["batched"] = batched_leg(R, WL, args.batch_streams, args.batch_slides, 2,
                                     autotune=not args.no_autotune)
        out["roofline_batched"] = batched_roofline(WL, out["batched"])
        log("batched: %s" % out["batched"])
        log("roofline (batched launch): %s" % out["roofline_batched"])
    if not args.no_serve:
        log("served leg: %d streams x %d slides" % (args.serve_streams, args.serve_slides))
        out["served"] = served_leg(R, WL, args.serve_streams, args.serve_slides,
                                   [float(x) for x in args.serve_rates.split(",")], 64, 2, 64)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
