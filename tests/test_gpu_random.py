"""Seeded random configurations against the oracle, through the C ABI.

Each case draws the dimension, window, stride, Eps, MinPts, stencil reach k
and kernel form from its seed, and the data from a mix of Gaussian blobs,
uniform noise and repeated points, so that the cases land on the cell-level
merge, the point-parallel regrouping of the count's and outputs' open
queries, the one-pass first cores, cells that cross 64-position waves and
strides from one point to the whole window. Bar: bit-exact `label` and `cid`
on every slot of every slide.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = ("0", "1", "2")  # neighbour kernel form: by size, cell-parallel, point-parallel


def _data(rng, n, dim):
    spread = 6.0
    nb = int(n * rng.uniform(0.4, 0.9))
    centres = rng.uniform(0, spread, (int(rng.integers(2, 9)), dim))
    sig = rng.uniform(0.05, 0.4)
    p = centres[rng.integers(0, len(centres), nb)] + rng.normal(0, sig, (nb, dim))
    q = rng.uniform(0, spread, (n - nb, dim))
    pts = np.concatenate([p, q]).astype(np.float32)
    pts = pts[rng.permutation(n)]
    dup = rng.random(n) < 0.05  # some points repeat an earlier one (distance 0)
    src = rng.integers(0, n, n)
    pts[dup] = pts[src[dup]]
    return pts


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    dim = int(rng.choice([1, 2, 3], p=[0.15, 0.35, 0.5]))
    # the window is a whole number of strides (the ring's units)
    units = int(rng.choice([1, 2, 5, 20, 300]))
    S = max(1, int(np.exp(rng.uniform(np.log(100), np.log(6000)))) // units)
    W = S * units
    eps = float(np.float32(rng.uniform(0.05, 0.5)))
    mp = int(rng.choice([1, 2, 4, 8, 16, 30]))
    k = int(rng.choice([0, 1, 2, 3]))
    mode = str(rng.choice(MODES))
    slides = 3
    data = _data(rng, W + slides * S, dim)
    return dim, W, S, eps, mp, k, mode, slides, data


@pytest.mark.parametrize("seed", range(64))
def test_random_stream_bit_exact(engine, oracle, monkeypatch, seed):
    dim, W, S, eps, mp, k, mode, slides, data = _case(seed)
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    ref_l, ref_c = oracle.stream(data, W, S, eps, mp, slides, method="grid")
    # a third of the cases give grid bounds inside the data's extent: the
    # points outside are clamped into border cells, which must not change a
    # label (the oracle has no grid bounds)
    rng = np.random.default_rng(7000 + seed)
    lo = hi = None
    if rng.random() < 1 / 3:
        dmin, dmax = data.min(0), data.max(0)
        ext = np.maximum(dmax - dmin, 1e-3)
        lo = (dmin + rng.uniform(0.0, 0.4, dim) * ext).tolist()
        hi = (np.array(lo) + rng.uniform(0.05, 0.6, dim) * ext).tolist()
    s = engine.Stream(dim, W, S, eps, mp, lo=lo, hi=hi, cell_div=k)
    s.init(data[:W])
    for t in range(slides):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        assert np.array_equal(l, ref_l[t]), "seed %d (dim %d W %d S %d k %d mode %s) slide %d labels" % (
            seed, dim, W, S, k, mode, t)
        assert np.array_equal(c, ref_c[t]), "seed %d slide %d ids" % (seed, t)
    s.close()


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("seed", range(3))
def test_random_mixed_batch_bit_exact(engine, oracle, monkeypatch, mode, seed):
    # six random streams of mixed dimension and reach in one launch per phase:
    # the kernels pick each stream's arrays, dimension and stencil per lane
    # group, cell or chunk
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    cases = [_case(100 + 10 * seed + i) for i in range(6)]
    refs, streams = [], []
    for dim, W, S, eps, mp, k, _, slides, data in cases:
        refs.append(oracle.stream(data, W, S, eps, mp, slides, method="grid"))
        st = engine.Stream(dim, W, S, eps, mp, cell_div=k)
        st.init(data[:W])
        streams.append(st)
    b = engine.Batch(streams)
    for t in range(3):
        strides = [c[8][c[1] + t * c[2]: c[1] + (t + 1) * c[2]] for c in cases]
        labels, cids = b.slide(strides)
        for i, (l, c) in enumerate(zip(labels, cids)):
            assert np.array_equal(l, refs[i][0][t]), "batch seed %d stream %d slide %d labels" % (seed, i, t)
            assert np.array_equal(c, refs[i][1][t]), "batch seed %d stream %d slide %d ids" % (seed, i, t)
    b.close()
    for st in streams:
        st.close()


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("W", [1, 2, 63, 64, 65, 511, 512, 513, 1025])
def test_boundary_window_sizes_bit_exact(engine, oracle, monkeypatch, W, dim, mode):
    # windows on either side of the 64-position waves and 512-position tiles
    # that the kernels index by, strides of one point and of the whole window
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    rng = np.random.default_rng(7 * W + dim)
    for S in sorted({1, W}):
        data = _data(rng, W + 3 * S, dim)
        ref_l, ref_c = oracle.stream(data, W, S, 0.3, 3, 3, method="grid")
        s = engine.Stream(dim, W, S, 0.3, 3)
        s.init(data[:W])
        for t in range(3):
            l, c = s.slide(data[W + t * S: W + (t + 1) * S])
            assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t]), "W %d S %d slide %d" % (W, S, t)
        s.close()


@pytest.mark.parametrize("seed", range(2))
def test_random_large_batch_bit_exact(engine, oracle, seed):
    # eight random streams of 100-300 K points, 1-D to 3-D and k = 1 to 3, in
    # one launch past 10^6 window points: the kernel set the engine picks by
    # size (point-parallel count and regrouped outputs, one-pass first cores,
    # XCD-ordered merge and outputs, 32 blocks per CU)
    rng = np.random.default_rng(500 + seed)
    cases, refs, streams = [], [], []
    for i in range(8):
        dim = int(rng.choice([1, 2, 3], p=[0.2, 0.4, 0.4]))
        units = int(rng.choice([10, 100]))
        S = int(rng.uniform(100_000, 300_000)) // units
        W = S * units
        # a few tens of points within eps at the blobs' density (k = 1 leaves
        # no cell safe, so the merge unites every pair of nearby cores)
        eps = float(np.float32(rng.uniform(*{1: (0.0005, 0.002), 2: (0.01, 0.03), 3: (0.03, 0.08)}[dim])))
        mp = int(rng.choice([4, 10, 20]))
        k = int(rng.choice([1, 2, 3], p=[0.2, 0.5, 0.3]))
        data = _data(rng, W + 2 * S, dim)
        cases.append((dim, W, S, data))
        refs.append(oracle.stream(data, W, S, eps, mp, 2, method="grid"))
        st = engine.Stream(dim, W, S, eps, mp, cell_div=k)
        st.init(data[:W])
        streams.append(st)
    assert sum(c[1] for c in cases) > 1_000_000
    b = engine.Batch(streams)
    for t in range(2):
        labels, cids = b.slide([c[3][c[1] + t * c[2]: c[1] + (t + 1) * c[2]] for c in cases])
        for i, (l, c) in enumerate(zip(labels, cids)):
            assert np.array_equal(l, refs[i][0][t]), "large batch seed %d stream %d slide %d labels" % (seed, i, t)
            assert np.array_equal(c, refs[i][1][t]), "large batch seed %d stream %d slide %d ids" % (seed, i, t)
    b.close()
    for st in streams:
        st.close()


@pytest.mark.parametrize("swap", [False, True])
def test_random_server_mix_bit_exact(engine, oracle, swap):
    # twelve random streams through the server from two submitting threads
    # (each stream's requests from one of them, so they run in order), small
    # batches so that the membership changes from launch to launch, and with
    # swap=True a budget of half the streams' bytes, so that streams are
    # evicted to host checkpoints and rebuilt between their requests
    import threading
    cases = [_case(200 + i) for i in range(12)]
    refs = [oracle.stream(c[8], c[1], c[2], c[3], c[4], c[7], method="grid") for c in cases]
    need = sum(engine.Stream.device_bytes(c[0], c[1], c[2], c[3], c[4], cell_div=c[5]) for c in cases)
    budget = need // 2 if swap else 8 << 30
    srv = engine.Server(hbm_budget_bytes=max(budget, max(
        engine.Stream.device_bytes(c[0], c[1], c[2], c[3], c[4], cell_div=c[5]) for c in cases)),
        max_batch=4, lanes=3, swap=swap)
    sids = [srv.open_stream(c[0], c[1], c[2], c[3], c[4], c[8][:c[1]], cell_div=c[5]) for c in cases]
    rng = np.random.default_rng(17 + int(swap))
    order = [i for i in range(12) for _ in range(3)]
    rng.shuffle(order)
    tickets = {i: [] for i in range(12)}

    def submit(part):
        nxt = {}
        for i in order:
            if i % 2 != part:
                continue
            t = nxt.get(i, 0)
            nxt[i] = t + 1
            dim, W, S = cases[i][0], cases[i][1], cases[i][2]
            tickets[i].append(srv.submit(sids[i], cases[i][8][W + t * S: W + (t + 1) * S]))

    th = [threading.Thread(target=submit, args=(p,)) for p in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for i in range(12):
        for t, tk in enumerate(tickets[i]):
            l, c = srv.wait(tk, 60000)
            assert np.array_equal(l, refs[i][0][t]), "server stream %d slide %d labels" % (i, t)
            assert np.array_equal(c, refs[i][1][t]), "server stream %d slide %d ids" % (i, t)
    srv.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_random_pipelined_batch_bit_exact(engine, oracle, use_graph):
    # six random streams of mixed dimension and reach slid by pipelined calls
    # from engine-pinned buffers: every slide writes its own label and id
    # arrays, which are complete once the next call (or drain) returns
    cases = [_case(300 + i) for i in range(6)]
    slides = 3
    refs = [oracle.stream(c[8], c[1], c[2], c[3], c[4], slides, method="grid") for c in cases]
    streams = []
    for dim, W, S, eps, mp, k, _, _, data in cases:
        st = engine.Stream(dim, W, S, eps, mp, cell_div=k)
        st.init(data[:W])
        streams.append(st)
    b = engine.Batch(streams, use_graph=use_graph)
    in_b = sum(c[2] * c[0] * 4 for c in cases) * slides
    out_b = sum(c[1] * 8 for c in cases) * slides
    hin, hout = engine.HostBuffer(in_b), engine.HostBuffer(out_b)
    oi = oo = 0
    ins, labs, cids = [], [], []
    for t in range(slides):
        row_i, row_l, row_c = [], [], []
        for dim, W, S, _, _, _, _, _, data in cases:
            a = hin.array(np.float32, S * dim, offset=oi)
            a[:] = data[W + t * S: W + (t + 1) * S].reshape(-1)
            row_i.append(a)
            oi += S * dim * 4
            row_l.append(hout.array(np.int32, W, offset=oo))
            row_c.append(hout.array(np.int32, W, offset=oo + W * 4))
            oo += W * 8
        ins.append(row_i)
        labs.append(row_l)
        cids.append(row_c)
    for t in range(slides):
        b.slide_pipelined(ins[t], labs[t], cids[t])
    b.drain()
    for t in range(slides):
        for i in range(len(cases)):
            assert np.array_equal(labs[t][i], refs[i][0][t]), "pipelined stream %d slide %d labels" % (i, t)
            assert np.array_equal(cids[t][i], refs[i][1][t]), "pipelined stream %d slide %d ids" % (i, t)
    b.close()
    for st in streams:
        st.close()
    hin.free()
    hout.free()
