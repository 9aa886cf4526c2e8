"""Parity of the HIP engine with the oracle, through the C ABI.

Bar: bit-exact `label` and `cid` on every slot of every slide (integer
output; the FP32 distance test is evaluated in the same pinned order on both
sides). Inputs are seeded; sizes stay where the oracle finishes in seconds.
"""
import numpy as np
import pytest

from rtdbstream_amd import workloads as WL
from tests.refcheck import reference_check, window_at

pytestmark = pytest.mark.gpu


def engine_slides(R, data, dim, W, S, eps, mp, slides, cell_div=0, lo=None, hi=None):
    s = R.Stream(dim, W, S, eps, mp, lo=lo, hi=hi, cell_div=cell_div)
    s.init(data[:W])
    out = []
    for t in range(slides):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        out.append((l.copy(), c.copy()))
    s.close()
    return out


def assert_parity(got, ref_labels, ref_cids, what):
    for t, (l, c) in enumerate(got):
        if not np.array_equal(l, ref_labels[t]):
            i = int(np.flatnonzero(l != ref_labels[t])[0])
            pytest.fail("%s slide %d: label[%d] = %d, oracle %d" % (what, t, i, l[i], ref_labels[t][i]))
        if not np.array_equal(c, ref_cids[t]):
            i = int(np.flatnonzero(c != ref_cids[t])[0])
            pytest.fail("%s slide %d: cid[%d] = %d, oracle %d (label %d)" % (what, t, i, c[i], ref_cids[t][i], l[i]))


def _mixed(n, dim, seed, spread=6.0, sigma=0.25, noise=0.3):
    rng = np.random.default_rng(seed)
    nb = int(n * (1 - noise))
    c = rng.uniform(0, spread, (6, dim))
    p = c[rng.integers(0, 6, nb)] + rng.normal(0, sigma, (nb, dim))
    q = rng.uniform(0, spread, (n - nb, dim))
    allp = np.concatenate([p, q]).astype(np.float32)
    return allp[rng.permutation(n)]


SMALL = [
    # dim, W, S, eps, min_pts, cell_div, seed
    (2, 512, 64, 0.2, 4, 0, 1),
    (2, 512, 512, 0.2, 4, 0, 2),     # S == W: whole window replaced
    (3, 512, 128, 0.4, 5, 0, 3),
    (1, 256, 32, 0.01, 3, 0, 4),
    (2, 2048, 256, 0.15, 6, 1, 5),   # k = 1: no dense shortcut
    (2, 2048, 256, 0.15, 6, 2, 6),
    (2, 2048, 256, 0.15, 6, 3, 7),
    (3, 2048, 512, 0.3, 10, 1, 8),
    (3, 2048, 512, 0.3, 10, 2, 9),
    (3, 2048, 512, 0.3, 10, 3, 10),
    (2, 1000, 100, 0.3, 1, 0, 11),   # min_pts 1: everything core
    (3, 1000, 250, 0.3, 2000, 0, 12),  # min_pts > W: all noise
    (2, 4096, 2048, 0.1, 5, 2, 13),  # stride sort across tiles
    (3, 6000, 3000, 0.25, 8, 2, 14), # merge passes
]


@pytest.mark.parametrize("dim,W,S,eps,mp,k,seed", SMALL)
def test_small_streams_bit_exact(engine, oracle, dim, W, S, eps, mp, k, seed):
    slides = 6
    data = _mixed(W + slides * S, dim, seed)
    ref_l, ref_c = oracle.stream(data, W, S, eps, mp, slides, method="grid")
    got = engine_slides(engine, data, dim, W, S, eps, mp, slides, cell_div=k)
    assert_parity(got, ref_l, ref_c, "dim%d W%d S%d k%d" % (dim, W, S, k))


def test_lattice_ties(engine, oracle):
    # points on a 0.5 lattice with eps = 0.5: every nearest pair is an exact tie
    rng = np.random.default_rng(21)
    W, S, slides = 1024, 128, 5
    data = (rng.integers(0, 24, (W + slides * S, 2)).astype(np.float32) * np.float32(0.5))
    for mp in (2, 3, 5):
        for k in (1, 2, 3):
            ref_l, ref_c = oracle.stream(data, W, S, 0.5, mp, slides, method="grid")
            got = engine_slides(engine, data, 2, W, S, 0.5, mp, slides, cell_div=k)
            assert_parity(got, ref_l, ref_c, "lattice mp%d k%d" % (mp, k))


def test_duplicate_points(engine, oracle):
    # many identical points (d2 = 0) and a dense cell far above min_pts
    rng = np.random.default_rng(22)
    W, S, slides = 800, 200, 4
    base = rng.uniform(0, 2, (40, 3)).astype(np.float32)
    data = base[rng.integers(0, 40, W + slides * S)]
    ref_l, ref_c = oracle.stream(data, W, S, 0.3, 6, slides, method="grid")
    got = engine_slides(engine, data, 3, W, S, 0.3, 6, slides)
    assert_parity(got, ref_l, ref_c, "duplicates")


def test_out_of_bounds_points_clamped_exactly(engine, oracle):
    # explicit grid bounds much smaller than the data: points outside are
    # clamped into border cells, which must not change any label
    data = _mixed(1200 + 5 * 300, 2, 23, spread=10.0)
    ref_l, ref_c = oracle.stream(data, 1200, 300, 0.35, 4, 5, method="grid")
    got = engine_slides(engine, data, 2, 1200, 300, 0.35, 4, 5, lo=[4.0, 4.0], hi=[5.0, 5.0])
    assert_parity(got, ref_l, ref_c, "clamped")


def test_blobs2d_full_config(engine, oracle):
    cfg = WL.WORKLOADS["blobs2d"]
    data = WL.generate("blobs2d")
    W, S, n = cfg["window"], cfg["stride"], cfg["slides"]
    ref_l, ref_c = oracle.stream(data, W, S, cfg["eps"], cfg["min_pts"], n, method="grid")
    got = engine_slides(engine, data, 2, W, S, cfg["eps"], cfg["min_pts"], n)
    assert_parity(got, ref_l, ref_c, "blobs2d")
    # the reference's own acceptance rules on the last window
    win = window_at(data, W, S, n - 1)
    assert reference_check(win, cfg["eps"], got[-1][0], got[-1][1], ref_l[-1], ref_c[-1]) == []


def test_repeat_runs_identical(engine):
    data = _mixed(3000 + 4 * 500, 3, 31)
    a = engine_slides(engine, data, 3, 3000, 500, 0.3, 6, 4)
    b = engine_slides(engine, data, 3, 3000, 500, 0.3, 6, 4)
    for (l1, c1), (l2, c2) in zip(a, b):
        assert np.array_equal(l1, l2) and np.array_equal(c1, c2)


def test_walks3d_full_config_first_slides(engine, oracle):
    cfg = WL.WORKLOADS["walks3d"]
    W, S, slides = cfg["window"], cfg["stride"], 4
    data = WL.generate("walks3d", n=W + slides * S)
    ref_l, ref_c = oracle.stream(data, W, S, cfg["eps"], cfg["min_pts"], slides, method="grid")
    got = engine_slides(engine, data, 3, W, S, cfg["eps"], cfg["min_pts"], slides)
    assert_parity(got, ref_l, ref_c, "walks3d")


def test_skewed3d_full_window(engine, oracle):
    """W = 1,000,000, S = 50,000, Eps 1, MinPts 20: two slides against the
    oracle, then size-independent checks on later slides."""
    cfg = WL.WORKLOADS["skewed3d"]
    W, S = cfg["window"], cfg["stride"]
    data = WL.generate("skewed3d", n=W + 4 * S)
    ref_l, ref_c = oracle.stream(data, W, S, cfg["eps"], cfg["min_pts"], 2, method="grid")
    s = engine.Stream(3, W, S, cfg["eps"], cfg["min_pts"])
    s.init(data[:W])
    for t in range(4):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        if t < 2:
            assert_parity([(l, c)], ref_l[t:t + 1], ref_c[t:t + 1], "skewed3d slide %d" % t)
        else:
            # invariants: labels in range; noise keeps its slot; a core's id is a
            # core slot no larger than its own; every id names a core
            assert set(np.unique(l)) <= {0, 1, 2}
            noise = np.flatnonzero(l == 2)
            assert np.array_equal(c[noise], noise)
            core = l == 0
            assert np.all(c[core] <= np.flatnonzero(core))
            assert np.all(l[c[l != 2]] == 0)
            assert np.all(c[c[core]] == c[core])
    s.close()


def test_autotune_choice_is_exact(engine, oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("RTDBS_TUNE_CACHE", str(tmp_path / "tune.txt"))
    W, S, slides = 4000, 400, 3
    data = _mixed(W + 8 * S, 3, 41)
    r = engine.autotune(3, W, S, 0.3, 6, data, slides=slides)
    assert r["cell_div"] in (1, 2, 3) and r["neighbor_mode"] in (1, 2)
    assert r["lanes_per_query"] in ((0,) if r["neighbor_mode"] == 1 else (1, 2, 4))
    assert r["cellpar_per_cu"] in (8, 12, 16, 24, 32, 48) and r["rows_per_cu"] in (4, 8, 16, 32)
    # 3 reaches and 4 kernel forms timed twice each, 6 cell-parallel and 4 rows grid sizes
    assert not r["from_cache"] and r["trials"] == 24 and r["ms_per_slide"] > 0
    r2 = engine.autotune(3, W, S, 0.3, 6, data, slides=slides)
    assert r2["from_cache"] and r2["trials"] == 0
    keys = ("cell_div", "neighbor_mode", "lanes_per_query", "cellpar_per_cu", "rows_per_cu")
    assert [r2[k] for k in keys] == [r[k] for k in keys]
    ref_l, ref_c = oracle.stream(data, W, S, 0.3, 6, 5, method="grid")
    s = engine.Stream(3, W, S, 0.3, 6, cell_div=r["cell_div"], lanes_per_query=r["lanes_per_query"],
                      neighbor_mode=r["neighbor_mode"], cellpar_per_cu=r["cellpar_per_cu"],
                      rows_per_cu=r["rows_per_cu"])
    s.init(data[:W])
    for t in range(5):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("cp,rows", [(1, 1), (8, 4), (48, 32), (1024, 1024)])
def test_grid_sizes_agree(engine, oracle, cp, rows):
    # workgroups per CU of the persistent kernels change how cells are dealt
    # out, never the result: from one workgroup per CU (every wave walks many
    # cells) to more workgroups than cells
    data = _mixed(4000 + 4 * 400, 3, 43)
    ref_l, ref_c = oracle.stream(data, 4000, 400, 0.3, 6, 4, method="grid")
    s = engine.Stream(3, 4000, 400, 0.3, 6, cellpar_per_cu=cp, rows_per_cu=rows)
    s.init(data[:4000])
    for t in range(4):
        l, c = s.slide(data[4000 + t * 400: 4000 + (t + 1) * 400])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("mode", ["1", "2"])
def test_xcd_grouped_work_order_agrees(engine, oracle, mode, monkeypatch):
    # the merge and the point-parallel output kernel with each XCD's
    # workgroups on one contiguous range of cells or positions; W = 3000 gives
    # grids that are not multiples of the 8 XCDs (750 cell-parallel blocks,
    # 47 chunks), so the regrouping must cover every block exactly once
    monkeypatch.setenv("RTDBS_XCD_GROUPS", "3")
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode)
    data = _mixed(3000 + 4 * 300, 3, 44)
    ref_l, ref_c = oracle.stream(data, 3000, 300, 0.3, 5, 4, method="grid")
    s = engine.Stream(3, 3000, 300, 0.3, 5)
    s.init(data[:3000])
    for t in range(4):
        l, c = s.slide(data[3000 + t * 300: 3000 + (t + 1) * 300])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("regroup", ["0", "1"])
@pytest.mark.parametrize("dim", [2, 3])
def test_point_outputs_regrouped_agree(engine, oracle, dim, regroup, monkeypatch):
    # point-parallel outputs with one lane per core and the non-core points
    # regrouped onto four lanes each (1), or four lanes for every position (0);
    # 30% noise puts more than 16 non-core points into some 64-position chunks
    # (several regrouping rounds), and W = 3000 ends in a partial chunk
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", "2")
    monkeypatch.setenv("RTDBS_BORDER_REGROUP", regroup)
    data = _mixed(3000 + 4 * 300, dim, 47)
    ref_l, ref_c = oracle.stream(data, 3000, 300, 0.3, 5, 4, method="grid")
    s = engine.Stream(dim, 3000, 300, 0.3, 5)
    s.init(data[:3000])
    for t in range(4):
        l, c = s.slide(data[3000 + t * 300: 3000 + (t + 1) * 300])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("k", [1, 2])
def test_point_form_first_cores_agree(engine, oracle, k, monkeypatch):
    # the point-parallel count's follow-up pass finds every cell's first core
    # from its wave's core flags or, for a cell that crosses a 64-position
    # wave, by scanning: dense blobs put well over 64 points into many cells
    # (k = 1: side eps, no cell safe; k = 2: dense cells)
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", "2")
    data = _mixed(6000 + 4 * 600, 3, 48, sigma=0.2)
    ref_l, ref_c = oracle.stream(data, 6000, 600, 0.3, 20, 4, method="grid")
    s = engine.Stream(3, 6000, 600, 0.3, 20, cell_div=k)
    s.init(data[:6000])
    for t in range(4):
        l, c = s.slide(data[6000 + t * 600: 6000 + (t + 1) * 600])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("lpq", [1, 2, 4])
def test_lanes_per_query_settings_agree(engine, oracle, lpq, monkeypatch):
    # lanes per query shape the point-parallel count and border kernels
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", "2")
    data = _mixed(3000 + 4 * 300, 2, 42)
    ref_l, ref_c = oracle.stream(data, 3000, 300, 0.2, 5, 4, method="grid")
    s = engine.Stream(2, 3000, 300, 0.2, 5, lanes_per_query=lpq)
    s.init(data[:3000])
    for t in range(4):
        l, c = s.slide(data[3000 + t * 300: 3000 + (t + 1) * 300])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("lpq", [1, 2, 4])
@pytest.mark.parametrize("case", ["long1d", "dense2d", "sparse3d"])
def test_flat_walks_agree(engine, oracle, case, lpq, monkeypatch):
    # the point-parallel count and outputs walk a group of stencil rows as one
    # sequence, 8 candidates per round across row boundaries: rows of dozens
    # of candidates (1-D, many rounds per row), high MinPts on dense blobs
    # (every group of rows walked), and sparse 3-D data (most rows empty,
    # groups without a candidate)
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", "2")
    W, S = 3000, 300
    if case == "long1d":
        dim, eps, mp = 1, 0.05, 30
        data = np.random.default_rng(61).uniform(0, 10, (W + 4 * S, 1)).astype(np.float32)
    elif case == "dense2d":
        dim, eps, mp = 2, 0.3, 40
        data = _mixed(W + 4 * S, 2, 62, sigma=0.4, noise=0.2)
    else:
        dim, eps, mp = 3, 0.5, 4
        data = np.random.default_rng(63).uniform(0, 12, (W + 4 * S, 3)).astype(np.float32)
    ref_l, ref_c = oracle.stream(data, W, S, eps, mp, 4, method="grid")
    assert 0 < int(np.sum(ref_l[-1] == 0)) < W  # cores and non-cores both
    s = engine.Stream(dim, W, S, eps, mp, lanes_per_query=lpq)
    s.init(data[:W])
    for t in range(4):
        l, c = s.slide(data[W + t * S: W + (t + 1) * S])
        assert np.array_equal(l, ref_l[t]) and np.array_equal(c, ref_c[t])
    s.close()


@pytest.mark.parametrize("mode", ["1", "2", "2u"])
@pytest.mark.parametrize("dim,W,S,eps,mp,k,seed", [SMALL[2], SMALL[6], SMALL[9], SMALL[13]])
def test_neighbour_kernel_forms_agree(engine, oracle, monkeypatch, mode, dim, W, S, eps, mp, k, seed):
    """Cell-parallel (1) and point-parallel (2) count/border kernels, and the
    cell-level merge kernels (2u), all reproduce the oracle (launches pick a
    form by size; each is forced here)."""
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", mode[0])
    if mode == "2u":
        monkeypatch.setenv("RTDBS_UNION_CELLS", "1")
    slides = 4
    data = _mixed(W + slides * S, dim, seed)
    ref_l, ref_c = oracle.stream(data, W, S, eps, mp, slides, method="grid")
    got = engine_slides(engine, data, dim, W, S, eps, mp, slides, cell_div=k)
    assert_parity(got, ref_l, ref_c, "mode %s dim%d W%d k%d" % (mode, dim, W, k))


@pytest.mark.parametrize("split", ["0", "1"])
@pytest.mark.parametrize("dim,W,S,eps,mp,seed", [
    (2, 20000, 5000, 0.04, 6, 61),   # 3 stride tiles, 10 runs, a partial last run and tile
    (3, 12288, 2048, 0.12, 8, 62),   # exactly one stride tile, 4 runs
    (3, 5600, 700, 0.2, 8, 63),      # two runs, the second partial
])
def test_stride_rank_forms_agree(engine, oracle, monkeypatch, split, dim, W, S, eps, mp, seed):
    """The stride rank's 256-key blocks (0) and 64-key split blocks (1) -
    launches pick one by size, each is forced here - reproduce the oracle,
    with strides of several sort runs and tiles, partial runs included."""
    monkeypatch.setenv("RTDBS_RANK_SPLIT", split)
    slides = 3
    data = _mixed(W + slides * S, dim, seed)
    ref_l, ref_c = oracle.stream(data, W, S, eps, mp, slides, method="grid")
    got = engine_slides(engine, data, dim, W, S, eps, mp, slides)
    assert_parity(got, ref_l, ref_c, "rank split %s dim%d W%d S%d" % (split, dim, W, S))


@pytest.mark.parametrize("blpq", ["1", "2", "4"])
def test_output_kernel_lanes_agree(engine, oracle, blpq, monkeypatch):
    """The point-parallel output kernel on 1, 2 or 4 lanes per query, with the
    count on its own setting (large batches pair a 1-lane count with a 4-lane
    output kernel), reproduces the oracle."""
    monkeypatch.setenv("RTDBS_NEIGHBOR_MODE", "2")
    monkeypatch.setenv("RTDBS_BORDER_LPQ", blpq)
    W, S, slides = 3000, 300, 4
    data = _mixed(W + slides * S, 3, 44)
    ref_l, ref_c = oracle.stream(data, W, S, 0.3, 6, slides, method="grid")
    got = engine_slides(engine, data, 3, W, S, 0.3, 6, slides)
    assert_parity(got, ref_l, ref_c, "output lanes %s" % blpq)
