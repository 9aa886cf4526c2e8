"""Native dataset loaders (rtdbs_dataset_load) for the reference runner's five
formats, and the command-line runner's argument handling (no GPU needed:
loading and parsing happen before any device call)."""
import os
import subprocess

import numpy as np
import pytest

import rtdbstream_amd as R
from rtdbstream_amd import datasets as D

RUNNER = os.path.join(os.path.dirname(R.library_path()), "rtdbs_run")


def _values(n, dim, seed):
    rng = np.random.default_rng(seed)
    return np.round(rng.uniform(-50, 50, (n, dim)), 4)


def _write(path, fmt, v):
    if fmt == "geolife":
        v.astype("<f8").tofile(path)
        return
    with open(path, "w") as f:
        if fmt != "tao":
            f.write({"rbf": "x,y", "eds": "id x y cid", "stock": "price"}[fmt] + "\n")
        for i, r in enumerate(v):
            r = [float(x) for x in r]
            if fmt == "tao":
                f.write("%r,%r,%r\n" % tuple(r))
            elif fmt == "rbf":
                f.write("%r,%r\n" % tuple(r))
            elif fmt == "eds":
                f.write("%d %r %r %d\n" % (i, r[0], r[1], i % 7))
            else:
                f.write("%r\n" % r[0])


DIMS = {"tao": 3, "geolife": 3, "rbf": 2, "eds": 2, "stock": 1}
FILES = {"tao": "tao_small.txt", "geolife": "geolife_small.bin", "rbf": "RBF_small.csv", "eds": "EDS_small.txt",
         "stock": "stock_small.csv"}


@pytest.mark.parametrize("fmt", sorted(DIMS))
def test_format_round_trip(tmp_path, fmt):
    v = _values(257, DIMS[fmt], 11)
    path = tmp_path / FILES[fmt]
    _write(path, fmt, v)
    assert D.format_of(path) == fmt
    pts, lo, hi, name = D.load(path)
    assert name == fmt and pts.shape == v.shape
    # stored as FP32 after reading as double, as the reference's float3 window
    assert np.array_equal(pts, v.astype(np.float64).astype(np.float32))
    assert np.array_equal(lo, pts.min(axis=0)) and np.array_equal(hi, pts.max(axis=0))
    first, _, _, _ = D.load(path, n=100)
    assert np.array_equal(first, pts[:100])


def test_explicit_format_overrides_name(tmp_path):
    v = _values(10, 2, 3)
    path = tmp_path / "points.csv"
    _write(path, "rbf", v)
    assert D.format_of(path) is None
    with pytest.raises(R.RtdbsError):
        D.load(path)
    pts, _, _, name = D.load(path, fmt="rbf")
    assert name == "rbf" and pts.shape == (10, 2)


def test_name_rule_follows_reference_order():
    # tao is tested before RBF, geolife before EDS (dbscan.cpp:2454-2469)
    assert D.format_of("RBF_tao.txt") == "tao"
    assert D.format_of("EDS_geolife.bin") == "geolife"
    assert D.format_of("rbf.csv") is None  # case-sensitive, as the reference


def test_load_errors(tmp_path):
    path = tmp_path / "tao_short.txt"
    _write(path, "tao", _values(5, 3, 1))
    with pytest.raises(R.RtdbsError, match="5 records"):
        D.load(path, n=6)
    bad = tmp_path / "tao_bad.txt"
    bad.write_text("1,2,3\n4,five,6\n")
    with pytest.raises(R.RtdbsError, match="line 2"):
        D.load(bad)
    with pytest.raises(R.RtdbsError):
        D.load(tmp_path / "tao_missing.txt")
    empty = tmp_path / "stock_empty.csv"
    empty.write_text("header\n")
    with pytest.raises(R.RtdbsError):
        D.load(empty)


def test_trailing_blank_line_ignored(tmp_path):
    path = tmp_path / "tao_blank.txt"
    path.write_text("1,2,3\n4,5,6\n\n")
    pts, _, _, _ = D.load(path)
    assert pts.shape == (2, 3)


def _run(*args):
    return subprocess.run([RUNNER] + [str(a) for a in args], capture_output=True, text=True, timeout=60)


def test_runner_usage_and_argument_errors(tmp_path):
    assert os.path.exists(RUNNER)
    r = _run("--help")
    assert r.returncode == 1 and "--window | -W" in r.stderr
    r = _run("-f", "x")  # window, stride, minPts and radius missing
    assert r.returncode == 1 and "Usage" in r.stderr
    r = _run("--bogus")
    assert r.returncode == 1 and "Unknown option '--bogus'" in r.stderr
    r = _run("-f", tmp_path / "tao_none.txt", "-W", 10, "-S", 5, "-K", 2, "-R", 0.5)
    assert r.returncode == 1 and "[Error] Dataset" in r.stdout
    path = tmp_path / "tao_tiny.txt"
    _write(path, "tao", _values(12, 3, 2))
    r = _run("-f", path, "-W", 10, "-S", 5, "-K", 2, "-R", 0.5)
    assert r.returncode == 1 and "no slide" in r.stdout and "DIM[2]" in r.stdout
