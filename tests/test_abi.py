"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/rtdbs.h declares, and rejects invalid parameters before touching a
device (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import rtdbstream_amd as R
from rtdbstream_amd import workloads as WL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "rtdbs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rtdbs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = R.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    bound = {name for name, _, _ in R.SIGNATURES}
    assert set(syms) == bound  # the Python binding covers the whole ABI


def test_version_and_error_string():
    L = R.lib()
    assert b"gfx950" in L.rtdbs_version()
    assert isinstance(L.rtdbs_last_error(), bytes)


@pytest.mark.parametrize("kw,msg", [
    (dict(dim=4), "dim"),
    (dict(dim=0), "dim"),
    (dict(window=0), "window"),
    (dict(stride=0), "stride"),
    (dict(stride=300), "divide"),      # 300 does not divide 1000
    (dict(stride=2000), "stride"),
    (dict(eps=0.0), "eps"),
    (dict(eps=-1.0), "eps"),
    (dict(eps=float("inf")), "eps"),
    (dict(min_pts=0), "min_pts"),
    (dict(cell_div=4), "cell_div"),
    (dict(lanes_per_query=3), "lanes_per_query"),
    (dict(neighbor_mode=3), "neighbor_mode"),
    (dict(cellpar_per_cu=-1), "cellpar_per_cu"),
    (dict(cellpar_per_cu=1025), "cellpar_per_cu"),
    (dict(rows_per_cu=-2), "rows_per_cu"),
    (dict(rows_per_cu=4096), "rows_per_cu"),
    (dict(window=(1 << 24) + 10, stride=10), "window"),
])
def test_invalid_params_rejected(kw, msg):
    base = dict(dim=2, window=1000, stride=100, eps=0.5, min_pts=5, cell_div=0)
    base.update(kw)
    with pytest.raises(R.RtdbsError) as e:
        R.Stream(**base)
    assert e.value.code == R.EINVAL
    assert msg in str(e.value)


def test_device_bytes_is_host_only():
    n3 = R.Stream.device_bytes(3, 100_000, 10_000, 0.5, 10)
    n2 = R.Stream.device_bytes(2, 100_000, 10_000, 0.5, 10)
    assert n3 > n2 > 100_000 * 40
    assert R.Stream.device_bytes(3, 100_000, 10_000, 0.5, 10, cell_div=1) < n3
    # grid sizes shape launches only, not the stream's memory
    assert R.Stream.device_bytes(3, 100_000, 10_000, 0.5, 10, cellpar_per_cu=48, rows_per_cu=4) == n3


def test_null_arguments_rejected():
    L = R.lib()
    assert L.rtdbs_stream_create(None, None) == R.EINVAL
    assert L.rtdbs_stream_init(None, None) == R.EINVAL
    assert L.rtdbs_stream_slide(None, None, None, None) == R.EINVAL
    assert L.rtdbs_batch_create(None, 0, 0, None) == R.EINVAL
    assert L.rtdbs_server_create(None, None) == R.EINVAL
    L.rtdbs_stream_destroy(None)
    L.rtdbs_batch_destroy(None)
    L.rtdbs_server_destroy(None)


def test_workloads_deterministic_and_shaped():
    a = WL.generate("blobs2d")
    b = WL.generate("blobs2d")
    assert a.dtype == np.float32 and a.shape == (30_000, 2) and np.array_equal(a, b)
    w = WL.walks3d(20_000, seed=1)
    assert w.shape == (20_000, 3)
    assert (w[:, 2] >= 0).all() and (w[:, 2] <= 10).all()
    specs = WL.multi_stream_specs()
    assert len(specs) == 64
    for sp in specs:
        assert 5_000 <= sp["window"] <= 200_000 and sp["window"] % 10 == 0
        assert sp["stride"] * 10 == sp["window"]
        assert 0.3 <= sp["eps"] <= 1.2 and sp["min_pts"] in (5, 10, 20)


def test_skewed_generator_is_skewed():
    p = WL.skewed3d(300_000, seed=2)
    assert WL.cell_skew(p, 1.0) > 30
