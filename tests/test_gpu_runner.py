"""The command-line runner end to end on every dataset format: warm-up pass
checked window by window against the on-device brute-force validator, then
the timed pass and the reference's "[Time] total:" line."""
import os
import re
import subprocess

import numpy as np
import pytest

import rtdbstream_amd as R
from rtdbstream_amd import workloads as WL

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(R.library_path()), "rtdbs_run")


def _write(path, fmt, v):
    if fmt == "geolife":
        v.astype("<f8").tofile(path)
        return
    with open(path, "w") as f:
        if fmt != "tao":
            f.write("header\n")
        for i, r in enumerate(v):
            r = [float(x) for x in r]
            if fmt == "tao":
                f.write("%r,%r,%r\n" % tuple(r))
            elif fmt == "rbf":
                f.write("%r,%r\n" % tuple(r))
            elif fmt == "eds":
                f.write("%d %r %r %d\n" % (i, r[0], r[1], 0))
            else:
                f.write("%r\n" % r[0])


CASES = [
    # format, file name, data, W, S, K, R
    ("tao", "tao_walks.txt", lambda: WL.walks3d(8000, seed=3).astype(np.float64), 4000, 500, 10, 0.5),
    ("geolife", "geolife_walks.bin", lambda: WL.walks3d(6000, seed=4).astype(np.float64), 3000, 500, 10, 0.5),
    ("rbf", "RBF_blobs.csv", lambda: WL.blobs2d(6000, seed=5).astype(np.float64), 3000, 300, 5, 0.8),
    ("eds", "EDS_blobs.txt", lambda: WL.blobs2d(5000, seed=6).astype(np.float64), 2000, 500, 5, 0.8),
    ("stock", "stock_series.csv",
     lambda: np.cumsum(np.random.default_rng(7).normal(0, 0.01, (6000, 1)), axis=0), 3000, 600, 4, 0.002),
]


@pytest.mark.parametrize("fmt,name,make,W,S,K,eps", CASES, ids=[c[0] for c in CASES])
def test_runner_checks_every_window(tmp_path, fmt, name, make, W, S, K, eps):
    path = tmp_path / name
    _write(path, fmt, make())
    r = subprocess.run([RUNNER, "-f", str(path), "-W", str(W), "-S", str(S), "-K", str(K), "-R", str(eps),
                        "--warmup", "2", "--check", "--phases", "--stats"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    n = len(make())
    slides = (n - W) // S
    assert "[Check] %d windows equal to the brute-force validator" % slides in r.stdout
    m = re.search(r"\[Time\] total: ([0-9.eE+-]+) ms", r.stdout)
    assert m and float(m.group(1)) > 0
    assert "[Step] Warmup 1" in r.stdout and "[Time] count:" in r.stdout
    assert "avg num dist calculations for each window:" in r.stdout
