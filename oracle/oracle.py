"""ctypes loader for the C oracle (oracle/rtdbs_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
CPU-baseline leg as the checker. The product path never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "rtdbs_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")

# -ffp-contract=off keeps ((dx*dx + dy*dy) + dz*dz) unfused, as in the engine.
CFLAGS = ["-O3", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c11"]

_lib = None


def build():
    tmp = LIB + ".tmp"
    subprocess.run(["gcc"] + CFLAGS + ["-o", tmp, SRC, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.oracle_bruteforce.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, P, P]
        L.oracle_grid.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, P, P]
        L.oracle_neighbor_counts.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, P]
        L.oracle_stream.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]
        L.oracle_eps2.argtypes = [ctypes.c_float]
        L.oracle_eps2.restype = ctypes.c_float
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _pts(points):
    a = np.ascontiguousarray(points, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return a


def _rc(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d" % rc)


def eps2(eps):
    return load().oracle_eps2(float(eps))


def neighbor_counts(points, eps):
    a = _pts(points)
    nn = np.empty(a.shape[0], np.int32)
    _rc(load().oracle_neighbor_counts(a.ctypes.data, a.shape[0], a.shape[1], float(eps), nn.ctypes.data))
    return nn


def bruteforce(points, eps, min_pts):
    a = _pts(points)
    n = a.shape[0]
    label = np.empty(n, np.int32)
    cid = np.empty(n, np.int32)
    _rc(load().oracle_bruteforce(a.ctypes.data, n, a.shape[1], float(eps), int(min_pts), label.ctypes.data,
                                 cid.ctypes.data))
    return label, cid


def grid(points, eps, min_pts, threads=0):
    a = _pts(points)
    n = a.shape[0]
    label = np.empty(n, np.int32)
    cid = np.empty(n, np.int32)
    _rc(load().oracle_grid(a.ctypes.data, n, a.shape[1], float(eps), int(min_pts), int(threads),
                           label.ctypes.data, cid.ctypes.data))
    return label, cid


def stream(data, window, stride, eps, min_pts, n_slides, method="grid", threads=0):
    """Labels of every slide of the reference's ring (dbscan.cpp:2259-2353)."""
    a = _pts(data)
    labels = np.empty((n_slides, window), np.int32)
    cids = np.empty((n_slides, window), np.int32)
    _rc(load().oracle_stream(a.ctypes.data, a.shape[0], a.shape[1], int(window), int(stride), float(eps),
                             int(min_pts), int(n_slides), 0 if method == "bruteforce" else 1, int(threads),
                             labels.ctypes.data, cids.ctypes.data))
    return labels, cids


def max_threads():
    return load().oracle_max_threads()
