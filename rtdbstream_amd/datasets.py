"""Dataset loaders for the reference runner's five on-disk formats, through
the native loader of librtdbs.so (rtdbs_dataset_load; util.cpp:12-254 in the
reference). The format comes from the file name as in the reference's main
(dbscan.cpp:2454-2469) unless given.

    tao      CSV "x,y,z", no header (3-D)
    geolife  raw little-endian double x,y,z records (3-D)
    RBF      header line, CSV "x,y" (2-D)
    EDS      header line, "id x y cid" (2-D)
    stock    header line, one value per line (1-D)
"""
import ctypes

import numpy as np

from . import Dataset, RtdbsError, _check, lib

FORMATS = {"auto": 0, "tao": 1, "geolife": 2, "rbf": 3, "eds": 4, "stock": 5}
NAMES = {v: k for k, v in FORMATS.items()}


def format_of(path):
    """Format name the file name selects, or None."""
    f = lib().rtdbs_dataset_format_of(str(path).encode())
    return NAMES[f] if f else None


def load(path, n=0, fmt="auto"):
    """Reads the first n records (all if n <= 0). Returns (points float32 of
    shape (n, dim), lo, hi, format name); lo/hi are the per-axis min/max."""
    L = lib()
    if fmt not in FORMATS:
        raise ValueError("unknown format %r" % (fmt,))
    ds = Dataset()
    _check(L.rtdbs_dataset_load(str(path).encode(), FORMATS[fmt], int(n), ctypes.byref(ds)))
    try:
        pts = np.ctypeslib.as_array(ds.points, shape=(ds.n * ds.dim,)).copy().reshape(ds.n, ds.dim)
        lo = np.array(ds.lo[:ds.dim], np.float32)
        hi = np.array(ds.hi[:ds.dim], np.float32)
        name = NAMES[ds.format]
    finally:
        L.rtdbs_dataset_free(ctypes.byref(ds))
    return pts, lo, hi, name


__all__ = ["FORMATS", "load", "format_of", "RtdbsError"]
