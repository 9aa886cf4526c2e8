"""MI355X streaming DBSCAN slide engine - Python bindings over the C ABI.

The engine (librtdbs.so: C++ host code + HIP kernels for gfx950) implements the
per-slide step of RTDBStream's sliding-window DBSCAN (reference
src/dbscan/dbscan.cpp:2258-2364): each slide inserts S points into a ring of W
slots, expires the S oldest, and returns slot-indexed `label` (0 core,
1 border, 2 noise) and `cid` arrays.

These bindings load the in-tree library and fail loudly if it is missing: there
is no CPU fallback on the product path.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "RtdbsError", "Params", "Stream", "Batch", "Server", "DeviceBuffer", "HostBuffer", "lib", "library_path",
    "LABEL_CORE", "LABEL_BORDER", "LABEL_NOISE", "device_count", "autotune",
]

_PKG = os.path.dirname(os.path.abspath(__file__))
# RTDBS_LIBRARY may point at another build of the same library (e.g. a
# profiling build); the default is the in-tree one.
_LIB_PATH = os.environ.get("RTDBS_LIBRARY") or os.path.join(_PKG, "librtdbs.so")

LABEL_CORE, LABEL_BORDER, LABEL_NOISE = 0, 1, 2

OK = 0
EINVAL, ENOMEM, EHIP, ESTATE, EBUDGET, ETIMEOUT, ECANCELLED, ENOTFOUND = -1, -2, -3, -4, -5, -6, -7, -8
_NAMES = {EINVAL: "EINVAL", ENOMEM: "ENOMEM", EHIP: "EHIP", ESTATE: "ESTATE", EBUDGET: "EBUDGET",
          ETIMEOUT: "ETIMEOUT", ECANCELLED: "ECANCELLED", ENOTFOUND: "ENOTFOUND"}


class RtdbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_NAMES.get(code, "E?"), code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("window", ctypes.c_int32), ("stride", ctypes.c_int32),
        ("eps", ctypes.c_float), ("min_pts", ctypes.c_int32),
        ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
        ("cell_div", ctypes.c_int32), ("lanes_per_query", ctypes.c_int32), ("neighbor_mode", ctypes.c_int32),
        ("cellpar_per_cu", ctypes.c_int32), ("rows_per_cu", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("cell_div", ctypes.c_int32), ("lanes_per_query", ctypes.c_int32), ("ms_per_slide", ctypes.c_double),
        ("from_cache", ctypes.c_int32), ("trials", ctypes.c_int32), ("key", ctypes.c_char * 96),
        ("neighbor_mode", ctypes.c_int32), ("cellpar_per_cu", ctypes.c_int32), ("rows_per_cu", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class StreamInfo(ctypes.Structure):
    _fields_ = [
        ("slides", ctypes.c_int64), ("cells", ctypes.c_int32), ("cell_div", ctypes.c_int32),
        ("rows", ctypes.c_int32), ("dense_shortcut", ctypes.c_int32), ("grid", ctypes.c_int64 * 3),
        ("cell_side", ctypes.c_double), ("lo", ctypes.c_double * 3), ("device_bytes", ctypes.c_size_t),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_int32), ("lanes", ctypes.c_int32),
        ("use_graph", ctypes.c_int32), ("admit_wait_ms", ctypes.c_int32), ("swap", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_int64), ("failed", ctypes.c_int64), ("batches", ctypes.c_int64),
        ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double), ("bytes_in_use", ctypes.c_size_t), ("streams", ctypes.c_int32),
        ("resident", ctypes.c_int32), ("swap_ins", ctypes.c_int64), ("swap_outs", ctypes.c_int64),
    ]


_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)

# (name, restype, argtypes) of every exported entry point of include/rtdbs.h
class Stats(ctypes.Structure):
    """rtdbs_stats: per-stream work counters summed over counted slides."""
    _fields_ = [("slides", ctypes.c_int64), ("cells", ctypes.c_int64), ("dense_points", ctypes.c_int64),
                ("count_queries", ctypes.c_int64), ("count_tests", ctypes.c_int64), ("merge_tests", ctypes.c_int64),
                ("border_tests", ctypes.c_int64), ("unites", ctypes.c_int64), ("merge_pairs", ctypes.c_int64),
                ("cell_parallel", ctypes.c_int32), ("pad", ctypes.c_int32), ("reserved", ctypes.c_int64 * 6)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "reserved")}


class Dataset(ctypes.Structure):
    """rtdbs_dataset: a loaded dataset (points owned by the library)."""
    _fields_ = [("format", ctypes.c_int32), ("dim", ctypes.c_int32), ("n", ctypes.c_int64),
                ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3), ("points", ctypes.POINTER(ctypes.c_float))]


SIGNATURES = [
    ("rtdbs_stream_create", ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    ("rtdbs_stream_device_bytes", ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_size_t)]),
    ("rtdbs_stream_init", ctypes.c_int, [_P, _P]),
    ("rtdbs_stream_restore", ctypes.c_int, [_P, _P, ctypes.c_int64]),
    ("rtdbs_stream_slide", ctypes.c_int, [_P, _P, _P, _P]),
    ("rtdbs_stream_slide_async", ctypes.c_int, [_P, _P, _P, _P]),
    ("rtdbs_stream_sync", ctypes.c_int, [_P]),
    ("rtdbs_stream_get_info", ctypes.c_int, [_P, ctypes.POINTER(StreamInfo)]),
    ("rtdbs_stream_get_window", ctypes.c_int, [_P, _P]),
    ("rtdbs_stream_destroy", None, [_P]),
    ("rtdbs_batch_create", ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    ("rtdbs_batch_slide", ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("rtdbs_batch_slide_async", ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("rtdbs_batch_sync", ctypes.c_int, [_P]),
    ("rtdbs_batch_slide_pipelined", ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("rtdbs_batch_drain", ctypes.c_int, [_P]),
    ("rtdbs_batch_destroy", None, [_P]),
    ("rtdbs_server_create", ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    ("rtdbs_server_open_stream", ctypes.c_int, [_P, ctypes.POINTER(Params), _P, _I32P]),
    ("rtdbs_server_close_stream", ctypes.c_int, [_P, ctypes.c_int32]),
    ("rtdbs_server_evict_stream", ctypes.c_int, [_P, ctypes.c_int32]),
    ("rtdbs_server_readmit_stream", ctypes.c_int, [_P, ctypes.c_int32]),
    ("rtdbs_server_submit", ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
    ("rtdbs_server_wait", ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32]),
    ("rtdbs_server_poll", ctypes.c_int, [_P, ctypes.c_int64]),
    ("rtdbs_server_cancel", ctypes.c_int, [_P, ctypes.c_int64]),
    ("rtdbs_server_get_stats", ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    ("rtdbs_server_reset_stats", ctypes.c_int, [_P]),
    ("rtdbs_server_destroy", None, [_P]),
    ("rtdbs_host_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    ("rtdbs_host_free", ctypes.c_int, [_P]),
    ("rtdbs_device_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    ("rtdbs_device_free", ctypes.c_int, [_P]),
    ("rtdbs_memcpy_h2d", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    ("rtdbs_memcpy_d2h", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    ("rtdbs_device_copy", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    ("rtdbs_dataset_format_of", ctypes.c_int, [ctypes.c_char_p]),
    ("rtdbs_dataset_load", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Dataset)]),
    ("rtdbs_dataset_free", None, [ctypes.POINTER(Dataset)]),
    ("rtdbs_bruteforce", ctypes.c_int, [ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _P, ctypes.c_int64, _P, _P,
                                        _P]),
    ("rtdbs_bruteforce_noalloc", ctypes.c_int, [ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _P, ctypes.c_int64,
                                                _P, _P, _P]),
    ("rtdbs_device_synchronize", ctypes.c_int, []),
    ("rtdbs_device_count", ctypes.c_int, [_I32P]),
    ("rtdbs_batch_set_stats", ctypes.c_int, [_P, ctypes.c_int]),
    ("rtdbs_batch_get_stats", ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(Stats)]),
    ("rtdbs_stream_set_stats", ctypes.c_int, [_P, ctypes.c_int]),
    ("rtdbs_stream_get_stats", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("rtdbs_batch_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("rtdbs_stream_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("rtdbs_batch_get_timing", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                              ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    ("rtdbs_stream_get_timing", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                               ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    ("rtdbs_test_fail_allocations", ctypes.c_int, [ctypes.c_int]),
    ("rtdbs_autotune", ctypes.c_int, [ctypes.POINTER(Params), _P, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.POINTER(TuneResult)]),
    ("rtdbs_last_error", ctypes.c_char_p, []),
    ("rtdbs_version", ctypes.c_char_p, []),
]

_lib = None


def library_path():
    return _LIB_PATH


def lib():
    """Loads librtdbs.so (built by __graft_entry__.build()). Raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("librtdbs.so is not built (%s); run __graft_entry__.build()" % _LIB_PATH)
        L = ctypes.CDLL(_LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise RtdbsError(rc, lib().rtdbs_last_error().decode(errors="replace"))
    return rc


def device_count():
    n = ctypes.c_int32(0)
    _check(lib().rtdbs_device_count(ctypes.byref(n)))
    return n.value


def make_params(dim, window, stride, eps, min_pts, lo=None, hi=None, cell_div=0, lanes_per_query=0,
                neighbor_mode=0, cellpar_per_cu=0, rows_per_cu=0):
    p = Params()
    p.dim, p.window, p.stride, p.eps, p.min_pts = int(dim), int(window), int(stride), float(eps), int(min_pts)
    for a in range(3):
        p.lo[a] = float(lo[a]) if lo is not None and a < len(lo) else 0.0
        p.hi[a] = float(hi[a]) if hi is not None and a < len(hi) else 0.0
    p.cell_div = int(cell_div)
    p.lanes_per_query = int(lanes_per_query)
    p.neighbor_mode = int(neighbor_mode)
    p.cellpar_per_cu = int(cellpar_per_cu)
    p.rows_per_cu = int(rows_per_cu)
    return p


def autotune(dim, window, stride, eps, min_pts, points, slides=3, lo=None, hi=None):
    """Stencil reach, kernel form and persistent grid sizes for this stream,
    from timed trial slides on `points` (W + (slides+1)*S points) or the
    tuning cache."""
    p = make_params(dim, window, stride, eps, min_pts, lo, hi)
    pts = np.ascontiguousarray(points, dtype=np.float32)
    r = TuneResult()
    _check(lib().rtdbs_autotune(ctypes.byref(p), _ptr(pts), pts.size // int(dim), int(slides), ctypes.byref(r)))
    return {"cell_div": r.cell_div, "lanes_per_query": r.lanes_per_query, "neighbor_mode": r.neighbor_mode,
            "cellpar_per_cu": r.cellpar_per_cu, "rows_per_cu": r.rows_per_cu,
            "ms_per_slide": r.ms_per_slide, "from_cache": bool(r.from_cache), "trials": r.trials,
            "key": r.key.decode()}


def bruteforce(points, eps, min_pts, dim=None):
    """On-device brute-force DBSCAN of one window (every pair tested):
    returns (label, cid, nn) with the slide step's rules, slot i = point i."""
    pts = np.asarray(points, dtype=np.float32)
    if dim is None:
        dim = 1 if pts.ndim == 1 else pts.shape[1]
    pts = np.ascontiguousarray(pts.reshape(-1), dtype=np.float32)
    n = pts.size // int(dim) if int(dim) > 0 else 0
    label = np.empty(max(n, 1), np.int32)
    cid = np.empty(max(n, 1), np.int32)
    nn = np.empty(max(n, 1), np.int32)
    _check(lib().rtdbs_bruteforce(int(dim), float(eps), int(min_pts), _ptr(pts), n, _ptr(label), _ptr(cid),
                                  _ptr(nn)))
    return label, cid, nn


def _f32(a, n, dim, what):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.size != n * dim:
        raise ValueError("%s must hold %d points of dim %d (got %d floats)" % (what, n, dim, a.size))
    return a


def _out_i32(a, n, what):
    """A caller-supplied output array (the engine writes n int32 into it), or
    a new one. Anything else - another dtype or size, a strided view, a
    read-only array - is refused: the library copies n*4 bytes into it."""
    if a is None:
        return np.empty(n, np.int32)
    if not isinstance(a, np.ndarray):
        raise ValueError("%s must be a numpy int32 array of %d elements" % (what, n))
    if a.dtype != np.int32 or a.size != n:
        raise ValueError("%s must be int32 with %d elements (got %s with %d)" % (what, n, a.dtype, a.size))
    if not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError("%s must be C-contiguous and writeable" % what)
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceBuffer:
    """Raw device allocation owned by Python (for device-resident inputs)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().rtdbs_device_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        _check(lib().rtdbs_memcpy_h2d(ctypes.c_void_p(self.ptr + offset), _ptr(arr), arr.nbytes))

    def download(self, out, offset=0):
        assert offset + out.nbytes <= self.nbytes
        _check(lib().rtdbs_memcpy_d2h(_ptr(out), ctypes.c_void_p(self.ptr + offset), out.nbytes))
        return out

    def addr(self, offset=0):
        return self.ptr + offset

    def free(self):
        if self.ptr:
            lib().rtdbs_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory from the engine (rtdbs_host_alloc). Host-buffer
    slides copy arrays that live here by DMA, without the engine's staging
    copy; `array` returns numpy views into it."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().rtdbs_host_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self._raw = (ctypes.c_char * self.nbytes).from_address(self.ptr)

    def array(self, dtype, count, offset=0):
        """A numpy view of `count` elements of `dtype` at byte `offset`."""
        dt = np.dtype(dtype)
        if offset < 0 or offset + dt.itemsize * count > self.nbytes:
            raise ValueError("view [%d, %d) outside the %d-byte buffer" % (offset, offset + dt.itemsize * count,
                                                                            self.nbytes))
        return np.frombuffer(self._raw, dtype=dt, count=count, offset=offset)

    def free(self):
        if self.ptr:
            self._raw = None
            lib().rtdbs_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """One sliding window (the reference's ScanState + slide loop body)."""

    def __init__(self, dim, window, stride, eps, min_pts, lo=None, hi=None, cell_div=0, lanes_per_query=0,
                 neighbor_mode=0, cellpar_per_cu=0, rows_per_cu=0):
        self.params = make_params(dim, window, stride, eps, min_pts, lo, hi, cell_div, lanes_per_query,
                                  neighbor_mode, cellpar_per_cu, rows_per_cu)
        self.dim, self.W, self.S = int(dim), int(window), int(stride)
        h = ctypes.c_void_p()
        _check(lib().rtdbs_stream_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h.value

    @staticmethod
    def device_bytes(dim, window, stride, eps, min_pts, cell_div=0, **kw):
        p = make_params(dim, window, stride, eps, min_pts, cell_div=cell_div, **kw)
        n = ctypes.c_size_t()
        _check(lib().rtdbs_stream_device_bytes(ctypes.byref(p), ctypes.byref(n)))
        return n.value

    def init(self, points):
        pts = _f32(points, self.W, self.dim, "first window")
        _check(lib().rtdbs_stream_init(self.h, _ptr(pts)))

    def restore(self, window_points, slides_done):
        pts = _f32(window_points, self.W, self.dim, "window")
        _check(lib().rtdbs_stream_restore(self.h, _ptr(pts), int(slides_done)))

    def slide(self, stride_points, label=None, cid=None):
        pts = _f32(stride_points, self.S, self.dim, "stride")
        label = _out_i32(label, self.W, "label")
        cid = _out_i32(cid, self.W, "cid")
        _check(lib().rtdbs_stream_slide(self.h, _ptr(pts), _ptr(label), _ptr(cid)))
        return label, cid

    def slide_async(self, stride_dev, label_dev, cid_dev):
        _check(lib().rtdbs_stream_slide_async(self.h, ctypes.c_void_p(stride_dev), ctypes.c_void_p(label_dev),
                                              ctypes.c_void_p(cid_dev)))

    def sync(self):
        _check(lib().rtdbs_stream_sync(self.h))

    def info(self):
        i = StreamInfo()
        _check(lib().rtdbs_stream_get_info(self.h, ctypes.byref(i)))
        return {"slides": i.slides, "cells": i.cells, "cell_div": i.cell_div, "rows": i.rows,
                "dense_shortcut": i.dense_shortcut, "grid": list(i.grid), "cell_side": i.cell_side,
                "lo": list(i.lo),
                "device_bytes": i.device_bytes}

    def window(self):
        out = np.empty(self.W * self.dim, np.float32)
        _check(lib().rtdbs_stream_get_window(self.h, _ptr(out)))
        return out.reshape(self.W, self.dim)

    def set_timing(self, enable=True):
        _check(lib().rtdbs_stream_set_timing(self.h, int(enable)))

    def set_stats(self, enable=True):
        """Turns the work counters on (zeroed) or off."""
        _check(lib().rtdbs_stream_set_stats(self.h, int(enable)))

    def stats(self):
        st = Stats()
        _check(lib().rtdbs_stream_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def timing(self):
        return _get_timing(lib().rtdbs_stream_get_timing, self.h)

    def close(self):
        if self.h:
            lib().rtdbs_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _get_timing(fn, h):
    ms = (ctypes.c_double * 32)()
    n = (ctypes.c_int64 * 32)()
    names = ctypes.create_string_buffer(1024)
    k = _check(fn(h, ms, n, 32, names, 1024))
    keys = names.value.decode().split(",")
    return {keys[i]: (ms[i], n[i]) for i in range(k)}


class Batch:
    """Many streams slid together: one launch per phase, optional HIP graph."""

    def __init__(self, streams, use_graph=True):
        self.streams = list(streams)
        arr = (ctypes.c_void_p * len(self.streams))(*[s.h for s in self.streams])
        h = ctypes.c_void_p()
        _check(lib().rtdbs_batch_create(arr, len(self.streams), int(use_graph), ctypes.byref(h)))
        self.h = h.value

    def slide(self, strides, labels=None, cids=None):
        n = len(self.streams)
        if len(strides) != n:
            raise ValueError("one stride per stream: %d streams, %d strides" % (n, len(strides)))
        strides = [_f32(x, s.S, s.dim, "stride") for x, s in zip(strides, self.streams)]
        if labels is None:
            labels = [None] * n
        if cids is None:
            cids = [None] * n
        if len(labels) != n or len(cids) != n:
            raise ValueError("one label and one cid array per stream")
        labels = [_out_i32(x, s.W, "label") for x, s in zip(labels, self.streams)]
        cids = [_out_i32(x, s.W, "cid") for x, s in zip(cids, self.streams)]
        A = (ctypes.c_void_p * n)(*[a.ctypes.data for a in strides])
        Lb = (ctypes.c_void_p * n)(*[a.ctypes.data for a in labels])
        C = (ctypes.c_void_p * n)(*[a.ctypes.data for a in cids])
        _check(lib().rtdbs_batch_slide(self.h, A, Lb, C))
        return labels, cids

    def slide_pipelined(self, strides, labels, cids):
        """Pipelined slide (rtdbs_batch_slide_pipelined): every array must
        live in a HostBuffer (engine-pinned memory). Returns once the previous
        pipelined call's labels and ids are in its arrays; this call's arrays
        must stay untouched until the next pipelined call, drain() or any
        other slide or sync of the batch."""
        n = len(self.streams)
        if len(strides) != n or len(labels) != n or len(cids) != n:
            raise ValueError("one stride, label and cid array per stream")
        for x, s in zip(strides, self.streams):
            if x.dtype != np.float32 or x.size != s.S * s.dim or not x.flags.c_contiguous:
                raise ValueError("stride arrays must be contiguous float32 of S*dim values")
        labels = [_out_i32(x, s.W, "label") for x, s in zip(labels, self.streams)]
        cids = [_out_i32(x, s.W, "cid") for x, s in zip(cids, self.streams)]
        A = (ctypes.c_void_p * n)(*[a.ctypes.data for a in strides])
        Lb = (ctypes.c_void_p * n)(*[a.ctypes.data for a in labels])
        C = (ctypes.c_void_p * n)(*[a.ctypes.data for a in cids])
        _check(lib().rtdbs_batch_slide_pipelined(self.h, A, Lb, C))

    def drain(self):
        """Waits for the last pipelined slide's outputs."""
        _check(lib().rtdbs_batch_drain(self.h))

    def slide_async(self, stride_ptrs, label_ptrs, cid_ptrs):
        n = len(self.streams)
        A = (ctypes.c_void_p * n)(*stride_ptrs)
        Lb = (ctypes.c_void_p * n)(*label_ptrs)
        C = (ctypes.c_void_p * n)(*cid_ptrs)
        _check(lib().rtdbs_batch_slide_async(self.h, A, Lb, C))

    def sync(self):
        _check(lib().rtdbs_batch_sync(self.h))

    def set_timing(self, enable=True):
        _check(lib().rtdbs_batch_set_timing(self.h, int(enable)))

    def set_stats(self, enable=True):
        _check(lib().rtdbs_batch_set_stats(self.h, int(enable)))

    def stats(self, index):
        st = Stats()
        _check(lib().rtdbs_batch_get_stats(self.h, int(index), ctypes.byref(st)))
        return st.as_dict()

    def timing(self):
        return _get_timing(lib().rtdbs_batch_get_timing, self.h)

    def close(self):
        if self.h:
            lib().rtdbs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching slide server with HBM-budget admission."""

    def __init__(self, hbm_budget_bytes, max_batch=64, lanes=3, use_graph=False, admit_wait_ms=0, swap=False):
        cfg = ServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_batch = int(max_batch)
        cfg.lanes = int(lanes)
        cfg.use_graph = int(use_graph)
        cfg.admit_wait_ms = int(admit_wait_ms)
        cfg.swap = int(bool(swap))
        h = ctypes.c_void_p()
        _check(lib().rtdbs_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h.value
        self._keep = {}
        self._meta = {}

    def open_stream(self, dim, window, stride, eps, min_pts, first_window, lo=None, hi=None, cell_div=0):
        p = make_params(dim, window, stride, eps, min_pts, lo, hi, cell_div)
        pts = _f32(first_window, window, dim, "first window")
        sid = ctypes.c_int32()
        _check(lib().rtdbs_server_open_stream(self.h, ctypes.byref(p), _ptr(pts), ctypes.byref(sid)))
        self._meta[sid.value] = (int(dim), int(window), int(stride))
        return sid.value

    def close_stream(self, sid):
        _check(lib().rtdbs_server_close_stream(self.h, int(sid)))
        self._meta.pop(sid, None)

    def evict(self, sid):
        _check(lib().rtdbs_server_evict_stream(self.h, int(sid)))

    def readmit(self, sid):
        _check(lib().rtdbs_server_readmit_stream(self.h, int(sid)))

    def submit(self, sid, stride_points, label=None, cid=None):
        dim, W, S = self._meta[sid]
        pts = _f32(stride_points, S, dim, "stride")
        label = _out_i32(label, W, "label")
        cid = _out_i32(cid, W, "cid")
        t = ctypes.c_int64()
        _check(lib().rtdbs_server_submit(self.h, int(sid), _ptr(pts), _ptr(label), _ptr(cid), ctypes.byref(t)))
        self._keep[t.value] = (pts, label, cid)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        try:
            _check(lib().rtdbs_server_wait(self.h, int(ticket), int(timeout_ms)))
        except RtdbsError as e:
            if e.code != ETIMEOUT:
                self._keep.pop(ticket, None)
            raise
        _, label, cid = self._keep.pop(ticket)
        return label, cid

    def poll(self, ticket):
        rc = lib().rtdbs_server_poll(self.h, int(ticket))
        if rc < 0:
            self._keep.pop(ticket, None)
            _check(rc)
        return rc == 1

    def result(self, ticket):
        _, label, cid = self._keep.pop(ticket)
        return label, cid

    def cancel(self, ticket):
        _check(lib().rtdbs_server_cancel(self.h, int(ticket)))

    def stats(self):
        s = ServerStats()
        _check(lib().rtdbs_server_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().rtdbs_server_reset_stats(self.h))

    def close(self):
        if self.h:
            lib().rtdbs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
