"""Python binding of the C ABI in include/mdsctk_knn.h (ctypes).

This is host plumbing for tests and benchmarks; the engine is the HIP
library mdsctk_amd/lib/libmdsctk_knn.so.  Loading fails loudly when the
library has not been built: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

PREC_FP32 = 0
PREC_SPLIT_BF16 = 1
PREC_SPLIT_FP16 = 2
PREC_EXACT_FP64 = 3
PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_SPLIT_BF16, "fp16": PREC_SPLIT_FP16,
              "exact": PREC_EXACT_FP64}

KNN_NOFIT = 1
KNN_FULL_ROWS = 2
KNN_KEEP_RANK0 = 4

METRIC_EUCLIDEAN = 1
METRIC_CORRELATION = 2
METRICS = {"euclidean": METRIC_EUCLIDEAN, "correlation": METRIC_CORRELATION}

OPT_HBM_BUDGET = 1
OPT_EPS_SCALE = 2
OPT_REF_SPLITS = 3
OPT_SERVE_SLOTS = 4
OPT_SERVE_GRAPHS = 6
OPT_SERVE_SHARE_CUS = 7
OPT_SELF_SYMMETRIC = 8
OPT_MIRROR_CAP = 9
OPT_SERVE_MAX_FRAMES = 5

ERR_ARG = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_IO = -4
ERR_FORMAT = -5
ERR_NATOMS = -6
ERR_STATE = -7
ERR_PENDING = -8


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_prep", ctypes.c_double),
        ("ms_contract", ctypes.c_double),
        ("ms_merge", ctypes.c_double),
        ("ms_fallback", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("n_pairs", ctypes.c_int64),
        ("n_candidates", ctypes.c_int64),
        ("n_fallback_rows", ctypes.c_int64),
        ("ref_splits", ctypes.c_int32),
        ("list_capacity", ctypes.c_int32),
        ("eps_rel", ctypes.c_double),
        ("n_pairs_computed", ctypes.c_int64),
        ("ms_landmark", ctypes.c_double),
        ("n_pairs_landmark", ctypes.c_int64),
        ("n_mirrored", ctypes.c_int64),
        ("n_mirrored_max", ctypes.c_int64),
        ("n_relisted_rows", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("n_batches", ctypes.c_int32),
        ("n_fallback_rows", ctypes.c_int64),
        ("t_submit_ms", ctypes.c_double),
        ("t_admit_ms", ctypes.c_double),
        ("t_done_ms", ctypes.c_double),
        ("batch_frames", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("requests_done", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("frames", ctypes.c_int64),
        ("pairs", ctypes.c_int64),
        ("max_batch_frames", ctypes.c_int64),
        ("fallback_rows", ctypes.c_int64),
        ("slots", ctypes.c_int32),
        ("slot_frames", ctypes.c_int32),
        ("busy_ms", ctypes.c_double),
        ("graph_launches", ctypes.c_int64),
        ("graph_builds", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib_path() -> str:
    return _build.LIB_PATH


def load():
    """Load the engine library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"mdsctk_amd: {path} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    vp, i32, i64, u32, dbl = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32,
                              ctypes.c_double)
    L.mdsctk_knn_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int]
    L.mdsctk_knn_create.restype = ctypes.c_int
    L.mdsctk_knn_destroy.argtypes = [vp]
    L.mdsctk_knn_destroy.restype = ctypes.c_int
    L.mdsctk_knn_last_error.argtypes = [vp]
    L.mdsctk_knn_last_error.restype = ctypes.c_char_p
    L.mdsctk_knn_set_option.argtypes = [vp, ctypes.c_int, dbl]
    L.mdsctk_knn_set_option.restype = ctypes.c_int
    L.mdsctk_knn_set_reference.argtypes = [vp, vp, i64, i32, vp]
    L.mdsctk_knn_set_reference.restype = ctypes.c_int
    L.mdsctk_knn_set_reference_device.argtypes = [vp, vp, i64, i32, vp]
    L.mdsctk_knn_set_reference_device.restype = ctypes.c_int
    L.mdsctk_knn_query.argtypes = [vp, vp, i64, i32, u32, vp, vp]
    L.mdsctk_knn_query.restype = ctypes.c_int
    L.mdsctk_knn_query_device.argtypes = [vp, vp, i64, i32, u32, vp, vp]
    L.mdsctk_knn_query_device.restype = ctypes.c_int
    L.mdsctk_knn_query_device_noalloc.argtypes = [vp, vp, i64, i32, u32, vp, vp]
    L.mdsctk_knn_query_device_noalloc.restype = ctypes.c_int
    L.mdsctk_knn_debug_scores.argtypes = [vp, vp, i64, u32, vp, vp]
    L.mdsctk_knn_debug_scores.restype = ctypes.c_int
    L.mdsctk_knn_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.mdsctk_knn_get_stats.restype = ctypes.c_int
    L.mdsctk_knn_row_width.argtypes = [vp, i32, u32]
    L.mdsctk_knn_row_width.restype = ctypes.c_int
    L.mdsctk_knn_submit.argtypes = [vp, vp, i64, i32, u32, vp, vp, ctypes.POINTER(i64)]
    L.mdsctk_knn_submit.restype = ctypes.c_int
    L.mdsctk_knn_poll.argtypes = [vp, i64, ctypes.POINTER(RequestInfo)]
    L.mdsctk_knn_poll.restype = ctypes.c_int
    L.mdsctk_knn_wait.argtypes = [vp, i64, ctypes.POINTER(RequestInfo)]
    L.mdsctk_knn_wait.restype = ctypes.c_int
    L.mdsctk_knn_serve_stats_get.argtypes = [vp, ctypes.POINTER(ServeStats)]
    L.mdsctk_knn_serve_stats_get.restype = ctypes.c_int
    L.mdsctk_knn_clock_ms.argtypes = [vp]
    L.mdsctk_knn_clock_ms.restype = dbl
    L.mdsctk_data_set_reference.argtypes = [vp, vp, i64, i32, ctypes.c_int]
    L.mdsctk_data_set_reference.restype = ctypes.c_int
    L.mdsctk_data_query.argtypes = [vp, vp, i64, i32, u32, vp, vp]
    L.mdsctk_data_query.restype = ctypes.c_int
    L.mdsctk_data_debug_scores.argtypes = [vp, vp, i64, vp, vp]
    L.mdsctk_data_debug_scores.restype = ctypes.c_int
    _bind_io(L)
    _lib = L
    return L


def _bind_io(L):
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    if hasattr(L, "mdsctk_xtc_info"):
        L.mdsctk_xtc_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(i64), ctypes.POINTER(i32),
                                      ctypes.POINTER(ctypes.c_float)]
        L.mdsctk_xtc_info.restype = ctypes.c_int
        L.mdsctk_xtc_read.argtypes = [ctypes.c_char_p, vp, i64, i32, ctypes.POINTER(i64), vp,
                                      vp, vp, ctypes.c_int]
        L.mdsctk_xtc_read.restype = ctypes.c_int
        L.mdsctk_xtc_write.argtypes = [ctypes.c_char_p, vp, i64, i32, ctypes.c_float, vp, vp, vp]
        L.mdsctk_xtc_write.restype = ctypes.c_int
        L.mdsctk_topology_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32), vp, i32]
        L.mdsctk_topology_read.restype = ctypes.c_int
        L.mdsctk_xtc_read_device.argtypes = [ctypes.c_char_p, i32, ctypes.c_int, ctypes.POINTER(vp),
                                             ctypes.POINTER(i64)]
        L.mdsctk_xtc_read_device.restype = ctypes.c_int
        L.mdsctk_xtc_free_device.argtypes = [vp]
        L.mdsctk_xtc_free_device.restype = ctypes.c_int
        L.mdsctk_xtc_natoms.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
        L.mdsctk_xtc_natoms.restype = ctypes.c_int
        L.mdsctk_io_last_error.argtypes = []
        L.mdsctk_io_last_error.restype = ctypes.c_char_p


class KnnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdsctk error {code}: {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class KnnContext:
    """One device context holding one resident reference set."""

    def __init__(self, device: int = 0, precision="bf16"):
        L = load()
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self.precision = precision
        h = ctypes.c_void_p()
        rc = L.mdsctk_knn_create(ctypes.byref(h), device, precision)
        if rc != 0:
            raise KnnError(rc, "mdsctk_knn_create failed")
        self._h = h
        self._L = L
        self.n_ref = 0
        self.natoms = 0

    def last_error(self) -> str:
        """The calling thread's last failure text on this context."""
        return self._L.mdsctk_knn_last_error(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdsctk_knn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise KnnError(rc, self._L.mdsctk_knn_last_error(self._h).decode())
        return rc

    def set_option(self, opt, value):
        self._check(self._L.mdsctk_knn_set_option(self._h, opt, float(value)))

    def set_reference(self, xyz, weights=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        n, natoms = xyz.shape[0], xyz.shape[1]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self._check(self._L.mdsctk_knn_set_reference(self._h, _ptr(xyz), n, natoms, _ptr(w)))
        self.n_ref, self.natoms = n, natoms

    def set_reference_device(self, d_ptr: int, n_ref: int, natoms: int, weights=None):
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self._check(self._L.mdsctk_knn_set_reference_device(self._h, ctypes.c_void_p(d_ptr),
                                                            n_ref, natoms, _ptr(w)))
        self.n_ref, self.natoms = n_ref, natoms

    def row_width(self, k, flags=0):
        return self._check(self._L.mdsctk_knn_row_width(self._h, k, flags))

    def query(self, xyz=None, k=0, flags=0):
        """Returns (dist_A[n_fit, width], idx[n_fit, width])."""
        if xyz is None:
            n_fit = self.n_ref
            xp = None
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float32)
            assert xyz.shape[1] == self.natoms and xyz.shape[2] == 3
            n_fit = xyz.shape[0]
            xp = _ptr(xyz)
        width = self.row_width(k, flags)
        d = np.zeros((n_fit, width), np.float64)
        i = np.zeros((n_fit, width), np.int32)
        got = self._check(self._L.mdsctk_knn_query(self._h, xp, n_fit, k, flags, _ptr(d), _ptr(i)))
        assert got == width
        return d, i

    def query_device(self, d_xyz, n_fit, k, flags, d_dist, d_idx):
        """Device-pointer query (ints are raw device addresses; d_xyz 0 = self)."""
        return self._check(self._L.mdsctk_knn_query_device(
            self._h, ctypes.c_void_p(d_xyz) if d_xyz else None, n_fit, k, flags,
            ctypes.c_void_p(d_dist), ctypes.c_void_p(d_idx)))

    def query_device_noalloc(self, d_xyz, n_fit, k, flags, d_dist, d_idx):
        """As query_device, through the entry point that allocates nothing
        (its workspace is the library's fixed per-GPU region)."""
        return self._check(self._L.mdsctk_knn_query_device_noalloc(
            self._h, ctypes.c_void_p(d_xyz) if d_xyz else None, n_fit, k, flags,
            ctypes.c_void_p(d_dist), ctypes.c_void_p(d_idx)))

    def debug_scores(self, xyz=None, flags=0, exact=True):
        if xyz is None:
            n_fit, xp = self.n_ref, None
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float32)
            n_fit, xp = xyz.shape[0], _ptr(xyz)
        a = np.zeros((n_fit, self.n_ref), np.float64)
        e = np.zeros((n_fit, self.n_ref), np.float64) if exact else None
        self._check(self._L.mdsctk_knn_debug_scores(self._h, xp, n_fit, flags, _ptr(a), _ptr(e)))
        return a, e

    # -- knn_data: data vectors ------------------------------------------------

    def set_data_reference(self, x, metric="euclidean"):
        """Resident FP64 data vectors x[n_ref, vdim] (knn_data)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n, vdim = x.shape
        self._check(self._L.mdsctk_data_set_reference(self._h, _ptr(x), n, vdim, METRICS[metric]))
        self.n_ref, self.vdim = n, vdim

    def data_query(self, x=None, k=0, flags=0):
        """knn_data rows: (dist[n_fit, width], idx[n_fit, width]); x None = self."""
        if x is None:
            n_fit, xp = self.n_ref, None
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            assert x.shape[1] == self.vdim
            n_fit, xp = x.shape[0], _ptr(x)
        width = self.row_width(k, flags)
        d = np.zeros((n_fit, width), np.float64)
        i = np.zeros((n_fit, width), np.int32)
        got = self._check(self._L.mdsctk_data_query(self._h, xp, n_fit, k, flags, _ptr(d), _ptr(i)))
        assert got == width
        return d, i

    def data_debug_scores(self, x=None, exact=True):
        if x is None:
            n_fit, xp = self.n_ref, None
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            n_fit, xp = x.shape[0], _ptr(x)
        a = np.zeros((n_fit, self.n_ref), np.float64)
        e = np.zeros((n_fit, self.n_ref), np.float64) if exact else None
        self._check(self._L.mdsctk_data_debug_scores(self._h, xp, n_fit, _ptr(a), _ptr(e)))
        return a, e

    def stats(self) -> dict:
        s = Stats()
        self._check(self._L.mdsctk_knn_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    # ---- serving (asynchronous requests, continuous batching)

    def submit(self, xyz, k, flags=0) -> "Request":
        """Queue a query; the returned Request owns the input copy and the
        output arrays until it is done."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        assert xyz.ndim == 3 and xyz.shape[1] == self.natoms and xyz.shape[2] == 3
        n_fit = xyz.shape[0]
        width = self.row_width(k, flags)
        d = np.zeros((n_fit, width), np.float64)
        i = np.zeros((n_fit, width), np.int32)
        rid = ctypes.c_int64()
        self._check(self._L.mdsctk_knn_submit(self._h, _ptr(xyz), n_fit, k, flags, _ptr(d), _ptr(i),
                                              ctypes.byref(rid)))
        return Request(rid.value, xyz, d, i)

    def _finish(self, req, rc, info):
        req.info = info.as_dict()
        if rc < 0:
            req.error = KnnError(rc, self._L.mdsctk_knn_last_error(self._h).decode())
            raise req.error
        req.done = True
        return req

    def poll(self, req) -> bool:
        """True when `req` is done (its outputs are final); raises on failure."""
        if req.done:
            return True
        info = RequestInfo()
        rc = self._L.mdsctk_knn_poll(self._h, req.id, ctypes.byref(info))
        if rc == ERR_PENDING:
            return False
        self._finish(req, rc, info)
        return True

    def wait(self, req) -> "Request":
        if req.done:
            return req
        info = RequestInfo()
        rc = self._L.mdsctk_knn_wait(self._h, req.id, ctypes.byref(info))
        return self._finish(req, rc, info)

    def serve_stats(self) -> dict:
        s = ServeStats()
        self._check(self._L.mdsctk_knn_serve_stats_get(self._h, ctypes.byref(s)))
        return s.as_dict()

    def clock_ms(self) -> float:
        return self._L.mdsctk_knn_clock_ms(self._h)


class Request:
    """A submitted query: `dist`/`idx` are final once `done` is True."""

    def __init__(self, rid, xyz, dist, idx):
        self.id = rid
        self.xyz = xyz
        self.dist = dist
        self.idx = idx
        self.done = False
        self.error = None
        self.info = None
