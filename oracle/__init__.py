"""CPU oracle for the knn_rms path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker or as the timed CPU baseline.  The product
(mdsctk_amd/, include/, the knn_rms executable) never imports, links or
calls it.

The numerics live in oracle.cpp (compiled to _build/liboracle.so by
oracle/build.py); this module is a thin numpy/ctypes wrapper.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        c_int, c_i64, c_dbl = ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.orc_qcp_lambda.argtypes = [f64p, c_dbl, c_dbl, ctypes.POINTER(c_int)]
        L.orc_qcp_lambda.restype = c_dbl
        L.orc_jacobi4_lambda.argtypes = [f64p]
        L.orc_jacobi4_lambda.restype = c_dbl
        L.orc_center.argtypes = [c_int, f32p, f32p, f64p, ctypes.POINTER(c_dbl)]
        L.orc_center.restype = c_dbl
        L.orc_cross_cov.argtypes = [c_int, f32p, f32p, f32p, f64p,
                                    ctypes.POINTER(c_dbl), ctypes.POINTER(c_dbl)]
        L.orc_cross_cov.restype = None
        L.orc_rmsd_nm.argtypes = [c_int, f32p, f32p, f32p, c_int]
        L.orc_rmsd_nm.restype = c_dbl
        L.orc_pairs_nm.argtypes = [f32p, f32p, c_int, f32p, c_i64, i64p, i64p, c_int, f64p, c_int]
        L.orc_pairs_nm.restype = None
        L.orc_full_exact.argtypes = [f32p, c_i64, f32p, c_i64, c_int, f32p, c_int, f64p, c_int]
        L.orc_full_exact.restype = c_int
        L.orc_knn_exact.argtypes = [f32p, c_i64, f32p, c_i64, c_int, f32p, c_int, c_int, c_int,
                                    f64p, i32p, c_int]
        L.orc_knn_exact.restype = c_int
        L.orc_port_center.argtypes = [f32p, c_i64, c_int, f32p]
        L.orc_port_center.restype = None
        L.orc_port_distance.argtypes = [c_int, f32p, f32p, f32p, c_int]
        L.orc_port_distance.restype = c_dbl
        L.orc_port_knn_rows.argtypes = [f32p, c_i64, f32p, c_i64, c_i64, c_int, f32p, c_int,
                                        c_int, c_int, f64p, i32p, c_int]
        L.orc_port_knn_rows.restype = c_int
        L.orc_euclidean.argtypes = [c_int, f64p, f64p]
        L.orc_euclidean.restype = c_dbl
        L.orc_correlation.argtypes = [c_int, f64p, f64p]
        L.orc_correlation.restype = c_dbl
        L.orc_knn_data.argtypes = [f64p, c_i64, f64p, c_i64, c_int, c_int, c_int, c_int, f64p,
                                   i32p, c_int]
        L.orc_knn_data.restype = c_int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def qcp_lambda(S, ga, gb):
    it = ctypes.c_int(0)
    lam = lib().orc_qcp_lambda(np.ascontiguousarray(S, dtype=np.float64).reshape(9), ga, gb,
                               ctypes.byref(it))
    return lam, it.value


def jacobi4_lambda(S):
    return lib().orc_jacobi4_lambda(np.ascontiguousarray(S, dtype=np.float64).reshape(9))


def center(x, w):
    """FP64 centring of one frame x[natoms,3]: returns (G, centroid[3], W)."""
    x = _f32(x)
    c = np.zeros(3, np.float64)
    W = ctypes.c_double(0)
    G = lib().orc_center(x.shape[0], x.reshape(-1), _f32(w), c, ctypes.byref(W))
    return G, c, W.value


def cross_cov(a, b, w):
    S = np.zeros(9, np.float64)
    ga, gb = ctypes.c_double(0), ctypes.c_double(0)
    lib().orc_cross_cov(a.shape[0], _f32(a).reshape(-1), _f32(b).reshape(-1), _f32(w), S,
                        ctypes.byref(ga), ctypes.byref(gb))
    return S.reshape(3, 3), ga.value, gb.value


def rmsd_nm(a, b, w, fit=True):
    a, b = _f32(a), _f32(b)
    return lib().orc_rmsd_nm(a.shape[0], a.reshape(-1), b.reshape(-1), _f32(w), int(fit))


def pairs_nm(ref, fit, w, ref_idx, fit_idx, dofit=True, nthreads=0):
    ref, fit = _f32(ref), _f32(fit)
    ri = np.ascontiguousarray(ref_idx, dtype=np.int64)
    fi = np.ascontiguousarray(fit_idx, dtype=np.int64)
    out = np.zeros(len(ri), np.float64)
    lib().orc_pairs_nm(ref.reshape(-1), fit.reshape(-1), ref.shape[1], _f32(w), len(ri), ri, fi,
                       int(dofit), out, nthreads)
    return out


def full_exact(ref, fit, w, nofit=False, nthreads=0):
    ref, fit = _f32(ref), _f32(fit)
    out = np.zeros((fit.shape[0], ref.shape[0]), np.float64)
    lib().orc_full_exact(ref.reshape(-1), ref.shape[0], fit.reshape(-1), fit.shape[0],
                         ref.shape[1], _f32(w), int(nofit), out, nthreads)
    return out


def knn_exact(ref, fit, w, k, nofit=False, keep_rank0=False, nthreads=0):
    """FP64 oracle kNN in the reference's file layout: (dist_A[n_fit,k], idx[n_fit,k])."""
    ref, fit = _f32(ref), _f32(fit)
    kk = min(k, ref.shape[0] - 1)
    d = np.zeros((fit.shape[0], max(kk, 0)), np.float64)
    i = np.zeros((fit.shape[0], max(kk, 0)), np.int32)
    kk2 = lib().orc_knn_exact(ref.reshape(-1), ref.shape[0], fit.reshape(-1), fit.shape[0],
                              ref.shape[1], _f32(w), k, int(nofit), int(keep_rank0), d, i,
                              nthreads)
    assert kk2 == kk
    return d, i


def port_center(x, w):
    """Float centring in place of a copy, as the reference does at load time."""
    x = _f32(x).copy()
    lib().orc_port_center(x.reshape(-1), x.shape[0], x.shape[1], _f32(w))
    return x


def port_knn_rows(ref_c, fit_c, w, k, row0=0, row1=None, sort=True, nofit=False, nthreads=0):
    """Reference-equivalent float knn over fit rows [row0,row1) of float-centred frames."""
    ref_c, fit_c = _f32(ref_c), _f32(fit_c)
    if row1 is None:
        row1 = fit_c.shape[0]
    n = row1 - row0
    n_ref = ref_c.shape[0]
    kk = min(k, n_ref - 1)
    width = kk if sort else n_ref
    d = np.zeros((n, width), np.float64)
    i = np.zeros((n, width), np.int32)
    lib().orc_port_knn_rows(ref_c.reshape(-1), n_ref, fit_c.reshape(-1), row0, row1,
                            ref_c.shape[1], _f32(w), k, int(sort), int(nofit), d, i, nthreads)
    return d, i


METRICS = {"euclidean": 1, "correlation": 2}


def data_distance(a, b, metric="euclidean"):
    """One knn_data distance, query vector a against reference vector b."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    f = lib().orc_euclidean if metric == "euclidean" else lib().orc_correlation
    return f(a.shape[0], a, b)


def knn_data(ref, fit, k, metric="euclidean", keep_rank0=False, nthreads=0):
    """FP64 knn_data in the reference's file layout: (dist[n_fit,k], idx[n_fit,k])."""
    ref = np.ascontiguousarray(ref, np.float64)
    fit = np.ascontiguousarray(fit, np.float64)
    kk = max(min(k, ref.shape[0] - 1), 0)
    d = np.zeros((fit.shape[0], kk), np.float64)
    i = np.zeros((fit.shape[0], kk), np.int32)
    kk2 = lib().orc_knn_data(ref.reshape(-1), ref.shape[0], fit.reshape(-1), fit.shape[0],
                             ref.shape[1], METRICS[metric], k, int(keep_rank0), d, i, nthreads)
    assert kk2 == kk
    return d, i
