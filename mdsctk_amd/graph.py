"""Python binding of include/mdsctk_graph.h: the .ssm kNN graph
(make_sysparse) and its leading spectral decomposition (decomp_sparse); the
.gsm matrix of out-of-sample rows (make_gesparse) and the Nystrom extension
(decomp_sparse_nystrom)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import knn

_bound = False


def _lib():
    global _bound
    L = knn.load()
    if not _bound:
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.mdsctk_ssm_build.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, ctypes.POINTER(i64)]
        L.mdsctk_ssm_build.restype = ctypes.c_int
        L.mdsctk_ssm_write.argtypes = [ctypes.c_char_p, i32, vp, vp, vp]
        L.mdsctk_ssm_write.restype = ctypes.c_int
        L.mdsctk_ssm_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32), ctypes.POINTER(i64),
                                      vp, vp, vp]
        L.mdsctk_ssm_read.restype = ctypes.c_int
        L.mdsctk_spectral_decomp.argtypes = [i32, vp, vp, vp, ctypes.c_double, i32, vp, vp, vp]
        L.mdsctk_spectral_decomp.restype = ctypes.c_int
        L.mdsctk_gsm_build.argtypes = [vp, vp, i64, i32, i32, ctypes.c_int, vp, vp, vp,
                                       ctypes.POINTER(i64)]
        L.mdsctk_gsm_build.restype = ctypes.c_int
        L.mdsctk_nystrom_decomp.argtypes = [i32, vp, vp, vp, i32, vp, vp, vp, ctypes.c_double, i32,
                                            vp, vp, vp, vp]
        L.mdsctk_nystrom_decomp.restype = ctypes.c_int
        L.mdsctk_io_last_error.restype = ctypes.c_char_p
        _bound = True
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(L, rc):
    if rc < 0:
        raise knn.KnnError(rc, L.mdsctk_io_last_error().decode())
    return rc


class SSM:
    """A CSC matrix as the .ssm / .gsm files hold it: n columns, row indices
    irow, values vals.  For an .ssm it is the strict lower triangle of a
    symmetric matrix; for a .gsm one column per out-of-sample row."""

    def __init__(self, n, pcol, irow, vals):
        self.n, self.pcol, self.irow, self.vals = int(n), pcol, irow, vals

    @property
    def nnz(self):
        return int(self.pcol[-1])


def ssm_build(dist, idx, k=None) -> SSM:
    """make_sysparse on kNN rows dist[n, maxk] (float64), idx[n, maxk] (int32)."""
    L = _lib()
    dist = np.ascontiguousarray(dist, np.float64)
    idx = np.ascontiguousarray(idx, np.int32)
    n, maxk = idx.shape
    k = maxk if k is None else k
    nnz = ctypes.c_int64()
    _check(L, L.mdsctk_ssm_build(_p(dist), _p(idx), n, maxk, k, None, None, None, ctypes.byref(nnz)))
    pcol = np.zeros(n + 1, np.int32)
    irow = np.zeros(nnz.value, np.int32)
    vals = np.zeros(nnz.value, np.float64)
    _check(L, L.mdsctk_ssm_build(_p(dist), _p(idx), n, maxk, k, _p(pcol), _p(irow), _p(vals),
                                 ctypes.byref(nnz)))
    return SSM(n, pcol, irow, vals)


def ssm_write(path, m: SSM):
    L = _lib()
    _check(L, L.mdsctk_ssm_write(str(path).encode(), m.n, _p(m.pcol), _p(m.irow), _p(m.vals)))


def ssm_read(path) -> SSM:
    L = _lib()
    n, nnz = ctypes.c_int32(), ctypes.c_int64()
    _check(L, L.mdsctk_ssm_read(str(path).encode(), ctypes.byref(n), ctypes.byref(nnz), None, None,
                                None))
    pcol = np.zeros(n.value + 1, np.int32)
    irow = np.zeros(nnz.value, np.int32)
    vals = np.zeros(nnz.value, np.float64)
    _check(L, L.mdsctk_ssm_read(str(path).encode(), ctypes.byref(n), ctypes.byref(nnz), _p(pcol),
                                _p(irow), _p(vals)))
    return SSM(n.value, pcol, irow, vals)


def spectral_decomp(m: SSM, sigma, nev, vectors=True):
    """Leading nev eigenpairs of D^-1/2 A D^-1/2, A = exp(-d^2/2 sigma^2) on the
    graph's edges.  Returns (evals descending, evecs[nev, n] or None,
    residuals, converged)."""
    L = _lib()
    ev = np.zeros(nev, np.float64)
    vec = np.zeros((nev, m.n), np.float64) if vectors else None
    res = np.zeros(nev, np.float64)
    got = _check(L, L.mdsctk_spectral_decomp(m.n, _p(m.pcol), _p(m.irow), _p(m.vals), float(sigma),
                                             nev, _p(ev), _p(vec), _p(res)))
    return ev, vec, res, got


def gsm_build(dist, idx, k=None, symmetric=False) -> SSM:
    """make_gesparse on kNN rows dist[m, maxk] (float64), idx[m, maxk] (int32):
    one column per row (``-s``: mirrored entries fill absent edges)."""
    L = _lib()
    dist = np.ascontiguousarray(dist, np.float64)
    idx = np.ascontiguousarray(idx, np.int32)
    n, maxk = idx.shape
    k = maxk if k is None else k
    sym = 1 if symmetric else 0
    nnz = ctypes.c_int64()
    _check(L, L.mdsctk_gsm_build(_p(dist), _p(idx), n, maxk, k, sym, None, None, None,
                                 ctypes.byref(nnz)))
    pcol = np.zeros(n + 1, np.int32)
    irow = np.zeros(nnz.value, np.int32)
    vals = np.zeros(nnz.value, np.float64)
    _check(L, L.mdsctk_gsm_build(_p(dist), _p(idx), n, maxk, k, sym, _p(pcol), _p(irow), _p(vals),
                                 ctypes.byref(nnz)))
    return SSM(n, pcol, irow, vals)


gsm_write = ssm_write
gsm_read = ssm_read


def nystrom_decomp(a: SSM, b: SSM, sigma, nev, vectors=True):
    """Leading nev eigenpairs of the in-sample graph ``a`` and their Nystrom
    extension to the out-of-sample columns of ``b``.  Returns (evals
    descending, evecs[nev, n] or None, ext[nev, m], residuals, converged)."""
    L = _lib()
    ev = np.zeros(nev, np.float64)
    vec = np.zeros((nev, a.n), np.float64) if vectors else None
    ext = np.zeros((nev, b.n), np.float64)
    res = np.zeros(nev, np.float64)
    got = _check(L, L.mdsctk_nystrom_decomp(a.n, _p(a.pcol), _p(a.irow), _p(a.vals), b.n,
                                            _p(b.pcol), _p(b.irow), _p(b.vals), float(sigma), nev,
                                            _p(ev), _p(vec), _p(ext), _p(res)))
    return ev, vec, ext, res, got
