"""Parity checks shared by the GPU tests: engine output vs the FP64 oracle."""
import numpy as np

import oracle

# 1e-5 nm expressed in the file unit (Angstrom).
TOL_A = 1e-4


def check_knn(ref, fit, w, d, i, d_o, i_o, nofit=False, tol=TOL_A):
    """d/i: engine rows, d_o/i_o: oracle rows, both ascending, same k.

    Passes when (1) the sorted distance rows agree within `tol`, (2) every
    returned index is distinct within its row and its exact FP64 distance
    equals the returned distance within `tol`, so index sets may differ only
    among ties closer than `tol`.
    """
    assert d.shape == d_o.shape and i.shape == i_o.shape
    if d.size == 0:
        return
    err = np.abs(d - d_o)
    assert err.max() <= tol, f"distance mismatch {err.max():.3e} A at {np.unravel_index(err.argmax(), err.shape)}"
    for r in range(i.shape[0]):
        assert len(set(i[r].tolist())) == i.shape[1], f"duplicate index in row {r}"
    rows = np.repeat(np.arange(i.shape[0]), i.shape[1])
    exact = 10.0 * oracle.pairs_nm(ref, fit, w, i.reshape(-1), rows, dofit=not nofit)
    e2 = np.abs(exact - d.reshape(-1))
    assert e2.max() <= tol, f"returned distance != exact distance of returned index ({e2.max():.3e} A)"
    same = (i == i_o)
    return float(err.max()), float(same.mean())
