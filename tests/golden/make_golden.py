"""Writes the golden vectors in tests/golden/golden_rmsd.npz.

Expected values come from an independent FP64 Kabsch solution: the SVD of
the weighted cross-covariance H = sum w (a-ca)(b-cb)^T, with the smallest
singular value's sign flipped when det(U V^T) < 0 (proper rotations only),
so RMSD^2 = (G_a + G_b - 2 (s1 + s2 + d s3)) / W.  numpy.linalg.svd is the
library routine; nothing here uses the oracle under test.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def kabsch_rmsd(a, b, w, fit=True):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    w = w.astype(np.float64)
    W = w.sum()
    A = a - (w[:, None] * a).sum(0) / W
    B = b - (w[:, None] * b).sum(0) / W
    G = (w[:, None] * A * A).sum() + (w[:, None] * B * B).sum()
    H = (w[:, None] * A).T @ B
    if fit:
        U, s, Vt = np.linalg.svd(H)
        d = np.sign(np.linalg.det(U @ Vt))
        lam = s[0] + s[1] + d * s[2]
    else:
        lam = np.trace(H)
    return np.sqrt(max(G - 2.0 * lam, 0.0) / W)


def matrix(ref, fit, w, fit_flag=True):
    return np.array([[kabsch_rmsd(r, f, w, fit_flag) for r in ref] for f in fit])


def knn_rows(M_nm, k, keep_rank0=False):
    """Reference file layout from a [n_fit, n_ref] nm matrix: ranks 1..k by
    (value, index), in Angstrom."""
    n_fit, n_ref = M_nm.shape
    k = min(k, n_ref - 1)
    off = 0 if keep_rank0 else 1
    d = np.zeros((n_fit, k))
    i = np.zeros((n_fit, k), np.int32)
    for r in range(n_fit):
        order = sorted(range(n_ref), key=lambda j: (M_nm[r, j], j))
        sel = order[off:off + k]
        d[r] = 10.0 * M_nm[r, sel]
        i[r] = sel
    return d, i


def main():
    rng = np.random.default_rng(20260101)
    out = {}
    # 1. random frames with random masses: 64 ref x 40 fit x 20 atoms
    base = np.cumsum(rng.normal(size=(20, 3)) * 0.3, axis=0)
    ref = (base[None] + 0.2 * rng.normal(size=(64, 20, 3))).astype(np.float32)
    fit = (base[None] + 0.2 * rng.normal(size=(40, 20, 3))).astype(np.float32)
    for F in (ref, fit):
        for x in F:
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            if np.linalg.det(q) < 0:
                q[:, 0] *= -1
            x[:] = (x @ q.T + rng.uniform(-5, 5, 3)).astype(np.float32)
    w = rng.uniform(1.0, 16.0, 20).astype(np.float32)
    out["rand_ref"], out["rand_fit"], out["rand_w"] = ref, fit, w
    out["rand_rmsd_fit"] = matrix(ref, fit, w, True)
    out["rand_rmsd_nofit"] = matrix(ref, fit, w, False)
    d, i = knn_rows(out["rand_rmsd_fit"], 7)
    out["rand_knn7_dist"], out["rand_knn7_idx"] = d, i

    # 2. special geometries (pairs a_j, b_j with expected nm)
    cases_a, cases_b, cases_w = [], [], []
    n = 12
    a = rng.normal(size=(n, 3))
    m = a.copy()
    m[:, 0] *= -1  # mirror image: needs a reflection, so proper RMSD > 0
    cases_a.append(a); cases_b.append(m)
    p = rng.normal(size=(n, 3)); p[:, 2] = 0.0  # planar frames
    cases_a.append(p); cases_b.append(p + 0.05 * rng.normal(size=(n, 3)))
    t = np.linspace(0, 1, n)[:, None] * np.array([[1.0, 2.0, 0.5]])  # collinear
    cases_a.append(t); cases_b.append(t[::-1] + 0.01 * rng.normal(size=(n, 3)))
    tri = np.zeros((n, 3)); tri[:3] = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]  # 3 atoms + 9 massless
    tri2 = tri.copy(); tri2[:3] = [[0, 0, 0], [1.1, 0, 0], [0, 0.9, 0.2]]
    cases_a.append(tri); cases_b.append(tri2)
    wt = np.ones(n, np.float32)
    ws = [wt, wt, wt, np.r_[np.ones(3), np.zeros(n - 3)].astype(np.float32)]
    A = np.array(cases_a, np.float32)
    B = np.array(cases_b, np.float32)
    out["geo_a"], out["geo_b"] = A, B
    out["geo_w"] = np.array(ws, np.float32)
    out["geo_rmsd"] = np.array([kabsch_rmsd(A[j], B[j], ws[j]) for j in range(len(A))])

    np.savez_compressed(os.path.join(HERE, "golden_rmsd.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_rmsd.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
