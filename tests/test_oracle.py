"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Independent checks: numpy SVD-Kabsch golden vectors (tests/golden), a
long-double Jacobi eigen-solve of the 4x4 key matrix, closed forms (rigid
copies, the 2- and 3-atom cases, symmetry, weight scaling, a small isotropic
perturbation) and the reference-equivalent float port."""
import os

import numpy as np
import pytest

import oracle
from mdsctk_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden_rmsd.npz"))


def test_golden_full_matrix():
    d = oracle.full_exact(G["rand_ref"], G["rand_fit"], G["rand_w"]) / 10.0
    assert np.abs(d - G["rand_rmsd_fit"]).max() < 1e-12
    dn = oracle.full_exact(G["rand_ref"], G["rand_fit"], G["rand_w"], nofit=True) / 10.0
    assert np.abs(dn - G["rand_rmsd_nofit"]).max() < 1e-12
    # the fitted value never exceeds the centred-only value
    assert (d <= dn + 1e-12).all()


def test_golden_knn_rows():
    d, i = oracle.knn_exact(G["rand_ref"], G["rand_fit"], G["rand_w"], 7)
    assert (i == G["rand_knn7_idx"]).all()
    assert np.abs(d - G["rand_knn7_dist"]).max() < 1e-11


def test_golden_special_geometries():
    for a, b, w, want in zip(G["geo_a"], G["geo_b"], G["geo_w"], G["geo_rmsd"]):
        got = oracle.rmsd_nm(a, b, w)
        assert abs(got - want) < 1e-12, (got, want)
    # mirror image: a proper rotation cannot superpose it
    assert G["geo_rmsd"][0] > 0.1


def test_qcp_matches_jacobi():
    rng = np.random.default_rng(3)
    for t in range(300):
        a = rng.normal(size=(rng.integers(3, 40), 3))
        b = a @ np.linalg.qr(rng.normal(size=(3, 3)))[0] + rng.normal(size=a.shape) * rng.choice([1e-6, 1e-3, 0.3, 3])
        S = a.T @ b
        ga, gb = (a * a).sum(), (b * b).sum()
        lam, it = oracle.qcp_lambda(S, ga, gb)
        lj = oracle.jacobi4_lambda(S)
        assert it > 0
        assert abs(lam - lj) <= 1e-12 * max(ga, gb)


def test_rigid_copy_is_zero_and_symmetric():
    rng = np.random.default_rng(4)
    a = synth.trajectory(1, 50, 4)[0]
    w = synth.masses(50, "random", 1)
    R = synth.random_rotations(1, rng)[0]
    b = (a.astype(np.float64) @ R.T + [1.0, -2.0, 3.0]).astype(np.float32)
    assert oracle.rmsd_nm(a, b, w) < 1e-6
    c = synth.trajectory(1, 50, 5, base_seed=4)[0]
    assert abs(oracle.rmsd_nm(a, c, w) - oracle.rmsd_nm(c, a, w)) < 1e-13


def test_two_atom_closed_form():
    # centred dumbbells of lengths L1, L2, equal weights: RMSD = |L1 - L2| / 2
    a = np.array([[0, 0, 0], [1.3, 0, 0]], np.float32)
    b = np.array([[0.5, 0.5, 0.5], [0.5, 0.5 + 0.8, 0.5]], np.float32)
    assert abs(oracle.rmsd_nm(a, b, np.ones(2, np.float32)) - abs(1.3 - 0.8) / 2) < 1e-7


def test_weight_scaling_invariance():
    x = synth.trajectory(2, 33, 6)
    w = synth.masses(33, "random", 2)
    r1 = oracle.rmsd_nm(x[0], x[1], w)
    r2 = oracle.rmsd_nm(x[0], x[1], w * 8.0)  # exact in float32
    assert abs(r1 - r2) < 1e-13


def test_knn_semantics_clamp_and_rank0():
    x = synth.trajectory(12, 15, 8)
    w = synth.masses(15)
    d, i = oracle.knn_exact(x, x, w, 40)        # k clamps to n_ref - 1
    assert d.shape == (12, 11)
    # self query: the dropped rank 0 is the self match, so self never appears
    assert all(r not in i[r] for r in range(12))
    d0, i0 = oracle.knn_exact(x, x, w, 3, keep_rank0=True)
    assert (i0[:, 0] == np.arange(12)).all() and np.abs(d0[:, 0]).max() < 1e-5
    assert (np.diff(d, axis=1) >= 0).all()


def test_port_agrees_with_oracle():
    """The reference-equivalent float path (cumulative in-place rotation of
    the fit copy, float rmsdev) stays within the 1e-5 nm tolerance."""
    ref = synth.trajectory(120, 304, 9)
    fit = synth.trajectory(30, 304, 10, base_seed=9)
    w = synth.masses(304, "backbone")
    k = 10
    d_o, i_o = oracle.knn_exact(ref, fit, w, k)
    d_p, i_p = oracle.port_knn_rows(oracle.port_center(ref, w), oracle.port_center(fit, w), w, k)
    assert np.abs(d_p - d_o).max() < 1e-4          # Angstrom = 1e-5 nm
    assert (i_p == i_o).mean() > 0.99


def test_port_full_rows_layout():
    ref = synth.trajectory(9, 12, 11)
    w = synth.masses(12)
    d, i = oracle.port_knn_rows(oracle.port_center(ref, w), oracle.port_center(ref, w), w, 3, sort=False)
    assert d.shape == (9, 9) and (i == np.arange(9)[None]).all()
    full = oracle.full_exact(ref, ref, w)
    assert np.abs(d - full).max() < 1e-4


# ---------------------------------------------------------------------------
# knn_data distances (mdsctk.cpp:329-356)


@pytest.mark.parametrize("kind", ["clusters", "dihedral", "offset"])
def test_data_distances_closed_forms(kind):
    from scipy.spatial.distance import cdist
    x = synth.data_vectors(60, 37, 3, kind)
    y = synth.data_vectors(25, 37, 4, kind)
    e = np.array([[oracle.data_distance(a, b) for b in x] for a in y])
    assert np.abs(e - cdist(y, x)).max() <= 1e-12 * cdist(y, x).max()
    c = np.array([[oracle.data_distance(a, b, "correlation") for b in x] for a in y])
    r = np.corrcoef(np.vstack([y, x]))[:len(y), len(y):]
    assert np.abs(c - np.sqrt(np.maximum((1.0 - r) / 2.0, 0.0))).max() < 1e-9
    # a vector against itself: 0; against its negation: correlation 1
    assert oracle.data_distance(x[0], x[0]) == 0.0
    assert abs(oracle.data_distance(x[0], -x[0], "correlation") - 1.0) < 1e-12
    # correlation ignores scale and offset
    assert abs(oracle.data_distance(x[0], 3.0 * x[1] + 7.0, "correlation")
               - oracle.data_distance(x[0], x[1], "correlation")) < 1e-12


@pytest.mark.parametrize("metric", ["euclidean", "correlation"])
def test_knn_data_rows(metric):
    from scipy.spatial.distance import cdist
    x = synth.data_vectors(300, 24, 5)
    y = synth.data_vectors(40, 24, 6)
    full = cdist(y, x) if metric == "euclidean" else np.sqrt(np.maximum(
        (1.0 - np.corrcoef(np.vstack([y, x]))[:40, 40:]) / 2.0, 0.0))
    d, i = oracle.knn_data(x, y, 7, metric)
    order = np.lexsort((np.broadcast_to(np.arange(300), full.shape), full), axis=1)
    assert (i == order[:, 1:8]).all()
    assert np.abs(d - np.take_along_axis(full, order[:, 1:8], 1)).max() < 1e-9
    d0, i0 = oracle.knn_data(x, y, 7, metric, keep_rank0=True)
    assert (i0 == order[:, :7]).all()
    # k clamps to n_ref - 1
    d, i = oracle.knn_data(x[:5], y, 30, metric)
    assert d.shape == (40, 4)


def test_three_atom_closed_form():
    # equilateral triangles of sides s1, s2 (equal weights, any pose): every
    # atom sits s/sqrt(3) from the centroid, the optimal rotation aligns the
    # two, so RMSD = |s1 - s2| / sqrt(3)
    rng = np.random.default_rng(7)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]])
    for s1, s2 in [(1.0, 1.0), (1.0, 1.5), (0.7, 2.3)]:
        R1, R2 = synth.random_rotations(2, rng)
        a = (s1 * tri @ R1.T + rng.normal(size=3)).astype(np.float32)
        b = (s2 * tri @ R2.T + rng.normal(size=3)).astype(np.float32)
        want = abs(s1 - s2) / np.sqrt(3.0)
        assert abs(oracle.rmsd_nm(a, b, np.ones(3, np.float32)) - want) < 1e-6


def test_small_isotropic_perturbation():
    # b = a + delta with a tiny isotropic delta: the fitted RMSD never exceeds
    # the centred-only one, and the rotation removes only about 3 of the 3N
    # degrees of freedom of delta, so it stays within a few % of it
    x = synth.trajectory(1, 400, 8)[0].astype(np.float64)
    rng = np.random.default_rng(8)
    delta = 1e-3 * rng.normal(size=x.shape)
    a, b = x.astype(np.float32), (x + delta).astype(np.float32)
    w = np.ones(400, np.float32)
    fit = oracle.rmsd_nm(a, b, w)
    nofit = oracle.rmsd_nm(a, b, w, fit=False)
    assert fit <= nofit + 1e-12
    assert fit >= 0.97 * nofit


def test_cross_covariance_and_qcp_compose_to_the_rmsd():
    """The pieces the device epilogue and re-score compute, one at a time:
    the FP64 cross-covariance S = sum w (a - ca)(b - cb)^T against numpy's
    weighted product of centred frames; lambda_max of Horn's key matrix from S
    (QCP) against numpy's eigen-solve of the same 4x4 matrix; and
    s = G_a + G_b - 2 lambda turned into an RMSD against the oracle's
    rmsd_nm and a numpy SVD-Kabsch fit with a proper rotation."""
    rng = np.random.default_rng(12)
    for natoms in (3, 17, 304):
        w = rng.uniform(1.0, 32.0, natoms).astype(np.float32)
        a = rng.normal(size=(natoms, 3)).astype(np.float32) * 2.0
        b = (a + 0.3 * rng.normal(size=(natoms, 3))).astype(np.float32)
        S, ga, gb = oracle.cross_cov(a, b, w)
        wd = w.astype(np.float64)
        ac = a - (wd[:, None] * a).sum(0) / wd.sum()
        bc = b - (wd[:, None] * b).sum(0) / wd.sum()
        S_np = (wd[:, None] * ac).T @ bc
        assert np.abs(S - S_np).max() <= 1e-9 * np.abs(S_np).max()
        assert abs(ga - (wd[:, None] * ac * ac).sum()) <= 1e-9 * ga
        assert abs(gb - (wd[:, None] * bc * bc).sum()) <= 1e-9 * gb
        xx, xy, xz, yx, yy, yz, zx, zy, zz = S_np.reshape(-1)
        K = np.array([[xx + yy + zz, yz - zy, zx - xz, xy - yx],
                      [yz - zy, xx - yy - zz, xy + yx, zx + xz],
                      [zx - xz, xy + yx, -xx + yy - zz, yz + zy],
                      [xy - yx, zx + xz, yz + zy, -xx - yy + zz]])
        lam, _ = oracle.qcp_lambda(S, ga, gb)
        lam_np = np.linalg.eigvalsh(K)[-1]
        assert abs(lam - lam_np) <= 1e-9 * abs(lam_np)
        rmsd = np.sqrt(max(ga + gb - 2.0 * lam, 0.0) / wd.sum())
        assert abs(rmsd - oracle.rmsd_nm(a, b, w)) <= 1e-9
        # Kabsch: the proper rotation maximising tr(R S) gives the same minimum
        U, sv, Vt = np.linalg.svd(S_np)
        dsign = np.sign(np.linalg.det(U @ Vt))
        tr = sv[0] + sv[1] + dsign * sv[2]
        assert abs(lam - tr) <= 1e-9 * abs(tr)
