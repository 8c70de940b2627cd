"""Graph stages after knn_rms on the host: the .ssm kNN graph (make_sysparse
semantics) and its leading spectral decomposition (decomp_sparse); the .gsm
matrix of out-of-sample rows (make_gesparse) and the Nystrom extension
(decomp_sparse_nystrom)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import oracle.graph as og
from mdsctk_amd import build, graph, knn, synth
from tests.test_formats import TRP_OOS, TRP_PDB, TRP_XTC, topology, xtc_read

HERE = os.path.dirname(os.path.abspath(__file__))


def random_knn(n, maxk, seed, dup_scale=0.0):
    """kNN-like rows: distinct random neighbours != row, ascending random
    distances; the same pair may carry different distances in its two rows."""
    rng = np.random.default_rng(seed)
    idx = np.zeros((n, maxk), np.int32)
    for i in range(n):
        cand = rng.choice(n - 1, size=maxk, replace=False)
        idx[i] = np.where(cand >= i, cand + 1, cand)
    dist = np.sort(rng.uniform(0.5, 5.0, (n, maxk)), axis=1)
    return dist, idx


@pytest.mark.parametrize("k", [None, 5])
def test_ssm_build_matches_restatement(k):
    dist, idx = random_knn(300, 12, 1)
    m = graph.ssm_build(dist, idx, k)
    n, pcol, irow, vals = og.ssm_build(dist, idx, k)
    assert m.n == n
    assert (m.pcol == pcol).all() and (m.irow == irow).all() and (m.vals == vals).all()
    # strict lower triangle, rows ascending within a column
    cols = np.repeat(np.arange(n), np.diff(m.pcol))
    assert (m.irow > cols).all()
    for c in range(n):
        assert (np.diff(m.irow[m.pcol[c]:m.pcol[c + 1]]) > 0).all()


def test_ssm_later_row_overwrites_and_empty_columns():
    # row 0 -> 2 (1.0), row 2 -> 0 (7.0): the later row wins; column 1 and
    # the last column are empty
    idx = np.array([[2], [2], [0], [1]], np.int32)
    dist = np.array([[1.0], [2.0], [7.0], [3.0]])
    m = graph.ssm_build(dist, idx)
    assert list(m.pcol) == [0, 1, 3, 3, 3]
    assert list(m.irow) == [2, 2, 3] and list(m.vals) == [7.0, 2.0, 3.0]
    with pytest.raises(knn.KnnError):
        graph.ssm_build(dist, np.array([[2], [2], [9], [1]], np.int32))
    with pytest.raises(knn.KnnError):
        graph.ssm_build(dist, idx, k=2)


def test_ssm_file_layout_and_round_trip(tmp_path):
    dist, idx = random_knn(120, 8, 2)
    m = graph.ssm_build(dist, idx)
    f = tmp_path / "g.ssm"
    graph.ssm_write(f, m)
    assert f.read_bytes() == og.ssm_bytes(*og.ssm_build(dist, idx))
    r = graph.ssm_read(f)
    assert r.n == m.n and (r.pcol == m.pcol).all() and (r.irow == m.irow).all() and (r.vals == m.vals).all()
    bad = tmp_path / "bad.ssm"
    bad.write_bytes(f.read_bytes()[:-5])
    with pytest.raises(knn.KnnError):
        graph.ssm_read(bad)


def trp_cage_graph(k=20):
    _, x, *_ = xtc_read(TRP_XTC)
    _, w = topology(TRP_PDB)
    d, i = oracle.knn_exact(x, x, w, k)
    return d, i


def test_spectral_matches_arpack_on_trp_cage():
    d, i = trp_cage_graph()
    m = graph.ssm_build(d, i)
    sigma = float(np.median(d))
    ev, vec, res, got = graph.spectral_decomp(m, sigma, 10)
    w, z = og.spectral_decomp(m.n, m.pcol, m.irow, m.vals, sigma, 10)
    assert got == 10
    assert np.abs(ev - w).max() < 1e-10, (ev, w)
    assert abs(ev[0] - 1.0) < 1e-12          # the normalised affinity's top eigenvalue
    assert res.max() < 1e-9
    for a, b in zip(vec, z):                 # distinct eigenvalues: vectors up to sign
        assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 1e-6


def test_spectral_matches_arpack_threaded():
    """A 12,000-node kNN graph: large enough that the solver splits its
    vector work over host threads; eigenpairs as scipy's ARPACK gives."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(4)
    x = rng.normal(size=(12000, 6))
    dd, ii = cKDTree(x).query(x, 17)
    d, i = dd[:, 1:].copy(), ii[:, 1:].astype(np.int32)
    m = graph.ssm_build(d, i)
    sigma = float(np.median(d))
    ev, vec, res, got = graph.spectral_decomp(m, sigma, 8)
    w, z = og.spectral_decomp(m.n, m.pcol, m.irow, m.vals, sigma, 8)
    assert got == 8
    assert np.abs(ev - w).max() < 1e-10, (ev, w)
    assert res.max() < 1e-9
    # the same answer again, bit for bit (fixed row blocks, fixed sum order)
    ev2, vec2, _, _ = graph.spectral_decomp(m, sigma, 8)
    assert (ev2 == ev).all() and (vec2 == vec).all()


def test_spectral_closed_forms():
    # two disconnected rings of 30 and 41 nodes with equal weights: the
    # normalised affinity of a ring of n nodes has eigenvalues cos(2 pi j / n),
    # so the leading ones are 1 (once per ring), then the doubly degenerate
    # cos(2 pi / 41) and cos(2 pi / 30)
    sizes = (30, 41)
    n = sum(sizes)
    idx = np.zeros((n, 2), np.int32)
    base = 0
    for size in sizes:
        for j in range(size):
            idx[base + j] = [base + (j + 1) % size, base + (j - 1) % size]
        base += size
    m = graph.ssm_build(np.ones((n, 2)), idx)
    ev, _, res, got = graph.spectral_decomp(m, 1.0, 6, vectors=False)
    assert got == 6
    want = np.sort(np.concatenate([np.cos(2 * np.pi * np.arange(s) / s) for s in sizes]))[::-1][:6]
    assert np.abs(ev - want).max() < 1e-10, (ev, want)
    assert res.max() < 1e-9
    # a small random graph against a dense eigen-solve
    dist, idx = random_knn(150, 6, 3)
    m = graph.ssm_build(dist, idx)
    M = og.normalized_affinity(m.n, m.pcol, m.irow, m.vals, 2.0).toarray()
    dense = np.sort(np.linalg.eigvalsh(M))[::-1][:8]
    ev, _, _, got = graph.spectral_decomp(m, 2.0, 8, vectors=False)
    assert got == 8 and np.abs(ev - dense).max() < 1e-10


def run(tool, args, cwd):
    return subprocess.run([tool] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=120)


def test_make_sysparse_and_decomp_sparse_cli(tmp_path):
    d, i = trp_cage_graph(k=16)
    d.tofile(tmp_path / "distances.dat")
    i.tofile(tmp_path / "indices.dat")
    p = run(build.MAKE_SYSPARSE_PATH, ["-k", "16", "-n", "12"], tmp_path)
    assert p.returncode == 0, p.stdout
    want = og.ssm_bytes(*og.ssm_build(d, i, 12))
    assert (tmp_path / "distances.ssm").read_bytes() == want
    assert run(build.MAKE_SYSPARSE_PATH, [], tmp_path).returncode == 255
    assert run(build.MAKE_SYSPARSE_PATH, ["-k", "4", "-n", "8"], tmp_path).returncode == 255
    assert run(build.MAKE_SYSPARSE_PATH, ["--help"], tmp_path).returncode == 1
    sigma = float(np.median(d[:, :12]))
    p = run(build.DECOMP_SPARSE_PATH, ["-q", repr(sigma), "-n", "5"], tmp_path)
    assert p.returncode == 0, p.stdout
    ev = np.loadtxt(tmp_path / "eigenvalues.dat")
    m = graph.ssm_read(tmp_path / "distances.ssm")
    w, _ = og.spectral_decomp(m.n, m.pcol, m.irow, m.vals, sigma, 5)
    assert np.abs(ev - w).max() < 1e-10
    vec = np.loadtxt(tmp_path / "eigenvectors.dat")
    assert vec.shape == (5, 1000)
    assert run(build.DECOMP_SPARSE_PATH, ["-n", "5"], tmp_path).returncode == 255


# ---------------------------------------------------------------------------
# make_gesparse / decomp_sparse_nystrom


def oos_knn(m, n_ref, maxk, seed, self_hits=0):
    """Out-of-sample-like rows: m rows of distinct neighbours in [0, n_ref),
    ascending random distances; the first self_hits rows list their own
    index (a query that is also a reference frame)."""
    rng = np.random.default_rng(seed)
    idx = np.zeros((m, maxk), np.int32)
    for i in range(m):
        idx[i] = rng.choice(n_ref, size=maxk, replace=False)
        if i < self_hits and i not in idx[i]:
            idx[i, 0] = i
    dist = np.sort(rng.uniform(0.5, 5.0, (m, maxk)), axis=1)
    return dist, idx


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("k", [None, 4])
def test_gsm_build_matches_restatement(symmetric, k):
    # references beyond the query rows (j >= m) and self entries both occur
    dist, idx = oos_knn(200, 500, 9, 5, self_hits=20)
    g = graph.gsm_build(dist, idx, k, symmetric=symmetric)
    m, pcol, irow, vals = og.gsm_build(dist, idx, k, symmetric=symmetric)
    assert g.n == m
    assert (g.pcol == pcol).all() and (g.irow == irow).all() and (g.vals == vals).all()
    for c in range(m):
        assert (np.diff(g.irow[g.pcol[c]:g.pcol[c + 1]]) > 0).all()
    if not symmetric:
        kk = idx.shape[1] if k is None else k
        assert (np.diff(g.pcol) == kk).all()


def test_gsm_entry_rules():
    # row 0 lists 1 (2.0) and itself; row 1 lists 0 (5.0) and 3 (beyond the
    # last row); row 2 lists 1 (4.0)
    idx = np.array([[1, 0], [0, 3], [1, 2]], np.int32)
    dist = np.array([[2.0, 9.0], [5.0, 6.0], [4.0, 1.0]])
    g = graph.gsm_build(dist, idx)
    assert list(g.pcol) == [0, 2, 4, 6]
    assert list(g.irow) == [0, 1, 0, 3, 1, 2] and list(g.vals) == [9.0, 2.0, 5.0, 6.0, 4.0, 1.0]
    # -s: (1, 0) keeps row 1's own 5.0 (the mirror of row 0 came first but
    # the row's own put replaces it); (1, 2) is the mirror of row 2 (4.0);
    # (0, 1) is row 0's own 2.0; the mirror (3, 1) has no column
    g = graph.gsm_build(dist, idx, symmetric=True)
    assert list(g.pcol) == [0, 2, 5, 7]
    assert list(g.irow) == [0, 1, 0, 2, 3, 1, 2]
    assert list(g.vals) == [9.0, 2.0, 5.0, 4.0, 6.0, 4.0, 1.0]
    with pytest.raises(knn.KnnError):
        graph.gsm_build(dist, np.array([[1, 0], [-1, 3], [1, 2]], np.int32))
    with pytest.raises(knn.KnnError):
        graph.gsm_build(dist, idx, k=3)


def symmetric_knn(n, k, seed):
    """kNN rows (self excluded) of a random symmetric distance matrix, so
    that both rows of a pair carry the same value."""
    rng = np.random.default_rng(seed)
    pts = rng.normal(size=(n, 3))
    D = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    D = np.triu(D, 1) + np.triu(D, 1).T
    np.fill_diagonal(D, np.inf)
    idx = np.argsort(D, axis=1, kind="stable")[:, :k].astype(np.int32)
    return np.take_along_axis(D, idx, 1), idx


def test_nystrom_reproduces_in_sample_vectors():
    # With the -s general matrix of the same rows, column i of B is the
    # symmetric row i of A and its column sum is d_i, so the extension
    # D^-1/2 A D^-1/2 z / l of an eigenvector is the eigenvector itself.
    dist, idx = symmetric_knn(400, 8, 7)
    a = graph.ssm_build(dist, idx)
    b = graph.gsm_build(dist, idx, symmetric=True)
    sigma = float(np.median(dist))
    ev, vec, ext, res, got = graph.nystrom_decomp(a, b, sigma, 6)
    assert got == 6 and res.max() < 1e-9
    assert np.abs(ext - vec).max() < 1e-10
    ev2, vec2, _, _ = graph.spectral_decomp(a, sigma, 6)
    assert (ev == ev2).all() and (vec == vec2).all()


def test_nystrom_matches_restatement_on_trp_cage():
    _, x, *_ = xtc_read(TRP_XTC)
    _, y, *_ = xtc_read(TRP_OOS)
    _, w = topology(TRP_PDB)
    y = y[:300]
    d, i = oracle.knn_exact(x, x, w, 16)
    do, io = oracle.knn_exact(x, y, w, 16, keep_rank0=True)
    a = graph.ssm_build(d, i)
    b = graph.gsm_build(do, io)
    sigma = float(np.median(d))
    ev, vec, ext, res, got = graph.nystrom_decomp(a, b, sigma, 5)
    assert got == 5 and res.max() < 1e-9
    w_o, z_o = og.spectral_decomp(a.n, a.pcol, a.irow, a.vals, sigma, 5)
    assert np.abs(ev - w_o).max() < 1e-10
    sign = np.sign((vec * z_o).sum(axis=1))
    want = og.nystrom_extend(a.n, a.pcol, a.irow, a.vals, b.n, b.pcol, b.irow, b.vals, sigma, w_o,
                             z_o * sign[:, None])
    assert np.abs(ext - want).max() < 1e-6 * np.abs(want).max()
    # the extension with our own vectors, evaluated by the restatement
    mine = og.nystrom_extend(a.n, a.pcol, a.irow, a.vals, b.n, b.pcol, b.irow, b.vals, sigma, ev, vec)
    assert np.abs(ext - mine).max() < 1e-12 * np.abs(mine).max()
    with pytest.raises(knn.KnnError):   # a general matrix indexing beyond the graph
        bad = graph.SSM(b.n, b.pcol, b.irow + a.n, b.vals)
        graph.nystrom_decomp(a, bad, sigma, 5)


def test_make_gesparse_and_nystrom_cli(tmp_path):
    _, x, *_ = xtc_read(TRP_XTC)
    _, y, *_ = xtc_read(TRP_OOS)
    _, w = topology(TRP_PDB)
    d, i = oracle.knn_exact(x, x, w, 12)
    d.tofile(tmp_path / "distances.dat")
    i.tofile(tmp_path / "indices.dat")
    assert run(build.MAKE_SYSPARSE_PATH, ["-k", "12"], tmp_path).returncode == 0
    do, io = oracle.knn_exact(x, y[:200], w, 12, keep_rank0=True)
    do.tofile(tmp_path / "oos_d.dat")
    io.tofile(tmp_path / "oos_i.dat")
    p = run(build.MAKE_GESPARSE_PATH, ["-k", "12", "-n", "10", "-d", "oos_d.dat", "-i", "oos_i.dat"],
            tmp_path)
    assert p.returncode == 0, p.stdout
    assert (tmp_path / "distances.gsm").read_bytes() == og.ssm_bytes(*og.gsm_build(do, io, 10))
    p = run(build.MAKE_GESPARSE_PATH, ["-k", "12", "-s", "-o", "sym.gsm", "-d", "oos_d.dat",
                                       "-i", "oos_i.dat"], tmp_path)
    assert p.returncode == 0, p.stdout
    assert (tmp_path / "sym.gsm").read_bytes() == og.ssm_bytes(*og.gsm_build(do, io, symmetric=True))
    assert run(build.MAKE_GESPARSE_PATH, [], tmp_path).returncode == 255
    assert run(build.MAKE_GESPARSE_PATH, ["-k", "4", "-n", "8"], tmp_path).returncode == 255
    assert run(build.MAKE_GESPARSE_PATH, ["-h"], tmp_path).returncode == 1
    sigma = float(np.median(d))
    p = run(build.DECOMP_SPARSE_NYSTROM_PATH, ["-q", repr(sigma), "-n", "4"], tmp_path)
    assert p.returncode == 0, p.stdout
    ev = np.loadtxt(tmp_path / "eigenvalues.dat")
    vec = np.loadtxt(tmp_path / "eigenvectors.dat")
    assert vec.shape == (4, 1000 + 200)
    a = graph.ssm_read(tmp_path / "distances.ssm")
    b = graph.gsm_read(tmp_path / "distances.gsm")
    ev2, vec2, ext2, _, _ = graph.nystrom_decomp(a, b, sigma, 4)
    assert np.abs(ev - ev2).max() < 1e-15
    assert np.abs(vec - np.hstack([vec2, ext2])).max() < 1e-15
    assert run(build.DECOMP_SPARSE_NYSTROM_PATH, ["-q", "1.0"], tmp_path).returncode == 255
    assert run(build.DECOMP_SPARSE_NYSTROM_PATH, ["--help"], tmp_path).returncode == 1
