"""Graph stages after knn_rms on the host: the .ssm kNN graph (make_sysparse
semantics) and its leading spectral decomposition (decomp_sparse)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import oracle.graph as og
from mdsctk_amd import build, graph, knn, synth
from tests.test_formats import TRP_PDB, TRP_XTC, topology, xtc_read

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
