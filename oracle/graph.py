"""Oracle for the graph stages after knn_rms (test infrastructure only; the
product is libmdsctk_knn's mdsctk_ssm_build / mdsctk_spectral_decomp /
mdsctk_gsm_build / mdsctk_nystrom_decomp).

make_sysparse (make_sysparse.cpp:226-279 with compare_edge / split_edges at
mdsctk.cpp:567-604) is restated with a dictionary of edges; decomp_sparse
(decomp_sparse.cpp:177-264) is restated with scipy.sparse.linalg.eigsh,
which drives the same ARPACK dsaupd/dseupd routines the reference calls,
with the reference's settings (which='LA', ncv = min(10 nev + 1, n),
tol = 0, maxiter = 100 nev).  make_gesparse (make_gesparse.cpp:226-333) is
restated as the reference's sequence of B-tree puts and gets on a
dictionary, and the Nystrom extension of decomp_sparse_nystrom
(decomp_sparse_nystrom.cpp:182-215, 297-304) with scipy sparse products."""
import numpy as np


def ssm_build(dist, idx, k=None):
    """(n, pcol, irow, vals) of the symmetric CSC matrix, strict lower triangle."""
    n, maxk = idx.shape
    k = maxk if k is None else k
    edges = {}
    for i in range(n):
        for x in range(k):
            j = int(idx[i, x])
            if i < j:
                edges[(i, j)] = float(dist[i, x])   # a later put overwrites
            elif j < i:
                edges[(j, i)] = float(dist[i, x])
    keys = sorted(edges)
    pcol = np.zeros(n + 1, np.int32)
    for (c, _r) in keys:
        pcol[c + 1] += 1
    pcol = np.cumsum(pcol).astype(np.int32)
    irow = np.array([r for (_c, r) in keys], np.int32)
    vals = np.array([edges[e] for e in keys], np.float64)
    return n, pcol, irow, vals


def ssm_bytes(n, pcol, irow, vals):
    """The .ssm file content: int32 n, pcol[n+1], irow[nnz], float64 vals[nnz]."""
    return (np.int32(n).tobytes() + pcol.astype(np.int32).tobytes() +
            irow.astype(np.int32).tobytes() + vals.astype(np.float64).tobytes())


def normalized_affinity(n, pcol, irow, vals, sigma):
    """D^-1/2 A D^-1/2 as a scipy CSR matrix (symmetric, no diagonal)."""
    import scipy.sparse as sp
    a = np.exp(-(vals * vals) / (2.0 * sigma * sigma))
    cols = np.repeat(np.arange(n), np.diff(pcol))
    low = sp.csr_matrix((a, (irow, cols)), shape=(n, n))
    A = (low + low.T).tocsr()
    d = 1.0 / np.sqrt(np.asarray(A.sum(axis=1)).ravel())
    return (sp.diags(d) @ A @ sp.diags(d)).tocsr()


def spectral_decomp(n, pcol, irow, vals, sigma, nev):
    """Leading eigenvalues (descending) and eigenvectors[nev, n] via ARPACK."""
    from scipy.sparse.linalg import eigsh
    M = normalized_affinity(n, pcol, irow, vals, sigma)
    ncv = min(10 * nev + 1, n)
    v0 = np.full(n, 1.0 / np.sqrt(n))
    w, z = eigsh(M, k=nev, which="LA", ncv=ncv, tol=0, maxiter=100 * nev, v0=v0)
    order = np.argsort(w)[::-1]
    return w[order], z[:, order].T


def gsm_build(dist, idx, k=None, symmetric=False):
    """(m, pcol, irow, vals) of the general CSC matrix: the reference's loop of
    db.put (overwrite) and, with -s, db.get + db.put of the mirrored edge when
    absent; then columns 0..m-1 of the (from, to)-sorted edges."""
    m, maxk = idx.shape
    k = maxk if k is None else k
    edges = {}
    for i in range(m):
        for x in range(k):
            j = int(idx[i, x])
            edges[(i, j)] = float(dist[i, x])
            if symmetric and (j, i) not in edges:
                edges[(j, i)] = float(dist[i, x])
    keys = sorted(e for e in edges if e[0] < m)
    pcol = np.zeros(m + 1, np.int64)
    for (c, _r) in keys:
        pcol[c + 1] += 1
    pcol = np.cumsum(pcol).astype(np.int32)
    irow = np.array([r for (_c, r) in keys], np.int32)
    vals = np.array([edges[e] for e in keys], np.float64)
    return m, pcol, irow, vals


def nystrom_extend(n, pcol_a, irow_a, vals_a, m, pcol_b, irow_b, vals_b, sigma, evals, evecs):
    """ext[nev, m]: the out-of-sample affinities exp(-d^2/2 sigma^2), scaled by
    the in-sample d_i^-1/2 of their rows and their own column sums^-1/2,
    applied to z / l."""
    import scipy.sparse as sp
    a = np.exp(-(vals_a * vals_a) / (2.0 * sigma * sigma))
    cols = np.repeat(np.arange(n), np.diff(pcol_a))
    low = sp.csr_matrix((a, (irow_a, cols)), shape=(n, n))
    deg = np.asarray((low + low.T).sum(axis=1)).ravel()
    b = np.exp(-(vals_b * vals_b) / (2.0 * sigma * sigma))
    bcols = np.repeat(np.arange(m), np.diff(pcol_b))
    B = sp.csc_matrix((b, (irow_b, bcols)), shape=(n, m))
    colsum = np.asarray(B.sum(axis=0)).ravel()
    Bn = sp.diags(1.0 / np.sqrt(deg)) @ B @ sp.diags(1.0 / np.sqrt(colsum))
    return np.asarray((Bn.T @ (evecs / evals[:, None]).T).T)
