"""The kNN graph built from the GPU engine's rows against the CPU pipeline:
FP64 oracle kNN -> make_sysparse restatement -> ARPACK (scipy eigsh, the
library decomp_sparse calls).  The leading eigenvalues of the normalised
Gaussian affinity (sigma = median kNN distance) must agree to 1e-6."""
import numpy as np
import pytest

import oracle
import oracle.graph as og
from mdsctk_amd import graph, knn, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["bf16"])
def test_knn_graph_spectrum_matches_cpu_pipeline(prec):
    n, natoms, k, nev = 8000, 150, 32, 10
    x = synth.trajectory(n, natoms, 31)
    w = synth.masses(natoms)
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(x, w)
        d, i = ctx.query(None, k)
    d_o, i_o = oracle.knn_exact(x, x, w, k)
    m = graph.ssm_build(d, i)
    n_o, pcol_o, irow_o, vals_o = og.ssm_build(d_o, i_o)
    # same edges (up to ties closer than the tolerance), same distances
    same = (m.pcol == pcol_o).all() and (m.irow == irow_o).all()
    if same:
        assert np.abs(m.vals - vals_o).max() < 1e-4
    sigma = float(np.median(d))
    sigma_o = float(np.median(d_o))
    assert abs(sigma - sigma_o) < 1e-4
    ev, _, res, got = graph.spectral_decomp(m, sigma, nev, vectors=False)
    w_o, _ = og.spectral_decomp(n_o, pcol_o, irow_o, vals_o, sigma_o, nev)
    print("GPU-graph eigenvalues", ev, "CPU pipeline", w_o, "same edges", same)
    assert got == nev
    assert np.abs(ev - w_o).max() < 1e-6
    assert res.max() < 1e-8
