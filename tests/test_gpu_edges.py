"""GPU edge cases of the knn_rms engine against the FP64 oracle: atom counts
at both ends (1-3 atoms, more atoms than the merge stages in LDS), the
largest fused list capacity and the exact path beyond it, exact duplicate
frames (ties at zero), tiny and empty sets.  All through the C ABI."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.parity import check_knn

pytestmark = pytest.mark.gpu


def _case(n_ref, n_fit, natoms, seed):
    ref = synth.trajectory(n_ref, natoms, seed)
    fit = synth.trajectory(n_fit, natoms, seed + 1000, base_seed=seed)
    w = synth.masses(natoms, "backbone", seed)
    return ref, fit, w


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_many_atoms_unstaged_merge(prec):
    # 2100 atoms: the query frame no longer fits the merge's LDS stage
    ref, fit, w = _case(160, 40, 2100, 31)
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 7)
        ds, is_ = ctx.query(None, 7)
        st = ctx.stats()
    check_knn(ref, fit, w, d, i, *oracle.knn_exact(ref, fit, w, 7))
    check_knn(ref, ref, w, ds, is_, *oracle.knn_exact(ref, ref, w, 7))
    assert st["n_fallback_rows"] == 0


@pytest.mark.parametrize("natoms", [1, 2, 3])
@pytest.mark.parametrize("nofit", [False, True])
def test_few_atoms(natoms, nofit):
    # one atom: every centred frame is the origin, all distances 0 (ties
    # by index); two atoms: collinear; three: planar
    ref, fit, w = _case(90, 30, natoms, 32)
    flags = knn.KNN_NOFIT if nofit else 0
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 6, flags)
    check_knn(ref, fit, w, d, i, *oracle.knn_exact(ref, fit, w, 6, nofit=nofit), nofit=nofit)
    if natoms == 1:
        assert (d == 0).all() and (i == np.arange(1, 7)[None, :]).all()


@pytest.mark.parametrize("k", [300, 500, 900])
def test_large_k(k):
    # k+1 = 301 and 501: lists of 512 and 1024 entries; 901 exceeds the
    # largest list and takes the exact path for every row
    ref, fit, w = _case(1200, 24, 20, 33)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, k)
        st = ctx.stats()
    check_knn(ref, fit, w, d, i, *oracle.knn_exact(ref, fit, w, k))
    fused = k + 1 <= 896
    assert st["n_fallback_rows"] == (0 if fused else fit.shape[0])
    if fused:
        assert st["list_capacity"] == (512 if k + 1 <= 384 else 1024)


def test_exact_duplicates_tie_by_index():
    ref, _, w = _case(64, 1, 30, 34)
    ref = np.concatenate([ref, ref[:16], ref[:16]])       # frames 64.. copy 0..15 twice
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(None, 3, knn.KNN_KEEP_RANK0)
    for f in range(16):
        assert (d[f, :3] < 1e-5).all()
        assert list(i[f]) == [f, 64 + f, 80 + f]          # equal distances: index order
    check_knn(ref, ref, w, d, i, *oracle.knn_exact(ref, ref, w, 3, keep_rank0=True))


def test_tiny_and_empty_sets():
    ref, fit, w = _case(2, 5, 12, 35)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref[:1], w)
        d, i = ctx.query(fit, 4)                          # k clamps to n_ref - 1 = 0
        assert d.shape == (5, 0)
        d, i = ctx.query(fit, 4, knn.KNN_KEEP_RANK0)       # the one reference
        assert d.shape == (5, 1) and (i == 0).all()
        exact = 10.0 * oracle.pairs_nm(ref[:1], fit, w, np.zeros(5, np.int64), np.arange(5))
        assert np.abs(d[:, 0] - exact).max() < 1e-4
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 4)
        assert d.shape == (5, 1)
        check_knn(ref, fit, w, d, i, *oracle.knn_exact(ref, fit, w, 4))
        d, i = ctx.query(fit[:0], 4)                      # no query frames
        assert d.shape == (0, 1)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_lists_full_of_equal_zero_keys(prec):
    """One atom per frame: every pair scores exactly 0, so each candidate
    list fills with more zero keys than it can hold.  The compaction keeps
    as many as fit and records the overflow; the merge then sends every
    row to the exact path, whose (distance, index) order gives indices
    1..k."""
    ref = synth.trajectory(20000, 1, 36)
    fit = synth.trajectory(64, 1, 37, base_seed=36)
    w = np.ones(1, np.float32)
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 5)
        st = ctx.stats()
    assert (d == 0).all() and (i == np.arange(1, 6)[None, :]).all()
    assert st["n_fallback_rows"] == fit.shape[0]
