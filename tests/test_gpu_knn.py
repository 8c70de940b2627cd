"""GPU parity of the fused knn_rms engine against the FP64 oracle.

All calls go through the C ABI (libmdsctk_knn.so via ctypes)."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.parity import check_knn, TOL_A

pytestmark = pytest.mark.gpu

PRECS = ["fp32", "bf16", "fp16", "exact"]


def _case(n_ref, n_fit, natoms, seed, wkind="uniform"):
    ref = synth.trajectory(n_ref, natoms, seed)
    fit = synth.trajectory(n_fit, natoms, seed + 1000, base_seed=seed)
    w = synth.masses(natoms, wkind, seed)
    return ref, fit, w


def _aligned(n, natoms, seed, dup=0.3):
    """Frames already superposed on one another (no random rigid motion), a
    third of them near-duplicates: the case where the split forms' dropped
    lo*lo' term adds up coherently."""
    rng = np.random.default_rng(seed)
    base = synth.random_walk(natoms, rng)
    x = base[None] + rng.uniform(0.02, 0.3, (n, 1, 1)) * rng.normal(size=(n, natoms, 3))
    for f in np.nonzero(rng.random(n) < dup)[0]:
        if f > 0:
            x[f] = x[rng.integers(0, f)] + 1e-4 * rng.normal(size=(natoms, 3))
    return x.astype(np.float32)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("gen", ["random", "metastable", "aligned"])
@pytest.mark.parametrize("natoms", [16, 57, 304, 1000])
@pytest.mark.parametrize("nofit", [False, True])
def test_approx_scores_within_window(prec, gen, natoms, nofit):
    """The contraction's approximate score s~ stays inside the candidate
    window kappa*(G_q + max G_ref) the engine uses, and the engine's exact
    FP64 score equals the oracle's."""
    if gen == "random":
        ref, fit, w = _case(96, 80, natoms, 11, "backbone")
    elif gen == "metastable":
        allf = synth.metastable_trajectory(176, natoms, natoms, n_centres=4, dup_fraction=0.2)
        ref, fit, w = allf[:96], allf[96:], synth.masses(natoms, "backbone")
    else:
        allf = _aligned(176, natoms, natoms)
        ref, fit, w = allf[:96], allf[96:], synth.masses(natoms, "uniform")
    flags = knn.KNN_NOFIT if nofit else 0
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        a, e = ctx.debug_scores(fit, flags=flags)
        a_self, e_self = ctx.debug_scores(None, flags=flags)   # the self-query operand
        ctx.query(fit, 3, flags)
        kappa = ctx.stats()["eps_rel"]
    W = w.astype(np.float64).sum()
    G = np.array([oracle.center(x, w)[0] for x in ref])
    Gf = np.array([oracle.center(x, w)[0] for x in fit])
    d_gpu = np.sqrt(np.maximum(e / W, 0))
    d_orc = oracle.full_exact(ref, fit, w, nofit=nofit) / 10.0
    assert np.abs(d_gpu - d_orc).max() < 1e-9
    rel = np.abs(a - e) / (Gf[:, None] + G.max())
    rel_self = np.abs(a_self - e_self) / (G[:, None] + G.max())
    print(prec, gen, natoms, nofit, "max rel", rel.max(), "self", rel_self.max(), "kappa", kappa)
    assert rel.max() < kappa and rel_self.max() < kappa


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_cross_query_on_straight_chains(prec):
    """Straight 2000-atom chains: G terms of about 1e8 nm^2 against scores of
    about 30, so approximate sums can exceed (G_a+G_b)/2 and scores go
    negative.  The candidate lists compact at 2(k+1) entries and set row
    thresholds, and the QCP's early exit is then live: it must not drop a
    pair on the way (Newton starts above the largest root).  The window
    covers every pair here, so 900 references keep the rows within the
    merge's 1024 candidates (the fused path, not the exact rows).  Rows
    against the FP64 oracle."""
    allf = synth.extended_chain(1000, 2000, 29)
    w = synth.masses(2000, "uniform")
    ref, fit = allf[:900], allf[900:]
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 5)
        st = ctx.stats()
    print(prec, "candidates per query", st["n_candidates"] / len(fit), "fallback", st["n_fallback_rows"])
    assert st["n_fallback_rows"] < len(fit) // 2
    d_o, i_o = oracle.knn_exact(ref, fit, w, 5)
    check_knn(ref, fit, w, d, i, d_o, i_o)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("gen", ["metastable-2000-wide", "extended-2000", "extended-1000-wide",
                                 "metastable-1000-dups"])
def test_window_at_large_frames(prec, gen):
    """The candidate window at the large-frame shapes: 2000-atom frames
    (the mixed serving workload's largest set) with masses spanning 1-200,
    nearly collinear extended chains (near-degenerate rotations), and the
    1000-atom metastable frames with 5 % near-duplicates at 1e-4 nm.  Every
    approximate score, cross and self, stays inside kappa (G_q + max G_ref);
    and queries at a window 10^4 times too narrow are caught by the run-time
    check (rows recomputed, outputs unchanged)."""
    kind, natoms, extra = (gen.split("-") + [""])[:3]
    natoms = int(natoms)
    if kind == "metastable":
        allf = synth.metastable_trajectory(176, natoms, natoms + 3, n_centres=4,
                                           dup_fraction=0.05 if extra == "dups" else 0.2)
    else:
        allf = synth.extended_chain(176, natoms, natoms)
    w = synth.masses(natoms, "wide", 5) if extra == "wide" else synth.masses(natoms, "uniform")
    ref, fit = allf[:96], allf[96:]
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        a, e = ctx.debug_scores(fit)
        a_self, e_self = ctx.debug_scores(None)
        d0, i0 = ctx.query(fit, 5)
        kappa = ctx.stats()["eps_rel"]
        ctx.set_option(knn.OPT_EPS_SCALE, 1e-4)
        d1, i1 = ctx.query(fit, 5)
        fb = ctx.stats()["n_fallback_rows"]
    G = np.array([oracle.center(x, w)[0] for x in ref])
    Gf = np.array([oracle.center(x, w)[0] for x in fit])
    rel = np.abs(a - e) / (Gf[:, None] + G.max())
    rel_self = np.abs(a_self - e_self) / (G[:, None] + G.max())
    print(prec, gen, "max rel", rel.max(), "self", rel_self.max(), "kappa", kappa,
          "narrow-window fallback rows", fb)
    assert rel.max() < kappa and rel_self.max() < kappa
    assert fb > 0
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(i1, i0)
    d_o, i_o = oracle.knn_exact(ref, fit[:8], w, 5)
    check_knn(ref, fit[:8], w, d0[:8], i0[:8], d_o, i_o)


@pytest.mark.parametrize("prec", PRECS)
def test_knn_small_vs_oracle(prec):
    ref, fit, w = _case(300, 170, 37, 3, "random")
    k = 12
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, k)
        st = ctx.stats()
    d_o, i_o = oracle.knn_exact(ref, fit, w, k)
    check_knn(ref, fit, w, d, i, d_o, i_o)
    assert st["n_fallback_rows"] == (fit.shape[0] if prec == "exact" else 0)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_knn_self_query(prec):
    ref, _, w = _case(257, 1, 61, 5)
    k = 9
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(None, k)
    d_o, i_o = oracle.knn_exact(ref, ref, w, k)
    check_knn(ref, ref, w, d, i, d_o, i_o)


@pytest.mark.parametrize("prec", ["bf16", "exact"])
def test_full_rows_and_nofit(prec):
    ref, fit, w = _case(70, 33, 25, 7, "backbone")
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 0, knn.KNN_FULL_ROWS)
        dn, iN = ctx.query(fit, 5, knn.KNN_NOFIT)
    full = oracle.full_exact(ref, fit, w)
    assert np.abs(d - full).max() < TOL_A
    assert (i == np.arange(ref.shape[0])[None, :]).all()
    d_o, i_o = oracle.knn_exact(ref, fit, w, 5, nofit=True)
    check_knn(ref, fit, w, dn, iN, d_o, i_o, nofit=True)


def test_k_beyond_2048_within_a_budget():
    """Exact rows under an HBM budget go through in batches that fit what it
    leaves, the segmented sort's scratch included: a budget about 20 MB
    below what one batch of all 300 rows holds (48 KB of scores and about
    72 KB of sort keys per row for k = 3000 over 6000 frames) gives the same
    rows, and the context stays within it."""
    ref, fit, w = _case(6000, 300, 24, 19, "backbone")
    k = 3000
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d0, i0 = ctx.query(fit, k)
        full = ctx.stats()["device_bytes"]
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        held = ctx.stats()["device_bytes"]
        budget = full - 20 * 2**20
        assert budget > held + 16 * 2**20, (held, full)  # outputs (10.8 MB) and a few rows fit
        ctx.set_option(knn.OPT_HBM_BUDGET, budget)
        d1, i1 = ctx.query(fit, k)
        assert ctx.stats()["device_bytes"] <= budget
        assert ctx.stats()["n_fallback_rows"] == fit.shape[0]
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(i1, i0)
    d_o, i_o = oracle.knn_exact(ref, fit[:4], w, k)
    check_knn(ref, fit[:4], w, d0[:4], i0[:4], d_o, i_o)


def test_cross_query_lists_within_a_budget():
    """A budget without room for a cross query's candidate lists: the query
    takes fewer reference phases (fewer lists per query) until they fit and,
    when even one phase's lists do not, the exact rows in batches that fit.
    Both give the rows of the unconstrained query, bit for bit."""
    ref, fit, w = _case(20_000, 4000, 40, 23, "backbone")
    k = 7
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d0, i0 = ctx.query(fit, k)
        st0 = ctx.stats()
    p0, cap = st0["ref_splits"], st0["list_capacity"]
    lists = lambda ph: fit.shape[0] * ph * 4 * cap * 8
    assert p0 >= 8 and st0["n_fallback_rows"] == 0, (p0, st0["n_fallback_rows"])
    full = st0["device_bytes"]
    # room for half the phases' lists; then for half of one phase's
    for budget, fewer in ((full - lists(p0) // 2, True), (full - lists(p0) + lists(1) // 2, False)):
        with knn.KnnContext(0, "bf16") as ctx:
            ctx.set_reference(ref, w)
            ctx.set_option(knn.OPT_HBM_BUDGET, budget)
            d1, i1 = ctx.query(fit, k)
            st1 = ctx.stats()
        print("budget", budget, "phases", p0, "->", st1["ref_splits"], "fallback rows",
              st1["n_fallback_rows"], "held", st1["device_bytes"])
        assert st1["device_bytes"] <= budget
        if fewer:
            assert 1 <= st1["ref_splits"] <= p0 // 2
        else:
            assert st1["n_fallback_rows"] == fit.shape[0]
        np.testing.assert_array_equal(d1, d0)
        np.testing.assert_array_equal(i1, i0)
    d_o, i_o = oracle.knn_exact(ref, fit[:6], w, k)
    check_knn(ref, fit[:6], w, d0[:6], i0[:6], d_o, i_o)


@pytest.mark.parametrize("served", [False, True])
def test_nofit_without_room_for_unrotated_operands(served):
    """A resident set held in its principal axes needs a second, unrotated
    copy of its operands for --nofit scores.  A budget with no room for that
    copy sends a --nofit query to the exact FP64 rows (in batches that fit the
    budget) instead of failing: the outputs equal those of the contraction
    with the copy, bit for bit; lifting the budget lets a later --nofit query
    build the copy and take the contraction again.  Both contexts first run a
    fitted query synchronously and served (one slot of up to 512 frames), so
    that the budget is set with every workspace but the copy already held."""
    ref, fit, w = _case(20_000, 300, 200, 17, "backbone")
    k, nf = 7, knn.KNN_NOFIT

    def _warm(ctx):
        ctx.set_option(knn.OPT_SERVE_SLOTS, 1)
        ctx.set_option(knn.OPT_SERVE_MAX_FRAMES, 512)
        ctx.set_reference(ref, w)
        ctx.query(fit, k)
        ctx.wait(ctx.submit(fit, k))
        return ctx.stats()["device_bytes"]

    with knn.KnnContext(0, "bf16") as ctx:
        held_fitted = _warm(ctx)
        d0, i0 = ctx.query(fit, k, nf)
        st0 = ctx.stats()
    copy = st0["device_bytes"] - held_fitted
    assert st0["n_fallback_rows"] == 0
    assert copy > 16 * 2**20, copy
    with knn.KnnContext(0, "bf16") as ctx:
        held = _warm(ctx)
        budget = held + copy // 2
        ctx.set_option(knn.OPT_HBM_BUDGET, budget)
        if served:
            r = ctx.wait(ctx.submit(fit, k, nf))
            d1, i1, fb = r.dist, r.idx, r.info["n_fallback_rows"]
        else:
            d1, i1 = ctx.query(fit, k, nf)
            fb = ctx.stats()["n_fallback_rows"]
        assert ctx.stats()["device_bytes"] <= budget
        assert fb == fit.shape[0]
        ctx.set_option(knn.OPT_HBM_BUDGET, 0)
        d2, i2 = ctx.query(fit, k, nf)
        assert ctx.stats()["n_fallback_rows"] == 0
    for d, i in ((d1, i1), (d2, i2)):
        np.testing.assert_array_equal(d, d0)
        np.testing.assert_array_equal(i, i0)
    rows = [0, 150, 299]
    d_o, i_o = oracle.knn_exact(ref, fit[rows], w, k, nofit=True)
    check_knn(ref, fit[rows], w, d0[rows], i0[rows], d_o, i_o, nofit=True)


def test_keep_rank0_and_k_clamp():
    ref, fit, w = _case(20, 9, 16, 9)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 50)          # k clamps to n_ref-1
        assert d.shape == (9, 19)
        d0, i0 = ctx.query(fit, 4, knn.KNN_KEEP_RANK0)
    d_o, i_o = oracle.knn_exact(ref, fit, w, 50)
    check_knn(ref, fit, w, d, i, d_o, i_o)
    d_o0, i_o0 = oracle.knn_exact(ref, fit, w, 4, keep_rank0=True)
    check_knn(ref, fit, w, d0, i0, d_o0, i_o0)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_c1_workload(prec):
    """2,000 x 2,000 frames x 304 atoms, k=30, self query (the reference's CPU case)."""
    ref = synth.trajectory(2000, 304, 1)
    w = synth.masses(304, "uniform")
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(None, 30)
        st = ctx.stats()
    d_o, i_o = oracle.knn_exact(ref, ref, w, 30)
    err, same = check_knn(ref, ref, w, d, i, d_o, i_o)
    print(prec, "C1 max err A", err, "same-index fraction", same, st)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_c2_workload_full_size(prec):
    """50,000 reference x 50,000 query frames (independent seed streams) x
    304 atoms, k=100: 16 sampled rows against the FP64 oracle, and every row
    sorted, with distinct indices in range."""
    ref = synth.trajectory(50_000, 304, 21)
    fit = synth.trajectory(50_000, 304, 22, base_seed=21)
    w = synth.masses(304, "backbone")
    k = 100
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, k)
        st = ctx.stats()
    assert d.shape == (50_000, k) and st["n_fallback_rows"] == 0
    assert np.all(np.diff(d, axis=1) >= 0)
    srt = np.sort(i, axis=1)
    assert np.all(srt[:, 1:] != srt[:, :-1]) and i.min() >= 0 and i.max() < 50_000
    rows = np.random.default_rng(5).choice(50_000, 16, replace=False)
    d_o, i_o = oracle.knn_exact(ref, fit[rows], w, k)
    check_knn(ref, fit[rows], w, d[rows], i[rows], d_o, i_o)
    print(prec, "C2", st["ms_total"], "ms", st["n_candidates"] / 50_000, "candidates per query")


@pytest.mark.parametrize("prec", ["bf16", "exact"])
def test_k_beyond_2048(prec):
    """k has no upper bound (knn_rms.cpp:131, 224-225, 242): k+1 above the
    LDS select's 2048 takes the compaction + segmented sort path; rows in
    (distance, index) order, equal to the oracle's, for cross and self
    queries, with rank 0 dropped or kept, and k clamped to n_ref - 1."""
    ref, fit, w = _case(3000, 40, 24, 17, "backbone")
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d, i = ctx.query(fit, 2500)
        d_s, i_s = ctx.query(None, 2047)           # k+1 = 2048: the LDS select
        d_t, i_t = ctx.query(None, 2048)           # k+1 = 2049: the sorted path
        d_k, i_k = ctx.query(fit[:5], 5000, knn.KNN_KEEP_RANK0)  # clamped to 3000
    assert d.shape == (40, 2500) and d_k.shape == (5, 3000)
    d_o, i_o = oracle.knn_exact(ref, fit, w, 2500)
    check_knn(ref, fit, w, d, i, d_o, i_o)
    np.testing.assert_array_equal(d_t[:, :2047], d_s)
    np.testing.assert_array_equal(i_t[:, :2047], i_s)
    rows = [0, 1500, 2999]
    d_o, i_o = oracle.knn_exact(ref, ref[rows], w, 2048)
    check_knn(ref, ref[rows], w, d_t[rows], i_t[rows], d_o, i_o)
    d_o, i_o = oracle.knn_exact(ref, fit[:5], w, 3000, keep_rank0=True)
    m = min(d_o.shape[1], d_k.shape[1])
    check_knn(ref, fit[:5], w, d_k[:, :m], i_k[:, :m], d_o[:, :m], i_o[:, :m])
    assert (np.sort(i_k, axis=1) == np.arange(3000)[None, :]).all()
