"""GPU parity of the symmetric self query: a landmark pass bounds every
row's threshold, the contraction scores the upper block triangle once into
each fit frame's candidate lists and each reference frame's mirrored list
(no score matrix), and the merge re-scores both.  From 512 frames on (fitted
RMSD) the contraction runs over a copy of the set in landmark order, each
frame in its own principal axes, and skips the QCP of pairs that a bound on
lambda_max rejects.  Its outputs must equal the candidate-list path's bit
for bit (both re-score a certified window in FP64), and the FP64 oracle's
within the distance tolerance.

All calls go through the C ABI (libmdsctk_knn.so via ctypes)."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.parity import check_knn

pytestmark = pytest.mark.gpu


def _self_query(prec, frames, w, k, flags=0, sym=True, budget=0, mirror_cap=0):
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_option(knn.OPT_SELF_SYMMETRIC, 1 if sym else 0)
        if budget:
            ctx.set_option(knn.OPT_HBM_BUDGET, budget)
        if mirror_cap:
            ctx.set_option(knn.OPT_MIRROR_CAP, mirror_cap)
        ctx.set_reference(frames, w)
        d, i = ctx.query(None, k, flags)
        return d, i, ctx.stats()


def _triangle_pairs(n, block):
    """Pairs of the upper block triangle: each fit block of `block` frames
    against every frame from its own first frame on."""
    return sum(min(block, n - f0) * (n - f0) for f0 in range(0, n, block))


def _both(prec, frames, w, k, flags=0):
    d_s, i_s, st_s = _self_query(prec, frames, w, k, flags, sym=True)
    d_l, i_l, st_l = _self_query(prec, frames, w, k, flags, sym=False)
    n = len(frames)
    print(prec, n, k, flags, "sym pairs", st_s["n_pairs_computed"], "of", st_s["n_pairs"],
          "landmark", st_s["n_pairs_landmark"], "mirrored", st_s["n_mirrored"], "max",
          st_s["n_mirrored_max"], "relisted", st_s["n_relisted_rows"],
          "cand", st_s["n_candidates"], "vs", st_l["n_candidates"],
          "fallback", st_s["n_fallback_rows"], st_l["n_fallback_rows"])
    assert st_l["n_pairs_computed"] == n * n
    assert st_s["n_pairs_computed"] == _triangle_pairs(n, 64 if prec == "fp32" else 128)
    np.testing.assert_array_equal(i_s, i_l)
    np.testing.assert_array_equal(d_s, d_l)
    return d_s, i_s, st_s


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
@pytest.mark.parametrize("n", [2000, 777, 512, 511, 130, 67])
def test_symmetric_equals_lists(prec, n):
    """Ragged sizes: n not a multiple of 4, 32, 64 or the fit block; 512 and
    511 frames on either side of the landmark-ordered copy in principal
    axes."""
    frames = synth.metastable_trajectory(n, 304, 7 + n, n_centres=6, dup_fraction=0.1)
    w = synth.masses(304, "backbone")
    _both(prec, frames, w, min(30, n - 1))


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("flags", [knn.KNN_NOFIT, knn.KNN_KEEP_RANK0,
                                   knn.KNN_NOFIT | knn.KNN_KEEP_RANK0])
@pytest.mark.parametrize("n", [450, 700])
def test_symmetric_flags(prec, flags, n):
    """450 frames: the set in its own order; 700: the landmark-ordered copy
    in principal axes (fitted scores only)."""
    frames = synth.trajectory(n, 57, 5)
    w = synth.masses(57, "random", 5)
    _both(prec, frames, w, 9, flags)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_symmetric_vs_oracle(prec):
    frames = synth.metastable_trajectory(600, 120, 3, n_centres=5, dup_fraction=0.2)
    w = synth.masses(120, "uniform")
    k = 16
    d, i, st = _both(prec, frames, w, k)
    d_o, i_o = oracle.knn_exact(frames, frames, w, k)
    check_knn(frames, frames, w, d, i, d_o, i_o)


def test_symmetric_duplicates_overflow_to_exact_rows():
    """More than the list capacity of equal (zero) keys in every row: the
    candidate lists overflow, the merge sends the rows through the lists
    again and then to the exact path, and the outputs still equal the list
    path's (ties by index)."""
    rng = np.random.default_rng(1)
    base = rng.normal(size=(1, 40, 3)).astype(np.float32)
    frames = np.repeat(base, 2200, axis=0)
    frames[2100:] += 0.05 * rng.normal(size=(100, 40, 3)).astype(np.float32)
    w = synth.masses(40, "uniform")
    d, i, st = _both("bf16", frames, w, 5)
    assert st["n_fallback_rows"] > 0
    assert np.abs(d[:2100]).max() < 1e-5


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_symmetric_many_duplicates(prec):
    """5000 equal frames: every one of them lies below every equal frame's
    threshold, so their mirrored lists fill with thousands of zero keys;
    outputs still equal the list path's."""
    rng = np.random.default_rng(2)
    base = rng.normal(size=(1, 40, 3)).astype(np.float32)
    frames = np.repeat(base, 5100, axis=0)
    frames[5000:] += 0.05 * rng.normal(size=(100, 40, 3)).astype(np.float32)
    w = synth.masses(40, "uniform")
    d, i, st = _both(prec, frames, w, 5)
    assert np.abs(d[:5000]).max() < 1e-5


def test_symmetric_large_set_with_duplicates():
    """34000 distinct frames (the landmark bound keeps their mirrored lists
    short) plus 6000 equal ones (their zero keys overflow the candidate
    lists, so those rows end on the exact path)."""
    frames = synth.trajectory(40000, 16, 21)
    frames[34000:] = frames[34000]
    w = synth.masses(16, "uniform")
    d, i, st = _both("bf16", frames, w, 7)
    assert st["n_fallback_rows"] >= 6000
    assert np.abs(d[34000:]).max() < 1e-5
    assert st["n_mirrored_max"] <= 8192


@pytest.mark.parametrize("natoms", [1, 2, 3])
def test_symmetric_degenerate_frames(natoms):
    """Frames of 1, 2 and 3 atoms (a point, a line, a plane: second-moment
    matrices of rank 0, 1 and 2, whose principal axes are not unique) in a
    set large enough for the landmark-ordered copy in principal axes."""
    frames = synth.trajectory(700, natoms, 41)
    w = synth.masses(natoms, "random", 3)
    d, i, _ = _both("bf16", frames, w, 12)
    d_o, i_o = oracle.knn_exact(frames, frames[:16], w, 12)
    check_knn(frames, frames[:16], w, d[:16], i[:16], d_o, i_o)


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
@pytest.mark.parametrize("gen", ["extended-2000", "extended-2000-wide"])
@pytest.mark.parametrize("n", [640, 500])
def test_symmetric_hard_frames(prec, gen, n):
    """Nearly collinear 2000-atom chains, uniform or 1-200 masses: G terms
    of about 1e8 nm^2 against scores of about 30, so the approximate scores
    scatter hundreds below zero (the window covers every pair).  A tile whose
    scores are not mirrored must not append any of them to a mirrored list
    (a negative score once passed a sentinel threshold there, and diagonal
    pairs were then counted twice).  640 frames take the landmark-ordered
    copy in principal axes, 500 the set in its own order.  Outputs must
    equal the list path's and the oracle's."""
    natoms = 2000
    frames = synth.extended_chain(n, natoms, 23)
    w = synth.masses(natoms, "wide", 5) if gen.endswith("wide") else synth.masses(natoms, "uniform")
    d, i, st = _both(prec, frames, w, 10)
    # mirrored appends come only from pairs whose reference frame lies past
    # the fit block
    blk = 64 if prec == "fp32" else 128
    mirrorable = sum(min(blk, n - f0) * max(0, n - f0 - blk) for f0 in range(0, n, blk))
    assert st["n_mirrored"] <= mirrorable
    d_o, i_o = oracle.knn_exact(frames, frames[:12], w, 10)
    check_knn(frames, frames[:12], w, d[:12], i[:12], d_o, i_o)


def test_symmetric_large_k():
    frames = synth.trajectory(1500, 64, 9)
    w = synth.masses(64, "uniform")
    _both("bf16", frames, w, 800)


def test_symmetric_budget_falls_back_to_lists():
    """A budget that holds the list path's candidate lists (about 1.3 GB
    here) but not the symmetric form's lists and mirrored lists (about
    2.0 GB) keeps the list path, with the same outputs."""
    n = 20000
    frames = synth.trajectory(n, 16, 4)
    w = synth.masses(16, "uniform")
    d_s, i_s, st_s = _self_query("bf16", frames, w, 10)
    d_b, i_b, st_b = _self_query("bf16", frames, w, 10, budget=1.8e9)
    assert st_s["n_pairs_computed"] == _triangle_pairs(n, 128)
    assert st_b["n_pairs_computed"] == n * n
    np.testing.assert_array_equal(i_s, i_b)
    np.testing.assert_array_equal(d_s, d_b)


@pytest.mark.parametrize("metric", ["euclidean", "correlation"])
def test_symmetric_data_vectors(metric):
    x = synth.data_vectors(900, 23, 12, kind="clusters")
    out = []
    for sym in (1, 0):
        with knn.KnnContext(0, "bf16") as ctx:
            ctx.set_option(knn.OPT_SELF_SYMMETRIC, sym)
            ctx.set_data_reference(x, metric)
            d, i = ctx.data_query(None, 11)
            out.append((d, i, ctx.stats()))
    assert out[0][2]["n_pairs_computed"] == _triangle_pairs(900, 128)
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][0], out[1][0])
    d_o, i_o = oracle.knn_data(x, x, 11, metric)
    assert np.abs(out[0][0] - d_o).max() < 1e-9


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_symmetric_data_vectors_along_a_line(prec):
    """Euclidean vectors spread over 2000 units along one direction with
    unit spacing and 1e-3 noise across it: squared norms of about 1e6
    against neighbour distances of about 1, so approximate scores go below
    zero (the mirrored-append and Newton-start cases of the frame tests).
    Symmetric form = list path = oracle."""
    rng = np.random.default_rng(31)
    n, dim = 900, 24
    u = rng.normal(size=dim)
    u /= np.linalg.norm(u)
    t = rng.permutation(np.linspace(-1000.0, 1000.0, n))
    x = t[:, None] * u[None, :] + 1e-3 * rng.normal(size=(n, dim))
    out = []
    for sym in (1, 0):
        with knn.KnnContext(0, prec) as ctx:
            ctx.set_option(knn.OPT_SELF_SYMMETRIC, sym)
            ctx.set_data_reference(x, "euclidean")
            d, i = ctx.data_query(None, 6)
            out.append((d, i, ctx.stats()))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][0], out[1][0])
    d_o, i_o = oracle.knn_data(x, x, 6, "euclidean")
    assert np.abs(out[0][0] - d_o).max() < 1e-6


def test_symmetric_then_cross_then_serving_on_one_context():
    """A self query (mirrored lists), a cross query (lists) and a served
    request share one context's workspace; each gives what it gives on a
    fresh context.  A cross query, and a self query ending while serving
    runs, free the mirrored lists."""
    frames = synth.metastable_trajectory(1500, 100, 8, n_centres=5, dup_fraction=0.1)
    other = synth.trajectory(400, 100, 9)
    w = synth.masses(100, "backbone")
    d_self, i_self, _ = _self_query("bf16", frames, w, 12)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(frames, w)
        d_cross, i_cross = ctx.query(other, 12)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(frames, w)
        for _ in range(2):
            d, i = ctx.query(None, 12)
            assert ctx.stats()["n_pairs_computed"] < 1500 * 1500
            np.testing.assert_array_equal(d, d_self)
            np.testing.assert_array_equal(i, i_self)
            d, i = ctx.query(other, 12)
            np.testing.assert_array_equal(d, d_cross)
            np.testing.assert_array_equal(i, i_cross)
        r = ctx.wait(ctx.submit(other, 12))
        np.testing.assert_array_equal(r.dist, d_cross)
        np.testing.assert_array_equal(r.idx, i_cross)
        # a self query while serving runs, then another request
        d, i = ctx.query(None, 12)
        np.testing.assert_array_equal(d, d_self)
        np.testing.assert_array_equal(i, i_self)
        r = ctx.wait(ctx.submit(other, 12))
        np.testing.assert_array_equal(r.dist, d_cross)
        # a new reference set after the self query
        ctx.set_reference(other, w)
        d, i = ctx.query(None, 5)
        d_o, i_o = oracle.knn_exact(other, other, w, 5)
        assert np.abs(d - d_o).max() < 1e-4


@pytest.mark.parametrize("n", [2, 3, 5, 33])
@pytest.mark.parametrize("flags", [0, knn.KNN_KEEP_RANK0])
def test_symmetric_tiny_sets(n, flags):
    """Sets smaller than one tile, with the largest k each flag allows."""
    frames = synth.trajectory(n, 24, 30 + n)
    w = synth.masses(24, "random", n)
    k = n if flags & knn.KNN_KEEP_RANK0 else n - 1
    d, i, st = _both("bf16", frames, w, k, flags)
    d_o, i_o = oracle.knn_exact(frames, frames, w, k, keep_rank0=bool(flags))
    # the oracle clamps k to n-1 even when rank 0 is kept
    m = d_o.shape[1]
    check_knn(frames, frames, w, d[:, :m], i[:, :m], d_o, i_o)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_self_query_wide_k_takes_the_exact_path(prec):
    """k+1 above the candidate lists' 896: a self query scores every pair
    exactly in FP64, with or without the symmetric option, and matches the
    oracle."""
    frames = synth.metastable_trajectory(3000, 48, 12, n_centres=4, dup_fraction=0.05)
    w = synth.masses(48, "backbone")
    k = 1200
    d_s, i_s, st_s = _self_query(prec, frames, w, k)
    d_e, i_e, st_e = _self_query(prec, frames, w, k, sym=False)
    assert st_s["n_fallback_rows"] == 3000 and st_e["n_fallback_rows"] == 3000
    np.testing.assert_array_equal(i_s, i_e)
    np.testing.assert_array_equal(d_s, d_e)
    rows = np.arange(0, 3000, 97)
    d_o, i_o = oracle.knn_exact(frames, frames[rows], w, k)
    # rank 0 (the frame itself) is dropped in both
    check_knn(frames, frames[rows], w, d_s[rows], i_s[rows], d_o, i_o)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("cap", [1, 40])
def test_symmetric_mirror_overflow_relists_rows(prec, cap):
    """Mirrored lists too short for what the rows receive: the merge flags
    those rows, they are computed again through the candidate lists, and the
    outputs equal the list path's."""
    frames = synth.metastable_trajectory(2500, 80, 19, n_centres=3, dup_fraction=0.05)
    w = synth.masses(80, "backbone")
    k = 12
    d_s, i_s, st_s = _self_query(prec, frames, w, k, mirror_cap=cap)
    d_l, i_l, st_l = _self_query(prec, frames, w, k, sym=False)
    print(prec, cap, "relisted", st_s["n_relisted_rows"], "exact", st_s["n_fallback_rows"],
          "mirrored", st_s["n_mirrored"], "max", st_s["n_mirrored_max"])
    assert st_s["n_pairs_computed"] < 2500 * 2500
    assert st_s["n_relisted_rows"] > 0
    np.testing.assert_array_equal(i_s, i_l)
    np.testing.assert_array_equal(d_s, d_l)


def test_symmetric_then_cross_then_serving_within_a_budget():
    """A budget that holds the symmetric self query's lists, but not those
    plus what a later cross query and serving need unless the mirrored lists
    are freed: self, cross and served queries all succeed, with the outputs
    of fresh contexts."""
    n = 20000
    frames = synth.trajectory(n, 16, 4)
    other = synth.trajectory(3000, 16, 5, base_seed=4)
    w = synth.masses(16, "uniform")
    d_s, i_s, st_s = _self_query("bf16", frames, w, 10)
    assert st_s["n_pairs_computed"] < n * n
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(frames, w)
        d_c, i_c = ctx.query(other, 10)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_option(knn.OPT_HBM_BUDGET, 2.3e9)
        ctx.set_reference(frames, w)
        d, i = ctx.query(None, 10)
        assert ctx.stats()["n_pairs_computed"] < n * n
        np.testing.assert_array_equal(d, d_s)
        np.testing.assert_array_equal(i, i_s)
        d, i = ctx.query(other, 10)
        np.testing.assert_array_equal(d, d_c)
        np.testing.assert_array_equal(i, i_c)
        r = ctx.wait(ctx.submit(other, 10))
        np.testing.assert_array_equal(r.dist, d_c)
        np.testing.assert_array_equal(r.idx, i_c)


def test_symmetric_400k_frames():
    """400,000 frames (a size whose n x n FP32 matrix, 640 GB, would not fit
    one GPU) take the symmetric form: lists only."""
    n, natoms, k = 400_000, 24, 16
    frames = synth.trajectory(n, natoms, 40)
    w = synth.masses(natoms, "uniform")
    d, i, st = _self_query("bf16", frames, w, k)
    print("400k self query", st)
    assert st["n_pairs_computed"] < n * n // 2 + n * 256
    assert st["n_fallback_rows"] == 0
    rows = np.array([0, 1234, 199_999, 399_999])
    d_o, i_o = oracle.knn_exact(frames, frames[rows], w, k)
    check_knn(frames, frames[rows], w, d[rows], i[rows], d_o, i_o)


@pytest.mark.parametrize("sym", [1, 0])
def test_run_time_window_check_catches_a_narrow_window(sym):
    """With the candidate window shrunk far below the measured deviation of
    the approximate scores (MDSCTK_OPT_EPS_SCALE), the merge's check of the
    re-scored candidates sends rows to the exact path, so the outputs stay
    exact; at the default window no row trips it."""
    frames = synth.metastable_trajectory(1200, 200, 17, n_centres=4, dup_fraction=0.1)
    w = synth.masses(200, "backbone")
    k = 10
    d_o, i_o = oracle.knn_exact(frames, frames, w, k)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_option(knn.OPT_SELF_SYMMETRIC, sym)
        ctx.set_reference(frames, w)
        d, i = ctx.query(None, k)
        assert ctx.stats()["n_fallback_rows"] == 0
        ctx.set_option(knn.OPT_EPS_SCALE, 1e-4)
        d2, i2 = ctx.query(None, k)
        st = ctx.stats()
    print("narrow window: fallback rows", st["n_fallback_rows"], "of 1200")
    assert st["n_fallback_rows"] > 0
    check_knn(frames, frames, w, d, i, d_o, i_o)
    check_knn(frames, frames, w, d2, i2, d_o, i_o)
    np.testing.assert_array_equal(d2, d)
    np.testing.assert_array_equal(i2, i)
