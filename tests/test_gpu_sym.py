"""GPU parity of the symmetric self query: the contraction scores the upper
block triangle into an n x n FP32 matrix, a select pass picks each row's
candidates, and the merge re-scores them.  Its outputs must equal the
candidate-list path's bit for bit (both re-score a certified window in
FP64), and the FP64 oracle's within the distance tolerance.

All calls go through the C ABI (libmdsctk_knn.so via ctypes)."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.parity import check_knn

pytestmark = pytest.mark.gpu


def _self_query(prec, frames, w, k, flags=0, sym=True, budget=0):
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_option(knn.OPT_SELF_SYMMETRIC, 1 if sym else 0)
        if budget:
            ctx.set_option(knn.OPT_HBM_BUDGET, budget)
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
          "cand", st_s["n_candidates"], "vs", st_l["n_candidates"],
          "fallback", st_s["n_fallback_rows"], st_l["n_fallback_rows"])
    assert st_l["n_pairs_computed"] == n * n
    assert st_s["n_pairs_computed"] == _triangle_pairs(n, 64 if prec == "fp32" else 128)
    np.testing.assert_array_equal(i_s, i_l)
    np.testing.assert_array_equal(d_s, d_l)
    return d_s, i_s, st_s


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32"])
@pytest.mark.parametrize("n", [2000, 777, 130, 67])
def test_symmetric_equals_lists(prec, n):
    """Ragged sizes: n not a multiple of 4, 32, 64 or the fit block."""
    frames = synth.metastable_trajectory(n, 304, 7 + n, n_centres=6, dup_fraction=0.1)
    w = synth.masses(304, "backbone")
    _both(prec, frames, w, min(30, n - 1))


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("flags", [knn.KNN_NOFIT, knn.KNN_KEEP_RANK0,
                                   knn.KNN_NOFIT | knn.KNN_KEEP_RANK0])
def test_symmetric_flags(prec, flags):
    frames = synth.trajectory(450, 57, 5)
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
    select pass overflows, the merge sends the rows to the exact path, and
    the outputs still equal the list path's (ties by index)."""
    rng = np.random.default_rng(1)
    base = rng.normal(size=(1, 40, 3)).astype(np.float32)
    frames = np.repeat(base, 2200, axis=0)
    frames[2100:] += 0.05 * rng.normal(size=(100, 40, 3)).astype(np.float32)
    w = synth.masses(40, "uniform")
    d, i, st = _both("bf16", frames, w, 5)
    assert st["n_fallback_rows"] > 0
    assert np.abs(d[:2100]).max() < 1e-5


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_symmetric_duplicates_beyond_stage(prec):
    """5000 equal frames: more keys than the select pass can stage in LDS
    fall near the (k+1)-th smallest, so it reads the row again for the fine
    histogram and the gather; outputs still equal the list path's."""
    rng = np.random.default_rng(2)
    base = rng.normal(size=(1, 40, 3)).astype(np.float32)
    frames = np.repeat(base, 5100, axis=0)
    frames[5000:] += 0.05 * rng.normal(size=(100, 40, 3)).astype(np.float32)
    w = synth.masses(40, "uniform")
    d, i, st = _both(prec, frames, w, 5)
    assert np.abs(d[:5000]).max() < 1e-5


def test_symmetric_sampled_select():
    """Rows of 32768 keys or more take the sampled select.  34000 distinct
    frames (the sampled bound holds and stages few keys) plus 6000 equal
    ones (their near-zero keys overflow the LDS stage, so those rows fall
    back to the full select and then to exact rows)."""
    frames = synth.trajectory(40000, 16, 21)
    frames[34000:] = frames[34000]
    w = synth.masses(16, "uniform")
    d, i, st = _both("bf16", frames, w, 7)
    assert st["n_fallback_rows"] >= 6000
    assert np.abs(d[34000:]).max() < 1e-5


def test_symmetric_large_k():
    frames = synth.trajectory(1500, 64, 9)
    w = synth.masses(64, "uniform")
    _both("bf16", frames, w, 800)


def test_symmetric_budget_falls_back_to_lists():
    """A budget that holds the candidate lists (about 1.3 GB here) but not
    the 1.6 GB score matrix keeps the list path, with the same outputs."""
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


def test_symmetric_then_cross_then_serving_on_one_context():
    """A self query (score matrix), a cross query (lists) and a served
    request share one context's workspace; each gives what it gives on a
    fresh context.  While serving runs, a self query frees its score matrix
    when it ends."""
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
def test_symmetric_wide_k_beyond_list_capacity(prec):
    """k+1 above the candidate lists' 896: a self query still takes the
    symmetric form (select list and merge hold 2048) and equals the
    all-pairs FP64 path bit for bit."""
    frames = synth.metastable_trajectory(3000, 48, 12, n_centres=4, dup_fraction=0.05)
    w = synth.masses(48, "backbone")
    k = 1200
    d_s, i_s, st_s = _self_query(prec, frames, w, k)
    d_e, i_e, st_e = _self_query(prec, frames, w, k, sym=False)
    print(prec, "wide k", st_s["n_pairs_computed"], st_s["n_candidates"], st_s["n_fallback_rows"],
          st_s["ms_total"], "ms vs exact", st_e["ms_total"], "ms")
    assert st_s["n_pairs_computed"] == _triangle_pairs(3000, 64 if prec == "fp32" else 128)
    assert st_e["n_fallback_rows"] == 3000        # the all-pairs FP64 path
    np.testing.assert_array_equal(i_s, i_e)
    np.testing.assert_array_equal(d_s, d_e)
    rows = np.arange(0, 3000, 97)
    d_o, i_o = oracle.knn_exact(frames, frames[rows], w, k)
    # rank 0 (the frame itself) is dropped in both
    check_knn(frames, frames[rows], w, d_s[rows], i_s[rows], d_o, i_o)


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
