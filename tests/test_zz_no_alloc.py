"""The query entry point that allocates nothing
(mdsctk_knn_query_device_noalloc) against mdsctk_knn_query_device: the same
arguments give the same rows, bit for bit, on two calls in a row (the second
with fewer frames, so that the fixed region still holds the first call's
contents), for the fused path, the exact form, full rows and the self query.
Sizes beyond the fixed region are refused with a clear error.

All calls go through the C ABI (libmdsctk_knn.so via ctypes); device buffers
come from the HIP runtime directly."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.devmem import DevBuf
from tests.parity import check_knn

pytestmark = pytest.mark.gpu


def _device_query(ctx, noalloc, xyz, k, flags=0):
    n_fit = ctx.n_ref if xyz is None else len(xyz)
    width = ctx.row_width(k, flags)
    d_x = DevBuf.upload(np.ascontiguousarray(xyz, np.float32)) if xyz is not None else None
    try:
        with DevBuf(n_fit * width * 8) as d_d, DevBuf(n_fit * width * 4) as d_i:
            fn = ctx.query_device_noalloc if noalloc else ctx.query_device
            got = fn(d_x.ptr if d_x else 0, n_fit, k, flags, d_d.ptr, d_i.ptr)
            assert got == width
            return (d_d.download((n_fit, width), np.float64),
                    d_i.download((n_fit, width), np.int32), ctx.stats())
    finally:
        if d_x:
            d_x.free()


def _both(ctx, xyz, k, flags=0):
    d_a, i_a, st_a = _device_query(ctx, False, xyz, k, flags)
    d_b, i_b, st_b = _device_query(ctx, True, xyz, k, flags)
    np.testing.assert_array_equal(i_b, i_a)
    np.testing.assert_array_equal(d_b, d_a)
    return d_b, i_b, st_b


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "exact"])
def test_noalloc_matches_query_device_twice(prec):
    ref = synth.metastable_trajectory(2500, 120, 11, n_centres=8, dup_fraction=0.1)
    w = synth.masses(120, "backbone")
    fit = synth.metastable_trajectory(700, 120, 12, n_centres=8, dup_fraction=0.1)
    k = 20
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(ref, w)
        d1, i1, st1 = _both(ctx, fit, k)
        d2, i2, st2 = _both(ctx, fit[:77], k)  # second call, smaller input
        print(prec, "phases", st1["ref_splits"], st2["ref_splits"], "fallback rows",
              st1["n_fallback_rows"], st2["n_fallback_rows"])
    rows = [0, 13, 400, 699]
    d_o, i_o = oracle.knn_exact(ref, fit[rows], w, k)
    check_knn(ref, fit[rows], w, d1[rows], i1[rows], d_o, i_o)
    np.testing.assert_array_equal(d2, d1[:77])
    np.testing.assert_array_equal(i2, i1[:77])


def test_noalloc_self_query_full_rows_and_flags():
    ref = synth.trajectory(1800, 64, 5)
    w = synth.masses(64, "random", 5)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        # self query: the symmetric form on one side, candidate lists on the other
        d, i, _ = _both(ctx, None, 12)
        _both(ctx, ref[:300], 12, knn.KNN_NOFIT | knn.KNN_KEEP_RANK0)
        _both(ctx, ref[:40], 0, knn.KNN_FULL_ROWS)
        _both(ctx, ref[:9], 12)
    d_o, i_o = oracle.knn_exact(ref, ref[:5], w, 12)
    check_knn(ref, ref[:5], w, d[:5], i[:5], d_o, i_o)


def test_noalloc_forced_exact_rows():
    """A window scale of 1e-4 sends every row to the exact path, whose row
    scratch then also comes from the fixed region."""
    ref = synth.trajectory(900, 48, 8)
    w = synth.masses(48, "backbone")
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_option(knn.OPT_EPS_SCALE, 1e-4)
        ctx.set_reference(ref, w)
        _, _, st = _both(ctx, ref[100:260], 7)
        assert st["n_fallback_rows"] == 160
        _, _, st = _both(ctx, ref[100:130], 7)
        assert st["n_fallback_rows"] == 30


def test_noalloc_refuses_oversize_and_data_contexts():
    ref = synth.trajectory(400, 1000, 3)
    w = synth.masses(1000)
    big = np.zeros((12000, 1000, 3), np.float32)  # 12000 x 12 KB of operand alone
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        with DevBuf.upload(big) as d_x, DevBuf(12000 * 5 * 8) as d_d, DevBuf(12000 * 5 * 4) as d_i:
            rc = ctx._L.mdsctk_knn_query_device_noalloc(ctx._h, d_x.ptr, 12000, 5, 0, d_d.ptr,
                                                         d_i.ptr)
            assert rc == knn.ERR_ARG
            assert b"fixed region" in ctx._L.mdsctk_knn_last_error(ctx._h)
        # the context still answers afterwards, through either entry point
        _both(ctx, ref[:50], 5)
    x = synth.data_vectors(200, 30, 4)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_data_reference(x, "euclidean")
        with DevBuf(8 * 5 * 8) as d_d, DevBuf(8 * 5 * 4) as d_i:
            rc = ctx._L.mdsctk_knn_query_device_noalloc(ctx._h, None, 0, 5, 0, d_d.ptr, d_i.ptr)
            assert rc == knn.ERR_STATE
