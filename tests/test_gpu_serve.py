"""Serving form: asynchronous requests with continuous batching.

Results must not depend on what a request is batched with: each request's
rows are compared bit-for-bit between running alone and running in a mix
(split over batches, sharing batches with other requests), and against the
FP64 oracle within 1e-5 nm."""
import numpy as np
import pytest

import oracle
from mdsctk_amd import knn, synth
from tests.parity import check_knn

pytestmark = pytest.mark.gpu


def _set(n_ref, natoms, seed):
    ref = synth.trajectory(n_ref, natoms, seed)
    w = synth.masses(natoms, "backbone", seed)
    return ref, w


def _queries(sizes, natoms, seed):
    pool = synth.trajectory(int(sum(sizes)), natoms, seed + 7, base_seed=seed)
    out, o = [], 0
    for n in sizes:
        out.append(pool[o:o + n])
        o += n
    return out


@pytest.mark.parametrize("graphs,share", [(True, False), (False, False), (True, True)])
def test_serve_matches_sync_and_oracle(graphs, share):
    ref, w = _set(700, 48, 21)
    qs = _queries([1, 5, 40, 129, 300], 48, 21)
    k = 9
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_option(knn.OPT_SERVE_GRAPHS, int(graphs))
        ctx.set_option(knn.OPT_SERVE_SHARE_CUS, int(share))
        ctx.set_reference(ref, w)
        sync = [ctx.query(q, k) for q in qs]
        reqs = [ctx.submit(q, k) for q in qs]
        for r in reqs:
            ctx.wait(r)
        st = ctx.serve_stats()
    for q, (d, i), r in zip(qs, sync, reqs):
        assert (r.dist == d).all() and (r.idx == i).all()
        d_o, i_o = oracle.knn_exact(ref, q, w, k)
        check_knn(ref, q, w, r.dist, r.idx, d_o, i_o)
    assert st["requests_done"] == len(qs)
    assert st["frames"] == sum(len(q) for q in qs)
    assert st["graph_launches"] == (st["batches"] if graphs else 0)


def test_graph_replays_match_direct_launches():
    """One executable graph per batch slot, updated and replayed for batch
    after batch of different sizes and k, with the serving loop's own copies
    and synchronisation between replays: every replay equals the direct
    synchronous query of the same frames."""
    ref, w = _set(900, 40, 24)
    sizes = [3, 64, 1, 200, 17, 96, 5, 130, 40, 2, 77, 31]
    qs = _queries(sizes, 40, 24)
    ks = [7, 7, 7, 12, 7, 7, 3, 7, 7, 7, 12, 7]
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_option(knn.OPT_SERVE_SLOTS, 1)
        ctx.set_reference(ref, w)
        sync = [ctx.query(q, k) for q, k in zip(qs, ks)]
        served = [ctx.wait(ctx.submit(q, k)) for q, k in zip(qs, ks)]   # one batch each
        st = ctx.serve_stats()
    for (d, i), r in zip(sync, served):
        assert (r.dist == d).all() and (r.idx == i).all()
    assert st["batches"] == len(qs) and st["graph_launches"] == len(qs)
    # the executable was updated in place, not rebuilt per batch
    assert st["graph_builds"] < len(qs)


@pytest.mark.parametrize("prec,graphs", [("bf16", True), ("fp32", True), ("bf16", False)])
def test_alone_vs_mix_bit_identical(prec, graphs):
    """The same requests alone (one at a time) and submitted together, with
    batches capped at 128 frames so that large requests are split over
    batches and small ones share them."""
    ref, w = _set(1500, 64, 22)
    sizes = [1, 3, 17, 64, 100, 250, 31, 2, 400, 77]
    qs = _queries(sizes, 64, 22)
    k = 12
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_option(knn.OPT_SERVE_MAX_FRAMES, 128)
        ctx.set_option(knn.OPT_SERVE_GRAPHS, int(graphs))
        ctx.set_reference(ref, w)
        alone = []
        for q in qs:
            alone.append(ctx.wait(ctx.submit(q, k)))
        mixed = [ctx.submit(q, k) for q in qs]
        for r in mixed:
            ctx.wait(r)
        st = ctx.serve_stats()
    for a, m in zip(alone, mixed):
        assert (a.dist == m.dist).all(), "distances differ alone vs in the mix"
        assert (a.idx == m.idx).all(), "indices differ alone vs in the mix"
    # the mix really was mixed: a 400-frame request spans >= 4 batches and
    # some small request shared a batch larger than itself
    big = mixed[sizes.index(400)]
    assert big.info["n_batches"] >= 4
    assert any(r.info["batch_frames"] > n for r, n in zip(mixed, sizes))
    assert st["max_batch_frames"] <= 128
    for q, r in zip(qs[:4], mixed[:4]):
        d_o, i_o = oracle.knn_exact(ref, q, w, k)
        check_knn(ref, q, w, r.dist, r.idx, d_o, i_o)


def test_mixed_k_and_flags_in_flight():
    ref, w = _set(400, 30, 23)
    qs = _queries([20, 33, 7, 50], 30, 23)
    shapes = [(5, 0), (11, knn.KNN_NOFIT), (5, 0), (3, knn.KNN_KEEP_RANK0)]
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        reqs = [ctx.submit(q, k, f) for q, (k, f) in zip(qs, shapes)]
        for r in reqs:
            ctx.wait(r)
    for q, (k, f), r in zip(qs, shapes, reqs):
        d_o, i_o = oracle.knn_exact(ref, q, w, k, nofit=bool(f & knn.KNN_NOFIT),
                                    keep_rank0=bool(f & knn.KNN_KEEP_RANK0))
        check_knn(ref, q, w, r.dist, r.idx, d_o, i_o, nofit=bool(f & knn.KNN_NOFIT))


def test_hbm_budget_caps_batches():
    """A tight HBM budget shrinks the per-batch frame limit instead of
    failing; results stay exact."""
    ref, w = _set(3000, 40, 24)
    qs = _queries([900, 5], 40, 24)
    k = 8
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        ctx.set_option(knn.OPT_HBM_BUDGET, 256 * 2**20)
        reqs = [ctx.submit(q, k) for q in qs]
        for r in reqs:
            ctx.wait(r)
        st = ctx.serve_stats()
    assert st["slot_frames"] < 2048
    assert reqs[0].info["n_batches"] >= 2
    for q, r in zip(qs, reqs):
        d_o, i_o = oracle.knn_exact(ref, q, w, k)
        check_knn(ref, q, w, r.dist, r.idx, d_o, i_o)


def test_serve_errors_and_retirement():
    ref, w = _set(100, 12, 25)
    q = _queries([4], 12, 25)[0]
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        with pytest.raises(knn.KnnError) as e:
            ctx.submit(q, 3, knn.KNN_FULL_ROWS)
        assert e.value.code == knn.ERR_ARG
        r = ctx.wait(ctx.submit(q, 3))
        assert r.done and r.dist.shape == (4, 3)
        # a retired id is unknown
        info = knn.RequestInfo()
        assert ctx._L.mdsctk_knn_poll(ctx._h, r.id, info) == knn.ERR_STATE
        # the reference set can be replaced once nothing is outstanding,
        # and serving resumes against the new set
        ref2, _ = _set(90, 12, 26)
        ctx.set_reference(ref2, w)
        r2 = ctx.wait(ctx.submit(q, 3))
    d_o, i_o = oracle.knn_exact(ref2, q, w, 3)
    check_knn(ref2, q, w, r2.dist, r2.idx, d_o, i_o)


def test_submit_poll_wait_from_several_threads():
    """Four host threads submit, poll and wait at the same time (the C ABI
    allows the serving calls from any thread), while the main thread runs a
    synchronous query on the same context; every result equals the same
    request run synchronously, and errors stay on the thread that caused
    them."""
    import threading
    ref, w = _set(1500, 40, 33)
    qs = _queries([1, 7, 33, 128, 257, 64, 3, 900], 40, 33)
    k = 6
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(ref, w)
        want = [ctx.query(q, k) for q in qs]
        got = {}
        errors = []

        def worker(t):
            try:
                mine = []
                for rep in range(6):
                    for j in range(t, len(qs), 4):
                        mine.append((j, ctx.submit(qs[j], k)))
                for j, r in mine:
                    if j % 2:
                        while not ctx.poll(r):
                            pass
                    else:
                        ctx.wait(r)
                    got.setdefault(j, []).append((r.dist.copy(), r.idx.copy()))
                # a failing call on this thread: its error text is this thread's
                import ctypes
                q0 = np.ascontiguousarray(qs[0], np.float32)
                out_d = np.zeros((1, k)), np.zeros((1, k), np.int32)
                rid = ctypes.c_int64()
                rc = ctx._L.mdsctk_knn_submit(ctx._h, q0.ctypes.data, 1, -1, 0,
                                              out_d[0].ctypes.data, out_d[1].ctypes.data,
                                              ctypes.byref(rid))
                assert rc == knn.ERR_ARG and "k must be" in ctx.last_error()
            except Exception as e:  # pragma: no cover - reported below
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
        for th in threads:
            th.start()
        d_sync, i_sync = ctx.query(qs[4], k)  # a synchronous query meanwhile
        for th in threads:
            th.join()
        assert not errors, errors
        assert ctx.last_error() == ""  # the workers' failures stayed on their threads
    np.testing.assert_array_equal(d_sync, want[4][0])
    for j, runs in got.items():
        assert len(runs) == 6
        for d, i in runs:
            np.testing.assert_array_equal(d, want[j][0])
            np.testing.assert_array_equal(i, want[j][1])
