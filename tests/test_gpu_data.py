"""GPU parity of knn_data (data vectors, Euclidean and correlation
distances) on the knn engine against the FP64 oracle restatement of
knn_data.cpp / mdsctk.cpp:329-356.  All calls go through the C ABI."""
import ctypes
import subprocess

import numpy as np
import pytest

import oracle
from mdsctk_amd import build, knn, synth

pytestmark = pytest.mark.gpu

# Distances are FP64 re-scores of the same formulas, so they agree with the
# oracle's sequential sums to rounding: Euclidean to 1e-9 relative.  The
# correlation distance sqrt(v), v = (1 - r) / 2, inherits the conditioning of
# the reference's one-pass standard deviation (n sum x^2 - (sum x)^2): with a
# large common offset (the "offset" vectors, mean 50, spread 1) its last
# ~4 digits are rounding noise in any summation order, so v is compared to
# 1e-11 absolute (d to 1e-9 where d is not near 0).
TOL = 1e-9
TOL_V = 1e-11


def close(a, b, metric):
    if metric == "euclidean":
        return np.abs(a - b) <= TOL * np.maximum(1.0, np.abs(b))
    return (np.abs(a - b) <= TOL) | (np.abs(a * a - b * b) <= TOL_V)


def check_rows(ref, fit, d, i, d_o, i_o, metric):
    """Sorted distances agree; every returned index is distinct and its own
    oracle distance equals the returned one, so index sets may differ only
    among ties closer than the tolerance."""
    assert d.shape == d_o.shape and i.shape == i_o.shape
    if d.size == 0:
        return
    ok = close(d, d_o, metric)
    assert ok.all(), f"distance mismatch {np.abs(d - d_o)[~ok].max():.3e}"
    for r in range(i.shape[0]):
        assert len(set(i[r].tolist())) == i.shape[1]
        own = np.array([oracle.data_distance(fit[r], ref[j], metric) for j in i[r]])
        assert close(d[r], own, metric).all()


@pytest.mark.parametrize("prec", ["bf16", "fp32", "exact"])
@pytest.mark.parametrize("metric", ["euclidean", "correlation"])
@pytest.mark.parametrize("kind", ["clusters", "dihedral", "offset"])
def test_data_knn_vs_oracle(prec, metric, kind):
    ref = synth.data_vectors(700, 45, 1, kind)
    fit = synth.data_vectors(230, 45, 2, kind)
    k = 15
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_data_reference(ref, metric)
        d, i = ctx.data_query(fit, k)
        st = ctx.stats()
        ds, is_ = ctx.data_query(None, k)
    check_rows(ref, fit, d, i, *oracle.knn_data(ref, fit, k, metric), metric)
    check_rows(ref, ref, ds, is_, *oracle.knn_data(ref, ref, k, metric), metric)
    if prec != "exact":
        assert st["n_fallback_rows"] == 0


@pytest.mark.parametrize("metric", ["euclidean", "correlation"])
def test_data_full_rows_rank0_and_clamp(metric):
    ref = synth.data_vectors(150, 64, 3)
    fit = synth.data_vectors(40, 64, 4)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_data_reference(ref, metric)
        full, ident = ctx.data_query(fit, 0, knn.KNN_FULL_ROWS)
        d0, i0 = ctx.data_query(fit, 9, knn.KNN_KEEP_RANK0)
        dc, ic = ctx.data_query(fit[:5], 500)
    want = np.array([[oracle.data_distance(f, r, metric) for r in ref] for f in fit])
    assert close(full, want, metric).all()
    assert (ident == np.arange(150)[None, :]).all()
    check_rows(ref, fit, d0, i0, *oracle.knn_data(ref, fit, 9, metric, keep_rank0=True), metric)
    assert dc.shape == (5, 149)
    check_rows(ref, fit[:5], dc, ic, *oracle.knn_data(ref, fit[:5], 500, metric), metric)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("kind", ["clusters", "dihedral", "offset"])
@pytest.mark.parametrize("vdim", [7, 96, 700])
def test_data_approx_scores_within_window(prec, kind, vdim):
    """The contraction's score of data vectors stays inside the window
    kappa (G_q + max G_ref) the engine reports (kappa includes the proven
    FP32-input term), for both metrics."""
    ref = synth.data_vectors(96, vdim, 5, kind)
    fit = synth.data_vectors(80, vdim, 6, kind)
    for metric in ("euclidean", "correlation"):
        with knn.KnnContext(0, prec) as ctx:
            ctx.set_data_reference(ref, metric)
            a, e = ctx.data_debug_scores(fit)
            ctx.data_query(fit, 3)
            kappa = ctx.stats()["eps_rel"]
        if metric == "euclidean":
            mu = ref.mean(axis=0)
            G = ((ref - mu) ** 2).sum(1)
            Gf = ((fit - mu) ** 2).sum(1)
            want = ((fit[:, None, :] - ref[None, :, :]) ** 2).sum(-1)
        else:
            G, Gf = np.ones(len(ref)), np.ones(len(fit))
            r = np.corrcoef(np.vstack([fit, ref]))[:len(fit), len(fit):]
            want = 2.0 * (1.0 - r)
        # exact scores: to rounding (Euclidean), to the one-pass form's
        # conditioning (correlation, s = 4 v)
        assert np.abs(e - want).max() <= (1e-9 * max(1.0, np.abs(want).max())
                                          if metric == "euclidean" else 4 * TOL_V)
        rel = np.abs(a - e) / (Gf[:, None] + G.max())
        print(prec, kind, vdim, metric, "max rel", rel.max(), "kappa", kappa)
        assert rel.max() < kappa


def test_data_errors():
    ref = synth.data_vectors(50, 10, 7)
    with knn.KnnContext(0, "fp16") as ctx:
        with pytest.raises(knn.KnnError):
            ctx.set_data_reference(ref)
    with knn.KnnContext(0, "bf16") as ctx:
        bad = ref.copy()
        bad[3] = 2.5
        with pytest.raises(knn.KnnError):
            ctx.set_data_reference(bad, "correlation")
        bad = ref.copy()
        bad[4, 2] = np.nan
        with pytest.raises(knn.KnnError):
            ctx.set_data_reference(bad)
        ctx.set_data_reference(ref)
        with pytest.raises(knn.KnnError):       # frames entry point on a data set
            ctx.query(None, 3)
        with pytest.raises(knn.KnnError):
            ctx.data_query(ref, 3, knn.KNN_NOFIT)
        # serving takes frames only (called through the C ABI directly: the
        # Python wrapper checks frame shapes first)
        x = np.zeros((2, 4, 3), np.float32)
        d, i = np.zeros((2, 3)), np.zeros((2, 3), np.int32)
        rid = ctypes.c_int64()
        rc = ctx._L.mdsctk_knn_submit(ctx._h, knn._ptr(x), 2, 3, 0, knn._ptr(d), knn._ptr(i),
                                      ctypes.byref(rid))
        assert rc == knn.ERR_STATE
        d, i = ctx.data_query(ref, 3)           # still usable after the errors
        check_rows(ref, ref, d, i, *oracle.knn_data(ref, ref, 3), "euclidean")
        # the same context then takes frames again
        frames = synth.trajectory(60, 20, 8)
        w = synth.masses(20)
        ctx.set_reference(frames, w)
        with pytest.raises(knn.KnnError):
            ctx.data_query(ref, 3)
        ctx.query(None, 3)


def run(args, cwd):
    return subprocess.run([build.KNN_DATA_PATH] + args, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=300)


@pytest.mark.parametrize("corr", [False, True])
def test_knn_data_cli_self_query_forms_agree(tmp_path, corr):
    """knn_data without -f (a self query through the score matrix) and with
    -f naming a copy of the reference file (a cross query through the
    candidate lists) write identical files."""
    ref = synth.data_vectors(700, 21, 13, "clusters")
    ref.tofile(tmp_path / "reference.pts")
    ref.tofile(tmp_path / "copy.pts")
    extra = ["-c"] if corr else []
    outs = []
    for tag, more in (("a", []), ("b", ["-f", "copy.pts"])):
        p = run(["-k", "9", "-v", "21", "-d", f"d{tag}.dat", "-i", f"i{tag}.dat"] + more + extra,
                tmp_path)
        assert p.returncode == 0, p.stdout + p.stderr
        outs.append(((tmp_path / f"d{tag}.dat").read_bytes(), (tmp_path / f"i{tag}.dat").read_bytes()))
    assert outs[0] == outs[1]


@pytest.mark.parametrize("corr", [False, True])
def test_knn_data_cli(tmp_path, corr):
    ref = synth.data_vectors(400, 33, 9, "dihedral")
    fit = synth.data_vectors(90, 33, 10, "dihedral")
    ref.tofile(tmp_path / "reference.pts")
    with open(tmp_path / "fit.pts", "wb") as f:   # a trailing partial vector is ignored
        f.write(fit.tobytes() + b"\0" * 16)
    metric = "correlation" if corr else "euclidean"
    extra = ["-c"] if corr else []
    p = run(["-k", "10", "-v", "33", "-f", "fit.pts"] + extra, tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d = np.fromfile(tmp_path / "distances.dat", np.float64).reshape(90, 10)
    i = np.fromfile(tmp_path / "indices.dat", np.int32).reshape(90, 10)
    check_rows(ref, fit, d, i, *oracle.knn_data(ref, fit, 10, metric), metric)
    p = run(["-k", "10", "-v", "33", "--precision", "fp32", "-d", "s.dat", "-i", "s.idx"] + extra,
            tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d = np.fromfile(tmp_path / "s.dat", np.float64).reshape(400, 10)
    i = np.fromfile(tmp_path / "s.idx", np.int32).reshape(400, 10)
    check_rows(ref, ref, d, i, *oracle.knn_data(ref, ref, 10, metric), metric)
    p = run(["-s", "false", "-v", "33", "-f", "fit.pts", "-d", "f.dat", "-i", "f.idx"] + extra,
            tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    full = np.fromfile(tmp_path / "f.dat", np.float64).reshape(90, 400)
    want = np.array([[oracle.data_distance(a, b, metric) for b in ref] for a in fit])
    assert close(full, want, metric).all()
