"""knn_rms on one MI355X: all-pairs fitted RMSD + per-query top-k.

Default workload (the precision-form configuration): 200,000 reference x
200,000 query frames (self query) x 1,000 atoms, k = 64, split-BF16
contraction with FP64 G terms, QCP and FP64 re-score.  Frames come from the
seeded metastable generator (50 centres, 0.1 nm noise, 5% near-duplicates at
1e-4 nm).  One step = one full query of the resident set (contraction, QCP,
top-k, FP64 re-score, outputs written to device memory).

Beside the headline number it reports, on the same data, the plain-FP32
contraction path, the serving form (Poisson arrivals at 200 requests/s of
1..4096 frames, log-uniform, against a 100,000 x 500 resident reference,
k = 50, continuous batching; p50/p99 latency), the contraction kernel's roofline
(HIP events on the engine's stream), accuracy against the FP64 oracle on a
sample, and the reference-equivalent CPU knn_rms timed on a bounded sample.

Prints ONE JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RMSD pairs/sec: plain FP32 vs split-BF16 MFMA at <=1e-5 nm; p99 query latency"
PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "fp32": 157.3}
FLOPS_PER_ATOM = {"bf16": 54, "fp16": 54, "fp32": 18}


class Hip:
    """Minimal HIP runtime access for device buffers (no torch needed)."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = ctypes.c_void_p()
        rc = self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({n}) failed: {rc}")
        return p.value

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))

    def d2h(self, host, dev, n):
        rc = self.L.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev),
                              ctypes.c_size_t(n), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")

    def copy_d2d(self, dst, src, n):
        rc = self.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(n), 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")

    def sync(self):
        self.L.hipDeviceSynchronize()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_fused(ctx, hip, n_fit, k, steps, warmup, first=None):
    """Self query of the resident set: `warmup` untimed steps, then `steps`
    timed ones.  first (a dict) receives the first call's wall time, which
    includes the engine's one-time workspace allocation."""
    d_dist = hip.malloc(n_fit * k * 8)
    d_idx = hip.malloc(n_fit * k * 4)
    try:
        for w in range(warmup):
            t0 = time.perf_counter()
            ctx.query_device(0, n_fit, k, 0, d_dist, d_idx)
            hip.sync()
            if w == 0 and first is not None:
                first["ms"] = (time.perf_counter() - t0) * 1e3
        hip.sync()
        stats = []
        t0 = time.perf_counter()
        for s in range(steps):
            ctx.query_device(0, n_fit, k, 0, d_dist, d_idx)
            stats.append(ctx.stats())
            st = stats[-1]
            log(f"  step {s}: {st['ms_total']:.1f} ms device, landmark {st['ms_landmark']:.1f} ms, "
                f"contraction {st['ms_contract']:.1f} ms, merge {st['ms_merge']:.1f} ms, "
                f"relisted rows {st['n_relisted_rows']} ({st['ms_fallback']:.1f} ms), "
                f"fallback rows {st['n_fallback_rows']}, mirrored {st['n_mirrored']} "
                f"(max {st['n_mirrored_max']} per row)")
        hip.sync()
        wall = (time.perf_counter() - t0) / steps
        dist = np.zeros((n_fit, k), np.float64)
        idx = np.zeros((n_fit, k), np.int32)
        hip.d2h(dist, d_dist, dist.nbytes)
        hip.d2h(idx, d_idx, idx.nbytes)
    finally:
        hip.free(d_dist)
        hip.free(d_idx)
    return wall, stats, dist, idx


def run_c4_batches(knn, synth, prec, args):
    """The serving form's batch kernels without the serving thread: the
    resident 100,000 x 500-atom set queried synchronously with batches of 1,
    1000 and 4096 frames (k = 50), args.c4_batches times each.  Each call
    runs the same prep, contraction and merge kernels a served batch of that
    size replays; the per-size stage times come from the library's HIP
    events, and a kernel trace of this run gives their per-kernel times."""
    n_res, na_res, kq = 100_000, 500, 50
    res = synth.trajectory(n_res, na_res, args.seed + 7)
    w = synth.masses(na_res)
    pool = synth.trajectory(4096, na_res, args.seed + 8, base_seed=args.seed + 7)
    out = {"workload": f"resident {n_res} x {na_res} atoms, synchronous batches, k={kq}",
           "by_batch_frames": {}}
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_reference(res, w)
        for nb in (1, 1000, 4096):
            ctx.query(pool[:nb], kq)  # warm-up: sizes the workspace
            st = []
            for _ in range(args.c4_batches):
                ctx.query(pool[:nb], kq)
                st.append(ctx.stats())
            out["by_batch_frames"][str(nb)] = {
                "calls": args.c4_batches,
                "ms_total": mean_stat(st, "ms_total"),
                "ms_contract": mean_stat(st, "ms_contract"),
                "ms_merge": mean_stat(st, "ms_merge"),
                "fallback_rows": int(sum(x["n_fallback_rows"] for x in st))}
            log(f"c4 batch {nb} frames: {out['by_batch_frames'][str(nb)]}")
    return out


def run_serving(knn, synth, oracle, prec, args):
    """Serving form: a resident 100,000 x 500-atom reference set; requests of
    log-uniform 1..4096 frames arrive as a Poisson process; each is submitted
    at its arrival time and latency runs from arrival to outputs written."""
    n_res, na_res, kq, pool_n = 100_000, 500, 50, 16384
    rng = np.random.default_rng(args.seed + 9)
    res = synth.trajectory(n_res, na_res, args.seed + 7)
    w = synth.masses(na_res)
    pool = synth.trajectory(pool_n, na_res, args.seed + 8, base_seed=args.seed + 7)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serving_rate, int(args.serving_rate * args.serving_seconds * 1.5) + 10))
    arrivals = arrivals[arrivals < args.serving_seconds]
    sizes = np.minimum(np.floor(np.exp(rng.uniform(0.0, np.log(4097.0), len(arrivals)))), 4096).astype(np.int64)
    starts = rng.integers(0, pool_n - sizes + 1)
    keep = set(rng.choice(len(arrivals), size=min(8, len(arrivals)), replace=False).tolist())
    log(f"serving: {len(arrivals)} requests over {args.serving_seconds:.0f} s, "
        f"{sizes.sum()} frames ({sizes.mean():.0f} per request)")
    lat, lat_frames, done_t, kept = [], [], [], {}
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_option(knn.OPT_SERVE_GRAPHS, args.serve_graphs)
        ctx.set_option(knn.OPT_SERVE_SHARE_CUS, args.serve_share_cus)
        ctx.set_option(knn.OPT_SERVE_SLOTS, args.serve_slots)
        if getattr(args, "serve_max_frames", 0):
            ctx.set_option(knn.OPT_SERVE_MAX_FRAMES, args.serve_max_frames)
        ctx.set_reference(res, w)
        for nw in (4096, 1, 700, 4096):  # warm-up: sizes the batch slots
            ctx.wait(ctx.submit(pool[:nw], kq))
        st0 = ctx.serve_stats()
        t0 = ctx.clock_ms() + 10.0
        outstanding = []

        def retire(block=False):
            still = []
            for j, r in outstanding:
                if block:
                    ctx.wait(r)
                if r.done or ctx.poll(r):
                    lat.append(r.info["t_done_ms"] - r.t_arrival)
                    lat_frames.append(int(sizes[j]))
                    done_t.append(r.info["t_done_ms"])
                    if j in keep:
                        kept[j] = (r.dist, r.idx, r.info["batch_frames"])
                    r.dist = r.idx = r.xyz = None
                else:
                    still.append((j, r))
            outstanding[:] = still

        for j, (a, nq, s0) in enumerate(zip(arrivals, sizes, starts)):
            target = t0 + a * 1e3
            while True:
                now = ctx.clock_ms()
                if now >= target:
                    break
                retire()
                if target - now > 1.5:
                    time.sleep(0.0005)
            r = ctx.submit(pool[s0:s0 + nq], kq)
            r.t_arrival = target
            outstanding.append((j, r))
        retire(block=True)
        st1 = ctx.serve_stats()
        st = {key: st1[key] - st0[key]
              for key in ("batches", "frames", "busy_ms", "fallback_rows", "graph_launches")}
        st["max_batch_frames"], st["slots"] = st1["max_batch_frames"], st1["slots"]
        # the kept requests again, each alone: results must be bit-identical
        same = True
        for j, (d, i, bf) in kept.items():
            r = ctx.wait(ctx.submit(pool[starts[j]:starts[j] + sizes[j]], kq))
            same &= bool((r.dist == d).all() and (r.idx == i).all())
    # accuracy of the kept requests' first rows vs the FP64 oracle
    err = 0.0
    for j, (d, i, _) in kept.items():
        q = pool[starts[j]:starts[j] + 1]
        d_o, _ = oracle.knn_exact(res, q, w, kq)
        err = max(err, float(np.abs(d[:1] - d_o).max()) / 10.0)
    lat = np.array(lat)
    lat_frames = np.array(lat_frames)
    span_s = (max(done_t) - t0) / 1e3
    # latency by request size; 900..1100 frames is the "1k-frame query"
    buckets = {}
    for lo, hi in ((1, 16), (17, 256), (257, 899), (900, 1100), (1101, 4096)):
        sel = lat[(lat_frames >= lo) & (lat_frames <= hi)]
        if len(sel):
            buckets[f"{lo}-{hi}"] = {"requests": int(len(sel)),
                                     "p50_ms": float(np.percentile(sel, 50)),
                                     "p99_ms": float(np.percentile(sel, 99)),
                                     "max_ms": float(sel.max())}
    frames = float(sizes.sum())
    return {"workload": f"resident {n_res} x {na_res} atoms, Poisson {args.serving_rate:.0f} req/s, "
                        f"1..4096 frames log-uniform, k={kq}",
            "duration_s": args.serving_seconds, "requests": int(len(lat)),
            "frames_per_request": frames / len(lat),
            "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "by_request_frames": buckets,
            "p99_ms_1k_frame_requests": buckets.get("900-1100", {}).get("p99_ms"),
            "throughput_pairs_s": frames * n_res / span_s,
            "offered_pairs_s": frames * n_res / args.serving_seconds,
            "gpu_busy_frac": st["busy_ms"] / (span_s * 1e3),
            "batches": st["batches"], "mean_batch_frames": st["frames"] / max(1, st["batches"]),
            "max_batch_frames": st["max_batch_frames"], "slots": st["slots"],
            "fallback_rows": st["fallback_rows"], "graph_launches": st["graph_launches"],
            "alone_vs_mix_identical": same, "alone_vs_mix_requests": len(kept),
            "sample_max_err_nm": err,
            "mode": "host buffers in and out (PCIe included), latency from scheduled arrival"}


def run_mixed_sets(knn, synth, oracle, args):
    """Mixed-size batched serving: three resident sets (30,000 x 150,
    100,000 x 500, 20,000 x 2,000 atoms), Poisson arrivals of 1..2048-frame
    queries (log-uniform) to a random set, k = 32, split-BF16; one context
    (serving thread, batch slots) per set.  Then the 30,000-frame set's
    self-kNN graph (make_sysparse semantics) and the 10 leading eigenvalues
    of its normalised Gaussian affinity (sigma = median kNN distance) against
    the CPU pipeline: FP64 oracle kNN, edge-dictionary make_sysparse, ARPACK."""
    import oracle.graph as og
    from mdsctk_amd import graph
    sets = [(30_000, 150), (100_000, 500), (20_000, 2_000)]
    kq, pool_n = 32, 2048
    rng = np.random.default_rng(args.seed + 21)
    data, pools, w = [], [], []
    for j, (n, na) in enumerate(sets):
        data.append(synth.trajectory(n, na, args.seed + 30 + j))
        pools.append(synth.trajectory(pool_n + 64, na, args.seed + 40 + j, base_seed=args.seed + 30 + j))
        w.append(synth.masses(na))
    rate, dur = args.mixed_rate, args.mixed_seconds
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * dur * 1.5) + 10))
    arrivals = arrivals[arrivals < dur]
    target = rng.integers(0, len(sets), len(arrivals))
    sizes = np.minimum(np.floor(np.exp(rng.uniform(0.0, np.log(2049.0), len(arrivals)))), 2048).astype(np.int64)
    starts = rng.integers(0, 64, len(arrivals))
    ctxs = [knn.KnnContext(0, "bf16") for _ in sets]
    lat = [[] for _ in sets]
    pairs = 0.0
    try:
        for c, x, ww in zip(ctxs, data, w):
            c.set_reference(x, ww)
        for c, pool in zip(ctxs, pools):       # warm-up: sizes the batch slots
            c.wait(c.submit(pool[:2048], kq))
            c.wait(c.submit(pool[:1], kq))
        t0 = ctxs[0].clock_ms() + 10.0
        outstanding = []
        for a, tg, nq, s0 in zip(arrivals, target, sizes, starts):
            tgt = t0 + a * 1e3
            while ctxs[0].clock_ms() < tgt:
                still = []
                for (j, r, t_arr) in outstanding:
                    if ctxs[j].poll(r):
                        lat[j].append(r.info["t_done_ms"] - t_arr)
                    else:
                        still.append((j, r, t_arr))
                outstanding = still
                if tgt - ctxs[0].clock_ms() > 1.5:
                    time.sleep(0.0005)
            r = ctxs[tg].submit(pools[tg][s0:s0 + nq], kq)
            outstanding.append((int(tg), r, tgt))
            pairs += float(nq) * sets[tg][0]
        for (j, r, t_arr) in outstanding:
            ctxs[j].wait(r)
            lat[j].append(r.info["t_done_ms"] - t_arr)
        # graph check on the 30,000-frame set
        t = time.time()
        d, i = ctxs[0].query(None, kq)
        gpu_s = time.time() - t
    finally:
        for c in ctxs:
            c.close()
    m = graph.ssm_build(d, i)
    sigma = float(np.median(d))
    ev, _, res, got = graph.spectral_decomp(m, sigma, 10, vectors=False)
    t = time.time()
    d_o, i_o = oracle.knn_exact(data[0], data[0], w[0], kq)
    cpu_s = time.time() - t
    n_o, pcol_o, irow_o, vals_o = og.ssm_build(d_o, i_o)
    w_o, _ = og.spectral_decomp(n_o, pcol_o, irow_o, vals_o, float(np.median(d_o)), 10)
    all_lat = np.concatenate([np.array(l) for l in lat if l])
    per_set = {f"{n}x{na}": {"requests": len(l), "p50_ms": float(np.percentile(l, 50)),
                              "p99_ms": float(np.percentile(l, 99))}
               for (n, na), l in zip(sets, lat) if l}
    return {"workload": f"3 resident sets {sets}, Poisson {rate:.0f} req/s over {dur:.0f} s, "
                        "1..2048 frames log-uniform, random set, k=32, split-BF16",
            "requests": int(len(all_lat)), "p50_ms": float(np.percentile(all_lat, 50)),
            "p99_ms": float(np.percentile(all_lat, 99)), "per_set": per_set,
            "pairs_per_s": pairs / (dur + 1e-3),
            "graph_check": {"set": "30000 x 150 self kNN, k=32", "sigma": sigma,
                            "eigenvalues": [float(v) for v in ev],
                            "cpu_pipeline_eigenvalues": [float(v) for v in w_o],
                            "max_abs_diff": float(np.abs(ev - w_o).max()), "tolerance": 1e-6,
                            "converged": int(got), "max_residual": float(res.max()),
                            "same_edges": bool((m.pcol == pcol_o).all() and (m.irow == irow_o).all()),
                            "gpu_knn_s": gpu_s, "cpu_knn_s": cpu_s}}


def run_cli_e2e(knn, synth, args):
    """The drop-in executable end to end on the precision-form workload: a
    200,000-frame x 1000-atom XTC at precision 1e5 (1e-5 nm steps keep the
    1e-4 nm near-duplicates) and a 1000-atom PDB are written to a scratch
    folder; knn_rms -k 64 runs as a child process (wall time from its start
    to its exit: topology, XTC read and GPU decode, GPU context, query,
    output files); its files must equal the library's query of the decoded
    frames bit for bit.  The XTC was just written, so it is read from the
    page cache."""
    import shutil
    import subprocess
    import tempfile
    from mdsctk_amd import build
    L = knn.load()
    work = tempfile.mkdtemp(prefix="mdsctk_cli_", dir="/tmp")
    try:
        n, natoms, k = args.n_frames, args.natoms, args.k
        x = synth.metastable_trajectory(n, natoms, args.seed)
        xtc = os.path.join(work, "traj.xtc")
        t = time.time()
        assert L.mdsctk_xtc_write(xtc.encode(), x.ctypes.data, n, natoms, 1e5, None, None, None) == 0
        log(f"cli: wrote {os.path.getsize(xtc) / 1e9:.2f} GB of XTC in {time.time() - t:.1f} s")
        del x
        pdb = os.path.join(work, "top.pdb")
        with open(pdb, "w") as fh:
            for a in range(natoms):
                fh.write("ATOM  %5d  CA  ALA A%4d    %8.3f%8.3f%8.3f  1.00  0.00           C\n"
                         % (a + 1, a % 10000, 0.0, 0.0, 0.0))
        cmd = [build.CLI_PATH, "-k", str(k), "-p", pdb, "-r", xtc]
        t = time.perf_counter()
        p = subprocess.run(cmd, cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        wall = time.perf_counter() - t
        if p.returncode != 0:
            return {"error": p.stdout[-2000:]}
        summary = [l for l in p.stdout.splitlines() if "pairs, k=" in l]
        # the library's query of the same decoded frames
        n_got = ctypes.c_int64()
        xd = np.zeros((n, natoms, 3), np.float32)
        assert L.mdsctk_xtc_read(xtc.encode(), xd.ctypes.data, n, natoms, ctypes.byref(n_got),
                                 None, None, None, 0) == 0
        w = np.full(natoms, 12.011, np.float32)
        with knn.KnnContext(0, "bf16") as ctx:
            ctx.set_reference(xd, w)
            d, i = ctx.query(None, k)
        same = (open(os.path.join(work, "distances.dat"), "rb").read() == d.tobytes() and
                open(os.path.join(work, "indices.dat"), "rb").read() == i.tobytes())
        return {"command": "knn_rms -k %d -p top.pdb -r traj.xtc" % k, "frames": n, "natoms": natoms,
                "xtc_bytes": os.path.getsize(xtc), "wall_s": wall,
                "files_equal_library_query": bool(same),
                "summary": summary[-1] if summary else p.stdout[-400:]}
    finally:
        shutil.rmtree(work, ignore_errors=True)


COPY_PROBE_BYTES = 1 << 30


def _pmc_child(counter, child_args, tag, workdir):
    """One `rocprofv3 --pmc <counter>` pass over a child bench.py run; returns
    {dispatch: (kernel name, summed counter value)} or None."""
    import csv
    import shutil
    import subprocess
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    out = os.path.join(workdir, tag)
    cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "--output-format", "csv",
           "-d", out, "-o", "run", "--", sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    path = os.path.join(out, "run_counter_collection.csv")
    if p.returncode != 0 or not os.path.exists(path):
        return None
    res = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        name, v = res.get(d, (r["Kernel_Name"], 0.0))
        res[d] = (name, v + float(r["Counter_Value"]))
    return res


def measure_traffic(args):
    """HBM bytes per launch of the headline contraction: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch, one rocprofv3 --pmc pass each), each scaled
    by what the same counter reads for a plain 1 GiB device copy."""
    import tempfile
    work = tempfile.mkdtemp(prefix="mdsctk_pmc_", dir="/tmp")
    probe = ["--copy-probe", str(COPY_PROBE_BYTES)]
    child = ["--traffic-child", "--n-frames", str(args.n_frames), "--natoms", str(args.natoms),
             "--k", str(args.k), "--precision", args.precision, "--seed", str(args.seed),
             "--self-symmetric", str(args.self_symmetric)]
    r = {}
    for tag, counter, a in (("copy_fetch", "FETCH_SIZE", probe), ("copy_write", "WRITE_SIZE", probe),
                            ("fetch", "FETCH_SIZE", child), ("write", "WRITE_SIZE", child)):
        log(f"traffic pass {tag} ...")
        r[tag] = _pmc_child(counter, a, tag, work)
        if r[tag] is None:
            log(f"traffic pass {tag} failed; traffic not reported")
            return None, None
    copy_f = [v for n, v in r["copy_fetch"].values() if "copyBuffer" in n]
    copy_w = [v for n, v in r["copy_write"].values() if "copyBuffer" in n]
    con_f = [v for n, v in r["fetch"].values() if "contract_kernel" in n]
    con_w = [v for n, v in r["write"].values() if "contract_kernel" in n]
    if not (copy_f and copy_w and con_f and con_w):
        return None, None
    f_scale = COPY_PROBE_BYTES / (max(copy_f) * 1024.0)
    w_scale = COPY_PROBE_BYTES / (max(copy_w) * 1024.0)
    fetch_b = max(con_f) * 1024.0 * f_scale
    write_b = max(con_w) * 1024.0 * w_scale
    info = {"fetch_bytes": fetch_b, "write_bytes": write_b,
            "fetch_scale": f_scale, "write_scale": w_scale,
            "method": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, KiB per dispatch, "
                      "scaled by a 1 GiB hipMemcpy device-to-device copy"}
    # the same passes' merge_kernel dispatch (the query's second kernel)
    mer_f = [v for n, v in r["fetch"].values() if "merge_kernel" in n]
    mer_w = [v for n, v in r["write"].values() if "merge_kernel" in n]
    if mer_f and mer_w:
        info["merge_kernel"] = {"fetch_bytes": max(mer_f) * 1024.0 * f_scale,
                                "write_bytes": max(mer_w) * 1024.0 * w_scale}
    return fetch_b + write_b, info


def traffic_child(args):
    """One headline query, nothing else: the launch the PMC passes read."""
    from mdsctk_amd import knn, synth
    hip = Hip()
    frames = synth.metastable_trajectory(args.n_frames, args.natoms, args.seed)
    w = synth.masses(args.natoms, "uniform")
    with knn.KnnContext(0, args.precision) as ctx:
        ctx.set_option(knn.OPT_SELF_SYMMETRIC, args.self_symmetric)
        ctx.set_reference(frames, w)
        time_fused(ctx, hip, args.n_frames, args.k, 1, 0)
    print(json.dumps({"traffic_child": True}), flush=True)


def mean_stat(stats, key):
    return float(np.mean([s[key] for s in stats]))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-frames", type=int, default=200_000)
    ap.add_argument("--natoms", type=int, default=1000)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp16", "fp32"])
    ap.add_argument("--fp32-steps", type=int, default=1, help="steps of the FP32 comparator (0 = skip)")
    ap.add_argument("--self-symmetric", type=int, default=1,
                    help="score each unordered pair once, the upper block triangle with mirrored "
                         "lists (1), or every ordered pair through the candidate lists (0)")
    ap.add_argument("--list-steps", type=int, default=1,
                    help="steps of the candidate-list path, reported beside a symmetric headline "
                         "(0 = skip)")
    ap.add_argument("--fp16-steps", type=int, default=1,
                    help="steps of the split-FP16 form, reported beside the headline (0 = skip)")
    ap.add_argument("--serving-seconds", type=float, default=60.0,
                    help="length of the Poisson serving run (0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=200.0, help="mean request arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=2, help="batches in flight when serving")
    ap.add_argument("--serve-max-frames", type=int, default=0,
                    help="query frames per served batch (0 = the library default)")
    ap.add_argument("--serve-share-cus", type=int, default=1,
                    help="each served batch's contraction takes 1/slots of the CUs (1) or all (0)")
    ap.add_argument("--serve-graphs", type=int, default=1,
                    help="replay each served batch's kernels as a hipGraph (1) or launch them (0)")
    ap.add_argument("--mixed-seconds", type=float, default=20.0,
                    help="run the three-set mixed serving workload and its graph check for this "
                         "many seconds (0 = skip; the CPU side of the graph check, the FP64 "
                         "oracle kNN of 30,000 frames, takes about 40 s on 16 threads)")
    ap.add_argument("--mixed-rate", type=float, default=100.0, help="mixed workload arrivals per second")
    ap.add_argument("--cli", type=int, default=0,
                    help="also time the drop-in knn_rms executable end to end on the workload "
                         "(writes a 1.6 GB XTC to /tmp first; 0 = skip)")
    ap.add_argument("--c4-batches", type=int, default=0, metavar="N",
                    help="only run the serving form's batch sizes (1, 1000, 4096 frames against "
                         "the 100k x 500 set) N times each as synchronous queries, print their "
                         "stage times and exit (for a kernel trace of the served batch kernels)")
    ap.add_argument("--cpu-rows", type=int, default=128, help="fit rows of the CPU baseline sample (0 = skip)")
    ap.add_argument("--check-rows", type=int, default=8, help="rows checked against the FP64 oracle")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--traffic", type=int, default=1,
                    help="measure HBM traffic of the contraction with rocprofv3 PMC passes (0 = skip)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device (HIP blit kernel) 3 times and exit: the "
                         "plain copy of known size that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe > 0:
        hip = Hip()
        a, b = hip.malloc(args.copy_probe), hip.malloc(args.copy_probe)
        for _ in range(3):
            hip.copy_d2d(b, a, args.copy_probe)
        hip.sync()
        hip.free(a)
        hip.free(b)
        print(json.dumps({"copy_probe_bytes": args.copy_probe}), flush=True)
        return

    if args.traffic_child:
        traffic_child(args)
        return

    if args.c4_batches > 0:
        from mdsctk_amd import knn, synth
        print(json.dumps({"c4_batches": run_c4_batches(knn, synth, args.precision, args)}),
              flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the knn_rms engine runs on a single "
              "MI355X; query rows shard across GPUs without a collective). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    # PMC passes first, in child processes, before this process touches the GPU.
    traffic, traffic_info = (None, None)
    if args.traffic:
        traffic, traffic_info = measure_traffic(args)

    from mdsctk_amd import knn, synth
    import oracle

    hip = Hip()
    n, natoms, k, prec = args.n_frames, args.natoms, args.k, args.precision
    t = time.time()
    frames = synth.metastable_trajectory(n, natoms, args.seed)
    w = synth.masses(natoms, "uniform")
    log(f"generated {n} x {natoms} frames in {time.time() - t:.1f} s")

    out = {"metric": METRIC, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": prec, "data": "synthetic"}
    out["config"] = {"workload": f"self kNN {n} x {n} frames x {natoms} atoms, k={k}, "
                                 f"metastable generator (50 centres, 5% near-duplicates)",
                     "n_ref": n, "n_fit": n, "natoms": natoms, "k": k,
                     "precision": {"bf16": "split-BF16 hi/lo, 3 MFMA products, FP32 accumulate",
                                   "fp16": "split-FP16 hi/lo, 3 MFMA products, FP32 accumulate",
                                   "fp32": "FP32 MFMA"}[prec],
                     "epilogue": "FP64 G terms + FP64 QCP Newton, FP64 re-score of the candidate window"}

    pairs = float(n) * n
    first = {}
    with knn.KnnContext(0, prec) as ctx:
        ctx.set_option(knn.OPT_SELF_SYMMETRIC, args.self_symmetric)
        t = time.time()
        ctx.set_reference(frames, w)
        log(f"reference resident in HBM ({time.time() - t:.1f} s)")
        wall, stats, dist, idx = time_fused(ctx, hip, n, k, args.steps, args.warmup, first)
    ms_contract = mean_stat(stats, "ms_contract")
    computed = mean_stat(stats, "n_pairs_computed")
    out["value"] = pairs / wall
    out["ms_per_step"] = wall * 1e3
    out["config"]["self_symmetric"] = bool(computed < pairs)
    out["config"]["pairs_per_step"] = pairs
    out["config"]["pairs_scored_per_step"] = computed
    if "ms" in first:
        out["first_query_ms"] = first["ms"]
    # algorithmic MFMA work of the pairs the contraction actually scored
    flops = FLOPS_PER_ATOM[prec] * natoms * computed
    achieved = flops / (ms_contract * 1e-3) / 1e12
    out["roofline"] = {"bound": "mfma", "kernel": "contract_kernel",
                       "achieved": achieved, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                       "frac": achieved / PEAK_TFLOPS[prec], "traffic": traffic,
                       "flops_per_launch": flops, "ms_per_launch": ms_contract,
                       "flops_per_pair": FLOPS_PER_ATOM[prec] * natoms}
    if traffic_info:
        out["roofline"]["traffic_detail"] = traffic_info
    out["stages_ms"] = {s: mean_stat(stats, "ms_" + s) for s in ("prep", "contract", "merge", "fallback", "total")}
    out["candidates_per_query"] = mean_stat(stats, "n_candidates") / n
    # the merge: HBM-bound, each re-scored candidate reads its reference
    # frame (12 N bytes of FP32 coordinates; the guard keys above the cut,
    # up to 32 per row, are re-scored too and not counted here)
    ms_merge = mean_stat(stats, "ms_merge")
    mbytes = mean_stat(stats, "n_candidates") * 12.0 * natoms
    m_traffic = None
    if traffic_info and "merge_kernel" in traffic_info:
        m_traffic = traffic_info["merge_kernel"]["fetch_bytes"] + traffic_info["merge_kernel"]["write_bytes"]
    out["merge_roofline"] = {"bound": "hbm", "kernel": "merge_kernel",
                             "achieved": mbytes / (ms_merge * 1e-3) / 1e9, "peak": 8000.0,
                             "unit": "GB/s", "frac": mbytes / (ms_merge * 1e-3) / 8e12,
                             "traffic": m_traffic, "bytes_per_launch": mbytes,
                             "ms_per_launch": ms_merge}
    out["fallback_rows"] = int(max(s["n_fallback_rows"] for s in stats))
    if computed < pairs:
        # the symmetric self query's landmark pass and mirrored lists
        out["symmetric"] = {"landmark_ms": mean_stat(stats, "ms_landmark"),
                            "landmark_pairs": mean_stat(stats, "n_pairs_landmark"),
                            "mirrored_per_row": mean_stat(stats, "n_mirrored") / n,
                            "mirrored_max_row": int(max(s["n_mirrored_max"] for s in stats)),
                            "relisted_rows": int(max(s["n_relisted_rows"] for s in stats))}

    # accuracy vs the FP64 oracle: 10k output pairs and a few full rows
    rng = np.random.default_rng(args.seed)
    rows = rng.integers(0, n, 10000)
    cols = rng.integers(0, k, 10000)
    ref_idx = idx[rows, cols].astype(np.int64)
    exact = oracle.pairs_nm(frames, frames, w, ref_idx, rows.astype(np.int64))
    err_nm = float(np.abs(dist[rows, cols] / 10.0 - exact).max())
    chk = rng.choice(n, size=args.check_rows, replace=False) if args.check_rows else np.zeros(0, int)
    match = 1.0
    if len(chk):
        t = time.time()
        d_o, i_o = oracle.knn_exact(frames, frames[chk], w, k)
        dd = np.abs(dist[chk] - d_o).max()
        sets = np.mean([len(set(idx[r]) & set(i_o[j])) / k for j, r in enumerate(chk)])
        match = float(sets)
        err_nm = max(err_nm, float(dd) / 10.0)
        log(f"oracle rows checked in {time.time() - t:.1f} s")
    out["accuracy"] = {"max_err_nm": err_nm, "tolerance_nm": 1e-5, "pairs_checked": 10000,
                       "rows_checked": int(len(chk)), "index_set_match": match}

    if args.list_steps > 0 and computed < pairs:
        # the same query through the candidate-list path (every pair scored)
        with knn.KnnContext(0, prec) as ctx:
            ctx.set_option(knn.OPT_SELF_SYMMETRIC, 0)
            ctx.set_reference(frames, w)
            wall_l, st_l, d_l, i_l = time_fused(ctx, hip, n, k, args.list_steps, 1)
        ms_l = mean_stat(st_l, "ms_contract")
        a_l = FLOPS_PER_ATOM[prec] * natoms * pairs / (ms_l * 1e-3) / 1e12
        out["list_path"] = {"value": pairs / wall_l, "unit": "pairs/s", "ms_per_step": wall_l * 1e3,
                            "contract_ms": ms_l, "merge_ms": mean_stat(st_l, "ms_merge"),
                            "achieved_tflops": a_l, "frac_peak": a_l / PEAK_TFLOPS[prec],
                            "outputs_equal_headline": bool((d_l == dist).all() and (i_l == idx).all())}

    if args.fp32_steps > 0 and prec != "fp32":
        with knn.KnnContext(0, "fp32") as ctx:
            ctx.set_option(knn.OPT_SELF_SYMMETRIC, args.self_symmetric)
            ctx.set_reference(frames, w)
            wall32, st32, _, _ = time_fused(ctx, hip, n, k, args.fp32_steps, 1)
        ms32 = mean_stat(st32, "ms_contract")
        a32 = 18 * natoms * mean_stat(st32, "n_pairs_computed") / (ms32 * 1e-3) / 1e12
        out["fp32_path"] = {"value": pairs / wall32, "unit": "pairs/s", "ms_per_step": wall32 * 1e3,
                            "contract_ms": ms32, "achieved_tflops": a32, "frac_fp32_peak": a32 / 157.3}
        out["speedup_vs_fp32"] = out["value"] / out["fp32_path"]["value"]

    if args.fp16_steps > 0 and prec != "fp16":
        # the other split form the precision form allows: same MFMA rate, an
        # 11-bit hi part, so a narrower candidate window
        with knn.KnnContext(0, "fp16") as ctx:
            ctx.set_option(knn.OPT_SELF_SYMMETRIC, args.self_symmetric)
            ctx.set_reference(frames, w)
            wall16, st16, d16, i16 = time_fused(ctx, hip, n, k, args.fp16_steps, 1)
        ms16 = mean_stat(st16, "ms_contract")
        a16 = 54 * natoms * mean_stat(st16, "n_pairs_computed") / (ms16 * 1e-3) / 1e12
        out["fp16_path"] = {"value": pairs / wall16, "unit": "pairs/s", "ms_per_step": wall16 * 1e3,
                            "contract_ms": ms16, "merge_ms": mean_stat(st16, "ms_merge"),
                            "achieved_tflops": a16, "frac_fp16_peak": a16 / 2500.0,
                            "candidates_per_query": mean_stat(st16, "n_candidates") / n,
                            "outputs_equal_headline": bool((d16 == dist).all() and (i16 == idx).all())}

    if args.serving_seconds > 0:
        out["serving"] = run_serving(knn, synth, oracle, prec, args)

    if args.mixed_seconds > 0:
        out["mixed_sets"] = run_mixed_sets(knn, synth, oracle, args)

    if args.cli:
        out["cli_end_to_end"] = run_cli_e2e(knn, synth, args)

    if args.cpu_rows > 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t = time.time()
        cref = oracle.port_center(frames, w)
        log(f"CPU centring {time.time() - t:.1f} s")
        t = time.perf_counter()
        oracle.port_knn_rows(cref, cref, w, k, 0, args.cpu_rows, nthreads=threads)
        cpu_s = time.perf_counter() - t
        cpu_pairs = args.cpu_rows * float(n)
        out["cpu_baseline"] = {"value": cpu_pairs / cpu_s, "unit": "pairs/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {args.cpu_rows} fit rows x all {n} refs "
                                         f"({cpu_pairs:.3g} pairs, {cpu_s:.1f} s), reference-"
                                         "equivalent float knn_rms (OpenMP over fit rows)"}
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
