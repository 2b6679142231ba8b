#!/usr/bin/env python3
"""Benchmark of the MI355X tree-ensemble build + predict path.

Workload (one GPU): random-forest classifier on 1,000,000 rows x 128 fp32
features, N(0,1) with 16 informative features, 10 classes from the argmax of
a random linear map plus Gumbel noise; 200 bootstrap trees, entropy
criterion, 11 features per node, max depth 20, min node size 5.

A step is one full fit of the forest from samples already resident in HBM
(dense-rank preprocessing, bootstrap, every level of every tree) with the
batched level-wise build.  `value` is trees/s over the timed steps.  Beside
it:
  - the plain per-node build (the reference's launch structure) timed on
    the same input, and the batched/plain speed-up (the forests must be
    identical);
  - predict rows/s of the fitted forest on the 1M training rows;
  - `roofline`: algorithmic bytes of the dominant split-search kernel over
    its average launch time, from HIP events on the library stream;
    `traffic` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of the same kernel (profiles/traffic_rf_1m128.json) when present;
  - `cpu_baseline`: the CPU oracle on the host cores, one tree per core, on
    the same data (a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RF build trees/s at 1M×128 + predict rows/s; p99 latency at fixed QPS"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_rf_1m128.json")
# family -> kernel symbol prefix as rocprofv3 names it (10 classes -> MAXC 16)
KERNEL_SYMBOL = {
    "search_seg16": "dte::(anonymous namespace)::k_search_rf_seg<16, 0, 16>",
    "search_seg32": "dte::(anonymous namespace)::k_search_rf_seg<16, 0, 32>",
    "search_seg64": "dte::(anonymous namespace)::k_search_rf_seg<16, 0, 64>",
    "search_team64": "dte::(anonymous namespace)::k_search_rf<16, 0, 64, false>",
    "search_block": "dte::(anonymous namespace)::k_search_rf<16, 0, 256, false>",
    "chunk_eval": "dte::(anonymous namespace)::k_v2_eval<16, 0>",
}


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(X, y, cores):
    """The oracle fits `cores` trees of the same workload on `cores` threads
    (trees are independent, so trees/s scales to the full forest)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    t0 = time.perf_counter()
    f = pyoracle.fit(X, y, algorithm=0, n_trees=cores, k=11, max_depth=20, min_node_size=5, n_classes=10,
                     seed=324123, n_threads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(cores / dt, 4), "unit": "trees/s", "cores": cores, "kind": "port",
            "sample": f"{cores} of the 200 trees (one per thread) of the same 1M x 128 RF fit, same seed; "
                      f"{dt:.1f} s wall, {f.n_nodes} nodes; CPU: {cpu_model()}"}


def fit_steps(capi, Xd, n_rows, n_features, y, params, steps, warmup, label):
    for i in range(warmup):
        t = time.perf_counter()
        f = capi.fit_device(Xd.ptr, n_rows, n_features, y, params)
        log(f"{label} warmup {i}: {time.perf_counter() - t:.3f} s, {f.n_nodes} nodes")
        f.close()
    capi.profile_reset()
    capi.profiling(True)
    times, forest = [], None
    for i in range(steps):
        if forest is not None:
            forest.close()
        t = time.perf_counter()
        forest = capi.fit_device(Xd.ptr, n_rows, n_features, y, params)
        times.append(time.perf_counter() - t)
        log(f"{label} step {i}: {times[-1]:.3f} s")
    capi.profiling(False)
    return times, capi.profile_stats(), forest


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain copy kernel of known size, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")  # fill kernel: writes nbytes
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)  # elementwise copy kernel: reads nbytes, writes nbytes
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} x {nbytes} bytes")


def serve_bench(capi, synth, np, n_requests, seed=0):
    """Four resident forests of 500 trees (depth <= 20, 256 features, 10
    classes, trained on a 200,000-row set of the same generator), requests
    of log-uniform 1..65,536 rows of N(0,1) features, Poisson arrivals.
    Capacity is measured first (all requests submitted at once); then two
    fixed QPS points at 30% and 70% of it."""
    import threading
    t0 = time.perf_counter()
    X, y = synth.gaussian_classes(seed=5, n=200_000, f=256, n_informative=16, n_classes=10)
    Xd = capi.DeviceArray.from_numpy(X)
    forests = []
    for i in range(4):
        p = capi.default_params(capi.RANDOM_FOREST, n_trees=500, max_depth=20, min_node_size=5, n_classes=10,
                                seed=1000 + i)
        forests.append(capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p))
    Xd.free()
    log(f"serving: 4 x 500-tree forests trained in {time.perf_counter() - t0:.1f} s")
    srv = capi.Server(hbm_budget_bytes=0, max_batch_rows=65536)
    slots = [srv.load(f) for f in forests]
    rng = np.random.default_rng(seed)
    pool = rng.standard_normal((1 << 17, 256), dtype=np.float32)
    sizes = synth.request_sizes(seed=seed + 1, n_requests=n_requests, max_rows=65536)
    offs = rng.integers(0, pool.shape[0] - 65536, n_requests)
    which = rng.integers(0, 4, n_requests)
    reqs = [(int(which[i]), pool[offs[i]:offs[i] + sizes[i]]) for i in range(n_requests)]
    total_rows = int(sizes.sum())

    def run(arrivals):
        done_t = [0.0] * n_requests
        tickets = [None] * n_requests
        ready = threading.Semaphore(0)

        def waiter():
            for i in range(n_requests):
                ready.acquire()
                srv.wait(tickets[i])
                done_t[i] = time.perf_counter()

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        arr_abs = [start + a for a in arrivals]
        for i, (slot, rows) in enumerate(reqs):
            dt = arr_abs[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[i] = srv.submit(slots[slot], rows)
            ready.release()
        th.join()
        lat = np.array([done_t[i] - arr_abs[i] for i in range(n_requests)]) * 1e3
        span = max(done_t) - start
        return lat, span

    lat, span = run([0.0] * n_requests)
    cap_rows = total_rows / span
    cap_qps = n_requests / span
    points = []
    for frac in (0.3, 0.7):
        qps = cap_qps * frac
        arrivals = np.cumsum(rng.exponential(1.0 / qps, n_requests)).tolist()
        lat, span = run(arrivals)
        points.append({"qps": round(qps, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3), "rows_per_s": round(total_rows / span, 1)})
        log(f"serving at {qps:.0f} QPS: p50 {points[-1]['p50_ms']} ms, p99 {points[-1]['p99_ms']} ms")
    st = srv.stats()
    srv.close()
    nodes = sum(f.n_nodes for f in forests)
    for f in forests:
        f.close()
    return {"forests": "4 x 500 trees, 256 features, 10 classes, depth <= 20, trained on 200,000 rows",
            "requests_per_point": n_requests, "rows_per_request": "log-uniform 1..65536",
            "capacity_rows_per_s": round(cap_rows, 1), "capacity_qps": round(cap_qps, 1), "points": points,
            "batches": st["batches_launched"], "traversal_launches": st["groups_launched"],
            "batch_time_ms": {k: round(st[k], 1) for k in ("stage_ms", "h2d_ms", "kernel_ms", "d2h_ms", "scatter_ms")},
            "resident_nodes": nodes, "hbm_in_use_bytes": st["hbm_in_use"]}


def rate(n, dt, digits):
    """n / dt rounded, or None when the leg did not run."""
    return round(n / dt, digits) if dt else None


def other_workload(name, args):
    """Measurement lines of the other workload shapes (not the headline).
    --gpu-path-only keeps the timed GPU path alone (no plain driver, CPU
    oracle, autotuner or per-forest comparisons), for profiling runs."""
    import ctypes

    import numpy as np
    from datamininglibs_amd import capi, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = args.cpu_cores or min(16, os.cpu_count() or 1)

    def timed_fit(X, y, p, reps):
        f = capi.fit(X, y, p)  # warm-up (host upload included in every call)
        f.close()
        Xd = capi.DeviceArray.from_numpy(X)
        t = time.perf_counter()
        for _ in range(reps):
            f = capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p)
            nodes = f.n_nodes
            f.close()
        dt = (time.perf_counter() - t) / reps
        Xd.free()
        log(f"{name}: {p.n_trees} trees ({'plain' if p.build_mode else 'batched'}) in {dt:.3f} s per fit")
        return dt, nodes

    if name == "c1":
        X, y = synth.spam_like(seed=0)
        kw = dict(n_trees=100, n_features_per_node=7, max_depth=25, min_node_size=10, n_classes=2, seed=0)
        dt, nodes = timed_fit(X, y, capi.default_params(capi.EXTRA_TREES, **kw), 5)
        dtp = dto = None
        if not args.gpu_path_only:
            dtp, _ = timed_fit(X, y, capi.default_params(capi.EXTRA_TREES, build_mode=capi.BUILD_PLAIN, **kw), 2)
            t = time.perf_counter()
            pyoracle.fit(X, y, algorithm=1, n_trees=100, k=7, max_depth=25, min_node_size=10, n_classes=2, seed=0,
                         n_threads=cores)
            dto = time.perf_counter() - t
        line = {"workload": "extra-trees classifier 4,601 x 57 (spam-like), 100 trees, k 7, depth 25, min node 10",
                "gpu_trees_per_s": round(100 / dt, 1), "gpu_plain_trees_per_s": rate(100, dtp, 1),
                "cpu_trees_per_s": rate(100, dto, 1), "cpu_cores": cores, "nodes": nodes}
    elif name == "c3":
        X, y = synth.sine_regression(seed=0, n=4_000_000, f=64)
        kw = dict(task=capi.REGRESSION, n_trees=100, n_features_per_node=21, max_depth=18, min_node_size=5,
                  n_classes=0, seed=7)
        capi.profile_reset()
        capi.profiling(True)
        dt, nodes = timed_fit(X, y, capi.default_params(capi.RANDOM_FOREST, **kw), 2)
        capi.profiling(False)
        st = capi.profile_stats()
        dtp = dto = None
        if not args.gpu_path_only:
            kw10 = dict(kw, n_trees=10)  # the plain driver on a 10-tree slice
            dtp, _ = timed_fit(X, y, capi.default_params(capi.RANDOM_FOREST, build_mode=capi.BUILD_PLAIN, **kw10), 1)
            dtp *= 10.0
            log(f"c3: cpu oracle, {cores} trees on {cores} threads")
            t = time.perf_counter()
            pyoracle.fit(X, y, algorithm=0, task=1, n_trees=cores, k=21, max_depth=18, min_node_size=5, seed=7,
                         n_threads=cores)
            dto = time.perf_counter() - t
        ss = st.get("split_search", {"total_ms": 0, "bytes": 0})
        lg = st.get("level_graph", {"total_ms": 0, "bytes": 0})
        ss = {"total_ms": ss["total_ms"] + lg["total_ms"], "bytes": ss["bytes"] + lg["bytes"]}
        line = {"workload": "random-forest regression 4,000,000 x 64, 100 trees, variance, k 21, depth 18",
                "gpu_trees_per_s": round(100 / dt, 2), "gpu_plain_trees_per_s": rate(100, dtp, 2),
                "plain_sample": "10 of the 100 trees",
                "cpu_trees_per_s": rate(cores, dto, 4), "cpu_cores": cores, "nodes": nodes,
                "split_search_GBps": round(ss["bytes"] / (ss["total_ms"] * 1e-3) / 1e9, 1) if ss["total_ms"] else None,
                "kernel_ms_per_fit": {k: round(v["total_ms"] / 3, 2) for k, v in st.items()}}
    else:  # c4
        specs = synth.uneven_forests(seed=0)
        data = [synth.uneven_forest_data(sp) for sp in specs]
        params = [capi.default_params(capi.EXTRA_TREES, n_trees=50, n_classes=sp["classes"], seed=sp["seed"])
                  for sp in specs]
        devs = [capi.DeviceArray.from_numpy(X) for X, _ in data]
        # one autotuner run on the largest forest; its extra-trees tier cut
        # serves every forest
        tuned, tune_s = None, None
        if not args.gpu_path_only:
            big = max(range(len(specs)), key=lambda i: specs[i]["n"] * specs[i]["f"])
            tuned, tune_s = capi.autotune(devs[big].ptr, data[big][0].shape[1], data[big][0].shape[0],
                                          data[big][1], params[big], trial_trees=10)
            for p in params:
                p.cut_ert = tuned.cut_ert
        dt_seq = None
        nodes = None
        if not args.gpu_path_only:
            for (X, y), Xd, p in zip(data, devs, params):
                capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p).close()
            t = time.perf_counter()
            nodes = 0
            for (X, y), Xd, p in zip(data, devs, params):
                f = capi.fit_device(Xd.ptr, X.shape[1], X.shape[0], y, p)
                nodes += f.n_nodes
                f.close()
            dt_seq = time.perf_counter() - t
        # all 64 forests in one joint build (dte_fit_many: every level step
        # serves the open nodes of every forest)
        args_j = ([Xd.ptr for Xd in devs], [X.shape[1] for X, _ in data], [X.shape[0] for X, _ in data],
                  [y for _, y in data], params)
        for f in capi.fit_many(*args_j):
            f.close()
        capi.profile_reset()
        capi.profiling(True)
        t = time.perf_counter()
        forests = capi.fit_many(*args_j)
        dt = time.perf_counter() - t
        capi.profiling(False)
        jprof = capi.profile_stats()
        replays = jprof.get("level_graph_replays", {"launches": 0})["launches"]
        joint_nodes = sum(f.n_nodes for f in forests)
        assert nodes is None or joint_nodes == nodes
        nodes = joint_nodes
        for f in forests:
            f.close()
        dt_streams = dto = None
        ntr = 0
        if not args.gpu_path_only:
            # the same 64 fits as separate fits on 8 concurrent streams, each
            # sizing its levels on the host (no joint build, no level graphs)
            hp = []
            for p in params:
                q = capi.default_params(capi.EXTRA_TREES)
                ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
                q.build_mode = capi.BUILD_BATCHED_HOST
                hp.append(q)
            hargs = args_j[:4] + (hp,)
            for f in capi.fit_many(*hargs, max_concurrent=8):
                f.close()
            t = time.perf_counter()
            hforests = capi.fit_many(*hargs, max_concurrent=8)
            dt_streams = time.perf_counter() - t
            assert sum(f.n_nodes for f in hforests) == nodes
            for f in hforests:
                f.close()
            small = sorted(range(64), key=lambda i: specs[i]["n"] * specs[i]["f"])[:8]
            t = time.perf_counter()
            for i in small:
                X, y = data[i]
                pyoracle.fit(X, y, algorithm=1, n_trees=50, n_classes=specs[i]["classes"], seed=specs[i]["seed"],
                             n_threads=cores)
            dto = time.perf_counter() - t
            ntr = sum(specs[i]["n"] for i in small)
        line = {"workload": "64 extra-trees forests, 2k..200k rows, 16..256 features, 2..8 classes, 50 deep trees each",
                "gpu_trees_per_s": round(64 * 50 / dt, 1), "gpu_seconds_all": round(dt, 2), "nodes": nodes,
                "gpu_path": "one joint build of the 64 forests (dte_fit_many), every level step one hipGraph replay",
                "level_graph_replays": replays,
                "joint_ms": {k: round(v["total_ms"], 2) for k, v in jprof.items() if v["total_ms"] > 0},
                "gpu_trees_per_s_8_streams_host_levels": rate(64 * 50, dt_streams, 1),
                "tuning": {"cut_ert": tuned.cut_ert if tuned else "library default",
                           "autotune_s": round(tune_s, 2) if tune_s is not None else None},
                "gpu_trees_per_s_one_after_another": rate(64 * 50, dt_seq, 1),
                "cpu_sample": f"8 smallest forests ({ntr} rows total), {dto:.1f} s on {cores} threads" if dto else None,
                "cpu_trees_per_s_sample": rate(8 * 50, dto, 1)}
    print(json.dumps({"workload_name": name, **line}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trees", type=int, default=200)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--features", type=int, default=128)
    ap.add_argument("--plain-steps", type=int, default=1, help="timed fits of the plain per-node path (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true", help="also run a 1 GiB copy kernel")
    ap.add_argument("--serve-requests", type=int, default=1000, help="requests per serving QPS point (0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="default tier cuts instead of dte_autotune's")
    ap.add_argument("--tuning", default="", help="explicit cut_team,cut_block,bucket_target (no autotuner run)")
    ap.add_argument("--gpu-path-only", action="store_true",
                    help="c1/c3/c4: only the timed GPU path (for profiling runs)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4"],
                    help="c2 (default): the headline random-forest line; c1/c3/c4: the other workload shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: this build measures the single-GPU share of the workload; "
                                   "multi-GPU tree sharding is a design note, not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload != "c2":
        other_workload(args.workload, args)
        return

    import ctypes

    import numpy as np

    from datamininglibs_amd import capi, synth

    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    X, y = synth.gaussian_classes(seed=0, n=args.rows, f=args.features, n_informative=16, n_classes=10)
    log(f"data generated in {time.perf_counter() - t0:.1f} s")
    Xd = capi.DeviceArray.from_numpy(X)
    common = dict(n_trees=args.trees, n_features_per_node=11, max_depth=20, min_node_size=5, n_classes=10,
                  seed=324123)

    params = capi.default_params(capi.RANDOM_FOREST, **common)
    tuning = None
    if args.tuning:
        params.cut_team, params.cut_block, params.bucket_target = (int(v) for v in args.tuning.split(","))
        tuning = {"cut_team": params.cut_team, "cut_block": params.cut_block,
                  "bucket_target": params.bucket_target, "autotune_s": None}
    elif not args.no_autotune:
        # split-search tier cuts and bucket size for this data set, from
        # trial fits of 16 trees (outside the timed steps)
        params, tune_s = capi.autotune(Xd.ptr, args.rows, args.features, y, params, trial_trees=16)
        tuning = {"cut_team": params.cut_team, "cut_block": params.cut_block,
                  "bucket_target": params.bucket_target, "autotune_s": round(tune_s, 2)}
        log(f"autotune: {tuning}")
    times, stats, forest = fit_steps(capi, Xd, args.rows, args.features, y, params, args.steps, args.warmup,
                                     "batched")
    total = sum(times)
    trees_per_s = args.trees * args.steps / total

    plain = None
    if args.plain_steps > 0:
        ptimes, _, pforest = fit_steps(capi, Xd, args.rows, args.features, y,
                                       capi.default_params(capi.RANDOM_FOREST, build_mode=capi.BUILD_PLAIN,
                                                           **common), args.plain_steps, 0, "plain")
        a, b = forest.canonical(), pforest.canonical()
        identical = all(np.array_equal(a[k], b[k]) for k in a)
        pforest.close()
        p_tps = args.trees * args.plain_steps / sum(ptimes)
        plain = {"trees_per_s": round(p_tps, 3), "ms_per_step": round(1000 * sum(ptimes) / args.plain_steps, 1),
                 "speedup_batched_vs_plain": round(trees_per_s / p_tps, 2), "identical_forest": identical}
        if not identical:
            raise SystemExit("plain and batched forests differ")

    # Predict throughput on the training rows (feature-major, resident).
    lib = capi.lib()
    n_pred = args.rows
    votes = capi.DeviceArray(4 * n_pred * 10)

    def predict_once():
        capi._check(lib.dte_predict_device(forest.handle, ctypes.c_void_p(Xd.ptr), n_pred, args.rows, 0,
                                           ctypes.c_void_p(votes.ptr), None, None))

    predict_once()
    capi.profile_reset()
    capi.profiling(True)
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        predict_once()
    pred_dt = (time.perf_counter() - t) / reps
    capi.profiling(False)
    pstats = capi.profile_stats().get("predict", {"total_ms": 0.0, "launches": 1})
    v = votes.to_numpy(np.int32, (n_pred, 10))
    assert (v.sum(1) == args.trees).all()
    acc = float((v.argmax(1) == y.astype(np.int64)).mean())
    pred_kernel_ms = pstats["total_ms"] / max(1, pstats["launches"])
    pred_bytes = n_pred * args.features * 4 + forest.n_nodes * 16 + n_pred * 10 * 4

    # Out-of-bag pass over the training rows: every (tree, row) pair left
    # out of the tree's bootstrap, plain and with one feature permuted.
    capi.oob_evaluate(forest, Xd.ptr, y, args.rows)
    t = time.perf_counter()
    oob_pairs, oob_ok = capi.oob_evaluate(forest, Xd.ptr, y, args.rows)
    oob_dt = time.perf_counter() - t
    t = time.perf_counter()
    _, perm_ok = capi.oob_evaluate(forest, Xd.ptr, y, args.rows, 0)
    perm_dt = time.perf_counter() - t
    oob = {"pairs": int(oob_pairs), "accuracy": round(oob_ok / oob_pairs, 4), "ms": round(1000 * oob_dt, 1),
           "pairs_per_s": round(oob_pairs / oob_dt, 1), "permuted_pass_ms": round(1000 * perm_dt, 1),
           "importance_feature_0": round((oob_ok - perm_ok) / oob_pairs, 5)}

    # Roofline of the dominant split-search kernel.
    # single-kernel families that carry algorithmic bytes (the chunk
    # pipeline's evaluation kernel carries its items' bytes)
    fam = {k: v_ for k, v_ in stats.items()
           if (k.startswith("search_") and k != "search_chunked") or k == "chunk_eval"}
    dom = max(fam, key=lambda k: fam[k]["total_ms"])
    ds = fam[dom]
    avg_ms = ds["total_ms"] / max(1, ds["launches"])
    bytes_per_launch = ds["bytes"] / max(1, ds["launches"])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    traffic_src = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            tj = json.load(fh)
        if tj.get("family") == dom and tj.get("rows") == args.rows and tj.get("trees_per_fit") == args.trees:
            traffic = tj["hbm_bytes_per_dispatch"]
            traffic_src = os.path.relpath(TRAFFIC_FILE, ROOT) + f" (measured at tuning {tj.get('tuning')})"
    ss = stats.get("split_search", {"total_ms": 0.0, "bytes": 0.0})
    # levels replayed from captured hipGraphs: their split-search bytes over
    # the whole level-step time (partition and bookkeeping included)
    lg = stats.get("level_graph", {"total_ms": 0.0, "bytes": 0.0, "launches": 0})
    lg_replays = stats.get("level_graph_replays", {"launches": 0})["launches"]
    ss_all_ms = ss["total_ms"] + lg["total_ms"]
    ss_all_bytes = ss["bytes"] + lg["bytes"]
    level_graphs = {"replays_per_fit": round(lg_replays / args.steps, 1),
                    "ms_per_fit": round(lg["total_ms"] / args.steps, 2),
                    "split_search_GB_per_fit": round(lg["bytes"] / args.steps / 1e9, 3)}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": KERNEL_SYMBOL.get(dom, dom), "family": dom, "launches": ds["launches"],
                "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": round(bytes_per_launch),
                "traffic_source": traffic_src,
                "split_search_all_tiers_GBps": round(ss_all_bytes / (ss_all_ms * 1e-3) / 1e9, 1)
                if ss_all_ms > 0 else None}
    serving = None
    if args.serve_requests > 0:
        serving = serve_bench(capi, synth, np, args.serve_requests)
    cpu = None
    if not args.no_cpu_baseline:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {cores} threads")
        cpu = cpu_baseline(X, y, cores)

    line = {
        "metric": METRIC,
        "value": round(trees_per_s, 3),
        "unit": "trees/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "random forest classifier, 1,000,000 x 128 fp32 N(0,1), 16 informative, 10 classes, "
                               "200 bootstrap trees, entropy, 11 features/node, max depth 20, min node 5",
                   "rows": args.rows, "features": args.features, "trees": args.trees, "parallelism": "1 GPU",
                   "tuning": tuning},
        "plain_path": plain,
        "predict": {"rows_per_s": round(n_pred / pred_dt, 1), "kernel_ms": round(pred_kernel_ms, 3),
                    "compulsory_GBps": round(pred_bytes / (pred_kernel_ms * 1e-3) / 1e9, 1) if pred_kernel_ms else None,
                    "train_accuracy": round(acc, 4)},
        "level_graphs": level_graphs,
        "oob": oob,
        "serving": serving,
        "kernel_ms_per_step": {k: round(v_["total_ms"] / args.steps, 3) for k, v_ in stats.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    forest.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
