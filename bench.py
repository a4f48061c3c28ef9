"""Benchmark of the distance-bound hot path on one MI355X.

Default workload: brute-force kNN, 4,000,000 x 128 f32 reference rows
(rows ~ N(0, I) scaled by per-row norms ~ LogNormal(0, 0.5)), 65,536 queries
of the same distribution, k = 10, Euclidean.  One step is one
sc_knn_search_device call over all 65,536 queries with the rows and queries
already resident in HBM.  `--workload kmeans` runs the k-means Lloyd loop on
1,000,000 x 64 f32 Gaussian blobs, k = 64, up to 100 iterations (one step =
one fit; the reported rate is Lloyd iterations per second).  `--workload
batch` runs 512 uneven k-means problems in one batched call, `--workload
serve` the continuous-batching server at fixed Poisson arrival rates,
`--workload dbscan` a DBSCAN fit over the device radius search, and the
default `all` reports them all beside the kNN line.

Prints one JSON line (rank 0) with the throughput, a roofline entry for the
dominant kernel (its duration measured with HIP events on the stream it
runs on, via the library's tracing hooks) and a single-core CPU baseline
from the oracle on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN queries/s (4M×128,k=10) & k-means iters/s vs FP64 form; serve p50/p99 ms"
PEAK_BF16 = 2.5e15      # dense BF16 MFMA, FLOP/s
PEAK_FP64 = 78.6e12     # FP64 vector, FLOP/s
PEAK_HBM = 8.0e12       # bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measured_traffic(kernel: str):
    """(HBM bytes per launch of `kernel`, the summary they come from): two
    rocprofv3 --pmc runs of the same workload (FETCH_SIZE, then WRITE_SIZE,
    scaled by what a plain copy of known size reads), filed in
    profiles/traffic.json; (None, None) when no measurement is on file."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None, None
    return (None, None) if not t else (t["bytes"], t.get("source"))


def profile_read(lib, name: str):
    ms = C.c_double()
    n = C.c_uint64()
    lib.sc_profile_read(name.encode(), C.byref(ms), C.byref(n))
    return ms.value, n.value


def bench_knn(args):
    import torch

    from smartcore_amd import _capi as S
    from smartcore_amd import synth

    lib = S.lib()
    lib.sc_profile_enable.argtypes = [C.c_int]
    lib.sc_profile_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    N, Q, D, K = args.n, args.q, 128, 10
    mode = S.SC_MIXED if args.mode == "mixed" else S.SC_PLAIN_FP64
    dev = torch.device("cuda:0")
    t0 = time.time()
    refs = synth.lognormal_rows_device(101, N, D, device=dev)
    qs = synth.lognormal_rows_device(202, Q, D, device=dev)
    torch.cuda.synchronize()
    log(f"[bench] generated refs {N}x{D} and queries {Q}x{D} on device in {time.time() - t0:.1f}s")
    ix = S.Index(device_ptr=refs.data_ptr(), n=N, d=D, dtype=np.float32)
    idx = torch.empty((Q, K), dtype=torch.int64, device=dev)
    dist = torch.empty((Q, K), dtype=torch.float64, device=dev)
    cnt = torch.empty((Q,), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ix.knn_device(qs.data_ptr(), Q, K, idx.data_ptr(), dist.data_ptr(), cnt.data_ptr(),
                      S.SC_EUCLIDIAN, 2, mode, stream)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    lib.sc_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step()
        torch.cuda.synchronize()
        log(f"[bench] step {s + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    lib.sc_profile_enable(0)
    qps = Q * args.steps / elapsed

    if mode == S.SC_MIXED:
        # the timed main scan is the last row range: on indexes of at least
        # 524,288 rows the mixed form scans the first 16,384 rows, then
        # ranges growing 8 times; the last one starts at the first
        # 16,384 * 8^j whose 8-fold reaches n (1,048,576 at 4M rows)
        kname = "knn_mixed_gemm"
        start = 0
        if N >= 8 * 65536:
            start = 16384
            while start * 8 < N:
                start *= 8
        flops_per_query = 2.0 * (N - start) * D
        e2e_flops_per_query = 2.0 * N * D
        peak, bound, unit_peak = PEAK_BF16, "mfma", "TFLOP/s"
    else:
        kname = "knn_scan"
        flops_per_query = 3.0 * N * D
        e2e_flops_per_query = flops_per_query
        peak, bound, unit_peak = PEAK_FP64, "fp64-valu", "TFLOP/s"
    kms, launches = profile_read(lib, kname)
    per_launch_q = Q * args.steps / max(launches, 1)
    avg_s = kms / 1e3 / max(launches, 1)
    achieved = flops_per_query * per_launch_q / avg_s if avg_s > 0 else 0.0
    others = {}
    for name in ("knn_scan", "knn_replay", "knn_serial", "knn_mixed_prefix", "knn_mixed_gemm",
                 "knn_mixed_rescore"):
        ms, n = profile_read(lib, name)
        if n:
            others[name] = {"avg_ms": ms / n, "launches": n}

    plain = None
    if mode == S.SC_MIXED and not args.no_compare:
        # the FP64 form on the same inputs: one untimed and one timed step
        step_plain = lambda: ix.knn_device(qs.data_ptr(), Q, K, idx.data_ptr(), dist.data_ptr(),
                                           cnt.data_ptr(), S.SC_EUCLIDIAN, 2, S.SC_PLAIN_FP64,
                                           stream)
        gi_mixed = idx.clone()
        gd_mixed = dist.clone()
        step_plain()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step_plain()
        torch.cuda.synchronize()
        pt = time.perf_counter() - t1
        mism = int((idx != gi_mixed).sum().item())
        dbits = int((dist.view(torch.int64) != gd_mixed.view(torch.int64)).sum().item())
        plain = {"fp64_form_queries_per_s": Q / pt, "speedup_vs_fp64_form": qps / (Q / pt),
                 "index_mismatches_vs_fp64_form": mism, "distance_bit_mismatches": dbits}
        log(f"[bench] FP64 form {Q / pt:.0f} q/s, mixed/FP64 = {qps / (Q / pt):.2f}, "
            f"mismatches {mism}")
        idx.copy_(gi_mixed)
    # parity spot check of this run's output against the oracle on a few queries
    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if mode == S.SC_MIXED else "f64",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"kNN {N:,}x{D} f32 rows, {Q:,} queries, k={K}, Euclidean",
                   "mode": args.mode},
        "roofline": {"bound": bound, "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": unit_peak, "frac": achieved / peak,
                     "traffic": measured_traffic(kname)[0],
                     "traffic_source": measured_traffic(kname)[1],
                     "kernel": kname, "avg_ms": avg_s * 1e3,
                     "work_per_launch": f"{flops_per_query:.4g} FLOP/query x {per_launch_q:.0f} queries"},
        # every kernel of a step (query preparation, all levels, rescoring,
        # replay) against the same peak: the cross-term FLOPs of the whole row
        # set per query times the step's query rate
        "roofline_end_to_end": {"bound": bound, "unit": unit_peak,
                                "achieved": e2e_flops_per_query * qps / 1e12, "peak": peak / 1e12,
                                "frac": e2e_flops_per_query * qps / peak,
                                "work": f"{e2e_flops_per_query:.4g} FLOP/query x {qps:.0f} queries/s"},
        "kernels": others,
    }
    if plain:
        out["vs_fp64_form"] = plain
    if not args.no_cpu:
        import oracle as O

        S_q = args.cpu_sample
        href = refs.cpu().numpy()
        hq = qs[:S_q].cpu().numpy()
        O.lib()
        t0 = time.perf_counter()
        oi, od, oc = O.linear_find_batch(href, hq, K)
        cpu_s = time.perf_counter() - t0
        gi = idx[:S_q].cpu().numpy()
        mismatch = int(np.sum(gi != oi))
        out["cpu_baseline"] = {"value": S_q / cpu_s, "unit": "queries/s", "cores": 1, "kind": "port",
                               "sample": f"first {S_q} of {Q:,} queries against all {N:,} rows "
                                         f"({cpu_s:.1f} s); host {cpu_model()}, nproc {os.cpu_count()}"}
        out["parity_sample"] = {"queries": S_q, "index_mismatches": mismatch}
    ix.close()
    return out


def bench_classifier(args):
    """KNNClassifier, k = 5, Euclidian, LinearSearch, 10,000 x 32 f64 rows from
    a 10-class Gaussian mixture, 2,000 queries from the same mixture: fit +
    predict through the C++ estimator mirror (one batched device search per
    predict).  The CPU baseline is the oracle's LinearKNNSearch::find per
    query plus the vote, on one core, over all 2,000 queries; the labels of
    both are compared."""
    import oracle as O
    from smartcore_amd import smartcore as sc
    from smartcore_amd import synth

    xtr, ytr, xq, _ = synth.knn_mixture(seed=1)
    p = sc.KNNClassifierParameters().with_k(5).with_algorithm(sc.KNNAlgorithmName.LinearSearch)
    for _ in range(max(args.warmup, 1)):
        sc.KNNClassifier.fit(xtr, ytr, p).predict(xq)
    reps = max(args.steps, 1)
    predict_s = 0.0
    t0 = time.perf_counter()
    for _ in range(reps):
        model = sc.KNNClassifier.fit(xtr, ytr, p)
        t1 = time.perf_counter()
        got = model.predict(xq)
        predict_s += time.perf_counter() - t1
    total_s = (time.perf_counter() - t0) / reps
    predict_s /= reps  # the mean over the timed repetitions
    out = {"workload": "KNNClassifier k=5, Euclidian, LinearSearch, 10,000x32 f64 train, "
                       "2,000 queries (10-class Gaussian mixture)",
           "value": xq.shape[0] / predict_s, "unit": "queries/s (predict)",
           "ms_fit_predict": total_s * 1e3, "dtype": "f64"}
    if not args.no_cpu:
        classes = np.unique(ytr)
        ycls = np.searchsorted(classes, ytr)
        O.lib()
        t0 = time.perf_counter()
        oi, od, oc = O.linear_find_batch(xtr, xq, 5)
        want = np.array([classes[O.classifier_predict_row(oi[r, :oc[r]], od[r, :oc[r]], ycls,
                                                          len(classes), 0)]
                         for r in range(xq.shape[0])])
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": xq.shape[0] / cpu_s, "unit": "queries/s", "cores": 1,
                               "kind": "port",
                               "sample": f"all {xq.shape[0]:,} queries, fit + predict "
                                         f"({cpu_s:.2f} s); host {cpu_model()}, nproc {os.cpu_count()}"}
        out["label_mismatches_vs_oracle"] = int(np.sum(np.asarray(got) != want))
    return out


def bench_dbscan(args):
    """DBSCAN::fit (dbscan.rs:274-354), eps = 1.5, min_samples = 10, Euclidian,
    on 200,000 x 8 f64 rows of a 10-class Gaussian mixture (unit variance,
    centres ~ N(0, 16 I); about 55 neighbours per row).  One step is one
    sc_dbscan_fit call on the resident index: the batched device radius
    search over every row, the neighbour lists copied to the host and the
    reference's expansion loop there.  The rate counts n^2 distance
    evaluations per fit, a unit that compares fits of different sizes.  The
    CPU baseline is the oracle's restatement of the reference loop on the
    first 40,000 rows, on one core; the device fit of the same rows must give
    its labels and cluster count."""
    from smartcore_amd import _capi as S
    from smartcore_amd import synth

    n, d, eps, min_samples, ns = 200_000, 8, 1.5, 10, 40_000
    x, _, _, _ = synth.knn_mixture(seed=7, n=n, q=16, d=d)
    lib = S.lib()
    lib.sc_profile_enable.argtypes = [C.c_int]
    lib.sc_profile_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    ix = S.Index(x)
    for _ in range(max(args.warmup, 1)):
        ix.dbscan_fit(eps, min_samples)
    reps = max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        lab, nc = ix.dbscan_fit(eps, min_samples)
    fit_s = (time.perf_counter() - t0) / reps
    # the dominant kernel (the capped radius pass) timed with HIP events on
    # its stream in one more fit
    lib.sc_profile_enable(1)
    ix.dbscan_fit(eps, min_samples)
    lib.sc_profile_enable(0)
    kms, launches = profile_read(lib, "radius_capped")
    k_s = kms / 1e3 / max(launches, 1)
    flops = 3.0 * d * float(n) * n  # per pair: d subtractions and d fused multiply-adds
    noise = int(np.sum(lab < 0))
    out = {"workload": f"DBSCAN eps={eps}, min_samples={min_samples}, Euclidian, {n:,}x{d} f64 "
                       "rows (10-class Gaussian mixture)",
           "value": float(n) * n / fit_s, "unit": "distance evaluations/s",
           "ms_per_fit": fit_s * 1e3, "clusters": int(nc), "noise_rows": noise, "dtype": "f64",
           "roofline": {"bound": "fp64-valu", "achieved": flops / k_s / 1e12 if k_s > 0 else None,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                        "frac": flops / k_s / PEAK_FP64 if k_s > 0 else None, "traffic": None,
                        "kernel": "radius_kernel (capped pass over all rows)",
                        "avg_ms": k_s * 1e3}}
    if not args.no_cpu:
        import oracle as O

        xs = np.ascontiguousarray(x[:ns])
        glab, gnc = S.Index(xs).dbscan_fit(eps, min_samples)
        t0 = time.perf_counter()
        olab, onc = O.dbscan_fit(xs, eps, min_samples)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": float(ns) * ns / cpu_s, "unit": "distance evaluations/s",
                               "cores": 1, "kind": "port",
                               "sample": f"one fit of the first {ns:,} rows ({cpu_s:.1f} s); host "
                                         f"{cpu_model()}, nproc {os.cpu_count()}"}
        out["vs_oracle"] = {"rows": ns, "label_mismatches": int(np.sum(glab != olab)),
                            "clusters_equal": int(gnc) == int(onc)}
    return out


def bench_serve(args):
    """Continuous-batching server over a resident 2,000,000 x 96 f32 index
    (rows ~ N(0, I)), k = 20.  Requests arrive by a Poisson process at fixed
    rates, each with round(exp(U[0, ln 4096])) query rows; 80% Euclidean
    (BF16 MFMA shortlist, FP64 rescoring), 20% Manhattan (bounded f32 sums,
    FP64 rescoring), both the SC_MIXED form.  Reports per rate the query
    throughput and the p50 / p99 submit-to-result latency, overall and per
    metric, and checks sampled requests against the same search run alone
    and a few rows of each metric against the oracle."""
    import threading

    import torch

    from smartcore_amd import _capi as S
    from smartcore_amd import synth  # noqa: F401

    N, D, K = args.serve_n, 96, 20
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(303)
    refs = torch.randn((N, D), generator=g, device=dev, dtype=torch.float32)
    ix = S.Index(device_ptr=refs.data_ptr(), n=N, d=D, dtype=np.float32)
    rng = np.random.default_rng(404)
    pool = rng.normal(size=(65_536, D)).astype(np.float32)
    srv = S.Server(ix, max_batch_queries=4096, max_inflight=3, batch_window_us=200)

    def make_request():
        q = int(round(np.exp(rng.uniform(0.0, np.log(4096.0)))))
        s0 = int(rng.integers(0, pool.shape[0] - q + 1))
        euclid = rng.random() < 0.8
        metric = S.SC_EUCLIDIAN if euclid else S.SC_MANHATTAN
        return pool[s0:s0 + q], metric, S.SC_MIXED

    # warm-up: both forms compiled and the BF16 copy built
    for metric, mode in ((S.SC_EUCLIDIAN, S.SC_MIXED), (S.SC_MANHATTAN, S.SC_MIXED)):
        srv.wait(srv.submit(pool[:64], K, metric, 2, mode))
    levels = []
    for rate in args.serve_rates:
        n_req = max(20, int(rate * args.serve_seconds))
        reqs = [make_request() for _ in range(n_req)]
        gaps = rng.exponential(1.0 / rate, size=n_req)
        tickets = [None] * n_req
        lat = [0.0] * n_req
        outs = {}
        sample = set(range(0, n_req, max(1, n_req // 12)))

        def waiter():
            for j in range(n_req):
                while tickets[j] is None:
                    time.sleep(0.0002)
                gi, gd, gc, lt = srv.wait(tickets[j])
                lat[j] = lt
                if j in sample:
                    outs[j] = (gi, gd)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for j, (qs, metric, mode) in enumerate(reqs):
            t_next += gaps[j]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[j] = srv.submit(qs, K, metric, 2, mode)
        th.join()
        wall = time.perf_counter() - t0
        mism = 0
        for j in sample:
            qs, metric, mode = reqs[j]
            ai, ad, ac = ix.knn(qs, K, metric, 2, mode)
            mism += int(np.sum(outs[j][0] != ai))
        lat_ms = np.array(lat)
        nq = sum(r[0].shape[0] for r in reqs)
        is_l1 = np.array([r[1] == S.SC_MANHATTAN for r in reqs])
        per_metric = {}
        for name, sel in (("euclidian", ~is_l1), ("manhattan", is_l1)):
            if sel.any():
                per_metric[name] = {"requests": int(sel.sum()),
                                    "p50_ms": float(np.percentile(lat_ms[sel], 50)),
                                    "p99_ms": float(np.percentile(lat_ms[sel], 99))}
        levels.append({"offered_requests_per_s": rate, "requests": n_req,
                       "queries_per_s": nq / wall, "p50_ms": float(np.percentile(lat_ms, 50)),
                       "p99_ms": float(np.percentile(lat_ms, 99)), "per_metric": per_metric,
                       "checked_requests": len(sample), "index_mismatches_vs_alone": mism})
        log(f"[bench] serve {rate} req/s: {nq / wall:.0f} q/s p50 {np.percentile(lat_ms, 50):.1f} ms "
            f"p99 {np.percentile(lat_ms, 99):.1f} ms")
    st = srv.stats()
    srv.close()
    # the oracle on a few rows of one Euclidean and one Manhattan request
    oracle_mism = None
    cpu = None
    if not args.no_cpu:
        import oracle as O

        href = refs.cpu().numpy()
        qs = pool[:2]
        ai, _, _ = ix.knn(qs, K, S.SC_EUCLIDIAN, 2, S.SC_MIXED)
        bi, _, _ = ix.knn(qs, K, S.SC_MANHATTAN, 1, S.SC_MIXED)
        oi, _, _ = O.linear_find_batch(href, qs, K, O.EUCLIDIAN)
        mi, _, _ = O.linear_find_batch(href, qs, K, O.MANHATTAN)
        oracle_mism = int(np.sum(ai != oi) + np.sum(bi != mi))
        t0 = time.perf_counter()
        O.linear_find_batch(href, pool[2:4], K, O.EUCLIDIAN)
        t1 = time.perf_counter()
        O.linear_find_batch(href, pool[4:6], K, O.MANHATTAN)
        t2 = time.perf_counter()
        cpu = {"value": (t1 - t0) / 2 * 1e3, "unit": "ms/query (Euclidian)",
               "manhattan_ms_per_query": (t2 - t1) / 2 * 1e3, "cores": 1, "kind": "port",
               "sample": f"2 Euclidian + 2 Manhattan queries against all {N:,} rows; host "
                         f"{cpu_model()}, nproc {os.cpu_count()}"}
    ix.close()
    return {"workload": f"kNN server, {N:,}x{D} f32 resident rows, k={K}, requests of 1-4096 "
                        "rows (log-uniform), Poisson arrivals, 80% Euclidean / 20% Manhattan, "
                        "both SC_MIXED (FP64-rescored)",
            "levels": levels, "batches": st["batches"], "hbm_peak_bytes": st["hbm_peak"],
            "oracle_index_mismatches_4_queries": oracle_mism, "cpu_baseline": cpu}


def bench_copy(args):
    """A plain 1 GiB device copy (torch), the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters into bytes."""
    import torch

    n = 1 << 28  # float32 elements = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(max(args.steps, 1)):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"workload": "copy", "bytes_read": n * 4, "bytes_written": n * 4,
            "launches": max(args.steps, 1)}


def bench_kmeans(args):
    import torch

    from smartcore_amd import _capi as S
    from smartcore_amd import synth

    lib = S.lib()
    lib.sc_profile_enable.argtypes = [C.c_int]
    lib.sc_profile_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    n, d, k, iters = args.km_n, 64, 64, 100
    mode = S.SC_MIXED if args.mode == "mixed" else S.SC_PLAIN_FP64
    t0 = time.time()
    x, _, init = synth.kmeans_blobs(seed=2, n=n, d=d, blobs=64, k=k)
    log(f"[bench] generated blobs {n}x{d} in {time.time() - t0:.1f}s")
    ix = S.Index(x)
    for w in range(args.warmup):
        ix.kmeans_run(init, iters, mode)
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    total_iters = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        cent, labels, size, dist, it = ix.kmeans_run(init, iters, mode)
        total_iters += it
        log(f"[bench] step {s + 1}/{args.steps}: {it} iterations")
    elapsed = time.perf_counter() - t0
    ips = total_iters / elapsed
    # The timed fits replay their Lloyd steps from one hipGraph, whose
    # launches HIP events cannot bracket one by one; one more fit with the
    # tracing hooks on launches the same kernels one by one on the same
    # stream, and its per-launch event timings give the kernel durations.
    lib.sc_profile_enable(1)
    pc_, pl_, ps_, pd_, pit_ = ix.kmeans_run(init, iters, mode)
    lib.sc_profile_enable(0)
    assert np.array_equal(pl_, labels) and pit_ == it
    vs_fp64 = None
    if mode == S.SC_MIXED:
        t1 = time.perf_counter()
        pc, pl, psz, pd, pit = ix.kmeans_run(init, iters, S.SC_PLAIN_FP64)
        pt = time.perf_counter() - t1
        vs_fp64 = {"fp64_form_iterations_per_s": pit / pt, "speedup_vs_fp64_form": ips / (pit / pt),
                   "label_mismatches": int(np.sum(pl != labels)),
                   "centroid_bit_mismatches": int(np.sum(pc.view(np.uint64) != cent.view(np.uint64))),
                   "iterations_equal": bool(pit == it)}
        log(f"[bench] k-means FP64 form {pit / pt:.0f} it/s; mixed/FP64 = {ips / (pit / pt):.2f}")
    kname = "kmeans_mixed_assign" if mode == S.SC_MIXED else "kmeans_assign"
    kms, launches = profile_read(lib, kname)
    avg_s = kms / 1e3 / max(launches, 1)
    if mode == S.SC_MIXED:
        bytes_per_launch = n * d * 4 + 4 * n + 2 * k * d * 8
        achieved = bytes_per_launch / avg_s
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": achieved / PEAK_HBM, "traffic": measured_traffic(kname)[0],
                "traffic_source": measured_traffic(kname)[1]}
    else:
        flops = 3.0 * n * k * d
        achieved = flops / avg_s
        roof = {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None}
    roof.update({"kernel": kname, "avg_ms": avg_s * 1e3,
                 "timing": "HIP events per launch on a profiled fit of the same workload run "
                           "right after the timed fits (those replay a hipGraph)"})
    kern = {}
    for name in ("kmeans_assign", "kmeans_partial", "kmeans_reduce", "kmeans_mixed_assign"):
        ms, c = profile_read(lib, name)
        if c:
            kern[name] = {"avg_ms": ms / c, "launches": c}
    out = {
        "metric": METRIC,
        "value": ips,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if mode == S.SC_MIXED else "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"k-means Lloyd {n:,}x{d} f32 blobs, k={k}, max_iter={iters}",
                   "mode": args.mode, "iterations_per_fit": int(total_iters / args.steps)},
        "roofline": roof,
        "kernels": kern,
    }
    if vs_fp64:
        out["vs_fp64_form"] = vs_fp64
    if not args.no_cpu:
        import oracle as O

        t0 = time.perf_counter()
        tree = O.BBDTree(x)
        build_s = time.perf_counter() - t0
        c = init.copy()
        t0 = time.perf_counter()
        nit = args.cpu_kmeans_iters
        for _ in range(nit):
            dd, sums, counts, memb = tree.clustering(c)
            c = np.where(counts[:, None] > 0, sums / np.maximum(counts, 1)[:, None], c)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": nit / cpu_s, "unit": "iterations/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{nit} BBD-tree Lloyd steps on the full {n:,}x{d} set "
                                         f"(tree build {build_s:.1f} s excluded); host {cpu_model()}, "
                                         f"nproc {os.cpu_count()}"}
        # the same steps on the device from the same initial centroids: labels
        # of the last step and the centroids after it against the oracle's
        gc, gl, gsz, gd, git = ix.kmeans_run(init, nit, mode)
        out["vs_oracle_bbd"] = {
            "steps": nit, "device_steps": int(git),
            "label_mismatches": int(np.sum(gl != memb)),
            "size_mismatches": int(np.sum(gsz != counts)),
            "centroid_rel_error": float(np.max(np.abs(gc - c)) / np.max(np.abs(c)))}
    ix.close()
    return out


def bench_batch(args):
    """512 independent k-means problems of uneven size (n ~ log-uniform
    [1,000, 200,000], d in {8, 16, 32, 64}, k ~ U{4..128}, f32 blobs), 20
    Lloyd iterations each, in one sc_kmeans_batch_run call.  The problems go
    in two groups of about equal row bytes: one group's rows are
    copied while the previous group iterates (persistent label launches per
    shape class, one sums launch over the group's row blocks, the reduces;
    the iterations replayed as one hipGraph per group).  The warmup call
    picks each group's launch shape by timing; the timed calls reuse it, so
    their graphs cover all iterations."""
    from smartcore_amd import _capi as S
    from smartcore_amd import synth

    lib = S.lib()
    lib.sc_profile_enable.argtypes = [C.c_int]
    lib.sc_profile_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    iters = 20
    t0 = time.time()
    probs = synth.kmeans_batch_problems(seed=4, count=args.batch_count)
    rows = sum(x.shape[0] for x, _ in probs)
    log(f"[bench] generated {len(probs)} problems, {rows:,} rows in {time.time() - t0:.1f}s")
    for w in range(max(args.warmup, 1)):
        S.kmeans_batch_run(probs, iters, S.SC_MIXED)
        log(f"[bench] batch warmup {w + 1} done")
    lib.sc_profile_enable(1)
    t0 = time.perf_counter()
    for s in range(args.steps):
        got = S.kmeans_batch_run(probs, iters, S.SC_MIXED)
    wall = (time.perf_counter() - t0) / args.steps
    lib.sc_profile_enable(0)
    st = S.last_stats()
    shape = {"waves_per_workgroup": st["launch_waves"], "rows_per_chunk": 32 * st["launch_waves"],
             "workgroups_per_cu": st["launch_groups_per_cu"],
             "upload_groups": st["upload_groups"],
             "how": "label grid timed on the warmup call's first six iterations per group of "
                    "problems (1, 2, 3, 4, 6 or 8 four-wave workgroups per CU), reused by the "
                    "timed calls; the shape shown is the largest group's"}
    gms, launches = profile_read(lib, "kmeans_batch")
    # one graph per group of problems and call: a call's device time is
    # the sum of its groups' graphs
    graph_s = gms / 1e3 / max(args.steps, 1)
    inst_iters = sum(g[4] for g in got)
    # algorithmic HBM bytes: every active problem's rows (f32, unpadded) and
    # labels once per iteration it runs
    byts = sum(g[4] * (x.shape[0] * x.shape[1] * 4 + 4 * x.shape[0]) for (x, _), g in zip(probs, got))
    vs_fp64 = None
    if not args.no_compare:
        t1 = time.perf_counter()
        ref = S.kmeans_batch_run(probs, iters, S.SC_PLAIN_FP64)
        fp64_s = time.perf_counter() - t1
        bad = [i for i, (a, b) in enumerate(zip(got, ref)) if not np.array_equal(a[1], b[1])]
        it_bad = sum(1 for a, b in zip(got, ref) if a[4] != b[4])
        vs_fp64 = {"fp64_form_ms_per_call": fp64_s * 1e3, "speedup_vs_fp64_form": fp64_s / wall,
                   "problems_with_label_mismatch": len(bad),
                   "problems_with_iteration_mismatch": it_bad}
        log(f"[bench] FP64 form {fp64_s * 1e3:.0f} ms; label-mismatched problems {len(bad)}")
    log(f"[bench] batch graph {graph_s * 1e3:.1f} ms, call {wall * 1e3:.1f} ms")
    cpu = None
    if not args.no_cpu:
        # the oracle's KMeans fit loop (BBD tree) over the first problems one
        # after another, on one core; its labels checked against the batch's
        import oracle as O

        t0 = time.perf_counter()
        done_iters, sample, bad_cpu = 0, 0, 0
        for (x, init), g in zip(probs, got):
            oc, ol, osz, odist, oit = O.kmeans_fit(x, init, iters, use_bbd=True)
            done_iters += oit
            sample += 1
            bad_cpu += int(not np.array_equal(ol, g[1]) or oit != g[4])
            if time.perf_counter() - t0 > args.cpu_batch_seconds:
                break
        cpu_s = time.perf_counter() - t0
        cpu = {"value": done_iters / cpu_s, "unit": "problem-iterations/s", "cores": 1,
               "kind": "port",
               "sample": f"first {sample} of {len(probs)} problems fitted one after another "
                         f"({cpu_s:.1f} s, BBD tree build included); host {cpu_model()}, "
                         f"nproc {os.cpu_count()}",
               "problems_differing_from_batch": bad_cpu}
    return {
        "value": inst_iters / graph_s,
        "unit": "problem-iterations/s",
        "ms_per_call_device": graph_s * 1e3,
        "ms_per_call_with_transfers": wall * 1e3,
        "launch_shape": shape,
        "config": {"workload": f"{len(probs)} k-means problems, n log-uniform[1e3, 2e5] "
                               f"({rows:,} rows), d in {{8,16,32,64}}, k in 4..128, "
                               f"max_iter={iters}, problems in groups of about equal row bytes (a group's rows "
                               f"go up while the previous group iterates), per group one hipGraph of "
                               f"its iterations"},
        "roofline": {"bound": "hbm", "achieved": byts / graph_s / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": byts / graph_s / PEAK_HBM,
                     "traffic": measured_traffic("kmeans_batch")[0],
                     "traffic_source": measured_traffic("kmeans_batch")[1],
                     "traffic_per": "call (the iteration kernels of both groups, 20 iterations)",
                     "kernel": "label + sums + reduce kernels (graphs of all iterations)"},
        "vs_fp64_form": vs_fp64,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["all", "knn", "classifier", "kmeans", "batch", "serve",
                                           "dbscan", "copy"],
                    default="all",
                    help="all = the kNN line plus the k-means and serving figures")
    ap.add_argument("--serve-n", type=int, default=2_000_000)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[25.0, 100.0, 200.0])
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--mode", choices=["plain", "mixed"], default="mixed")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64-form comparison steps of the mixed kNN and batched "
                         "k-means benches")
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--q", type=int, default=65_536)
    ap.add_argument("--km-n", type=int, default=1_000_000)
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-kmeans-iters", type=int, default=2)
    ap.add_argument("--cpu-batch-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the distributed form is design-only)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.workload == "copy":
        out = bench_copy(args)
    elif args.workload == "kmeans":
        out = bench_kmeans(args)
    elif args.workload == "serve":
        out = bench_serve(args)
    elif args.workload == "classifier":
        out = bench_classifier(args)
    elif args.workload == "dbscan":
        out = bench_dbscan(args)
    elif args.workload == "batch":
        out = bench_batch(args)
        out.update({"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "dtype": "bf16", "data": "synthetic (seeded)"})
    else:
        out = bench_knn(args)
        if args.workload == "all":
            # the other two parts of the metric, reported beside the kNN line
            km_args = argparse.Namespace(**vars(args))
            km_args.steps, km_args.warmup = 1, 1
            km = bench_kmeans(km_args)
            out["kmeans"] = {k: km[k] for k in ("value", "unit", "ms_per_step", "dtype", "config",
                                                "roofline", "kernels", "cpu_baseline",
                                                "vs_fp64_form", "vs_oracle_bbd") if k in km}
            out["knn_classifier"] = bench_classifier(args)
            b_args = argparse.Namespace(**vars(args))
            b_args.steps, b_args.warmup = 1, 1
            out["kmeans_batch"] = bench_batch(b_args)
            d_args = argparse.Namespace(**vars(args))
            d_args.steps, d_args.warmup = 3, 1
            out["dbscan"] = bench_dbscan(d_args)
            out["serve"] = bench_serve(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
