"""Benchmark: batched SimHash sketch search with exact rerank on one MI355X.

Workload (one step = one search of the whole query batch, inputs resident in
HBM): N = 1,200,000 points, d = 100, float32, drawn from a mixture of 1,000
Gaussian clusters (sigma 0.3) and L2-normalised; 10,000 queries from the same
mixture; 1024-bit SimHash sketches (32 words) from the reference's
hyperplane stream; top-1024 candidates by Hamming distance, exact rerank to
k = 32 by angular distance.

Beside it the same run times the plain FP32 exact scan on the same GPU
(recall@32 is measured against its results, which are FP64-verified exact)
and the reference's CPU exact k-NN (GoldStandard) on a bounded query sample.

python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HBM bytes per launch of the dominant kernel measured with FETCH_SIZE /
# WRITE_SIZE counters at the default workload, recorded in these summaries:
# the pivot-cell search's appending pass (with the approximate distances
# fused into its keys, the default) and the rerank's filter.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "rerank_kernel_simhash1024_N1.2M_Q10k.md")
PASS2_PROFILE = os.path.join(ROOT, "profiles", "cell_scan_kernel_fused_simhash1024_N1.2M_Q10k.md")
DEFAULT_WORKLOAD = dict(n=1_200_000, queries=10_000, words=32, candidates=1024, k=32)

# MI355X published peaks.
HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def measured_traffic(args, path=TRAFFIC_PROFILE):
    """Counter-measured bytes per launch, when this run is the profiled workload."""
    if any(getattr(args, key) != val for key, val in DEFAULT_WORKLOAD.items()):
        return None
    try:
        for line in open(path):
            if line.startswith("TRAFFIC_BYTES_PER_LAUNCH:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def roofline(args, bytes_launch: float, avg_launch_ms: float) -> dict:
    """The rerank filter against HBM.  Queries grouped by nearest cell share
    candidate rows in L2 and the MALL, so the bytes the algorithm gathers
    (each query's candidate rows) exceed what reaches HBM: `achieved` is the
    counter-measured HBM traffic of one launch over its live launch time, and
    the gather rate is reported beside it."""
    traffic = measured_traffic(args)
    gathered_gbs = bytes_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    if traffic is not None and avg_launch_ms > 0:
        achieved = traffic / (avg_launch_ms * 1e-3) / 1e9
        basis = "counter-measured HBM bytes per launch (traffic) / avg_launch_ms"
    else:
        achieved = gathered_gbs
        basis = "no counter profile for this workload: algorithmic (gathered) bytes / avg_launch_ms"
    return {
        "kernel": "rerank_kernel (filter: gather of the candidate rows, approximate angular distance, error band)",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "achieved_basis": basis,
        "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain copy kernel; "
                          "profiles/rerank_kernel_simhash1024_N1.2M_Q10k.md",
        "algorithmic_bytes_per_launch": bytes_launch,
        "gathered_gbs": gathered_gbs,
        "gathered_note": "every candidate row of every query in the filter's format, over the launch time; above "
                         "HBM peak because rows shared by queries of one cell and its neighbours are served from "
                         "L2; the limit is the rate of the random row gather through L2, not HBM bandwidth (the "
                         "launch time is the same with 6, 7 or 8 resident workgroups per CU, and 36% fewer bytes "
                         "past L2 took 4% off it)",
        "avg_launch_ms": avg_launch_ms,
        "launches_per_step": 1,
    }


def roofline_pass2(args, avg_launch_ms: float, tiles: float, bits: int, d: int) -> dict:
    """The pivot-cell search's appending pass with fused distances against
    HBM: `achieved` is the counter-measured HBM traffic of one launch over
    its live launch time.  The algorithmic bytes are the data of every tile
    product (its 32 points' packed sketches and FP16 unit rows, read once per
    chunk of up to 32 queries); the appended keys it writes are in the
    counters but not in that figure.  Its matrix-core rate (FP4 Hamming
    products and FP16 hi/lo angular products) is reported beside it."""
    traffic = measured_traffic(args, PASS2_PROFILE)
    ks = (d + 15) // 16
    tile_bytes = 32.0 * (bits / 8.0 + 2.0 * ((d + 7) // 8 * 8))
    algo = tiles * tile_bytes
    secs = avg_launch_ms * 1e-3
    if traffic is not None and avg_launch_ms > 0:
        achieved = traffic / secs / 1e9
        basis = "counter-measured HBM bytes per launch (traffic) / avg_launch_ms"
    else:
        achieved = algo / secs / 1e9 if secs > 0 else 0.0
        basis = "no counter profile for this workload: algorithmic bytes / avg_launch_ms"
    fp4_flops = tiles * 32 * 32 * 2.0 * bits
    fp16_flops = tiles * 32 * 32 * 2.0 * 2 * 16 * ks
    return {
        "kernel": "cell_scan_kernel, appending pass (FP4 Hamming tile products, threshold appends, FP16 "
                  "approximate angular distances fused into the keys)",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "achieved_basis": basis,
        "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain copy kernel; "
                          "profiles/cell_scan_kernel_fused_simhash1024_N1.2M_Q10k.md",
        "algorithmic_bytes_per_launch": algo,
        "tile_products_per_launch": tiles,
        "mfma": {"fp4_tflops": fp4_flops / secs / 1e12 if secs > 0 else 0.0,
                 "fp4_frac_of_dense_peak": fp4_flops / secs / 10.1e15 if secs > 0 else 0.0,
                 "fp16_tflops": fp16_flops / secs / 1e12 if secs > 0 else 0.0,
                 "fp16_frac_of_dense_peak": fp16_flops / secs / 2.5e15 if secs > 0 else 0.0},
        "bound_note": "neither: each chunk of up to 32 queries x 8 tiles lives 11-25 us, most of it dependent "
                      "set-up loads, operand staging and the appends' atomics; 4 workgroups a CU (128 registers)",
        "avg_launch_ms": avg_launch_ms,
        "launches_per_step": 1,
    }


FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix-core peak


def exact_scan(lsh, torch, data, queries, k, out_like, stream, valu=False, reps=3) -> dict:
    """The plain FP32 exact angular scan of `queries` over `data` on this GPU:
    whole-search time (median of `reps` searches after one warm-up), and the
    device time of its scan launches alone (events around each, one profiled
    search).  valu=True builds the index with the scan on the FP32 vector
    units (LSHGPU_EXACT_VALU=1) instead of the FP32 matrix cores.  Returns the
    results of the last search as `out`."""
    old = os.environ.get("LSHGPU_EXACT_VALU")
    if valu:
        os.environ["LSHGPU_EXACT_VALU"] = "1"
    else:
        os.environ.pop("LSHGPU_EXACT_VALU", None)
    try:
        ex = lsh.Index(data, implementation=2, distanceFunc=1)
    finally:
        if old is None:
            os.environ.pop("LSHGPU_EXACT_VALU", None)
        else:
            os.environ["LSHGPU_EXACT_VALU"] = old
    eout = torch.empty_like(out_like)
    ex.search(queries, k, out=eout, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ex.search(queries, k, out=eout, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ex.set_profiling(1)
    ex.search(queries, k, out=eout, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    pt = ex.phase_times()
    ex.set_profiling(False)
    ex.close()
    Q, d = queries.shape
    n = data.shape[0]
    ms = sorted(times)[len(times) // 2] * 1e3
    flops = 2.0 * n * Q * d
    return {"queries_per_s": Q / (ms * 1e-3), "ms": ms, "scan_kernels_ms": pt["scan_ms"],
            "select_ms": pt["select_ms"], "finalize_ms": pt["final_ms"],
            "scan_tflops": flops / (pt["scan_ms"] * 1e-3) / 1e12 if pt["scan_ms"] > 0 else None,
            "search_tflops": flops / (ms * 1e-3) / 1e12, "out": eout}


def exact_roofline(ex: dict, n: int, q: int, d: int) -> dict:
    """The exact scan against the FP32 matrix-core peak: algorithmic flops
    2*N*Q*d (one multiply-add per point, query and dimension) over the device
    time of the scan launches of one search."""
    achieved = ex["scan_tflops"] or 0.0
    return {"kernel": "exact_mfma_kernel (v_mfma_f32_32x32x2f32 dot products, distances, threshold appends)",
            "bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
            "algorithmic_flops_per_search": 2.0 * n * q * d, "scan_kernels_ms": ex["scan_kernels_ms"],
            "whole_search_tflops": ex["search_tflops"],
            "whole_search_frac": ex["search_tflops"] / FP32_PEAK_TFLOPS}


def cpu_baseline(args, data_h, queries_h) -> dict:
    """goldstandard.cpp:136-173 restated (oracle.cpp, with the by-value point
    copies and the full std::sort per query): (a) one thread over a
    time-bounded prefix of the first 100 queries, (b) every host thread the
    run may use over the first 100 queries."""
    import oracle

    nproc = os.cpu_count() or 1
    threads = max(1, min(args.cpu_threads, nproc))
    nq = min(args.cpu_queries, len(queries_h))
    secs, _ = oracle.goldstandard_port(data_h, queries_h[:nq], 32, threads=threads)
    # (a) single thread, in chunks until the 100 queries are done or the
    # time budget is spent
    done, st_secs, chunk = 0, 0.0, 4
    while done < nq and st_secs < args.cpu_single_budget_s:
        m = min(chunk, nq - done)
        t, _ = oracle.goldstandard_port(data_h, queries_h[done:done + m], 32, threads=1)
        st_secs += t
        done += m
    model = cpu_model()
    return {"value": nq / secs, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {nq} of the {len(queries_h)} queries against all {len(data_h)} points; "
                      f"goldstandard.cpp:136-173 restated with its by-value point copies and full std::sort per "
                      f"query; {threads} threads (queries split across std::threads) on {model}",
            "nproc": nproc, "cpu_model": model,
            "single_thread": {"value": done / st_secs if st_secs > 0 else None, "unit": "queries/s", "cores": 1,
                              "queries_timed": done, "seconds": st_secs,
                              "sample": f"first {done} queries, one thread, same algorithm"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_200_000)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--words", type=int, default=32)
    ap.add_argument("--candidates", type=int, default=1024)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--cpu-queries", type=int, default=100, help="query sample of the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the all-cores CPU baseline")
    ap.add_argument("--cpu-single-budget-s", type=float, default=20.0,
                    help="time budget of the single-thread CPU baseline (whole chunks of 4 queries)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-fp32-projection", action="store_true",
                    help="skip the comparison run with the SimHash projections on the FP32 vector units")
    ap.add_argument("--no-batches", action="store_true",
                    help="skip the uneven-batch run (the queries in batches log-uniform in 1..4096, through "
                         "captured graphs and as one search call per batch)")
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 3, 4],
                    help="1: the benchmark workload (default); 0, 2, 3, 4: the other configurations, measured for "
                         "information (one JSON line, no roofline or CPU baseline)")
    ap.add_argument("--serve-rates", default="50,200,800,2000,4000,8000,16000,32000",
                    help="configuration 4: offered request rates (requests/s) to sweep")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="configuration 4: seconds per rate")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB on the device 3 times with an elementwise kernel (a "
                         "plain copy of known size that scales FETCH_SIZE/WRITE_SIZE counter readings)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(sharding the index over several GPUs is not implemented)"}))
        return 2

    import numpy as np
    import torch  # device memory and streams only

    import lshongpu_amd as lsh
    from lshongpu_amd import synth

    torch.cuda.init()
    dev = torch.device("cuda:0")
    if args.config != 1:
        return other_config(args, np, torch, lsh, synth, dev)
    t0 = time.time()
    data_h, queries_h = synth.config(1, n=args.n, q=args.queries)
    gen_s = time.time() - t0
    data = torch.from_numpy(data_h).to(dev)
    queries = torch.from_numpy(queries_h).to(dev)
    Q, d, k = queries.shape[0], data.shape[1], args.k
    kq = lsh.calculateK(k)
    # A stream of its own (default priority): a search queued on a given
    # stream with device results returns while its last kernels run, so the
    # host queues the next step meanwhile.  (The default stream's handle is
    # 0, which the library reads as "its own stream, wait for the results".)
    stream = torch.cuda.Stream(device=dev)

    # ---- index build (constructionTime of the reference; not in the step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix = lsh.Index(data, implementation=3, sketchDim=args.words, candidates=args.candidates)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    out = torch.empty((Q, kq * 2), dtype=torch.int32, device=dev)

    def step():
        ix.search(queries, k, out=out, stream=stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    # The dominant kernel: with the approximate distances fused into the cell
    # scan's keys (the default for k <= 64 over the FP16 image) the appending
    # pass of the cell search, else the rerank's filter.
    half_filter = os.environ.get("LSHGPU_RERANK_F32", "0") != "1"
    fused = half_filter and kq <= 64 and os.environ.get("LSHGPU_FUSED", "1") != "0" and d <= 128
    # Timed region: HIP events around that kernel's launch only (profiling
    # mode 3 or 2), read once after the steps.
    ix.set_profiling(3 if fused else 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    live = ix.phase_times()
    dominant_ms_total = live["pass2_ms"] if fused else live["final_ms"]
    # Per-phase breakdown from a separate, fully profiled pass (events around
    # every phase cost time, so it stays out of the timed region).
    nb = max(1, min(args.steps, 5))
    ix.set_profiling(1)
    for _ in range(nb):
        step()
    torch.cuda.synchronize()
    pt = ix.phase_times()
    phases = {key: pt[key] / nb for key in pt}
    ix.set_profiling(False)
    ms_per_step = elapsed * 1e3 / args.steps
    qps = Q * args.steps / elapsed
    res = out.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)

    # ---- dominant kernel: the rerank's filter (one launch per step), a
    # gather of the candidate rows bound by HBM; its FP64 recheck of the
    # error band runs as a separate, much shorter kernel.  Algorithmic bytes
    # per launch, per query: every candidate row once in the filter's format
    # (FP16 unit rows of hs = d rounded up to 8 halves, or FP32 rows), the
    # candidate keys (8*C), the query row and the band it writes (at least k
    # entries of 8 bytes).
    steps = args.steps
    C = min(args.candidates, args.n)
    avg_launch_ms = dominant_ms_total / steps
    row_bytes = 2.0 * ((d + 7) // 8 * 8) if half_filter else 4.0 * d
    bytes_launch = float(Q) * (row_bytes * C + 8.0 * C + 4.0 * d + 8.0 * kq)
    # The threshold pass over the pivot cells: matrix-core tile products
    # (32 queries x 32 points x all sketch bits) against a pass over all points.
    bits = 32 * ((args.words + 7) // 8 * 8)
    pair_products = phases["mfma_points"] * Q  # (query, point) pairs multiplied per step
    threshold_pass = {
        "ms": phases["scan_ms"],
        "pairs_multiplied": pair_products,
        "fraction_of_all_pairs": pair_products / (float(Q) * args.n),
        "mfma_tflops": 2.0 * bits * pair_products / (phases["mfma_scan_ms"] * 1e-3) / 1e12
        if phases["mfma_scan_ms"] > 0 else 0.0,
    }

    # ---- the same search with the SimHash projections on the FP32 vector
    # units (projection 1) beside the default bf16 matrix cores with the FP32
    # recheck (projection 0): identical sketches, so identical results.
    projection = None
    if not args.no_fp32_projection:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ix32 = lsh.Index(data, implementation=3, sketchDim=args.words, candidates=args.candidates, projection=1)
        torch.cuda.synchronize()
        build32_ms = (time.perf_counter() - t0) * 1e3
        out32 = torch.empty_like(out)

        def step32():
            ix32.search(queries, k, out=out32, stream=stream.cuda_stream)

        for _ in range(args.warmup):
            step32()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step32()
        torch.cuda.synchronize()
        el32 = time.perf_counter() - t0
        ix32.set_profiling(1)
        for _ in range(nb):
            step32()
        torch.cuda.synchronize()
        ph32 = ix32.phase_times()
        ix32.close()
        same = bool(torch.equal(out32, out))
        projection = {
            "bf16_mfma_fp32_recheck": {"queries_per_s": qps, "index_build_ms": build_ms,
                                       "query_sketch_ms": phases["prep_ms"]},
            "fp32_valu": {"queries_per_s": Q * args.steps / el32, "index_build_ms": build32_ms,
                          "query_sketch_ms": ph32["prep_ms"] / nb},
            "results_identical": same,
        }

    # ---- the same search with host buffers: the queries copied in and the
    # results copied out over PCIe inside every step (reported beside the
    # device-resident rate, never as it).
    nh = max(1, min(args.steps, 5))
    host_buffers = {"note": "host queries in and host results out, the PCIe copies and the host wait inside the "
                            "step: pageable numpy arrays (staged by the driver) and pinned host tensors (DMA)"}
    hq = np.ascontiguousarray(queries_h)
    pq = torch.from_numpy(hq).pin_memory()
    pres = torch.empty((Q, kq * 2), dtype=torch.int32, pin_memory=True)
    for name, fn in (("pageable", lambda: ix.search(hq, k)), ("pinned", lambda: ix.search(pq, k, out=pres))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nh):
            hres = fn()
        host_s = time.perf_counter() - t0
        if name == "pinned":
            hres = pres.numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
        host_buffers[name] = {"queries_per_s": Q * nh / host_s, "ms_per_step": host_s * 1e3 / nh,
                              "results_equal_device": bool(np.array_equal(hres, res))}

    # ---- the same queries in uneven batches (sizes log-uniform in 1..4096):
    # each batch replays the captured graph of its size class with no host
    # wait between batches, against one ordinary search call per batch.
    batches = None
    if not args.no_batches:
        batches = {}
        brng = np.random.default_rng(7)
        for label, top, nq in (("log_uniform_1_4096", 4096.0, Q), ("log_uniform_1_64", 64.0, min(Q, 2000))):
            sizes, left = [], nq
            while left > 0:
                b = min(left, int(round(np.exp(brng.uniform(0.0, np.log(top))))))
                sizes.append(b)
                left -= b
            bounds = np.cumsum([0] + sizes)
            outb = torch.empty_like(out)

            def step_graphs():
                ix.search_batches(queries[:nq], sizes, k, out=outb, stream=stream.cuda_stream)

            def step_calls():
                for a, b in zip(bounds[:-1], bounds[1:]):
                    ix.search(queries[a:b], k, out=outb[a:b], stream=stream.cuda_stream)

            reps = max(1, min(args.steps, 5))
            timing = {}
            for name, fn in (("graphs", step_graphs), ("one_search_call_per_batch", step_calls)):
                fn()  # warm-up (captures the graphs of the classes)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                torch.cuda.synchronize()
                timing[name] = nq * reps / (time.perf_counter() - t0)
                if name == "graphs":
                    bstats = ix.batch_stats()
                    same_b = bool(torch.equal(outb[:nq], out[:nq]))
            batches[label] = {"queries": nq, "batches": len(sizes), "largest": int(max(sizes)),
                              "graphs_qps": timing["graphs"],
                              "one_search_call_per_batch_qps": timing["one_search_call_per_batch"],
                              "graph_batches": bstats["graph_batches"], "redone_queries": bstats["redone_queries"],
                              "results_equal_one_batch": same_b}

    # ---- plain FP32 exact scan on the same GPU (and the truth for recall)
    exact = {}
    recall = None
    if not args.no_exact:
        em = exact_scan(lsh, torch, data, queries, k, out, stream)
        truth = em.pop("out").cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
        ev = exact_scan(lsh, torch, data, queries, k, out, stream, valu=True, reps=1)
        same = bool(np.array_equal(ev.pop("out").cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq), truth))
        inter = [len(set(a[:k]) & set(b[:k])) / k for a, b in zip(res["ID"], truth["ID"])]
        recall = float(np.mean(inter))
        # The scan's ceiling: every query against every point at the FP32
        # matrix-core peak, 2*N*d flops a query.
        ceiling = FP32_PEAK_TFLOPS * 1e12 / (2.0 * args.n * d)
        exact = {"exact_fp32_scan_qps": em["queries_per_s"], "exact_fp32_scan_ms": em["ms"],
                 "speedup_vs_exact_scan": qps / em["queries_per_s"],
                 "exact_scan_ceiling_qps": ceiling,
                 "speedup_vs_exact_scan_ceiling": qps / ceiling,
                 "exact_fp32_scan": {"matrix_cores": em, "vector_units": ev, "results_identical": same,
                                     "roofline": exact_roofline(em, args.n, Q, d)}}

    # ---- reference CPU path (GoldStandard exact k-NN) on query samples
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, data_h, queries_h)

    line = {
        "metric": "k-NN queries/sec at recall@32>=0.9, N=1.2M d=100, vs exact scan; % HBM peak",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 Hamming distances; approximate angular distances from f16 matrix-core products of unit rows "
                 "(f32 sums, 16-bit fixed point) with a proven error band, f64 recheck of the band" if fused else
                 "u32 Hamming distances; angular rerank filter in f32 arithmetic over f16 unit rows with a proven "
                 "error band, f64 recheck of the band" if half_filter else
                 "u32 Hamming distances; f32 rerank filter, f64 recheck of the error band",
        "data": "synthetic (seeded Gaussian mixture), generated in-process",
        "config": {"workload": f"SimHash {32 * args.words}-bit sketch scan, top-{args.candidates} by Hamming, "
                               f"exact angular rerank to k={k}",
                   "N": args.n, "d": d, "queries": Q, "clusters": 1000, "sigma": 0.3,
                   "candidates": args.candidates, "k": k},
        "recall_at_32": recall,
        **exact,
        "index_build_ms": build_ms,
        "phase_ms_per_step": phases,
        "phase_breakdown_note": f"from {nb} fully profiled steps after the timed region",
        "roofline": roofline_pass2(args, avg_launch_ms, phases["pass2_tiles"], bits, d) if fused
        else roofline(args, bytes_launch, avg_launch_ms),
        # The rerank's filter beside it, timed in the fully profiled steps.
        "rerank_filter": {"ms": phases["final_ms"], "fused_distances": fused,
                          "note": "with fused distances the select does the filter's work from the keys (the "
                                  "k-th 16-bit distance and the band) and no filter kernel runs; without, the "
                                  "filter gathers the candidate rows (the roofline of that form is the "
                                  "rerank_kernel profile's)"},
        "threshold_pass": threshold_pass,
        "simhash_projection": projection,
        "uneven_batches": batches,
        "host_buffers": host_buffers,
        "cpu_baseline": cpu,
        "data_gen_s": gen_s,
    }
    ix.close()
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes n*4 bytes
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = n * 4
    print(json.dumps(line))
    return 0


def other_config(args, np, torch, lsh, synth, dev) -> int:
    """One informational line for configuration 0, 2, 3 or 4: queries/s of
    its path on one GPU (inputs resident in HBM), recall@k against the exact
    scan of the same data on the same GPU."""
    cfg = args.config
    k = 32
    kq = lsh.calculateK(k)
    t0 = time.time()
    # configuration 4 serves requests drawn from a pool of 20,000 queries
    data_h, queries_h = synth.config(cfg, q=20000 if cfg == 4 else None)
    gen_s = time.time() - t0
    if cfg == 4:
        return serve_config4(args, np, torch, lsh, dev, data_h, queries_h, gen_s)
    rng = np.random.default_rng(77 + cfg)
    if cfg == 2:
        # The match-count scan has no matrix-core form: a 512-query sample.
        queries_h = queries_h[np.sort(rng.choice(len(queries_h), 512, replace=False))]
    data = torch.from_numpy(data_h).to(dev)
    queries = torch.from_numpy(queries_h).to(dev)
    Q, d = queries.shape
    stream = torch.cuda.Stream(device=dev)
    if cfg == 0:
        return exact_config0(args, np, torch, lsh, dev, data_h, queries_h, gen_s)
    if cfg == 2:
        workload = ("weighted MinHash, 256 values per point, top-32 by matching values, generalized-Jaccard rerank; "
                    "N=1M d=128 (70% zeros), 512 of the 5,000 queries")
        kw = dict(implementation=4, sketchDim=256, distanceFunc=2, candidates=32)
        dfunc = 2
    elif cfg == 3:
        workload = ("(K,L) framework, N=4M d=128 clustered: L=32 tables of K=2 dense-rotation cross-polytope hashes "
                    "(bf16 matrix cores, FP32 recheck), keys sorted on the device, exact angular rescan of the bucket "
                    "union, query batches log-uniform in 1..4096")
        kw = dict(implementation=2, distanceFunc=1, framework=1, keysImplementation=8, bucketKeyBits=16, tables=32)
        dfunc = 1
    else:
        workload = "SimHash 512-bit sketch scan, top-1024 by Hamming, exact angular rerank; N=10M d=100, batch of 2,048"
        kw = dict(implementation=3, sketchDim=16, distanceFunc=1, candidates=1024)
        dfunc = 1
    tuning = None
    if cfg == 3:
        # (K, L) by the autotuner: every (K hashes, L tables) candidate is built
        # over the data and measured on a 256-query sample (recall@32 against
        # the exact scan, search time); the fastest one at recall >= 0.95 runs.
        grid = [(K_, L_) for K_ in (2, 3, 4) for L_ in (16, 24, 32)]
        cands = [dict(kw, bucketKeyBits=8 * K_, tables=L_) for K_, L_ in grid]
        sample = queries_h[rng.choice(len(queries_h), 256, replace=False)]
        t0 = time.perf_counter()
        best, reports = lsh.autotune(data, sample, k, cands, target_recall=0.95)
        tune_s = time.perf_counter() - t0
        kw = cands[best]
        tuning = {"chosen": {"K": grid[best][0], "L": grid[best][1]}, "target_recall": 0.95, "sample_queries": 256,
                  "seconds": tune_s,
                  "candidates": [{"K": g[0], "L": g[1], "recall": r["recall"], "search_ms": r["search_ms"],
                                  "build_ms": r["build_ms"]} for g, r in zip(grid, reports)]}
        workload = workload.replace("L=32 tables of K=2", f"L={grid[best][1]} tables of K={grid[best][0]} (autotuned)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix = lsh.Index(data, **kw)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    out = torch.empty((Q, kq * 2), dtype=torch.int32, device=dev)
    if cfg == 3:
        sizes, left = [], Q
        while left > 0:
            b = min(left, int(round(np.exp(rng.uniform(0.0, np.log(4096.0))))))
            sizes.append(b)
            left -= b
        bounds = np.cumsum([0] + sizes)

        def step_eager():
            for a, b in zip(bounds[:-1], bounds[1:]):
                ix.search(queries[a:b], k, out=out[a:b], stream=stream.cuda_stream)

        def step():
            # every batch its own search, replaying the captured graph of its
            # size class, with no host wait between batches
            ix.search_batches(queries, sizes, k, out=out, stream=stream.cuda_stream)
    else:
        def step():
            ix.search(queries, k, out=out, stream=stream.cuda_stream)
    steps = max(1, min(args.steps, 3 if cfg in (2, 3) else args.steps))
    for _ in range(max(1, min(args.warmup, 1))):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = out.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
    extra = {}
    if cfg == 3:
        extra["batches"] = len(sizes)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step_eager()
        torch.cuda.synchronize()
        extra["one_search_call_per_batch_qps"] = Q / (time.perf_counter() - t1)
        same = np.array_equal(out.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq), res)
        extra["graph_results_equal_eager"] = bool(same)
    ix.close()
    ex = lsh.Index(data, implementation=2, distanceFunc=dfunc)
    eout = torch.empty_like(out)
    ex.search(queries, k, out=eout, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    truth = eout.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
    ex.close()
    recall = float(np.mean([len(set(a[:k]) & set(b[:k])) / k for a, b in zip(res["ID"], truth["ID"])]))
    print(json.dumps({
        "metric": "queries/s (informational configuration)", "value": Q * steps / elapsed, "unit": "queries/s",
        "n_gpus": 1, "steps": steps, "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True,
        "data": "synthetic, generated in-process", "config": {"workload": workload, "N": int(data_h.shape[0]),
                                                              "d": int(d), "queries": int(Q), "k": k},
        "recall_at_32": recall, "index_build_ms": build_ms, "data_gen_s": gen_s, "autotune": tuning, **extra}))
    return 0


def exact_config0(args, np, torch, lsh, dev, data_h, queries_h, gen_s) -> int:
    """Configuration 0, the plain exact k-NN: N=100k d=100 unit rows, 1,000
    queries, k=32 by angular distance, on the FP32 matrix cores and on the
    FP32 vector units, checked against each other, with the FP32 roofline of
    the scan launches."""
    data = torch.from_numpy(data_h).to(dev)
    queries = torch.from_numpy(queries_h).to(dev)
    Q, d = queries.shape
    n = data.shape[0]
    k = 32
    kq = lsh.calculateK(k)
    stream = torch.cuda.Stream(device=dev)
    out = torch.empty((Q, kq * 2), dtype=torch.int32, device=dev)
    reps = max(3, min(args.steps, 20))
    em = exact_scan(lsh, torch, data, queries, k, out, stream, reps=reps)
    ev = exact_scan(lsh, torch, data, queries, k, out, stream, valu=True, reps=reps)
    a = em.pop("out").cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
    b = ev.pop("out").cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, kq)
    cpu = None if args.no_cpu else cpu_baseline(args, data_h, queries_h)
    print(json.dumps({
        "metric": "queries/s (informational configuration)", "value": em["queries_per_s"], "unit": "queries/s",
        "n_gpus": 1, "steps": reps, "ms_per_step": em["ms"], "higher_is_better": True, "dtype": "f32",
        "data": "synthetic, generated in-process",
        "config": {"workload": "exact angular scan, N=100k d=100, 1,000 queries, k=32", "N": int(n), "d": int(d),
                   "queries": int(Q), "k": k},
        "matrix_cores": em, "vector_units": ev, "results_identical": bool(np.array_equal(a, b)),
        "roofline": exact_roofline(em, n, Q, d), "cpu_baseline": cpu, "data_gen_s": gen_s}))
    return 0


def serve_config4(args, np, torch, lsh, dev, data_h, queries_h, gen_s) -> int:
    """Configuration 4 as a serving load: the 10M-point index with 512-bit
    sketches resident in HBM behind the request loop (continuous batching,
    memory admission); requests of 1..2,048 queries (log-uniform) with k from
    {1, 10, 32, 100} arrive as a Poisson process at each swept rate; reported
    per rate: queries/s served, p50/p99 request latency; recall@k of the
    served path on a query sample against the exact scan."""
    import threading

    rng = np.random.default_rng(4004)
    data = torch.from_numpy(data_h).to(dev)
    N, d = data.shape
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix = lsh.Index(data, implementation=3, sketchDim=16, distanceFunc=1, candidates=1024)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    # recall@k on a 512-query sample (the served results equal the index's)
    sample = queries_h[rng.choice(len(queries_h), 512, replace=False)]
    recall = {}
    with lsh.Index(data, implementation=2, distanceFunc=1) as ex:
        truth = ex.search(sample, 100)["ID"]
    for kk in (1, 10, 32, 100):
        got = ix.search(sample, kk)["ID"][:, :kk]
        recall[str(kk)] = float(np.mean([len(set(a) & set(b[:kk])) / kk for a, b in zip(got, truth)]))
    rates = [float(r) for r in args.serve_rates.split(",")]
    sweep = []
    with lsh.Server(ix, max_batch_queries=4096) as sv:
        # warm-up: a few requests of each size class
        for n in (1, 64, 2048):
            sv.wait(sv.submit(queries_h[:n], 32))
        for lam in rates:
            nreq = max(20, int(lam * args.serve_seconds))
            gaps = rng.exponential(1.0 / lam, nreq)
            sizes = np.minimum(2048, np.ceil(np.exp(rng.uniform(0, np.log(2048), nreq)))).astype(int)
            ks = rng.choice([1, 10, 32, 100], nreq)
            starts = rng.integers(0, len(queries_h) - 2048, nreq)
            tickets = [None] * nreq
            lat = [0.0] * nreq
            ready = threading.Semaphore(0)

            def producer():
                t = time.perf_counter()
                for i in range(nreq):
                    t += gaps[i]
                    dt = t - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    tickets[i] = sv.submit(queries_h[starts[i]:starts[i] + sizes[i]], int(ks[i]))
                    ready.release()

            def collector():
                for i in range(nreq):
                    ready.acquire()
                    lat[i] = sv.wait(tickets[i])[1]

            b0 = sv.stats()
            th = [threading.Thread(target=producer), threading.Thread(target=collector)]
            t0 = time.perf_counter()
            for t in th:
                t.start()
            for t in th:
                t.join()
            wall = time.perf_counter() - t0
            b1 = sv.stats()
            served = int(sizes.sum())
            sweep.append({"offered_requests_per_s": lam, "requests": nreq, "queries": served,
                          "queries_per_s": served / wall, "requests_per_s": nreq / wall,
                          "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                          "batches": b1["batches"] - b0["batches"],
                          "mean_batch_queries": served / max(1, b1["batches"] - b0["batches"])})
        max_batch = sv.stats()["max_batch"]
    ix.close()
    best = max(sweep, key=lambda r: r["queries_per_s"])
    print(json.dumps({
        "metric": "queries/s served (informational configuration)", "value": best["queries_per_s"],
        "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic, generated in-process",
        "config": {"workload": "online serving: SimHash 512-bit index of N=10M d=100 (5,000 clusters) resident in HBM, "
                               "top-1024 by Hamming, exact angular rerank; Poisson arrivals, 1..2048 queries per "
                               "request (log-uniform), k in {1,10,32,100}; continuous batching, memory admission",
                   "N": int(N), "d": int(d), "max_batch_queries": max_batch},
        "recall_at_k": recall, "sweep": sweep, "index_build_ms": build_ms, "data_gen_s": gen_s}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
