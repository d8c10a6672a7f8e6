"""Benchmark: batched overlap search on one MI355X.

Workload (log-normal intervals, 10M x 10M): 10M int32 intervals, starts uniform on
[0, 2.5e8), lengths clamp(round(LogNormal(6, 1.5)), 1, 1e6); 10M queries, lengths
max(1, round(LogNormal(7, 1))).  Inputs are synthetic (seeded generator) and resident
in HBM before timing.  One step = search_idxs for every query in two passes: plan
(sort by end + upper_bound) -> offsets (rank-form count + exclusive scan) -> fill of the
CSR index lists (descending order).

Prints one JSON line (the metric contract) with the roofline of the dominant kernel (the
fill pass) and a CPU baseline (the reference's own code, compiled from its headers into
oracle/_ref when available, else the oracle's restatement) on a bounded query sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=10_000_000)
    p.add_argument("--nq", type=int, default=10_000_000)
    p.add_argument("--cpu-sample", type=int, default=2_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-tune", action="store_true", help="keep the library's default launch choices")
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also launch one 1 GiB device copy (reference kernel for PMC byte counters)")
    p.add_argument("--learned", action="store_true",
                   help="learned vs plain upper_bound (200M piecewise-density intervals, 50M queries)")
    p.add_argument("--learned-n", type=int, default=200_000_000)
    p.add_argument("--learned-nq", type=int, default=50_000_000)
    p.add_argument("--learned-steps", type=int, default=3000)
    p.add_argument("--serving", action="store_true",
                   help="continuous-batching server: 24 resident indexes, Poisson request stream")
    p.add_argument("--serving-sets", type=int, default=24)
    p.add_argument("--serving-max-n", type=int, default=50_000_000)
    p.add_argument("--serving-rates", type=str, default="50,200,800,3200")
    p.add_argument("--serving-seconds", type=float, default=3.0)
    p.add_argument("--skewed", action="store_true",
                   help="int64 hotspot workload: 50M intervals, 20M queries (1%% wide); count + streamed search")
    p.add_argument("--skewed-n", type=int, default=50_000_000)
    p.add_argument("--skewed-nq", type=int, default=20_000_000)
    p.add_argument("--skewed-batch", type=int, default=1_000_000)
    p.add_argument("--skewed-check", type=int, default=200,
                   help="queries of the skewed run whose lists are checked against the oracle (0: none)")
    p.add_argument("--graph", action="store_true",
                   help="small batches: eager plan/offsets/fill vs one captured graph replay")
    p.add_argument("--graph-nq", type=int, default=4096)
    p.add_argument("--single", action="store_true",
                   help="microseconds per single-query call (the C-ABI's countOverlaps / searchIdxs path)")
    p.add_argument("--walk-once", action="store_true",
                   help="one launch of each kernel of the 10M x 10M step (plan, offsets, rank count, count walk, "
                        "fill) with the --tuning knobs, for PMC counters")
    p.add_argument("--tuning", type=str, default=None,
                   help="JSON object of launch knobs to apply instead of autotuning (e.g. the 'launch' of a run)")
    return p.parse_args()


def events(n):
    import torch
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def bench_learned(a):
    """Learned upper_bound (RMI: FP32 root MLP 1-16-16-1, 4096 linear leaves, exact error
    windows) against the plain binary search on the same device index and queries."""
    import numpy as np
    import torch
    import superintervals_amd as sia
    dev = torch.device("cuda:0")
    torch.cuda.init()
    n, nq = a.learned_n, a.learned_nq
    t0 = time.perf_counter()
    s, e = sia.workload_intervals(4, n, seed=41, sorted=True)
    gen_s = time.perf_counter() - t0
    print("[learned] generated %d intervals in %.1f s" % (n, gen_s), file=sys.stderr, flush=True)
    ix = sia.DeviceIndex(np.int32).build(s, e)
    del s, e
    print("[learned] built (%.1f ms on the device)" % ix.last_build_ms, file=sys.stderr, flush=True)
    qs, qe = sia.workload_queries(4, nq, seed=42)
    xs = torch.from_numpy(qe).to(dev)       # upper_bound of the query ends (point and short ranges)
    del qs, qe
    stream = torch.cuda.current_stream()
    lib = sia._native.lib

    def time_it(fn, reps):
        fn()
        torch.cuda.synchronize()
        ev = events(2)
        ev[0].record(stream)
        for _ in range(reps):
            fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    from superintervals_amd import autotune
    plain = torch.empty(nq, dtype=torch.int32, device=dev)
    ub = lambda: sia._native.check(lib.siUpperBound(ix.handle, xs.data_ptr(), nq, plain.data_ptr(), None),
                                   "siUpperBound")
    autotune.apply({"ub_tree": 0})
    t_plain = time_it(ub, a.steps)           # plain halving over all starts
    ref = plain.clone()
    autotune.apply({"ub_tree": 1})
    t_tree = time_it(ub, a.steps)            # Eytzinger top in LDS + block separators
    tree_same = bool(torch.equal(plain, ref))
    res = {}
    for name, steps in (("untrained", 0), ("trained", a.learned_steps)):
        m = sia.LearnedIndex(ix, 4096)
        ev = events(2)
        ev[0].record(stream)
        info = m.train(steps=steps, batch=65536, lr=3e-3, stream=stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        t_train = ev[0].elapsed_time(ev[1])
        out = torch.empty(nq, dtype=torch.int32, device=dev)
        st = torch.zeros(3, dtype=torch.int64, device=dev)
        m.upper_bound_device(xs, out, st, stream=stream)
        torch.cuda.synchronize()
        fb, w, corr = st.cpu().tolist()
        same = bool(torch.equal(out, ref))
        t = time_it(lambda: m.upper_bound_device(xs, out, None, stream=stream), a.steps)
        res[name] = {"ms": t, "queries_per_s": nq / (t / 1e3), "identical_to_plain": same,
                     "fallbacks": fb, "mean_window": w / nq, "mean_leaf_correction": corr / nq,
                     "train_ms": t_train, **info}
        print("[learned] %s: %.3f ms (plain %.3f ms, tree %.3f ms), identical %s, max error %d" %
              (name, t, t_plain, t_tree, same, info["max_abs_err"]), file=sys.stderr, flush=True)
        m.close()
    # the batched search (plan -> offsets -> fill, descending) with the plan's bounds found by
    # plain halving, by the tree and by the trained model: identical output, timed per stage
    m = sia.LearnedIndex(ix, 4096)
    m.train(steps=a.learned_steps, batch=65536, lr=3e-3, stream=stream)
    qs, qe = sia.workload_queries(4, nq, seed=42)
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    del qs, qe
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    ix.plan_device(tqs, tqe, plan, stream=stream)
    ix.offsets_planned(plan, nq, off, stream=stream)
    torch.cuda.synchronize()
    total = int(off[-1].item())
    idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    ref_idx = ref_off = None
    search = {}
    saved = autotune.current()
    try:
        for name, knobs in (("halving", {"ub_tree": 0, "ub_learned": 0}), ("tree", {"ub_tree": 1, "ub_learned": 0}),
                            ("learned", {"ub_tree": 1, "ub_learned": 1})):
            autotune.apply(dict(knobs, plan_sync=0))
            ts = {"plan": 0.0, "offsets": 0.0, "fill": 0.0}
            for rep in range(a.steps + 1):
                ev = events(4)
                ev[0].record(stream)
                ix.plan_device(tqs, tqe, plan, stream=stream)
                ev[1].record(stream)
                ix.offsets_planned(plan, nq, off, stream=stream)
                ev[2].record(stream)
                ix.fill_planned(plan, nq, off, idx, order="desc", stream=stream)
                ev[3].record(stream)
                torch.cuda.synchronize()
                if rep:                                  # the first pass is a warm-up
                    ts["plan"] += ev[0].elapsed_time(ev[1]) / a.steps
                    ts["offsets"] += ev[1].elapsed_time(ev[2]) / a.steps
                    ts["fill"] += ev[2].elapsed_time(ev[3]) / a.steps
            if ref_idx is None:
                ref_off, ref_idx = off.clone(), idx.clone()
                same = True
            else:
                same = bool(torch.equal(off, ref_off)) and bool(torch.equal(idx, ref_idx))
            step = ts["plan"] + ts["offsets"] + ts["fill"]
            search[name] = dict({k: round(v, 4) for k, v in ts.items()}, step_ms=round(step, 4),
                                queries_per_s=nq / (step / 1e3), identical_to_halving=same)
            print("[learned] batched search, bounds by %s: plan %.3f ms, offsets %.3f ms, fill %.3f ms, identical %s"
                  % (name, ts["plan"], ts["offsets"], ts["fill"], same), file=sys.stderr, flush=True)
    finally:
        autotune.apply(saved)
    m.close()
    del plan, idx, ref_idx
    tr = res["trained"]
    out = {
        "metric": "learned upper_bound queries/s vs plain binary search (200M idx x 50M point/short qry)",
        "value": tr["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
        "higher_is_better": True, "dtype": "int32 keys, fp32 model",
        "data": "synthetic (seeded piecewise-density generator, inputs resident in HBM)",
        "config": {"workload": "upper_bound of 50M query ends over 200M int32 starts (10 density segments)",
                   "intervals": n, "queries": nq, "leaves": 4096, "train_steps": a.learned_steps, "batch": 65536,
                   "host_generation_s": round(gen_s, 2), "build_ms": round(ix.last_build_ms, 2)},
        "plain": {"ms": t_plain, "queries_per_s": nq / (t_plain / 1e3)},
        "tree": {"ms": t_tree, "queries_per_s": nq / (t_tree / 1e3), "identical_to_plain": tree_same},
        "learned": res,
        "speedup_over_plain": t_plain / tr["ms"],
        "speedup_over_tree": t_tree / tr["ms"],
        "max_abs_err": tr["max_abs_err"],
        "batched_search": dict(search, hits=total,
                               learned_over_halving=search["halving"]["step_ms"] / search["learned"]["step_ms"],
                               learned_over_tree=search["tree"]["step_ms"] / search["learned"]["step_ms"]),
    }
    print(json.dumps(out))


def load_traffic():
    """HBM bytes per launch of the walk kernels from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_search_lognormal_10Mx10M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibrate_copy(dev):
    """One 1 GiB device-to-device copy: the reference kernel for the PMC byte counters
    (its FETCH_SIZE / WRITE_SIZE readings convert counter KiB to bytes)."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def cpu_baseline(s, e, qs, qe, sample, check=None):
    """Reference CPU path on a bounded sample of the same queries, all host cores: search_idxs
    and count.  With check = (offsets, positions, query numbers) of the GPU's descending
    output, the same CPU code also computes those queries' lists and they must match."""
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    sq, se = qs[:sample], qe[:sample]

    def median3(fn):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[1]

    O = oracle.Oracle()
    if oracle.Reference.available():
        R = oracle.Reference()
        m = oracle.RefMap(R, s, e)
        m.search_batch_total(sq[:10000], se[:10000], threads)        # warm-up
        t = median3(lambda: m.search_batch_total(sq, se, threads))
        tc = median3(lambda: m.count_batch(sq, se, threads))
        kind = "reference"
        desc = "si::IntervalMap<int,int>::search_idxs(s, e, std::vector&) and count(s, e) from the reference headers"
        lists = lambda q: np.asarray(m.search(qs[q], qe[q], 0), np.int64)
    else:
        ix = O.build_index(s, e)
        t = median3(lambda: O.search_batch(ix, sq, se, mode=1, threads=threads, lists=False))
        tc = median3(lambda: O.count_batch(ix, sq, se, threads=threads))
        kind = "port"
        desc = "oracle restatement of search_idxs and count"
        lists = lambda q: O.search(ix, qs[q], qe[q], 0).astype(np.int64)
    out = {"value": sample / t, "unit": "queries/s", "cores": threads, "kind": kind,
           "cpu_model": cpu_model(), "count_queries_per_s": sample / tc,
           "sample": f"first {sample} of the {qs.size} queries, median of 3; {desc}; "
                     f"static chunks over {threads} threads"}
    if check is not None:
        off, idx, picks = check
        bad = 0
        for q in picks:
            got = idx[int(off[q]):int(off[q + 1])].astype(np.int64)
            bad += int(not np.array_equal(got, lists(int(q))))
        out["checked_queries"] = int(len(picks))
        out["mismatched_queries"] = bad
    return out


def cpu_baseline_sampled(s, e, qs, qe, picks, gpu_lists):
    """The reference's own search_idxs (its headers, oracle/_ref; else the oracle's restatement)
    on the sampled queries picks of a workload, all host cores, timed (median of 3), and each
    picked query's descending list compared with the GPU's (gpu_lists: query -> positions)."""
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    sq, se = qs[picks], qe[picks]
    t0 = time.perf_counter()
    if oracle.Reference.available():
        m = oracle.RefMap(oracle.Reference(), s, e)
        kind, desc = "reference", "si::IntervalMap<int64_t,int>::search_idxs(s, e, std::vector&) from the reference headers"
        run = lambda: m.search_batch_total(sq, se, threads)
        lists = lambda k: np.asarray(m.search(sq[k], se[k], 0), np.int64)
    else:
        O = oracle.Oracle()
        ix = O.build_index(s, e)
        kind, desc = "port", "oracle restatement of search_idxs"
        run = lambda: O.search_batch(ix, sq, se, mode=1, threads=threads, lists=False)
        lists = lambda k: O.search(ix, sq[k], se[k], 0).astype(np.int64)
    build_s = time.perf_counter() - t0
    ts = []
    for _ in range(3):
        t1 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t1)
    t = sorted(ts)[1]
    bad = sum(int(not np.array_equal(gpu_lists[int(q)], lists(k))) for k, q in enumerate(picks))
    return {"value": len(picks) / t, "unit": "queries/s", "cores": threads, "kind": kind, "cpu_model": cpu_model(),
            "sample": f"{len(picks)} queries spread evenly over the {qs.size}, median of 3; {desc}; "
                      f"static chunks over {threads} threads (build {build_s:.1f} s, not timed)",
            "checked_queries": int(len(picks)), "mismatched_queries": int(bad)}


def verify_output(ix, tqs, tqe, s, e, off, idx, nq, stream):
    """Checks the timed configuration's own output on the device: counts against the closed
    form (torch.searchsorted over the starts and the sorted ends, every query has s <= e), the
    walk-form count against the offsets (two different device algorithms), every emitted
    position overlaps its query, every list strictly descending.  Returns a dict of results."""
    import torch
    from superintervals_amd import autotune
    dev = tqs.device
    res = {}
    torch.cuda.synchronize()
    cnt = (off[1:] - off[:-1])
    ss = torch.from_numpy(s).to(dev)
    se = torch.sort(torch.from_numpy(e).to(dev)).values
    closed = torch.searchsorted(ss, tqe, right=True) - torch.searchsorted(se, tqs, right=False)
    res["closed_form_counts"] = bool(torch.equal(closed.to(torch.int64), cnt))
    del se, closed
    saved = autotune.current()
    walk = torch.empty(nq, dtype=torch.int32, device=dev)
    try:
        autotune.apply({"count_rank": 0})
        ix.count_device(tqs, tqe, walk, stream=stream)
    finally:
        autotune.apply(saved)
    torch.cuda.synchronize()
    res["walk_counts_equal_offsets"] = bool(torch.equal(walk.to(torch.int64), cnt))
    del walk
    total = int(off[-1].item())
    qid = torch.repeat_interleave(torch.arange(nq, device=dev, dtype=torch.int32), cnt)
    pos = idx[:total].to(torch.int64)
    # positions index the built (sorted) arrays: equal starts are ordered by end descending
    arr = ix.download()
    ss = torch.from_numpy(arr["starts"]).to(dev)
    ends_t = torch.from_numpy(arr["ends"]).to(dev)
    ok = bool((ss[pos] <= tqe[qid]).all()) and bool((ends_t[pos] >= tqs[qid]).all())
    res["positions_overlap_query"] = ok
    same = qid[1:] == qid[:-1]
    res["lists_strictly_descending"] = bool((pos[1:][same] < pos[:-1][same]).all())
    del qid, pos, same, ends_t, ss
    torch.cuda.empty_cache()
    return res


def bench_serving(a):
    """Continuous batching over resident indexes: sets of log-uniform sizes (log-normal
    interval distributions), each an int32 index in a slot of a set table; Poisson request
    streams at several rates, each request a query batch of log-uniform size in [1, 1e6]
    against a random set; batches replay captured graphs per shape; results (descending
    position lists) return to pinned host memory.  Reports sustained queries/s and p50/p99
    request latency per rate."""
    import numpy as np
    import superintervals_amd as sia
    import torch
    from superintervals_amd import serving
    torch.cuda.init()
    rng = np.random.default_rng(11)
    sizes = serving.log_uniform_int(rng, 1_000_000, a.serving_max_n, a.serving_sets)
    t0 = time.perf_counter()
    sets = [sia.workload_intervals(2, int(n), seed=1000 + k, sorted=True) for k, n in enumerate(sizes)]
    gen_s = time.perf_counter() - t0
    print("[serving] generated %d sets (%d intervals) in %.1f s" % (len(sets), int(sizes.sum()), gen_s),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = serving.ResidentSets(sets, max_sets=max(64, len(sets)))
    pack_build_s = time.perf_counter() - t0
    print("[serving] built %d resident sets in %.1f s" % (len(sets), pack_build_s), file=sys.stderr, flush=True)
    del sets
    pool = 16_000_000
    pqs, pqe = sia.workload_queries(2, pool, seed=77)
    rows = []
    srv = None
    for rate in [float(x) for x in a.serving_rates.split(",")]:
        nreq = max(1, int(rate * a.serving_seconds))
        m = serving.log_uniform_int(rng, 1, 1_000_000, nreq)
        off = rng.integers(0, pool - m)
        tgt = rng.integers(0, len(sizes), nreq)
        reqs = [(int(tgt[i]), pqs[off[i]:off[i] + m[i]], pqe[off[i]:off[i] + m[i]], "search") for i in range(nreq)]
        arrivals = serving.poisson_arrivals(rate, nreq, seed=int(rate))
        if srv is None:
            srv = serving.Server(res, hbm_budget_bytes=64 << 30, streams=4, max_batch_queries=1 << 20,
                                 keep_results=False)
        srv.completed.clear()
        b1, b2 = srv.stats["batches1"], srv.stats["batches2"]
        lat, wall = serving.serve_stream(srv, reqs, arrivals, timeout_s=120.0)
        lat_ms = np.sort(lat) * 1e3
        rows.append({"rate_req_per_s": rate, "requests": nreq, "queries": int(m.sum()),
                     "wall_s": round(wall, 3), "queries_per_s": float(m.sum() / wall),
                     "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
                     "batches_stage1": srv.stats["batches1"] - b1, "batches_stage2": srv.stats["batches2"] - b2,
                     "graphs_captured_so_far": srv.stats["graphs"],
                     "max_hbm_in_use_gb": round(srv.stats["max_in_use"] / 2**30, 2)})
        print("[serving] rate %.0f/s: %d requests, %.3g q/s, p50 %.2f ms, p99 %.2f ms" %
              (rate, nreq, rows[-1]["queries_per_s"], rows[-1]["p50_ms"], rows[-1]["p99_ms"]),
              file=sys.stderr, flush=True)
    best = max(rows, key=lambda r: r["queries_per_s"])
    out = {
        "metric": "serving: sustained queries/s and p50/p99 request latency (resident indexes, Poisson arrivals)",
        "value": best["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "int32", "data": "synthetic (seeded log-normal generator)",
        "config": {"workload": "%d resident sets of %d..%d int32 intervals (log-uniform), one int32 index per slot "
                               "of a set table; request batches log-uniform in [1, 1e6] queries, batched across "
                               "sets, graph replays per shape; results to pinned host memory"
                               % (a.serving_sets, int(sizes.min()), int(sizes.max())),
                   "sets": a.serving_sets, "intervals_total": int(res.n), "host_generation_s": round(gen_s, 2),
                   "build_s": round(pack_build_s, 2), "streams": 4, "hbm_budget_gb": 64},
        "rates": rows,
    }
    print(json.dumps(out))


def bench_skewed(a):
    """Uneven int64 batch: 50M intervals (70 % uniform, 30 % from 200 Gaussian hotspots, 20 %
    of lengths log-uniform in [1e5, 1e7]) and 20M queries of which 1 % are wide.  The counts
    of all queries in one launch (walk with the heavy-query path, and the rank form), then
    search_idxs streamed in batches (the full CSR, ~1e11 positions, does not fit in HBM)."""
    import numpy as np
    import torch
    import superintervals_amd as sia
    from superintervals_amd import autotune
    dev = torch.device("cuda:0")
    torch.cuda.init()
    if a.calibrate_copy:
        calibrate_copy(dev)
    n, nq, bq = a.skewed_n, a.skewed_nq, a.skewed_batch
    t0 = time.perf_counter()
    s, e = sia.workload_intervals(3, n, seed=31, sorted=True)
    qs, qe = sia.workload_queries(3, nq, seed=32)
    gen_s = time.perf_counter() - t0
    print("[skewed] generated in %.1f s" % gen_s, file=sys.stderr, flush=True)
    ix = sia.DeviceIndex(np.int64).build(s, e)
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    stream = torch.cuda.current_stream()
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)
    saved = autotune.current()

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        e0, e1 = events(2)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # the plan of all 20M queries (sort by end, bounds and rank counts), warm
    t_plan = timed(lambda: ix.plan_device(tqs, tqe, plan, stream=stream))
    try:
        autotune.apply({"count_rank": 0, "heavy_ends32": 0})
        t_walk64 = timed(lambda: ix.count_planned(plan, nq, counts, stream=stream))
        autotune.apply({"heavy_ends32": saved["heavy_ends32"]})
        t_walk = timed(lambda: ix.count_planned(plan, nq, counts, stream=stream))
        c_walk = counts.clone()
        autotune.apply({"count_rank": 1})
        t_rank = timed(lambda: ix.count_planned(plan, nq, counts, stream=stream))
        same = bool(torch.equal(counts, c_walk))
        # count() from the queries, the whole call (plan with its rank counts + scatter)
        counts.zero_()
        t_call = timed(lambda: ix.count_device(tqs, tqe, counts, stream=stream))
        same_call = bool(torch.equal(counts, c_walk))
    finally:
        autotune.apply(saved)
    hits = int(c_walk.to(torch.int64).sum().item())
    wide = int((c_walk > 100_000).sum().item())
    print("[skewed] count: walk %.2f ms (int64 ends in the wave walk: %.2f ms), rank counts of the plan %.2f ms, "
          "count() whole call %.2f ms, identical %s/%s, %.3g hits"
          % (t_walk, t_walk64, t_rank, t_call, same, same_call, hits), file=sys.stderr, flush=True)
    # streamed search: batches of bq queries through plan -> offsets -> fill (descending)
    nb = (nq + bq - 1) // bq
    bplan = torch.empty(ix.plan_bytes(bq), dtype=torch.uint8, device=dev)
    off = torch.empty(bq + 1, dtype=torch.int64, device=dev)
    # the per-lane step budget before a walk is parked for the one-wave kernel, tuned on the
    # first batch: these walks are long but sparse, which the lane walk serves better
    m0 = min(nq, bq)
    ix.plan_device(tqs[:m0], tqe[:m0], bplan, stream=stream)
    ix.offsets_planned(bplan, m0, off, stream=stream)
    torch.cuda.synchronize()
    cap = int(int(off[m0].item()) * 1.1) + 1
    idx = torch.empty(cap, dtype=torch.int32, device=dev)
    sweep, csweep = {}, {}
    try:
        for ls in (32, 64, 256, 1024, 4096):
            autotune.apply({"lane_steps": ls})
            sweep[ls] = timed(lambda: ix.fill_planned(bplan, m0, off, idx, order="desc", stream=stream), reps=1)
            autotune.apply({"count_rank": 0})
            csweep[ls] = timed(lambda: ix.count_planned(bplan, m0, counts, stream=stream), reps=1)
            autotune.apply({"count_rank": saved["count_rank"]})
            print("[skewed] lane_steps %d: fill of one batch %.1f ms, count walk %.1f ms"
                  % (ls, sweep[ls], csweep[ls]), file=sys.stderr, flush=True)
    finally:
        autotune.apply(saved)
    best_ls = min(sweep, key=sweep.get)
    autotune.apply({"lane_steps": best_ls})
    # the wave walk's ends: the 32-bit copy or the int64 array, by the fill of one batch
    ends32_ms = {}
    try:
        for e32 in (1, 0):
            autotune.apply({"heavy_ends32": e32})
            ends32_ms[e32] = timed(lambda: ix.fill_planned(bplan, m0, off, idx, order="desc", stream=stream), reps=1)
    finally:
        autotune.apply({"heavy_ends32": saved["heavy_ends32"]})
    best_e32 = min(ends32_ms, key=ends32_ms.get)
    autotune.apply({"heavy_ends32": best_e32})
    print("[skewed] fill of one batch: 32-bit ends in the wave walk %.1f ms, int64 ends %.1f ms"
          % (ends32_ms[1], ends32_ms[0]), file=sys.stderr, flush=True)
    t_fill = 0.0
    tot = 0
    # a sample of the queries (spread over all batches) is kept for the oracle check below,
    # and every batch's counts are compared with the walk's of the whole-batch count
    sample = np.unique(np.linspace(0, nq - 1, a.skewed_check).astype(np.int64)) if a.skewed_check else np.zeros(0, np.int64)
    kept = {}
    counts_equal = True
    t0 = time.perf_counter()
    for b in range(nb):
        lo, hi = b * bq, min(nq, (b + 1) * bq)
        m = hi - lo
        ix.plan_device(tqs[lo:hi], tqe[lo:hi], bplan, stream=stream)
        ix.offsets_planned(bplan, m, off, stream=stream)
        torch.cuda.synchronize()
        t = int(off[m].item())
        if t > cap:
            idx = None
            torch.cuda.empty_cache()
            cap = int(t * 1.1) + 1
            idx = torch.empty(cap, dtype=torch.int32, device=dev)
        e0, e1 = events(2)
        e0.record(stream)
        ix.fill_planned(bplan, m, off, idx, order="desc", stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        t_fill += e0.elapsed_time(e1)
        tot += t
        counts_equal = counts_equal and bool(torch.equal((off[1:m + 1] - off[:m]).to(torch.int32), c_walk[lo:hi]))
        for q in sample[(sample >= lo) & (sample < hi)]:
            o0, o1 = int(off[q - lo].item()), int(off[q - lo + 1].item())
            kept[int(q)] = idx[o0:o1].cpu().numpy().astype(np.int64)
    wall = time.perf_counter() - t0
    autotune.apply(saved)
    assert tot == hits, (tot, hits)
    verified = {"streamed_total_equals_count": tot == hits, "batch_offsets_equal_walk_counts": counts_equal,
                "rank_counts_equal_walk": same, "count_call_equals_walk": same_call}
    cpu = None
    if len(sample) and not a.no_cpu:
        cpu = cpu_baseline_sampled(s, e, qs, qe, sample, kept)
        verified.update({"cpu_sample_queries": cpu["checked_queries"],
                         "cpu_sample_lists_equal": cpu["mismatched_queries"] == 0})
        print("[skewed] CPU check of %d sampled queries: %d mismatched" % (cpu["checked_queries"],
              cpu["mismatched_queries"]), file=sys.stderr, flush=True)
    del s, e
    print("[skewed] streamed search: %d batches, %.3g positions, fill %.1f ms, wall %.2f s" % (nb, tot, t_fill, wall),
          file=sys.stderr, flush=True)
    # roofline of the streamed fill: the bytes it must move (positions written, plan entry and
    # offset read per query) per second; the counter-measured traffic of one batch's fill
    # launches, when the summary of a PMC run is committed
    alg = 4 * tot + (2 * 8 + 4 + 8 + 8) * nq
    ach = alg / (t_fill / 1e3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_walk<long, desc fill> + k_walk_heavy<long, desc fill> (streamed fill)",
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": None,
            "basis": "achieved = (4 B per position written + 36 B of plan entry, offsets and query per query) / "
                     "fill time; traffic = FETCH_SIZE + WRITE_SIZE of one 1M-query batch's fill launches"}
    tp = os.path.join(ROOT, "profiles", "pmc_traffic_search_skewed_50Mx20M.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tr = json.load(f)
        per_batch = tr["fill_lane"]["bytes_per_launch"] + tr["fill_heavy"]["bytes_per_launch"]
        roof["traffic"] = per_batch
        roof["measured_GBps"] = round(per_batch * nb / (t_fill / 1e3) / 1e9, 1)
        roof["measured_frac"] = round(roof["measured_GBps"] / HBM_PEAK_GBS, 4)
        roof["traffic_profile"] = "profiles/pmc_traffic_search_skewed_50Mx20M.json"
    out = {"metric": "skewed int64 batch: count queries/s (one launch) and streamed search_idxs positions/s",
           "value": nq / (t_walk / 1e3), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "int64", "data": "synthetic (seeded hotspot-mixture generator, inputs resident in HBM)",
           "config": {"workload": "%dM int64 intervals (hotspot mixture, 20 %% long) x %dM queries (1 %% wide)" % (n // 10**6, nq // 10**6),
                      "intervals": n, "queries": nq, "search_batch": bq, "host_generation_s": round(gen_s, 2),
                      "build_ms": round(ix.last_build_ms, 2)},
           "plan_ms": t_plan, "count_walk_ms": t_walk, "count_walk_int64_ends_ms": t_walk64,
           "count_rank_ms": t_rank, "count_identical": same, "count_call_ms": t_call,
           "count_note": "plan_ms: sort by end, bounds and rank-form counts of all queries (warm); count_rank_ms: "
                         "the scatter of the plan's counts; count_call_ms: count() from the queries (plan + scatter); "
                         "count_walk_ms: the walk form on the planned batch",
           "hits": hits, "queries_over_1e5_hits": wide,
           "lane_steps_sweep_ms": {str(k): v for k, v in sweep.items()}, "lane_steps": best_ls,
           "lane_steps_count_walk_ms": {str(k): v for k, v in csweep.items()},
           "fill_batch_ms_ends32": ends32_ms.get(1), "fill_batch_ms_int64_ends": ends32_ms.get(0),
           "search": {"batches": nb, "fill_ms_total": t_fill, "positions_per_s": tot / (t_fill / 1e3),
                      "wall_s_including_plan_offsets_sync": wall},
           "roofline": roof, "cpu_baseline": cpu, "verified": verified}
    print(json.dumps(out))
    if not all(v for k, v in verified.items() if isinstance(v, bool)):
        print("[skewed] VERIFICATION FAILED: %s" % verified, file=sys.stderr, flush=True)
        sys.exit(3)


def bench_graph(a):
    """Launch-bound small batches on the 10M-interval index: search_idxs of graph_nq queries
    per batch, eager (plan with and without the digit read-back) against the replay of one
    captured graph of plan -> offsets -> bounded fill."""
    import numpy as np
    import torch
    import superintervals_amd as sia
    from superintervals_amd import autotune
    dev = torch.device("cuda:0")
    torch.cuda.init()
    s, e = sia.workload_intervals(2, a.n, seed=1, sorted=True)
    ix = sia.DeviceIndex(np.int32).build(s, e)
    nq = a.graph_nq
    qs, qe = sia.workload_queries(2, nq, seed=5)
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    ix.plan_device(tqs, tqe, plan, stream=stream)
    ix.offsets_planned(plan, nq, off, stream=stream)
    torch.cuda.synchronize()
    cap = int(off[-1].item())
    idx = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    reps = 200
    saved = autotune.current()

    def eager():
        ix.plan_device(tqs, tqe, plan, stream=stream)
        ix.offsets_planned(plan, nq, off, stream=stream)
        ix.fill_planned(plan, nq, off, idx, capacity=cap, stream=stream)

    def timed(fn):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e6

    res = {}
    try:
        autotune.apply({"plan_sync": 1})
        res["eager_plan_readback_us"] = timed(eager)
        autotune.apply({"plan_sync": 0})
        res["eager_no_readback_us"] = timed(eager)
        ref = idx.clone()
        g = torch.cuda.CUDAGraph()
        cs = torch.cuda.Stream()
        with torch.cuda.graph(g, stream=cs):
            st = torch.cuda.current_stream()
            ix.plan_device(tqs, tqe, plan, stream=st)
            ix.offsets_planned(plan, nq, off, stream=st)
            ix.fill_planned(plan, nq, off, idx, capacity=cap, stream=st)
        res["graph_replay_us"] = timed(g.replay)
        res["graph_output_identical"] = bool(torch.equal(idx, ref))
        for ls in (64, 16):
            autotune.apply({"lane_steps": ls})
            res["eager_no_readback_lane_steps%d_us" % ls] = timed(eager)
            res["lane_steps%d_output_identical" % ls] = bool(torch.equal(idx, ref))
    finally:
        autotune.apply(saved)
    out = {"metric": "search_idxs latency per small batch: eager vs captured graph (10M idx)",
           "value": nq / (res["graph_replay_us"] / 1e6), "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "int32", "data": "synthetic (seeded log-normal generator)",
           "config": {"workload": "batches of %d queries on 10M int32 intervals, back to back" % nq,
                      "intervals": a.n, "queries_per_batch": nq, "hits_per_batch": cap, "reps": reps},
           **res}
    print(json.dumps(out))


def tuning_arg(a):
    """The --tuning knobs (launch-choice keys only), or None."""
    if not a.tuning:
        return None
    from superintervals_amd import autotune
    knobs = json.loads(a.tuning)
    return {k: v for k, v in knobs.items() if k in autotune.KNOBS}


def walk_once(a):
    """One launch of each kernel of the 10M x 10M step with the --tuning knobs: plan, offsets
    (rank count + scan), the rank count alone, the count walk, the descending fill."""
    import numpy as np
    import torch
    import superintervals_amd as sia
    from superintervals_amd import autotune
    dev = torch.device("cuda:0")
    torch.cuda.init()
    if a.calibrate_copy:
        calibrate_copy(dev)
    knobs = tuning_arg(a)
    if knobs:
        autotune.apply(knobs)
    s, e = sia.workload_intervals(2, a.n, seed=1, sorted=True)
    qs, qe = sia.workload_queries(2, a.nq, seed=2)
    ix = sia.DeviceIndex(np.int32).build(s, e)
    nq = a.nq
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)
    ix.plan_device(tqs, tqe, plan)
    ix.offsets_planned(plan, nq, off)
    torch.cuda.synchronize()
    idx = torch.empty(int(off[-1].item()), dtype=torch.int32, device=dev)
    autotune.apply({"count_rank": 1})
    ix.count_planned(plan, nq, counts)
    autotune.apply({"count_rank": 0})
    ix.count_planned(plan, nq, counts)
    ix.fill_planned(plan, nq, off, idx)
    torch.cuda.synchronize()
    print(json.dumps({"walk_once": True, "queries": nq, "hits": int(off[-1].item()), "tuning": autotune.current()}))


def cpu_baseline_single(s, e, qs, qe, per_call):
    """The reference's own count and search_idxs (its headers, oracle/_ref), one query per
    call through ctypes, on one host thread; None when that build is absent."""
    import oracle
    if not oracle.Reference.available():
        return None
    m = oracle.RefMap(oracle.Reference(), s, e)
    return {"count": round(per_call(lambda x, y: m.count(x, y), qs, qe), 2),
            "search_idxs": round(per_call(lambda x, y: m.search(x, y, 1), qs, qe), 2)}


def bench_single(a):
    """Latency of single-query calls, the path of the C-ABI's countOverlaps / searchIdxs /
    searchValues / upperBound / anyOverlaps: one wave on the index's scratch, results in
    pinned host memory.  Through the C-ABI on 100K intervals (geometric lengths) and through
    the host-buffer calls on the 10M-interval log-normal index; per call, ctypes included;
    the reference's own count and search_idxs per call through the same ctypes layer."""
    import ctypes as C
    import numpy as np
    import superintervals_amd as sia
    import superintervals_amd._native as N
    L = N.lib
    out = {"metric": "single-query call latency (microseconds per call, ctypes included)", "unit": "us",
           "higher_is_better": False, "n_gpus": 1, "data": "synthetic (seeded generators)"}
    reps = 20000

    def per_call(fn, qs, qe):
        for k in range(200):
            fn(int(qs[k]), int(qe[k]))
        t0 = time.perf_counter()
        for k in range(reps):
            fn(int(qs[k]), int(qe[k]))
        return (time.perf_counter() - t0) / reps * 1e6

    # the C-ABI on a cSuperIntervals of 100K intervals
    s, e = sia.workload_intervals(1, 100_000, seed=1, sorted=True)
    qs, qe = sia.workload_queries(1, reps, seed=2)
    h = L.createSuperIntervals()
    for i in range(s.size):
        L.addInterval(h, int(s[i]), int(e[i]), i)
    L.indexSuperIntervals(h)
    res = N.cIndexResult()

    def search_idxs(x, y):
        L.clearIndexResult(C.byref(res))
        L.searchIdxs(h, x, y, C.byref(res))

    capi = {"countOverlaps": per_call(lambda x, y: L.countOverlaps(h, x, y), qs, qe),
            "searchIdxs": per_call(search_idxs, qs, qe),
            "anyOverlaps": per_call(lambda x, y: L.anyOverlaps(h, x, y), qs, qe),
            "upperBound": per_call(lambda x, y: L.upperBound(h, y), qs, qe)}
    L.destroyIndexResult(C.byref(res))
    L.destroySuperIntervals(h)
    out["capi_100k"] = {k: round(v, 2) for k, v in capi.items()}
    print("[single] C-ABI, 100K intervals: %s" % out["capi_100k"], file=sys.stderr, flush=True)
    # host-buffer calls on the 10M-interval index
    s, e = sia.workload_intervals(2, a.n, seed=1, sorted=True)
    qs, qe = sia.workload_queries(2, reps, seed=2)
    ix = sia.DeviceIndex(np.int32).build(s, e)
    one_s = np.zeros(1, np.int32)
    one_e = np.zeros(1, np.int32)
    cnt = np.zeros(1, np.uint32)
    off = np.zeros(2, np.uint64)
    buf = np.zeros(1 << 16, np.uint32)
    tot = C.c_uint64(0)

    def count(x, y):
        one_s[0], one_e[0] = x, y
        L.siCountHost(ix.handle, one_s.ctypes.data, one_e.ctypes.data, 1, cnt.ctypes.data)

    def search(x, y):
        one_s[0], one_e[0] = x, y
        L.siSearchHost(ix.handle, one_s.ctypes.data, one_e.ctypes.data, 1, 0, off.ctypes.data, buf.ctypes.data,
                       None, buf.size, C.byref(tot))

    dev = {"count": per_call(count, qs, qe), "search_desc": per_call(search, qs, qe)}
    out["host_calls_10m"] = {k: round(v, 2) for k, v in dev.items()}
    print("[single] host-buffer calls, 10M intervals: %s" % out["host_calls_10m"], file=sys.stderr, flush=True)
    ref = cpu_baseline_single(s, e, qs, qe, per_call)
    if ref:
        out["reference_cpu_10m"] = ref
        print("[single] reference (CPU, 1 thread), 10M intervals: %s" % ref, file=sys.stderr, flush=True)
    out["value"] = out["host_calls_10m"]["count"]
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    if a.learned:
        return bench_learned(a)
    if a.serving:
        return bench_serving(a)
    if a.walk_once:
        return walk_once(a)
    if a.single:
        return bench_single(a)
    if a.graph:
        return bench_graph(a)
    if a.skewed:
        return bench_skewed(a)
    import numpy as np
    import torch
    import superintervals_amd as sia

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.calibrate_copy:
        calibrate_copy(dev)
    s, e = sia.workload_intervals(2, a.n, seed=1, sorted=True)
    qs, qe = sia.workload_queries(2, a.nq, seed=2)
    ix = sia.DeviceIndex(np.int32).build(s, e)
    build_ms = ix.last_build_ms
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    nq = a.nq
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    ix.plan_device(tqs, tqe, plan, stream=stream)
    ix.offsets_planned(plan, nq, off, stream=stream)
    torch.cuda.synchronize()
    total = int(off[-1].item())
    idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)

    # reference-walk statistics for the algorithmic byte model
    hits, V, J, found = ix.walk_stats_device(tqs, tqe)

    # launch choices: autotuned on this index and batch before the timed region
    tuning = None
    if tuning_arg(a):
        from superintervals_amd import autotune
        tuning = tuning_arg(a)
        autotune.apply(tuning)
    elif not a.no_tune:
        from superintervals_amd import autotune
        tuning = autotune.load_or_tune(ix, plan, nq, off, idx, total, stream=stream, retune=True,
                                       qs_t=tqs, qe_t=tqe, verbose=a.verbose)

    def step(evs=None):
        if evs:
            evs[0].record(stream)
        ix.plan_device(tqs, tqe, plan, stream=stream)
        if evs:
            evs[1].record(stream)
        ix.offsets_planned(plan, nq, off, stream=stream)
        if evs:
            evs[2].record(stream)
        ix.fill_planned(plan, nq, off, idx, order="desc", stream=stream)
        if evs:
            evs[3].record(stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(a.steps)]
    t0 = time.perf_counter()
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_plan = sum(ev[0].elapsed_time(ev[1]) for ev in evs) / a.steps
    t_off = sum(ev[1].elapsed_time(ev[2]) for ev in evs) / a.steps
    t_fill = sum(ev[2].elapsed_time(ev[3]) for ev in evs) / a.steps
    # batch latency: plan -> offsets -> fill of one batch of nq queries, per timed step
    lat_ms = np.array([ev[0].elapsed_time(ev[3]) for ev in evs])
    ms_step = wall * 1e3 / a.steps

    # the timed configuration's own output (the last timed step's offsets and lists)
    checks = verify_output(ix, tqs, tqe, s, e, off, idx, nq, stream)
    picks = np.unique(np.linspace(0, nq - 1, 2000).astype(np.int64))
    off_h = idx_h = None
    if not a.no_cpu:
        off_h = off.cpu().numpy()
        idx_h = idx[:total].cpu().numpy()

    # count walk alone on the planned batch (one launch per pass); the walk form is forced
    # here whatever the tuner chose, since the byte model below is the walk's
    from superintervals_amd import autotune
    tuned_count_rank = autotune.current()["count_rank"]
    autotune.apply({"count_rank": 0})
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(2):
        ix.count_planned(plan, nq, counts, stream=stream)
    c0.record(stream)
    for _ in range(a.steps):
        ix.count_planned(plan, nq, counts, stream=stream)
    c1.record(stream)
    torch.cuda.synchronize()
    t_count = c0.elapsed_time(c1) / a.steps
    # rank form of count (upper_bound from the plan + one search of the sorted ends)
    autotune.apply({"count_rank": 1})
    ix.count_planned(plan, nq, counts, stream=stream)
    c0.record(stream)
    for _ in range(a.steps):
        ix.count_planned(plan, nq, counts, stream=stream)
    c1.record(stream)
    torch.cuda.synchronize()
    autotune.apply({"count_rank": tuned_count_rank})
    t_count_rank = c0.elapsed_time(c1) / a.steps
    # count() from the queries, the whole call (plan + the counts it keeps, scattered)
    ix.count_device(tqs, tqe, counts, stream=stream)
    c0.record(stream)
    for _ in range(a.steps):
        ix.count_device(tqs, tqe, counts, stream=stream)
    c1.record(stream)
    torch.cuda.synchronize()
    t_count_call = c0.elapsed_time(c1) / a.steps
    checks["count_call_equals_offsets"] = bool(torch.equal(counts.to(torch.int64), off[1:] - off[:-1]))

    # Algorithmic bytes (per query: c = coordinate bytes, V/J = the reference walk's ends
    # reads and branch jumps, h = hits).  Each launch is credited with its own work only:
    #   count walk  12 + c*V + 4*J + 4          (plan entry q/s/idx, walk, count)
    #   fill walk   12 + c*V + 4*J + 4*h + 8    (plan entry, walk, positions, offset)
    #   whole step  2c + c*ceil(log2 N) + c*V + 4*J + 4 + 4*h + 8  (the search_idxs figure)
    c = 4
    logn = math.ceil(math.log2(a.n))
    b_count_walk = nq * (12 + 4) + V * c + 4 * J
    b_fill = nq * (12 + 8) + V * c + 4 * J + 4 * hits
    b_search = nq * (2 * c + 4 + 8) + found * c * logn + V * c + 4 * J + 4 * hits
    kernels = {
        "count": {"ms": t_count, "alg_bytes": b_count_walk, "GBps": b_count_walk / t_count / 1e6},
        "count_rank": {"ms": t_count_rank, "queries_per_s": nq / (t_count_rank / 1e3)},
        "count_call": {"ms": t_count_call, "queries_per_s": nq / (t_count_call / 1e3),
                       "note": "count() from the queries, the whole call: plan (sort, bounds, rank counts) + scatter"},
        "plan": {"ms": t_plan},
        "search_offsets": {"ms": t_off},
        "search_fill": {"ms": t_fill, "alg_bytes": b_fill, "GBps": b_fill / t_fill / 1e6},
        "search_total": {"ms": t_plan + t_off + t_fill, "alg_bytes": b_search,
                         "GBps": b_search / (t_plan + t_off + t_fill) / 1e6},
    }
    dom = max(("plan", t_plan), ("search_offsets", t_off), ("search_fill", t_fill), key=lambda x: x[1])[0]
    ach = b_fill / t_fill / 1e6
    tr = load_traffic()
    roof = {"bound": "hbm", "kernel": "k_walk<int, desc fill> (search_idxs fill pass, one launch over all queries)",
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": tr["fill"]["bytes_per_launch"] if tr else None,
            "basis": "achieved = algorithmic bytes per launch (per query: 12 B plan entry + c*V ends + 4*J "
                     "branch reads of the reference walk + 4*h positions + 8 B offset) / launch time; the "
                     "walk serves most ends reads from its LDS window, so the counter-measured HBM traffic "
                     "(traffic, measured_GBps) is below the algorithmic bytes"}
    if tr:
        roof["measured_GBps"] = round(tr["fill"]["bytes_per_launch"] / t_fill / 1e6, 1)
        roof["measured_frac"] = round(roof["measured_GBps"] / HBM_PEAK_GBS, 4)
        roof["traffic_profile"] = tr.get("profile", "profiles/pmc_traffic_search_lognormal_10Mx10M.json")
        kernels["count"]["hbm_bytes_pmc"] = tr["count"]["bytes_per_launch"]
        kernels["search_fill"]["hbm_bytes_pmc"] = tr["fill"]["bytes_per_launch"]
        if "count_rank" in tr:
            kernels["count_rank"]["hbm_bytes_pmc"] = tr["count_rank"]["bytes_per_launch"]
            kernels["count_rank"]["measured_GBps"] = round(tr["count_rank"]["bytes_per_launch"] / t_count_rank / 1e6, 1)

    out = {
        "metric": "overlap queries/sec + achieved HBM GB/s (10M idx x 10M qry); p50/p99 batch latency",
        "value": nq / (ms_step / 1e3),
        "unit": "queries/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded log-normal generator, inputs resident in HBM)",
        "config": {"workload": "search_idxs two-pass, 10M int32 intervals x 10M queries (log-normal lengths)",
                   "intervals": a.n, "queries": nq, "hits": total, "order": "descending",
                   "build_ms": round(build_ms, 2), "dominant_search_kernel": dom,
                   "launch": tuning if tuning else "library defaults"},
        "roofline": roof,
        "kernels": kernels,
        # count() as tuned (walk or rank form), and the walk alone
        "batch_latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                             "batch_queries": nq, "batches": int(lat_ms.size)},
        "count_queries_per_s": nq / ((t_count_rank if tuned_count_rank else t_count) / 1e3),
        "count_walk_queries_per_s": nq / (t_count / 1e3),
        "walk": {"hits": hits, "ends_reads": V, "branch_reads": J, "found": found},
    }
    if not a.no_cpu:
        cb = cpu_baseline(s, e, qs, qe, min(a.cpu_sample, nq), check=(off_h, idx_h, picks))
        checks["cpu_sample_lists_equal"] = cb.pop("mismatched_queries") == 0
        checks["cpu_sample_queries"] = cb.pop("checked_queries")
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    out["verified"] = checks
    print(json.dumps(out))
    bad = [k for k, v in checks.items() if v is False]
    if bad:
        print("bench output check failed: %s" % ", ".join(bad), file=sys.stderr, flush=True)
        sys.exit(3)


if __name__ == "__main__":
    main()
