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
    return p.parse_args()


def load_traffic():
    """HBM bytes per launch of the walk kernels from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(s, e, qs, qe, sample):
    """Reference CPU path on a bounded sample of the same queries, all host cores."""
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    sq, se = qs[:sample], qe[:sample]
    if oracle.Reference.available():
        R = oracle.Reference()
        m = oracle.RefMap(R, s, e)
        m.search_batch_total(sq[:10000], se[:10000], threads)        # warm-up
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            m.search_batch_total(sq, se, threads)
            ts.append(time.perf_counter() - t0)
        kind = "reference"
        desc = "si::IntervalMap<int,int>::search_idxs(s, e, std::vector&) from the reference headers"
    else:
        O = oracle.Oracle()
        ix = O.build_index(s, e)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.search_batch(ix, sq, se, mode=1, threads=threads, lists=False)
            ts.append(time.perf_counter() - t0)
        kind = "port"
        desc = "oracle restatement of search_idxs"
    t = sorted(ts)[1]
    return {"value": sample / t, "unit": "queries/s", "cores": threads, "kind": kind,
            "sample": f"first {sample} of the {qs.size} queries, median of 3; {desc}; "
                      f"static chunks over {threads} threads"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    import superintervals_amd as sia

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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
    if not a.no_tune:
        from superintervals_amd import autotune
        tuning = autotune.load_or_tune(ix, plan, nq, off, idx, total, stream=stream, retune=True,
                                       verbose=a.verbose)

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
    ms_step = wall * 1e3 / a.steps

    # count kernel alone on the planned batch (one launch per pass)
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
    from superintervals_amd import autotune
    autotune.apply({"count_rank": 1})
    ix.count_planned(plan, nq, counts, stream=stream)
    c0.record(stream)
    for _ in range(a.steps):
        ix.count_planned(plan, nq, counts, stream=stream)
    c1.record(stream)
    torch.cuda.synchronize()
    autotune.apply({"count_rank": 0})
    t_count_rank = c0.elapsed_time(c1) / a.steps

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
            "traffic": tr["fill"]["bytes_per_launch"] if tr else None}
    if tr:
        kernels["count"]["hbm_bytes_pmc"] = tr["count"]["bytes_per_launch"]
        kernels["search_fill"]["hbm_bytes_pmc"] = tr["fill"]["bytes_per_launch"]

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
        "count_queries_per_s": nq / (t_count / 1e3),
        "walk": {"hits": hits, "ends_reads": V, "branch_reads": J, "found": found},
    }
    if not a.no_cpu:
        cb = cpu_baseline(s, e, qs, qe, min(a.cpu_sample, nq))
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
