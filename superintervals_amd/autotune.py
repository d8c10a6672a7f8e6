"""Autotuner for the walk kernels' launch choices.

The kernels' results never depend on these choices (tile size, resident workgroups per
CU for the count and the fill pass, the per-lane step budget before a long walk is
parked for the one-wave-per-query kernel, walk or rank form of the count); only their
speed does.  tune() times every candidate on the caller's own index and planned batch
with HIP events: first tile x workgroups for count + fill, then, with those fixed, the
lane step budget by the fill's time, the count form by the count's and, given the
queries, the plan's constant-digit check by the plan's and the unsorted low end digit
by plan + fill.  Results are
cached per workload signature (coordinate width, log2 of the index size, log2 of the
batch size, log2 of the mean hits per query) in a JSON file, so later runs of the same
shape reuse them.
"""
from __future__ import annotations

import json
import math
import os
import time

from . import _native as N

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "superintervals_amd", "tune.json")
KNOBS = ("tile", "count_wgs", "fill_wgs", "count_rank", "offsets_rank", "lane_steps", "plan_sync", "ub_tree",
         "plan_skip_low", "heavy_ends32", "range_skip", "ub_learned", "learned_shrink", "long_walk", "long_len")
CANDIDATE_TILES = (512, 1024, 2048, 4096)
CANDIDATE_WGS = (1, 0)     # 0 = the occupancy limit
CANDIDATE_LANE_STEPS = (64, 256, 1024)


def current() -> dict:
    return {k: int(N.lib.siTuneGet(k.encode())) for k in KNOBS}


def knobs_of(record: dict) -> dict:
    """The knob entries of a tuning record (which also holds measured times)."""
    return {k: record[k] for k in KNOBS if k in record}


def apply(params: dict) -> None:
    unknown = set(params) - set(KNOBS)
    if unknown:
        raise KeyError("unknown tuning knob(s): %s" % ", ".join(sorted(unknown)))
    for k in KNOBS:
        if k in params:
            N.check(N.lib.siTuneSet(k.encode(), int(params[k])), f"siTuneSet({k})")


def signature(coord_bytes: int, n: int, nq: int, hits: int) -> str:
    mean = hits / max(nq, 1)
    return "c%d_n%d_q%d_h%d" % (coord_bytes, int(math.log2(max(n, 1))), int(math.log2(max(nq, 1))),
                                int(math.log2(max(mean, 1.0))))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time_ms(fn, reps, stream):
    """Mean ms of fn() over reps calls on stream (HIP events), after one warm call."""
    import torch
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(ix, plan_t, nq, offsets_t, idx_t, stream=None, reps=3, tiles=CANDIDATE_TILES,
         wgs=CANDIDATE_WGS, lane_steps=CANDIDATE_LANE_STEPS, qs_t=None, qe_t=None, verbose=False):
    """Time count + fill for every candidate on a planned batch; apply and return the best.

    ix is a DeviceIndex, plan_t a planned batch (plan_device), offsets_t the batch's
    offsets (offsets_planned already run) and idx_t an output buffer of offsets[nq]
    entries.  With the batch's query tensors (qs_t, qe_t) the plan's constant-digit check
    (plan_sync: a digit histogram and one host read-back, which pay off when key digits
    are constant, as in int64 coordinates) is timed both ways too; the plan is then redone
    into plan_t.  Returns (best params, table of measured ms)."""
    import torch
    stream = stream or torch.cuda.current_stream()
    counts = torch.empty(nq, dtype=torch.int32, device=offsets_t.device)
    saved = current()
    count = lambda: ix.count_planned(plan_t, nq, counts, stream=stream)
    fill = lambda: ix.fill_planned(plan_t, nq, offsets_t, idx_t, stream=stream)
    results = []
    try:
        # the tile and the workgroup counts are judged on the walk form of count (the rank
        # form reads neither); the form itself is chosen afterwards
        apply({"count_rank": 0})
        for tile in tiles:
            for cw in wgs:
                for fw in wgs:
                    apply({"tile": tile, "count_wgs": cw, "fill_wgs": fw})
                    tc, tf = _time_ms(count, reps, stream), _time_ms(fill, reps, stream)
                    results.append({"tile": tile, "count_wgs": cw, "fill_wgs": fw, "count_ms": tc, "fill_ms": tf})
                    if verbose:
                        print("tune tile=%d count_wgs=%d fill_wgs=%d: count %.3f ms fill %.3f ms"
                              % (tile, cw, fw, tc, tf), flush=True)
        # count and fill share the tile; pick the tile by their sum, the wgs of each by its own time
        by_tile = {}
        for r in results:
            c = by_tile.setdefault(r["tile"], {"count": r, "fill": r})
            if r["count_ms"] < c["count"]["count_ms"]:
                c["count"] = r
            if r["fill_ms"] < c["fill"]["fill_ms"]:
                c["fill"] = r
        tile = min(by_tile, key=lambda t: by_tile[t]["count"]["count_ms"] + by_tile[t]["fill"]["fill_ms"])
        best = {"tile": tile, "count_wgs": by_tile[tile]["count"]["count_wgs"],
                "fill_wgs": by_tile[tile]["fill"]["fill_wgs"],
                "count_ms": by_tile[tile]["count"]["count_ms"], "fill_ms": by_tile[tile]["fill"]["fill_ms"]}
        apply(knobs_of(best))
        # the lane step budget (shared by the count walk and the fills), by the fill's time
        for ls in lane_steps:
            apply({"lane_steps": ls})
            tf = _time_ms(fill, reps, stream)
            results.append(dict(knobs_of(best), lane_steps=ls, fill_ms=tf))
            if verbose:
                print("tune lane_steps=%d: fill %.3f ms" % (ls, tf), flush=True)
            if "lane_steps" not in best or tf < best["fill_ms"]:
                best["lane_steps"], best["fill_ms"] = ls, tf
        apply({"lane_steps": best["lane_steps"]})
        # the count's form: the walk (best for short lists) or the rank form (O(log n))
        for cr in (0, 1):
            apply({"count_rank": cr})
            tc = _time_ms(count, reps, stream)
            results.append(dict(knobs_of(best), count_rank=cr, count_ms=tc))
            if verbose:
                print("tune count_rank=%d: count %.3f ms" % (cr, tc), flush=True)
            if "count_rank" not in best or tc < best["count_ms"]:
                best["count_rank"], best["count_ms"] = cr, tc
        if qs_t is not None and qe_t is not None:
            plan = lambda: ix.plan_device(qs_t, qe_t, plan_t, stream=stream)
            for ps in (1, 0):
                apply({"plan_sync": ps})
                tp = _time_ms(plan, reps, stream)
                results.append(dict(knobs_of(best), plan_sync=ps, plan_ms=tp))
                if verbose:
                    print("tune plan_sync=%d: plan %.3f ms" % (ps, tp), flush=True)
                if "plan_sync" not in best or tp < best["plan_ms"]:
                    best["plan_sync"], best["plan_ms"] = ps, tp
            # the lowest end digit left unsorted: one radix pass fewer, but coarser tiles;
            # judged by plan + fill (the offsets do not depend on the plan's order)
            apply({"plan_sync": best["plan_sync"]})
            base = None
            for psl in (0, 1):
                apply({"plan_skip_low": psl})
                tp = _time_ms(plan, reps, stream)
                tf = _time_ms(fill, reps, stream)
                results.append(dict(knobs_of(best), plan_skip_low=psl, plan_ms=tp, fill_ms=tf))
                if verbose:
                    print("tune plan_skip_low=%d: plan %.3f ms fill %.3f ms" % (psl, tp, tf), flush=True)
                if base is None or tp + tf < base:
                    base = tp + tf
                    best["plan_skip_low"], best["plan_ms"], best["fill_ms"] = psl, tp, tf
            apply({"plan_skip_low": best["plan_skip_low"]})
            plan()   # plan_t holds the batch planned with the chosen knobs
    finally:
        apply(saved)
    apply(knobs_of(best))
    return best, results


def load_or_tune(ix, plan_t, nq, offsets_t, idx_t, hits, cache=DEFAULT_CACHE, stream=None, retune=False,
                 qs_t=None, qe_t=None, verbose=False):
    """Apply cached choices for this workload signature, tuning (and caching) if absent."""
    sig = signature(ix.dtype.itemsize, ix.n, nq, hits)
    table = _load(cache)
    if not retune and sig in table:
        apply(knobs_of(table[sig]))
        return dict(table[sig], cached=True, signature=sig)
    t0 = time.perf_counter()
    best, _ = tune(ix, plan_t, nq, offsets_t, idx_t, stream=stream, qs_t=qs_t, qe_t=qe_t, verbose=verbose)
    best = dict(best, tuned_s=round(time.perf_counter() - t0, 2))
    table[sig] = best
    try:
        _save(cache, table)
    except OSError:
        pass
    return dict(best, cached=False, signature=sig)
