"""Writes the golden fixtures under tests/golden/ from the reference's own code.

Runs in the development container only: it needs oracle/_ref/libsiref.so, which
oracle/Makefile compiles in place from the reference checkout
(src/superintervals.hpp and src/c_superintervals.h of kcleal/superintervals v1.0.0).
The fixtures hold inputs (or the SHA-256 of generated inputs) and the reference's
outputs: sorted arrays, branch, counts, upper_bound, has_overlaps, coverage and
order-sensitive per-query hashes of search_idxs (both orders) and searchValues.

    python tests/golden/make_golden.py                 (every fixture)
    python tests/golden/make_golden.py --extras-only   (extras.npz only)
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import superintervals_amd as sia  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
NPREFIX = 2048     # per-query hashes / values kept explicitly for the first queries
NLISTS = 64        # explicit index lists kept for the first queries


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def cases():
    """(name, meta, starts, ends, qs, qe, store_inputs)"""
    s, e = sia.workload_intervals(1, 100_000, seed=1, sorted=True)
    qs, qe = sia.workload_queries(1, 100_000, seed=2)
    yield "c1_sorted", dict(config=1, n=100_000, nq=100_000, seed=1, qseed=2, sorted=1, transform="none"), s, e, qs, qe, False

    s, e = sia.workload_intervals(1, 100_000, seed=3, sorted=False)
    qs, qe = sia.workload_queries(1, 100_000, seed=4)
    yield "c1_shuffled", dict(config=1, n=100_000, nq=100_000, seed=3, qseed=4, sorted=0, transform="none"), s, e, qs, qe, False

    # int64 coordinates beyond 2^32, unsorted input
    s, e = sia.workload_intervals(1, 100_000, seed=5, sorted=False, coord_dtype=np.int64)
    qs, qe = sia.workload_queries(1, 100_000, seed=6, coord_dtype=np.int64)
    off = np.int64(1) << 33
    yield "c1_i64", dict(config=1, n=100_000, nq=100_000, seed=5, qseed=6, sorted=0, transform="plus2^33"), \
        s + off, e + off, qs + off, qe + off, False

    # dense exact duplicates (stable payload order), unsorted input; inputs stored
    rng = np.random.default_rng(7)
    n = 20_000
    ds = rng.integers(0, 2_000, n).astype(np.int32)
    de = (ds + rng.integers(0, 6, n)).astype(np.int32)
    dq = rng.integers(0, 2_000, 20_000).astype(np.int32)
    dqe = (dq + rng.integers(0, 4, 20_000)).astype(np.int32)
    yield "dups", dict(config=0, n=n, nq=20_000, sorted=0, transform="none"), ds, de, dq, dqe, True

    s, e = sia.workload_intervals(2, 1_000_000, seed=8, sorted=True)
    qs, qe = sia.workload_queries(2, 100_000, seed=9)
    yield "c2_1m", dict(config=2, n=1_000_000, nq=100_000, seed=8, qseed=9, sorted=1, transform="none"), s, e, qs, qe, False


def csr(lists, dtype=np.int64):
    off = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.uint64)
    flat = np.concatenate(lists).astype(dtype) if any(len(x) for x in lists) else np.zeros(0, dtype)
    return off, flat


def extras(R):
    """tests/golden/extras.npz: reference outputs for upper_bound_range, an index holding
    intervals with start > end, and the set operations (C and C++).  Inputs are stored."""
    rec = {}
    rng = np.random.default_rng(2024)
    # upper_bound_range (superintervals.hpp:524-543) on built maps, int32 and int64
    for sfx, dt, off in (("i32", np.int32, 0), ("i64", np.int64, np.int64(1) << 33)):
        n = 3000
        s = np.sort(rng.integers(0, 200_000, n)).astype(dt) + dt(off)
        e = (s + rng.integers(0, 500, n)).astype(dt)
        m = oracle.RefMap(R, s, e)
        a = m.arrays()
        k = 6_000
        right = rng.integers(0, n, k)
        left = np.maximum(right - rng.integers(0, 3, k) * rng.integers(0, 60, k), 0)
        pick = rng.integers(0, n, k)
        value = np.where(rng.random(k) < 0.5, a["starts"][pick],
                         rng.integers(-1000, 201_000, k).astype(dt) + dt(off)).astype(dt)
        out = np.array([m.upper_bound_range(v, l_, r_) for v, l_, r_ in zip(value, left, right)], np.int64)
        rec[f"ubr_{sfx}_starts"] = a["starts"]
        rec[f"ubr_{sfx}_value"], rec[f"ubr_{sfx}_left"], rec[f"ubr_{sfx}_right"] = value, left.astype(np.int64), right.astype(np.int64)
        rec[f"ubr_{sfx}_out"] = out
    # an index in which 15 % of the intervals have start > end (unsorted input), queries with
    # 5 % inverted: C++ counts, both C++ orders, C searchIdxs and searchValues
    n, nq = 3000, 2000
    s = rng.integers(0, 50_000, n).astype(np.int32)
    ln = rng.integers(0, 300, n)
    inv = rng.random(n) < 0.15
    e = np.where(inv, s - 1 - rng.integers(0, 300, n), s + ln).astype(np.int32)
    qs = rng.integers(-100, 50_500, nq).astype(np.int32)
    qe = np.where(rng.random(nq) < 0.05, qs - rng.integers(1, 200, nq), qs + rng.integers(0, 1500, nq)).astype(np.int32)
    data = np.arange(n, dtype=np.int32)
    m = oracle.RefMap(R, s, e, data)
    c = oracle.RefCMap(R, s, e, data)
    rec["inv_starts"], rec["inv_ends"], rec["inv_qs"], rec["inv_qe"] = s, e, qs, qe
    rec["inv_counts"] = np.array([m.count(x, y) for x, y in zip(qs, qe)], np.int64)
    rec["inv_desc_off"], rec["inv_desc_idx"] = csr([m.search(x, y, 0) for x, y in zip(qs, qe)], np.int32)
    rec["inv_vec_off"], rec["inv_vec_idx"] = csr([m.search(x, y, 1) for x, y in zip(qs, qe)], np.int32)
    rec["inv_cidx_off"], rec["inv_cidx"] = csr([c.search(x, y, 0) for x, y in zip(qs, qe)], np.int32)
    rec["inv_cval_off"], rec["inv_cval"] = csr([c.search(x, y, 1) for x, y in zip(qs, qe)], np.int32)
    # set operations: A in insertion order (10 % inverted), B indexed (5 % inverted)
    for sfx, dt in (("i32", np.int32), ("i64", np.int64)):
        na, nb = 700, 900
        base = dt(0) if dt == np.int32 else dt(3) << 35
        a_s = (rng.integers(0, 40_000, na) + base).astype(dt)
        a_e = np.where(rng.random(na) < 0.10, a_s - rng.integers(1, 50, na), a_s + rng.integers(0, 400, na)).astype(dt)
        b_s = (rng.integers(0, 40_000, nb) + base).astype(dt)
        b_e = np.where(rng.random(nb) < 0.05, b_s - rng.integers(1, 50, nb), b_s + rng.integers(0, 80, nb)).astype(dt)
        a_d = (1000 + np.arange(na)).astype(np.int32)
        b_d = (5000 + np.arange(nb)).astype(np.int32)
        rec[f"set_{sfx}_a"] = np.stack([a_s, a_e]).astype(dt)
        rec[f"set_{sfx}_b"] = np.stack([b_s, b_e]).astype(dt)
        rec[f"set_{sfx}_ad"], rec[f"set_{sfx}_bd"] = a_d, b_d
        ma = oracle.RefMap.raw(R, a_s, a_e, a_d)
        mb = oracle.RefMap(R, b_s, b_e, b_d)
        for op in ("intersection", "difference"):
            os_, oe, od = ma.set_op(mb, op)
            rec[f"set_{sfx}_cpp_{op}"] = np.stack([os_.astype(np.int64), oe.astype(np.int64), od.astype(np.int64)])
        if dt == np.int32:
            ca = oracle.RefCMap.raw(R, a_s, a_e, a_d)
            cb = oracle.RefCMap(R, b_s, b_e, b_d)
            for op in ("intersection", "difference"):
                os_, oe, od = ca.set_op(cb, op)
                rec[f"set_{sfx}_c_{op}"] = np.stack([os_.astype(np.int64), oe.astype(np.int64), od.astype(np.int64)])
    np.savez_compressed(os.path.join(OUT, "extras.npz"), **rec)


def main():
    if not oracle.Reference.available():
        raise SystemExit("oracle/_ref/libsiref.so is missing: run make -f oracle/Makefile ref")
    R = oracle.Reference()
    O = oracle.Oracle()
    extras(R)
    if "--extras-only" in sys.argv:
        print("done")
        return
    for name, meta, s, e, qs, qe, store in cases():
        print("case", name, flush=True)
        data = np.arange(s.size, dtype=np.int32)
        m = oracle.RefMap(R, s, e, data)
        a = m.arrays()
        rec = {k: np.array(v) for k, v in meta.items()}
        rec["input_sha256"] = np.array(sha(s, e, qs, qe))
        rec["dtype"] = np.array(str(s.dtype))
        rec["sorted_sha256"] = np.array(sha(a["starts"], a["ends"]))
        rec["branch_sha256"] = np.array(sha(a["branch"]))
        rec["cpp_data_sha256"] = np.array(sha(a["data"]))
        nq = qs.size
        counts, _ = m.count_batch(qs, qe, 8)
        rec["counts"] = counts
        rec["upper_bound"] = np.array([m.upper_bound(x) for x in qe[:NPREFIX]], np.int64)
        rec["has_overlaps"] = np.packbits(np.array([m.has_overlaps(a_, b_) for a_, b_ in zip(qs, qe)], bool))
        hd = np.empty(nq, np.uint64)
        hv = np.empty(nq, np.uint64)
        lists = []
        for q in range(nq):
            d_ = m.search(qs[q], qe[q], 0)
            v_ = m.search(qs[q], qe[q], 1)
            hd[q] = O.hash_list(d_)
            hv[q] = O.hash_list(v_)
            if q < NLISTS:
                lists.append(d_)
        rec["hash_desc_digest"] = np.array(O.digest(hd), np.uint64)
        rec["hash_vec_digest"] = np.array(O.digest(hv), np.uint64)
        rec["hash_desc_prefix"] = hd[:NPREFIX]
        rec["hash_vec_prefix"] = hv[:NPREFIX]
        rec["lists_offsets"] = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.uint64)
        rec["lists_idx"] = np.concatenate(lists).astype(np.int64) if lists else np.zeros(0, np.int64)
        if s.dtype == np.int32:
            c = oracle.RefCMap(R, s, e, data)
            ca = c.arrays()
            assert np.array_equal(ca["starts"], a["starts"]) and np.array_equal(ca["ends"], a["ends"])
            assert np.array_equal(ca["branch"], a["branch"])
            rec["c_data_sha256"] = np.array(sha(ca["data"]))
            hval = np.empty(nq, np.uint64)
            for q in range(nq):
                hval[q] = O.hash_list(c.search(qs[q], qe[q], 1).astype(np.int64))
            rec["hash_cvalues_digest"] = np.array(O.digest(hval), np.uint64)
            rec["hash_cvalues_prefix"] = hval[:NPREFIX]
            cov = np.array([c.coverage(qs[q], qe[q]) for q in range(NPREFIX)], np.int64)
            rec["coverage_prefix"] = cov
        if store:
            rec["in_starts"], rec["in_ends"], rec["in_qs"], rec["in_qe"] = s, e, qs, qe
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **rec)
    print("done")


if __name__ == "__main__":
    main()
