"""Writes the golden fixtures under tests/golden/ from the reference's own code.

Runs in the development container only: it needs oracle/_ref/libsiref.so, which
oracle/Makefile compiles in place from the reference checkout
(src/superintervals.hpp and src/c_superintervals.h of kcleal/superintervals v1.0.0).
The fixtures hold inputs (or the SHA-256 of generated inputs) and the reference's
outputs: sorted arrays, branch, counts, upper_bound, has_overlaps, coverage and
order-sensitive per-query hashes of search_idxs (both orders) and searchValues.

    python tests/golden/make_golden.py
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


def main():
    if not oracle.Reference.available():
        raise SystemExit("oracle/_ref/libsiref.so is missing: run make -f oracle/Makefile ref")
    R = oracle.Reference()
    O = oracle.Oracle()
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
