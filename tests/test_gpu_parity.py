"""GPU parity: the HIP library against the golden vectors and the CPU oracle.

Everything here calls libsuperintervals_hip.so through its C-ABI (ctypes) and
needs a GPU.  Integer results must be bit-exact: same sorted arrays, same branch,
same counts, same index lists in the same order.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN_CASES, load_golden, sha

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sia():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import superintervals_amd
    return superintervals_amd


def _dev_index(sia, s, e, values=None):
    ix = sia.DeviceIndex(s.dtype)
    ix.build(s, e, values)
    return ix


def _hash(orc, off, idx):
    return orc.hash_csr(off, idx)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_build_and_queries(sia, orc, name):
    print(f"[gpu] golden {name}", flush=True)
    rec, s, e, qs, qe = load_golden(name)
    data = np.arange(s.size, dtype=np.int32)
    ix = _dev_index(sia, s, e, data)
    arr = ix.download(values=True)
    assert sha(arr["starts"], arr["ends"]) == str(rec["sorted_sha256"])
    assert sha(arr["branch"]) == str(rec["branch_sha256"])
    if "c_data_sha256" in rec:
        assert sha(arr["values"]) == str(rec["c_data_sha256"])
        assert sha(arr["perm"].astype(np.int32)) == str(rec["c_data_sha256"])
    counts = ix.count(qs, qe)
    assert np.array_equal(counts, rec["counts"])
    off, idx = ix.search(qs, qe, order="desc")
    assert np.array_equal(np.diff(off), rec["counts"].astype(np.uint64))
    hd = orc.hash_csr(off, idx)
    assert np.array_equal(hd[:rec["hash_desc_prefix"].size], rec["hash_desc_prefix"])
    assert orc.digest(hd) == int(rec["hash_desc_digest"])
    offv, idxv = ix.search(qs, qe, order="vector")
    assert np.array_equal(offv, off)
    hv = orc.hash_csr(offv, idxv)
    assert orc.digest(hv) == int(rec["hash_vec_digest"])
    lo = rec["lists_offsets"]
    for q in range(lo.size - 1):
        assert np.array_equal(idx[off[q]:off[q + 1]].astype(np.int64), rec["lists_idx"][lo[q]:lo[q + 1]])
    k = rec["upper_bound"].size
    ub = ix.upper_bound(qe[:k]).astype(np.int64)
    ub[ub == 0xFFFFFFFF] = -1
    assert np.array_equal(ub, rec["upper_bound"])
    assert np.array_equal(np.packbits(ix.has_overlaps(qs, qe)), rec["has_overlaps"])
    if "hash_cvalues_digest" in rec:
        offval, vals = ix.search(qs, qe, values=True)
        assert np.array_equal(offval, off)
        assert orc.digest(orc.hash_csr(offval, vals)) == int(rec["hash_cvalues_digest"])
        kc = rec["coverage_prefix"].shape[0]
        c, v = ix.coverage(qs[:kc], qe[:kc])
        assert np.array_equal(np.stack([c.astype(np.int64), v.astype(np.int64)], 1), rec["coverage_prefix"])
    ix.close()


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_random_vs_oracle(sia, orc, seed, dtype):
    print(f"[gpu] random seed={seed} dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(100 + seed)
    n = int(rng.choice([1, 2, 63, 64, 65, 4095, 4096, 4097, 20_000, 150_000]))
    span = int(rng.choice([50, 5_000, 1_000_000]))
    s = rng.integers(-span, span, n).astype(dtype)
    L = rng.geometric(1.0 / max(2, span // 50), n)
    e = (s + L - 1).astype(dtype)
    if seed % 2:   # presorted input takes the no-sort path
        o = np.lexsort((-e.astype(np.int64), s.astype(np.int64)))
        s, e = s[o], e[o]
    if dtype == np.int64:
        s = s + (dtype(3) << 40)
        e = e + (dtype(3) << 40)
    nq = 30_000
    qs = rng.integers(-span - 10, span + 10, nq).astype(dtype)
    qe = (qs + rng.integers(-5, span // 20 + 5, nq)).astype(dtype)   # a few inverted
    if dtype == np.int64:
        qs = qs + (dtype(3) << 40)
        qe = qe + (dtype(3) << 40)
    oix = orc.build_index(s, e)
    ix = _dev_index(sia, s, e, np.arange(n, dtype=np.int32))
    a = ix.download(values=True)
    assert np.array_equal(a["starts"], oix["starts"]) and np.array_equal(a["ends"], oix["ends"])
    assert np.array_equal(a["branch"], oix["branch"])
    assert np.array_equal(a["values"], oix["data"])
    ocount = orc.count_batch(oix, qs, qe)
    from superintervals_amd import autotune
    saved = autotune.current()
    try:
        for cr in (0, 1):                   # the walk, and the rank form (walk for inverted queries)
            autotune.apply({"count_rank": cr})
            assert np.array_equal(ix.count(qs, qe), ocount), cr
    finally:
        autotune.apply(saved)
    for order, mode in (("desc", 0), ("vector", 1)):
        off, idx = ix.search(qs, qe, order=order)
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=mode)
        assert np.array_equal(off, ooff)
        assert np.array_equal(idx.astype(np.int64), oidx)
    off, vals = ix.search(qs, qe, values=True)
    assert np.array_equal(vals, oix["data"][orc.search_batch(oix, qs, qe, mode=0)[1]])
    ub = ix.upper_bound(qe).astype(np.int64)
    ub[ub == 0xFFFFFFFF] = -1
    assert np.array_equal(ub, np.searchsorted(oix["starts"], qe, side="right") - 1)
    ho = np.array([orc.has_overlaps(oix, x, y) for x, y in zip(qs[:3000], qe[:3000])])
    assert np.array_equal(ix.has_overlaps(qs[:3000], qe[:3000]), ho)
    if dtype == np.int32:
        c, v = ix.coverage(qs[:3000], qe[:3000])
        want = [orc.coverage(oix, x, y) for x, y in zip(qs[:3000], qe[:3000])]
        assert np.array_equal(np.stack([c, v], 1).astype(np.int64), np.array(want, np.int64))
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_upper_bound_tree_and_halving(sia, dtype):
    """Both upper_bound forms (Eytzinger top + block separators, and plain halving) against
    numpy on sizes around the 64-start blocks and the 4095-key top, long duplicate runs
    across block boundaries, and keys at the type's extremes."""
    from superintervals_amd import autotune
    info = np.iinfo(dtype)
    rng = np.random.default_rng(91)
    saved = autotune.current()
    try:
        for n in (1, 63, 64, 65, 4095 * 64, 4096 * 64 + 1, 300_001):
            for kind in range(3):
                if kind == 0:
                    s = np.sort(rng.integers(info.min, info.max, n, dtype=np.int64)).astype(dtype)
                elif kind == 1:
                    s = np.sort(rng.integers(0, max(2, n // 50), n)).astype(dtype)
                else:
                    s = np.sort(np.concatenate([rng.integers(0, 10, n // 2),
                                                np.full(n - n // 2, info.max, np.int64)])).astype(dtype)
                ix = _dev_index(sia, s, s)
                st = ix.download()["starts"]
                xs = np.concatenate([st[rng.integers(0, n, 500)].astype(np.int64),
                                     st[rng.integers(0, n, 200)].astype(np.int64) - 1,
                                     st[rng.integers(0, n, 200)].astype(np.int64) + 1,
                                     [info.min, info.max, int(st[0]) - 1],
                                     rng.integers(info.min, info.max, 300, dtype=np.int64)])
                xs = np.clip(xs, info.min, info.max).astype(dtype)
                want = np.searchsorted(st, xs, side="right").astype(np.int64) - 1
                for tree in (1, 0):
                    autotune.apply({"ub_tree": tree})
                    got = ix.upper_bound(xs).astype(np.int64)
                    got[got == 0xFFFFFFFF] = -1
                    assert np.array_equal(got, want), (n, kind, tree)
                ix.close()
    finally:
        autotune.apply(saved)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_search_keys_and_items(sia, orc, dtype):
    """Batched search_keys / search_items (superintervals.hpp:913-971): the descending walk's
    hits as (start, end) and (start, end, value), against the oracle's descending lists."""
    print(f"[gpu] search_keys/items dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(77)
    n = 30_000
    s = rng.integers(0, 2_000_000, n).astype(dtype)
    e = (s + rng.geometric(1 / 400, n) - 1).astype(dtype)
    if dtype == np.int64:
        s, e = s + (dtype(5) << 36), e + (dtype(5) << 36)
    vals = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    qs = rng.integers(0, 2_000_000, 20_000).astype(dtype)
    qe = (qs + rng.integers(0, 3_000, qs.size)).astype(dtype)
    if dtype == np.int64:
        qs, qe = qs + (dtype(5) << 36), qe + (dtype(5) << 36)
    ix = _dev_index(sia, s, e, vals)
    oix = orc.build_index(s, e, vals)
    ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
    off, ks, ke = ix.search_keys(qs, qe)
    assert np.array_equal(off, ooff.astype(np.uint64))
    assert np.array_equal(ks.astype(np.int64), oix["starts"][oidx].astype(np.int64))
    assert np.array_equal(ke.astype(np.int64), oix["ends"][oidx].astype(np.int64))
    off2, ks2, ke2, kv2 = ix.search_keys(qs, qe, items=True)
    assert np.array_equal(off2, off) and np.array_equal(ks2, ks) and np.array_equal(ke2, ke)
    assert np.array_equal(kv2, oix["data"][oidx])
    off0, ks0, ke0 = ix.search_keys(qs[:0], qe[:0])
    assert off0.tolist() == [0] and ks0.size == 0
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_planned_batch_reuse(sia, orc, dtype):
    """One plan serves count, offsets and every fill mode, repeatedly, on one stream."""
    import torch
    print(f"[gpu] planned reuse dtype={np.dtype(dtype).name}", flush=True)
    rec, s, e, qs, qe = load_golden("c1_shuffled")
    s, e, qs, qe = (x.astype(dtype) for x in (s, e, qs, qe))
    ix = _dev_index(sia, s, e, np.arange(s.size, dtype=np.int32))
    oix = orc.build_index(s, e)
    dev = torch.device("cuda:0")
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    nq = qs.size
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    ix.plan_device(tqs, tqe, plan)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    for _ in range(2):
        ix.count_planned(plan, nq, counts)
        assert np.array_equal(counts.cpu().numpy().astype(np.uint32), rec["counts"])
        ix.offsets_planned(plan, nq, off, counts)
        assert np.array_equal(np.diff(off.cpu().numpy()), rec["counts"].astype(np.int64))
    total = int(off[-1].item())
    idx = torch.empty(total, dtype=torch.int32, device=dev)
    offh = off.cpu().numpy().astype(np.uint64)
    for order, key in (("desc", "hash_desc_digest"), ("vector", "hash_vec_digest"), ("desc", "hash_desc_digest")):
        idx.fill_(-7)
        ix.fill_planned(plan, nq, off, idx, order=order)
        h = orc.hash_csr(offh, idx.cpu().numpy().view(np.uint32))
        assert orc.digest(h) == int(rec[key]), order
    vals = torch.empty(total, dtype=torch.int32, device=dev)
    ix.values_fill_planned(plan, nq, off, vals)
    ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
    assert np.array_equal(vals.cpu().numpy(), oix["data"][oidx])
    ix.close()


def test_launch_choices_do_not_change_results(sia, orc):
    """Every tile size / residency choice the autotuner may pick gives identical output."""
    from superintervals_amd import autotune
    rec, s, e, qs, qe = load_golden("c2_1m")
    ix = _dev_index(sia, s, e)
    saved = autotune.current()
    try:
        for tile, cw, fw, cr, orank, psl, rs in ((64, 1, 1, 0, 1, 0, 1), (512, 0, 1, 1, 0, 1, 0),
                                                 (4096, 1, 0, 1, 1, 2, 1), (100000, 0, 0, 0, 0, 1, 0),
                                                 (1024, 0, 0, 1, 1, 2, 1), (1024, 0, 0, 0, 0, 0, 0)):
            print(f"[gpu] launch tile={tile} count_wgs={cw} fill_wgs={fw} rank={cr}/{orank} skip_low={psl} "
                  f"range_skip={rs}", flush=True)
            autotune.apply({"tile": tile, "count_wgs": cw, "fill_wgs": fw, "count_rank": cr,
                            "offsets_rank": orank, "plan_skip_low": psl, "range_skip": rs})
            assert np.array_equal(ix.count(qs, qe), rec["counts"])
            off, idx = ix.search(qs, qe, order="vector")
            assert orc.digest(orc.hash_csr(off, idx)) == int(rec["hash_vec_digest"])
    finally:
        autotune.apply(saved)
    with pytest.raises(sia.SIError):
        autotune.apply({"tile": 3})
    ix.close()


@pytest.mark.parametrize("lane_steps", [1, 3, 17])
def test_heavy_path_parity(sia, orc, lane_steps):
    """Walks parked after lane_steps lines finish in the wave-cooperative kernel with the
    same results, in every mode (count by walk, descending, vector, values)."""
    from superintervals_amd import autotune
    print(f"[gpu] heavy path lane_steps={lane_steps}", flush=True)
    rec, s, e, qs, qe = load_golden("c1_shuffled")
    ix = _dev_index(sia, s, e, np.arange(s.size, dtype=np.int32))
    oix = orc.build_index(s, e)
    saved = autotune.current()
    try:
        autotune.apply({"lane_steps": lane_steps, "offsets_rank": 0, "count_rank": 0})
        assert np.array_equal(ix.count(qs, qe), rec["counts"])
        off, idx = ix.search(qs, qe, order="desc")
        assert orc.digest(orc.hash_csr(off, idx)) == int(rec["hash_desc_digest"])
        off, idx = ix.search(qs, qe, order="vector")
        assert orc.digest(orc.hash_csr(off, idx)) == int(rec["hash_vec_digest"])
        off, vals = ix.search(qs, qe, values=True)
        assert orc.digest(orc.hash_csr(off, vals)) == int(rec["hash_cvalues_digest"])
    finally:
        autotune.apply(saved)
    ix.close()


@pytest.mark.parametrize("lane_steps,far", [(1, False), (40, False), (1, True)])
def test_heavy_path_int64_long_walks(sia, orc, lane_steps, far):
    """int64 hotspot mixture (workload 3 distributions): every walk parked after lane_steps
    lines and finished by the wave-cooperative kernel, which fetches four lines at a time
    and leaves a line through exitp lookups.  Ends within 2^32 of each other are read from
    the 32-bit copy; far = True adds intervals 2^40 away, so the index has no such copy.
    Counts, both orders and values equal the oracle's."""
    from superintervals_amd import autotune
    print(f"[gpu] heavy path int64 lane_steps={lane_steps} far={far}", flush=True)
    s, e = sia.workload_intervals(3, 1_000_000, seed=33, sorted=False)
    if far:
        extra = (np.int64(1) << 40) + np.arange(0, 2000 * 97, 97, dtype=np.int64)
        s = np.concatenate([s, extra])
        e = np.concatenate([e, extra + 500])
    qs, qe = sia.workload_queries(3, 6_000, seed=34)
    # ragged edges: queries before, after and around the whole index, and a point query
    qs = np.concatenate([qs, np.array([-5, 0, int(e.max()), int(s.min()), 10**9], np.int64)])
    qe = np.concatenate([qe, np.array([-1, int(e.max()) + 10, int(e.max()) + 10, int(s.min()), 10**9], np.int64)])
    vals = (np.arange(s.size, dtype=np.int64) * 7 % 1_000_003).astype(np.int32)
    ix = _dev_index(sia, s, e, vals)
    oix = orc.build_index(s, e, vals)
    saved = autotune.current()
    try:
        autotune.apply({"lane_steps": lane_steps, "offsets_rank": 0, "count_rank": 0})
        ocounts = orc.count_batch(oix, qs, qe)
        assert np.array_equal(ix.count(qs, qe), ocounts)
        for order, mode in (("desc", 0), ("vector", 1)):
            off, idx = ix.search(qs, qe, order=order)
            ooff, oidx = orc.search_batch(oix, qs, qe, mode=mode)
            assert np.array_equal(off, ooff), order
            assert np.array_equal(idx.astype(np.int64), oidx), order
        off, v = ix.search(qs, qe, values=True)
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
        assert np.array_equal(v, oix["data"][oidx])
        # a walk with more than 4 x 64 hits spans more lines than one fetch holds
        assert (ocounts > 4 * 64).sum() > 100
    finally:
        autotune.apply(saved)
    ix.close()


def test_autotune_picks_and_keeps_results(sia, orc):
    """tune() on a planned batch picks tile, residency, lane step budget and count form,
    applies them, and the library's output with them equals the oracle's."""
    import torch
    from superintervals_amd import autotune
    print("[gpu] autotune on workload 3 (300K int64 intervals)", flush=True)
    s, e = sia.workload_intervals(3, 300_000, seed=35, sorted=False)
    qs, qe = sia.workload_queries(3, 4_000, seed=36)
    ix = _dev_index(sia, s, e)
    oix = orc.build_index(s, e)
    dev = torch.device("cuda:0")
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    nq = qs.size
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    saved = autotune.current()
    try:
        ix.plan_device(tqs, tqe, plan)
        ix.offsets_planned(plan, nq, off)
        total = int(off[-1].item())
        idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        best, table = autotune.tune(ix, plan, nq, off, idx, reps=1, tiles=(512, 2048), wgs=(0,),
                                    lane_steps=(16, 256), qs_t=tqs, qe_t=tqe)
        assert best["lane_steps"] in (16, 256) and best["count_rank"] in (0, 1)
        assert best["plan_sync"] in (0, 1) and best["plan_skip_low"] in (0, 1)
        now = autotune.current()
        for k, v in autotune.knobs_of(best).items():
            assert now[k] == v, k
        assert len(table) == 2 + 2 + 2 + 2 + 2
        # the plan redone during tuning is still this batch's plan
        ix.offsets_planned(plan, nq, off)
        ix.fill_planned(plan, nq, off, idx, order="desc")
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
        assert np.array_equal(off.cpu().numpy().astype(np.uint64), ooff)
        assert np.array_equal(idx[:total].cpu().numpy().view(np.uint32).astype(np.int64), oidx)
        counts = ix.count(qs, qe)
        assert np.array_equal(counts, orc.count_batch(oix, qs, qe))
        o, i = ix.search(qs, qe, order="desc")
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
        assert np.array_equal(o, ooff) and np.array_equal(i.astype(np.int64), oidx)
    finally:
        autotune.apply(saved)
    ix.close()


def test_workload3_wide_queries(sia, orc):
    """int64 hotspot mixture with 1 % wide queries (workload 3 distributions, 2M intervals),
    unsorted input: full CSR against the oracle in both orders."""
    print("[gpu] workload 3 (2M int64 intervals, 20K queries)", flush=True)
    s, e = sia.workload_intervals(3, 2_000_000, seed=31, sorted=False)
    qs, qe = sia.workload_queries(3, 20_000, seed=32)
    ix = _dev_index(sia, s, e, np.arange(s.size, dtype=np.int32))
    oix = orc.build_index(s, e)
    a = ix.download(values=True)
    assert np.array_equal(a["starts"], oix["starts"]) and np.array_equal(a["branch"], oix["branch"])
    assert np.array_equal(a["values"], oix["data"])
    counts = ix.count(qs, qe)
    ocounts = orc.count_batch(oix, qs, qe)
    assert np.array_equal(counts, ocounts)
    assert ocounts.max() > 5_000           # the wide queries exceed the lane step budget
    for order, mode in (("desc", 0), ("vector", 1)):
        off, idx = ix.search(qs, qe, order=order)
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=mode)
        assert np.array_equal(off, ooff)
        assert np.array_equal(idx.astype(np.int64), oidx)
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_wide_queries_guaranteed_run(sia, orc, dtype):
    """Queries much wider than the intervals: most hits are the run [lower_bound(s),
    upper_bound(e)] taken without reading ends.  Every mode and order, the run taken or
    walked, lane and wave walks, against the oracle; some runs reach position 0."""
    from superintervals_amd import autotune
    print(f"[gpu] wide queries, guaranteed run dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(77)
    n = 150_000
    s = rng.integers(0, 3_000_000, n)
    e = s + rng.integers(0, 200, n)
    long_ = rng.random(n) < 0.02
    e = np.where(long_, s + rng.integers(1000, 200_000, n), e)
    qs = rng.integers(-5_000, 3_000_000, 6000)
    qe = qs + rng.integers(20_000, 400_000, qs.size)
    qs[:50] = -10                        # runs that reach position 0
    base = (np.int64(9) << 34) if dtype == np.int64 else 0
    s, e, qs, qe = ((x + base).astype(dtype) for x in (s, e, qs, qe))
    ix = _dev_index(sia, s, e, np.arange(n, dtype=np.int32))
    oix = orc.build_index(s, e)
    ocount = orc.count_batch(oix, qs, qe)
    st = oix["starts"]
    run = np.searchsorted(st, qe, side="right") - np.searchsorted(st, qs, side="left")
    assert run.sum() > 0.8 * ocount.sum()      # the run dominates
    want = {o: orc.search_batch(oix, qs, qe, mode=m) for o, m in (("desc", 0), ("vector", 1))}
    saved = autotune.current()
    try:
        for rs in (1, 0):
            for ls in (1024, 2):
                autotune.apply({"range_skip": rs, "lane_steps": ls, "count_rank": 0, "offsets_rank": 0})
                assert np.array_equal(ix.count(qs, qe), ocount), (rs, ls)
                for order in ("desc", "vector"):
                    off, idx = ix.search(qs, qe, order=order)
                    assert np.array_equal(off, want[order][0]) and np.array_equal(idx.astype(np.int64), want[order][1]), (rs, ls, order)
                off, vals = ix.search(qs, qe, values=True)
                assert np.array_equal(vals, oix["data"][want["desc"][1]])
    finally:
        autotune.apply(saved)
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_range_lengths_around_wave_written_threshold(sia, orc, dtype):
    """Range segments of 64 or more values are written by the whole wave, shorter ones in
    the shared 64-slot chunks.  Runs of exactly 1..1000 values (63, 64, 65, 127, 128, 129
    among them), next to line masks from a few longer intervals, in every mode and order,
    against the oracle."""
    print(f"[gpu] range lengths around the wave-written threshold dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(5)
    n = 100_000
    s = np.arange(n, dtype=np.int64)
    e = s + 1
    longer = rng.random(n) < 0.05
    e = np.where(longer, s + rng.integers(2, 300, n), e)
    lens = np.array([1, 2, 31, 32, 33, 63, 64, 65, 66, 127, 128, 129, 191, 192, 193, 1000])
    L = np.repeat(lens, 300)
    qs = rng.integers(0, n - 1200, L.size)
    qe = qs + L - 1
    base = (np.int64(5) << 33) if dtype == np.int64 else 0
    s, e, qs, qe = ((x + base).astype(dtype) for x in (s, e, qs, qe))
    ix = _dev_index(sia, s, e, np.arange(n, dtype=np.int32))
    oix = orc.build_index(s, e)
    assert np.array_equal(ix.count(qs, qe), orc.count_batch(oix, qs, qe))
    for order, mode in (("desc", 0), ("vector", 1)):
        off, idx = ix.search(qs, qe, order=order)
        ooff, oidx = orc.search_batch(oix, qs, qe, mode=mode)
        assert np.array_equal(off, ooff) and np.array_equal(idx.astype(np.int64), oidx), order
    off, vals = ix.search(qs, qe, values=True)
    ooff, oidx = orc.search_batch(oix, qs, qe, mode=0)
    assert np.array_equal(vals, oix["data"][oidx])
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_plan_sort_forms_order_queries_by_end(sia, dtype):
    """The plan sorts the queries by end, stably.  Both forms of its radix sort (plan_sync
    1: a host read-back of the digit histograms skips the digits that are equal for every
    query; 0: every digit is sorted, with no read-back) must give numpy's stable argsort of
    the ends, for one item, one tile, a partial last tile and thousands of tiles, with ends
    that are random, spread over the whole coordinate range, few distinct values (long runs
    of equal digits, and digits constant over the batch), runs of equal ends (digit counts
    taken once per wave where a wave's digits agree, per lane elsewhere) or all equal (the
    read-back form then runs no pass)."""
    import torch
    from superintervals_amd import autotune
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(91)
    s, e = sia.workload_intervals(2, 100_000, seed=92, sorted=True)
    ix = _dev_index(sia, s.astype(dtype), e.astype(dtype))
    info = np.iinfo(dtype)
    saved = autotune.current()
    try:
        for nq in (1, 2, 2047, 2048, 2049, 65_537, 3_000_001):
            for kind in ("random", "wide", "few", "runs", "equal"):
                print(f"[gpu] plan sort dtype={np.dtype(dtype).name} nq={nq} ends={kind}", flush=True)
                if kind == "random":
                    qe = rng.integers(0, 2_000_000, nq)
                elif kind == "wide":
                    qe = rng.integers(info.min // 2, info.max // 2, nq)
                elif kind == "few":
                    qe = rng.choice(np.array([-3, 0, 255, 256, 70_000], dtype=np.int64), nq)
                elif kind == "runs":   # runs of ~100 equal ends: most waves see one digit value
                    qe = np.repeat(rng.integers(-5_000_000, 5_000_000, nq // 100 + 1), 100)[:nq]
                else:   # every digit constant: the read-back form runs no pass
                    qe = np.full(nq, 12_345, dtype=np.int64)
                qe = qe.astype(dtype)
                qs = (qe.astype(np.int64) - rng.integers(0, 1000, nq)).astype(dtype)
                want = np.argsort(qe, kind="stable")
                tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
                plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
                for ps in (1, 0):
                    autotune.apply({"plan_sync": ps, "plan_skip_low": 0})
                    plan.fill_(0xA5)
                    ix.plan_device(tqs, tqe, plan)
                    torch.cuda.synchronize()
                    q = plan[256:256 + 4 * nq].view(torch.int32).cpu().numpy()
                    assert np.array_equal(q, want), (nq, kind, ps)
    finally:
        autotune.apply(saved)
    ix.close()


@pytest.mark.parametrize("dtype,long_len", [(np.int64, 0), (np.int64, 4096), (np.int32, 64), (np.int32, 1)])
def test_long_list_walk(sia, orc, dtype, long_len):
    """The wave walk with a long list: hits below the first start >= s - len come from the
    list of intervals longer than len.  Built with the threshold chosen from the lengths
    (0) or forced, queries on the wave walk (lane_steps 2) and not, the list used or not,
    every mode and order against the oracle."""
    from superintervals_amd import autotune
    print(f"[gpu] long list dtype={np.dtype(dtype).name} long_len={long_len}", flush=True)
    if dtype == np.int64:
        s, e = sia.workload_intervals(3, 2_000_000, seed=33, sorted=False)
        qs, qe = sia.workload_queries(3, 6000, seed=34)
    else:
        rng = np.random.default_rng(12)
        n = 300_000
        s = rng.integers(0, 20_000_000, n)
        e = s + np.where(rng.random(n) < 0.1, rng.integers(5000, 2_000_000, n), rng.integers(-3, 300, n))
        qs = rng.integers(0, 20_000_000, 6000)
        qe = qs + rng.integers(-2, 3000, qs.size)
        s, e, qs, qe = (x.astype(np.int32) for x in (s, e, qs, qe))
    saved = autotune.current()
    try:
        autotune.apply({"long_len": long_len})
        ix = _dev_index(sia, s, e, np.arange(s.size, dtype=np.int32))
    finally:
        autotune.apply(saved)
    oix = orc.build_index(s, e)
    wc = orc.count_batch(oix, qs, qe)
    want = {o: orc.search_batch(oix, qs, qe, mode=m) for o, m in (("desc", 0), ("vector", 1))}
    try:
        for lw in (1, 0):
            for ls in (2, 1024):
                autotune.apply({"long_walk": lw, "lane_steps": ls, "count_rank": 0, "offsets_rank": 0})
                assert np.array_equal(ix.count(qs, qe), wc), (lw, ls)
                for order in ("desc", "vector"):
                    off, idx = ix.search(qs, qe, order=order)
                    assert np.array_equal(off, want[order][0]), (lw, ls, order)
                    assert np.array_equal(idx.astype(np.int64), want[order][1]), (lw, ls, order)
                off, vals = ix.search(qs, qe, values=True)
                assert np.array_equal(vals, oix["data"][want["desc"][1]]), (lw, ls)
        # single queries take the wave walk with the list too
        autotune.apply({"long_walk": 1})
        for q in range(0, 200, 7):
            assert ix.count(qs[q:q + 1], qe[q:q + 1])[0] == wc[q]
            off, idx = ix.search(qs[q:q + 1], qe[q:q + 1], order="vector")
            assert np.array_equal(idx.astype(np.int64), orc.search(oix, qs[q], qe[q], 1))
    finally:
        autotune.apply(saved)
    ix.close()


def test_learned_index(sia, orc):
    """Learned upper_bound (workload 4 distributions, 2M intervals): GPU gradients against the
    float64 oracle, output identical to the plain search, and the trained root's gain over the
    untrained (identity-CDF) one in leaf corrections."""
    import torch
    from oracle import rmi
    print("[gpu] learned index", flush=True)
    s, e = sia.workload_intervals(4, 2_000_000, seed=41, sorted=False)
    ix = _dev_index(sia, s, e)
    starts = ix.download()["starts"]
    n = starts.size
    dev = torch.device("cuda:0")

    untrained = sia.LearnedIndex(ix, 4096)
    info0 = untrained.train(steps=0)
    model = sia.LearnedIndex(ix, 4096)
    info1 = model.train(steps=400, batch=65536, lr=3e-3)
    assert info1["loss"] < 1e-3

    # gradients of the training kernels against the oracle, at the initial and trained points
    kmin, inv = model.norm()
    pos = np.random.default_rng(3).integers(0, n, 4096).astype(np.uint32)
    x = (np.float64(starts[pos].astype(np.int64) - kmin) * np.float64(np.float32(inv))).astype(np.float32)
    t = (pos.astype(np.float32) / np.float32(n)).astype(np.float32)
    for p in (rmi.identity_params().astype(np.float32), model.params()):
        g_gpu = model.gradient(pos, p)
        g_cpu, l_cpu = rmi.gradient(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        scale = np.abs(g_cpu).max()
        assert np.all(np.abs(g_gpu[:-1] - g_cpu) <= 1e-4 * scale + 2e-3 * np.abs(g_cpu)), \
            np.abs(g_gpu[:-1] - g_cpu).max() / scale
        assert abs(g_gpu[-1] - l_cpu) <= 1e-5 + 1e-3 * l_cpu

    # identical to the plain upper_bound on point, short-range and out-of-range queries
    qs, qe = sia.workload_queries(4, 1_000_000, seed=42)
    xs = np.concatenate([qe, qs, starts[:1000], starts[-1000:], np.array([np.iinfo(np.int32).min,
                         np.iinfo(np.int32).max, int(starts[0]) - 1, int(starts[-1]) + 1], np.int32)])
    xt = torch.from_numpy(xs).to(dev)
    plain = torch.empty(xs.size, dtype=torch.int32, device=dev)
    assert sia._native.lib.siUpperBound(ix.handle, xt.data_ptr(), xs.size, plain.data_ptr(), None) == 0
    want = np.searchsorted(starts, xs, side="right").astype(np.int64) - 1
    got_plain = plain.cpu().numpy().view(np.uint32).astype(np.int64)
    got_plain[got_plain == 0xFFFFFFFF] = -1
    assert np.array_equal(got_plain, want)
    corr = {}
    for name, m in (("untrained", untrained), ("trained", model)):
        out = torch.empty(xs.size, dtype=torch.int32, device=dev)
        st = torch.zeros(3, dtype=torch.int64, device=dev)
        m.upper_bound_device(xt, out, st)
        got = out.cpu().numpy().view(np.uint32).astype(np.int64)
        got[got == 0xFFFFFFFF] = -1
        assert np.array_equal(got, want), name
        fb, w, c = st.cpu().tolist()
        corr[name] = c / xs.size
        print(f"[gpu] {name}: fallbacks {fb}, mean window {w / xs.size:.1f}, mean leaf correction "
              f"{c / xs.size:.2f}", flush=True)
    # Leaves own fixed position ranges, so the windows do not depend on the root; what the
    # trained network gains is naming the right leaf: fewer correction steps than the
    # identity-CDF root on a density that varies 100x between segments.
    print(f"[gpu] max leaf error bound {info1['max_abs_err']}", flush=True)
    assert info0["max_abs_err"] == info1["max_abs_err"]
    assert corr["trained"] < 0.25 * corr["untrained"]
    model.close()
    untrained.close()
    ix.close()


def test_learned_plan_forms_and_forced_fallback(sia, orc):
    """The learned model as the plan's bound search (knob ub_learned): the batched search and
    count through it equal the plain-halving and tree forms and the oracle.  With every leaf
    window narrowed (test knob learned_shrink) the queries take the whole-leaf fallback and
    the output is still identical.  A rebuilt index drops the model; destroying the index
    before the model is safe."""
    import torch
    from superintervals_amd import autotune
    print("[gpu] learned plan forms and forced fallback", flush=True)
    s, e = sia.workload_intervals(4, 1_000_000, seed=43, sorted=False)
    ix = _dev_index(sia, s, e)
    model = sia.LearnedIndex(ix, 4096)
    model.train(steps=200, batch=65536, lr=3e-3)
    qs, qe = sia.workload_queries(4, 200_000, seed=44)
    qs[:100] = np.iinfo(np.int32).min          # lower bound of the smallest value
    oix = orc.build_index(s, e)
    wd = orc.search_batch(oix, qs, qe, mode=0)
    wv = orc.search_batch(oix, qs, qe, mode=1)
    wc = orc.count_batch(oix, qs, qe)
    saved = autotune.current()
    try:
        for knobs in ({"ub_tree": 0, "ub_learned": 0}, {"ub_tree": 1, "ub_learned": 0}, {"ub_learned": 1},
                      {"ub_learned": 1, "range_skip": 0}, {"ub_learned": 1, "learned_shrink": 1 << 20},
                      {"ub_learned": 1, "learned_shrink": 3}):
            autotune.apply(dict({"ub_tree": 1, "range_skip": 1, "learned_shrink": 0}, **knobs))
            print(f"[gpu] plan bounds {knobs}", flush=True)
            assert np.array_equal(ix.count(qs, qe), wc), knobs
            off, idx = ix.search(qs, qe, order="desc")
            assert np.array_equal(off, wd[0]) and np.array_equal(idx.astype(np.int64), wd[1]), knobs
            off, idx = ix.search(qs, qe, order="vector")
            assert np.array_equal(off, wv[0]) and np.array_equal(idx.astype(np.int64), wv[1]), knobs
        # the forced fallback, counted: identical to upper_bound with fallbacks > 0
        starts = ix.download()["starts"]
        want = np.searchsorted(starts, qe, side="right").astype(np.int64) - 1
        xt = torch.from_numpy(qe).cuda()
        for shrink in (0, 3, 1 << 20):
            autotune.apply({"learned_shrink": shrink})
            out = torch.empty(qe.size, dtype=torch.int32, device="cuda")
            st = torch.zeros(3, dtype=torch.int64, device="cuda")
            model.upper_bound_device(xt, out, st)
            got = out.cpu().numpy().view(np.uint32).astype(np.int64)
            got[got == 0xFFFFFFFF] = -1
            assert np.array_equal(got, want), shrink
            fb = int(st[0].item())
            print(f"[gpu] learned_shrink {shrink}: fallbacks {fb}", flush=True)
            if shrink == 0:
                assert fb == 0
            if shrink == 1 << 20:
                assert fb >= int((want >= 0).sum())    # every query with a position took it
    finally:
        autotune.apply(saved)
    # rebuilding the index drops the model: ub_learned then plans through the tree
    ix.build(s, e)
    with pytest.raises(sia.SIError):
        model.upper_bound_device(xt, torch.empty(qe.size, dtype=torch.int32, device="cuda"))
    try:
        autotune.apply({"ub_learned": 1})
        assert np.array_equal(ix.count(qs, qe), wc)
    finally:
        autotune.apply(saved)
    ix.close()                   # the index first: the model forgets it
    model.close()


def test_walk_stats_match_oracle(sia, orc):
    import torch
    rec, s, e, qs, qe = load_golden("c1_sorted")
    ix = _dev_index(sia, s, e)
    tot = ix.walk_stats_device(torch.from_numpy(qs).cuda(), torch.from_numpy(qe).cuda())
    oix = orc.build_index(s, e)
    h, v, j = orc.walk_stats_batch(oix, qs, qe)
    assert tot[:3] == [h, v, j]
    assert tot[0] == int(rec["counts"].sum())


def test_empty_and_edge_cases(sia, orc):
    # empty index: every query is well defined and finds nothing
    ix = sia.DeviceIndex(np.int32).build(np.zeros(0, np.int32), np.zeros(0, np.int32))
    q = np.array([1, 5], np.int32)
    assert ix.count(q, q).tolist() == [0, 0]
    off, idx = ix.search(q, q)
    assert off.tolist() == [0, 0, 0] and idx.size == 0
    assert ix.has_overlaps(q, q).tolist() == [False, False]
    # empty query batch
    ix1 = sia.DeviceIndex(np.int32).build(np.array([1], np.int32), np.array([10], np.int32))
    assert ix1.count(np.zeros(0, np.int32), np.zeros(0, np.int32)).size == 0
    off, idx = ix1.search(np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert off.tolist() == [0] and idx.size == 0
    assert ix1.count([1, 11], [5, 20]).tolist() == [1, 0]
    # int32 extremes
    big, small = np.iinfo(np.int32).max, np.iinfo(np.int32).min
    s = np.array([small, small, 0, big - 5, big], np.int32)
    e = np.array([small, big, 10, big, big], np.int32)
    qs = np.array([small, small, 5, big, big, 0, big], np.int32)
    qe = np.array([small, big, 5, big, big - 1, small, small], np.int32)
    ixe = sia.DeviceIndex(np.int32).build(s, e)
    oix = orc.build_index(s, e)
    assert np.array_equal(ixe.count(qs, qe), orc.count_batch(oix, qs, qe))
    # the rank form (#{ends < s} through the sorted-ends tree; s = min has none below) at
    # both widths' extremes, for count and for the offsets of a search
    from superintervals_amd import autotune
    saved = autotune.current()
    try:
        for dt in (np.int32, np.int64):
            lo, hi = np.iinfo(dt).min, np.iinfo(dt).max
            s2 = np.array([lo, lo, 0, hi - 5, hi, lo + 1], dt)
            e2 = np.array([lo, hi, 10, hi, hi, lo + 1], dt)
            qs2 = np.array([lo, lo, 5, hi, hi, 0, hi, lo + 1, lo], dt)
            qe2 = np.array([lo, hi, 5, hi, hi - 1, lo, lo, lo + 1, lo + 1], dt)
            ix2 = sia.DeviceIndex(dt).build(s2, e2)
            oix2 = orc.build_index(s2, e2)
            want = orc.count_batch(oix2, qs2, qe2)
            for ub_tree in (1, 0):
                autotune.apply({"count_rank": 1, "offsets_rank": 1, "ub_tree": ub_tree})
                assert np.array_equal(ix2.count(qs2, qe2), want), (dt, ub_tree)
                off, idx = ix2.search(qs2, qe2)
                ooff, oidx = orc.search_batch(oix2, qs2, qe2, mode=0)
                assert np.array_equal(off, ooff) and np.array_equal(idx.astype(np.int64), oidx), (dt, ub_tree)
            ix2.close()
    finally:
        autotune.apply(saved)
    # query before build
    with pytest.raises(sia.SIError):
        sia.DeviceIndex(np.int32).count(q, q)


def _capi():
    import superintervals_amd._native as N
    return N


def test_capi_reference_functions(sia, orc):
    """The reference C-ABI (c_superintervals.h names) on the device index."""
    N = _capi()
    L = N.lib
    rec, s, e, qs, qe = load_golden("c1_shuffled")
    h = L.createSuperIntervals()
    L.reserveSuperIntervals(h, 16)
    for i in range(s.size):
        L.addInterval(h, int(s[i]), int(e[i]), i)
    st = N.cSuperIntervals.from_address(h)
    assert st.size == s.size and not st.startSorted
    L.indexSuperIntervals(h)
    st = N.cSuperIntervals.from_address(h)
    assert st.startSorted and st.endSorted and st.idx == 0
    n = s.size
    hs = np.ctypeslib.as_array(st.starts, (n,)).copy()
    he = np.ctypeslib.as_array(st.ends, (n,)).copy()
    hd = np.ctypeslib.as_array(st.data, (n,)).copy()
    hb = np.ctypeslib.as_array(st.branch, (n,)).copy().view(np.int64)
    assert sha(hs, he) == str(rec["sorted_sha256"])
    assert sha(hb) == str(rec["branch_sha256"])
    assert sha(hd) == str(rec["c_data_sha256"])
    oix = orc.build_index(s, e)
    res = N.cIndexResult()
    for q in range(0, 300):
        a, b = int(qs[q]), int(qe[q])
        assert L.countOverlaps(h, a, b) == rec["counts"][q]
        L.clearIndexResult(C.byref(res))
        L.searchIdxs(h, a, b, C.byref(res))
        got = [res.data[k] for k in range(res.size)]
        assert got == orc.search(oix, a, b, 0).tolist()
        L.clearIndexResult(C.byref(res))
        L.searchValues(h, a, b, C.byref(res))
        assert [res.data[k] for k in range(res.size)] == orc.values(oix, a, b).tolist()
        ub = L.upperBound(h, b)
        want = orc.upper_bound(oix, b)
        assert (ub == 2 ** 64 - 1 and want == -1) or ub == want
        assert N.cSuperIntervals.from_address(h).idx == ub
        assert L.anyOverlaps(h, a, b) == orc.has_overlaps(oix, a, b)
        cnt = C.c_size_t(0)
        cov = C.c_int32(0)
        L.coverage(h, a, b, C.byref(cnt), C.byref(cov))
        assert (cnt.value, cov.value) == orc.coverage(oix, a, b)
    # appending semantics: results accumulate until cleared (c_superintervals.h:209-210)
    L.clearIndexResult(C.byref(res))
    L.searchIdxs(h, int(qs[0]), int(qe[0]), C.byref(res))
    L.searchIdxs(h, int(qs[1]), int(qe[1]), C.byref(res))
    assert res.size == rec["counts"][0] + rec["counts"][1]
    L.destroyIndexResult(C.byref(res))
    keys = N.cKeyResult()
    items = N.cItemResult()
    L.searchKeys(h, int(qs[5]), int(qe[5]), C.byref(keys))
    L.searchItems(h, int(qs[5]), int(qe[5]), C.byref(items))
    strict = orc.search(oix, qs[5], qe[5], 0)
    assert [(keys.data[k].start, keys.data[k].end) for k in range(keys.size)] == \
        [(int(oix["starts"][j]), int(oix["ends"][j])) for j in strict]
    assert [items.data[k].data for k in range(items.size)] == [int(oix["data"][j]) for j in strict]
    L.destroyKeyResult(C.byref(keys))
    L.destroyItemResult(C.byref(items))
    # batch extensions
    counts = np.empty(qs.size, np.uint32)
    assert L.siCountBatch(h, qs.ctypes.data, qe.ctypes.data, qs.size, counts.ctypes.data) == 0
    assert np.array_equal(counts, rec["counts"])
    off = np.empty(qs.size + 1, np.uint64)
    tot = C.c_uint64(0)
    rc = L.siSearchIdxsBatch(h, qs.ctypes.data, qe.ctypes.data, qs.size, off.ctypes.data, None, 0, C.byref(tot))
    assert rc == N.SI_ERR_CAPACITY and tot.value == int(rec["counts"].sum())
    idx = np.empty(tot.value, np.int32)
    assert L.siSearchIdxsBatch(h, qs.ctypes.data, qe.ctypes.data, qs.size, off.ctypes.data, idx.ctypes.data,
                               idx.size, C.byref(tot)) == 0
    assert orc.digest(orc.hash_csr(off, idx.astype(np.int64))) == int(rec["hash_desc_digest"])
    vals = np.empty(tot.value, np.int32)
    assert L.siSearchValuesBatch(h, qs.ctypes.data, qe.ctypes.data, qs.size, off.ctypes.data,
                                 vals.ctypes.data, vals.size, C.byref(tot)) == 0
    assert orc.digest(orc.hash_csr(off, vals)) == int(rec["hash_cvalues_digest"])
    L.destroySuperIntervals(h)


def test_capi_empty_and_unindexed(sia):
    N = _capi()
    L = N.lib
    h = L.createSuperIntervals()
    L.indexSuperIntervals(h)              # empty: no-op, as the reference
    assert L.countOverlaps(h, 1, 5) == 0
    assert not L.anyOverlaps(h, 1, 5)
    L.addInterval(h, 1, 10, 7)
    assert L.countOverlaps(h, 1, 5) == 0   # not indexed yet: no result, error recorded
    assert b"indexSuperIntervals" in L.siGetLastError()
    L.indexSuperIntervals(h)
    assert L.countOverlaps(h, 1, 5) == 1 and L.countOverlaps(h, 11, 20) == 0
    L.destroySuperIntervals(h)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_single_query_path(sia, orc, dtype):
    """One-query host calls run one wave on the index's scratch: every mode against the
    batched path and the oracle, repeated calls (the capacity probe), a result larger than
    the scratch (the batched path takes it), an index with start > end intervals."""
    import superintervals_amd._native as N
    print(f"[gpu] single-query path dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(8)
    n = 300_000
    s = rng.integers(0, 3_000_000, n)
    e = s + rng.geometric(1 / 200, n) - 1
    inv = rng.random(n) < 0.01
    base = (np.int64(5) << 33) if dtype == np.int64 else 0
    for with_inv in (False, True):
        ee = np.where(inv, s - 3, e) if with_inv else e
        ss_, ee_ = (s + base).astype(dtype), (ee + base).astype(dtype)
        ix = _dev_index(sia, ss_, ee_, np.arange(n, dtype=np.int32))
        oix = orc.build_index(ss_, ee_)
        qs = rng.integers(-10, 3_000_000, 400)
        qe = qs + rng.integers(0, 5000, qs.size)
        qe[::13] = qs[::13] - 2                                  # inverted queries
        qs[0], qe[0] = 0, 3_000_000                               # every interval: > 65536 results
        qs, qe = (qs + base).astype(dtype), (qe + base).astype(dtype)
        for q in range(qs.size):
            a, b = qs[q:q + 1], qe[q:q + 1]
            assert ix.count(a, b)[0] == orc.count(oix, a[0], b[0])
            for order, mode in (("desc", 0), ("vector", 1)):
                off, idx = ix.search(a, b, order=order)   # probe + fill: the second call reuses the first
                assert np.array_equal(idx.astype(np.int64), orc.search(oix, a[0], b[0], mode)), (q, order)
            off, vals = ix.search(a, b, values=True)
            assert np.array_equal(vals, oix["data"][orc.search(oix, a[0], b[0], 0)])
            ub = int(ix.upper_bound(b)[0])
            want = orc.upper_bound(oix, b[0])
            assert (ub == 0xFFFFFFFF and want == -1) or ub == want
            assert bool(ix.has_overlaps(a, b)[0]) == orc.has_overlaps(oix, a[0], b[0])
        # capacity probe through the C-ABI-style call: too small, then exact, both right
        tot = C.c_uint64(0)
        o2 = np.zeros(2, np.uint64)
        small = np.zeros(1, np.uint32)
        a, b = qs[5:6], qe[5:6]
        want = orc.search(oix, a[0], b[0], 0)
        if want.size > 1:
            rc = N.lib.siSearchHost(ix.handle, a.ctypes.data, b.ctypes.data, 1, 0, o2.ctypes.data, small.ctypes.data,
                                    None, 1, C.byref(tot))
            assert rc == N.SI_ERR_CAPACITY and tot.value == want.size and o2[1] == want.size
            big = np.zeros(want.size, np.uint32)
            assert N.lib.siSearchHost(ix.handle, a.ctypes.data, b.ctypes.data, 1, 0, o2.ctypes.data, big.ctypes.data,
                                      None, big.size, C.byref(tot)) == 0
            assert np.array_equal(big.astype(np.int64), want)
        ix.close()


def test_capi_stale_handle_after_clear_and_add(sia):
    """A handle changed after indexSuperIntervals has no device index until it is indexed
    again, even when the new contents have the old size."""
    N = _capi()
    L = N.lib
    h = L.createSuperIntervals()
    for s0, e0 in [(1, 10), (5, 20), (30, 40)]:
        L.addInterval(h, s0, e0, 0)
    L.indexSuperIntervals(h)
    assert L.countOverlaps(h, 8, 9) == 2
    L.clearSuperIntervals(h)
    for s0, e0 in [(100, 110), (105, 120), (130, 140)]:
        L.addInterval(h, s0, e0, 0)
    assert L.countOverlaps(h, 8, 9) == 0 and L.countOverlaps(h, 106, 107) == 0
    assert b"indexSuperIntervals" in L.siGetLastError()
    L.indexSuperIntervals(h)
    assert L.countOverlaps(h, 106, 107) == 2 and L.countOverlaps(h, 8, 9) == 0
    L.addInterval(h, 1, 500, 0)           # one more after indexing: stale again
    assert L.countOverlaps(h, 106, 107) == 0
    L.indexSuperIntervals(h)
    assert L.countOverlaps(h, 106, 107) == 3
    L.destroySuperIntervals(h)


def _extras():
    import os
    from conftest import GOLDEN
    return dict(np.load(os.path.join(GOLDEN, "extras.npz")))


@pytest.mark.parametrize("lane_steps", [256, 2])
def test_index_with_inverted_intervals_golden(sia, orc, lane_steps):
    """An index in which 15 % of the intervals have start > end: every count form and both
    orders against the reference's outputs (the rank forms are not exact there, so the
    library walks)."""
    from superintervals_amd import autotune
    print(f"[gpu] inverted intervals lane_steps={lane_steps}", flush=True)
    x = _extras()
    s, e, qs, qe = x["inv_starts"], x["inv_ends"], x["inv_qs"], x["inv_qe"]
    ix = _dev_index(sia, s, e, np.arange(s.size, dtype=np.int32))
    saved = autotune.current()
    try:
        autotune.apply({"lane_steps": lane_steps})
        for cr in (0, 1):
            for orank in (0, 1):
                autotune.apply({"count_rank": cr, "offsets_rank": orank})
                assert np.array_equal(ix.count(qs, qe).astype(np.int64), x["inv_counts"]), (cr, orank)
                off, idx = ix.search(qs, qe, order="desc")
                assert np.array_equal(off, x["inv_desc_off"]) and np.array_equal(idx.astype(np.int32), x["inv_desc_idx"])
                off, idx = ix.search(qs, qe, order="vector")
                assert np.array_equal(off, x["inv_vec_off"]) and np.array_equal(idx.astype(np.int32), x["inv_vec_idx"])
                off, vals = ix.search(qs, qe, values=True)
                assert np.array_equal(off, x["inv_cval_off"]) and np.array_equal(vals, x["inv_cval"])
    finally:
        autotune.apply(saved)
    ix.close()
    # the C-ABI on the same intervals
    N = _capi()
    L = N.lib
    h = L.createSuperIntervals()
    for i in range(s.size):
        L.addInterval(h, int(s[i]), int(e[i]), i)
    L.indexSuperIntervals(h)
    res = N.cIndexResult()
    off = x["inv_cidx_off"].astype(np.int64)
    for q in range(0, qs.size, 7):
        L.clearIndexResult(C.byref(res))
        L.searchIdxs(h, int(qs[q]), int(qe[q]), C.byref(res))
        assert [res.data[k] for k in range(res.size)] == x["inv_cidx"][off[q]:off[q + 1]].tolist()
        assert L.countOverlaps(h, int(qs[q]), int(qe[q])) == x["inv_counts"][q]
    L.destroyIndexResult(C.byref(res))
    L.destroySuperIntervals(h)


@pytest.mark.parametrize("order", ["desc", "vector"])
def test_fill_reports_offsets_that_disagree(sia, orc, order):
    """A fill never writes outside a query's segment; with offsets that disagree with the
    index (one query given a slot too many, or one too few) siPlanCheck reports it."""
    import torch
    N = _capi()
    print(f"[gpu] fill/offsets mismatch order={order}", flush=True)
    rng = np.random.default_rng(5)
    s = np.sort(rng.integers(0, 100_000, 20_000)).astype(np.int32)
    e = (s + rng.integers(0, 500, s.size)).astype(np.int32)
    qs = rng.integers(0, 100_000, 5000).astype(np.int32)
    qe = (qs + rng.integers(0, 300, qs.size)).astype(np.int32)
    ix = _dev_index(sia, s, e)
    dev = torch.device("cuda:0")
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    nq = qs.size
    plan = torch.empty(ix.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    ix.plan_device(tqs, tqe, plan)
    ix.offsets_planned(plan, nq, off)
    torch.cuda.synchronize()
    counts = np.diff(off.cpu().numpy())
    q = int(np.flatnonzero(counts >= 2)[0])
    total = int(off[-1].item())
    idx = torch.full((total + 1,), -1, dtype=torch.int32, device=dev)
    ix.fill_planned(plan, nq, off, idx, order=order)
    assert N.lib.siPlanCheck(ix.handle, plan.data_ptr(), None) == N.SI_OK
    ref = idx[:total].clone()
    # query q given one slot too many: its list leaves a gap
    bad = off.clone()
    bad[q + 1:] += 1
    idx.fill_(-1)
    ix.fill_planned(plan, nq, bad, idx, order=order)
    assert N.lib.siPlanCheck(ix.handle, plan.data_ptr(), None) == N.SI_ERR_MISMATCH
    # query q given one slot too few (q + 1 one more): q's list is cut, never overrun
    bad = off.clone()
    bad[q + 1] -= 1
    idx.fill_(-1)
    ix.fill_planned(plan, nq, bad, idx, order=order)
    assert N.lib.siPlanCheck(ix.handle, plan.data_ptr(), None) == N.SI_ERR_MISMATCH
    o0, o1 = int(off[q].item()), int(off[q + 1].item())
    assert torch.equal(idx[o0:o1 - 1], ref[o0:o1 - 1])
    # and a correct fill clears the report
    ix.fill_planned(plan, nq, off, idx, order=order)
    assert N.lib.siPlanCheck(ix.handle, plan.data_ptr(), None) == N.SI_OK
    assert torch.equal(idx[:total], ref)
    ix.close()


def test_python_intervalmap_mirror(sia, orc):
    """The Python IntervalMap mirrors intervalmap.pyx on the reference's examples."""
    m = sia.IntervalMap()
    for s, e, v in [(10, 20, "a"), (11, 12, None), (13, 14, None), (15, 16, None), (25, 29, "d")]:
        m.add(s, e, v)
    m.build()
    assert m.has_overlaps(17, 30) and not m.has_overlaps(1, 3)
    assert m.search_values(17, 30) == ["d", "a"]
    assert m.count(17, 30) == 2
    assert m.search_idxs(17, 30) == [4, 0]
    assert m.coverage(10, 29)[1] == 17
    m2 = sia.IntervalMap.from_arrays([1, 2, 4, 5, 6, 60], [100, 3, 50, 50, 50, 70])
    assert m2.search_idxs(10, 65) == [2, 3, 4, 5, 0]          # vector order (pyx uses it)
    assert [x.tolist() for x in m2.search_idxs_batch([10, 1], [65, 1])] == [[2, 3, 4, 5, 0], [0]]
    assert m2.count_batch([10, 200], [65, 300]).tolist() == [5, 0]
    with pytest.raises(ValueError):
        m2.count_batch([1, 2], [3])


def test_cpp_api_program(sia):
    """include/superintervals.hpp (si::IntervalMap<S, T>) through a compiled C++ program."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "cpp", "test_intervalmap")
    assert os.path.exists(exe), "build the C++ test first (__graft_entry__.build())"
    print("[gpu] C++ API program", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


@pytest.mark.parametrize("knobs", [None, {"tile": 512, "count_wgs": 0, "fill_wgs": 0, "lane_steps": 1024,
                                          "count_rank": 1, "plan_sync": 0, "plan_skip_low": 1}],
                         ids=["defaults", "benched"])
def test_full_size_lognormal_10Mx10M_properties(sia, orc, knobs):
    """10M intervals x 10M queries (log-normal lengths): size-independent properties of the whole
    output plus an oracle check of a query sample, with the library's default launch choices
    and with those the headline benchmark's autotuner picks (512-query tiles, no plan
    read-back, the lowest end digit unsorted)."""
    import torch
    from superintervals_amd import autotune
    print(f"[gpu] full-size log-normal 10M x 10M knobs={knobs}", flush=True)
    saved = autotune.current()
    if knobs:
        autotune.apply(knobs)
    try:
        _full_size_lognormal(sia, orc)
    finally:
        autotune.apply(saved)


def _full_size_lognormal(sia, orc):
    import torch
    n = nq = 10_000_000
    s, e = sia.workload_intervals(2, n, seed=1, sorted=True)
    qs, qe = sia.workload_queries(2, nq, seed=2)
    ix = _dev_index(sia, s, e)
    dev = torch.device("cuda:0")
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)
    ix.offsets_device(tqs, tqe, off, counts)
    torch.cuda.synchronize()
    total = int(off[-1].item())
    # closed form: count = #starts <= e - #ends < s (all queries have s <= e)
    ss = torch.from_numpy(s).to(dev)
    se = torch.sort(torch.from_numpy(e).to(dev)).values
    closed = torch.searchsorted(ss, tqe, right=True) - torch.searchsorted(se, tqs, right=False)
    assert torch.equal(closed.to(torch.int32), counts)
    assert torch.equal(off[1:] - off[:-1], counts.to(torch.int64))
    idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    ix.fill_device(tqs, tqe, off, idx, order="desc")
    torch.cuda.synchronize()
    d = ix.device_arrays()
    ends_t = torch.from_numpy(ix.download()["ends"]).to(dev)
    starts_t = ss
    # every emitted position overlaps its query
    qid = torch.repeat_interleave(torch.arange(nq, device=dev), counts.to(torch.int64))
    pos = idx[:total].to(torch.int64)
    assert bool((starts_t[pos] <= tqe[qid]).all()) and bool((ends_t[pos] >= tqs[qid]).all())
    # strictly descending inside every list
    same = qid[1:] == qid[:-1]
    assert bool((pos[1:][same] < pos[:-1][same]).all())
    # oracle on a sample
    oix = orc.build_index(s, e)
    sample = np.arange(0, nq, 4999)
    ooff, oidx = orc.search_batch(oix, qs[sample], qe[sample], mode=0)
    offh = off.cpu().numpy().astype(np.uint64)
    idxh = idx[:total].cpu().numpy()
    for k, q in enumerate(sample):
        assert np.array_equal(idxh[offh[q]:offh[q + 1]].astype(np.int64), oidx[ooff[k]:ooff[k + 1]])
    del d
    ix.close()


def test_output_offsets_beyond_2_31(sia):
    """Lists whose output offsets pass 2^31 entries (8.7 GB of positions) on n = 2^24 + 5
    intervals [i, n + i], with queries in triples: [0, n - 1] hits all n as one guaranteed
    run longer than 2^23 (written on its own by the whole wave), [0, 2^21 - 1] hits 2^21 as
    a run of whole 64-slot chunks, and the point 2^23, whose 2^23 hits below its run are
    walked by the wave-cooperative kernel.  Both orders; every entry is checked against the
    closed form."""
    import torch
    print("[gpu] output offsets beyond 2^31", flush=True)
    n, r2, p3, trip = (1 << 24) + 5, 1 << 21, 1 << 23, 80
    s = np.arange(n, dtype=np.int32)
    e = (s + n).astype(np.int32)
    ix = _dev_index(sia, s, e)
    qs = np.zeros(3 * trip, np.int32)
    qe = np.tile(np.array([n - 1, r2 - 1, p3], np.int32), trip)
    qs[2::3] = p3
    dev = torch.device("cuda:0")
    tqs, tqe = torch.from_numpy(qs).to(dev), torch.from_numpy(qe).to(dev)
    off = torch.empty(qs.size + 1, dtype=torch.int64, device=dev)
    ix.offsets_device(tqs, tqe, off)
    torch.cuda.synchronize()
    per = n + r2 + p3 + 1
    total = int(off[-1].item())
    assert total == trip * per and total > (1 << 31)
    idx = torch.empty(total, dtype=torch.int32, device=dev)
    rows = idx.view(trip, per)
    desc = torch.cat([torch.arange(n - 1, -1, -1), torch.arange(r2 - 1, -1, -1), torch.arange(p3, -1, -1)])
    vec = torch.cat([torch.arange(n), torch.arange(r2), torch.arange(p3 + 1)])
    for order, row in (("desc", desc), ("vector", vec)):
        idx.fill_(-1)
        ix.fill_device(tqs, tqe, off, idx, order=order)
        torch.cuda.synchronize()
        assert bool((rows == row.to(torch.int32).to(dev)).all()), order
    del rows, idx
    ix.close()
