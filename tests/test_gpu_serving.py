"""GPU tests of the continuous-batching server (superintervals_amd/serving.py).

Each request's result in a Poisson mix of requests over several resident interval sets
must equal the result of the same request alone (tolerance 0: integer outputs), here the
CPU oracle's answer on that set.  Sets are int32 indexes in the slots of a set table;
batches replay captured graphs per shape.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import oracle
    import superintervals_amd as sia
    from superintervals_amd import serving
    rng = np.random.default_rng(5)
    sizes = serving.log_uniform_int(rng, 20_000, 200_000, 6)
    sets = [sia.workload_intervals(2, int(n), seed=100 + k, sorted=bool(k % 2)) for k, n in enumerate(sizes)]
    res = serving.ResidentSets(sets, max_sets=16)
    O = oracle.Oracle()
    oix = {k: O.build_index(s, e) for k, (s, e) in enumerate(sets)}
    reqs = []
    for k in range(60):
        i = int(rng.integers(0, len(sets)))
        m = int(serving.log_uniform_int(rng, 1, 20_000))
        qs, qe = sia.workload_queries(2, m, seed=500 + k)
        reqs.append((i, qs, qe, "count" if k % 7 == 3 else "search"))
    return sia, serving, res, O, oix, reqs


def _check(O, oix, reqs, done, rids=None):
    by_id = {r.rid: r for r in done}
    rids = list(range(len(reqs))) if rids is None else rids
    assert len(done) == len(rids)
    for rid, (i, qs, qe, kind) in zip(rids, reqs):
        r = by_id[rid]
        assert r.index_id == i
        if kind == "count":
            assert np.array_equal(r.counts, O.count_batch(oix[i], qs, qe))
        else:
            off, idx = O.search_batch(oix[i], qs, qe, mode=0)
            assert np.array_equal(r.offsets, off), rid
            assert np.array_equal(r.positions().astype(np.int64), idx), rid


def test_mix_equals_alone(setup):
    sia, serving, res, O, oix, reqs = setup
    print("[gpu] serving: Poisson mix, large budget", flush=True)
    srv = serving.Server(res, hbm_budget_bytes=4 << 30, streams=3)
    arrivals = serving.poisson_arrivals(2000.0, len(reqs), seed=9)
    lat, wall = serving.serve_stream(srv, reqs, arrivals, timeout_s=120)
    _check(O, oix, reqs, srv.completed)
    assert srv.stats["batches1"] >= 1 and srv.in_use == 0
    assert srv.stats["graphs"] >= 2
    assert np.all(lat >= 0)


def test_tight_budget_queues_but_stays_exact(setup):
    sia, serving, res, O, oix, reqs = setup
    print("[gpu] serving: tight budget", flush=True)
    biggest = res.stage1_bytes(max(q.size for _, q, _, _ in reqs))
    srv = serving.Server(res, hbm_budget_bytes=biggest + 4096, streams=2)
    for i, qs, qe, kind in reqs:
        srv.submit(i, qs, qe, kind)
    srv.run(until_idle=True)
    _check(O, oix, reqs, srv.completed)
    assert srv.stats["batches1"] > 1          # the budget forced several batches
    assert srv.in_use == 0
    with pytest.raises(MemoryError):
        srv.submit(0, np.zeros(1 << 24, np.int32), np.zeros(1 << 24, np.int32))


def test_larger_mix_with_sets_added_and_evicted(setup):
    """More and larger sets (up to 1.5M intervals, one unsorted with duplicates), larger
    requests, and sets added and evicted while the server runs; every result exact."""
    sia, serving, res0, O, oix0, _ = setup
    print("[gpu] serving: larger mix, add / evict", flush=True)
    rng = np.random.default_rng(77)
    sizes = serving.log_uniform_int(rng, 50_000, 1_500_000, 10)
    sets = {k: sia.workload_intervals(2, int(n), seed=300 + k, sorted=bool(k % 3)) for k, n in enumerate(sizes)}
    res = serving.ResidentSets(max_sets=32)
    for k, (s, e) in sets.items():
        res.add(k, s, e)
    oix = {k: O.build_index(s, e) for k, (s, e) in sets.items()}
    srv = serving.Server(res, hbm_budget_bytes=2 << 30, streams=4)

    def make(k, slots):
        i = int(rng.choice(slots))
        m = int(serving.log_uniform_int(rng, 1, 200_000))
        qs, qe = sia.workload_queries(2, m, seed=900 + k)
        return (i, qs, qe, "count" if k % 11 == 5 else "search")

    first = [make(k, list(sets)) for k in range(40)]
    arrivals = serving.poisson_arrivals(400.0, len(first), seed=3)
    serving.serve_stream(srv, first, arrivals, timeout_s=300)
    _check(O, oix, first, srv.completed)
    # evict two sets, add a new one in a free slot and replace another
    srv.evict_set(3)
    srv.evict_set(7)
    with pytest.raises(ValueError):
        srv.submit(3, np.zeros(4, np.int32), np.ones(4, np.int32))
    s_new, e_new = sia.workload_intervals(2, 800_000, seed=999, sorted=False)
    srv.add_set(20, s_new, e_new)
    s_rep, e_rep = sia.workload_intervals(2, 300_000, seed=998, sorted=True)
    srv.add_set(1, s_rep, e_rep)
    oix[20] = O.build_index(s_new, e_new)
    oix[1] = O.build_index(s_rep, e_rep)
    live = [k for k in list(sets) + [20] if k not in (3, 7)]
    second = [make(100 + k, live) for k in range(40)]
    n0 = len(srv.completed)
    srv.completed.clear()
    arrivals = serving.poisson_arrivals(400.0, len(second), seed=4)
    serving.serve_stream(srv, second, arrivals, timeout_s=300)
    _check(O, oix, second, srv.completed, rids=list(range(n0, n0 + len(second))))
    assert srv.in_use == 0
    res.table.close()
