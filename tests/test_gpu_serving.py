"""GPU tests of the continuous-batching server (superintervals_amd/serving.py).

Each request's result in a Poisson mix of requests over several resident interval sets
must equal the result of the same request alone (tolerance 0: integer outputs), here the
CPU oracle's answer on that set.
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
    res = serving.ResidentSet(sets)
    O = oracle.Oracle()
    oix = [O.build_index(s, e) for s, e in sets]
    reqs = []
    for k in range(60):
        i = int(rng.integers(0, len(sets)))
        m = int(serving.log_uniform_int(rng, 1, 20_000))
        qs, qe = sia.workload_queries(2, m, seed=500 + k)
        reqs.append((i, qs, qe, "count" if k % 7 == 3 else "search"))
    return sia, serving, res, O, oix, reqs


def _check(O, oix, reqs, done):
    assert len(done) == len(reqs)
    by_id = {r.rid: r for r in done}
    for rid, (i, qs, qe, kind) in enumerate(reqs):
        r = by_id[rid]
        assert r.index_id == i
        if kind == "count":
            assert np.array_equal(r.counts, O.count_batch(oix[i], qs, qe))
        else:
            off, idx = O.search_batch(oix[i], qs, qe, mode=0)
            assert np.array_equal(r.offsets, off), rid
            assert np.array_equal(r.positions().astype(np.int64), idx), rid


def test_resident_set_positions(setup):
    sia, serving, res, O, oix, reqs = setup
    a = res.index.download()
    for k in range(len(oix)):
        b0, b1 = res.base[k], res.base[k + 1]
        assert np.array_equal(a["starts"][b0:b1] - k * serving.SHIFT, oix[k]["starts"])
        br = a["branch"][b0:b1]
        want = oix[k]["branch"]
        assert np.array_equal(np.where(br < 0, -1, br - b0), want)


def test_mix_equals_alone(setup):
    sia, serving, res, O, oix, reqs = setup
    print("[gpu] serving: Poisson mix, large budget", flush=True)
    srv = serving.Server(res, hbm_budget_bytes=4 << 30, streams=3)
    arrivals = serving.poisson_arrivals(2000.0, len(reqs), seed=9)
    lat, wall = serving.serve_stream(srv, reqs, arrivals, timeout_s=120)
    _check(O, oix, reqs, srv.completed)
    assert srv.stats["batches1"] >= 1 and srv.in_use == 0
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
        srv.submit(0, np.zeros(1 << 20, np.int32), np.zeros(1 << 20, np.int32))
