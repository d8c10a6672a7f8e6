"""Set tables (siSetTable): one batch over many resident int32 indexes.  Every query's count
and position list must equal its own set's DeviceIndex results (and the oracle's), for any
mix of slots, both orders, the lane and wave walks, an empty slot, out-of-range slot numbers,
a set holding intervals with start > end, put / evict / replace, and graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sia():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import superintervals_amd
    return superintervals_amd


def _make_sets(rng):
    sets = []
    for n, span, inv in ((200_000, 5_000_000, 0.0), (30_000, 200_000, 0.0), (1, 100, 0.0), (70_000, 1_000_000, 0.1),
                         (5000, 50_000, 0.0)):
        s = rng.integers(0, span, n)
        e = s + rng.geometric(1 / 300, n) - 1
        e = np.where(rng.random(n) < inv, s - rng.integers(1, 100, n), e)
        sets.append((s.astype(np.int32), e.astype(np.int32)))
    return sets


def _queries(rng, nq, nslots):
    sid = rng.integers(0, nslots + 2, nq).astype(np.uint32)    # a few ids past the table
    qs = rng.integers(-100, 5_000_000, nq)
    qe = qs + rng.integers(0, 20_000, nq)
    qe[::97] = qs[::97] - 5                                      # some inverted queries
    return sid, qs.astype(np.int32), qe.astype(np.int32)


def _expect(orc, oixs, sid, qs, qe, mode):
    """Per-query lists from each set's oracle index (empty for empty / unknown slots)."""
    counts = np.zeros(qs.size, np.int64)
    lists = [None] * qs.size
    for g in np.unique(sid):
        sel = np.flatnonzero(sid == g)
        if g < len(oixs) and oixs[g] is not None:
            off, idx = orc.search_batch(oixs[g], qs[sel], qe[sel], mode=mode)
            for j, q in enumerate(sel):
                lists[q] = idx[off[j]:off[j + 1]]
                counts[q] = off[j + 1] - off[j]
        else:
            for q in sel:
                lists[q] = np.zeros(0, np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    return off, np.concatenate(lists).astype(np.int64)


@pytest.mark.parametrize("lane_steps", [1024, 3])
def test_set_table_matches_each_set(sia, orc, lane_steps):
    from superintervals_amd import autotune
    print(f"[gpu] set table lane_steps={lane_steps}", flush=True)
    rng = np.random.default_rng(21)
    sets = _make_sets(rng)
    nslots = 8
    table = sia.SetTable(nslots)
    idxs, oixs = [], [None] * nslots
    for g, (s, e) in enumerate(sets):
        ix = sia.DeviceIndex(np.int32).build(s, e)
        idxs.append(ix)
        table.put(g, ix)
        oixs[g] = orc.build_index(s, e)
    sid, qs, qe = _queries(rng, 120_000, nslots)
    saved = autotune.current()
    try:
        autotune.apply({"lane_steps": lane_steps})
        for order, mode in (("desc", 0), ("vector", 1)):
            off, idx = table.search(sid, qs, qe, order=order)
            woff, widx = _expect(orc, oixs, sid, qs, qe, mode)
            assert np.array_equal(off, woff)
            assert np.array_equal(idx.astype(np.int64), widx)
        # each set's queries alone through its own index give the same lists
        off, idx = table.search(sid, qs, qe, order="desc")
        for g in range(len(sets)):
            sel = np.flatnonzero(sid == g)[:3000]
            o1, i1 = idxs[g].search(qs[sel], qe[sel], order="desc")
            for j, q in enumerate(sel):
                assert np.array_equal(idx[off[q]:off[q + 1]], i1[o1[j]:o1[j + 1]])
        # evict slot 1, replace slot 0 with set 4's index: other slots unchanged
        table.put(1, None)
        table.put(0, idxs[4])
        oixs2 = list(oixs)
        oixs2[1] = None
        oixs2[0] = oixs[4]
        off, idx = table.search(sid, qs, qe, order="desc")
        woff, widx = _expect(orc, oixs2, sid, qs, qe, 0)
        assert np.array_equal(off, woff) and np.array_equal(idx.astype(np.int64), widx)
    finally:
        autotune.apply(saved)
    # a rebuilt index must be put again
    idxs[2].build(sets[2][0], sets[2][1])
    with pytest.raises(sia.SIError):
        table.search(sid[:10], qs[:10], qe[:10])
    table.put(2, idxs[2])
    table.search(sid[:10], qs[:10], qe[:10])
    table.close()
    for ix in idxs:
        ix.close()


def test_set_table_graph_replays(sia, orc):
    """Offsets and fill captured once with fixed buffers, replayed for new query contents of
    the same shape (eager work between replays); every replay equals the oracle."""
    import torch
    from superintervals_amd import autotune
    print("[gpu] set table graph replays", flush=True)
    rng = np.random.default_rng(33)
    sets = _make_sets(rng)[:3]
    table = sia.SetTable(4)
    idxs, oixs = [], [None] * 4
    for g, (s, e) in enumerate(sets):
        ix = sia.DeviceIndex(np.int32).build(s, e)
        idxs.append(ix)
        table.put(g, ix)
        oixs[g] = orc.build_index(s, e)
    nq = 8192
    dev = torch.device("cuda:0")
    sid_t = torch.zeros(nq, dtype=torch.int32, device=dev)
    qs_t = torch.zeros(nq, dtype=torch.int32, device=dev)
    qe_t = torch.zeros(nq, dtype=torch.int32, device=dev)
    plan = torch.empty(table.plan_bytes(nq), dtype=torch.uint8, device=dev)
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    cap = 4_000_000
    idx = torch.empty(cap, dtype=torch.int32, device=dev)

    def load(seed):
        r = np.random.default_rng(seed)
        sid, qs, qe = _queries(r, nq, 4)
        sid_t.copy_(torch.from_numpy(sid.view(np.int32)))
        qs_t.copy_(torch.from_numpy(qs))
        qe_t.copy_(torch.from_numpy(qe))
        return sid, qs, qe

    load(1)
    s0 = torch.cuda.Stream()
    with torch.cuda.stream(s0):                     # warm-up outside the capture
        table.offsets_device(sid_t, qs_t, qe_t, plan, off, stream=s0)
        table.fill_device(plan, nq, off, idx, capacity=cap, stream=s0)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s0):
        table.offsets_device(sid_t, qs_t, qe_t, plan, off, stream=s0)
        table.fill_device(plan, nq, off, idx, capacity=cap, stream=s0)
    for seed in (2, 3, 4, 5):
        sid, qs, qe = load(seed)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        total = int(off[-1].item())
        assert total <= cap
        woff, widx = _expect(orc, oixs, sid, qs, qe, 0)
        assert np.array_equal(off.cpu().numpy().astype(np.uint64), woff)
        assert np.array_equal(idx[:total].cpu().numpy().astype(np.int64), widx)
        _ = (idx[:1000].to(torch.int64) * 3).sum().item()   # eager work between replays
    del g
    table.close()
    for ix in idxs:
        ix.close()
