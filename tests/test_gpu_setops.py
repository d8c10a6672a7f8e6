"""Set operations over a batched search: intersection and difference
(superintervals.hpp:1158-1211) on the GPU against the oracle's search and brute force.

intersection: every overlapping (A interval k, B interval h) pair gives the piece
[max(a_s, b_s), min(a_e, b_e)], for k in A's order and h in search_idxs (vector) order.
Pinned by brute force: the (k, h) pairs are exactly the overlapping pairs, each piece is the
pair's geometric overlap, and the order is the reference's (the oracle's vector-order
search).

difference: the parts of each A interval that no B interval covers.  Pinned by brute force
on small coordinates: the pieces of interval k cover exactly its points that no B interval
holds, are disjoint, ascending and maximal (separated by covered points), and their order
is the reference's sweep over the oracle's descending search.
"""
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


def _sets(seed, dtype, na=300, nb=400, span=3000):
    rng = np.random.default_rng(seed)
    a_s = rng.integers(0, span, na)
    a_e = a_s + rng.integers(0, 120, na)
    b_s = rng.integers(0, span, nb)
    b_e = b_s + rng.integers(0, 60, nb)
    base = (np.int64(5) << 36) if dtype == np.int64 else 0
    return tuple((x + base).astype(dtype) for x in (a_s, a_e, b_s, b_e))


def _reference_difference(a_s, a_e, ob_s, ob_e, hits_desc):
    """The reference sweep (superintervals.hpp:1188-1211) over B's keys sorted ascending."""
    out = []
    for k in range(a_s.size):
        cover = sorted((int(ob_s[h]), int(ob_e[h])) for h in hits_desc[k])
        s, e = int(a_s[k]), int(a_e[k])
        cursor = s
        for cs0, ce0 in cover:
            cs, ce = max(cs0, s), min(ce0, e)
            if cs > cursor:
                out.append((cursor, cs - 1, k))
            if ce + 1 > cursor:
                cursor = ce + 1
        if cursor <= e:
            out.append((cursor, e, k))
    return out


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("seed", [1, 2])
def test_intersection_pieces(sia, orc, dtype, seed):
    print(f"[gpu] intersection dtype={np.dtype(dtype).name} seed={seed}", flush=True)
    a_s, a_e, b_s, b_e = _sets(seed, dtype)
    ix = sia.DeviceIndex(dtype).build(b_s, b_e)
    oix = orc.build_index(b_s, b_e)
    ps, pe, pk, ph = ix.intersection_pieces(a_s, a_e)
    ob_s, ob_e = oix["starts"], oix["ends"]
    # brute force: the (k, h) pairs are exactly the overlapping pairs
    want_pairs = {(k, h) for k in range(a_s.size) for h in range(ob_s.size)
                  if ob_s[h] <= a_e[k] and a_s[k] <= ob_e[h]}
    assert set(zip(pk.tolist(), ph.tolist())) == want_pairs and len(want_pairs) == pk.size
    # each piece is the pair's overlap
    assert np.array_equal(ps, np.maximum(a_s[pk], ob_s[ph])) and np.array_equal(pe, np.minimum(a_e[pk], ob_e[ph]))
    # order: k ascending, hits in the vector order of the oracle's search
    ooff, oidx = orc.search_batch(oix, a_s, a_e, mode=1)
    assert np.array_equal(pk, np.repeat(np.arange(a_s.size), np.diff(ooff).astype(np.int64)))
    assert np.array_equal(ph, oidx)
    ix.close()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("seed", [3, 4])
def test_difference_pieces(sia, orc, dtype, seed):
    print(f"[gpu] difference dtype={np.dtype(dtype).name} seed={seed}", flush=True)
    a_s, a_e, b_s, b_e = _sets(seed, dtype)
    ix = sia.DeviceIndex(dtype).build(b_s, b_e)
    oix = orc.build_index(b_s, b_e)
    ps, pe, pk = ix.difference_pieces(a_s, a_e)
    base = int(min(a_s.min(), b_s.min()))
    span = int(max(a_e.max(), b_e.max())) - base + 2
    covered = np.zeros(span, bool)
    for s0, e0 in zip(b_s.tolist(), b_e.tolist()):
        covered[s0 - base:e0 - base + 1] = True
    for k in range(a_s.size):
        sel = pk == k
        s, e = int(a_s[k]) - base, int(a_e[k]) - base
        got = np.zeros(span, bool)
        prev_end = None
        for x, y in zip((ps[sel] - base).tolist(), (pe[sel] - base).tolist()):
            assert s <= x <= y <= e
            assert not got[x:y + 1].any()            # disjoint
            if prev_end is not None:
                assert x > prev_end + 1               # ascending and maximal
            got[x:y + 1] = True
            prev_end = y
        want = np.zeros(span, bool)
        want[s:e + 1] = ~covered[s:e + 1]
        assert np.array_equal(got, want), k
    # order and pieces exactly as the reference's sweep
    ooff, oidx = orc.search_batch(oix, a_s, a_e, mode=0)
    hits = [oidx[ooff[k]:ooff[k + 1]] for k in range(a_s.size)]
    ref = _reference_difference(a_s, a_e, oix["starts"], oix["ends"], hits)
    assert list(zip(ps.tolist(), pe.tolist(), pk.tolist())) == ref
    ix.close()


def test_set_ops_edges(sia, orc):
    print("[gpu] set operations: empty sets and the int32 maximum", flush=True)
    big = np.iinfo(np.int32).max
    b_s = np.array([big - 2, 10], np.int32)
    b_e = np.array([big, 20], np.int32)
    ix = sia.DeviceIndex(np.int32).build(b_s, b_e)
    # an interval covered to the type's maximum leaves only its lower part
    ps, pe, pk = ix.difference_pieces(np.array([big - 5, 5, 12], np.int32), np.array([big, 30, 15], np.int32))
    assert list(zip(ps.tolist(), pe.tolist(), pk.tolist())) == [(big - 5, big - 3, 0), (5, 9, 1), (21, 30, 1)]
    ps, pe, pk, ph = ix.intersection_pieces(np.array([big - 5], np.int32), np.array([big], np.int32))
    assert (ps.tolist(), pe.tolist()) == ([big - 2], [big])
    # an empty first set
    assert ix.difference_pieces(np.zeros(0, np.int32), np.zeros(0, np.int32))[0].size == 0
    assert ix.intersection_pieces(np.zeros(0, np.int32), np.zeros(0, np.int32))[0].size == 0
    ix.close()


def test_intervalmap_set_ops(sia, orc):
    """The Python mirror: intersection keeps this set's values; difference returns the
    uncovered pieces with their source values; both unindexed, in the reference's order."""
    print("[gpu] IntervalMap intersection / difference", flush=True)
    IM = sia.IntervalMap
    a = IM()
    for s0, e0, v in [(1, 100, "a0"), (40, 60, "a1"), (200, 210, "a2"), (95, 130, "a3")]:
        a.add(s0, e0, v)
    b = IM.from_arrays([10, 50, 55, 205], [20, 70, 58, 400], ["b0", "b1", "b2", "b3"])
    inter = a.intersection(b)     # a is not built: its insertion order
    got = [inter.at(i) for i in range(inter.size())]
    ab = [(int(s0), int(e0)) for s0, e0 in zip(b._s.tolist(), b._e.tolist())]
    want = []
    for k, (s0, e0, v) in enumerate([(1, 100, "a0"), (40, 60, "a1"), (200, 210, "a2"), (95, 130, "a3")]):
        for h in b.search_idxs(s0, e0):
            want.append((max(s0, ab[h][0]), min(e0, ab[h][1]), v))
    assert got == want
    diff = a.difference(b)
    got = [diff.at(i) for i in range(diff.size())]
    assert got == [(1, 9, "a0"), (21, 49, "a0"), (71, 100, "a0"), (40, 49, "a1"), (200, 204, "a2"),
                   (95, 130, "a3")]
    empty = IM.from_arrays([], [])
    assert [a.difference(empty).at(i) for i in range(4)] == [a.at(i) for i in range(4)]
    assert a.intersection(empty).size() == 0
