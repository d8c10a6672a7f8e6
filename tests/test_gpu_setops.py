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

import ctypes as C
import os
import subprocess
import tempfile

from conftest import GOLDEN, ROOT, overlap_pieces, sweep_difference

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
    ref = sweep_difference(a_s, a_e, oix["starts"], oix["ends"], hits)
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


def _extras():
    return dict(np.load(os.path.join(GOLDEN, "extras.npz")))


def _capi_handle(N, s, e, d, index):
    L = N.lib
    h = L.createSuperIntervals()
    for i in range(s.size):
        L.addInterval(h, int(s[i]), int(e[i]), int(d[i]))
    if index:
        L.indexSuperIntervals(h)
    return h


def _capi_arrays(N, h):
    st = N.cSuperIntervals.from_address(h)
    n = st.size
    if n == 0:
        return np.zeros((3, 0), np.int64)
    return np.stack([np.ctypeslib.as_array(st.starts, (n,)), np.ctypeslib.as_array(st.ends, (n,)),
                     np.ctypeslib.as_array(st.data, (n,))]).astype(np.int64)


def test_capi_intersection_difference_golden(sia):
    """intersection / difference of the C-ABI (c_superintervals.h:286, :292) against the
    reference C implementation's results; first operand in insertion order with 10 % of its
    intervals inverted, the second indexed."""
    import superintervals_amd._native as N
    print("[gpu] C-ABI intersection / difference", flush=True)
    x = _extras()
    a_s, a_e = x["set_i32_a"]
    b_s, b_e = x["set_i32_b"]
    a = _capi_handle(N, a_s, a_e, x["set_i32_ad"], False)
    b = _capi_handle(N, b_s, b_e, x["set_i32_bd"], True)
    r = N.lib.intersection(a, b, None)
    assert np.array_equal(_capi_arrays(N, r), x["set_i32_c_intersection"])
    assert not N.cSuperIntervals.from_address(r).dev     # results are unindexed
    N.lib.destroySuperIntervals(r)
    r = N.lib.difference(a, b)
    assert np.array_equal(_capi_arrays(N, r), x["set_i32_c_difference"])
    N.lib.destroySuperIntervals(r)
    # a combine callback sees (first operand's data, second operand's data)
    CB = C.CFUNCTYPE(C.c_int32, C.c_int32, C.c_int32)
    comb = CB(lambda p, q: p * 7 + q)
    r = N.lib.intersection(a, b, C.cast(comb, C.c_void_p))
    got = _capi_arrays(N, r)
    N.lib.destroySuperIntervals(r)
    sb = N.cSuperIntervals.from_address(b)
    bd_sorted = np.ctypeslib.as_array(sb.data, (sb.size,)).astype(np.int64)
    bs_sorted = np.ctypeslib.as_array(sb.starts, (sb.size,)).astype(np.int64)
    be_sorted = np.ctypeslib.as_array(sb.ends, (sb.size,)).astype(np.int64)
    res = N.cIndexResult()
    hits = []
    for k in range(a_s.size):
        N.lib.clearIndexResult(C.byref(res))
        N.lib.searchIdxs(b, int(a_s[k]), int(a_e[k]), C.byref(res))
        hits.append([res.data[j] for j in range(res.size)])
    N.lib.destroyIndexResult(C.byref(res))
    want = [(s0, e0, int(x["set_i32_ad"][k]) * 7 + int(bd_sorted[h]))
            for s0, e0, k, h in overlap_pieces(a_s, a_e, bs_sorted, be_sorted, hits)]
    assert np.array_equal(got, np.array(want, np.int64).T)
    # an empty second operand covers nothing; an unindexed one gives no result
    e_ = N.lib.createSuperIntervals()
    r = N.lib.difference(a, e_)
    keep = a_s <= a_e
    assert np.array_equal(_capi_arrays(N, r), np.stack([a_s[keep], a_e[keep], x["set_i32_ad"][keep]]).astype(np.int64))
    N.lib.destroySuperIntervals(r)
    N.lib.addInterval(e_, 1, 2, 3)
    r = N.lib.intersection(a, e_, None)
    assert N.cSuperIntervals.from_address(r).size == 0
    assert b"indexSuperIntervals" in N.lib.siGetLastError()
    for hh in (r, e_, a, b):
        N.lib.destroySuperIntervals(hh)


def test_python_and_cpp_set_ops_golden(sia):
    """The Python IntervalMap (int32) and the C++ header (int32, int64) against the reference
    C++ intersection / difference on the same inputs."""
    print("[gpu] Python / C++ set operations against the reference", flush=True)
    x = _extras()
    a_s, a_e = x["set_i32_a"]
    b_s, b_e = x["set_i32_b"]
    IM = sia.IntervalMap
    a = IM()
    for s0, e0, d0 in zip(a_s.tolist(), a_e.tolist(), x["set_i32_ad"].tolist()):
        a.add(s0, e0, d0)
    b = IM.from_arrays(b_s, b_e, x["set_i32_bd"].tolist())
    for op in ("intersection", "difference"):
        r = getattr(a, op)(b)
        got = np.array([r.at(i) for i in range(r.size())], np.int64).T
        assert np.array_equal(got, x[f"set_i32_cpp_{op}"]), op
    exe = os.path.join(ROOT, "tests", "cpp", "test_intervalmap")
    for sfx, width in (("i32", 4), ("i64", 8)):
        a_s, a_e = x[f"set_{sfx}_a"].astype(np.int64)
        b_s, b_e = x[f"set_{sfx}_b"].astype(np.int64)
        words = np.concatenate([[width, a_s.size], a_s, a_e, x[f"set_{sfx}_ad"], [b_s.size], b_s, b_e,
                                x[f"set_{sfx}_bd"]]).astype(np.int64)
        with tempfile.TemporaryDirectory() as d:
            fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            words.tofile(fin)
            rr = subprocess.run([exe, "setops", fin, fout], capture_output=True, text=True, timeout=240)
            assert rr.returncode == 0, rr.stderr[-2000:]
            w = np.fromfile(fout, np.int64)
        ni = int(w[0])
        inter = w[1:1 + 3 * ni].reshape(3, ni)
        nd = int(w[1 + 3 * ni])
        diff = w[2 + 3 * ni:2 + 3 * ni + 3 * nd].reshape(3, nd)
        # the C++ header drops empty intersection pieces on the host
        assert np.array_equal(inter, x[f"set_{sfx}_cpp_intersection"]), sfx
        assert np.array_equal(diff, x[f"set_{sfx}_cpp_difference"]), sfx
