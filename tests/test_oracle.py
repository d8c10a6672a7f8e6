"""Pins the CPU oracle (oracle/si_oracle.c) -- CPU only.

Each oracle function is checked against something that does not restate it:
golden vectors written by the reference's own compiled headers, the reference's
known-answer tests (test/tests.cpp), brute force on small inputs and closed forms.
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, load_golden, sha


def brute_count(s, e, qs, qe):
    s = s.astype(np.int64)[None, :]
    e = e.astype(np.int64)[None, :]
    return ((s <= qe.astype(np.int64)[:, None]) & (e >= qs.astype(np.int64)[:, None])).sum(1)


def brute_pge(ends):
    out = np.full(ends.size, -1, np.int64)
    for i in range(ends.size):
        for j in range(i - 1, -1, -1):
            if ends[j] >= ends[i]:
                out[i] = j
                break
    return out


# ---- reference known answers (test/tests.cpp) ---------------------------------------
def _ix(orc, pairs, data=None):
    s = np.array([p[0] for p in pairs], np.int32)
    e = np.array([p[1] for p in pairs], np.int32)
    return orc.build_index(s, e, data)


def test_known_basics(orc):
    # tests.cpp:63-107
    ix = _ix(orc, [(10, 20), (11, 12), (13, 14), (15, 16), (25, 29)], [0, -1, -1, -1, 4])
    assert orc.has_overlaps(ix, 17, 30) and not orc.has_overlaps(ix, 1, 3)
    assert list(orc.values(ix, 17, 30)) == [4, 0]
    assert orc.count(ix, 17, 30) == 2
    assert list(orc.search(ix, 17, 30, 1)) == [4, 0] and list(orc.search(ix, 17, 30, 0)) == [4, 0]
    keys = [(int(ix["starts"][i]), int(ix["ends"][i])) for i in orc.search(ix, 17, 30, 0)]
    assert keys[0][0] == 25 and keys[1][1] == 20
    assert orc.coverage(ix, 10, 29)[1] == 17


def test_known_iteration(orc):
    # tests.cpp:113-133
    ix = _ix(orc, [(1, 10)], [0])
    assert list(orc.search(ix, 5, 11, 0)) == [0] and list(orc.values(ix, 5, 11)) == [0]


def test_known_overlaps(orc):
    # tests.cpp:146-156: container + 17 intervals
    pairs = [(0, 250000)] + [(s, s + 1000) for s in (55, 115, 130, 281, 639, 842, 999, 1094, 1157, 1161,
                                                      1265, 1532, 1590, 1665, 1945, 2384, 2515)]
    ix = _ix(orc, pairs, [0] + [-1] * 17)
    a = orc.values(ix, 1377, 2377)
    assert a[-1] == 0 and a.size == 12
    # :161-171 shared endpoints
    ix = _ix(orc, [(3, 40), (4, 5), (6, 7), (10, 31), (31, 32)], [0, 4, 4, 5, 5])
    assert orc.values(ix, 31, 32).size == 3
    assert orc.values(ix, 10, 11).size == 2
    assert orc.values(ix, 4, 7).size == 3
    # :175-186 duplicates
    ix = _ix(orc, [(3, 40), (3, 40), (3, 40), (3, 4), (35, 50), (40, 400), (40, 400)], [0, 4, 4, 4, 4, 5, 4])
    assert orc.values(ix, 38, 41).size == 6
    assert orc.values(ix, 41, 42).size == 3


def test_known_coverage_and_edges(orc):
    # tests.cpp:194-213
    ix = _ix(orc, [(1, 100), (30, 200), (40, 50), (60, 70)], [0, 7, 6, 7])
    assert orc.values(ix, 55, 65).size == 3
    assert orc.coverage(ix, 55, 65) == (3, 25)
    # :219-238
    empty = orc.build_index(np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert orc.count(empty, 1, 5) == 0 and not orc.has_overlaps(empty, 1, 5)
    one = _ix(orc, [(1, 10)], [1])
    assert orc.count(one, 1, 5) == 1 and orc.count(one, 11, 20) == 0


def test_vector_order_example(orc):
    # The C++ vector form inserts the leading run ascending (superintervals.hpp:891-895).
    ix = _ix(orc, [(1, 100), (2, 3), (4, 50), (5, 50), (6, 50), (60, 70)])
    assert list(ix["branch"]) == [-1, 0, 0, 2, 3, 0]
    assert list(orc.search(ix, 10, 65, 1)) == [2, 3, 4, 5, 0]
    assert list(orc.search(ix, 10, 65, 0)) == [5, 4, 3, 2, 0]
    assert orc.count(ix, 10, 65) == 5


def test_has_overlaps_checks_last_start_only(orc):
    # has_overlaps looks only at the last interval starting at or before end
    # (superintervals.hpp:865-871): a covering earlier interval is not seen.
    ix = _ix(orc, [(1, 100), (50, 60)])
    assert orc.count(ix, 70, 80) == 1
    assert not orc.has_overlaps(ix, 70, 80)


# ---- brute force and closed forms -----------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_small_random_brute_force(orc, seed, dtype):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    span = int(rng.choice([20, 200, 5000]))
    s = rng.integers(-span, span, n).astype(dtype)
    e = (s + rng.integers(0, span // 2 + 2, n)).astype(dtype)
    if dtype == np.int64:
        s += dtype(1) << 40
        e += dtype(1) << 40
    qs = rng.integers(-span - 10, span + 10, 400).astype(dtype)
    qe = (qs + rng.integers(-3, span // 3 + 2, 400)).astype(dtype)   # includes inverted queries
    if dtype == np.int64:
        qs += dtype(1) << 40
        qe += dtype(1) << 40
    ix = orc.build_index(s, e)
    # sorted order: any sort by (start asc, end desc) gives the same pair sequence
    order = np.lexsort((-e.astype(np.int64), s.astype(np.int64)))
    assert np.array_equal(ix["starts"], s[order]) and np.array_equal(ix["ends"], e[order])
    assert np.array_equal(ix["data"], order.astype(np.int32))          # stable
    assert np.array_equal(ix["branch"], brute_pge(ix["ends"]))
    bc = brute_count(ix["starts"], ix["ends"], qs, qe)
    assert np.array_equal(orc.count_batch(ix, qs, qe, threads=2), bc)
    ss, ee = ix["starts"].astype(np.int64), ix["ends"].astype(np.int64)
    for q in range(qs.size):
        ub = np.searchsorted(ss, int(qe[q]), side="right") - 1
        assert orc.upper_bound(ix, qe[q]) == ub
        strict = [j for j in range(ub, -1, -1) if ee[j] >= qs[q]]
        assert list(orc.search(ix, qs[q], qe[q], 0)) == strict
        # vector form: [i*+1 .. ub] ascending then the rest of the strict list
        i = ub
        while i >= 0 and ee[i] >= qs[q]:
            i -= 1
        vec = list(range(i + 1, ub + 1)) + [j for j in strict if j <= i]
        assert list(orc.search(ix, qs[q], qe[q], 1)) == vec
        assert list(orc.values(ix, qs[q], qe[q])) == [int(ix["data"][j]) for j in strict]
        assert orc.has_overlaps(ix, qs[q], qe[q]) == (ub >= 0 and ee[ub] >= qs[q])
        if qs[q] <= qe[q]:   # rank difference closed form
            assert bc[q] == (ub + 1) - np.searchsorted(np.sort(ee), int(qs[q]), side="left")
        if dtype == np.int32:
            cov = sum(min(int(ee[j]), int(qe[q])) - max(int(ss[j]), int(qs[q])) for j in strict)
            assert orc.coverage(ix, qs[q], qe[q]) == (len(strict), cov)
        h, v, j = orc.walk_stats(ix, qs[q], qe[q])
        assert h == len(strict)


def test_walk_stats_rewalk(orc):
    """Visit/jump counts against an independent Python walk of the same index."""
    rng = np.random.default_rng(11)
    s = np.sort(rng.integers(0, 2000, 500)).astype(np.int32)
    e = (s + rng.geometric(1 / 40, 500)).astype(np.int32)
    ix = orc.build_index(s, e)
    br, en = ix["branch"], ix["ends"]
    for q in range(200):
        a = int(rng.integers(0, 2000)); b = a + int(rng.integers(0, 60))
        i = orc.upper_bound(ix, b)
        V = J = H = 0
        while i != -1:
            V += 1
            if a <= en[i]:
                H += 1; i -= 1
            else:
                J += 1; i = int(br[i])
        assert orc.walk_stats(ix, a, b) == (H, V, J)


def test_block_thresholds(orc):
    """count's 16-wide block scan: runs of hits around the block length (hpp:688-700)."""
    for n in (15, 16, 17, 31, 32, 33, 64, 65, 200):
        s = np.arange(n, dtype=np.int32)
        e = np.full(n, 1000, np.int32)
        e[::7] = 3    # short intervals break the run
        ix = orc.build_index(s, e)
        for a in (0, 2, 4, 500):
            for b in (0, n // 2, n - 1, n + 5):
                assert orc.count(ix, a, b) == brute_count(ix["starts"], ix["ends"], np.array([a]), np.array([b]))[0]


def test_extremes_int32(orc):
    big, small = np.iinfo(np.int32).max, np.iinfo(np.int32).min
    s = np.array([small, small, 0, big - 5, big], np.int32)
    e = np.array([small, big, 10, big, big], np.int32)
    ix = orc.build_index(s, e)
    qs = np.array([small, small, 5, big, big, 0, big], np.int32)
    qe = np.array([small, big, 5, big, big - 1, small, small], np.int32)
    bc = brute_count(ix["starts"], ix["ends"], qs, qe)
    for q in range(qs.size):
        assert orc.count(ix, qs[q], qe[q]) == bc[q]


# ---- golden vectors from the reference ---------------------------------------------------
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden(orc, name):
    rec, s, e, qs, qe = load_golden(name)
    ix = orc.build_index(s, e)
    assert sha(ix["starts"], ix["ends"]) == str(rec["sorted_sha256"])
    assert sha(ix["branch"]) == str(rec["branch_sha256"])
    if "c_data_sha256" in rec:
        assert sha(ix["data"]) == str(rec["c_data_sha256"])
    assert np.array_equal(orc.count_batch(ix, qs, qe), rec["counts"])
    c, hd = orc.search_batch(ix, qs, qe, mode=0, lists=False)
    assert np.array_equal(c, rec["counts"])
    assert orc.digest(hd) == int(rec["hash_desc_digest"])
    _, hv = orc.search_batch(ix, qs, qe, mode=1, lists=False)
    assert orc.digest(hv) == int(rec["hash_vec_digest"])
    k = rec["upper_bound"].size
    assert [orc.upper_bound(ix, x) for x in qe[:k]] == rec["upper_bound"].tolist()
    ho = np.array([orc.has_overlaps(ix, a, b) for a, b in zip(qs, qe)], bool)
    assert np.array_equal(np.packbits(ho), rec["has_overlaps"])
    off = rec["lists_offsets"]
    for q in range(off.size - 1):
        assert np.array_equal(orc.search(ix, qs[q], qe[q], 0), rec["lists_idx"][off[q]:off[q + 1]])
    if "hash_cvalues_digest" in rec:
        hv = [orc.hash_list(orc.values(ix, a, b).astype(np.int64)) for a, b in zip(qs, qe)]
        assert orc.digest(np.array(hv, np.uint64)) == int(rec["hash_cvalues_digest"])
        cov = [orc.coverage(ix, qs[q], qe[q]) for q in range(rec["coverage_prefix"].shape[0])]
        assert np.array_equal(np.array(cov, np.int64), rec["coverage_prefix"])


def test_planted_mistake_is_caught(orc):
    """A wrong index in an oracle-like result must fail the golden digest."""
    rec, s, e, qs, qe = load_golden("c1_sorted")
    ix = orc.build_index(s, e)
    off, idx = orc.search_batch(ix, qs[:2048], qe[:2048], mode=0)
    h = orc.hash_csr(off, idx)
    assert np.array_equal(h, rec["hash_desc_prefix"])
    bad = idx.copy()
    bad[len(bad) // 2] += 1
    assert not np.array_equal(orc.hash_csr(off, bad), rec["hash_desc_prefix"])
    swapped = idx.copy()
    q = int(np.argmax(np.diff(off) >= 2))
    a0 = int(off[q])
    swapped[a0], swapped[a0 + 1] = swapped[a0 + 1], swapped[a0]
    assert not np.array_equal(orc.hash_csr(off, swapped), rec["hash_desc_prefix"])


def test_reference_build_cross_check(orc):
    """Where the reference build is present (development container), compare the
    restatement with it on a fresh random case."""
    import oracle
    if not oracle.Reference.available():
        pytest.skip("oracle/_ref not built here")
    R = oracle.Reference()
    rng = np.random.default_rng(5)
    s = rng.integers(0, 50_000, 20_000).astype(np.int32)
    e = (s + rng.geometric(1 / 200, 20_000)).astype(np.int32)
    qs = rng.integers(0, 50_000, 5_000).astype(np.int32)
    qe = (qs + rng.integers(0, 400, 5_000)).astype(np.int32)
    m = oracle.RefMap(R, s, e)
    c = oracle.RefCMap(R, s, e)
    ix = orc.build_index(s, e)
    ra, ca = m.arrays(), c.arrays()
    assert np.array_equal(ra["starts"], ix["starts"]) and np.array_equal(ra["branch"], ix["branch"])
    assert np.array_equal(ca["data"], ix["data"])
    for q in range(0, 5_000, 7):
        assert m.count(qs[q], qe[q]) == orc.count(ix, qs[q], qe[q])
        assert np.array_equal(m.search(qs[q], qe[q], 1), orc.search(ix, qs[q], qe[q], 1))
        assert np.array_equal(c.search(qs[q], qe[q], 0), orc.search(ix, qs[q], qe[q], 0))


# ---- learned-index model oracle (oracle/rmi.py), pinned by finite differences ----------
def test_rmi_oracle_gradient_finite_differences():
    from oracle import rmi
    rng = np.random.default_rng(4)
    x = rng.random(300)
    t = np.sort(rng.random(300))
    for p in (rmi.identity_params(), rmi.identity_params() + 0.05 * rng.standard_normal(rmi.NPARAM)):
        g, l0 = rmi.gradient(p, x, t)
        assert abs(l0 - rmi.loss(p, x, t)) < 1e-14
        h = 1e-6
        for k in rng.choice(rmi.NPARAM, 60, replace=False).tolist() + [0, 16, 32, 288, 304, 320]:
            e = np.zeros(rmi.NPARAM)
            e[k] = h
            fd = (rmi.loss(p + e, x, t) - rmi.loss(p - e, x, t)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_rmi_identity_init_is_identity_cdf():
    from oracle import rmi
    x = np.linspace(0, 1, 101)
    assert np.allclose(rmi.forward(rmi.identity_params(), x), x)
    assert rmi.leaf_of(np.array([-0.5, 0.0, 0.5, 0.99999, 1.5, np.nan]), 4096).tolist() == [0, 0, 2048, 4095, 4095, 0]
