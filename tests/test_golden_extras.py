"""Reference outputs in tests/golden/extras.npz (written by the compiled reference, see
tests/golden/make_golden.py) checked against the code that needs no GPU:

* the oracle on an index in which 15 % of the intervals have start > end (the C++ counts,
  both C++ result orders, the C searchIdxs / searchValues lists);
* upper_bound_range of include/superintervals.hpp (host code) against the reference's
  exit positions (superintervals.hpp:524-543);
* the set-operation checkers the GPU tests use (tests/conftest.py: the difference sweep
  and the intersection pieces over the oracle's search) against the reference's C and C++
  intersection / difference results.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, overlap_pieces, sweep_difference


@pytest.fixture(scope="module")
def ext():
    return dict(np.load(os.path.join(GOLDEN, "extras.npz")))


def _lists(off, flat):
    off = off.astype(np.int64)
    return [flat[off[q]:off[q + 1]] for q in range(off.size - 1)]


def test_oracle_on_index_with_inverted_intervals(orc, ext):
    s, e, qs, qe = ext["inv_starts"], ext["inv_ends"], ext["inv_qs"], ext["inv_qe"]
    assert (s > e).mean() > 0.1 and (qs > qe).any()
    oix = orc.build_index(s, e, np.arange(s.size, dtype=np.int32))
    assert np.array_equal(orc.count_batch(oix, qs, qe).astype(np.int64), ext["inv_counts"])
    off, idx = orc.search_batch(oix, qs, qe, mode=0)
    assert np.array_equal(off, ext["inv_desc_off"]) and np.array_equal(idx, ext["inv_desc_idx"])
    assert np.array_equal(off, ext["inv_cidx_off"]) and np.array_equal(idx, ext["inv_cidx"])
    assert np.array_equal(oix["data"][idx], ext["inv_cval"])
    off, idx = orc.search_batch(oix, qs, qe, mode=1)
    assert np.array_equal(off, ext["inv_vec_off"]) and np.array_equal(idx, ext["inv_vec_idx"])
    # the rank form (upper_bound + 1 - #{ends < s}) is not the count on such an index
    st, en = oix["starts"], oix["ends"]
    ub = np.searchsorted(st, qe, side="right")
    rank = ub - np.searchsorted(np.sort(en), qs, side="left")
    ok = qs <= qe
    assert (rank[ok] != ext["inv_counts"][ok]).any()


def _run_cpp(mode, words):
    exe = os.path.join(ROOT, "tests", "cpp", "test_intervalmap")
    assert os.path.exists(exe), "build the C++ test first (__graft_entry__.build())"
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.asarray(words, np.int64).tofile(fin)
        r = subprocess.run([exe, mode, fin, fout], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        return np.fromfile(fout, np.int64)


@pytest.mark.parametrize("sfx,width", [("i32", 4), ("i64", 8)])
def test_upper_bound_range_cpp_header(ext, sfx, width):
    starts = ext[f"ubr_{sfx}_starts"].astype(np.int64)
    v, lo, hi = (ext[f"ubr_{sfx}_{k}"].astype(np.int64) for k in ("value", "left", "right"))
    words = np.concatenate([[width, starts.size], starts, [v.size], v, lo, hi])
    got = _run_cpp("ubr", words)
    assert np.array_equal(got, ext[f"ubr_{sfx}_out"])
    # every position above the exit (up to the range) starts at or after value: the exit is
    # the position just below the first such start, or one inside the run that starts
    # exactly at value (-1 when position 0 qualifies)
    lb = np.searchsorted(starts, v, side="left")
    g = np.maximum(got, 0)
    inside = (got >= lb) & (starts[g] >= v) & ((g == 0) | (starts[np.maximum(g - 1, 0)] <= v))
    ok = (got == lb - 1) | inside | ((got == -1) & (lb == 0)) | (lb > hi)
    assert ok.all()


@pytest.mark.parametrize("sfx", ["i32", "i64"])
def test_set_op_checkers_against_reference(orc, ext, sfx):
    a_s, a_e = ext[f"set_{sfx}_a"]
    b_s, b_e = ext[f"set_{sfx}_b"]
    a_d, b_d = ext[f"set_{sfx}_ad"], ext[f"set_{sfx}_bd"]
    oix = orc.build_index(b_s, b_e, b_d)
    ob_s, ob_e, ob_d = oix["starts"], oix["ends"], oix["data"]
    desc = _lists(*orc.search_batch(oix, a_s, a_e, mode=0))
    vec = _lists(*orc.search_batch(oix, a_s, a_e, mode=1))
    # C++: intersection over search_idxs (vector order), difference over search_keys
    want = np.array([(s, e, a_d[k]) for s, e, k, _ in overlap_pieces(a_s, a_e, ob_s, ob_e, vec)], np.int64).T
    assert np.array_equal(want, ext[f"set_{sfx}_cpp_intersection"])
    diff = np.array([(s, e, a_d[k]) for s, e, k in sweep_difference(a_s, a_e, ob_s, ob_e, desc)], np.int64).T
    assert np.array_equal(diff, ext[f"set_{sfx}_cpp_difference"])
    if sfx == "i32":   # C: intersection over searchIdxs (descending)
        want = np.array([(s, e, a_d[k]) for s, e, k, _ in overlap_pieces(a_s, a_e, ob_s, ob_e, desc)], np.int64).T
        assert np.array_equal(want, ext["set_i32_c_intersection"])
        assert np.array_equal(diff, ext["set_i32_c_difference"])
    assert (a_s > a_e).any() and (b_s > b_e).any()
