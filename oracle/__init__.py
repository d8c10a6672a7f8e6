"""CPU oracle for the superset-index hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings of
  oracle/liboracle.so     the plain-C restatement in si_oracle.c, and
  oracle/_ref/libsiref.so the reference's own headers compiled in place
                          (present only where the reference checkout was available
                          when oracle/Makefile ran).
Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only,
always as the checker or the baseline, never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsiref.so")

_vp, _sz, _i = C.c_void_p, C.c_size_t, C.c_int


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference checkout is present)."""
    out = subprocess.run(["make", "-f", os.path.join(HERE, "Makefile"), "all"],
                         capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """The restatement (si_oracle.c).  Coordinates int32 or int64."""

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            build()
        self.lib = C.CDLL(ORACLE_SO)
        L = self.lib
        for sfx, ct in (("i32", C.c_int32), ("i64", C.c_int64)):
            getattr(L, f"or_flags_{sfx}").restype = _i
            getattr(L, f"or_flags_{sfx}").argtypes = [_sz, _vp, _vp]
            getattr(L, f"or_sort_{sfx}").restype = None
            getattr(L, f"or_sort_{sfx}").argtypes = [_sz, _vp, _vp, _vp, _i, _i]
            getattr(L, f"or_branch_{sfx}").restype = None
            getattr(L, f"or_branch_{sfx}").argtypes = [_sz, _vp, _vp]
            getattr(L, f"or_upper_bound_{sfx}").restype = C.c_int64
            getattr(L, f"or_upper_bound_{sfx}").argtypes = [_sz, _vp, ct]
            getattr(L, f"or_count_{sfx}").restype = C.c_uint64
            getattr(L, f"or_count_{sfx}").argtypes = [_sz, _vp, _vp, _vp, ct, ct]
            getattr(L, f"or_search_{sfx}").restype = _sz
            getattr(L, f"or_search_{sfx}").argtypes = [_sz, _vp, _vp, _vp, ct, ct, _i, _vp]
            getattr(L, f"or_values_{sfx}").restype = _sz
            getattr(L, f"or_values_{sfx}").argtypes = [_sz, _vp, _vp, _vp, _vp, ct, ct, _vp]
            getattr(L, f"or_has_overlaps_{sfx}").restype = _i
            getattr(L, f"or_has_overlaps_{sfx}").argtypes = [_sz, _vp, _vp, ct, ct]
            getattr(L, f"or_walk_stats_{sfx}").restype = None
            getattr(L, f"or_walk_stats_{sfx}").argtypes = [_sz, _vp, _vp, _vp, ct, ct, _vp, _vp, _vp]
            getattr(L, f"or_count_batch_{sfx}").restype = C.c_uint64
            getattr(L, f"or_count_batch_{sfx}").argtypes = [_i, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp]
            getattr(L, f"or_search_batch_{sfx}").restype = C.c_uint64
            getattr(L, f"or_search_batch_{sfx}").argtypes = [_i, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _i,
                                                             _vp, _vp, _vp, _vp]
            getattr(L, f"or_walk_stats_batch_{sfx}").restype = None
            getattr(L, f"or_walk_stats_batch_{sfx}").argtypes = [_i, _sz, _vp, _vp, _vp, _sz, _vp, _vp, _vp]
        L.or_coverage_i32.restype = None
        L.or_coverage_i32.argtypes = [_sz, _vp, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp]
        L.or_hash_list.restype = C.c_uint64
        L.or_hash_list.argtypes = [_vp, _sz]
        L.or_hash_list_u32.restype = C.c_uint64
        L.or_hash_list_u32.argtypes = [_vp, _sz]
        for nm in ("or_hash_csr_u32", "or_hash_csr_i64", "or_hash_csr_i32"):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [_sz, _vp, _vp, _vp]

    @staticmethod
    def _sfx(a):
        return "i64" if np.asarray(a).dtype == np.int64 else "i32"

    def build_index(self, starts, ends, data=None, flags=None):
        """Sort (with the add() flags) and build branch.  Returns dict of arrays."""
        s = np.array(starts, copy=True)
        e = np.array(ends, dtype=s.dtype, copy=True)
        d = (np.arange(s.size, dtype=np.int32) if data is None
             else np.array(data, dtype=np.int32, copy=True))
        sfx = self._sfx(s)
        if flags is None:
            f = getattr(self.lib, f"or_flags_{sfx}")(s.size, _p(s), _p(e))
            flags = (bool(f & 1), bool(f & 2))
        getattr(self.lib, f"or_sort_{sfx}")(s.size, _p(s), _p(e), _p(d), int(flags[0]), int(flags[1]))
        b = np.empty(s.size, np.int64)
        getattr(self.lib, f"or_branch_{sfx}")(s.size, _p(e), _p(b))
        return {"starts": s, "ends": e, "data": d, "branch": b}

    def upper_bound(self, ix, x):
        return int(getattr(self.lib, f"or_upper_bound_{self._sfx(ix['starts'])}")(
            ix["starts"].size, _p(ix["starts"]), int(x)))

    def count(self, ix, s, e):
        return int(getattr(self.lib, f"or_count_{self._sfx(ix['starts'])}")(
            ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), int(s), int(e)))

    def search(self, ix, s, e, mode=0):
        sfx = self._sfx(ix["starts"])
        f = getattr(self.lib, f"or_search_{sfx}")
        n = ix["starts"].size
        h = f(n, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), int(s), int(e), mode, None)
        out = np.empty(h, np.int64)
        f(n, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), int(s), int(e), mode, _p(out))
        return out

    def values(self, ix, s, e):
        sfx = self._sfx(ix["starts"])
        f = getattr(self.lib, f"or_values_{sfx}")
        n = ix["starts"].size
        args = (n, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), _p(ix["data"]), int(s), int(e))
        h = f(*args, None)
        out = np.empty(h, np.int32)
        f(*args, _p(out))
        return out

    def has_overlaps(self, ix, s, e):
        return bool(getattr(self.lib, f"or_has_overlaps_{self._sfx(ix['starts'])}")(
            ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), int(s), int(e)))

    def coverage(self, ix, s, e):
        c = np.zeros(1, np.uint64)
        v = np.zeros(1, np.int32)
        self.lib.or_coverage_i32(ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]),
                                 int(s), int(e), _p(c), _p(v))
        return int(c[0]), int(v[0])

    def walk_stats(self, ix, s, e):
        h = np.zeros(1, np.uint64); v = np.zeros(1, np.uint64); j = np.zeros(1, np.uint64)
        getattr(self.lib, f"or_walk_stats_{self._sfx(ix['starts'])}")(
            ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), int(s), int(e),
            _p(h), _p(v), _p(j))
        return int(h[0]), int(v[0]), int(j[0])

    def count_batch(self, ix, qs, qe, threads=8):
        qs = np.ascontiguousarray(qs, ix["starts"].dtype)
        qe = np.ascontiguousarray(qe, ix["starts"].dtype)
        out = np.empty(qs.size, np.uint32)
        getattr(self.lib, f"or_count_batch_{self._sfx(qs)}")(
            threads, ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]),
            qs.size, _p(qs), _p(qe), _p(out))
        return out

    def search_batch(self, ix, qs, qe, mode=0, threads=8, lists=True):
        """(offsets uint64[nq+1], idx int64[...]) if lists else (counts, hashes)."""
        qs = np.ascontiguousarray(qs, ix["starts"].dtype)
        qe = np.ascontiguousarray(qe, ix["starts"].dtype)
        sfx = self._sfx(qs)
        f = getattr(self.lib, f"or_search_batch_{sfx}")
        n = ix["starts"].size
        counts = np.empty(qs.size, np.uint32)
        hashes = np.empty(qs.size, np.uint64)
        f(threads, n, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), qs.size, _p(qs), _p(qe), mode,
          _p(counts), None, None, _p(hashes))
        if not lists:
            return counts, hashes
        off = np.zeros(qs.size + 1, np.uint64)
        np.cumsum(counts, out=off[1:])
        out = np.empty(int(off[-1]), np.int64)
        f(threads, n, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]), qs.size, _p(qs), _p(qe), mode,
          None, _p(off), _p(out), None)
        return off, out

    def walk_stats_batch(self, ix, qs, qe, threads=8):
        qs = np.ascontiguousarray(qs, ix["starts"].dtype)
        qe = np.ascontiguousarray(qe, ix["starts"].dtype)
        st = np.zeros(3, np.uint64)
        getattr(self.lib, f"or_walk_stats_batch_{self._sfx(qs)}")(
            threads, ix["starts"].size, _p(ix["starts"]), _p(ix["ends"]), _p(ix["branch"]),
            qs.size, _p(qs), _p(qe), _p(st))
        return [int(x) for x in st]

    def hash_csr(self, offsets, idx) -> np.ndarray:
        """Per-query order-sensitive hashes of a CSR result.  Index lists (uint32 or
        int64) hash by value; int32 payload lists hash their sign-extended values."""
        off = np.ascontiguousarray(offsets, np.uint64)
        nq = off.size - 1
        out = np.empty(nq, np.uint64)
        idx = np.asarray(idx)
        if idx.dtype == np.uint32:
            self.lib.or_hash_csr_u32(nq, _p(off), _p(np.ascontiguousarray(idx)), _p(out))
        elif idx.dtype == np.int32:
            self.lib.or_hash_csr_i32(nq, _p(off), _p(np.ascontiguousarray(idx)), _p(out))
        else:
            self.lib.or_hash_csr_i64(nq, _p(off), _p(np.ascontiguousarray(idx, np.int64)), _p(out))
        return out

    def digest(self, hashes) -> int:
        """One 64-bit digest of a per-query hash array."""
        return self.hash_list(np.asarray(hashes, np.uint64).view(np.int64))

    def hash_list(self, v) -> int:
        v = np.ascontiguousarray(v, np.int64)
        return int(self.lib.or_hash_list(_p(v), v.size))

    def hash_list_u32(self, v) -> int:
        v = np.ascontiguousarray(v, np.uint32)
        return int(self.lib.or_hash_list_u32(_p(v), v.size))


class Reference:
    """The reference's own headers, compiled in place (oracle/_ref/libsiref.so)."""

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def __init__(self):
        self.lib = C.CDLL(REF_SO)
        L = self.lib
        for sfx, ct in (("i32", C.c_int32), ("i64", C.c_int64)):
            getattr(L, f"ref_build_{sfx}").restype = _vp
            getattr(L, f"ref_build_{sfx}").argtypes = [_sz, _vp, _vp, _vp]
            getattr(L, f"ref_free_{sfx}").argtypes = [_vp]
            getattr(L, f"ref_size_{sfx}").restype = _sz
            getattr(L, f"ref_size_{sfx}").argtypes = [_vp]
            getattr(L, f"ref_arrays_{sfx}").argtypes = [_vp, _vp, _vp, _vp, _vp]
            getattr(L, f"ref_upper_bound_{sfx}").restype = C.c_int64
            getattr(L, f"ref_upper_bound_{sfx}").argtypes = [_vp, ct]
            getattr(L, f"ref_count_{sfx}").restype = C.c_uint64
            getattr(L, f"ref_count_{sfx}").argtypes = [_vp, ct, ct]
            getattr(L, f"ref_has_overlaps_{sfx}").restype = _i
            getattr(L, f"ref_has_overlaps_{sfx}").argtypes = [_vp, ct, ct]
            getattr(L, f"ref_search_{sfx}").restype = _sz
            getattr(L, f"ref_search_{sfx}").argtypes = [_vp, ct, ct, _i, _vp, _sz]
            getattr(L, f"ref_values_{sfx}").restype = _sz
            getattr(L, f"ref_values_{sfx}").argtypes = [_vp, ct, ct, _vp, _sz]
            getattr(L, f"ref_count_batch_{sfx}").restype = C.c_uint64
            getattr(L, f"ref_count_batch_{sfx}").argtypes = [_vp, _i, _sz, _vp, _vp, _vp]
            getattr(L, f"ref_search_batch_{sfx}").restype = C.c_uint64
            getattr(L, f"ref_search_batch_{sfx}").argtypes = [_vp, _i, _sz, _vp, _vp, _vp]
        L.ref_coverage_i32.argtypes = [_vp, C.c_int32, C.c_int32, _vp, _vp]
        for sfx, ct in (("i32", C.c_int32), ("i64", C.c_int64)):
            getattr(L, f"ref_upper_bound_range_{sfx}").restype = C.c_int64
            getattr(L, f"ref_upper_bound_range_{sfx}").argtypes = [_vp, ct, _sz, _sz]
            getattr(L, f"ref_raw_{sfx}").restype = _vp
            getattr(L, f"ref_raw_{sfx}").argtypes = [_sz, _vp, _vp, _vp]
            getattr(L, f"ref_raw_arrays_{sfx}").argtypes = [_vp, _vp, _vp, _vp]
            getattr(L, f"ref_intersection_{sfx}").restype = _vp
            getattr(L, f"ref_intersection_{sfx}").argtypes = [_vp, _vp]
            getattr(L, f"ref_difference_{sfx}").restype = _vp
            getattr(L, f"ref_difference_{sfx}").argtypes = [_vp, _vp]
        L.refc_raw.restype = _vp
        L.refc_raw.argtypes = [_sz, _vp, _vp, _vp]
        L.refc_size.restype = _sz
        L.refc_size.argtypes = [_vp]
        L.refc_raw_arrays.argtypes = [_vp, _vp, _vp, _vp]
        L.refc_intersection.restype = _vp
        L.refc_intersection.argtypes = [_vp, _vp]
        L.refc_difference.restype = _vp
        L.refc_difference.argtypes = [_vp, _vp]
        L.refc_build.restype = _vp
        L.refc_build.argtypes = [_sz, _vp, _vp, _vp]
        L.refc_free.argtypes = [_vp]
        L.refc_arrays.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.refc_upper_bound.restype = C.c_int64
        L.refc_upper_bound.argtypes = [_vp, C.c_int32]
        L.refc_count.restype = C.c_uint64
        L.refc_count.argtypes = [_vp, C.c_int32, C.c_int32]
        L.refc_any.restype = _i
        L.refc_any.argtypes = [_vp, C.c_int32, C.c_int32]
        L.refc_search.restype = _sz
        L.refc_search.argtypes = [_vp, C.c_int32, C.c_int32, _i, _vp, _sz]
        L.refc_coverage.argtypes = [_vp, C.c_int32, C.c_int32, _vp, _vp]


class RefMap:
    """A built reference si::IntervalMap<S, int32_t> (C++ template)."""

    def __init__(self, ref: Reference, starts, ends, data=None):
        self.L = ref.lib
        s = np.ascontiguousarray(starts)
        e = np.ascontiguousarray(ends, dtype=s.dtype)
        d = None if data is None else np.ascontiguousarray(data, np.int32)
        self.sfx = "i64" if s.dtype == np.int64 else "i32"
        self.dtype = s.dtype
        self.h = getattr(self.L, f"ref_build_{self.sfx}")(s.size, _p(s), _p(e), _p(d))
        self.n = s.size

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.L, f"ref_free_{self.sfx}")(self.h)
            self.h = None

    def arrays(self):
        s = np.empty(self.n, self.dtype); e = np.empty(self.n, self.dtype)
        d = np.empty(self.n, np.int32); b = np.empty(self.n, np.int64)
        getattr(self.L, f"ref_arrays_{self.sfx}")(self.h, _p(s), _p(e), _p(d), _p(b))
        return {"starts": s, "ends": e, "data": d, "branch": b}

    def upper_bound(self, x):
        return int(getattr(self.L, f"ref_upper_bound_{self.sfx}")(self.h, int(x)))

    def count(self, s, e):
        return int(getattr(self.L, f"ref_count_{self.sfx}")(self.h, int(s), int(e)))

    def has_overlaps(self, s, e):
        return bool(getattr(self.L, f"ref_has_overlaps_{self.sfx}")(self.h, int(s), int(e)))

    def search(self, s, e, mode):
        f = getattr(self.L, f"ref_search_{self.sfx}")
        h = f(self.h, int(s), int(e), mode, None, 0)
        out = np.empty(h, np.int64)
        f(self.h, int(s), int(e), mode, _p(out), h)
        return out

    def values(self, s, e):
        f = getattr(self.L, f"ref_values_{self.sfx}")
        h = f(self.h, int(s), int(e), None, 0)
        out = np.empty(h, np.int32)
        f(self.h, int(s), int(e), _p(out), h)
        return out

    def coverage(self, s, e):
        c = np.zeros(1, np.uint64); v = np.zeros(1, np.int32)
        self.L.ref_coverage_i32(self.h, int(s), int(e), _p(c), _p(v))
        return int(c[0]), int(v[0])

    def count_batch(self, qs, qe, threads):
        qs = np.ascontiguousarray(qs, self.dtype); qe = np.ascontiguousarray(qe, self.dtype)
        out = np.empty(qs.size, np.uint32)
        tot = getattr(self.L, f"ref_count_batch_{self.sfx}")(self.h, threads, qs.size, _p(qs), _p(qe), _p(out))
        return out, int(tot)

    def upper_bound_range(self, value, left, right):
        """Exit value of `left` (-1 for SIZE_MAX), superintervals.hpp:524-543."""
        return int(getattr(self.L, f"ref_upper_bound_range_{self.sfx}")(self.h, int(value), int(left), int(right)))

    def set_op(self, other, op):
        """C++ intersection (default combine) or difference of this map with other:
        (starts, ends, data) of the unindexed result in its order."""
        h = getattr(self.L, f"ref_{op}_{self.sfx}")(self.h, other.h)
        n = int(getattr(self.L, f"ref_size_{self.sfx}")(h))
        s = np.empty(n, self.dtype); e = np.empty(n, self.dtype); d = np.empty(n, np.int32)
        getattr(self.L, f"ref_raw_arrays_{self.sfx}")(h, _p(s), _p(e), _p(d))
        getattr(self.L, f"ref_free_{self.sfx}")(h)
        return s, e, d

    @classmethod
    def raw(cls, ref, starts, ends, data=None):
        """An unbuilt map holding the intervals in insertion order."""
        m = cls.__new__(cls)
        m.L = ref.lib
        s = np.ascontiguousarray(starts)
        e = np.ascontiguousarray(ends, dtype=s.dtype)
        d = None if data is None else np.ascontiguousarray(data, np.int32)
        m.sfx = "i64" if s.dtype == np.int64 else "i32"
        m.dtype = s.dtype
        m.h = getattr(m.L, f"ref_raw_{m.sfx}")(s.size, _p(s), _p(e), _p(d))
        m.n = s.size
        return m

    def search_batch_total(self, qs, qe, threads):
        qs = np.ascontiguousarray(qs, self.dtype); qe = np.ascontiguousarray(qe, self.dtype)
        return int(getattr(self.L, f"ref_search_batch_{self.sfx}")(self.h, threads, qs.size, _p(qs), _p(qe), None))


class RefCMap:
    """A built reference cSuperIntervals (the C implementation)."""

    def __init__(self, ref: Reference, starts, ends, data=None):
        self.L = ref.lib
        s = np.ascontiguousarray(starts, np.int32)
        e = np.ascontiguousarray(ends, np.int32)
        d = None if data is None else np.ascontiguousarray(data, np.int32)
        self.h = self.L.refc_build(s.size, _p(s), _p(e), _p(d))
        self.n = s.size

    def __del__(self):
        if getattr(self, "h", None):
            self.L.refc_free(self.h)
            self.h = None

    def arrays(self):
        s = np.empty(self.n, np.int32); e = np.empty(self.n, np.int32)
        d = np.empty(self.n, np.int32); b = np.empty(self.n, np.int64)
        self.L.refc_arrays(self.h, _p(s), _p(e), _p(d), _p(b))
        return {"starts": s, "ends": e, "data": d, "branch": b}

    def count(self, s, e):
        return int(self.L.refc_count(self.h, int(s), int(e)))

    def any(self, s, e):
        return bool(self.L.refc_any(self.h, int(s), int(e)))

    def upper_bound(self, x):
        return int(self.L.refc_upper_bound(self.h, int(x)))

    def search(self, s, e, which=0):
        h = self.L.refc_search(self.h, int(s), int(e), which, None, 0)
        out = np.empty(h, np.int32)
        self.L.refc_search(self.h, int(s), int(e), which, _p(out), h)
        return out

    def coverage(self, s, e):
        c = np.zeros(1, np.uint64); v = np.zeros(1, np.int32)
        self.L.refc_coverage(self.h, int(s), int(e), _p(c), _p(v))
        return int(c[0]), int(v[0])

    @classmethod
    def raw(cls, ref, starts, ends, data=None):
        """An unindexed handle holding the intervals in insertion order."""
        m = cls.__new__(cls)
        m.L = ref.lib
        s = np.ascontiguousarray(starts, np.int32)
        e = np.ascontiguousarray(ends, np.int32)
        d = None if data is None else np.ascontiguousarray(data, np.int32)
        m.h = m.L.refc_raw(s.size, _p(s), _p(e), _p(d))
        m.n = s.size
        return m

    def set_op(self, other, op):
        """C intersection (NULL combine) or difference with other (indexed):
        (starts, ends, data) of the unindexed result in its order."""
        h = (self.L.refc_intersection if op == "intersection" else self.L.refc_difference)(self.h, other.h)
        n = int(self.L.refc_size(h))
        s = np.empty(n, np.int32); e = np.empty(n, np.int32); d = np.empty(n, np.int32)
        if n:
            self.L.refc_raw_arrays(h, _p(s), _p(e), _p(d))
        self.L.refc_free(h)
        return s, e, d
