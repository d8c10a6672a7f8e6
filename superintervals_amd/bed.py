"""BED ingestion and a per-chromosome set of device indexes.

read_bed follows the reference's BED convention: BED intervals are half-open [start, end)
and the index is end-inclusive, so each interval is added as (start, end - 1)
(test/bench.cpp:210, :220).  ChromIndex is the per-sequence map of indexes the reference's
examples build (examples/bed-intersect-si.rs:100-123: one IntervalMap per sequence name),
with queries batched per chromosome on the GPU.
"""
from __future__ import annotations

import numpy as np


def parse_bed(lines, half_open: bool = True):
    """{chrom: (starts int64, ends int64, names)} from BED text lines.

    Header lines (`#`, `track`, `browser`) and lines with fewer than three fields are
    skipped; so are intervals whose inclusive end lies before their start (empty BED
    intervals), as the reference's reader skips last < first.  names holds column 4 when
    present, else None."""
    acc = {}
    for line in lines:
        if isinstance(line, bytes):
            line = line.decode()
        line = line.rstrip("\r\n")
        if not line or line.startswith(("#", "track", "browser")):
            continue
        f = line.split("\t") if "\t" in line else line.split()
        if len(f) < 3:
            continue
        try:
            s, e = int(f[1]), int(f[2])
        except ValueError:
            continue
        if half_open:
            e -= 1
        if e < s:
            continue
        ss, ee, nn = acc.setdefault(f[0], ([], [], []))
        ss.append(s)
        ee.append(e)
        nn.append(f[3] if len(f) > 3 else None)
    out = {}
    for c, (ss, ee, nn) in acc.items():
        out[c] = (np.asarray(ss, np.int64), np.asarray(ee, np.int64), nn if any(x is not None for x in nn) else None)
    return out


def read_bed(path: str, half_open: bool = True):
    """parse_bed over the lines of a BED file."""
    with open(path, "rb") as f:
        return parse_bed(f, half_open)


class ChromIndex:
    """One device index per chromosome, built from {chrom: (starts, ends, names)}.

    Coordinates are int32 when every chromosome fits, else int64.  Queries are
    (chroms, starts, ends) batches in any order of chromosomes; each chromosome's queries
    go to its index in one batched call, and results come back in input order.  A query on
    a chromosome with no intervals finds nothing."""

    def __init__(self, intervals: dict):
        from . import DeviceIndex
        big = max((int(max(np.abs(s).max(), np.abs(e).max())) for s, e, _ in intervals.values() if s.size),
                  default=0)
        self.dtype = np.int32 if big < 2**31 - 1 else np.int64
        self.names = {}
        self._ix = {}
        for c, (s, e, nn) in intervals.items():
            ix = DeviceIndex(self.dtype).build(np.asarray(s, self.dtype), np.asarray(e, self.dtype))
            self._ix[c] = ix
            if nn is not None:
                perm = ix.download()["perm"]
                self.names[c] = [nn[p] for p in perm.tolist()]   # names in sorted position order

    @classmethod
    def from_bed(cls, path: str, half_open: bool = True):
        return cls(read_bed(path, half_open))

    def chroms(self):
        return list(self._ix)

    def index(self, chrom: str):
        return self._ix[chrom]

    def _cast(self, x):
        info = np.iinfo(self.dtype)
        return np.clip(x, info.min, info.max).astype(self.dtype)   # no interval lies outside

    def _groups(self, chroms):
        chroms = np.asarray(chroms)
        for c in np.unique(chroms):
            yield str(c), np.nonzero(chroms == c)[0]

    def count(self, chroms, starts, ends, half_open: bool = False) -> np.ndarray:
        """counts[q] of intervals on chroms[q] overlapping [starts[q], ends[q]] (end - 1 with
        half_open queries)."""
        starts = np.asarray(starts, np.int64)
        ends = np.asarray(ends, np.int64) - (1 if half_open else 0)
        out = np.zeros(starts.size, np.uint32)
        for c, sel in self._groups(chroms):
            if c in self._ix:
                out[sel] = self._ix[c].count(self._cast(starts[sel]), self._cast(ends[sel]))
        return out

    def search(self, chroms, starts, ends, half_open: bool = False):
        """CSR in input order: offsets[nq+1] and, per query, the descending positions of its
        hits in its chromosome's index (names[chrom][position] gives a hit's name)."""
        starts = np.asarray(starts, np.int64)
        ends = np.asarray(ends, np.int64) - (1 if half_open else 0)
        nq = starts.size
        lists = [None] * nq
        empty = np.zeros(0, np.uint32)
        for c, sel in self._groups(chroms):
            if c not in self._ix:
                continue
            off, idx = self._ix[c].search(self._cast(starts[sel]), self._cast(ends[sel]))
            for j, q in enumerate(sel.tolist()):
                lists[q] = idx[off[j]:off[j + 1]]
        counts = np.array([0 if x is None else x.size for x in lists], np.uint64)
        offsets = np.zeros(nq + 1, np.uint64)
        np.cumsum(counts, out=offsets[1:])
        idx = np.concatenate([empty if x is None else x for x in lists]) if nq else empty
        return offsets, idx

    def close(self):
        for ix in self._ix.values():
            ix.close()
        self._ix = {}
