import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_CASES = ["c1_sorted", "c1_shuffled", "c1_i64", "dups", "c2_1m"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


def load_golden(name):
    """(record dict, starts, ends, qs, qe) with the inputs regenerated or stored."""
    import hashlib
    import superintervals_amd as sia
    rec = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    if "in_starts" in rec:
        s, e, qs, qe = rec["in_starts"], rec["in_ends"], rec["in_qs"], rec["in_qe"]
    else:
        cfg = int(rec["config"])
        dt = np.dtype(str(rec["dtype"]))
        s, e = sia.workload_intervals(cfg, int(rec["n"]), seed=int(rec["seed"]), sorted=bool(rec["sorted"]),
                                      coord_dtype=dt)
        qs, qe = sia.workload_queries(cfg, int(rec["nq"]), seed=int(rec["qseed"]), coord_dtype=dt)
        if str(rec["transform"]) == "plus2^33":
            off = np.int64(1) << 33
            s, e, qs, qe = s + off, e + off, qs + off, qe + off
    h = hashlib.sha256()
    for a in (s, e, qs, qe):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == str(rec["input_sha256"]), f"{name}: regenerated inputs differ from the fixture's"
    return rec, s, e, qs, qe


def sha(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


def sweep_difference(a_s, a_e, ob_s, ob_e, hits):
    """The reference's difference sweep (superintervals.hpp:1188-1211, c_superintervals.h:
    950-975) in Python integers: for each interval k of A, B's hit keys sorted ascending by
    (start, end), pieces (start, end, k) of what they leave uncovered."""
    out = []
    for k in range(len(a_s)):
        cover = sorted((int(ob_s[h]), int(ob_e[h])) for h in hits[k])
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


def overlap_pieces(a_s, a_e, ob_s, ob_e, hits):
    """The reference's intersection (superintervals.hpp:1158-1174, c_superintervals.h:
    921-940): for each interval k of A and each hit h in the given order, the piece
    (max(starts), min(ends), k, h) when it is not empty."""
    out = []
    for k in range(len(a_s)):
        for h in hits[k]:
            s, e = max(int(a_s[k]), int(ob_s[h])), min(int(a_e[k]), int(ob_e[h]))
            if s <= e:
                out.append((s, e, k, int(h)))
    return out
