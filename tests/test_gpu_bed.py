"""ChromIndex: one device index per chromosome from BED records, queries batched per
chromosome and returned in input order, against the oracle on each chromosome."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_chrom_index_matches_oracle(orc, tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from superintervals_amd import bed
    print("[gpu] ChromIndex over a BED file of three chromosomes", flush=True)
    rng = np.random.default_rng(7)
    lines, want = [], {}
    for c, n in (("chr1", 3000), ("chr2", 1500), ("chrM", 40)):
        s = rng.integers(0, 200_000, n)
        ln = rng.integers(1, 2_000, n)   # half-open length >= 1
        for k in range(n):
            lines.append("%s\t%d\t%d\tiv_%s_%d" % (c, s[k], s[k] + ln[k], c, k))
        want[c] = (s.astype(np.int32), (s + ln - 1).astype(np.int32))
    rng.shuffle(lines)
    path = tmp_path / "x.bed"
    path.write_text("track name=test\n" + "\n".join(lines) + "\n")
    ci = bed.ChromIndex.from_bed(str(path))
    assert sorted(ci.chroms()) == ["chr1", "chr2", "chrM"] and ci.dtype == np.int32
    nq = 4000
    qc = rng.choice(np.array(["chr1", "chr2", "chrM", "chrUn"]), nq)
    qs = rng.integers(0, 200_000, nq)
    qe = qs + rng.integers(1, 5_000, nq)
    counts = ci.count(qc, qs, qe, half_open=True)
    off, idx = ci.search(qc, qs, qe, half_open=True)
    assert np.array_equal(np.diff(off).astype(np.uint32), counts)
    for c in ("chr1", "chr2", "chrM"):
        # the oracle index over this chromosome's intervals in file order
        s_c = np.array([int(l.split("\t")[1]) for l in lines if l.startswith(c + "\t")], np.int32)
        e_c = np.array([int(l.split("\t")[2]) - 1 for l in lines if l.startswith(c + "\t")], np.int32)
        oix = orc.build_index(s_c, e_c)
        sel = np.nonzero(qc == c)[0]
        ocount = orc.count_batch(oix, qs[sel].astype(np.int32), (qe[sel] - 1).astype(np.int32))
        assert np.array_equal(counts[sel], ocount), c
        ooff, oidx = orc.search_batch(oix, qs[sel].astype(np.int32), (qe[sel] - 1).astype(np.int32), mode=0)
        for j, q in enumerate(sel.tolist()):
            assert np.array_equal(idx[off[q]:off[q + 1]].astype(np.int64), oidx[ooff[j]:ooff[j + 1]]), (c, q)
        # names follow the sorted positions
        names = ci.names[c]
        arr = ci.index(c).download()
        for p in range(0, len(names), 97):
            k = int(names[p].rsplit("_", 1)[1])
            assert (int(arr["starts"][p]), int(arr["ends"][p])) == (int(want[c][0][k]), int(want[c][1][k]))
    assert counts[qc == "chrUn"].sum() == 0
    ci.close()
