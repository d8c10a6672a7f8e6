"""BED parsing (host logic, no GPU): the half-open to end-inclusive convention of the
reference's BED readers (test/bench.cpp:210, :220: add(start, end - 1)), header and
malformed lines skipped, empty intervals dropped as the reference's reader drops
last < first (examples/bed-intersect-si.rs:100-123)."""
import numpy as np


def test_parse_bed_convention():
    import importlib.util, os
    from conftest import ROOT
    spec = importlib.util.spec_from_file_location("bed", os.path.join(ROOT, "superintervals_amd", "bed.py"))
    bed = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bed)
    text = [
        "# a comment",
        "track name=x",
        "browser position chr1:1-100",
        "chr1\t10\t20\tgeneA",
        "chr1\t5\t6\tgeneB",
        "chr2\t100\t100\tempty",          # empty half-open interval: dropped
        "chr2\t7\t9",
        "chr1 30 40 spaced",             # whitespace-separated fallback
        "chrX\tabc\t10",                 # malformed: skipped
        "chr3\t1",                       # too few fields: skipped
        b"chr3\t0\t1\tbytes",
    ]
    d = bed.parse_bed(text)
    assert sorted(d) == ["chr1", "chr2", "chr3"]
    s, e, n = d["chr1"]
    assert s.tolist() == [10, 5, 30] and e.tolist() == [19, 5, 39] and n == ["geneA", "geneB", "spaced"]
    s, e, n = d["chr2"]
    assert s.tolist() == [7] and e.tolist() == [8] and n is None
    assert d["chr3"][0].tolist() == [0] and d["chr3"][1].tolist() == [0]
    # end-inclusive input kept as is
    d2 = bed.parse_bed(["chr1\t10\t20"], half_open=False)
    assert d2["chr1"][1].tolist() == [20]
