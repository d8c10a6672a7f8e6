// C++ API test: si::IntervalMap<S, T> from include/superintervals.hpp on the GPU.
//
// Known answers are the ones the reference's own test suite asserts for the query path
// (test/tests.cpp:63-238); the randomized part checks every query family against closed
// forms computed here on the host from the sorted arrays.  Exit status 0 = pass.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "superintervals.hpp"

static int failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);    \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

using Map = si::IntervalMap<int32_t, int>;

static void known_answers() {
    {   // tests.cpp:63-107
        Map m;
        CHECK(!m.has_overlaps(0, 1));
        m.add(10, 20, 0); m.add(11, 12, -1); m.add(13, 14, -1); m.add(15, 16, -1); m.add(25, 29, 4);
        m.build();
        CHECK(m.has_overlaps(17, 30));
        CHECK(!m.has_overlaps(1, 3));
        std::vector<int> v;
        m.search_values(17, 30, v);
        CHECK(v.size() == 2 && v[0] == 4 && v[1] == 0);
        CHECK(m.count(17, 30) == 2);
        std::vector<size_t> ix;
        m.search_idxs(17, 30, ix);
        CHECK(ix.size() == 2 && ix[0] == 4 && ix[1] == 0);
        std::vector<std::pair<int32_t, int32_t>> keys;
        m.search_keys(17, 30, keys);
        CHECK(keys.size() == 2 && keys[0].first == 25 && keys[1].second == 20);
        std::pair<size_t, int32_t> cov{0, 0};
        m.coverage(10, 29, cov);
        CHECK(cov.second == 17);
    }
    {   // tests.cpp:113-133
        Map m;
        m.add(1, 10, 0);
        m.build();
        size_t n = 0;
        int last = -1;
        for (size_t i : m.search_idxs(5, 11)) { last = m.data[i]; ++n; }
        CHECK(n == 1 && last == 0);
        for (const auto& it : m.search_items(5, 11)) CHECK(it.start == 1 && it.end == 10 && it.data == 0);
    }
    {   // tests.cpp:146-156
        Map m;
        m.add(0, 250000, 0);
        for (int s : {55, 115, 130, 281, 639, 842, 999, 1094, 1157, 1161, 1265, 1532, 1590, 1665, 1945, 2384, 2515})
            m.add(s, s + 1000, -1);
        m.build();
        std::vector<int> a;
        m.search_values(1377, 2377, a);
        CHECK(a.size() == 12 && a.back() == 0);
    }
    {   // tests.cpp:161-171
        Map m;
        m.add(3, 40, 0); m.add(4, 5, 4); m.add(6, 7, 4); m.add(10, 31, 5); m.add(31, 32, 5);
        m.build();
        std::vector<int> a;
        m.search_values(31, 32, a); CHECK(a.size() == 3); a.clear();
        m.search_values(10, 11, a); CHECK(a.size() == 2); a.clear();
        m.search_values(4, 7, a); CHECK(a.size() == 3);
    }
    {   // tests.cpp:175-186
        Map m;
        m.add(3, 40, 0); m.add(3, 40, 4); m.add(3, 40, 4); m.add(3, 4, 4); m.add(35, 50, 4);
        m.add(40, 400, 5); m.add(40, 400, 4);
        m.build();
        std::vector<int> a;
        m.search_values(38, 41, a); CHECK(a.size() == 6); a.clear();
        m.search_values(41, 42, a); CHECK(a.size() == 3);
    }
    {   // tests.cpp:194-213
        Map m;
        m.add(1, 100, 0); m.add(30, 200, 7); m.add(40, 50, 6); m.add(60, 70, 7);
        m.build();
        std::vector<int> a;
        m.search_values(55, 65, a);
        CHECK(a.size() == 3);
        std::pair<size_t, int32_t> cov{0, 0};
        m.coverage(55, 65, cov);
        CHECK(cov.first == 3 && cov.second == 25);
    }
    {   // tests.cpp:219-238
        Map e;
        e.build();
        CHECK(e.count(1, 5) == 0 && !e.has_overlaps(1, 5));
        Map one;
        one.add(1, 10, 1);
        one.build();
        CHECK(one.count(1, 5) == 1 && one.count(11, 20) == 0);
    }
    {   // vector order: the leading run is inserted ascending (superintervals.hpp:891-895)
        Map m;
        int s[] = {1, 2, 4, 5, 6, 60}, e[] = {100, 3, 50, 50, 50, 70};
        for (int k = 0; k < 6; ++k) m.add(s[k], e[k], 100 + k);
        m.build();
        std::vector<size_t> v;
        m.search_idxs(10, 65, v);
        CHECK((v == std::vector<size_t>{2, 3, 4, 5, 0}));
        std::vector<size_t> it;
        for (size_t i : m.search_idxs(10, 65)) it.push_back(i);
        CHECK((it == std::vector<size_t>{5, 4, 3, 2, 0}));
        std::vector<size_t> br(m.branch.begin(), m.branch.end());
        CHECK((br == std::vector<size_t>{SIZE_MAX, 0, 0, 2, 3, 0}));
        std::vector<int> vals;
        m.search_values(10, 65, vals);
        CHECK((vals == std::vector<int>{105, 104, 103, 102, 100}));
    }
}

template <typename S>
static void randomized(uint64_t seed) {
    std::mt19937_64 rng(seed);
    for (int round = 0; round < 4; ++round) {
        const size_t n = 1 + rng() % 5000;
        const int64_t span = 1000 + (int64_t)(rng() % 200000);
        const S off = (sizeof(S) == 8) ? (S)((int64_t)1 << 36) : 0;
        si::IntervalMap<S, int> m;
        for (size_t i = 0; i < n; ++i) {
            S s = off + (S)(rng() % span);
            m.add(s, s + (S)(rng() % (1 + span / 50)), (int)i);
        }
        m.build();
        // sorted by (start asc, end desc), branch = previous greater-or-equal end
        for (size_t i = 1; i < n; ++i)
            CHECK(m.starts[i - 1] < m.starts[i] || (m.starts[i - 1] == m.starts[i] && m.ends[i - 1] >= m.ends[i]));
        for (size_t i = 0; i < n; i += 7) {
            size_t want = SIZE_MAX;
            for (size_t j = i; j-- > 0;) if (m.ends[j] >= m.ends[i]) { want = j; break; }
            CHECK(m.branch[i] == want);
        }
        std::vector<S> qs(300), qe(300);
        for (size_t q = 0; q < qs.size(); ++q) {
            qs[q] = off + (S)(rng() % (span + 100)) - 50;
            qe[q] = qs[q] + (S)(rng() % (span / 20 + 1));
        }
        std::vector<uint32_t> counts(qs.size());
        CHECK(m.count_batch(qs.data(), qe.data(), qs.size(), counts.data()));
        std::vector<uint64_t> off_v, off_d;
        std::vector<uint32_t> idx_v, idx_d;
        CHECK(m.search_idxs_batch(qs.data(), qe.data(), qs.size(), off_v, idx_v, true));
        CHECK(m.search_idxs_batch(qs.data(), qe.data(), qs.size(), off_d, idx_d, false));
        for (size_t q = 0; q < qs.size(); ++q) {
            const S a = qs[q], b = qe[q];
            size_t ub = (size_t)(std::upper_bound(m.starts.begin(), m.starts.end(), b) - m.starts.begin());
            std::vector<size_t> strict;
            for (size_t j = ub; j-- > 0;) if (m.ends[j] >= a) strict.push_back(j);
            size_t i = ub;
            while (i > 0 && m.ends[i - 1] >= a) --i;   // [i, ub) is the leading run
            std::vector<size_t> vec;
            for (size_t j = i; j < ub; ++j) vec.push_back(j);
            for (size_t j : strict) if (j < i) vec.push_back(j);
            CHECK(m.count(a, b) == strict.size());
            CHECK(counts[q] == strict.size());
            CHECK(m.upper_bound(b) == (ub == 0 ? SIZE_MAX : ub - 1));
            std::vector<size_t> got;
            m.search_idxs(a, b, got);
            CHECK(got == vec);
            std::vector<size_t> it;
            for (size_t j : m.search_idxs(a, b)) it.push_back(j);
            CHECK(it == strict);
            CHECK(off_v[q + 1] - off_v[q] == vec.size() && off_d[q + 1] - off_d[q] == strict.size());
            for (size_t k = 0; k < vec.size(); ++k) CHECK(idx_v[off_v[q] + k] == vec[k]);
            for (size_t k = 0; k < strict.size(); ++k) CHECK(idx_d[off_d[q] + k] == strict[k]);
            std::vector<int> vals;
            m.search_values(a, b, vals);
            CHECK(vals.size() == strict.size());
            for (size_t k = 0; k < vals.size() && k < strict.size(); ++k) CHECK(vals[k] == m.data[strict[k]]);
            CHECK(m.has_overlaps(a, b) == (ub > 0 && m.ends[ub - 1] >= a));
        }
        // payload follows its interval through the sort
        for (size_t k = 0; k < n; ++k) CHECK(m.data[k] >= 0 && (size_t)m.data[k] < n);
    }
}

// intersection / difference (superintervals.hpp:1158-1211) on a hand-checked case
void set_ops() {
    si::IntervalMap<int32_t, int> a, b;
    a.add(1, 100, 0); a.add(40, 60, 1); a.add(200, 210, 2); a.add(95, 130, 3);   // not built
    b.add(10, 20, 10); b.add(50, 70, 11); b.add(55, 58, 12); b.add(205, 400, 13);
    b.build();
    si::IntervalMap<int32_t, int> d = a.difference(b);
    const int want[][3] = {{1, 9, 0}, {21, 49, 0}, {71, 100, 0}, {40, 49, 1}, {200, 204, 2}, {95, 130, 3}};
    CHECK(d.size() == 6);
    for (size_t i = 0; i < 6 && i < d.size(); ++i)
        CHECK(d.starts[i] == want[i][0] && d.ends[i] == want[i][1] && d.data[i] == want[i][2]);
    si::IntervalMap<int32_t, int> x = a.intersection(b, [](const int& p, const int& q) { return p * 100 + q; });
    // pieces per a interval, other's hits in search_idxs order
    size_t n = 0;
    for (size_t k = 0; k < a.starts.size(); ++k) {
        std::vector<size_t> h;
        b.search_idxs(a.starts[k], a.ends[k], h);
        for (size_t j : h) {
            CHECK(n < x.size());
            if (n >= x.size()) break;
            CHECK(x.starts[n] == std::max(a.starts[k], b.starts[j]) && x.ends[n] == std::min(a.ends[k], b.ends[j]));
            CHECK(x.data[n] == a.data[k] * 100 + b.data[j]);
            ++n;
        }
    }
    CHECK(n == x.size() && n == 6);   // [10,20] [50,70] [55,58] in a0; [50,60] [55,58] in a1; [205,210] in a2
    // the type's maximum: covered to the end leaves only the lower part
    si::IntervalMap<int32_t, int> c, e;
    const int32_t big = 2147483647;
    c.add(big - 5, big, 7);
    e.add(big - 2, big, 8);
    e.build();
    si::IntervalMap<int32_t, int> cd = c.difference(e);
    CHECK(cd.size() == 1 && cd.starts[0] == big - 5 && cd.ends[0] == big - 3 && cd.data[0] == 7);
}

// ---- file-driven checks (vectors written by the Python tests; int64 words) ---------------
static std::vector<int64_t> read_words(const char* path) {
    std::vector<int64_t> w;
    FILE* f = std::fopen(path, "rb");
    if (!f) return w;
    int64_t x;
    while (std::fread(&x, sizeof(x), 1, f) == 1) w.push_back(x);
    std::fclose(f);
    return w;
}

static bool write_words(const char* path, const std::vector<int64_t>& w) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(w.data(), sizeof(int64_t), w.size(), f) == w.size();
    return std::fclose(f) == 0 && ok;
}

// upper_bound_range on a map whose starts are given (host code only, no GPU):
// in = [n, starts[n], k, value[k], left[k], right[k]], out = exit left[k] (-1 = SIZE_MAX).
template <typename S>
static int ubr_mode(const std::vector<int64_t>& w, std::vector<int64_t>& out) {
    size_t p = 1;
    const size_t n = (size_t)w[p++];
    si::IntervalMap<S, int> m;
    for (size_t i = 0; i < n; ++i) m.starts.push_back((S)w[p++]);
    const size_t k = (size_t)w[p++];
    const int64_t* value = &w[p];
    const int64_t* left = value + k;
    const int64_t* right = left + k;
    for (size_t q = 0; q < k; ++q) {
        size_t l = (size_t)left[q];
        m.upper_bound_range((S)value[q], l, (size_t)right[q]);
        out.push_back(l == SIZE_MAX ? -1 : (int64_t)l);
    }
    return 0;
}

// intersection and difference of A (in order) with B (built), on the GPU:
// in = [na, a_s, a_e, a_d, nb, b_s, b_e, b_d], out = [ni, s, e, d, nd, s, e, d].
template <typename S>
static int setops_mode(const std::vector<int64_t>& w, std::vector<int64_t>& out) {
    size_t p = 1;
    si::IntervalMap<S, int> a, b;
    const size_t na = (size_t)w[p++];
    for (size_t i = 0; i < na; ++i) a.add((S)w[p + i], (S)w[p + na + i], (int)w[p + 2 * na + i]);
    p += 3 * na;
    const size_t nb = (size_t)w[p++];
    for (size_t i = 0; i < nb; ++i) b.add((S)w[p + i], (S)w[p + nb + i], (int)w[p + 2 * nb + i]);
    b.build();
    for (auto* r : {new si::IntervalMap<S, int>(a.intersection(b)), new si::IntervalMap<S, int>(a.difference(b))}) {
        out.push_back((int64_t)r->starts.size());
        for (S x : r->starts) out.push_back((int64_t)x);
        for (S x : r->ends) out.push_back((int64_t)x);
        for (int x : r->data) out.push_back((int64_t)x);
        delete r;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4) {   // <ubr|setops> <in> <out>; word 0 of in = coordinate bytes
        const std::string mode = argv[1];
        const std::vector<int64_t> w = read_words(argv[2]);
        if (w.empty()) return 3;
        std::vector<int64_t> out;
        int rc;
        if (mode == "ubr") rc = (w[0] == 4) ? ubr_mode<int32_t>(w, out) : ubr_mode<int64_t>(w, out);
        else if (mode == "setops") rc = (w[0] == 4) ? setops_mode<int32_t>(w, out) : setops_mode<int64_t>(w, out);
        else return 3;
        return (rc == 0 && write_words(argv[3], out)) ? 0 : 1;
    }
    siIndex* probe = siIndexCreate(4);
    if (!probe) {
        std::fprintf(stderr, "no GPU: %s\n", siGetLastError());
        return 2;
    }
    siIndexDestroy(probe);
    known_answers();
    randomized<int32_t>(1);
    randomized<int64_t>(2);
    set_ops();
    if (failures) {
        std::fprintf(stderr, "%d failure(s)\n", failures);
        return 1;
    }
    std::printf("superintervals.hpp C++ API: all checks passed\n");
    return 0;
}
