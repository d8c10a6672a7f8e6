// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" wrapper around the reference's own C++ template
// (src/superintervals.hpp of kcleal/superintervals v1.0.0), compiled in place
// from the reference checkout by oracle/Makefile into oracle/_ref/.  It is used
// to write golden vectors, to cross-check the restatement in si_oracle.c, and
// as the "reference" CPU baseline.  No reference source is copied here: the
// header is included from the reference tree at build time.
#include "superintervals.hpp"
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
template <typename S> using Map = si::IntervalMap<S, int32_t>;

template <typename S>
Map<S>* build(size_t n, const S* starts, const S* ends, const int32_t* data) {
    auto* m = new Map<S>();
    m->reserve(n);
    for (size_t i = 0; i < n; ++i) m->add(starts[i], ends[i], data ? data[i] : (int32_t)i);
    m->build();
    return m;
}

template <typename S>
void arrays(const Map<S>* m, S* starts, S* ends, int32_t* data, int64_t* branch) {
    size_t n = m->starts.size();
    for (size_t i = 0; i < n; ++i) {
        if (starts) starts[i] = m->starts[i];
        if (ends) ends[i] = m->ends[i];
        if (data) data[i] = m->data[i];
        if (branch) branch[i] = (m->branch[i] == SIZE_MAX) ? -1 : (int64_t)m->branch[i];
    }
}

// mode 0: lazy IndexRange iterator (strict descending), mode 1: vector form.
template <typename S>
size_t search(const Map<S>* m, S s, S e, int mode, int64_t* out, size_t cap) {
    size_t h = 0;
    if (mode == 0) {
        for (size_t i : m->search_idxs(s, e)) { if (out && h < cap) out[h] = (int64_t)i; ++h; }
    } else {
        std::vector<size_t> v;
        m->search_idxs(s, e, v);
        for (size_t i : v) { if (out && h < cap) out[h] = (int64_t)i; ++h; }
    }
    return h;
}

template <typename S>
size_t values(const Map<S>* m, S s, S e, int32_t* out, size_t cap) {
    std::vector<int32_t> v;
    m->search_values(s, e, v);
    for (size_t k = 0; k < v.size() && out && k < cap; ++k) out[k] = v[k];
    return v.size();
}

// Multi-threaded batch count: static contiguous chunks over a shared const map.
template <typename S>
uint64_t count_batch(const Map<S>* m, int nthreads, size_t nq, const S* qs, const S* qe,
                     uint32_t* counts) {
    if (nthreads < 1) nthreads = 1;
    std::vector<uint64_t> part(nthreads, 0);
    std::vector<std::thread> ws;
    for (int t = 0; t < nthreads; ++t) {
        ws.emplace_back([&, t] {
            size_t a = nq * t / nthreads, b = nq * (t + 1) / nthreads;
            uint64_t c = 0;
            for (size_t q = a; q < b; ++q) {
                size_t k = m->count(qs[q], qe[q]);
                if (counts) counts[q] = (uint32_t)k;
                c += k;
            }
            part[t] = c;
        });
    }
    for (auto& w : ws) w.join();
    uint64_t tot = 0;
    for (auto p : part) tot += p;
    return tot;
}

// Multi-threaded batch search_idxs (vector form, one reused vector per thread,
// as test/bench.cpp:217-227 does for search_values).  Returns the hit total.
template <typename S>
uint64_t search_batch(const Map<S>* m, int nthreads, size_t nq, const S* qs, const S* qe,
                      uint32_t* counts) {
    if (nthreads < 1) nthreads = 1;
    std::vector<uint64_t> part(nthreads, 0);
    std::vector<std::thread> ws;
    for (int t = 0; t < nthreads; ++t) {
        ws.emplace_back([&, t] {
            size_t a = nq * t / nthreads, b = nq * (t + 1) / nthreads;
            uint64_t c = 0;
            std::vector<size_t> f;
            f.reserve(1 << 16);
            for (size_t q = a; q < b; ++q) {
                f.clear();
                m->search_idxs(qs[q], qe[q], f);
                if (counts) counts[q] = (uint32_t)f.size();
                c += f.size();
            }
            part[t] = c;
        });
    }
    for (auto& w : ws) w.join();
    uint64_t tot = 0;
    for (auto p : part) tot += p;
    return tot;
}
}  // namespace

#define REF_INSTANCE(S, SFX)                                                                    \
    REF_API void* ref_build_##SFX(size_t n, const S* st, const S* en, const int32_t* d) {     \
        return build<S>(n, st, en, d);                                                         \
    }                                                                                          \
    REF_API void ref_free_##SFX(void* m) { delete (Map<S>*)m; }                                \
    REF_API size_t ref_size_##SFX(void* m) { return ((Map<S>*)m)->starts.size(); }             \
    REF_API void ref_arrays_##SFX(void* m, S* st, S* en, int32_t* d, int64_t* br) {            \
        arrays<S>((Map<S>*)m, st, en, d, br);                                                  \
    }                                                                                          \
    REF_API int64_t ref_upper_bound_##SFX(void* m, S x) {                                      \
        size_t i = ((Map<S>*)m)->upper_bound(x);                                               \
        return i == SIZE_MAX ? -1 : (int64_t)i;                                                \
    }                                                                                          \
    REF_API uint64_t ref_count_##SFX(void* m, S s, S e) { return ((Map<S>*)m)->count(s, e); }  \
    REF_API int ref_has_overlaps_##SFX(void* m, S s, S e) {                                    \
        return ((Map<S>*)m)->has_overlaps(s, e) ? 1 : 0;                                       \
    }                                                                                          \
    REF_API size_t ref_search_##SFX(void* m, S s, S e, int mode, int64_t* out, size_t cap) {   \
        return search<S>((Map<S>*)m, s, e, mode, out, cap);                                    \
    }                                                                                          \
    REF_API size_t ref_values_##SFX(void* m, S s, S e, int32_t* out, size_t cap) {             \
        return values<S>((Map<S>*)m, s, e, out, cap);                                          \
    }                                                                                          \
    REF_API uint64_t ref_count_batch_##SFX(void* m, int t, size_t nq, const S* qs,             \
                                           const S* qe, uint32_t* c) {                         \
        return count_batch<S>((Map<S>*)m, t, nq, qs, qe, c);                                   \
    }                                                                                          \
    REF_API uint64_t ref_search_batch_##SFX(void* m, int t, size_t nq, const S* qs,            \
                                            const S* qe, uint32_t* c) {                        \
        return search_batch<S>((Map<S>*)m, t, nq, qs, qe, c);                                  \
    }

// upper_bound_range (superintervals.hpp:524-543): the exit value of `left`, -1 for SIZE_MAX.
#define REF_EXTRA(S, SFX)                                                                       \
    REF_API int64_t ref_upper_bound_range_##SFX(void* m, S value, size_t left, size_t right) {  \
        ((Map<S>*)m)->upper_bound_range(value, left, right);                                   \
        return left == SIZE_MAX ? -1 : (int64_t)left;                                          \
    }                                                                                          \
    REF_API void* ref_raw_##SFX(size_t n, const S* st, const S* en, const int32_t* d) {         \
        auto* m = new Map<S>();                                                                \
        for (size_t i = 0; i < n; ++i) m->add(st[i], en[i], d ? d[i] : (int32_t)i);             \
        return m;                                                                              \
    }                                                                                          \
    REF_API void ref_raw_arrays_##SFX(void* m, S* st, S* en, int32_t* d) {                      \
        arrays<S>((Map<S>*)m, st, en, d, nullptr);                                             \
    }                                                                                          \
    REF_API void* ref_intersection_##SFX(void* a, void* b) {                                   \
        return new Map<S>(((Map<S>*)a)->intersection(*(Map<S>*)b));                            \
    }                                                                                          \
    REF_API void* ref_difference_##SFX(void* a, void* b) {                                     \
        return new Map<S>(((Map<S>*)a)->difference(*(Map<S>*)b));                              \
    }

REF_INSTANCE(int32_t, i32)
REF_INSTANCE(int64_t, i64)
REF_EXTRA(int32_t, i32)
REF_EXTRA(int64_t, i64)

REF_API void ref_coverage_i32(void* m, int32_t s, int32_t e, uint64_t* c, int32_t* cov) {
    std::pair<size_t, int> r{0, 0};
    ((Map<int32_t>*)m)->coverage(s, e, r);
    *c = r.first;
    *cov = r.second;
}
