// superintervals.hpp -- si::IntervalMap<S, T> on the MI355X device index.
//
// Drop-in for the reference's C++ template (kcleal/superintervals v1.0.0,
// src/superintervals.hpp:45-1454) on the query path: the same public members
// (starts, ends, branch, data, start_sorted, end_sorted; :49-61) and the same
// add / build / query methods with the same results and result order.  build()
// sorts and indexes on the GPU (libsuperintervals_hip.so, si_device.h) and mirrors
// the sorted arrays and the branch array back into the public members; every query
// runs on the GPU.  Batch overloads (count_batch, search_idxs_batch,
// search_values_batch) answer many queries per call and are the fast path.
//
// S must be int32_t or int64_t (the device index's coordinate types).  T is any
// copyable payload; it stays on the host and is gathered by position.
//
// Differences a caller can observe:
//   * Exact (start, end) duplicates keep their insertion order after build() (the
//     reference's std::sort leaves their payload order unspecified).
//   * Single-query calls are GPU round trips; prefer the batch overloads.
//   * Device failures are reported by si::IntervalMap::last_error() (and
//     siGetLastError()); the query then returns an empty / zero result.
//   * Of the set operations (:1037-1390) intersection and difference are provided (one
//     batched device search each, below); merge_overlaps, gaps, union_with,
//     symmetric_difference, span, expand, flank and unique are host-side sorts and
//     merges and are not.
//   * count_large and search_values_large return count's and search_values' results.
//     The reference's versions count the run [first start >= query start, upper_bound]
//     without reading ends, which differs from count() only on an index holding an
//     interval with start > end.
//   * IntervalMapEytz (:1457-1535) is provided as the same GPU index; its
//     upper_bound returns SIZE_MAX when the value is below every start, like the base.
//
// Link with -lsuperintervals_hip (superintervals_amd/libsuperintervals_hip.so).
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "si_device.h"

namespace si {

// superintervals.hpp:23-29
template <typename S, typename T>
struct Interval {
    S start, end;
    T data;
    Interval() = default;
    Interval(S s, S e, T d) : start(s), end(e), data(d) {}
};

template <typename S, typename T>
class IntervalMap {
    static_assert(std::is_same<S, int32_t>::value || std::is_same<S, int64_t>::value,
                  "the device index supports int32_t and int64_t coordinates");

  public:
    std::vector<S> starts;
    std::vector<S> ends;
    std::vector<size_t> branch;
    std::vector<T> data;
    bool start_sorted, end_sorted;

    IntervalMap() : start_sorted(true), end_sorted(true) {}
    virtual ~IntervalMap() { siIndexDestroy(dev_); }
    IntervalMap(const IntervalMap& o)
        : starts(o.starts), ends(o.ends), branch(o.branch), data(o.data),
          start_sorted(o.start_sorted), end_sorted(o.end_sorted) {
        if (o.dev_ && siIndexSize(o.dev_) == o.starts.size()) rebuild_device();
    }
    IntervalMap& operator=(const IntervalMap& o) {
        if (this != &o) {
            starts = o.starts; ends = o.ends; branch = o.branch; data = o.data;
            start_sorted = o.start_sorted; end_sorted = o.end_sorted;
            siIndexDestroy(dev_);
            dev_ = nullptr;
            if (o.dev_ && siIndexSize(o.dev_) == o.starts.size()) rebuild_device();
        }
        return *this;
    }

    // ---- storage (superintervals.hpp:69-105) ---------------------------------------
    void clear() noexcept {
        data.clear(); starts.clear(); ends.clear(); branch.clear();
        built_ = false;
    }
    void reserve(size_t n) { data.reserve(n); starts.reserve(n); ends.reserve(n); }
    size_t size() { return starts.size(); }

    void add(S start, S end, const T& value) {
        if (start_sorted && !starts.empty()) {
            start_sorted = !(start < starts.back());
            if (start_sorted && start == starts.back() && end > ends.back()) end_sorted = false;
        }
        starts.push_back(start);
        ends.push_back(end);
        data.emplace_back(value);
        built_ = false;
    }

    // ---- build (superintervals.hpp:112-130) -------------------------------------------
    virtual void build() {
        if (starts.empty()) return;
        if (!dev_) dev_ = siIndexCreate((int)sizeof(S));
        if (!dev_) return;
        const size_t n = starts.size();
        if (siIndexBuild(dev_, starts.data(), ends.data(), nullptr, n, start_sorted, end_sorted, 0,
                         nullptr) != SI_OK)
            return;
        std::vector<uint32_t> perm(n);
        branch.resize(n);
        static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t must be 64-bit");
        if (siIndexDownload(dev_, starts.data(), ends.data(), nullptr, perm.data(),
                            reinterpret_cast<uint64_t*>(branch.data())) != SI_OK)
            return;
        if (!(start_sorted && end_sorted)) {
            std::vector<T> sorted;
            sorted.reserve(n);
            for (size_t k = 0; k < n; ++k) sorted.push_back(data[perm[k]]);
            data.swap(sorted);
        }
        start_sorted = end_sorted = true;
        built_ = true;
    }

    Interval<S, T> at(size_t index) const { return Interval<S, T>{starts[index], ends[index], data[index]}; }
    void at(size_t index, Interval<S, T>& itv) {
        itv.start = starts[index]; itv.end = ends[index]; itv.data = data[index];
    }

    // ---- result ranges (superintervals.hpp:153-494) ---------------------------------
    // Each range holds the query's strictly descending positions, computed on the GPU
    // when the range is created; iteration yields what the reference's iterators yield.
    template <typename Get>
    class Range {
      public:
        using value_type = decltype(std::declval<Get>()(size_t(0)));
        class iterator {
          public:
            iterator(const Range* r, size_t k) : r_(r), k_(k) {}
            value_type operator*() const { return r_->get_(r_->pos_[k_]); }
            iterator& operator++() { ++k_; return *this; }
            bool operator!=(const iterator& o) const { return k_ != o.k_; }
            bool operator==(const iterator& o) const { return k_ == o.k_; }
          private:
            const Range* r_;
            size_t k_;
        };
        Range(std::vector<size_t> pos, Get g) : pos_(std::move(pos)), get_(g) {}
        iterator begin() const { return iterator(this, 0); }
        iterator end() const { return iterator(this, pos_.size()); }
        size_t size() const { return pos_.size(); }
      private:
        std::vector<size_t> pos_;
        Get get_;
    };

    struct GetIdx { size_t operator()(size_t i) const { return i; } };
    struct GetKey {
        const IntervalMap* m;
        std::pair<S, S> operator()(size_t i) const { return {m->starts[i], m->ends[i]}; }
    };
    struct GetVal {
        const IntervalMap* m;
        T operator()(size_t i) const { return m->data[i]; }
    };
    struct GetItem {
        const IntervalMap* m;
        Interval<S, T> operator()(size_t i) const { return Interval<S, T>{m->starts[i], m->ends[i], m->data[i]}; }
    };

    Range<GetIdx> search_idxs(S start, S end) const noexcept {
        return Range<GetIdx>(positions(start, end, SI_ORDER_DESC), GetIdx{});
    }
    Range<GetKey> search_keys(S start, S end) const noexcept {
        return Range<GetKey>(positions(start, end, SI_ORDER_DESC), GetKey{this});
    }
    Range<GetVal> search_values(S start, S end) const noexcept {
        return Range<GetVal>(positions(start, end, SI_ORDER_DESC), GetVal{this});
    }
    Range<GetItem> search_items(S start, S end) const noexcept {
        return Range<GetItem>(positions(start, end, SI_ORDER_DESC), GetItem{this});
    }

    // ---- single queries ---------------------------------------------------------------
    // superintervals.hpp:501-516: last position with starts[i] <= value, else SIZE_MAX.
    virtual size_t upper_bound(const S value) const noexcept {
        if (!ready()) return SIZE_MAX;
        uint32_t p = SI_INDEX_NONE;
        if (siUpperBoundHost(dev_, &value, 1, &p) != SI_OK) return SIZE_MAX;
        return p == SI_INDEX_NONE ? SIZE_MAX : (size_t)p;
    }

    // superintervals.hpp:518-543: moves left down towards the first position whose start is
    // >= value, for a caller scanning the run (left, right] of positions that start at or
    // after value.  It gallops down from left (steps 1, 2, 4, ... while the start there is
    // above value), then halves the last step's range [left, hi): a probe that starts below
    // value moves left just past it.  On exit every position above left (up to right)
    // starts at or after value: left is the position below the first such start, or one
    // in the run starting exactly at value; SIZE_MAX when position 0 starts at or after
    // value.
    // The exact exit position follows these steps, as callers see it.  Host code on the
    // public starts array.
    inline void upper_bound_range(const S value, size_t& left, const size_t right) const noexcept {
        size_t hi = right;
        for (size_t step = 1; left > 0 && starts[left] > value; step *= 2) {
            hi = left;
            left = (step <= left) ? left - step : 0;
        }
        for (size_t len = hi - left; len > 1;) {
            const size_t keep = len / 2;
            if (starts[left + keep] < value) left += len - keep;
            len = keep;
        }
        if (left == 0 && starts[0] >= value) left = SIZE_MAX;
    }

    // superintervals.hpp:551-579 (descending data order).
    void search_values(const S start, const S end, std::vector<T>& found) const {
        for (size_t i : positions(start, end, SI_ORDER_DESC)) found.push_back(data[i]);
    }
    // superintervals.hpp:588-615: same result as search_values.
    void search_values_large(const S start, const S end, std::vector<T>& found) const {
        search_values(start, end, found);
    }
    // superintervals.hpp:623-649, 651-825, 834-857: all three return the overlap count.
    size_t count_linear(const S start, const S end) const noexcept { return count(start, end); }
    size_t count(const S start, const S end) const noexcept {
        if (!ready()) return 0;
        uint32_t c = 0;
        if (siCountHost(dev_, &start, &end, 1, &c) != SI_OK) return 0;
        return c;
    }
    size_t count_large(const S start, const S end) const noexcept { return count(start, end); }

    // superintervals.hpp:865-871: only the last interval starting at or before end is
    // tested, as in the reference.
    bool has_overlaps(const S start, const S end) const noexcept {
        if (!ready()) return false;
        uint8_t r = 0;
        if (siHasOverlapsHost(dev_, &start, &end, 1, &r) != SI_OK) return false;
        return r != 0;
    }

    // superintervals.hpp:879-905: leading contiguous run ascending, then descending.
    void search_idxs(const S start, const S end, std::vector<size_t>& found) const {
        std::vector<size_t> p = positions(start, end, SI_ORDER_VECTOR);
        found.insert(found.end(), p.begin(), p.end());
    }
    // superintervals.hpp:913-938
    void search_keys(const S start, const S end, std::vector<std::pair<S, S>>& found) const {
        for (size_t i : positions(start, end, SI_ORDER_DESC)) found.emplace_back(starts[i], ends[i]);
    }
    // superintervals.hpp:946-971
    void search_items(const S start, const S end, std::vector<Interval<S, T>>& found) const {
        for (size_t i : positions(start, end, SI_ORDER_DESC)) found.emplace_back(starts[i], ends[i], data[i]);
    }
    // superintervals.hpp:979-1006: adds to cov_result (count, sum of min(e,end)-max(s,start)).
    void coverage(const S start, const S end, std::pair<size_t, S>& cov_result) const {
        std::vector<size_t> p = positions(start, end, SI_ORDER_DESC);
        for (size_t i : p) {
            ++cov_result.first;
            cov_result.second += (ends[i] < end ? ends[i] : end) - (starts[i] > start ? starts[i] : start);
        }
    }
    // superintervals.hpp:1013-1035
    void search_point(const S point, std::vector<T>& found) const { search_values(point, point, found); }

    // ---- batch queries (shaped after intervalmap.pyx:363-494) ---------------------------
    // counts[q] = count(qs[q], qe[q]) for q < n.  Returns false on a device failure.
    bool count_batch(const S* qs, const S* qe, size_t n, uint32_t* counts) const {
        if (n == 0) return true;
        if (starts.empty()) { for (size_t q = 0; q < n; ++q) counts[q] = 0; return true; }
        return ready() && siCountHost(dev_, qs, qe, n, counts) == SI_OK;
    }
    // CSR of search_idxs (vector order) or of the iterator order (descending).
    bool search_idxs_batch(const S* qs, const S* qe, size_t n, std::vector<uint64_t>& offsets,
                           std::vector<uint32_t>& idx, bool vector_order = true) const {
        return batch(qs, qe, n, vector_order ? SI_ORDER_VECTOR : SI_ORDER_DESC, offsets, idx);
    }
    // CSR of search_values: values[offsets[q] .. offsets[q+1]) in descending position order.
    bool search_values_batch(const S* qs, const S* qe, size_t n, std::vector<uint64_t>& offsets,
                             std::vector<T>& values) const {
        std::vector<uint32_t> idx;
        if (!batch(qs, qe, n, SI_ORDER_DESC, offsets, idx)) return false;
        values.clear();
        values.reserve(idx.size());
        for (uint32_t i : idx) values.push_back(data[i]);
        return true;
    }

    // ---- set operations (superintervals.hpp:1158-1211) ---------------------------------
    // Both run one batched search of this set's intervals (in their current order) in
    // other's index on the GPU, then assemble the pieces with this set's / other's data on
    // the host (T is any host type).  other must be built; the results are unindexed.
    //
    // superintervals.hpp:1158-1174: every overlapping piece, combine(this data, other data).
    template <typename Combine>
    IntervalMap<S, T> intersection(const IntervalMap<S, T>& other, Combine combine) const {
        IntervalMap<S, T> out;
        std::vector<uint64_t> off;
        std::vector<uint32_t> idx;
        if (!other.search_idxs_batch(starts.data(), ends.data(), starts.size(), off, idx, true)) return out;
        for (size_t k = 0; k < starts.size(); ++k) {
            for (uint64_t j = off[k]; j < off[k + 1]; ++j) {
                const size_t h = idx[j];
                const S s = starts[k] > other.starts[h] ? starts[k] : other.starts[h];
                const S e = ends[k] < other.ends[h] ? ends[k] : other.ends[h];
                if (s <= e) out.add(s, e, combine(data[k], other.data[h]));
            }
        }
        return out;
    }
    IntervalMap<S, T> intersection(const IntervalMap<S, T>& other) const {
        return intersection(other, [](const T& a, const T&) { return a; });
    }
    // superintervals.hpp:1188-1211: the parts of each interval that other does not cover.
    // other's hits in descending position order, read backwards, are its keys ascending by
    // (start, end desc); the sweep's pieces do not depend on the order of equal starts.  A
    // hit ending at the type's maximum covers the rest (the reference's ce + 1 overflows).
    IntervalMap<S, T> difference(const IntervalMap<S, T>& other) const {
        IntervalMap<S, T> out;
        std::vector<uint64_t> off;
        std::vector<uint32_t> idx;
        if (!other.search_idxs_batch(starts.data(), ends.data(), starts.size(), off, idx, false)) return out;
        for (size_t k = 0; k < starts.size(); ++k) {
            S cursor = starts[k];
            bool past = false;
            for (uint64_t j = off[k + 1]; j-- > off[k];) {
                const size_t h = idx[j];
                const S cs = other.starts[h] > starts[k] ? other.starts[h] : starts[k];
                const S ce = other.ends[h] < ends[k] ? other.ends[h] : ends[k];
                if (cs > cursor) out.add(cursor, cs - 1, data[k]);
                if (ce == std::numeric_limits<S>::max()) {
                    past = true;
                    break;
                }
                if (ce + 1 > cursor) cursor = ce + 1;
            }
            if (!past && cursor <= ends[k]) out.add(cursor, ends[k], data[k]);
        }
        return out;
    }

    // Device-resident index (valid after build()); for the device entry points.
    const siIndex* device_index() const { return dev_; }
    static std::string last_error() { return siGetLastError(); }

  protected:
    siIndex* dev_ = nullptr;
    bool built_ = false;

    bool ready() const {
        return !starts.empty() && dev_ && built_ && siIndexSize(dev_) == starts.size();
    }

    void rebuild_device() {
        dev_ = siIndexCreate((int)sizeof(S));
        if (dev_ && siIndexBuild(dev_, starts.data(), ends.data(), nullptr, starts.size(), 1, 1, 0,
                                 nullptr) == SI_OK)
            built_ = true;
    }

    std::vector<size_t> positions(S start, S end, int order) const {
        std::vector<size_t> out;
        if (!ready()) return out;
        uint64_t off[2] = {0, 0}, total = 0;
        std::vector<uint32_t> idx(64);
        int rc = siSearchHost(dev_, &start, &end, 1, order, off, idx.data(), nullptr, idx.size(), &total);
        if (rc == SI_ERR_CAPACITY) {
            idx.resize(total);
            rc = siSearchHost(dev_, &start, &end, 1, order, off, idx.data(), nullptr, idx.size(), &total);
        }
        if (rc != SI_OK) return out;
        out.assign(idx.begin(), idx.begin() + (size_t)total);
        return out;
    }

    bool batch(const S* qs, const S* qe, size_t n, int order, std::vector<uint64_t>& offsets,
               std::vector<uint32_t>& idx) const {
        offsets.assign(n + 1, 0);
        idx.clear();
        if (n == 0 || starts.empty()) return true;
        if (!ready()) return false;
        uint64_t total = 0;
        int rc = siSearchHost(dev_, qs, qe, n, order, offsets.data(), nullptr, nullptr, 0, &total);
        if (rc != SI_OK && rc != SI_ERR_CAPACITY) return false;
        idx.resize(total);
        if (total == 0) return true;
        return siSearchHost(dev_, qs, qe, n, order, offsets.data(), idx.data(), nullptr, idx.size(),
                            &total) == SI_OK;
    }
};

// superintervals.hpp:1457-1535: same index and results as IntervalMap.
template <typename S, typename T>
class IntervalMapEytz : public IntervalMap<S, T> {};

}  // namespace si
