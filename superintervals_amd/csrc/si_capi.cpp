// si_capi.cpp -- the reference's C-ABI (src/c_superintervals.h) on the device index.
//
// Host storage follows the reference exactly (malloc/realloc growth, the
// startSorted/endSorted flags of addInterval, capacity doubling of result
// buffers).  indexSuperIntervals builds the superset index on the GPU and
// mirrors the sorted arrays and the branch array back into the public fields;
// every query runs on the GPU through si_device.h.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/c_superintervals.h"
#include "si_common.h"

namespace {

// Grows an int32 result buffer by doubling from 16 (c_superintervals.h:585-588) until
// it holds at least need entries.
bool ensure_i32(cIndexResult* r, size_t need) {
    if (need <= r->capacity) return true;
    size_t cap = r->capacity ? r->capacity : 16;
    while (cap < need) cap *= 2;
    int32_t* p = (int32_t*)realloc(r->data, cap * sizeof(int32_t));
    if (!p) return false;
    r->data = p;
    r->capacity = cap;
    return true;
}

template <typename R, typename E>
bool ensure_any(R* r, size_t need) {
    if (need <= r->capacity) return true;
    size_t cap = r->capacity ? r->capacity : 16;
    while (cap < need) cap *= 2;
    E* p = (E*)realloc(r->data, cap * sizeof(E));
    if (!p) return false;
    r->data = p;
    r->capacity = cap;
    return true;
}

// The index must exist and be built for queries; the reference leaves querying an
// unindexed handle undefined, here it reports an error and returns no result.  addInterval
// and clearSuperIntervals drop the device index, so a handle changed since its last
// indexSuperIntervals has none.
bool ready(const cSuperIntervals* si) {
    if (!si) { si_set_error(SI_ERR_ARG, "null handle"); return false; }
    if (si->size == 0) return false;
    if (!si->dev || siIndexSize(si->dev) != si->size) {
        si_set_error(SI_ERR_NOT_BUILT, "indexSuperIntervals() has not been called since the last add");
        return false;
    }
    return true;
}

// One query's strict-descending positions (values == false) or payloads into out,
// appended at out->size.
void search_one(cSuperIntervals* si, int32_t s, int32_t e, bool values, cIndexResult* out) {
    if (!ready(si)) return;
    uint64_t off[2];
    uint64_t total = 0;
    size_t room = out->capacity - out->size;
    int32_t* dst = out->data ? out->data + out->size : nullptr;
    int rc = siSearchHost(si->dev, &s, &e, 1, SI_ORDER_DESC, off, values ? nullptr : (uint32_t*)dst,
                          values ? dst : nullptr, dst ? room : 0, &total);
    if (rc == SI_ERR_CAPACITY) {
        if (!ensure_i32(out, out->size + total)) { si_set_error(SI_ERR_NOMEM, "result realloc"); return; }
        dst = out->data + out->size;
        rc = siSearchHost(si->dev, &s, &e, 1, SI_ORDER_DESC, off, values ? nullptr : (uint32_t*)dst,
                          values ? dst : nullptr, out->capacity - out->size, &total);
    }
    if (rc == SI_OK) out->size += total;
}

void sort_and_mirror(cSuperIntervals* si, bool with_branch) {
    if (si->size == 0) return;
    if (!si->dev) {
        si->dev = siIndexCreate(4);
        if (!si->dev) return;
    }
    int rc = siIndexBuild(si->dev, si->starts, si->ends, si->data, si->size, si->startSorted,
                          si->endSorted, 0, nullptr);
    if (rc != SI_OK) return;
    size_t* br = nullptr;
    if (with_branch) {
        br = (size_t*)realloc(si->branch, si->size * sizeof(size_t));
        if (!br) { si_set_error(SI_ERR_NOMEM, "branch realloc"); return; }
        si->branch = br;
    }
    static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t must be 64-bit");
    rc = siIndexDownload(si->dev, si->starts, si->ends, si->data, nullptr, (uint64_t*)br);
    if (rc != SI_OK) return;
    si->startSorted = true;
    si->endSorted = true;
}

}  // namespace

// ---- Lifecycle ---------------------------------------------------------------
SI_EXPORT cSuperIntervals* createSuperIntervals(void) {
    cSuperIntervals* si = (cSuperIntervals*)malloc(sizeof(cSuperIntervals));
    if (!si) return nullptr;
    si->starts = nullptr;
    si->ends = nullptr;
    si->data = nullptr;
    si->branch = nullptr;
    si->size = 0;
    si->capacity = 0;
    si->idx = 0;
    si->startSorted = true;
    si->endSorted = true;
    si->dev = nullptr;
    return si;
}

SI_EXPORT void destroySuperIntervals(cSuperIntervals* si) {
    if (!si) return;
    free(si->starts);
    free(si->ends);
    free(si->data);
    free(si->branch);
    siIndexDestroy(si->dev);
    free(si);
}

// The device index no longer matches the host arrays once they change.
static void drop_device_index(cSuperIntervals* si) {
    if (si->dev) {
        siIndexDestroy(si->dev);
        si->dev = nullptr;
    }
}

SI_EXPORT void clearSuperIntervals(cSuperIntervals* si) {
    drop_device_index(si);
    si->size = 0;
    si->idx = 0;
    si->startSorted = true;
    si->endSorted = true;
}

SI_EXPORT void reserveSuperIntervals(cSuperIntervals* si, size_t n) {
    if (n > si->capacity) {
        si->capacity = n;
        si->starts = (int32_t*)realloc(si->starts, n * sizeof(int32_t));
        si->ends = (int32_t*)realloc(si->ends, n * sizeof(int32_t));
        si->data = (int32_t*)realloc(si->data, n * sizeof(int32_t));
    }
}

SI_EXPORT void addInterval(cSuperIntervals* si, int32_t start, int32_t end, int32_t value) {
    drop_device_index(si);
    if (si->size >= si->capacity) reserveSuperIntervals(si, si->capacity == 0 ? 1 : si->capacity * 2);
    // sortedness flags, c_superintervals.h:408-413
    if (si->startSorted && si->size > 0) {
        si->startSorted = !(start < si->starts[si->size - 1]);
        if (si->startSorted && start == si->starts[si->size - 1] && end > si->ends[si->size - 1])
            si->endSorted = false;
    }
    si->starts[si->size] = start;
    si->ends[si->size] = end;
    si->data[si->size] = value;
    si->size++;
}

SI_EXPORT size_t sizeSuperIntervals(const cSuperIntervals* si) { return si->size; }

// ---- Indexing ------------------------------------------------------------------
SI_EXPORT void sortIntervals(cSuperIntervals* si) { sort_and_mirror(si, false); }

SI_EXPORT void indexSuperIntervals(cSuperIntervals* si) {
    if (si->size == 0) return;
    sort_and_mirror(si, true);
    si->idx = 0;
}

// ---- Element access --------------------------------------------------------------
SI_EXPORT bool intervalAt(const cSuperIntervals* si, size_t index, Interval* out) {
    if (index >= si->size) return false;
    out->start = si->starts[index];
    out->end = si->ends[index];
    out->data = si->data[index];
    return true;
}
SI_EXPORT int32_t startAt(const cSuperIntervals* si, size_t index) { return si->starts[index]; }
SI_EXPORT int32_t endAt(const cSuperIntervals* si, size_t index) { return si->ends[index]; }
SI_EXPORT int32_t dataAt(const cSuperIntervals* si, size_t index) { return si->data[index]; }

// ---- Cursor ------------------------------------------------------------------------
SI_EXPORT size_t upperBound(cSuperIntervals* si, int32_t value) {
    si->idx = SI_NONE;
    if (!ready(si)) return SI_NONE;
    uint32_t pos = SI_INDEX_NONE;
    if (siUpperBoundHost(si->dev, &value, 1, &pos) != SI_OK) return SI_NONE;
    si->idx = (pos == SI_INDEX_NONE) ? SI_NONE : (size_t)pos;
    return si->idx;
}

// ---- Queries ------------------------------------------------------------------------
SI_EXPORT bool anyOverlaps(cSuperIntervals* si, int32_t start, int32_t end) {
    if (!ready(si)) return false;
    uint8_t r = 0;
    if (siHasOverlapsHost(si->dev, &start, &end, 1, &r) != SI_OK) return false;
    return r != 0;
}

SI_EXPORT size_t countOverlaps(cSuperIntervals* si, int32_t start, int32_t end) {
    if (!ready(si)) return 0;
    uint32_t c = 0;
    if (siCountHost(si->dev, &start, &end, 1, &c) != SI_OK) return 0;
    return c;
}

SI_EXPORT void searchValues(cSuperIntervals* si, int32_t start, int32_t end, cIndexResult* found) {
    search_one(si, start, end, true, found);
}

SI_EXPORT void searchIdxs(cSuperIntervals* si, int32_t start, int32_t end, cIndexResult* found) {
    search_one(si, start, end, false, found);
}

SI_EXPORT void searchKeys(cSuperIntervals* si, int32_t start, int32_t end, cKeyResult* found) {
    cIndexResult r = {nullptr, 0, 0};
    search_one(si, start, end, false, &r);
    if (ensure_any<cKeyResult, KeyPair>(found, found->size + r.size)) {
        for (size_t k = 0; k < r.size; ++k) {
            found->data[found->size].start = si->starts[r.data[k]];
            found->data[found->size].end = si->ends[r.data[k]];
            found->size++;
        }
    }
    free(r.data);
}

SI_EXPORT void searchItems(cSuperIntervals* si, int32_t start, int32_t end, cItemResult* found) {
    cIndexResult r = {nullptr, 0, 0};
    search_one(si, start, end, false, &r);
    if (ensure_any<cItemResult, Interval>(found, found->size + r.size)) {
        for (size_t k = 0; k < r.size; ++k) {
            Interval& it = found->data[found->size++];
            it.start = si->starts[r.data[k]];
            it.end = si->ends[r.data[k]];
            it.data = si->data[r.data[k]];
        }
    }
    free(r.data);
}

SI_EXPORT void searchPoint(cSuperIntervals* si, int32_t point, cIndexResult* found) {
    searchValues(si, point, point, found);
}

SI_EXPORT void coverage(cSuperIntervals* si, int32_t start, int32_t end, size_t* count_out,
                        int32_t* coverage_out) {
    *count_out = 0;
    *coverage_out = 0;
    if (!ready(si)) return;
    uint32_t c = 0;
    int32_t v = 0;
    if (siCoverageHost(si->dev, &start, &end, 1, &c, &v) != SI_OK) return;
    *count_out = c;
    *coverage_out = v;
}

SI_EXPORT void findOverlaps(cSuperIntervals* si, int32_t start, int32_t end, int32_t* found,
                            size_t* found_size) {
    *found_size = 0;
    if (!ready(si)) return;
    uint64_t off[2], total = 0;
    // The caller guarantees room for every result (c_superintervals.h:236-240).
    if (siSearchHost(si->dev, &start, &end, 1, SI_ORDER_DESC, off, nullptr, found, (size_t)-1, &total) == SI_OK)
        *found_size = total;
}

// ---- Result helpers -------------------------------------------------------------------
SI_EXPORT cIndexResult createIndexResult(void) { return cIndexResult{nullptr, 0, 0}; }
SI_EXPORT void clearIndexResult(cIndexResult* r) { r->size = 0; }
SI_EXPORT void destroyIndexResult(cIndexResult* r) { free(r->data); r->data = nullptr; r->size = 0; r->capacity = 0; }
SI_EXPORT cKeyResult createKeyResult(void) { return cKeyResult{nullptr, 0, 0}; }
SI_EXPORT void clearKeyResult(cKeyResult* r) { r->size = 0; }
SI_EXPORT void destroyKeyResult(cKeyResult* r) { free(r->data); r->data = nullptr; r->size = 0; r->capacity = 0; }
SI_EXPORT cItemResult createItemResult(void) { return cItemResult{nullptr, 0, 0}; }
SI_EXPORT void clearItemResult(cItemResult* r) { r->size = 0; }
SI_EXPORT void destroyItemResult(cItemResult* r) { free(r->data); r->data = nullptr; r->size = 0; r->capacity = 0; }

// ---- Batch extensions ---------------------------------------------------------------------
SI_EXPORT int siCountBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                           uint32_t* counts) {
    if (!si) return si_set_error(SI_ERR_ARG, "null handle");
    if (si->size == 0) {
        if (n) memset(counts, 0, n * sizeof(uint32_t));
        return SI_OK;
    }
    if (!ready(si)) return SI_ERR_NOT_BUILT;
    return siCountHost(si->dev, qs, qe, n, counts);
}

static int batch_search(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                        uint64_t* offsets, int32_t* out, bool values, size_t cap, uint64_t* total) {
    if (!si || !offsets) return si_set_error(SI_ERR_ARG, "null argument");
    if (si->size == 0) {
        memset(offsets, 0, (n + 1) * sizeof(uint64_t));
        if (total) *total = 0;
        return SI_OK;
    }
    if (!ready(si)) return SI_ERR_NOT_BUILT;
    return siSearchHost(si->dev, qs, qe, n, SI_ORDER_DESC, offsets, values ? nullptr : (uint32_t*)out,
                        values ? out : nullptr, cap, total);
}

SI_EXPORT int siSearchIdxsBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                                uint64_t* offsets, int32_t* idx, size_t cap, uint64_t* total) {
    return batch_search(si, qs, qe, n, offsets, idx, false, cap, total);
}

SI_EXPORT int siSearchValuesBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                                  uint64_t* offsets, int32_t* values, size_t cap, uint64_t* total) {
    return batch_search(si, qs, qe, n, offsets, values, true, cap, total);
}

// ---- Set operations (c_superintervals.h:286, :292) -------------------------------------
static int32_t keep_first(int32_t a, int32_t) { return a; }

// c_superintervals.h:921-940: for each interval k of si in its current order and each hit h
// of searchIdxs(other, ...) (descending), the piece [max(starts), min(ends)] when it is not
// empty, with combine(si->data[k], other->data[h]).  The search and the pieces run on the
// device; combine is the caller's host function.
SI_EXPORT cSuperIntervals* intersection(const cSuperIntervals* si, cSuperIntervals* other, cCombineFn combine) {
    if (!combine) combine = keep_first;
    cSuperIntervals* out = createSuperIntervals();
    if (!out || !si || si->size == 0 || !other || other->size == 0 || !ready(other)) return out;
    std::vector<uint64_t> off;
    std::vector<uint32_t> pos;
    std::vector<int32_t> ps, pe;
    if (si_intersect_host(other->dev, si->starts, si->ends, si->size, off, pos, ps, pe) != SI_OK) return out;
    for (size_t k = 0; k < si->size; ++k) {
        for (uint64_t j = off[k]; j < off[k + 1]; ++j) {
            if (ps[j] <= pe[j]) addInterval(out, ps[j], pe[j], combine(si->data[k], other->data[pos[j]]));
        }
    }
    return out;
}

// c_superintervals.h:950-975: the parts of each interval of si (current order) that no
// interval of other covers, with its data.  An empty other covers nothing.
SI_EXPORT cSuperIntervals* difference(const cSuperIntervals* si, cSuperIntervals* other) {
    cSuperIntervals* out = createSuperIntervals();
    if (!out || !si || si->size == 0 || !other) return out;
    if (other->size == 0) {
        for (size_t k = 0; k < si->size; ++k)
            if (si->starts[k] <= si->ends[k]) addInterval(out, si->starts[k], si->ends[k], si->data[k]);
        return out;
    }
    if (!ready(other)) return out;
    std::vector<int32_t> ps, pe, pd;
    if (si_difference_host(other->dev, si->starts, si->ends, si->data, si->size, ps, pe, pd) != SI_OK) return out;
    for (size_t j = 0; j < ps.size(); ++j) addInterval(out, ps[j], pe[j], pd[j]);
    return out;
}
