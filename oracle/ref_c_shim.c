/* ref_c_shim.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiles the reference's C implementation (src/c_superintervals.h of
 * kcleal/superintervals v1.0.0) in place from the reference checkout, with its
 * symbols hidden, and re-exports it under refc_* names.  Used to write golden
 * vectors for the C-ABI path (int32 indices, strict descending order, qsort
 * payload order) and to cross-check the restatement.
 */
#include "c_superintervals.h"

#define REFC_API __attribute__((visibility("default")))

REFC_API void* refc_build(size_t n, const int32_t* st, const int32_t* en, const int32_t* d) {
    cSuperIntervals* si = createSuperIntervals();
    for (size_t i = 0; i < n; ++i) addInterval(si, st[i], en[i], d ? d[i] : (int32_t)i);
    indexSuperIntervals(si);
    return si;
}
REFC_API void refc_free(void* p) { destroySuperIntervals((cSuperIntervals*)p); }
REFC_API void refc_arrays(void* p, int32_t* st, int32_t* en, int32_t* d, int64_t* br) {
    cSuperIntervals* si = (cSuperIntervals*)p;
    for (size_t i = 0; i < si->size; ++i) {
        if (st) st[i] = si->starts[i];
        if (en) en[i] = si->ends[i];
        if (d) d[i] = si->data[i];
        if (br) br[i] = (si->branch[i] == SI_NONE) ? -1 : (int64_t)si->branch[i];
    }
}
REFC_API int64_t refc_upper_bound(void* p, int32_t x) {
    size_t i = upperBound((cSuperIntervals*)p, x);
    return i == SI_NONE ? -1 : (int64_t)i;
}
REFC_API uint64_t refc_count(void* p, int32_t s, int32_t e) {
    return countOverlaps((cSuperIntervals*)p, s, e);
}
REFC_API int refc_any(void* p, int32_t s, int32_t e) {
    return anyOverlaps((cSuperIntervals*)p, s, e) ? 1 : 0;
}
/* which: 0 = searchIdxs, 1 = searchValues.  Returns the hit count. */
REFC_API size_t refc_search(void* p, int32_t s, int32_t e, int which, int32_t* out, size_t cap) {
    cIndexResult r = createIndexResult();
    if (which == 0) searchIdxs((cSuperIntervals*)p, s, e, &r);
    else searchValues((cSuperIntervals*)p, s, e, &r);
    for (size_t k = 0; k < r.size && out && k < cap; ++k) out[k] = r.data[k];
    size_t h = r.size;
    destroyIndexResult(&r);
    return h;
}
REFC_API void refc_coverage(void* p, int32_t s, int32_t e, uint64_t* c, int32_t* cov) {
    size_t cc = 0;
    coverage((cSuperIntervals*)p, s, e, &cc, cov);
    *c = cc;
}

/* An unindexed handle (insertion order), for the first operand of a set operation. */
REFC_API void* refc_raw(size_t n, const int32_t* st, const int32_t* en, const int32_t* d) {
    cSuperIntervals* si = createSuperIntervals();
    for (size_t i = 0; i < n; ++i) addInterval(si, st[i], en[i], d ? d[i] : (int32_t)i);
    return si;
}
REFC_API size_t refc_size(void* p) { return ((cSuperIntervals*)p)->size; }
/* Arrays in their current order (no branch: set-operation results are unindexed). */
REFC_API void refc_raw_arrays(void* p, int32_t* st, int32_t* en, int32_t* d) {
    cSuperIntervals* si = (cSuperIntervals*)p;
    for (size_t i = 0; i < si->size; ++i) {
        st[i] = si->starts[i];
        en[i] = si->ends[i];
        d[i] = si->data[i];
    }
}
/* intersection with the default combine (keep the first operand's data), difference. */
REFC_API void* refc_intersection(void* a, void* b) {
    return intersection((const cSuperIntervals*)a, (cSuperIntervals*)b, NULL);
}
REFC_API void* refc_difference(void* a, void* b) {
    return difference((const cSuperIntervals*)a, (cSuperIntervals*)b);
}
