// si_internal.h -- the device index object shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "si_common.h"
#include "si_kernels.h"

using namespace sik;

// Search tree over a sorted key array: level l + 1 holds every 64th key of level l
// (level 0 = the keys themselves) up to a top level of at most 4095 keys, which is also
// kept in Eytzinger (breadth-first) order with max(S) padding to 2^top_h - 1 keys.
struct KeyTree {
    void* d_lv = nullptr;            // levels 1 .. nlv, contiguous
    void* d_top = nullptr;
    int nlv = 0;                     // levels above the keys
    size_t lv_off[MAX_LEVELS] = {0}; // element offsets of levels 1 .. nlv in d_lv
    uint32_t lv_size[MAX_LEVELS] = {0};
    uint32_t top_size = 0, top_h = 0;

    void release() {
        hipFree(d_lv); hipFree(d_top);
        d_lv = d_top = nullptr;
        nlv = 0;
        top_size = top_h = 0;
    }
};

// Scratch of the single-query calls (the host-buffer forms with one query): a pinned host
// buffer mapped into the device, which the one-wave kernel writes directly (header + up to
// cap results), and a stream.  Made on the first such call, kept until the index is
// destroyed; the last query's results stay in it, so a repeated identical call (a caller
// that sized its buffer from the first) is answered without a launch.
struct OneQuery {
    std::mutex mu;
    void* h = nullptr;          // pinned host memory: [0] total, [1] spare, then cap uint32 results
    void* d = nullptr;          // its device address
    hipStream_t st = nullptr;
    uint64_t cap = 0;
    bool last_valid = false;
    int last_mode = -1;
    int64_t last_s = 0, last_e = 0;
    uint64_t last_total = 0;
    void release() {
        if (h) (void)hipHostFree(h);
        if (st) (void)hipStreamDestroy(st);
        h = d = nullptr;
        st = nullptr;
        cap = 0;
        last_valid = false;
    }
};

// The long list of an index (LongView): device arrays and the max tree over its ends.
struct LongList {
    int64_t len = 0;              // threshold; 0 = no list
    uint32_t n = 0;
    uint32_t* d_pos = nullptr;    // n, padded to 64 entries
    void* d_ends = nullptr;       // n, padded with the minimum
    uint32_t* d_ends32 = nullptr; // int64 indexes with a 32-bit copy of the ends
    uint32_t* d_lrank = nullptr;  // index size + 1
    void* d_tree = nullptr;       // levels 1.. of the max tree over d_ends
    int nlev = 0;
    size_t lvl_off[MAX_LEVELS] = {0};
    uint32_t lvl_size[MAX_LEVELS] = {0};
    void release() {
        hipFree(d_pos); hipFree(d_ends); hipFree(d_ends32); hipFree(d_lrank); hipFree(d_tree);
        d_pos = d_ends32 = d_lrank = nullptr;
        d_ends = d_tree = nullptr;
        len = 0;
        n = 0;
        nlev = 0;
    }
};

// ============================================================================
// siIndex
// ============================================================================
struct siIndex {
    int cb = 4;
    size_t n = 0;
    bool built = false;
    int device = 0;
    void* d_starts = nullptr;
    void* d_ends = nullptr;
    uint32_t* d_branch = nullptr;
    uint32_t* d_exit = nullptr;      // line exit targets (IndexView::exitp)
    uint32_t* d_perm = nullptr;
    int32_t* d_values = nullptr;
    void* d_tree = nullptr;   // levels 1.. of the max tree, contiguous
    void* d_ends_sorted = nullptr;   // ends in ascending order (rank count)
    uint32_t* d_ends32 = nullptr;    // int64 only, ends - ends32_base when they span < 2^32
    int64_t ends32_base = 0;
    // intervals with start > end: the rank form of count and the guaranteed-hit run
    // [first start >= s, upper_bound] both assume start <= end, so queries on an index
    // holding any such interval take the walk
    size_t n_inverted = 0;
    // the learned position model trained on this build (siLearnedTrain attaches it,
    // siLearnedDestroy and the next build detach it); generation counts builds
    siLearned* learned = nullptr;
    uint64_t generation = 0;
    std::vector<siLearned*> models;   // every model created on this index (siIndexDestroy unlinks them)
    OneQuery one;                     // single-query scratch (see OneQuery)
    LongList lng;                     // intervals longer than a threshold (LongView)
    KeyTree ub;   // search tree over the starts (upper_bound)
    KeyTree eb;   // search tree over the sorted ends (rank count)
    int nlev = 0;
    size_t lvl_off[MAX_LEVELS] = {0};   // element offsets into d_tree (level >= 1)
    uint32_t lvl_size[MAX_LEVELS] = {0};
    double last_build_ms = 0.0;

    void release() {
        hipFree(d_starts); hipFree(d_ends); hipFree(d_branch); hipFree(d_exit); hipFree(d_perm);
        hipFree(d_values); hipFree(d_tree); hipFree(d_ends_sorted); hipFree(d_ends32);
        d_starts = d_ends = d_tree = d_ends_sorted = nullptr;
        d_ends32 = nullptr;
        ends32_base = 0;
        n_inverted = 0;
        learned = nullptr;
        one.last_valid = false;   // the cached last query belongs to the old build
        lng.release();
        ++generation;
        ub.release();
        eb.release();
        d_branch = d_exit = d_perm = nullptr;
        d_values = nullptr;
        n = 0;
        nlev = 0;
        built = false;
    }

    template <typename S>
    IndexView<S> view() const {
        IndexView<S> v;
        v.starts = (const S*)d_starts;
        v.ends = (const S*)d_ends;
        v.branch = d_branch;
        v.exitp = d_exit;
        v.ends32 = d_ends32;
        v.ends32_base = ends32_base;
        v.n = (uint32_t)n;
        v.npad = (uint32_t)((n + 63) & ~(size_t)63);
        return v;
    }

    template <typename S>
    LongView<S> long_view() const {
        LongView<S> v;
        v.pos = lng.d_pos;
        v.ends = (const S*)lng.d_ends;
        v.ends32 = lng.d_ends32;
        v.ends32_base = ends32_base;
        v.lrank = lng.d_lrank;
        v.len = lng.len;
        v.n = lng.n;
        v.td.nlev = lng.nlev;
        v.td.lvl[0] = (const S*)lng.d_ends;
        v.td.size[0] = lng.n;
        for (int l = 1; l < lng.nlev; ++l) {
            v.td.lvl[l] = (const S*)lng.d_tree + lng.lvl_off[l];
            v.td.size[l] = lng.lvl_size[l];
        }
        for (int l = (lng.nlev > 0 ? lng.nlev : 1); l < MAX_LEVELS; ++l) { v.td.lvl[l] = nullptr; v.td.size[l] = 0; }
        return v;
    }

    template <typename S>
    TreeDesc<S> tree() const {
        TreeDesc<S> t;
        t.nlev = nlev;
        t.lvl[0] = (const S*)d_ends;
        t.size[0] = (uint32_t)n;
        for (int l = 1; l < nlev; ++l) {
            t.lvl[l] = (const S*)d_tree + lvl_off[l];
            t.size[l] = lvl_size[l];
        }
        for (int l = nlev; l < MAX_LEVELS; ++l) { t.lvl[l] = nullptr; t.size[l] = 0; }
        return t;
    }
};

inline int check_ix(const siIndex* ix) {
    if (!ix) return si_set_error(SI_ERR_ARG, "null index");
    if (!ix->built) return si_set_error(SI_ERR_NOT_BUILT, "index has not been built");
    return SI_OK;
}
