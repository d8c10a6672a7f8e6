// si_query.hip -- batched overlap queries on the device index.
//
// Replaces superintervals.hpp:501-516 (upper_bound), :651-825 (count), :865-871
// (has_overlaps), :879-905 (search_idxs), :551-579 (search_values) and
// c_superintervals.h:537-756 (upperBound, anyOverlaps, searchValues, searchIdxs,
// countOverlaps) with batch kernels.
//
// A batch runs in three stages.
//  1. Plan: the queries are sorted by end (stable LSD radix sort, payload = query
//     number; optionally only down to a coarser digit) so the lanes of a wave walk
//     neighbouring parts of the index and share cache lines; each sorted query gets
//     its upper_bound position.
//  2. Count: a persistent kernel walks every query.  A workgroup takes a tile of
//     TILE consecutive sorted queries from a global counter, stages the window of
//     ends[] and exitp[] just below the tile's largest upper_bound in LDS (26-27 KB),
//     and walks from LDS; the few reads below the window go to HBM.  One lane owns
//     one query; when it finishes, the lane takes the next query from a pool of 64
//     the wave grabs from the tile's LDS counter, so lanes stay busy however uneven
//     the hit counts are.  counts[q] / offsets are written in input order.
//  3. Fill: the same kernel writes each query's positions at its CSR offset.
//
// The walk is the reference's backward superset walk taken one aligned 64-byte line of
// ends at a time.  In the line [b0, b0 + W) holding the current position cur, every
// j in [b0, cur] with ends[j] >= s is a hit (the hit set is {j <= upper_bound :
// ends[j] >= s}).  The walk leaves the line to b0 - 1 when b0 itself is a hit.
// Otherwise the reference walk would enter the run of misses [b0, x] at its top x
// (just below the lowest hit, or cur when there is none) and follow branch[] from
// there; that chain's first element below b0 is the nearest j < b0 with
// ends[j] >= max(ends[b0..x]), which is branch[a] for the leftmost maximum a of
// ends[b0..x].  That target depends on x only, so the build stores it per position
// (exitp[x], see si_index.hip k_exit).  One line load and at most one exitp load replace
// the reference's per-element steps while visiting the same hits in the same order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "si_internal.h"
#include "si_prims.h"

namespace {

constexpr int Q_THREADS = 256;
constexpr uint32_t POOL = 64;

enum WalkMode { W_COUNT = 0, W_DESC = 1, W_VECTOR = 2, W_VALUES = 3 };

// ============================================================================
// Plan
// ============================================================================
struct HeavyEntry;

template <typename S>
struct PlanView {
    uint32_t* counter;   // [0] tiles taken, [1] heavy entries written, [3] set by a fill pass
                         // when a query's hits differ from its offsets, [4] 1 when the plan
                         // holds lower_bound positions (lb), [5] 1 when it holds the counts
                         // (cnt), [8 + c] entries taken from chunk c of the heavy list
                         // (HEAVY_CHUNKS chunks)
    uint32_t* q;         // input query number of sorted position k
    S* s;                // query start of sorted position k
    S* e;                // query end of sorted position k
    uint32_t* idx;       // upper_bound of sorted position k
    uint32_t* lb;        // first position whose start is >= the query start (lower_bound)
    uint32_t* cnt;       // the count of sorted position k (rank form), when counter[5] is set
    uint32_t* set;       // plans over a set table: the set of sorted position k
    HeavyEntry* heavy;   // parked walks (see HeavyEntry)
};

// Byte layout of a plan of nq queries (coordinate width cb):
//   [256 B counters][q: 4 nq][s: cb nq][e: cb nq][idx: 4 nq][lb: 4 nq][cnt: 4 nq][scratch: keys 2 x 8 nq,
//   vals 2 x 4 nq, radix temporaries]; the scratch is reused as the offsets scan workspace.
// A query whose walk exceeds the per-lane step budget, parked for the wave-cooperative
// kernel: sorted position, current position, leading-run top and phase, the count so far
// (count mode) or the next output offset (fill modes), and in fill modes the hits still
// expected by the query's offsets.
struct HeavyEntry {
    uint32_t k;
    uint32_t cur;
    uint32_t top;
    uint32_t lead;
    uint64_t acc;
    uint32_t left;
    uint32_t pad;
};

struct PlanLayout {
    size_t off_q, off_s, off_e, off_idx, off_lb, off_cnt, off_scratch, off_k0, off_k1, off_v0, off_v1, off_rs, off_heavy,
        total;
    constexpr PlanLayout(int cb, size_t nq) : off_q(0), off_s(0), off_e(0), off_idx(0), off_lb(0), off_cnt(0), off_scratch(0),
                                             off_k0(0), off_k1(0), off_v0(0), off_v1(0), off_rs(0), off_heavy(0), total(0) {
        off_q = 256;
        off_s = off_q + align256(4 * nq);
        off_e = off_s + align256((size_t)cb * nq);
        off_idx = off_e + align256((size_t)cb * nq);
        off_lb = off_idx + align256(4 * nq);
        off_cnt = off_lb + align256(4 * nq);
        off_scratch = off_cnt + align256(4 * nq);
        off_k0 = off_scratch;
        off_k1 = off_k0 + align256(8 * nq);
        off_v0 = off_k1 + align256(8 * nq);
        off_v1 = off_v0 + align256(4 * nq);
        off_rs = off_v1 + align256(4 * nq);
        size_t scan = align256(scan_ws_items(nq) * 8);
        // heavy list: after the scan workspace; dead once a walk launch has finished
        off_heavy = off_scratch + scan;
        total = std::max(std::max(off_rs + rs_temp_bytes(nq), off_scratch + scan),
                         off_heavy + align256(sizeof(HeavyEntry) * nq));
    }
};

// Levels of the upper_bound search tree (level 0 = starts), passed by value.
template <typename S>
struct UbTree {
    const S* lv[MAX_LEVELS];     // lv[0] = starts
    uint32_t size[MAX_LEVELS];
    int nlv;                     // levels above the starts
};

// One resident index of a set table (siSetTable): what the batch kernels read of it.
template <typename S>
struct SetView {
    IndexView<S> v;               // sorted arrays (n = 0: an empty slot)
    const int32_t* values;        // payload in sorted order, or null
    UbTree<S> ub;                 // search tree over the starts
    const S* ub_top;
    uint32_t ub_h;
    UbTree<S> eb;                 // search tree over the sorted ends
    const S* eb_top;
    uint32_t eb_h;
    TreeDesc<S> td;               // max tree over the ends
    uint32_t rank_exact;          // 1 when no interval has start > end
};

// Tiles of a plan over several sets: the sorted queries of set g are [sbeg[g], sbeg[g] +
// scnt[g]) and its tiles are tbeg[g] .. tbeg[g + 1] - 1; tbeg[nsets] is the number of tiles.
struct MultiTiles {
    const uint32_t* scnt;
    const uint32_t* sbeg;
    const uint32_t* tbeg;
    uint32_t nsets;
};

template <typename S>
PlanView<S> plan_view(void* base, size_t nq) {
    PlanLayout L(sizeof(S), nq);
    char* p = (char*)base;
    PlanView<S> v;
    v.counter = (uint32_t*)p;
    v.q = (uint32_t*)(p + L.off_q);
    v.s = (S*)(p + L.off_s);
    v.e = (S*)(p + L.off_e);
    v.idx = (uint32_t*)(p + L.off_idx);
    v.lb = (uint32_t*)(p + L.off_lb);
    v.cnt = (uint32_t*)(p + L.off_cnt);
    v.set = nullptr;
    v.heavy = (HeavyEntry*)(p + L.off_heavy);
    return v;
}

// int32: key = (biased end << 32) | query number, payload = start bits.  Only the high
// four digits are sorted, so the key's low half carries the query number.
__global__ void k_plan_keys_i32(const int32_t* __restrict__ qs, const int32_t* __restrict__ qe,
                                size_t nq, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nq) {
        key[k] = ((uint64_t)((uint32_t)qe[k] ^ 0x80000000u) << 32) | (uint32_t)k;
        val[k] = (uint32_t)qs[k];
    }
}

// The plan's sort written straight into the plan's arrays by its last pass (the unpacking
// of k_plan_finish, done where the items are written).
struct PlanOutI32 {
    uint32_t* __restrict__ q;
    int32_t* __restrict__ s;
    int32_t* __restrict__ e;
    __device__ __forceinline__ void put(size_t dst, uint64_t key, uint32_t val) const {
        q[dst] = (uint32_t)key;
        e[dst] = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
        s[dst] = (int32_t)val;
    }
};
struct PlanOutI64 {
    uint32_t* __restrict__ q;
    int64_t* __restrict__ s;
    int64_t* __restrict__ e;
    const int64_t* __restrict__ qs;
    __device__ __forceinline__ void put(size_t dst, uint64_t key, uint32_t val) const {
        q[dst] = val;
        e[dst] = (int64_t)(key ^ 0x8000000000000000ull);
        s[dst] = qs[val];
    }
};

// The plan's flags, when its last sort pass wrote the arrays (k_plan_finish sets them otherwise).
template <typename S>
__global__ void k_plan_counters(PlanView<S> pv, bool with_lb, bool with_rank) {
    if (threadIdx.x == 0) {
        pv.counter[0] = 0;
        pv.counter[4] = with_lb ? 1u : 0u;
        pv.counter[5] = with_rank ? 1u : 0u;
    }
}

// int64: key = biased end, payload = query number.
__global__ void k_plan_keys_i64(const int64_t* __restrict__ qe, size_t nq, uint64_t* __restrict__ key,
                                uint32_t* __restrict__ val) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nq) {
        key[k] = (uint64_t)qe[k] ^ 0x8000000000000000ull;
        val[k] = (uint32_t)k;
    }
}

// lower_bound from upper_bound: the first position whose start is >= s is upper_bound(s - 1)
// + 1 (NONE + 1 wraps to 0), and 0 when s is the smallest value.
template <typename S>
__device__ __forceinline__ S smin_of() { return (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN; }

template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_plan_finish(IndexView<S> ix, const uint64_t* __restrict__ key,
                                                           const uint32_t* __restrict__ val,
                                                           const S* __restrict__ qs, size_t nq,
                                                           PlanView<S> pv, bool with_ub, bool with_lb,
                                                           bool with_rank) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        pv.counter[0] = 0;
        pv.counter[4] = with_lb ? 1u : 0u;
        pv.counter[5] = with_rank ? 1u : 0u;
    }
    if (k >= nq) return;
    uint64_t kk = key[k];
    uint32_t q;
    S s, e;
    if (sizeof(S) == 4) {
        q = (uint32_t)kk;
        e = (S)(int32_t)((uint32_t)(kk >> 32) ^ 0x80000000u);
        s = (S)(int32_t)val[k];
    } else {
        q = val[k];
        e = (S)(int64_t)(kk ^ 0x8000000000000000ull);
        s = qs[q];
    }
    pv.q[k] = q;
    pv.s[k] = s;
    pv.e[k] = e;
    if (with_ub) {
        pv.idx[k] = upper_bound_pos(ix.starts, ix.n, e);
        if (with_lb) pv.lb[k] = (s == smin_of<S>()) ? 0u : upper_bound_pos(ix.starts, ix.n, (S)(s - 1)) + 1u;
    }
}

// upper_bound of n keys through the index's search tree over the starts (defined with its
// kernel below).
template <typename S>
int launch_ub_tree(const siIndex* ix, const S* xs, size_t n, uint32_t* out, hipStream_t st);

// The plan's upper_bound (and lower_bound) positions, and its rank-form counts, through the
// search trees (below).
template <typename S>
int launch_plan_bounds(const siIndex* ix, PlanView<S> pv, size_t nq, bool with_lb, bool with_rank, hipStream_t st);

// The rank form needs no interval with start > end (defined with the launch helpers below).
inline bool rank_exact(const siIndex* ix);

// Number of sorted ends below x (branchless lower_bound).
template <typename S>
__device__ __forceinline__ uint32_t count_below(const S* __restrict__ v, uint32_t n, S x) {
    if (n == 0) return 0;
    const S* base = v;
    uint32_t len = n;
    while (len > 1) {
        uint32_t half = len >> 1;
        base = (base[half] < x) ? base + half : base;
        len -= half;
    }
    return (uint32_t)(base - v) + ((*base < x) ? 1u : 0u);
}

// Rank form of count for the planned queries: for s <= e every interval with end < s
// starts before e, so #{starts <= e && ends >= s} = (upper_bound + 1) - #{ends < s}.
// Inverted queries (s > e) take the reference walk.
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_count_rank(IndexView<S> ix, const S* __restrict__ ends_sorted,
                                                          PlanView<S> pv, size_t nq,
                                                          uint32_t* __restrict__ counts,
                                                          uint64_t* __restrict__ counts64) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const uint32_t i = pv.idx[k];
    const S s = pv.s[k], e = pv.e[k];
    uint32_t c = 0;
    if (i != NONE) {
        if (s <= e) c = i + 1u - count_below(ends_sorted, ix.n, s);
        else c = walk_count(ix, i, s);
    }
    const uint32_t q = pv.q[k];
    if (counts) counts[q] = c;
    if (counts64) counts64[q] = c;
}

// Queries per tile and resident workgroups of the persistent walk kernel.
// Launch choices of the walk kernels, set through siTuneSet (the autotuner's knobs).
//   tile         sorted queries per workgroup tile
//   count_wgs    resident workgroups per CU for the count pass (0 = occupancy limit)
//   fill_wgs     resident workgroups per CU for the fill passes (0 = occupancy limit)
//   count_rank   siCount / siCountPlanned use the rank form (default) instead of the walk
//   offsets_rank siSearchOffsets* use the rank form (default) instead of the walk
//   lane_steps   line steps a lane spends on one query before parking it for the
//                wave-cooperative kernel
//   ub_tree      1: siUpperBound descends a 64-ary tree of first keys (its top level in
//                Eytzinger order in LDS); 0: plain halving over all starts
//   plan_sync    1: the plan's radix sort reads the digit histograms back to skip constant
//                digits (one stream synchronisation); 0: it sorts every key digit without
//                a host read-back, so plan -> offsets -> fill can be captured in a graph
//   heavy_ends32 1: the wave-cooperative walk of an int64 index reads the 32-bit copy of
//                the ends when the build made one (ends spanning < 2^32); 0: the ends
//   plan_skip_low  lowest 8-bit digits of the query end the plan leaves unsorted (0..2):
//                one radix pass fewer each; queries are then in end order only up to
//                those bits, which the walk's tiles allow (their LDS window ends at the
//                tile's largest upper_bound) at some cost in window hits
//   range_skip   1: the plan also finds each query's lower_bound and the walks take the
//                run [lower_bound, upper_bound] (every position there starts inside the
//                query) as hits without reading ends; 0: they walk it
//   ub_learned   1: the plan's upper_bound / lower_bound go through the learned model
//                trained on the index (siLearnedTrain), when it has one; 0: the tree
//   learned_shrink  test only: the learned model's leaf windows narrowed by this many
//                positions per side, which forces its whole-leaf fallback (0 normally)
//   long_walk    1: the wave walk takes the hits below s - len from the index's long list
//                when it has one; 0: it walks the index's arrays all the way
//   long_len     read by the build: > 0 the long list's length threshold, 0 chosen from the
//                lengths (64 x the median, a power of two), < 0 no long list
struct Tuning {
    uint32_t tile = 1024;
    int count_wgs = 0;
    int fill_wgs = 0;
    int count_rank = 1;
    int offsets_rank = 1;
    uint32_t lane_steps = 256;
    int plan_sync = 1;
    int ub_tree = 1;
    int plan_skip_low = 0;
    int heavy_ends32 = 1;
    int range_skip = 1;
    int ub_learned = 0;
    int learned_shrink = 0;
    int long_walk = 1;
    int64_t long_len = 0;
};
Tuning g_tune;

// The index's long list as the walks see it (none when the knob long_walk is 0).
template <typename S>
LongView<S> long_view_of(const siIndex* ix) {
    LongView<S> v = ix->long_view<S>();
    if (!g_tune.long_walk) v.len = 0;
    return v;
}

// The run [lower_bound(s), upper_bound(e)] holds only hits when every interval has
// start <= end (each starts at or after s, so ends at or after s).
inline bool range_skip_on(const siIndex* ix) { return g_tune.range_skip && ix->n_inverted == 0; }

template <typename S>
int plan_impl(const siIndex* ix, const S* qs, const S* qe, size_t nq, void* plan, hipStream_t st, bool read_back) {
    PlanLayout L(sizeof(S), nq);
    char* p = (char*)plan;
    uint64_t* k0 = (uint64_t*)(p + L.off_k0);
    uint64_t* k1 = (uint64_t*)(p + L.off_k1);
    uint32_t* v0 = (uint32_t*)(p + L.off_v0);
    uint32_t* v1 = (uint32_t*)(p + L.off_v1);
    RsTemp rt = rs_temp_carve(p + L.off_rs, nq);
    const unsigned g = grid_for(nq, Q_THREADS);
    // upper_bound of the sorted ends (and lower_bound of the starts when the walks take the
    // guaranteed run): through the search tree when the index has one
    const bool learned = g_tune.ub_learned && ix->learned;
    const bool tree = !learned && g_tune.ub_tree && ix->ub.d_top && ix->n > 0;
    const bool with_lb = range_skip_on(ix);
    // with both trees of the same shape the bound search also finds each query's rank-form
    // count, which the offsets and count passes then take from the plan
    const bool with_rank = tree && rank_exact(ix) && ix->eb.d_top && ix->eb.nlv == ix->ub.nlv &&
                           ix->eb.top_h == ix->ub.top_h;
    const PlanView<S> pv = plan_view<S>(plan, nq);
    uint64_t* ks;
    uint32_t* vs;
    int rc;
    bool unpacked = false;   // the sort's last pass wrote q, s, e (when the bounds come from a search below)
    if (sizeof(S) == 4) {
        k_plan_keys_i32<<<g, Q_THREADS, 0, st>>>((const int32_t*)qs, (const int32_t*)qe, nq, k0, v0);
        LAUNCH_CHECK();
        const PlanOutI32 out{pv.q, (int32_t*)pv.s, (int32_t*)pv.e};
        rc = radix_sort_pairs(k0, v0, k1, v1, nq, 4 + g_tune.plan_skip_low, 8, rt, &ks, &vs, st, read_back,
                              (tree || learned) ? &out : nullptr, &unpacked);
    } else {
        k_plan_keys_i64<<<g, Q_THREADS, 0, st>>>((const int64_t*)qe, nq, k0, v0);
        LAUNCH_CHECK();
        const PlanOutI64 out{pv.q, (int64_t*)pv.s, (int64_t*)pv.e, (const int64_t*)qs};
        rc = radix_sort_pairs(k0, v0, k1, v1, nq, g_tune.plan_skip_low, 8, rt, &ks, &vs, st, read_back,
                              (tree || learned) ? &out : nullptr, &unpacked);
    }
    if (rc != SI_OK) return rc;
    if (unpacked) k_plan_counters<S><<<1, 64, 0, st>>>(pv, with_lb, with_rank);
    else k_plan_finish<S><<<g, Q_THREADS, 0, st>>>(ix->view<S>(), ks, vs, qs, nq, pv, !tree && !learned, with_lb, with_rank);
    LAUNCH_CHECK();
    if (learned) return si_learned_plan_bounds(ix, pv.e, pv.s, nq, pv.idx, with_lb ? pv.lb : nullptr, st);
    if (tree) return launch_plan_bounds<S>(ix, pv, nq, with_lb, with_rank, st);
    return SI_OK;
}

// ============================================================================
// The block walk
// ============================================================================
template <typename S>
struct Line {
    static constexpr int W = 64 / sizeof(S);
    S e[W];
};

template <typename S>
__device__ __forceinline__ void load_line(const S* __restrict__ ends, uint32_t b0, Line<S>& L) {
    const uint4* p = reinterpret_cast<const uint4*>(ends + b0);
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = p[k];
    memcpy(L.e, v, 64);
}

// Bit k set when position b0 + k (k <= kc) is a hit.  The miss bits are shifted in
// from the top (miss + miss + bit: one compare and one add-with-carry per element).
template <typename S>
__device__ __forceinline__ uint32_t hit_mask(const Line<S>& L, uint32_t kc, S s) {
    uint32_t miss = 0;
#pragma unroll
    for (int k = Line<S>::W - 1; k >= 0; --k) miss = miss + miss + (uint32_t)(L.e[k] < s);
    return ~miss & ((2u << kc) - 1u);   // kc < W <= 16
}

// Offset in the line of the top of the trailing misses [b0, b0 + x]: just below the
// lowest hit, or kc when the line has none.  Only used when b0 itself is not a hit; the
// walk then continues at exitp[b0 + x] (branch[] of the leftmost maximum of that run).
__device__ __forceinline__ uint32_t miss_top(uint32_t kc, uint32_t m) {
    return m ? (uint32_t)(__builtin_ctz(m) - 1) : kc;
}

// One slot of a wave's emission scratch in LDS: the segment that starts at this output
// slot of the current 64-slot chunk, 16 bytes (one LDS access each way).  kv = m | dir << 16
// | pre << 18, never 0 for a segment (a mask segment has m != 0, a range dir != 0); 0 when
// the slot is free.  dir (a signed 2-bit field): +1 for the ascending range base ..
// base + c - 1, -1 for the descending range base .. base - c + 1, 0 for the set bits of m,
// highest first, offset by base (a range's m is 0).  pre: the segment's first slot in the
// emission's concatenation (< 2^14).  a_lo / a_hi: the address of the segment's first
// output element (so a slot's store address is one add away).
struct __align__(16) EmitSlot {
    uint32_t kv, base;
    uint32_t a_lo, a_hi;
};

// Orders this wave's LDS accesses for the compiler; LDS executes a wave's accesses in
// issue order, so no wait is needed between lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Wave-cooperative output.  Every lane holds this step's segment (c values at out
// offset o); the wave scans the counts and lane x writes element x of the concatenated
// segments, so one store instruction covers 64 consecutive output slots of a few lists
// instead of one slot in each of 64 lists.  The slot -> segment map of each 64-slot
// chunk goes through the wave's 64 LDS slots (sd): a lane whose segment starts in the
// chunk writes its descriptor at its start slot, a ballot gives the chunk's start mask,
// and slot x takes the last start at or before x (or the segment running in from the
// previous chunk).  The segments are contiguous in the concatenation, so every slot
// below the total lies in the segment found for it: no slot is written outside the
// segment its owner was given.  All lanes execute the ballots.
#ifndef SI_EMIT_LUT
#define SI_EMIT_LUT 1
#endif

// Table of the set bits of a byte, highest first: nibble k of kbit_entry(b) is the bit
// position of the (k + 1)-th highest set bit of b (unused nibbles are 0).
__device__ __forceinline__ uint32_t kbit_entry(uint32_t b) {
    uint32_t e = 0, k = 0;
    for (int p = 7; p >= 0; --p) {
        if ((b >> p) & 1u) {
            e |= (uint32_t)p << (4u * k);
            ++k;
        }
    }
    return e;
}

// A 32-bit word in global memory: output addresses rebuilt from integers keep their address
// space, so their stores stay global stores (a generic pointer would make them flat ones).
typedef __attribute__((address_space(1))) uint32_t global_u32;

// One output segment of a lane: c values at output offset o, either the ascending range
// base .. base + c - 1 (asc), the descending range base .. base - c + 1 (desc), or the set
// bits of m (descending) offset by base.
struct Seg {
    uint32_t c, mk, base;   // mk = m | asc << 16 | desc << 17
    uint64_t o;
};

// Segment descriptor written at its start slot of the current chunk (a: the address of the
// segment's first output element): the kind bits of mk become the signed step dir
// (asc 01 -> +1, desc 10 -> 11 = -1, mask 00 -> 0).
__device__ __forceinline__ void put_segment(EmitSlot* __restrict__ sd, uint32_t slot, uint32_t pre, uint32_t mk,
                                            uint32_t base, uint64_t a) {
    const uint32_t kind = (mk >> 16) & 3u;
    EmitSlot e;
    e.kv = (mk & 0xFFFFu) | (kind | (kind >> 1)) << 16 | pre << 18;
    e.base = base;
    e.a_lo = (uint32_t)a;
    e.a_hi = (uint32_t)(a >> 32);
    sd[slot] = e;
}

// Steps whose segments the fill walk holds and emits together (their segments of one
// query are adjacent in its output list, so each emission writes longer runs).
#ifndef SI_EMIT_EVERY
#define SI_EMIT_EVERY 2
#endif
constexpr int EMIT_EVERY = SI_EMIT_EVERY;
static_assert(EMIT_EVERY >= 2 && EMIT_EVERY <= 4, "2 to 4 segments per emission");

// Range segments of at least WAVE_RANGE values are written on their own, one at a time by
// the whole wave (see emit_segments): a few instructions per 64 values there, against the
// slot search and the position per slot in the shared chunks.  At most 64, so that every
// segment left for the chunks is shorter than a chunk (at most 16 values for a line mask).
#ifndef SI_WAVE_RANGE
#define SI_WAVE_RANGE 64
#endif
constexpr uint32_t WAVE_RANGE = SI_WAVE_RANGE;
static_assert(WAVE_RANGE >= 16 && WAVE_RANGE <= 64, "chunk segments are shorter than a chunk");
// the concatenation of one emission's chunk segments fits EmitSlot's 14-bit pre
static_assert(64 * (EMIT_EVERY + 1) * (WAVE_RANGE - 1) < (1 << 14), "emission slots fit 14 bits");

// Every lane holds up to N segments; they occupy consecutive slots of the concatenation.
// Range segments of at least WAVE_RANGE values are first written one at a time by the whole
// wave; what is left for the shared chunks are line masks (at most W values each) and ranges
// shorter than WAVE_RANGE, at most 64 x N x 64 slots.
template <typename S, int MODE, int N>
__device__ __forceinline__ void emit_segments(const Seg (&sg_in)[N], int lane, EmitSlot* __restrict__ sd,
                                              const uint32_t* __restrict__ kbit,
                                              uint32_t* __restrict__ out_idx,
                                              const int32_t* __restrict__ vals,
                                              int32_t* __restrict__ out_val) {
    constexpr int W = Line<S>::W;
    // the output array, as 32-bit words (positions, or values)
    uint32_t* const dst = (MODE == W_VALUES) ? reinterpret_cast<uint32_t*>(out_val) : out_idx;
    Seg sg[N];
#pragma unroll
    for (int j = 0; j < N; ++j) sg[j] = sg_in[j];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint64_t big = __builtin_amdgcn_ballot_w64(((sg[j].mk >> 16) & 3u) != 0u && sg[j].c >= WAVE_RANGE);
        while (big) {
            const int L = __builtin_ctzll(big);
            big &= big - 1ull;
            // (readlane returns int: every word is taken as uint32_t before it is widened)
            const uint32_t bc = (uint32_t)__builtin_amdgcn_readlane(sg[j].c, L);
            const uint32_t bmk = (uint32_t)__builtin_amdgcn_readlane(sg[j].mk, L);
            const uint32_t bbase = (uint32_t)__builtin_amdgcn_readlane(sg[j].base, L);
            const uint64_t bo = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(sg[j].o >> 32), L) << 32) |
                                (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)sg[j].o, L);
            const bool asc = (bmk >> 16) & 1u;
            // rounds of 64 values (counted so that x never wraps, whatever bc is)
            const uint32_t rounds = (uint32_t)(((uint64_t)bc + WAVE - 1u) / WAVE);
            for (uint32_t t = 0; t < rounds; ++t) {
                const uint32_t x = t * (uint32_t)WAVE + (uint32_t)lane;
                if (x < bc) {
                    const uint32_t p = asc ? bbase + x : bbase - x;
                    if (MODE == W_VALUES) out_val[bo + x] = vals[p];
                    else out_idx[bo + x] = p;
                }
            }
            if (lane == L) sg[j].c = 0u;
        }
    }
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) c += sg[j].c;
    if (__ballot(c != 0) == 0) return;
    const uint32_t inc = wave_incl_scan_dpp(c);
    const uint32_t total = __builtin_amdgcn_readlane(inc, WAVE - 1);
    const uint32_t pre0 = inc - c;
    const uint64_t le = (lane == WAVE - 1) ? ~0ull : ((2ull << lane) - 1ull);   // lanes <= this
    for (uint32_t x0 = 0; x0 < total; x0 += WAVE) {
        uint32_t pre = pre0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (sg[j].c != 0 && pre - x0 < (uint32_t)WAVE)   // (pre < x0 wraps to a large value)
                put_segment(sd, pre - x0, pre, sg[j].mk, sg[j].base, (uint64_t)(uintptr_t)(dst + sg[j].o));
            pre += sg[j].c;
        }
        wave_lds_sync();
        const uint64_t mine = __builtin_amdgcn_ballot_w64(sd[lane].kv != 0u) & le;
        // slot WAVE holds the segment running in from the previous chunk (never read in
        // the first chunk, whose slot 0 always starts a segment)
        const EmitSlot d = sd[mine ? 63u - (uint32_t)__builtin_clzll(mine) : (uint32_t)WAVE];
        wave_lds_sync();
        sd[lane].kv = 0u;   // free the slot for the next chunk (every read above is done)
        // the segment holding the chunk's last slot runs into the next chunk
        if (lane == WAVE - 1) sd[WAVE] = d;
        const uint32_t x = x0 + (uint32_t)lane;
        const uint32_t r = x - (d.kv >> 18);
        const uint32_t om = d.kv & 0xFFFFu;
        // The mask form's position (for a range, m = 0 gives base), plus r * dir for a
        // range: one expression, no branch or select.  (A chunk segment spans at most two
        // chunks, so a stored slot has r < 128, and r * dir fits the 24-bit multiply.)
        uint32_t mpos;
        if (SI_EMIT_LUT) {
            // the (r+1)-th highest set bit of om from the byte table
            uint32_t byte = om & 0xFFu, k = r, add = 0;
            if (W == 16) {
                const uint32_t hi = om >> 8;
                const uint32_t ph = (uint32_t)__popc(hi);
                const bool up = r < ph;
                byte = up ? hi : byte;
                k = up ? r : r - ph;
                add = up ? 8u : 0u;
            }
            mpos = d.base + add + ((kbit[byte] >> ((4u * k) & 31u)) & 7u);
        } else {
            // the (r+1)-th highest set bit of om: largest p with popc(om >> p) >= r + 1
            uint32_t p = 0;
#pragma unroll
            for (int sh = W / 2; sh >= 1; sh >>= 1)
                p = ((uint32_t)__popc(om >> (p + sh)) >= r + 1u) ? p + sh : p;
            mpos = d.base + p;
        }
        const int32_t dir = __builtin_amdgcn_sbfe((int32_t)d.kv, 16, 2);
        const uint32_t pos = (uint32_t)__mul24((int32_t)(r & 127u), dir) + mpos;
        if (x < total) {
            global_u32* const a = reinterpret_cast<global_u32*>((uintptr_t)(((uint64_t)d.a_hi << 32) | d.a_lo));
            a[r] = (MODE == W_VALUES) ? (uint32_t)vals[pos] : pos;
        }
    }
}

// LDS window of a tile: ends and exitp of positions [wlo, wlo + WIN), which always
// ends at or above the tile's largest upper_bound.  26 KB (int32) / 27 KB (int64) per
// workgroup, so that with the 12 KB of emission slots and the 1 KB bit table four
// 512-thread workgroups fit in a CU's 160 KB of LDS and the walk runs 8 waves per SIMD
// (its registers are held under 64 for that).  On 10M int32 intervals x 10M queries this beats an 8128-entry window at
// 4 waves per SIMD by 20 % (fill) with tiles of 1024 queries.  The sizes (multiples of
// 16) and the waves per SIMD can be set at compile time (SI_WALK_WIN32, SI_WALK_WIN64,
// SI_WALK_WAVES_PER_EU).
#ifndef SI_WALK_WIN32
#define SI_WALK_WIN32 3328
#endif
#ifndef SI_WALK_WIN64
#define SI_WALK_WIN64 2272
#endif
#ifndef SI_WALK_WAVES_PER_EU
#define SI_WALK_WAVES_PER_EU 8
#endif
#define SI_WALK_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SI_WALK_WAVES_PER_EU, SI_WALK_WAVES_PER_EU)))
template <typename S> struct TileCfg;
template <> struct TileCfg<int32_t> { static constexpr uint32_t WIN = SI_WALK_WIN32; };   // x 8 B (ends, exitp)
template <> struct TileCfg<int64_t> { static constexpr uint32_t WIN = SI_WALK_WIN64; };   // x 12 B
static_assert(SI_WALK_WIN32 % 16 == 0 && SI_WALK_WIN64 % 16 == 0, "window sizes are multiples of 16");

constexpr int T_THREADS = 512;   // 8 waves per workgroup, two workgroups per CU

// MULTI: the plan covers several resident indexes (a set table): each tile holds queries of
// one set, whose arrays the tile reads (sets[g]), and positions are the set's own.
template <typename S, int MODE, bool MULTI = false>
__global__ __launch_bounds__(T_THREADS) SI_WALK_OCCUPANCY void k_walk(IndexView<S> ix_arg, PlanView<S> pv, uint32_t nq,
                                                    uint32_t tile, uint32_t ntiles, uint32_t lane_steps,
                                                    bool use_lb,
                                                    uint32_t* __restrict__ tile_counter,
                                                    uint32_t* __restrict__ counts,
                                                    uint64_t* __restrict__ counts64,
                                                    const uint64_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ out_idx,
                                                    const int32_t* __restrict__ vals_arg,
                                                    int32_t* __restrict__ out_val, uint64_t cap,
                                                    const SetView<S>* __restrict__ sets, MultiTiles mt) {
    constexpr int W = Line<S>::W;
    constexpr uint32_t WIN = TileCfg<S>::WIN;
    __shared__ __align__(16) S s_end[WIN];
    __shared__ __align__(16) uint32_t s_ex[WIN];   // exit targets (IndexView::exitp)
    __shared__ uint32_t s_tile, s_next, s_whi1, s_set, s_t0, s_t1;
    IndexView<S> ix = ix_arg;
    const int32_t* __restrict__ vals = vals_arg;
    // per-wave emission slots: 64 + one for the segment carried between chunks
    __shared__ EmitSlot s_emit[(MODE == W_COUNT) ? 1 : T_THREADS / WAVE * (WAVE + 1)];
    __shared__ uint32_t s_kbit[(MODE == W_COUNT || !SI_EMIT_LUT) ? 1 : 256];   // byte set-bit table
    const int lane = lane_id();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the guaranteed runs need the plan's lower_bound positions
    use_lb = use_lb && __builtin_amdgcn_readfirstlane(pv.counter[4]) != 0u;
    EmitSlot* const my_emit = s_emit + (MODE == W_COUNT ? 0u : (threadIdx.x / WAVE) * (WAVE + 1));
    if (MODE != W_COUNT) my_emit[lane].kv = 0u;   // emission slots start free (own wave's)
    // segments held from the previous steps (emission runs every EMIT_EVERY steps)
    Seg held[EMIT_EVERY - 1];
#pragma unroll
    for (int j = 0; j < EMIT_EVERY - 1; ++j) held[j] = Seg{0u, 0u, 0u, 0ull};
    int phase = 0;
    if (MODE != W_COUNT && SI_EMIT_LUT) {
        for (uint32_t b = threadIdx.x; b < 256u; b += T_THREADS) s_kbit[b] = kbit_entry(b);
    }   // read only after the first tile's __syncthreads

    while (true) {
        // ---- take the next tile (workgroup-level work stealing)
        if (threadIdx.x == 0) {
            const uint32_t tt = atomicAdd(tile_counter, 1u);
            s_tile = tt;
            s_whi1 = 0u;
            if (MULTI) {
                // the set of tile tt: the last g with tbeg[g] <= tt (sets without queries
                // have no tiles); tiles past the last are NONE
                const uint32_t total = mt.tbeg[mt.nsets];
                uint32_t lo = 0, hi = mt.nsets;   // tbeg[lo] <= tt < tbeg[hi]
                if (tt < total) {
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (mt.tbeg[mid] <= tt) lo = mid;
                        else hi = mid;
                    }
                    s_set = lo;
                    s_t0 = mt.sbeg[lo] + (tt - mt.tbeg[lo]) * tile;
                    s_t1 = min(s_t0 + tile, mt.sbeg[lo] + mt.scnt[lo]);
                } else {
                    s_tile = NONE;
                }
            }
        }
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= ntiles) break;
        uint32_t t0, t1;
        if (MULTI) {
            const uint32_t g = __builtin_amdgcn_readfirstlane(s_set);
            t0 = __builtin_amdgcn_readfirstlane(s_t0);
            t1 = __builtin_amdgcn_readfirstlane(s_t1);
            ix = sets[g].v;
            vals = sets[g].values;
        } else {
            t0 = t * tile;
            t1 = min(t0 + tile, nq);
        }
        // the window ends at the tile's largest upper_bound (the plan's order by end need
        // not be exact): max of idx + 1 over the tile, where NONE + 1 = 0 drops queries
        // that have no start <= their end
        {
            uint32_t mx = 0;
            for (uint32_t k = t0 + threadIdx.x; k < t1; k += T_THREADS) mx = max(mx, pv.idx[k] + 1u);
            if (mx) atomicMax(&s_whi1, mx);
        }
        __syncthreads();
        const uint32_t whi = s_whi1 - 1u;
        uint32_t wlo = 0, wcnt = 0;
        if (whi != NONE) {
            const int64_t lo = (int64_t)(whi & ~15u) + 16 - (int64_t)WIN;
            wlo = lo < 0 ? 0u : (uint32_t)lo;
            wcnt = min(WIN, ix.npad - wlo);   // multiple of 16
        }
        {
            const uint4* ge = reinterpret_cast<const uint4*>(ix.ends + wlo);
            const uint4* gb = reinterpret_cast<const uint4*>(ix.exitp + wlo);
            uint4* le = reinterpret_cast<uint4*>(s_end);
            uint4* lb = reinterpret_cast<uint4*>(s_ex);
            const uint32_t ne = wcnt * (uint32_t)sizeof(S) / 16u, nb = wcnt / 4u;
            for (uint32_t j = threadIdx.x; j < ne; j += T_THREADS) le[j] = ge[j];
            for (uint32_t j = threadIdx.x; j < nb; j += T_THREADS) lb[j] = gb[j];
        }
        if (threadIdx.x == 0) s_next = t0;
        __syncthreads();

        // ---- walk the tile; each wave refills idle lanes from pools of sorted queries
        uint32_t pool_base = 0, pool_left = 0;
        bool exhausted = false, has = false, lead = false;
        uint32_t q = 0, cur = NONE, cnt = 0, top = 0, kq = 0, steps = 0;
        uint32_t rng = NONE;   // lower end of the query's guaranteed run [rng, cur], or NONE
        S s = 0;
        uint64_t o = 0, oend = 0;   // output cursor and end of the query's CSR segment
        while (true) {
            const bool need = !has;
            const uint64_t nm = __ballot(need);
            if (nm) {
                if (pool_left == 0 && !exhausted) {
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(&s_next, POOL);
                    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, WAVE));   // pool state is wave-uniform
                    if (b >= t1) exhausted = true;
                    else { pool_base = b; pool_left = min(POOL, t1 - b); }
                }
                const uint32_t r = (uint32_t)__popcll(nm & lt);
                if (need && r < pool_left) {
                    const uint32_t k = pool_base + r;
                    q = pv.q[k];
                    s = pv.s[k];
                    cur = pv.idx[k];
                    cnt = 0;
                    kq = k;
                    steps = 0;
                    has = true;
                    if (MODE != W_COUNT) {
                        // cnt counts down the hits the offsets promise; the output cursor and
                        // segment end are cut at the capacity
                        const uint64_t o0 = offsets[q], o1 = offsets[q + 1];
                        cnt = (uint32_t)(o1 - o0);
                        o = min(o0, cap);
                        oend = min(o1, cap);
                    }
                    if (MODE == W_VECTOR) { top = cur; lead = true; }
                    if (use_lb) {
                        const uint32_t b = pv.lb[k];
                        rng = (cur != NONE && b <= cur) ? b : NONE;
                    }
                }
                const uint32_t take = min((uint32_t)__popcll(nm), pool_left);
                pool_base += take;
                pool_left -= take;
            }
            if (!__any(has)) {
                if (exhausted) break;
                continue;
            }
            // this step's output segment of the lane: c_emit values at e_o, either the
            // ascending range e_base .. e_base + c_emit - 1 or the set bits of e_m
            // (descending) offset by e_base
            uint32_t c_emit = 0, e_m = 0, e_base = 0;
            bool e_asc = false, e_desc = false;
            uint64_t e_o = 0;
            if (has && rng != NONE) {
                // [rng, cur] = positions starting inside the query (lower_bound .. upper_bound):
                // all hits, taken without reading ends; the walk goes on below them
                const uint32_t c = cur - rng + 1u;
                if (MODE == W_COUNT) {
                    cnt += c;
                } else if (MODE != W_VECTOR) {
                    cnt -= c;
                    c_emit = (uint32_t)min((uint64_t)c, oend - o);
                    e_base = cur;
                    e_desc = true;
                    e_o = o;
                    o += c_emit;
                }
                cur = rng - 1u;   // rng == 0 wraps to NONE
                rng = NONE;
                if (MODE == W_VECTOR && cur == NONE) {
                    // the leading run reaches position 0: written ascending now
                    cnt -= top + 1u;
                    c_emit = (uint32_t)min((uint64_t)(top + 1u), oend - o);
                    e_base = 0;
                    e_asc = true;
                    e_o = o;
                    o += c_emit;
                    lead = false;
                }
            } else if (has && cur != NONE) {
                const uint32_t b0 = cur & ~(uint32_t)(W - 1);
                const uint32_t kc = cur - b0;
                const bool in_win = b0 >= wlo;   // every position visited is <= whi
                Line<S> L;
                if (in_win) load_line(s_end, b0 - wlo, L);
                else load_line(ix.ends, b0, L);
                const uint32_t m = hit_mask(L, kc, s);
                if (MODE == W_VECTOR && lead) {
                    // leading contiguous run of hits, written ascending once its bottom is known
                    const uint32_t inv = ~m & ((2u << kc) - 1u);
                    uint32_t lo;
                    if (inv == 0) {
                        cur = b0 - 1u;
                        lo = 0;
                    } else {
                        const uint32_t hm = 31u - (uint32_t)__builtin_clz(inv);
                        cur = b0 + hm;   // the first miss below the run; the walk resumes there
                        lo = cur + 1u;
                    }
                    if (inv != 0 || cur == NONE) {
                        cnt -= top - lo + 1u;
                        c_emit = (uint32_t)min((uint64_t)(top - lo + 1u), oend - o);   // never past the segment
                        e_base = lo;
                        e_asc = true;
                        e_o = o;
                        o += c_emit;
                        lead = false;
                    }
                } else {
                    if (MODE == W_COUNT) {
                        cnt += (uint32_t)__popc(m);
                    } else {
                        cnt -= (uint32_t)__popc(m);
                        c_emit = (uint32_t)min((uint64_t)__popc(m), oend - o);   // never past the segment
                        e_m = m;
                        e_base = b0;
                        e_o = o;
                        o += c_emit;
                    }
                    if (m & 1u) {
                        cur = b0 - 1u;   // b0 == 0 wraps to NONE
                    } else {
                        const uint32_t p = b0 + miss_top(kc, m);
                        // LDS read always (clamped), global read only below the window; the
                        // empty asm pins the LDS value so the two loads are not folded into
                        // one generic (flat) load through a selected pointer
                        uint32_t nb = s_ex[in_win ? p - wlo : 0u];
                        asm volatile("" : "+v"(nb));
                        if (!in_win) nb = ix.exitp[p];
                        cur = nb;
                    }
                }
            }
            if (MODE != W_COUNT) {
                // hold this step's segment; every EMIT_EVERY steps emit the held ones with
                // it (phase is wave-uniform: every lane runs every step)
                const Seg now = Seg{c_emit, e_m | ((e_asc ? 1u : 0u) << 16) | ((e_desc ? 1u : 0u) << 17), e_base, e_o};
                if (phase == EMIT_EVERY - 1) {
                    Seg all[EMIT_EVERY];
#pragma unroll
                    for (int j = 0; j < EMIT_EVERY - 1; ++j) all[j] = held[j];
                    all[EMIT_EVERY - 1] = now;
                    emit_segments<S, MODE, EMIT_EVERY>(all, lane, my_emit, s_kbit, out_idx, vals, out_val);
#pragma unroll
                    for (int j = 0; j < EMIT_EVERY - 1; ++j) held[j].c = 0;
                    phase = 0;
                } else {
#pragma unroll
                    for (int j = 0; j < EMIT_EVERY - 1; ++j) {
                        if (phase == j) held[j] = now;
                    }
                    ++phase;
                }
            }
            if (has && cur == NONE) {
                if (MODE == W_COUNT) {
                    if (counts) counts[q] = cnt;
                    if (counts64) counts64[q] = cnt;
                } else if (cnt != 0u) {
                    atomicOr(&pv.counter[3], 1u);   // the walk's hits differ from the offsets
                }
                lead = false;
                has = false;
            } else if (has && ++steps >= lane_steps) {
                // over budget: park the walk for the wave-cooperative kernel, take another query
                const uint32_t h = atomicAdd(&pv.counter[1], 1u);
                HeavyEntry he;
                he.k = kq;
                he.cur = cur;
                he.top = top;
                he.lead = lead ? 1u : 0u;
                he.acc = (MODE == W_COUNT) ? (uint64_t)cnt : o;
                he.left = cnt;
                he.pad = 0u;
                pv.heavy[h] = he;
                lead = false;
                has = false;
            }
        }
        __syncthreads();   // LDS window and counters are reused by the next tile
    }
    // the segments held from the last steps (all lanes of the workgroup left the tile loop)
    if (MODE != W_COUNT) emit_segments<S, MODE, EMIT_EVERY - 1>(held, lane, my_emit, s_kbit, out_idx, vals, out_val);
}

// Wave-cooperative walk of one query, guided by the max tree.  The hits at and below
// position cur are {j <= cur : ends[j] >= s}, and an aligned 64-position line holds one iff
// its maximum (level 1 of the max tree, one entry per line) is >= s.  The wave visits exactly
// those lines, highest first, one 64-position line per step with hits by ballot: the lines of
// the current group of 64 lines come from one ballot over their level-1 maxima (kept as a
// wave-uniform mask), and when the group has none left, the nearest group to the left with a
// line >= s is found through the levels above (prev_line, the search build uses for branch).
// Up to WAVE_LINES such lines (LONG_LINES in the long list) are fetched per round trip.
// The walk reads no branch or exit array: ends (or the 32-bit copy of int64 ends) and the
// small level-1 array only, plus, in the long list, the position of each entry, fetched in
// the same round trip as its end (a whole line costs the same cache lines as its hits).
//
// In the C++ vector order the leading run [i* + 1, top] is written ascending once its bottom
// i* (the first miss below top) is found: the lead phase steps through consecutive lines
// from cur while they are all hits.
//
// All 64 lanes must reach every ballot together, so every value that steers a loop is
// wave-uniform (ballot results, or readfirstlane).  (A kernel that held two lane == 0
// blocks in one loop once had its tests merged by the compiler, which sent lanes 1..63 back
// to a shuffle without lane 0; the loops here hold no lane-dependent branch.)
#ifndef SI_WAVE_LINES
#define SI_WAVE_LINES 4
#endif
#ifndef SI_LONG_LINES
#define SI_LONG_LINES 8
#endif
constexpr uint32_t WAVE_LINES = SI_WAVE_LINES;   // lines with hits fetched per round trip
constexpr uint32_t LONG_LINES = SI_LONG_LINES;   // the same in the long list (dense hits)

// Last position whose key is <= x through a search tree (UbTree), by one wave, or NONE:
// the lanes sample every 64th key of the top level (at most 4095 keys), then read the 64
// keys of the sampled chunk, then at each level below the 64 keys under the entry found
// (whose first key is that entry).  Keys are sorted, so each ballot is a prefix of lanes.
template <typename S>
__device__ __forceinline__ uint32_t wave_tree_ub(const UbTree<S>& t, S x, int lane) {
    const S* __restrict__ topl = t.lv[t.nlv];
    const uint32_t m = t.size[t.nlv];
    uint32_t i = (uint32_t)lane * 64u;
    uint64_t b = __ballot(i < m && topl[i] <= x);
    if (!b) return NONE;
    const uint32_t c = (uint32_t)__popcll(b) - 1u;
    i = c * 64u + (uint32_t)lane;
    b = __ballot(i < m && topl[i] <= x);
    uint32_t j = c * 64u + (uint32_t)__popcll(b) - 1u;
    for (int l = t.nlv - 1; l >= 0; --l) {
        i = j * 64u + (uint32_t)lane;
        b = __ballot(i < t.size[l] && t.lv[l][i] <= x);
        j = j * 64u + (uint32_t)__popcll(b) - 1u;
    }
    return j;
}

// An entry of an ends array, through its 32-bit copy where one is kept (int64 ends within
// 2^32 of their minimum).
template <typename S>
__device__ __forceinline__ S end_at(const S* __restrict__ ends, const uint32_t* __restrict__ ends32, int64_t base,
                                    uint32_t i) {
    if (sizeof(S) == 8 && ends32) return (S)((int64_t)ends32[i] + base);
    return ends[i];
}

template <typename S>
__device__ __forceinline__ S wave_end(const IndexView<S>& ix, uint32_t i) {
    return end_at(ix.ends, ix.ends32, ix.ends32_base, i);
}

// Highest line l in [lo_line, L0) whose level-1 maximum is >= s, or NONE: the entries
// [g, pos) of a level's group by one ballot, else the level above over [0, g); once found,
// down to the highest child >= s at each level (a parent's maximum is its children's).
template <typename S>
__device__ __forceinline__ uint32_t prev_line(const TreeDesc<S>& td, uint32_t L0, uint32_t lo_line, S s, int lane) {
    constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
    uint32_t pos = L0;   // exclusive bound at level l
    for (int l = 1; l < td.nlev; ++l) {
        if (pos == 0) return NONE;
        const uint32_t g = (pos - 1u) & ~63u;
        const uint32_t i = g + (uint32_t)lane;
        const S v = (i < pos) ? td.lvl[l][i] : SMIN;
        const uint64_t m = __ballot(v >= s);
        if (m) {
            uint32_t node = g + 63u - (uint32_t)__builtin_clzll(m);
            for (int d = l - 1; d >= 1; --d) {
                const uint32_t c = node * 64u + (uint32_t)lane;
                const S w = (c < td.size[d]) ? td.lvl[d][c] : SMIN;
                const uint64_t mm = __ballot(w >= s);
                if (!mm) return NONE;   // cannot happen: the parent's maximum is a child's
                node = node * 64u + 63u - (uint32_t)__builtin_clzll(mm);
            }
            return node >= lo_line ? node : NONE;
        }
        // the lines left are [0, g * 64^(l-1)): none at or above lo_line, or the level above
        if (((uint64_t)g << (6 * (l - 1))) <= (uint64_t)lo_line) return NONE;
        pos = g / 64u;   // level l + 1 entries [0, g / 64) cover level l entries [0, g)
    }
    return NONE;
}

// Mask of the lines of group g1 (lines g1 .. g1 + 63) whose maximum is >= s.
template <typename S>
__device__ __forceinline__ uint64_t group_mask(const TreeDesc<S>& td, uint32_t g1, S s, int lane) {
    const uint32_t i = g1 + (uint32_t)lane;
    return __ballot(i < td.size[1] && td.lvl[1][i] >= s);
}

// Cache policy bits of the wave walk's position stores: 2 = non-temporal (the lists are
// written once and not read back by the walk, so they need not hold L2 lines the walk's
// reads of ends and positions could use); 0 = default.
#ifndef SI_OUT_AUX
#define SI_OUT_AUX 2
#endif

// Bits 0 .. x of a 64-bit mask (x <= 63).
__device__ __forceinline__ uint64_t mask_upto(uint32_t x) { return (x >= 63u) ? ~0ull : ((2ull << x) - 1ull); }

// The hits among entries [lo, cur] of an ends array with max tree td (td.lvl[0] = the ends),
// highest first: one 64-entry line per step, the lines with a maximum >= s only, NL lines
// per round trip.  The line of cur and the line of lo are taken on their own, clipped to
// [lo, cur]; the lines between need no clipping.  The output value of entry i is posmap[i]
// (LONG: the long list), or i itself.  E32: the ends are read through their 32-bit copy
// (end - base32), compared with s - base32 (0 when s <= base32; no hit at all when s is
// more than 2^32 - 1 above base32), which selects the same entries.  acc, oend, left and the modes as wave_walk; the output offset of a hit within
// this walk is a 32-bit count (a walk finds at most one hit per position).
template <typename S, int MODE, bool LONG, bool E32, uint32_t NL>
__device__ __forceinline__ void guided_walk(const S* __restrict__ ends, const uint32_t* __restrict__ ends32,
                                            int64_t base32, const uint32_t* __restrict__ posmap,
                                            const TreeDesc<S>& td, int lane, S s, uint32_t cur, uint32_t lo,
                                            uint64_t& acc, uint64_t oend, uint32_t& left,
                                            uint32_t* __restrict__ out_idx, const int32_t* __restrict__ vals,
                                            int32_t* __restrict__ out_val) {
    if (cur == NONE || cur < lo) return;
    uint32_t s32 = 0;
    if (E32 && (int64_t)s > base32) {
        const uint64_t d = (uint64_t)s - (uint64_t)base32;
        if (d > 0xFFFFFFFFull) return;   // s is above every end of the array: no hit
        s32 = (uint32_t)d;
    }
    // fill modes: this walk writes out[acc + k] for k < room
    uint32_t* const outq = (MODE == W_DESC || MODE == W_VECTOR) ? out_idx + acc : nullptr;
    int32_t* const outv = (MODE == W_VALUES) ? out_val + acc : nullptr;
    const uint32_t room = (MODE == W_COUNT) ? 0u : (uint32_t)min(oend - acc, (uint64_t)0xFFFFFFFFu);
    uint32_t k = 0;   // hits found by this walk
    // this lane's entry of line b0 (the long list's arrays are padded to whole lines)
    auto load = [&](uint32_t b0, uint32_t& e32, S& ev, uint32_t& pv) {
        if (E32) e32 = ends32[b0 + (uint32_t)lane];
        else ev = ends[b0 + (uint32_t)lane];
        if (MODE != W_COUNT) pv = LONG ? posmap[b0 + (uint32_t)lane] : b0 + (uint32_t)lane;
    };
    // Position lists are written through a buffer resource over out[acc + k0 ..] (k0: the
    // hits found before the current round of lines) that holds the room left, at most 2^28
    // values: every lane stores, and a lane without a hit (or past the room) stores at an
    // offset out of the resource's range (0x80000000 at least), which the hardware drops, so
    // the lines' takes need no branch.  The payload mode writes with a lane test instead.
    auto out_rsrc = [&](uint32_t k0) {
        const uint32_t nv = (k0 < room) ? min(room - k0, 1u << 28) : 0u;
        return __builtin_amdgcn_make_buffer_rsrc((MODE == W_DESC || MODE == W_VECTOR) ? (void*)(outq + k0) : nullptr,
                                                 (short)0, (int)(nv * 4u), 0x00020000);
    };
    // h: this lane's entry is a hit; the wave's hits are written highest lane first
    auto take = [&](bool h, uint32_t pv, __amdgpu_buffer_rsrc_t rs, uint32_t k0) {
        const uint64_t m = __builtin_amdgcn_ballot_w64(h);
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (MODE != W_COUNT) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const uint32_t r = k + cnt - 1u - below;
            if (MODE == W_VALUES) {
                if (h && r < room) outv[r] = vals[pv];
            } else {
                // (r - k0) * 4 < 2^31 for a hit; bit 31 set for the others
                __builtin_amdgcn_raw_buffer_store_b32(pv, rs, ((r - k0) << 2) | (h ? 0u : 0x80000000u), 0, SI_OUT_AUX);
            }
        }
        k += cnt;
    };
    auto is_hit = [&](uint32_t e32, S ev) -> bool { return E32 ? (e32 >= s32) : (ev >= s); };
    // one line on its own, clipped to entries [klo, kc] of it
    auto take_clipped = [&](uint32_t line, uint32_t kc, uint32_t klo) {
        uint32_t e32 = 0, pv = 0;
        S ev = 0;
        load(line * 64u, e32, ev, pv);
        take(((uint32_t)lane <= kc) && ((uint32_t)lane >= klo) && is_hit(e32, ev), pv, out_rsrc(k), k);
    };
    const uint32_t L0 = cur >> 6;        // the first line, walked at and below cur only
    const uint32_t lo_line = lo >> 6;    // the last line, walked at and above lo only
    take_clipped(L0, cur - L0 * 64u, (L0 == lo_line) ? lo - L0 * 64u : 0u);
    if (L0 > lo_line + 1u && td.nlev >= 2) {
        // the lines strictly between lo_line and L0 whose maximum is >= s, highest first
        const uint32_t hi_line = L0 - 1u;
        uint32_t g1 = hi_line & ~63u;    // group of 64 lines in hand
        uint64_t rem = group_mask(td, g1, s, lane) & mask_upto(hi_line - g1);
        while (true) {
            if (lo_line >= g1) rem &= ~mask_upto(lo_line - g1);   // lines up to lo_line are out
            if (rem == 0) {
                if (g1 <= lo_line + 1u) break;
                const uint32_t l = prev_line(td, g1, lo_line + 1u, s, lane);
                if (l == NONE) break;
                g1 = l & ~63u;
                rem = group_mask(td, g1, s, lane) & mask_upto(l - g1);   // lines of the group up to l
                continue;
            }
            // The next NL lines with hits, highest first: line g1 + b[j].  The group's arrays
            // are read through buffer resources based at its first entry (64 lines of 64), at
            // byte b * 64 * width + lane * width.  A round with fewer lines left repeats its
            // last line in the free slots, whose takes find no hit (valid false).
            const uint32_t nv = min((uint32_t)__popcll(rem), NL);
            uint32_t b[NL];
#pragma unroll
            for (uint32_t j = 0; j < NL; ++j) {
                if (j < nv) {
                    b[j] = 63u - (uint32_t)__builtin_clzll(rem);
                    rem &= ~(1ull << b[j]);
                } else {
                    b[j] = b[j - 1u];
                }
            }
            constexpr uint32_t EW = E32 ? 4u : (uint32_t)sizeof(S);   // bytes per end read
            const __amdgpu_buffer_rsrc_t re = __builtin_amdgcn_make_buffer_rsrc(
                E32 ? (void*)(ends32 + (size_t)g1 * 64u) : (void*)(ends + (size_t)g1 * 64u), (short)0, (int)(4096u * EW),
                0x00020000);
            const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
                (LONG && MODE != W_COUNT) ? (void*)(posmap + (size_t)g1 * 64u) : nullptr, (short)0, (int)(4096u * 4u),
                0x00020000);
            uint32_t e32[NL], pv[NL];
            S ev[NL];
#pragma unroll
            for (uint32_t j = 0; j < NL; ++j) {
                if (E32) {
                    e32[j] = __builtin_amdgcn_raw_buffer_load_b32(re, (uint32_t)lane * 4u, b[j] * 256u, 0);
                } else if (sizeof(S) == 8) {
                    const auto w = __builtin_amdgcn_raw_buffer_load_b64(re, (uint32_t)lane * 8u, b[j] * 512u, 0);
                    ev[j] = (S)(((uint64_t)(uint32_t)w[1] << 32) | (uint64_t)(uint32_t)w[0]);
                } else {
                    ev[j] = (S)__builtin_amdgcn_raw_buffer_load_b32(re, (uint32_t)lane * 4u, b[j] * 256u, 0);
                }
                if (MODE != W_COUNT)
                    pv[j] = LONG ? __builtin_amdgcn_raw_buffer_load_b32(rp, (uint32_t)lane * 4u, b[j] * 256u, 0)
                                 : (g1 + b[j]) * 64u + (uint32_t)lane;
            }
            // every line is taken, so no branch holds back a load
            const uint32_t k0 = k;
            const __amdgpu_buffer_rsrc_t rs = out_rsrc(k0);
            if (nv == NL) {
#pragma unroll
                for (uint32_t j = 0; j < NL; ++j) take(is_hit(e32[j], ev[j]), pv[j], rs, k0);
            } else {
#pragma unroll
                for (uint32_t j = 0; j < NL; ++j) take(j < nv && is_hit(e32[j], ev[j]), pv[j], rs, k0);
            }
        }
    }
    if (L0 > lo_line) take_clipped(lo_line, 63u, lo - lo_line * 64u);
    if (MODE == W_COUNT) {
        acc += k;
    } else {
        left -= k;
        acc += min(k, room);
    }
}

// Walks the hits at and below cur (see above).  acc: the count so far (count mode) or the
// next output offset (fill modes, writes below oend only); left: in fill modes the hits
// still expected, counted down.  With a long list (lv.len > 0) the positions below P0 = the
// first position starting at or after s - lv.len hold no hit but a long interval, so the
// walk reads the index's arrays down to P0 only and takes the rest from the long list,
// whose entries all lie below P0: the output order is unchanged.
template <typename S, int MODE>
__device__ __forceinline__ void wave_walk(const IndexView<S>& ix, const TreeDesc<S>& td, const UbTree<S>& ub,
                                          const LongView<S>& lv, int lane, S s, uint32_t cur, uint32_t top, bool lead,
                                          uint64_t& acc, uint64_t oend, uint32_t& left, uint32_t* __restrict__ out_idx,
                                          const int32_t* __restrict__ vals, int32_t* __restrict__ out_val) {
    // leading run (vector order): consecutive lines while every position down from cur is a hit
    while (MODE == W_VECTOR && lead && cur != NONE) {
        const uint32_t b0 = cur & ~63u;
        const uint32_t kc = cur - b0;
        const bool hit = ((uint32_t)lane <= kc) && (wave_end(ix, b0 + lane) >= s);
        const uint64_t m = __ballot(hit);
        const uint64_t lim = (kc == 63) ? ~0ull : ((2ull << kc) - 1ull);
        const uint64_t inv = ~m & lim;
        uint32_t lo = 0;
        if (inv == 0) {
            cur = b0 - 1u;   // b0 == 0 wraps to NONE
        } else {
            cur = b0 + 63u - (uint32_t)__builtin_clzll(inv);   // the first miss below the run
            lo = cur + 1u;
        }
        if (inv != 0 || cur == NONE) {
            left -= top - lo + 1u;
            const uint64_t c = min((uint64_t)(top - lo + 1u), oend - acc);
            for (uint64_t x = lane; x < c; x += WAVE) out_idx[acc + x] = lo + (uint32_t)x;
            acc += c;
            lead = false;
        }
    }
    if (cur == NONE) return;
    uint32_t P0 = 0;
    if (lv.len > 0) {
        constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
        // x = s - len - 1 (upper_bound(x) + 1 = the first position starting at >= s - len)
        if ((int64_t)s > (int64_t)SMIN + lv.len) P0 = wave_tree_ub(ub, (S)((int64_t)s - lv.len - 1), lane) + 1u;
    }
    if (sizeof(S) == 8 && ix.ends32)
        guided_walk<S, MODE, false, true, WAVE_LINES>(ix.ends, ix.ends32, ix.ends32_base, nullptr, td, lane, s, cur, P0,
                                                      acc, oend, left, out_idx, vals, out_val);
    else
        guided_walk<S, MODE, false, false, WAVE_LINES>(ix.ends, nullptr, 0, nullptr, td, lane, s, cur, P0, acc, oend,
                                                       left, out_idx, vals, out_val);
    if (P0 > 0) {
        const uint32_t J = lv.lrank[min(P0, cur + 1u)];   // long entries below that position
        if (J > 0 && sizeof(S) == 8 && lv.ends32)
            guided_walk<S, MODE, true, true, LONG_LINES>(lv.ends, lv.ends32, lv.ends32_base, lv.pos, lv.td, lane, s,
                                                         J - 1u, 0u, acc, oend, left, out_idx, vals, out_val);
        else if (J > 0)
            guided_walk<S, MODE, true, false, LONG_LINES>(lv.ends, nullptr, 0, lv.pos, lv.td, lane, s, J - 1u, 0u, acc,
                                                          oend, left, out_idx, vals, out_val);
    }
}

// Continuation of walks the lane walk parked (over its step budget): one wave per parked
// query.  The heavy list holds the parked queries roughly in end order; it is cut into
// HEAVY_CHUNKS contiguous chunks, each with its own counter, and workgroup w starts on chunk
// w % HEAVY_CHUNKS, then takes what is left of the others.  Workgroups are dispatched round
// robin over the chip's 8 XCDs, so with 8 chunks each XCD's L2 serves the queries of one
// stretch of the coordinates (the results never depend on which wave takes an entry).
#ifndef SI_HEAVY_CHUNKS
#define SI_HEAVY_CHUNKS 8
#endif
constexpr uint32_t HEAVY_CHUNKS = SI_HEAVY_CHUNKS;
static_assert(HEAVY_CHUNKS >= 1 && HEAVY_CHUNKS <= 8, "the plan holds 8 chunk counters");

// Clears the counters a walk launch uses (tiles, heavy list, mismatch flag, chunk counters).
__global__ void k_walk_counters_reset(uint32_t* __restrict__ c) {
    const uint32_t i = threadIdx.x;
    if (i < 4u || (i >= 8u && i < 8u + HEAVY_CHUNKS)) c[i] = 0u;
}

template <typename S, int MODE, bool MULTI = false>
__global__ __launch_bounds__(Q_THREADS) void k_walk_heavy(IndexView<S> ix_arg, TreeDesc<S> td_arg, UbTree<S> ub_arg,
                                                          LongView<S> lv, PlanView<S> pv,
                                                          uint32_t* __restrict__ counts,
                                                          uint64_t* __restrict__ counts64,
                                                          const uint64_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ out_idx,
                                                          const int32_t* __restrict__ vals_arg,
                                                          int32_t* __restrict__ out_val, uint64_t cap,
                                                          const SetView<S>* __restrict__ sets) {
    IndexView<S> ix = ix_arg;
    TreeDesc<S> td = td_arg;
    UbTree<S> ub = ub_arg;
    const int32_t* __restrict__ vals = vals_arg;
    const int lane = lane_id();
    const uint32_t nheavy = __builtin_amdgcn_readfirstlane(pv.counter[1]);
    if (nheavy == 0) return;   // nothing parked: no wave touches the shared counters
    const uint32_t chunk = (nheavy + HEAVY_CHUNKS - 1u) / HEAVY_CHUNKS;
    uint32_t c = blockIdx.x % HEAVY_CHUNKS, tried = 0;
    while (true) {
        uint32_t h = 0;
        if (lane == 0) h = atomicAdd(&pv.counter[8 + c], 1u);
        h = __builtin_amdgcn_readfirstlane(__shfl(h, 0, WAVE));
        const uint32_t beg = c * chunk, len = (beg < nheavy) ? min(chunk, nheavy - beg) : 0u;
        if (h >= len) {   // chunk c is exhausted: the next one
            if (++tried == HEAVY_CHUNKS) break;
            c = (c + 1u) % HEAVY_CHUNKS;
            continue;
        }
        const HeavyEntry he = pv.heavy[beg + h];
        const uint32_t k = __builtin_amdgcn_readfirstlane(he.k);
        if (MULTI) {   // the parked query's set
            const uint32_t g = __builtin_amdgcn_readfirstlane(pv.set[k]);
            ix = sets[g].v;
            td = sets[g].td;
            ub = sets[g].ub;
            vals = sets[g].values;
        }
        const uint32_t q = __builtin_amdgcn_readfirstlane(pv.q[k]);
        const S s = pv.s[k];
        const uint32_t cur = __builtin_amdgcn_readfirstlane(he.cur);
        const uint32_t top = __builtin_amdgcn_readfirstlane(he.top);
        const bool lead = __builtin_amdgcn_readfirstlane(he.lead) != 0;
        uint32_t left = __builtin_amdgcn_readfirstlane(he.left);   // fill modes: hits still expected
        uint64_t acc = he.acc;
        const uint64_t oend = (MODE != W_COUNT) ? min(offsets[q + 1], cap) : 0;
        wave_walk<S, MODE>(ix, td, ub, lv, lane, s, cur, top, lead, acc, oend, left, out_idx, vals, out_val);
        if (MODE == W_COUNT) {
            // every lane stores the same value (no lane-dependent branch in the loop)
            if (counts) counts[q] = (uint32_t)acc;
            if (counts64) counts64[q] = acc;
        } else if (left != 0u) {
            atomicOr(&pv.counter[3], 1u);   // wave-uniform test: the walk's hits differ from the offsets
        }
    }
}

// ============================================================================
// One query, one wave (the single-query host calls)
// ============================================================================
enum OneMode { ONE_UB = 4, ONE_ANY = 5 };

// out (pinned host memory mapped into the device): out[0] = the count (or upper_bound /
// has_overlaps), then up to cap results as uint32 (positions) or int32 (payload) from
// out + 2.  The count is exact whatever cap is.
template <typename S, int MODE>
__global__ __launch_bounds__(WAVE) void k_one_query(IndexView<S> ix, UbTree<S> ub, UbTree<S> eb, TreeDesc<S> td,
                                                    LongView<S> lv,
                                                    const int32_t* __restrict__ vals, uint32_t rank_exact, S s, S e,
                                                    uint64_t* __restrict__ out, uint64_t cap) {
    const int lane = lane_id();
    const uint32_t idx = ix.n ? wave_tree_ub(ub, e, lane) : NONE;
    if (MODE == ONE_UB) {
        if (lane == 0) out[0] = idx;
        return;
    }
    if (MODE == ONE_ANY) {   // superintervals.hpp:865-871: the last interval starting at or before e
        if (lane == 0) out[0] = (idx != NONE && s <= ix.ends[idx]) ? 1u : 0u;
        return;
    }
    if (idx == NONE) {
        if (lane == 0) out[0] = 0;
        return;
    }
    constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
    if (MODE == W_COUNT && rank_exact && s <= e) {   // the rank form
        const uint32_t below = (s == SMIN) ? 0u : wave_tree_ub(eb, (S)(s - 1), lane) + 1u;
        if (lane == 0) out[0] = idx + 1u - below;
        return;
    }
    uint32_t* res = reinterpret_cast<uint32_t*>(out + 2);
    int32_t* resv = reinterpret_cast<int32_t*>(out + 2);
    uint64_t acc = 0, found = 0;
    uint32_t cur = idx, left = 0;
    bool lead = (MODE == W_VECTOR);
    if (rank_exact) {
        // [lower_bound(s), idx]: every position there starts inside the query, all hits
        const uint32_t lb = (s == SMIN) ? 0u : wave_tree_ub(ub, (S)(s - 1), lane) + 1u;
        if (lb <= cur) {
            const uint32_t c = cur - lb + 1u;
            if (MODE == W_COUNT) {
                acc += c;
            } else if (MODE == W_DESC || MODE == W_VALUES) {
                found += c;
                const uint64_t w = min((uint64_t)c, cap - acc);
                for (uint64_t x = lane; x < w; x += WAVE) {
                    const uint32_t p = cur - (uint32_t)x;
                    if (MODE == W_VALUES) resv[acc + x] = vals[p];
                    else res[acc + x] = p;
                }
                acc += w;
            }
            cur = lb - 1u;   // lb == 0 wraps to NONE
            if (MODE == W_VECTOR && cur == NONE) {
                // the leading run reaches position 0: written ascending here (otherwise the
                // walk below writes it, with the run, once its bottom is found)
                found += idx + 1u;
                const uint64_t w = min((uint64_t)(idx + 1u), cap - acc);
                for (uint64_t x = lane; x < w; x += WAVE) res[acc + x] = (uint32_t)x;
                acc += w;
                lead = false;
            }
        }
    }
    // the rest by the wave walk; left counts down from 0, so -left is what it found
    wave_walk<S, MODE>(ix, td, ub, lv, lane, s, cur, idx, lead, acc, (MODE == W_COUNT) ? 0 : cap, left, res, vals, resv);
    const uint64_t total = (MODE == W_COUNT) ? acc : found + (uint64_t)(uint32_t)(0u - left);
    if (lane == 0) out[0] = total;
}

// ============================================================================
// Per-lane kernels over unsorted queries (single-purpose queries)
// ============================================================================
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_upper_bound(IndexView<S> ix, const S* __restrict__ xs,
                                                           size_t n, uint32_t* __restrict__ out) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = upper_bound_pos(ix.starts, ix.n, xs[q]);
}

// Ranks of N keys in Eytzinger tops of h levels (LDS), together: the N reads of a level are
// issued at once.  Each descent ends at k = 2^h + r, r = the number of top keys <= x[i];
// padding keys are max(S) and a key equal to max(S) may count padding, so r is clamped to
// the level's size.
template <typename S, int N>
__device__ __forceinline__ void top_ranks(const S* (&tp)[N], uint32_t h, const uint32_t (&size)[N], const S (&x)[N],
                                          uint32_t (&r)[N]) {
    uint32_t k[N];
#pragma unroll
    for (int i = 0; i < N; ++i) k[i] = 1u;
    for (uint32_t lvl = 0; lvl < h; ++lvl) {
        S t[N];
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = tp[i][k[i] - 1u];
#pragma unroll
        for (int i = 0; i < N; ++i) k[i] = 2u * k[i] + (t[i] <= x[i] ? 1u : 0u);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = min(k[i] - (1u << h), size[i]);
}

// Last position p in [b[i], hi[i]] (one node: at most 64 keys) with v[i][p] <= x[i], or
// b[i] - 1, for N searches together: fixed halving steps 32, 16, .., 1, so the N probes of
// a step are issued at once, and the node's first key is read beside the first step.
template <typename S, int N>
__device__ __forceinline__ void node_last_le(const S* (&v)[N], const uint32_t (&b)[N], const uint32_t (&hi)[N],
                                             const S (&x)[N], int64_t (&out)[N]) {
    uint32_t p[N];
    bool first_le[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        p[i] = b[i];
        first_le[i] = v[i][b[i]] <= x[i];
    }
#pragma unroll
    for (uint32_t step = 32u; step >= 1u; step >>= 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const uint32_t q = p[i] + step;
            if (q <= hi[i] && v[i][q] <= x[i]) p[i] = q;
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (p[i] == b[i] && !first_le[i]) ? (int64_t)b[i] - 1 : (int64_t)p[i];
}

// Last r in [lo, hi] with v[r] <= x, or lo - 1 (branchless halving over len elements: for
// one key alone this reads no more than it must).
template <typename S>
__device__ __forceinline__ int64_t last_le_in(const S* __restrict__ v, int64_t lo, int64_t hi, S x) {
    const S* base = v + lo;
    uint32_t len = (uint32_t)(hi - lo + 1);
    while (len > 1) {
        const uint32_t half = len >> 1;
        base = (base[half] <= x) ? base + half : base;
        len -= half;
    }
    return (int64_t)(base - v) - ((*base <= x) ? 0 : 1);
}

// Last position with key <= x, or NONE, through the tree (s_top = its top level in LDS).
template <typename S>
__device__ __forceinline__ uint32_t ub_tree_pos(const UbTree<S>& t, const S* s_top, uint32_t h, S x) {
    const S* tp[1] = {s_top};
    const uint32_t tsz[1] = {t.size[t.nlv]};
    const S xx[1] = {x};
    uint32_t r[1];
    top_ranks<S, 1>(tp, h, tsz, xx, r);
    int64_t j = (int64_t)r[0] - 1;
    for (int l = t.nlv - 1; l >= 0 && j >= 0; --l) {
        const int64_t b = j * 64, e = std::min<int64_t>(b + 63, (int64_t)t.size[l] - 1);
        j = last_le_in(t.lv[l], b, e, x);
    }
    return (j < 0) ? NONE : (uint32_t)j;
}

// upper_bound through the search tree: descend the Eytzinger top (LDS) to r = the number of
// top-level keys <= x, then at each level below search the one 64-key node under the entry
// found (its first key is that entry, so it is <= x and the node holds the answer for the
// level).  Persistent blocks stage the top once.  Same result as upper_bound_pos
// (superintervals.hpp:501-516).
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_upper_bound_tree(UbTree<S> t, const S* __restrict__ top, uint32_t h,
                                                                const S* __restrict__ xs, size_t nx,
                                                                uint32_t* __restrict__ out) {
    extern __shared__ __align__(16) char s_mem[];
    S* s_top = reinterpret_cast<S*>(s_mem);
    const uint32_t T = (1u << h) - 1u;
    for (uint32_t k = threadIdx.x; k < T; k += Q_THREADS) s_top[k] = top[k];
    __syncthreads();
    for (size_t q0 = (size_t)blockIdx.x * Q_THREADS; q0 < nx; q0 += (size_t)gridDim.x * Q_THREADS) {
        const size_t q = q0 + threadIdx.x;
        if (q >= nx) continue;
        out[q] = ub_tree_pos(t, s_top, h, xs[q]);
    }
}

// Rank form of count for the planned queries (as k_count_rank) with #{ends < s} found
// through the search tree over the sorted ends: #{ends < s} = #{ends <= s - 1} =
// upper_bound_pos(s - 1) + 1 (NONE + 1 wraps to 0), and 0 when s is the smallest value.
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_count_rank_tree(IndexView<S> ix, UbTree<S> t,
                                                               const S* __restrict__ top, uint32_t h,
                                                               PlanView<S> pv, size_t nq,
                                                               uint32_t* __restrict__ counts,
                                                               uint64_t* __restrict__ counts64) {
    extern __shared__ __align__(16) char s_mem[];
    S* s_top = reinterpret_cast<S*>(s_mem);
    const uint32_t T = (1u << h) - 1u;
    const bool planned = __builtin_amdgcn_readfirstlane(pv.counter[5]) != 0u;   // the plan holds the counts
    if (!planned)
        for (uint32_t k = threadIdx.x; k < T; k += Q_THREADS) s_top[k] = top[k];
    __syncthreads();
    constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
    for (size_t k0 = (size_t)blockIdx.x * Q_THREADS; k0 < nq; k0 += (size_t)gridDim.x * Q_THREADS) {
        const size_t k = k0 + threadIdx.x;
        if (k >= nq) continue;
        const uint32_t i = pv.idx[k];
        const S s = pv.s[k], e = pv.e[k];
        uint32_t c = 0;
        if (planned) {
            c = pv.cnt[k];
        } else if (i != NONE) {
            if (s <= e) {
                const uint32_t below = (s == SMIN) ? 0u : ub_tree_pos(t, s_top, h, (S)(s - 1)) + 1u;
                c = i + 1u - below;
            } else {
                c = walk_count(ix, i, s);
            }
        }
        const uint32_t q = pv.q[k];
        if (counts) counts[q] = c;
        if (counts64) counts64[q] = c;
    }
}

template <typename S>
UbTree<S> ub_tree_of(const KeyTree& kt, const S* keys, size_t n) {
    UbTree<S> t;
    t.nlv = kt.nlv;
    t.lv[0] = keys;
    t.size[0] = (uint32_t)n;
    for (int l = 1; l < MAX_LEVELS; ++l) {
        t.lv[l] = (l <= kt.nlv) ? (const S*)kt.d_lv + kt.lv_off[l] : nullptr;
        t.size[l] = (l <= kt.nlv) ? kt.lv_size[l] : 0u;
    }
    return t;
}

// Persistent grid of the tree kernels: enough blocks to fill the GPU, each stages the top once.
inline unsigned tree_grid(size_t n) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cus = std::max(cus, 1);
    }
    return (unsigned)std::min<size_t>(grid_for(n, Q_THREADS), (size_t)cus * 8);
}

template <typename S>
int launch_ub_tree(const siIndex* ix, const S* xs, size_t n, uint32_t* out, hipStream_t st) {
    const KeyTree& kt = ix->ub;
    const size_t lds = (size_t)((1u << kt.top_h) - 1u) * ix->cb;
    k_upper_bound_tree<S><<<tree_grid(n), Q_THREADS, lds, st>>>(ub_tree_of<S>(kt, (const S*)ix->d_starts, ix->n),
                                                                 (const S*)kt.d_top, kt.top_h, xs, n, out);
    LAUNCH_CHECK();
    return SI_OK;
}

// The plan's bounds and counts: idx[k] = upper_bound(e[k]); with LB, lb[k] = the first
// position whose start is >= s[k] (upper_bound(s - 1) + 1, 0 when s is the smallest value);
// with RANK, cnt[k] = the rank-form count (upper_bound + 1 - #{ends < s}, #{ends < s} =
// upper_bound(ends_sorted, s - 1) + 1), or the reference walk's count for s > e.  One thread
// per sorted query; its searches descend the starts tree (twice) and the sorted-ends tree
// level by level together, so their reads overlap.  (Reading each node's 8-key boundaries
// in one round and then one 8-key segment, in place of halving, was slower: neighbouring
// sorted queries share nodes, so the halving reads mostly hit L1, and the wider rounds cost
// registers and instructions.)
constexpr int PB_THREADS = 512;

template <typename S, bool LB, bool RANK>
__global__ __launch_bounds__(PB_THREADS) void k_plan_bounds_tree(IndexView<S> ix, UbTree<S> ts,
                                                                 const S* __restrict__ tops, UbTree<S> te,
                                                                 const S* __restrict__ tope, uint32_t h,
                                                                 PlanView<S> pv, size_t nq) {
    constexpr int N = 1 + (LB ? 1 : 0) + (RANK ? 1 : 0);
    constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
    extern __shared__ __align__(16) char s_mem[];
    const uint32_t T = (1u << h) - 1u;
    S* s_ts = reinterpret_cast<S*>(s_mem);
    S* s_te = s_ts + T;
    for (uint32_t k = threadIdx.x; k < T; k += PB_THREADS) {
        s_ts[k] = tops[k];
        if (RANK) s_te[k] = tope[k];
    }
    __syncthreads();
    for (size_t k0 = (size_t)blockIdx.x * PB_THREADS; k0 < nq; k0 += (size_t)gridDim.x * PB_THREADS) {
        const size_t k = k0 + threadIdx.x;
        if (k >= nq) continue;
        const S e = pv.e[k], s = pv.s[k];
        const bool s1 = s != SMIN;   // s - 1 exists
        const S x1 = s1 ? (S)(s - 1) : s;
        // searches: 0 = upper_bound(e), then lower bound (starts at s - 1), then #{ends < s}
        const S* v[N];
        UbTree<S> const* tr[N];
        const S* tp[N];
        S x[N];
        int n = 0;
        tr[n] = &ts; tp[n] = s_ts; x[n] = e; ++n;
        if (LB) { tr[n] = &ts; tp[n] = s_ts; x[n] = x1; ++n; }
        if (RANK) { tr[n] = &te; tp[n] = s_te; x[n] = x1; ++n; }
        uint32_t tsz[N], r0[N];
#pragma unroll
        for (int i = 0; i < N; ++i) tsz[i] = tr[i]->size[tr[i]->nlv];
        top_ranks<S, N>(tp, h, tsz, x, r0);
        int64_t j[N];
#pragma unroll
        for (int i = 0; i < N; ++i) j[i] = (int64_t)r0[i] - 1;
        for (int l = ts.nlv - 1; l >= 0; --l) {
            uint32_t b[N], hi[N];
            int64_t out[N];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                v[i] = tr[i]->lv[l];
                b[i] = (uint32_t)(j[i] < 0 ? 0 : j[i]) * 64u;   // a search with no key <= x reads node 0, unused
                hi[i] = min(b[i] + 63u, tr[i]->size[l] - 1u);
            }
            node_last_le<S, N>(v, b, hi, x, out);
#pragma unroll
            for (int i = 0; i < N; ++i) j[i] = (j[i] < 0) ? j[i] : out[i];
        }
        const uint32_t ub = (j[0] < 0) ? NONE : (uint32_t)j[0];
        pv.idx[k] = ub;
        if (LB) pv.lb[k] = (!s1 || j[LB ? 1 : 0] < 0) ? 0u : (uint32_t)j[LB ? 1 : 0] + 1u;
        if (RANK) {
            const int64_t jr = j[N - 1];
            uint32_t c = 0;
            if (ub != NONE) {
                if (s <= e) c = ub + 1u - ((!s1 || jr < 0) ? 0u : (uint32_t)jr + 1u);
                else c = walk_count(ix, ub, s);
            }
            pv.cnt[k] = c;
        }
    }
}

template <typename S>
int launch_plan_bounds(const siIndex* ix, PlanView<S> pv, size_t nq, bool with_lb, bool with_rank, hipStream_t st) {
    const KeyTree& kt = ix->ub;
    const KeyTree& ke = ix->eb;
    const uint32_t T = (1u << kt.top_h) - 1u;
    const size_t lds = (size_t)T * ix->cb * (with_rank ? 2u : 1u);
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cus = std::max(cus, 1);
    }
    const UbTree<S> ts = ub_tree_of<S>(kt, (const S*)ix->d_starts, ix->n);
    const UbTree<S> te = with_rank ? ub_tree_of<S>(ke, (const S*)ix->d_ends_sorted, ix->n) : ts;
    const S* tope = with_rank ? (const S*)ke.d_top : nullptr;
    // persistent: the resident blocks (registers and the staged tops allow) stage the tops once
#define SI_PLAN_BOUNDS(LBV, RKV)                                                                                   \
    do {                                                                                                           \
        int per = 0;                                                                                               \
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_plan_bounds_tree<S, LBV, RKV>, PB_THREADS, lds);     \
        const unsigned grid = (unsigned)std::min<size_t>(grid_for(nq, PB_THREADS), (size_t)cus * std::max(per, 1)); \
        k_plan_bounds_tree<S, LBV, RKV><<<grid, PB_THREADS, lds, st>>>(ix->view<S>(), ts, (const S*)kt.d_top, te,   \
                                                                       tope, kt.top_h, pv, nq);                     \
    } while (0)
    if (with_lb && with_rank) SI_PLAN_BOUNDS(true, true);
    else if (with_lb) SI_PLAN_BOUNDS(true, false);
    else if (with_rank) SI_PLAN_BOUNDS(false, true);
    else SI_PLAN_BOUNDS(false, false);
#undef SI_PLAN_BOUNDS
    LAUNCH_CHECK();
    return SI_OK;
}

template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_has_overlaps(IndexView<S> ix, const S* __restrict__ qs,
                                                            const S* __restrict__ qe, size_t nq,
                                                            uint8_t* __restrict__ out) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t i = upper_bound_pos(ix.starts, ix.n, qe[q]);
    out[q] = (i != NONE && qs[q] <= ix.ends[i]) ? 1 : 0;
}

// Reference walk, element by element, counting its reads (c_superintervals.h:610-643).
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_walk_stats(IndexView<S> ix, const S* __restrict__ qs,
                                                          const S* __restrict__ qe, size_t nq,
                                                          unsigned long long* __restrict__ totals) {
    __shared__ unsigned long long s_t[4];
    if (threadIdx.x < 4) s_t[threadIdx.x] = 0;
    __syncthreads();
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long h = 0, v = 0, j = 0, f = 0;
    if (q < nq) {
        const S s = qs[q];
        uint32_t i = upper_bound_pos(ix.starts, ix.n, qe[q]);
        f = (i != NONE) ? 1 : 0;
        while (i != NONE) {
            ++v;
            if (s <= ix.ends[i]) { ++h; --i; }
            else { i = ix.branch[i]; ++j; }
        }
    }
    h = wave_incl_scan(h); v = wave_incl_scan(v); j = wave_incl_scan(j); f = wave_incl_scan(f);
    if (lane_id() == WAVE - 1) {
        atomicAdd(&s_t[0], h); atomicAdd(&s_t[1], v); atomicAdd(&s_t[2], j); atomicAdd(&s_t[3], f);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_t[threadIdx.x]) atomicAdd(&totals[threadIdx.x], s_t[threadIdx.x]);
}

// Keys (and payload) of result positions: out_s[k] = starts[idx[k]], out_e[k] =
// ends[idx[k]], out_v[k] = values[idx[k]] (each output may be NULL).
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_gather_keys(IndexView<S> ix, const int32_t* __restrict__ vals,
                                                           const uint32_t* __restrict__ idx, size_t n,
                                                           S* __restrict__ out_s, S* __restrict__ out_e,
                                                           int32_t* __restrict__ out_v) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = idx[k];
    if (i >= ix.n) return;   // not a position of this index: left as is
    if (out_s) out_s[k] = ix.starts[i];
    if (out_e) out_e[k] = ix.ends[i];
    if (out_v) out_v[k] = vals[i];
}

__global__ __launch_bounds__(Q_THREADS) void k_coverage_i32(IndexView<int32_t> ix,
                                                            const int32_t* __restrict__ qs,
                                                            const int32_t* __restrict__ qe, size_t nq,
                                                            uint32_t* __restrict__ cnt,
                                                            int32_t* __restrict__ cov) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int32_t s = qs[q], e = qe[q];
    uint32_t i = upper_bound_pos(ix.starts, ix.n, e);
    uint32_t c = 0, acc = 0;   // unsigned accumulation = int32 wraparound
    while (i != NONE) {
        int32_t en = ix.ends[i];
        if (s <= en) {
            int32_t hi = en < e ? en : e;
            int32_t st = ix.starts[i];
            int32_t lo = st > s ? st : s;
            ++c;
            acc += (uint32_t)hi - (uint32_t)lo;
            --i;
        } else {
            i = ix.branch[i];
        }
    }
    if (cnt) cnt[q] = c;
    if (cov) cov[q] = (int32_t)acc;
}

// ============================================================================
// Launch helpers
// ============================================================================
#define DISPATCH_S(ix, ...)                                                  \
    ((ix)->cb == 4 ? [&] { using S = int32_t; return __VA_ARGS__; }()        \
                   : [&] { using S = int64_t; return __VA_ARGS__; }())


template <typename S>
int launch_rank(const siIndex* ix, void* plan, size_t nq, uint32_t* counts, uint64_t* counts64,
                hipStream_t st) {
    if (nq == 0) return SI_OK;
    if (g_tune.ub_tree && ix->eb.d_top && ix->n > 0) {
        const KeyTree& kt = ix->eb;
        const size_t lds = (size_t)((1u << kt.top_h) - 1u) * ix->cb;
        k_count_rank_tree<S><<<tree_grid(nq), Q_THREADS, lds, st>>>(
            ix->view<S>(), ub_tree_of<S>(kt, (const S*)ix->d_ends_sorted, ix->n), (const S*)kt.d_top, kt.top_h,
            plan_view<S>(plan, nq), nq, counts, counts64);
        LAUNCH_CHECK();
        return SI_OK;
    }
    k_count_rank<S><<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>(
        ix->view<S>(), (const S*)ix->d_ends_sorted, plan_view<S>(plan, nq), nq, counts, counts64);
    LAUNCH_CHECK();
    return SI_OK;
}

// Launch shape of the persistent walk: tile size, number of tiles and resident workgroups.
// A batch too small to give every resident workgroup a full tile gets smaller tiles (at
// least one 64-query pool each), so its queries spread over all CUs instead of queueing in
// the lanes of a few workgroups.
template <typename S, int MODE>
void walk_shape(size_t nq, uint32_t* tile, uint32_t* ntiles, unsigned* grid) {
    static int cus = 0, per = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_walk<S, MODE>, T_THREADS, 0);
        cus = std::max(cus, 1);
        per = std::max(per, 1);
    }
    const int want = (MODE == W_COUNT) ? g_tune.count_wgs : g_tune.fill_wgs;
    const int wgs = (want > 0) ? std::min(want, per) : per;
    const size_t resident = (size_t)cus * (size_t)wgs;
    size_t t = g_tune.tile;
    if (nq < t * resident) t = std::max<size_t>(POOL, (nq + resident - 1) / resident);
    t = (t + POOL - 1) / POOL * POOL;
    *tile = (uint32_t)std::min<size_t>(t, g_tune.tile);
    *ntiles = (uint32_t)((nq + *tile - 1) / *tile);
    *grid = (unsigned)std::max<size_t>(1, std::min<size_t>(resident, *ntiles));
}

template <typename S, int MODE>
int launch_walk(const siIndex* ix, void* plan, size_t nq, uint32_t* counts, uint64_t* counts64,
                const uint64_t* offsets, uint32_t* out_idx, int32_t* out_val, hipStream_t st,
                uint64_t cap = ~0ull) {
    if (nq == 0) return SI_OK;
    PlanView<S> pv = plan_view<S>(plan, nq);
    uint32_t tile = 0, ntiles = 0;
    unsigned grid = 0;
    walk_shape<S, MODE>(nq, &tile, &ntiles, &grid);
    k_walk_counters_reset<<<1, 64, 0, st>>>(pv.counter);
    LAUNCH_CHECK();
    k_walk<S, MODE><<<grid, T_THREADS, 0, st>>>(
        ix->view<S>(), pv, (uint32_t)nq, tile, ntiles, g_tune.lane_steps, range_skip_on(ix), pv.counter, counts,
        counts64, offsets, out_idx, ix->d_values, out_val, cap, nullptr, MultiTiles{nullptr, nullptr, nullptr, 0u});
    LAUNCH_CHECK();
    // parked (heavy) walks: one wave each, resident grid, exits at once when none were parked
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cus = std::max(cus, 1);
    }
    IndexView<S> hv = ix->view<S>();
    LongView<S> lv = long_view_of<S>(ix);
    if (!g_tune.heavy_ends32) { hv.ends32 = nullptr; lv.ends32 = nullptr; }
    k_walk_heavy<S, MODE><<<(unsigned)cus * 8u, Q_THREADS, 0, st>>>(
        hv, ix->tree<S>(), ub_tree_of<S>(ix->ub, (const S*)ix->d_starts, ix->n), lv, pv, counts, counts64, offsets,
        out_idx, ix->d_values, out_val, cap, nullptr);
    LAUNCH_CHECK();
    return SI_OK;
}

// The rank form counts (upper_bound + 1) - #{ends < s}, which assumes that every interval
// ending before s starts before e: true unless the index holds an interval with start > end.
inline bool rank_exact(const siIndex* ix) { return ix->n_inverted == 0; }

int check_plan(const siIndex* ix, const void* plan, size_t plan_bytes, size_t nq) {
    if (!plan) return si_set_error(SI_ERR_ARG, "null plan");
    if (nq >= (size_t)0x7FFFFFFF) return si_set_error(SI_ERR_ARG, "batch of %zu queries exceeds 2^31-1", nq);
    const size_t need = PlanLayout(ix->cb, nq).total;
    if (plan_bytes < need) return si_set_error(SI_ERR_ARG, "plan buffer too small: %zu < %zu bytes", plan_bytes, need);
    return SI_OK;
}

// Stream-ordered scratch plan for the unplanned entry points.
struct ScopedPlan {
    void* p = nullptr;
    hipStream_t st;
    explicit ScopedPlan(hipStream_t s) : st(s) {}
    int alloc(size_t bytes) {
        hipError_t e = hipMallocAsync(&p, bytes, st);
        if (e != hipSuccess) { p = nullptr; return si_set_error(SI_ERR_NOMEM, "hipMallocAsync(%zu): %s", bytes, hipGetErrorString(e)); }
        return SI_OK;
    }
    ~ScopedPlan() { if (p) (void)hipFreeAsync(p, st); }
};

}  // namespace

// ============================================================================
// Tuning
// ============================================================================
int si_tune_learned_shrink() { return g_tune.learned_shrink; }
int64_t si_tune_long_len() { return g_tune.long_len; }

SI_EXPORT int siTuneSet(const char* name, int64_t value) {
    if (!name) return si_set_error(SI_ERR_ARG, "null tuning name");
    const std::string k(name);
    if (k == "tile") {
        if (value < 64 || value > (1 << 20)) return si_set_error(SI_ERR_ARG, "tile must be in [64, 2^20]");
        g_tune.tile = (uint32_t)value;
    } else if (k == "count_wgs") {
        if (value < 0 || value > 64) return si_set_error(SI_ERR_ARG, "count_wgs must be in [0, 64]");
        g_tune.count_wgs = (int)value;
    } else if (k == "fill_wgs") {
        if (value < 0 || value > 64) return si_set_error(SI_ERR_ARG, "fill_wgs must be in [0, 64]");
        g_tune.fill_wgs = (int)value;
    } else if (k == "lane_steps") {
        if (value < 1 || value > (1 << 30)) return si_set_error(SI_ERR_ARG, "lane_steps must be in [1, 2^30]");
        g_tune.lane_steps = (uint32_t)value;
    } else if (k == "ub_tree") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "ub_tree must be 0 or 1");
        g_tune.ub_tree = (int)value;
    } else if (k == "plan_sync") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "plan_sync must be 0 or 1");
        g_tune.plan_sync = (int)value;
    } else if (k == "heavy_ends32") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "heavy_ends32 must be 0 or 1");
        g_tune.heavy_ends32 = (int)value;
    } else if (k == "long_walk") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "long_walk must be 0 or 1");
        g_tune.long_walk = (int)value;
    } else if (k == "long_len") {
        g_tune.long_len = value;
    } else if (k == "ub_learned") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "ub_learned must be 0 or 1");
        g_tune.ub_learned = (int)value;
    } else if (k == "learned_shrink") {
        if (value < 0 || value > (1 << 20)) return si_set_error(SI_ERR_ARG, "learned_shrink must be in [0, 2^20]");
        g_tune.learned_shrink = (int)value;
    } else if (k == "range_skip") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "range_skip must be 0 or 1");
        g_tune.range_skip = (int)value;
    } else if (k == "plan_skip_low") {
        if (value < 0 || value > 2) return si_set_error(SI_ERR_ARG, "plan_skip_low must be in [0, 2]");
        g_tune.plan_skip_low = (int)value;
    } else if (k == "count_rank" || k == "offsets_rank") {
        if (value != 0 && value != 1) return si_set_error(SI_ERR_ARG, "%s must be 0 or 1", name);
        (k == "count_rank" ? g_tune.count_rank : g_tune.offsets_rank) = (int)value;
    } else {
        return si_set_error(SI_ERR_ARG, "unknown tuning parameter '%s'", name);
    }
    return SI_OK;
}

SI_EXPORT int64_t siTuneGet(const char* name) {
    if (!name) return -1;
    const std::string k(name);
    if (k == "tile") return g_tune.tile;
    if (k == "count_wgs") return g_tune.count_wgs;
    if (k == "fill_wgs") return g_tune.fill_wgs;
    if (k == "lane_steps") return g_tune.lane_steps;
    if (k == "plan_sync") return g_tune.plan_sync;
    if (k == "ub_tree") return g_tune.ub_tree;
    if (k == "plan_skip_low") return g_tune.plan_skip_low;
    if (k == "heavy_ends32") return g_tune.heavy_ends32;
    if (k == "range_skip") return g_tune.range_skip;
    if (k == "ub_learned") return g_tune.ub_learned;
    if (k == "long_walk") return g_tune.long_walk;
    if (k == "long_len") return g_tune.long_len;
    if (k == "learned_shrink") return g_tune.learned_shrink;
    if (k == "count_rank") return g_tune.count_rank;
    if (k == "offsets_rank") return g_tune.offsets_rank;
    return -1;
}

// ============================================================================
// Planned batch entry points
// ============================================================================
SI_EXPORT size_t siPlanBytes(const siIndex* ix, size_t nq) {
    return PlanLayout(ix ? ix->cb : 8, nq).total;
}

SI_EXPORT int siPlanQueries(const siIndex* ix, const void* qs, const void* qe, size_t nq, void* plan,
                            size_t plan_bytes, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if ((rc = check_plan(ix, plan, plan_bytes, nq))) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe) return si_set_error(SI_ERR_ARG, "null queries");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, plan_impl<S>(ix, (const S*)qs, (const S*)qe, nq, plan, st, g_tune.plan_sync != 0));
}

SI_EXPORT int siCountPlanned(const siIndex* ix, void* plan, size_t nq, uint32_t* counts, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!plan) return si_set_error(SI_ERR_ARG, "null plan");
    if (nq && !counts) return si_set_error(SI_ERR_ARG, "null counts");
    hipStream_t st = (hipStream_t)stream;
    if (g_tune.count_rank && rank_exact(ix))
        return DISPATCH_S(ix, (launch_rank<S>(ix, plan, nq, counts, nullptr, st)));
    return DISPATCH_S(ix, (launch_walk<S, W_COUNT>(ix, plan, nq, counts, nullptr, nullptr, nullptr, nullptr, st)));
}

SI_EXPORT int siSearchOffsetsPlanned(const siIndex* ix, void* plan, size_t nq, uint32_t* counts,
                                     uint64_t* offsets, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!plan || !offsets) return si_set_error(SI_ERR_ARG, "null plan/offsets");
    hipStream_t st = (hipStream_t)stream;
    const PlanLayout L(ix->cb, nq);
    uint64_t* ws = (uint64_t*)((char*)plan + L.off_scratch);
    if (g_tune.offsets_rank && rank_exact(ix)) {
        // 32-bit counts (the caller's, or the heavy-list area of the plan, which the rank
        // form does not use), then one scan to 64-bit offsets: half the scattered bytes
        uint32_t* c32 = counts ? counts : (uint32_t*)((char*)plan + L.off_heavy);
        rc = DISPATCH_S(ix, (launch_rank<S>(ix, plan, nq, c32, nullptr, st)));
        if (rc) return rc;
        return device_excl_scan<uint32_t, uint64_t>(c32, nq, offsets, offsets + nq, ws, st);
    }
    rc = DISPATCH_S(ix, (launch_walk<S, W_COUNT>(ix, plan, nq, counts, offsets, nullptr, nullptr, nullptr, st)));
    if (rc) return rc;
    return device_excl_scan<uint64_t, uint64_t>(offsets, nq, offsets, offsets + nq, ws, st);
}

SI_EXPORT int siSearchFillPlannedBounded(const siIndex* ix, void* plan, size_t nq, int order,
                                         const uint64_t* offsets, uint32_t* idx, uint64_t capacity,
                                         void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (nq && (!plan || !offsets || (!idx && capacity))) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (order == SI_ORDER_DESC)
        return DISPATCH_S(ix, (launch_walk<S, W_DESC>(ix, plan, nq, nullptr, nullptr, offsets, idx, nullptr, st,
                                                      capacity)));
    return DISPATCH_S(ix, (launch_walk<S, W_VECTOR>(ix, plan, nq, nullptr, nullptr, offsets, idx, nullptr, st,
                                                    capacity)));
}

SI_EXPORT int siPlanCheck(const siIndex* ix, const void* plan, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!plan) return si_set_error(SI_ERR_ARG, "null plan");
    hipStream_t st = (hipStream_t)stream;
    uint32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, (const uint32_t*)plan + 3, sizeof(flag), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (flag) return si_set_error(SI_ERR_MISMATCH, "a fill pass found a different number of hits than its offsets give");
    return SI_OK;
}

SI_EXPORT int siSearchFillPlanned(const siIndex* ix, void* plan, size_t nq, int order,
                                  const uint64_t* offsets, uint32_t* idx, void* stream) {
    if (nq && !idx) return si_set_error(SI_ERR_ARG, "null buffer");
    return siSearchFillPlannedBounded(ix, plan, nq, order, offsets, idx, ~0ull, stream);
}

SI_EXPORT int siSearchValuesFillPlannedBounded(const siIndex* ix, void* plan, size_t nq, const uint64_t* offsets,
                                               int32_t* values, uint64_t capacity, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq && (!plan || !offsets || (!values && capacity))) return si_set_error(SI_ERR_ARG, "null buffer");
    if (!ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, (launch_walk<S, W_VALUES>(ix, plan, nq, nullptr, nullptr, offsets, nullptr, values, st,
                                                    capacity)));
}

SI_EXPORT int siSearchValuesFillPlanned(const siIndex* ix, void* plan, size_t nq, const uint64_t* offsets,
                                        int32_t* values, void* stream) {
    if (nq && !values) return si_set_error(SI_ERR_ARG, "null buffer");
    return siSearchValuesFillPlannedBounded(ix, plan, nq, offsets, values, ~0ull, stream);
}

// ============================================================================
// Unplanned batch entry points (plan into stream-ordered scratch, then run)
// ============================================================================
SI_EXPORT int siCount(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts,
                      void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !counts) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    ScopedPlan P(st);
    const size_t pb = siPlanBytes(ix, nq);
    if ((rc = P.alloc(pb))) return rc;
    if ((rc = siPlanQueries(ix, qs, qe, nq, P.p, pb, stream))) return rc;
    return siCountPlanned(ix, P.p, nq, counts, stream);
}

SI_EXPORT size_t siSearchWorkspaceBytes(size_t nq) { return PlanLayout(8, nq).total; }

SI_EXPORT int siSearchOffsets(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                              uint32_t* counts, uint64_t* offsets, void* workspace,
                              size_t workspace_bytes, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!offsets || (nq && (!qs || !qe))) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    ScopedPlan P(st);
    void* plan = workspace;
    size_t pb = workspace_bytes;
    if (!plan) {
        pb = siPlanBytes(ix, nq);
        if ((rc = P.alloc(pb))) return rc;
        plan = P.p;
    }
    if ((rc = siPlanQueries(ix, qs, qe, nq, plan, pb, stream))) return rc;
    if (nq == 0) {
        HIP_TRY(hipMemsetAsync(offsets, 0, sizeof(uint64_t), st));
        return SI_OK;
    }
    return siSearchOffsetsPlanned(ix, plan, nq, counts, offsets, stream);
}

SI_EXPORT int siSearchFill(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                           const uint64_t* offsets, uint32_t* idx, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !offsets || !idx) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    ScopedPlan P(st);
    const size_t pb = siPlanBytes(ix, nq);
    if ((rc = P.alloc(pb))) return rc;
    if ((rc = siPlanQueries(ix, qs, qe, nq, P.p, pb, stream))) return rc;
    return siSearchFillPlanned(ix, P.p, nq, order, offsets, idx, stream);
}

SI_EXPORT int siSearchValuesFill(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                                 const uint64_t* offsets, int32_t* values, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !offsets || !values) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    ScopedPlan P(st);
    const size_t pb = siPlanBytes(ix, nq);
    if ((rc = P.alloc(pb))) return rc;
    if ((rc = siPlanQueries(ix, qs, qe, nq, P.p, pb, stream))) return rc;
    return siSearchValuesFillPlanned(ix, P.p, nq, offsets, values, stream);
}

SI_EXPORT int siUpperBound(const siIndex* ix, const void* xs, size_t n, uint32_t* out, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (n == 0) return SI_OK;
    if (!xs || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (g_tune.ub_tree && ix->ub.d_top && ix->n > 0)
        return DISPATCH_S(ix, (launch_ub_tree<S>(ix, (const S*)xs, n, out, st)));
    return DISPATCH_S(ix, ([&]() -> int {
        k_upper_bound<S><<<grid_for(n, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), (const S*)xs, n, out);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

SI_EXPORT int siHasOverlaps(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint8_t* out,
                            void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, ([&]() -> int {
        k_has_overlaps<S><<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), (const S*)qs,
                                                                        (const S*)qe, nq, out);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

SI_EXPORT int siGatherKeys(const siIndex* ix, const uint32_t* idx, size_t n, void* out_starts, void* out_ends,
                           int32_t* out_values, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (n == 0) return SI_OK;
    if (!idx) return si_set_error(SI_ERR_ARG, "null positions");
    if (out_values && !ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, ([&]() -> int {
        k_gather_keys<S><<<grid_for(n, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), ix->d_values, idx, n,
                                                                      (S*)out_starts, (S*)out_ends, out_values);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

// ---- Set operations over a batched search (superintervals.hpp:1158-1211) -------------
// Query k is interval k of the first set A (its starts, ends, data as the caller passes
// them); positions are the hits of a search of A's intervals in index B.

// Intersection pieces (superintervals.hpp:1158-1174): hit j of query k (search_idxs
// vector order) gives [max(a_s, b_s), min(a_e, b_e)] with A's data (the default combine)
// and B's, one thread per query.
template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_intersect_pieces(IndexView<S> ix, const int32_t* __restrict__ bvals,
                                                                const S* __restrict__ as, const S* __restrict__ ae,
                                                                const int32_t* __restrict__ ad, size_t na,
                                                                const uint64_t* __restrict__ off,
                                                                const uint32_t* __restrict__ idx,
                                                                S* __restrict__ os, S* __restrict__ oe,
                                                                int32_t* __restrict__ oad, int32_t* __restrict__ obd) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= na) return;
    const S s = as[k], e = ae[k];
    const int32_t d = ad ? ad[k] : 0;
    for (uint64_t j = off[k]; j < off[k + 1]; ++j) {
        const uint32_t h = idx[j];
        const S bs = ix.starts[h], be = ix.ends[h];
        if (os) os[j] = s > bs ? s : bs;
        if (oe) oe[j] = e < be ? e : be;
        if (oad) oad[j] = d;
        if (obd) obd[j] = bvals ? bvals[h] : 0;
    }
}

// Difference pieces (superintervals.hpp:1188-1211): the uncovered parts of A's interval k.
// The reference sorts B's hits ascending by (start, end); a descending search lists them
// in the reverse of position order, i.e. ascending by (start, end desc), and the sweep's
// pieces do not depend on the order of hits with equal starts (the first of them emits any
// gap, the cursor ends at the largest end + 1 either way), so the sweep reads the hits
// backwards.  The cursor is kept in 64 bits, so end + 1 never wraps for int32; for int64 a
// hit ending at the type's maximum covers the rest of the interval and ends the sweep.
// FILL = false counts the pieces, true writes them.
template <typename S, bool FILL>
__global__ __launch_bounds__(Q_THREADS) void k_difference_pieces(IndexView<S> ix, const S* __restrict__ as,
                                                                 const S* __restrict__ ae,
                                                                 const int32_t* __restrict__ ad, size_t na,
                                                                 const uint64_t* __restrict__ off,
                                                                 const uint32_t* __restrict__ idx,
                                                                 uint64_t* __restrict__ pieces,
                                                                 S* __restrict__ os, S* __restrict__ oe,
                                                                 int32_t* __restrict__ od) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= na) return;
    constexpr int64_t SMAX = (sizeof(S) == 4) ? (int64_t)INT32_MAX : INT64_MAX;
    const int64_t s = (int64_t)as[k], e = (int64_t)ae[k];
    const int32_t d = ad ? ad[k] : 0;
    uint64_t w = FILL ? pieces[k] : 0;
    int64_t cursor = s;
    bool past = false;   // cursor is beyond the type's maximum
    for (uint64_t j = off[k + 1]; j-- > off[k];) {
        const uint32_t h = idx[j];
        const int64_t cs = max((int64_t)ix.starts[h], s), ce = min((int64_t)ix.ends[h], e);
        if (cs > cursor) {
            if (FILL) {
                if (os) os[w] = (S)cursor;
                if (oe) oe[w] = (S)(cs - 1);
                if (od) od[w] = d;
            }
            ++w;
        }
        if (sizeof(S) == 8 && ce == SMAX) {
            past = true;
            break;
        }
        if (ce + 1 > cursor) cursor = ce + 1;
    }
    if (!past && cursor <= e) {
        if (FILL) {
            if (os) os[w] = (S)cursor;
            if (oe) oe[w] = (S)e;
            if (od) od[w] = d;
        }
        ++w;
    }
    if (!FILL) pieces[k] = w;
}

SI_EXPORT int siIntersectPieces(const siIndex* ix, const void* as, const void* ae, const int32_t* ad, size_t na,
                                const uint64_t* offsets, const uint32_t* idx, void* out_starts, void* out_ends,
                                int32_t* out_adata, int32_t* out_bdata, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (na == 0) return SI_OK;
    if (!as || !ae || !offsets) return si_set_error(SI_ERR_ARG, "null buffer");
    if (out_bdata && !ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, ([&]() -> int {
        k_intersect_pieces<S><<<grid_for(na, Q_THREADS), Q_THREADS, 0, st>>>(
            ix->view<S>(), ix->d_values, (const S*)as, (const S*)ae, ad, na, offsets, idx, (S*)out_starts,
            (S*)out_ends, out_adata, out_bdata);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

SI_EXPORT int siDifferenceOffsets(const siIndex* ix, const void* as, const void* ae, size_t na,
                                  const uint64_t* offsets, const uint32_t* idx, uint64_t* piece_offsets,
                                  void* workspace, size_t workspace_bytes, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!piece_offsets) return si_set_error(SI_ERR_ARG, "null piece offsets");
    if (na && (!as || !ae || !offsets)) return si_set_error(SI_ERR_ARG, "null buffer");
    if (!workspace || workspace_bytes < siSearchWorkspaceBytes(na))
        return si_set_error(SI_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes,
                            siSearchWorkspaceBytes(na));
    hipStream_t st = (hipStream_t)stream;
    if (na) {
        rc = DISPATCH_S(ix, ([&]() -> int {
            k_difference_pieces<S, false><<<grid_for(na, Q_THREADS), Q_THREADS, 0, st>>>(
                ix->view<S>(), (const S*)as, (const S*)ae, nullptr, na, offsets, idx, piece_offsets, nullptr,
                nullptr, nullptr);
            LAUNCH_CHECK();
            return SI_OK;
        }()));
        if (rc) return rc;
    }
    return device_excl_scan<uint64_t, uint64_t>(piece_offsets, na, piece_offsets, piece_offsets + na,
                                                (uint64_t*)workspace, st);
}

SI_EXPORT int siDifferenceFill(const siIndex* ix, const void* as, const void* ae, const int32_t* ad, size_t na,
                               const uint64_t* offsets, const uint32_t* idx, const uint64_t* piece_offsets,
                               void* out_starts, void* out_ends, int32_t* out_data, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (na == 0) return SI_OK;
    if (!as || !ae || !offsets || !piece_offsets) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, ([&]() -> int {
        k_difference_pieces<S, true><<<grid_for(na, Q_THREADS), Q_THREADS, 0, st>>>(
            ix->view<S>(), (const S*)as, (const S*)ae, ad, na, offsets, idx, (uint64_t*)piece_offsets,
            (S*)out_starts, (S*)out_ends, out_data);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

SI_EXPORT int siCoverage(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* count_out,
                         int32_t* cov_out, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->cb != 4) return si_set_error(SI_ERR_UNSUPPORTED, "coverage is defined for int32 coordinates");
    if (nq == 0) return SI_OK;
    if (!qs || !qe) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    k_coverage_i32<<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<int32_t>(), (const int32_t*)qs,
                                                                 (const int32_t*)qe, nq, count_out, cov_out);
    LAUNCH_CHECK();
    return SI_OK;
}

SI_EXPORT int siWalkStats(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint64_t* totals,
                          void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!totals) return si_set_error(SI_ERR_ARG, "null totals");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemsetAsync(totals, 0, 4 * sizeof(uint64_t), st));
    if (nq == 0) return SI_OK;
    return DISPATCH_S(ix, ([&]() -> int {
        k_walk_stats<S><<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), (const S*)qs, (const S*)qe,
                                                                      nq, (unsigned long long*)totals);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

// ============================================================================
// Set tables: batches over many resident indexes
// ============================================================================
// A set table holds the device views of up to max_sets built indexes (slots), plus one
// always-empty slot for out-of-range set numbers.  A batch gives every query its slot; the
// plan sorts the batch by (slot, end), finds each query's bounds and its rank-form count
// through its own set's search trees, and lays out tiles that never mix sets, so the walk
// kernels run unchanged on each tile's set and write the set's own positions.  Putting or
// evicting a set rewrites one slot; the other sets and their indexes are untouched.
struct siSetTable {
    int cb = 4;
    uint32_t nslots = 0;                       // max_sets + 1 (the last is always empty)
    void* d_views = nullptr;                   // nslots SetView<int32_t>
    std::vector<const siIndex*> sets;          // the index in each slot (not owned)
    std::vector<uint64_t> gens;                // its build generation when it was put
};

namespace {

SetView<int32_t> set_view_of(const siIndex* ix) {
    SetView<int32_t> v;
    std::memset(&v, 0, sizeof(v));
    if (!ix || ix->n == 0) return v;
    v.v = ix->view<int32_t>();
    v.values = ix->d_values;
    v.ub = ub_tree_of<int32_t>(ix->ub, (const int32_t*)ix->d_starts, ix->n);
    v.ub_top = (const int32_t*)ix->ub.d_top;
    v.ub_h = ix->ub.top_h;
    v.eb = ub_tree_of<int32_t>(ix->eb, (const int32_t*)ix->d_ends_sorted, ix->n);
    v.eb_top = (const int32_t*)ix->eb.d_top;
    v.eb_h = ix->eb.top_h;
    v.td = ix->tree<int32_t>();
    v.rank_exact = ix->n_inverted == 0 ? 1u : 0u;
    return v;
}

// Byte layout of a set-table plan: a plan of nq queries, then per sorted query its slot, and
// per slot its query count, its first sorted query and its first tile (nslots + 1 each).
struct MultiLayout {
    PlanLayout base;
    size_t off_set, off_scnt, off_sbeg, off_tbeg, total;
    MultiLayout(size_t nq, uint32_t nslots) : base(4, nq) {
        off_set = base.total;
        off_scnt = off_set + align256(4 * nq);
        off_sbeg = off_scnt + align256(4 * ((size_t)nslots + 1));
        off_tbeg = off_sbeg + align256(4 * ((size_t)nslots + 1));
        total = off_tbeg + align256(4 * ((size_t)nslots + 1));
    }
};

constexpr int MK_THREADS = 256;
constexpr int MK_ITEMS = 8;   // queries per thread in the key kernel

// Sort keys (slot << 32 | biased end), payload = input number; per-slot query counts through
// an LDS histogram.  Slots past the table map to the empty slot.
__global__ __launch_bounds__(MK_THREADS) void k_multi_keys(const uint32_t* __restrict__ set_ids,
                                                           const int32_t* __restrict__ qe, size_t nq,
                                                           uint32_t nslots, uint64_t* __restrict__ key,
                                                           uint32_t* __restrict__ val, uint32_t* __restrict__ scnt) {
    extern __shared__ uint32_t s_hist[];
    for (uint32_t b = threadIdx.x; b < nslots; b += MK_THREADS) s_hist[b] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * MK_THREADS * MK_ITEMS;
    for (int j = 0; j < MK_ITEMS; ++j) {
        const size_t k = base + (size_t)j * MK_THREADS + threadIdx.x;
        if (k < nq) {
            uint32_t g = set_ids[k];
            if (g >= nslots - 1) g = nslots - 1;
            key[k] = ((uint64_t)g << 32) | (uint32_t)((uint32_t)qe[k] ^ 0x80000000u);
            val[k] = (uint32_t)k;
            atomicAdd(&s_hist[g], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nslots; b += MK_THREADS)
        if (s_hist[b]) atomicAdd(&scnt[b], s_hist[b]);
}

// Per sorted query: its slot, start, end; upper_bound and lower_bound through its set's
// starts tree; its count in input order, the rank form through the set's sorted-ends tree
// (the walk when the set or the query has start > end).
__global__ __launch_bounds__(Q_THREADS) void k_multi_finish(const SetView<int32_t>* __restrict__ sets,
                                                            const uint64_t* __restrict__ key,
                                                            const uint32_t* __restrict__ val,
                                                            const int32_t* __restrict__ qs, size_t nq,
                                                            PlanView<int32_t> pv, uint32_t* __restrict__ counts) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        pv.counter[0] = 0;
        pv.counter[4] = 1u;   // lower bounds present (NONE where the set has start > end)
        pv.counter[5] = 0u;   // counts are not kept in the plan
    }
    if (k >= nq) return;
    const uint64_t kk = key[k];
    const uint32_t q = val[k];
    const uint32_t g = (uint32_t)(kk >> 32);
    const int32_t e = (int32_t)((uint32_t)kk ^ 0x80000000u);
    const int32_t s = qs[q];
    pv.q[k] = q;
    pv.s[k] = s;
    pv.e[k] = e;
    pv.set[k] = g;
    const SetView<int32_t>& sv = sets[g];
    uint32_t ub = NONE, lb = NONE, c = 0;
    if (sv.v.n) {
        ub = ub_tree_pos(sv.ub, sv.ub_top, sv.ub_h, e);
        if (sv.rank_exact) lb = (s == INT32_MIN) ? 0u : ub_tree_pos(sv.ub, sv.ub_top, sv.ub_h, s - 1) + 1u;
        if (ub != NONE) {
            if (sv.rank_exact && s <= e) {
                const uint32_t below = (s == INT32_MIN) ? 0u : ub_tree_pos(sv.eb, sv.eb_top, sv.eb_h, s - 1) + 1u;
                c = ub + 1u - below;
            } else {
                c = walk_count(sv.v, ub, s);
            }
        }
    }
    pv.idx[k] = ub;
    pv.lb[k] = lb;
    counts[q] = c;
}

// First sorted query and first tile of every slot (one thread: at most a few thousand slots).
__global__ void k_multi_tiles(const uint32_t* __restrict__ scnt, uint32_t nslots, uint32_t tile,
                              uint32_t* __restrict__ sbeg, uint32_t* __restrict__ tbeg) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t a = 0, b = 0;
    for (uint32_t g = 0; g < nslots; ++g) {
        sbeg[g] = a;
        tbeg[g] = b;
        a += scnt[g];
        b += (scnt[g] + tile - 1u) / tile;
    }
    sbeg[nslots] = a;
    tbeg[nslots] = b;
}

PlanView<int32_t> multi_view(void* plan, size_t nq, uint32_t nslots) {
    const MultiLayout L(nq, nslots);
    PlanView<int32_t> pv = plan_view<int32_t>(plan, nq);
    pv.set = (uint32_t*)((char*)plan + L.off_set);
    return pv;
}

MultiTiles multi_tiles(void* plan, size_t nq, uint32_t nslots) {
    const MultiLayout L(nq, nslots);
    char* p = (char*)plan;
    return MultiTiles{(const uint32_t*)(p + L.off_scnt), (const uint32_t*)(p + L.off_sbeg),
                      (const uint32_t*)(p + L.off_tbeg), nslots};
}

int check_table(const siSetTable* t) {
    if (!t) return si_set_error(SI_ERR_ARG, "null set table");
    for (uint32_t g = 0; g + 1 < t->nslots; ++g) {
        const siIndex* ix = t->sets[g];
        if (ix && (!ix->built || ix->generation != t->gens[g]))
            return si_set_error(SI_ERR_NOT_BUILT, "the index in slot %u was rebuilt or cleared since it was put", g);
    }
    return SI_OK;
}

// Tile size of the multi walk: the tuned tile (tiles never mix sets).
uint32_t multi_tile() { return std::max<uint32_t>(POOL, (g_tune.tile + POOL - 1) / POOL * POOL); }

template <int MODE>
int launch_walk_multi(const siSetTable* tb, void* plan, size_t nq, const uint64_t* offsets, uint32_t* out_idx,
                      hipStream_t st, uint64_t cap) {
    if (nq == 0) return SI_OK;
    PlanView<int32_t> pv = multi_view(plan, nq, tb->nslots);
    const MultiTiles mt = multi_tiles(plan, nq, tb->nslots);
    static int cus = 0, per = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_walk<int32_t, MODE, true>, T_THREADS, 0);
        cus = std::max(cus, 1);
        per = std::max(per, 1);
    }
    const int wgs = (g_tune.fill_wgs > 0) ? std::min(g_tune.fill_wgs, per) : per;
    const uint32_t tile = multi_tile();
    const size_t tiles_max = (nq + tile - 1) / tile + tb->nslots;
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)cus * wgs, tiles_max));
    k_walk_counters_reset<<<1, 64, 0, st>>>(pv.counter);
    LAUNCH_CHECK();
    const SetView<int32_t>* sets = (const SetView<int32_t>*)tb->d_views;
    k_walk<int32_t, MODE, true><<<grid, T_THREADS, 0, st>>>(
        IndexView<int32_t>{}, pv, (uint32_t)nq, tile, NONE - 1u, g_tune.lane_steps, true, pv.counter, nullptr,
        nullptr, offsets, out_idx, nullptr, nullptr, cap, sets, mt);
    LAUNCH_CHECK();
    LongView<int32_t> none{};   // the sets' walks read their own arrays only
    k_walk_heavy<int32_t, MODE, true><<<(unsigned)cus * 8u, Q_THREADS, 0, st>>>(
        IndexView<int32_t>{}, TreeDesc<int32_t>{}, UbTree<int32_t>{}, none, pv, nullptr, nullptr, offsets, out_idx,
        nullptr, nullptr, cap, sets);
    LAUNCH_CHECK();
    return SI_OK;
}

}  // namespace

SI_EXPORT siSetTable* siSetTableCreate(int coord_bytes, uint32_t max_sets) {
    if (coord_bytes != 4) {
        si_set_error(SI_ERR_UNSUPPORTED, "set tables hold int32 indexes");
        return nullptr;
    }
    if (max_sets < 1 || max_sets > 4095) {   // the key kernel's per-slot histogram lives in LDS
        si_set_error(SI_ERR_ARG, "max_sets must be in [1, 4095]");
        return nullptr;
    }
    siSetTable* t = new siSetTable();
    t->cb = 4;
    t->nslots = max_sets + 1;
    t->sets.assign(t->nslots, nullptr);
    t->gens.assign(t->nslots, 0);
    const size_t bytes = (size_t)t->nslots * sizeof(SetView<int32_t>);
    std::vector<char> zero(bytes, 0);
    if (hipMalloc(&t->d_views, bytes) != hipSuccess ||
        hipMemcpy(t->d_views, zero.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        si_set_error(SI_ERR_NOMEM, "set table allocation");
        (void)hipFree(t->d_views);
        delete t;
        return nullptr;
    }
    return t;
}

SI_EXPORT void siSetTableDestroy(siSetTable* t) {
    if (!t) return;
    (void)hipFree(t->d_views);
    delete t;
}

SI_EXPORT int siSetTablePut(siSetTable* t, uint32_t slot, const siIndex* ix) {
    if (!t) return si_set_error(SI_ERR_ARG, "null set table");
    if (slot + 1 >= t->nslots) return si_set_error(SI_ERR_ARG, "slot %u out of range (max_sets %u)", slot, t->nslots - 1);
    if (ix) {
        int rc = check_ix(ix);
        if (rc) return rc;
        if (ix->cb != t->cb) return si_set_error(SI_ERR_ARG, "the index's coordinate width differs from the table's");
    }
    const SetView<int32_t> v = set_view_of(ix);
    HIP_TRY(hipMemcpy((char*)t->d_views + (size_t)slot * sizeof(v), &v, sizeof(v), hipMemcpyHostToDevice));
    t->sets[slot] = ix;
    t->gens[slot] = ix ? ix->generation : 0;
    return SI_OK;
}

SI_EXPORT size_t siSetTablePlanBytes(const siSetTable* t, size_t nq) {
    return t ? MultiLayout(nq, t->nslots).total : 0;
}

SI_EXPORT int siSetTableSearchOffsets(const siSetTable* t, const uint32_t* set_ids, const void* qs, const void* qe,
                                      size_t nq, uint32_t* counts, uint64_t* offsets, void* plan, size_t plan_bytes,
                                      void* stream) {
    int rc = check_table(t);
    if (rc) return rc;
    if (!offsets || !plan) return si_set_error(SI_ERR_ARG, "null offsets/plan");
    if (nq >= (size_t)0x7FFFFFFF) return si_set_error(SI_ERR_ARG, "batch of %zu queries exceeds 2^31-1", nq);
    const MultiLayout L(nq, t->nslots);
    if (plan_bytes < L.total) return si_set_error(SI_ERR_ARG, "plan buffer too small: %zu < %zu bytes", plan_bytes, L.total);
    hipStream_t st = (hipStream_t)stream;
    if (nq == 0) {
        k_zero_u64<<<1, 64, 0, st>>>(offsets, 1);
        LAUNCH_CHECK();
        return SI_OK;
    }
    if (!set_ids || !qs || !qe) return si_set_error(SI_ERR_ARG, "null queries");
    char* p = (char*)plan;
    uint64_t* k0 = (uint64_t*)(p + L.base.off_k0);
    uint64_t* k1 = (uint64_t*)(p + L.base.off_k1);
    uint32_t* v0 = (uint32_t*)(p + L.base.off_v0);
    uint32_t* v1 = (uint32_t*)(p + L.base.off_v1);
    RsTemp rt = rs_temp_carve(p + L.base.off_rs, nq);
    uint32_t* scnt = (uint32_t*)(p + L.off_scnt);
    k_zero_u32<<<grid_for(t->nslots, 256), 256, 0, st>>>(scnt, t->nslots);
    LAUNCH_CHECK();
    const unsigned gk = grid_for(nq, MK_THREADS * MK_ITEMS);
    k_multi_keys<<<gk, MK_THREADS, t->nslots * sizeof(uint32_t), st>>>(set_ids, (const int32_t*)qe, nq, t->nslots, k0,
                                                                       v0, scnt);
    LAUNCH_CHECK();
    // sort by (slot, end): the end's digits (the lowest plan_skip_low left unsorted) and the
    // slot's one or two; no host read-back, so the batch can be captured in a graph
    const int slot_digits = (t->nslots <= 256) ? 1 : 2;
    uint64_t* ks;
    uint32_t* vs;
    rc = radix_sort_pairs(k0, v0, k1, v1, nq, g_tune.plan_skip_low, 4 + slot_digits, rt, &ks, &vs, st, false);
    if (rc) return rc;
    PlanView<int32_t> pv = multi_view(plan, nq, t->nslots);
    uint32_t* c32 = counts ? counts : (uint32_t*)(p + L.base.off_heavy);
    k_multi_finish<<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>((const SetView<int32_t>*)t->d_views, ks, vs,
                                                                 (const int32_t*)qs, nq, pv, c32);
    LAUNCH_CHECK();
    k_multi_tiles<<<1, 64, 0, st>>>(scnt, t->nslots, multi_tile(), (uint32_t*)(p + L.off_sbeg),
                                    (uint32_t*)(p + L.off_tbeg));
    LAUNCH_CHECK();
    uint64_t* ws = (uint64_t*)(p + L.base.off_scratch);
    return device_excl_scan<uint32_t, uint64_t>(c32, nq, offsets, offsets + nq, ws, st);
}

SI_EXPORT int siSetTableSearchFill(const siSetTable* t, void* plan, size_t nq, int order, const uint64_t* offsets,
                                   uint32_t* idx, uint64_t capacity, void* stream) {
    int rc = check_table(t);
    if (rc) return rc;
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (nq && (!plan || !offsets || (!idx && capacity))) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (order == SI_ORDER_DESC) return launch_walk_multi<W_DESC>(t, plan, nq, offsets, idx, st, capacity);
    return launch_walk_multi<W_VECTOR>(t, plan, nq, offsets, idx, st, capacity);
}

// ============================================================================
// Search without allocation
// ============================================================================
// siSearchNoAlloc keeps its plan in a file-scope device array instead of memory it
// allocates: sized for SI_NOALLOC_MAX_QUERIES (si_device.h) queries of either coordinate
// width (int64 plans are the larger), held by every device the library runs on until the process ends.
// In host code `scratch` names a host-side stand-in; its device address, one per device,
// comes from hipGetSymbolAddress on the first call on that device.
static constexpr size_t SI_NOALLOC_BYTES = PlanLayout(8, SI_NOALLOC_MAX_QUERIES).total;
static_assert(PlanLayout(4, SI_NOALLOC_MAX_QUERIES).total <= SI_NOALLOC_BYTES, "int32 plans fit");
static __device__ __align__(256) unsigned char scratch[SI_NOALLOC_BYTES];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_DEVICES] = {nullptr};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_ptr(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_DEVICES) return si_set_error(SI_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return si_set_error(SI_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SI_OK;
}

}  // namespace

SI_EXPORT int siSearchNoAlloc(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                              uint32_t* counts, uint64_t* offsets, uint32_t* idx, uint64_t capacity, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (nq > SI_NOALLOC_MAX_QUERIES)
        return si_set_error(SI_ERR_ARG, "siSearchNoAlloc takes at most %u queries per call (got %zu)",
                            (unsigned)SI_NOALLOC_MAX_QUERIES, nq);
    if (!offsets || (nq && (!qs || !qe)) || (capacity && !idx)) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (nq == 0) {
        k_zero_u64<<<1, 64, 0, st>>>(offsets, 1);
        LAUNCH_CHECK();
        return SI_OK;
    }
    void* plan = nullptr;
    if ((rc = scratch_ptr(&plan))) return rc;
    // no host read-back in the plan, so the call can also be captured in a graph
    rc = DISPATCH_S(ix, plan_impl<S>(ix, (const S*)qs, (const S*)qe, nq, plan, st, false));
    if (rc) return rc;
    if ((rc = siSearchOffsetsPlanned(ix, plan, nq, counts, offsets, stream))) return rc;
    if (capacity == 0) return SI_OK;
    if (order == SI_ORDER_DESC)
        return DISPATCH_S(ix, (launch_walk<S, W_DESC>(ix, plan, nq, nullptr, nullptr, offsets, idx, nullptr, st,
                                                      capacity)));
    return DISPATCH_S(ix, (launch_walk<S, W_VECTOR>(ix, plan, nq, nullptr, nullptr, offsets, idx, nullptr, st,
                                                    capacity)));
}

// ============================================================================
// Host-buffer forms
// ============================================================================
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t bytes) {
        if (bytes == 0) bytes = 8;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) { p = nullptr; return si_set_error(SI_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e)); }
        return SI_OK;
    }
};

// ---- single queries: one wave, results straight into pinned host memory ------------------
constexpr uint64_t ONE_CAP = 1u << 16;   // results the single-query scratch holds

// Runs one query (mode W_COUNT / W_DESC / W_VECTOR / W_VALUES / ONE_UB / ONE_ANY) on the
// index's single-query scratch, made on first use, and waits; *total = its count (or
// upper_bound / has_overlaps), and up to ONE_CAP results stay in the scratch after the
// 16-byte header.  The caller holds ix->one.mu.  A repeat of the last call launches nothing.
template <typename S>
int one_query_run(siIndex* ix, int mode, S s, S e, uint64_t* total) {
    OneQuery& o = ix->one;
    if (!o.h) {
        HIP_TRY(hipHostMalloc(&o.h, 16 + ONE_CAP * 4, hipHostMallocMapped));
        HIP_TRY(hipHostGetDevicePointer(&o.d, o.h, 0));
        HIP_TRY(hipStreamCreateWithFlags(&o.st, hipStreamNonBlocking));
        o.cap = ONE_CAP;
        o.last_valid = false;
    }
    if (o.last_valid && o.last_mode == mode && o.last_s == (int64_t)s && o.last_e == (int64_t)e) {
        *total = o.last_total;
        return SI_OK;
    }
    o.last_valid = false;
    const IndexView<S> v = ix->view<S>();
    const UbTree<S> ub = ub_tree_of<S>(ix->ub, (const S*)ix->d_starts, ix->n);
    const UbTree<S> eb = ub_tree_of<S>(ix->eb, (const S*)ix->d_ends_sorted, ix->n);
    const TreeDesc<S> td = ix->tree<S>();
    const LongView<S> lv = long_view_of<S>(ix);
    const uint32_t rx = ix->n_inverted == 0 ? 1u : 0u;
    uint64_t* out = (uint64_t*)o.d;
    switch (mode) {
        case W_COUNT: k_one_query<S, W_COUNT><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
        case W_DESC: k_one_query<S, W_DESC><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
        case W_VECTOR: k_one_query<S, W_VECTOR><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
        case W_VALUES: k_one_query<S, W_VALUES><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
        case ONE_UB: k_one_query<S, ONE_UB><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
        default: k_one_query<S, ONE_ANY><<<1, WAVE, 0, o.st>>>(v, ub, eb, td, lv, ix->d_values, rx, s, e, out, o.cap); break;
    }
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(o.st));
    *total = ((volatile uint64_t*)o.h)[0];
    o.last_valid = true;
    o.last_mode = mode;
    o.last_s = (int64_t)s;
    o.last_e = (int64_t)e;
    o.last_total = *total;
    return SI_OK;
}

// One query of a host-buffer call: the coordinates read from qs[0], qe[0].
int one_query(const siIndex* cix, int mode, const void* qs, const void* qe, uint64_t* total) {
    siIndex* ix = const_cast<siIndex*>(cix);   // the scratch is the index's mutable state
    if (ix->cb == 4) return one_query_run<int32_t>(ix, mode, qs ? *(const int32_t*)qs : 0, *(const int32_t*)qe, total);
    return one_query_run<int64_t>(ix, mode, qs ? *(const int64_t*)qs : 0, *(const int64_t*)qe, total);
}

int upload_queries(const siIndex* ix, const void* qs, const void* qe, size_t nq, DevBuf& dq) {
    const size_t qb = nq * ix->cb;
    int rc = dq.alloc(2 * qb);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dq.p, qs, qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)dq.p + qb, qe, qb, hipMemcpyHostToDevice));
    return SI_OK;
}
}  // namespace

SI_EXPORT int siCountHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !counts) return si_set_error(SI_ERR_ARG, "null buffer");
    if (nq == 1) {
        std::lock_guard<std::mutex> lock(const_cast<siIndex*>(ix)->one.mu);
        uint64_t c = 0;
        if ((rc = one_query(ix, W_COUNT, qs, qe, &c))) return rc;
        counts[0] = (uint32_t)c;
        return SI_OK;
    }
    DevBuf dq, dc;
    if ((rc = upload_queries(ix, qs, qe, nq, dq)) || (rc = dc.alloc(nq * 4))) return rc;
    rc = siCount(ix, dq.p, (char*)dq.p + nq * ix->cb, nq, (uint32_t*)dc.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(counts, dc.p, nq * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

SI_EXPORT int siUpperBoundHost(const siIndex* ix, const void* xs, size_t n, uint32_t* out) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (n == 0) return SI_OK;
    if (!xs || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    if (n == 1) {
        std::lock_guard<std::mutex> lock(const_cast<siIndex*>(ix)->one.mu);
        uint64_t u = 0;
        if ((rc = one_query(ix, ONE_UB, nullptr, xs, &u))) return rc;
        out[0] = (uint32_t)u;
        return SI_OK;
    }
    DevBuf dx, dout;
    if ((rc = dx.alloc(n * ix->cb)) || (rc = dout.alloc(n * 4))) return rc;
    HIP_TRY(hipMemcpy(dx.p, xs, n * ix->cb, hipMemcpyHostToDevice));
    rc = siUpperBound(ix, dx.p, n, (uint32_t*)dout.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, dout.p, n * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

SI_EXPORT int siHasOverlapsHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint8_t* out) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    if (nq == 1) {
        std::lock_guard<std::mutex> lock(const_cast<siIndex*>(ix)->one.mu);
        uint64_t a = 0;
        if ((rc = one_query(ix, ONE_ANY, qs, qe, &a))) return rc;
        out[0] = a ? 1 : 0;
        return SI_OK;
    }
    DevBuf dq, dout;
    if ((rc = upload_queries(ix, qs, qe, nq, dq)) || (rc = dout.alloc(nq))) return rc;
    rc = siHasOverlaps(ix, dq.p, (char*)dq.p + nq * ix->cb, nq, (uint8_t*)dout.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, dout.p, nq, hipMemcpyDeviceToHost));
    return SI_OK;
}

SI_EXPORT int siCoverageHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* count_out,
                             int32_t* cov_out) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !count_out || !cov_out) return si_set_error(SI_ERR_ARG, "null buffer");
    DevBuf dq, dc, dv;
    if ((rc = upload_queries(ix, qs, qe, nq, dq)) || (rc = dc.alloc(nq * 4)) || (rc = dv.alloc(nq * 4))) return rc;
    rc = siCoverage(ix, dq.p, (char*)dq.p + nq * ix->cb, nq, (uint32_t*)dc.p, (int32_t*)dv.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(count_out, dc.p, nq * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cov_out, dv.p, nq * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

SI_EXPORT int siSearchHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                           uint64_t* offsets, uint32_t* idx, int32_t* values, size_t cap, uint64_t* total) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!offsets) return si_set_error(SI_ERR_ARG, "null offsets");
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (values && order != SI_ORDER_DESC) return si_set_error(SI_ERR_ARG, "values are emitted in descending order only");
    if (total) *total = 0;
    if (nq == 0) { offsets[0] = 0; return SI_OK; }
    if (!qs || !qe) return si_set_error(SI_ERR_ARG, "null queries");
    if (values && !ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
    if (nq == 1) {   // one wave on the index's scratch, unless the result outgrows it
        std::lock_guard<std::mutex> lock(const_cast<siIndex*>(ix)->one.mu);
        const int mode = values ? W_VALUES : (order == SI_ORDER_DESC ? W_DESC : W_VECTOR);
        uint64_t tot = 0;
        if ((rc = one_query(ix, mode, qs, qe, &tot))) return rc;
        if (tot <= ONE_CAP) {
            offsets[0] = 0;
            offsets[1] = tot;
            if (total) *total = tot;
            if (tot > cap) return si_set_error(SI_ERR_CAPACITY, "%llu results exceed capacity %zu", (unsigned long long)tot, cap);
            if (tot == 0) return SI_OK;
            if (!values && !idx) return si_set_error(SI_ERR_ARG, "null idx");
            std::memcpy(values ? (void*)values : (void*)idx, (const char*)ix->one.h + 16, tot * 4);
            return SI_OK;
        }
    }
    const size_t pb = siPlanBytes(ix, nq);
    DevBuf dq, doff, dplan;
    if ((rc = upload_queries(ix, qs, qe, nq, dq)) || (rc = doff.alloc((nq + 1) * 8)) || (rc = dplan.alloc(pb)))
        return rc;
    if ((rc = siPlanQueries(ix, dq.p, (char*)dq.p + nq * ix->cb, nq, dplan.p, pb, nullptr))) return rc;
    if ((rc = siSearchOffsetsPlanned(ix, dplan.p, nq, nullptr, (uint64_t*)doff.p, nullptr))) return rc;
    HIP_TRY(hipMemcpy(offsets, doff.p, (nq + 1) * 8, hipMemcpyDeviceToHost));
    const uint64_t tot = offsets[nq];
    if (total) *total = tot;
    if (tot > cap) return si_set_error(SI_ERR_CAPACITY, "%llu results exceed capacity %zu", (unsigned long long)tot, cap);
    if (tot == 0) return SI_OK;
    if (!values && !idx) return si_set_error(SI_ERR_ARG, "null idx");
    DevBuf dout;
    if ((rc = dout.alloc(tot * 4))) return rc;
    if (values) rc = siSearchValuesFillPlanned(ix, dplan.p, nq, (const uint64_t*)doff.p, (int32_t*)dout.p, nullptr);
    else rc = siSearchFillPlanned(ix, dplan.p, nq, order, (const uint64_t*)doff.p, (uint32_t*)dout.p, nullptr);
    if (rc) return rc;
    if ((rc = siPlanCheck(ix, dplan.p, nullptr))) return rc;
    HIP_TRY(hipMemcpy(values ? (void*)values : (void*)idx, dout.p, tot * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

// ---- Set operations on host buffers (the C-ABI's intersection / difference) ----------
namespace {
// Uploads A's int32 intervals as queries and runs plan -> offsets -> descending fill on
// the device; off receives the CSR offsets (host), didx the positions (device).
int search_desc_host(const siIndex* ix, const int32_t* as, const int32_t* ae, size_t na, DevBuf& dq,
                     DevBuf& dplan, DevBuf& doff, DevBuf& didx, std::vector<uint64_t>& off) {
    const size_t pb = siPlanBytes(ix, na);
    int rc;
    if ((rc = upload_queries(ix, as, ae, na, dq)) || (rc = doff.alloc((na + 1) * 8)) || (rc = dplan.alloc(pb)))
        return rc;
    if ((rc = siPlanQueries(ix, dq.p, (char*)dq.p + na * 4, na, dplan.p, pb, nullptr))) return rc;
    if ((rc = siSearchOffsetsPlanned(ix, dplan.p, na, nullptr, (uint64_t*)doff.p, nullptr))) return rc;
    off.resize(na + 1);
    HIP_TRY(hipMemcpy(off.data(), doff.p, (na + 1) * 8, hipMemcpyDeviceToHost));
    if ((rc = didx.alloc(off[na] * 4))) return rc;
    if (off[na] == 0) return SI_OK;
    if ((rc = siSearchFillPlanned(ix, dplan.p, na, SI_ORDER_DESC, (const uint64_t*)doff.p, (uint32_t*)didx.p,
                                  nullptr)))
        return rc;
    return siPlanCheck(ix, dplan.p, nullptr);
}
}  // namespace

int si_intersect_host(const siIndex* ix, const int32_t* as, const int32_t* ae, size_t na,
                      std::vector<uint64_t>& off, std::vector<uint32_t>& pos, std::vector<int32_t>& ps,
                      std::vector<int32_t>& pe) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->cb != 4) return si_set_error(SI_ERR_UNSUPPORTED, "the C-ABI set operations take int32 indexes");
    off.assign(na + 1, 0);
    pos.clear(); ps.clear(); pe.clear();
    if (na == 0) return SI_OK;
    DevBuf dq, dplan, doff, didx, dps, dpe;
    if ((rc = search_desc_host(ix, as, ae, na, dq, dplan, doff, didx, off))) return rc;
    const uint64_t total = off[na];
    if (total == 0) return SI_OK;
    if ((rc = dps.alloc(total * 4)) || (rc = dpe.alloc(total * 4))) return rc;
    if ((rc = siIntersectPieces(ix, dq.p, (char*)dq.p + na * 4, nullptr, na, (const uint64_t*)doff.p,
                                (const uint32_t*)didx.p, dps.p, dpe.p, nullptr, nullptr, nullptr)))
        return rc;
    pos.resize(total); ps.resize(total); pe.resize(total);
    HIP_TRY(hipMemcpy(pos.data(), didx.p, total * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ps.data(), dps.p, total * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pe.data(), dpe.p, total * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

int si_difference_host(const siIndex* ix, const int32_t* as, const int32_t* ae, const int32_t* ad, size_t na,
                       std::vector<int32_t>& ps, std::vector<int32_t>& pe, std::vector<int32_t>& pd) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->cb != 4) return si_set_error(SI_ERR_UNSUPPORTED, "the C-ABI set operations take int32 indexes");
    ps.clear(); pe.clear(); pd.clear();
    if (na == 0) return SI_OK;
    std::vector<uint64_t> off;
    DevBuf dq, dplan, doff, didx, dpoff, dws, dad, dps, dpe, dpd;
    if ((rc = search_desc_host(ix, as, ae, na, dq, dplan, doff, didx, off))) return rc;
    const size_t wsb = siSearchWorkspaceBytes(na);
    if ((rc = dpoff.alloc((na + 1) * 8)) || (rc = dws.alloc(wsb)) || (rc = dad.alloc(na * 4))) return rc;
    HIP_TRY(hipMemcpy(dad.p, ad, na * 4, hipMemcpyHostToDevice));
    const void* dqs = dq.p;
    const void* dqe = (const char*)dq.p + na * 4;
    if ((rc = siDifferenceOffsets(ix, dqs, dqe, na, (const uint64_t*)doff.p, (const uint32_t*)didx.p,
                                  (uint64_t*)dpoff.p, dws.p, wsb, nullptr)))
        return rc;
    uint64_t np_ = 0;
    HIP_TRY(hipMemcpy(&np_, (const uint64_t*)dpoff.p + na, 8, hipMemcpyDeviceToHost));
    if (np_ == 0) return SI_OK;
    if ((rc = dps.alloc(np_ * 4)) || (rc = dpe.alloc(np_ * 4)) || (rc = dpd.alloc(np_ * 4))) return rc;
    if ((rc = siDifferenceFill(ix, dqs, dqe, (const int32_t*)dad.p, na, (const uint64_t*)doff.p,
                               (const uint32_t*)didx.p, (const uint64_t*)dpoff.p, dps.p, dpe.p, (int32_t*)dpd.p,
                               nullptr)))
        return rc;
    ps.resize(np_); pe.resize(np_); pd.resize(np_);
    HIP_TRY(hipMemcpy(ps.data(), dps.p, np_ * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pe.data(), dpe.p, np_ * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pd.data(), dpd.p, np_ * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}
