// si_kernels.h -- device kernels of the superset index (gfx950, wave64).
//
// Layout in HBM of a built index of n intervals (coordinate type S):
//   starts[n], ends[n]   S, sorted by (start asc, end desc), padded to a multiple of
//                        64 entries (starts with the maximum, ends with the minimum)
//   branch[n]            uint32, nearest j < i with ends[j] >= ends[i], or NONE (padded)
//   exitp[n]             uint32, the walk's exit from position i's 64-byte line of ends when
//                        [line start, i] are all misses: branch[leftmost argmax of ends
//                        over that range] (padded with NONE)
//   perm[n]              uint32, input position of the interval at sorted position k
//   values[n]            int32 payload in sorted order (optional)
//   tree levels          S, a 64-ary max tree over ends: level l+1 holds the max of
//                        each group of 64 entries of level l (level 0 is ends)
//   ends32[n]            uint32, int64 indexes whose ends span less than 2^32 only:
//                        ends[i] - ends32_base, read by the wave-cooperative walk (padded)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sik {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int WAVE = 64;
constexpr int FAN = 64;          // max-tree fan-out (one wave-wide load per node)
constexpr int MAX_LEVELS = 8;

template <typename S>
struct TreeDesc {
    const S* lvl[MAX_LEVELS];    // lvl[0] = ends
    uint32_t size[MAX_LEVELS];
    int nlev;
};

template <typename S>
struct IndexView {
    const S* __restrict__ starts;
    const S* __restrict__ ends;
    const uint32_t* __restrict__ branch;
    const uint32_t* __restrict__ exitp;   // line exit target of each position (see k_exit)
    const uint32_t* __restrict__ ends32;  // ends - ends32_base, or null
    int64_t ends32_base;
    uint32_t n;
    uint32_t npad;   // allocated length of starts/ends/branch (multiple of 64)
};

// Intervals longer than len (the long list), in position order: their positions, their ends
// (and ends - ends32_base as uint32 where the index keeps a 32-bit copy), a max tree over
// those ends (td.lvl[0] = ends), and lrank[p] = the number of them before position p.  The
// wave walk takes every hit below the first position starting at or after s - len from this
// list: an interval starting before s - len reaches s only if it is longer than len.
// len == 0: no list.
template <typename S>
struct LongView {
    const uint32_t* __restrict__ pos;
    const S* __restrict__ ends;
    const uint32_t* __restrict__ ends32;
    int64_t ends32_base;
    const uint32_t* __restrict__ lrank;
    TreeDesc<S> td;
    int64_t len;
    uint32_t n;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---- upper_bound ------------------------------------------------------------
// Last position with starts[pos] <= x, or NONE (superintervals.hpp:501-516
// semantics; written as the branchless "base/len" halving).
template <typename S>
__device__ __forceinline__ uint32_t upper_bound_pos(const S* __restrict__ st, uint32_t n, S x) {
    if (n == 0) return NONE;
    const S* base = st;
    uint32_t len = n;
    while (len > 1) {
        uint32_t half = len >> 1;
        base = (base[half] <= x) ? base + half : base;
        len -= half;
    }
    uint32_t c = (uint32_t)(base - st) + ((*base <= x) ? 1u : 0u);
    return c - 1u;   // c == 0 wraps to NONE
}

// ---- the backward superset walk ---------------------------------------------
// From position i (upper_bound of the query end) walk backwards: a hit (s <= ends[i])
// steps to i - 1, a miss jumps to branch[i].  This single loop is the two-phase walk
// of c_superintervals.h:618-642 (leading run, then branch jumps).
template <typename S>
__device__ __forceinline__ uint32_t walk_count(const IndexView<S>& ix, uint32_t i, S s) {
    uint32_t c = 0;
    while (i != NONE) {
        if (s <= ix.ends[i]) { ++c; --i; }
        else i = ix.branch[i];
    }
    return c;
}

}  // namespace sik
