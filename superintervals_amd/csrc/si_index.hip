// si_index.hip -- device build of the superset index.
//
// Replaces superintervals.hpp:112-130 (build) + :1415-1439 (sort_intervals) and
// c_superintervals.h:462-521 (sortIntervals / indexSuperIntervals):
//   1. unless the add() flags promise canonical order, a stable LSD radix sort of
//      (start asc, end desc) keys with the input position as payload;
//   2. gather of starts/ends/values into sorted order (arrays padded to a multiple of
//      64 entries: ends with the minimum coordinate, starts with the maximum, branch
//      with NONE, so aligned 64-byte line loads never leave the allocation);
//   3. a 64-ary max tree over ends;
//   4. branch[i] = previous-greater-or-equal of ends[i], one lane per element: search
//      the element's own 64-group, then climb the max tree to the nearest group on the
//      left whose max is >= ends[i] and descend to its rightmost such entry.  The
//      previous-greater-or-equal is unique, so this equals the reference's stack.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "si_internal.h"
#include "si_prims.h"

// ============================================================================
// Error state
// ============================================================================
namespace {
thread_local char g_err[1024] = {0};
}

int si_set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

SI_EXPORT const char* siGetLastError(void) { return g_err; }
SI_EXPORT const char* siVersion(void) { return "superintervals-mi355x 0.2.0 (reference API 1.0.0)"; }

namespace {
// ============================================================================
// Build kernels
// ============================================================================
__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// Padding beyond n: ends = minimum (never a hit), starts = maximum, branch = NONE.
template <typename S>
__global__ void k_pad(S* __restrict__ st, S* __restrict__ en, uint32_t* __restrict__ br,
                      int32_t* __restrict__ vals, size_t n, size_t npad) {
    size_t i = n + threadIdx.x;
    if (i < npad) {
        st[i] = (sizeof(S) == 4) ? (S)INT32_MAX : (S)INT64_MAX;
        en[i] = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
        br[i] = NONE;
        if (vals) vals[i] = 0;
    }
}

// Order-preserving unsigned keys of ends (and back), for the sorted copy of ends.
template <typename S>
__global__ void k_end_keys(const S* __restrict__ e, uint64_t* __restrict__ key, uint32_t* __restrict__ val,
                           size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        key[i] = (sizeof(S) == 4) ? (uint64_t)((uint32_t)e[i] ^ 0x80000000u)
                                  : ((uint64_t)e[i] ^ 0x8000000000000000ull);
        val[i] = (uint32_t)i;
    }
}

template <typename S>
__global__ void k_end_unkey(const uint64_t* __restrict__ key, S* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = (sizeof(S) == 4) ? (S)(int32_t)((uint32_t)key[i] ^ 0x80000000u)
                                  : (S)(int64_t)(key[i] ^ 0x8000000000000000ull);
    }
}

// int32 composite key: (start asc, end desc) in one 64-bit unsigned order.
__global__ void k_key_i32(const int32_t* __restrict__ s, const int32_t* __restrict__ e,
                          uint64_t* __restrict__ key, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint32_t hs = (uint32_t)s[i] ^ 0x80000000u;
        uint32_t le = ~((uint32_t)e[i] ^ 0x80000000u);
        key[i] = ((uint64_t)hs << 32) | le;
    }
}

// int64 first pass key: end descending.
__global__ void k_key_end_desc_i64(const int64_t* __restrict__ e, uint64_t* __restrict__ key,
                                   size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = ~((uint64_t)e[i] ^ 0x8000000000000000ull);
}

// int64 second pass key: start ascending, gathered through the first pass's order.
__global__ void k_key_start_i64(const int64_t* __restrict__ s, const uint32_t* __restrict__ order,
                                uint64_t* __restrict__ key, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = (uint64_t)s[order[i]] ^ 0x8000000000000000ull;
}

template <typename S>
__global__ void k_gather(const S* __restrict__ s_in, const S* __restrict__ e_in,
                         const int32_t* __restrict__ v_in, const uint32_t* __restrict__ perm,
                         S* __restrict__ s_out, S* __restrict__ e_out, int32_t* __restrict__ v_out,
                         size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint32_t p = perm[i];
        s_out[i] = s_in[p];
        e_out[i] = e_in[p];
        if (v_out) v_out[i] = v_in[p];
    }
}

// One wave per output node: out[b] = max(in[64b .. 64b+63]).
template <typename S>
__global__ __launch_bounds__(256) void k_tree_level(const S* __restrict__ in, size_t n_in,
                                                    S* __restrict__ out, size_t n_out) {
    size_t b = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (b >= n_out) return;
    size_t i = b * FAN + lane_id();
    S v = (i < n_in) ? in[i] : (sizeof(S) == 4 ? (S)INT32_MIN : (S)INT64_MIN);
    v = wave_max(v);
    if (lane_id() == 0) out[b] = v;
}

// Search tree for upper_bound.  Level l + 1 holds every 64th key of level l (out[j] =
// in[64 j]), so the 64 keys of level l below entry j of level l + 1 are one 256-byte
// (int32) node whose first key is that entry.
template <typename S>
__global__ __launch_bounds__(256) void k_every64(const S* __restrict__ in, uint32_t n_out, S* __restrict__ out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_out) out[j] = in[(size_t)j * 64];
}

// The top level in Eytzinger order: node k (1-based, k < 2^h) at depth d holds in-order
// element i = ((2 (k - 2^d) + 1) << (h - 1 - d)) - 1 of the level, or max(S) past its end.
template <typename S>
__global__ __launch_bounds__(256) void k_eytz(const S* __restrict__ lvl, uint32_t m, uint32_t h, S* __restrict__ top) {
    const uint32_t k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (1u << h)) return;
    const uint32_t d = 31u - (uint32_t)__builtin_clz(k);
    const uint32_t i = ((2u * (k - (1u << d)) + 1u) << (h - 1u - d)) - 1u;
    top[k - 1] = (i < m) ? lvl[i] : ((sizeof(S) == 4) ? (S)INT32_MAX : (S)INT64_MAX);
}

// exitp[i]: where the line walk leaves position i's line [b0, b0 + W) (W = 64 bytes of
// ends) when every position of [b0, i] is a miss.  The reference walk entering that run of
// misses at i follows branch[] until it drops below b0; the first such element is the
// nearest j < b0 with ends[j] >= max(ends[b0..i]), i.e. branch[a] for the leftmost maximum
// a of ends[b0..i] (all of [b0, a) is below ends[a]).  It depends on i only, not on the
// query, so the build computes it once and the walk replaces a W-element argmax by one read.
template <typename S>
__global__ __launch_bounds__(256) void k_exit(const S* __restrict__ ends, const uint32_t* __restrict__ branch,
                                              uint32_t n, uint32_t npad, uint32_t* __restrict__ ex) {
    constexpr uint32_t W = 64 / sizeof(S);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        ex[i] = NONE;
        return;
    }
    const uint32_t b0 = i & ~(W - 1);
    uint32_t a = b0;
    S best = ends[b0];
    for (uint32_t j = b0 + 1; j <= i; ++j) {
        const S v = ends[j];
        if (v > best) { best = v; a = j; }
    }
    ex[i] = branch[a];
}

// Number of intervals with start > end (one atomic per wave that holds any).
template <typename S>
__global__ __launch_bounds__(256) void k_count_inverted(const S* __restrict__ st, const S* __restrict__ en, size_t n,
                                                        unsigned long long* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool inv = i < n && st[i] > en[i];
    const uint64_t m = __ballot(inv);
    if (lane_id() == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

// branch[i] = nearest j < i with ends[j] >= ends[i], or NONE.
template <typename S>
__global__ __launch_bounds__(256) void k_branch(TreeDesc<S> td, uint32_t* __restrict__ branch) {
    const uint32_t n = td.size[0];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const S* __restrict__ ends = td.lvl[0];
    const S v = ends[i];
    uint32_t res = NONE;
    // own group at level 0
    const uint32_t lo0 = i & ~(uint32_t)(FAN - 1);
    for (uint32_t j = i; j > lo0;) {
        --j;
        if (ends[j] >= v) { res = j; break; }
    }
    if (res == NONE) {
        uint32_t pos = i / FAN;
        for (int l = 1; l < td.nlev; ++l) {
            const S* __restrict__ L = td.lvl[l];
            const uint32_t glo = pos & ~(uint32_t)(FAN - 1);
            uint32_t found = NONE;
            for (uint32_t k = pos; k > glo;) {
                --k;
                if (L[k] >= v) { found = k; break; }
            }
            if (found != NONE) {
                uint32_t node = found;
                for (int d = l - 1; d >= 0; --d) {
                    const S* __restrict__ D = td.lvl[d];
                    const uint32_t cb = node * FAN;
                    uint32_t k = min(cb + (uint32_t)(FAN - 1), td.size[d] - 1);
                    while (k > cb && D[k] < v) --k;
                    node = k;
                }
                res = node;
                break;
            }
            pos /= FAN;
        }
    }
    branch[i] = res;
}

}  // namespace

namespace {

// ends[i] - base as uint32 (the caller checked that every end fits), 0 in the padding.
__global__ __launch_bounds__(256) void k_ends32(const int64_t* __restrict__ ends, size_t n, size_t npad, int64_t base,
                                                uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) out[i] = (i < n) ? (uint32_t)((uint64_t)ends[i] - (uint64_t)base) : 0u;
}

// ---- the long list (LongView) ---------------------------------------------------------
// end - start + 1 > len, in 64-bit arithmetic that cannot overflow (start > end: not long).
template <typename S>
__device__ __forceinline__ bool longer_than(S st, S en, int64_t len) {
    if (en < st) return false;
    return (uint64_t)en - (uint64_t)st >= (uint64_t)len;   // (end - start + 1) > len
}

// Lengths of m evenly spaced intervals (for the threshold).
template <typename S>
__global__ __launch_bounds__(256) void k_len_sample(const S* __restrict__ st, const S* __restrict__ en, uint32_t n,
                                                    uint32_t m, int64_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t p = (uint32_t)((uint64_t)i * n / m);
    const uint64_t d = (uint64_t)en[p] - (uint64_t)st[p];
    out[i] = (en[p] < st[p]) ? 0 : (d >= (uint64_t)INT64_MAX ? INT64_MAX : (int64_t)d + 1);
}

template <typename S>
__global__ __launch_bounds__(256) void k_long_flags(const S* __restrict__ st, const S* __restrict__ en, uint32_t n,
                                                    int64_t len, uint32_t* __restrict__ flags) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) flags[p] = longer_than(st[p], en[p], len) ? 1u : 0u;
}

// Entry j = lrank[p] of every long position p; the padding past nlong never hits.
template <typename S>
__global__ __launch_bounds__(256) void k_long_fill(const S* __restrict__ en, const uint32_t* __restrict__ en32,
                                                   const uint32_t* __restrict__ flags, const uint32_t* __restrict__ lrank,
                                                   uint32_t n, uint32_t nlong, uint32_t npad, uint32_t* __restrict__ pos,
                                                   S* __restrict__ lends, uint32_t* __restrict__ lends32) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n && flags[p]) {
        const uint32_t j = lrank[p];
        pos[j] = p;
        lends[j] = en[p];
        if (lends32) lends32[j] = en32[p];
    }
    if (p >= nlong && p < npad) {
        pos[p] = NONE;
        lends[p] = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
        if (lends32) lends32[p] = 0u;
    }
}

// Search tree over n sorted keys: 64-ary levels of first keys up to a top level of at
// most 4095 keys, which is also stored in Eytzinger order (k_every64, k_eytz).
template <typename S>
int build_key_tree(const S* keys, size_t n, KeyTree& t, hipStream_t st) {
    uint32_t sz[MAX_LEVELS + 1];
    sz[0] = (uint32_t)n;
    int nl = 0;
    size_t tot = 0;
    while (sz[nl] > 4095 && nl < MAX_LEVELS - 1) {
        sz[nl + 1] = (uint32_t)((sz[nl] + 63) / 64);
        ++nl;
        t.lv_off[nl] = tot;
        t.lv_size[nl] = sz[nl];
        tot += sz[nl];
    }
    t.nlv = nl;
    if (tot) HIP_TRY(hipMalloc(&t.d_lv, tot * sizeof(S)));
    for (int l = 1; l <= nl; ++l) {
        const S* in = (l == 1) ? keys : (const S*)t.d_lv + t.lv_off[l - 1];
        k_every64<S><<<grid_for(sz[l], 256), 256, 0, st>>>(in, sz[l], (S*)t.d_lv + t.lv_off[l]);
        LAUNCH_CHECK();
    }
    const uint32_t m = sz[nl];
    uint32_t h = 0;
    while (((1u << h) - 1u) < m) ++h;
    t.top_size = m;
    t.top_h = h;
    HIP_TRY(hipMalloc(&t.d_top, (size_t)((1u << h) - 1u) * sizeof(S)));
    const S* topl = (nl == 0) ? keys : (const S*)t.d_lv + t.lv_off[nl];
    k_eytz<S><<<grid_for((1u << h) - 1u, 256), 256, 0, st>>>(topl, m, h, (S*)t.d_top);
    LAUNCH_CHECK();
    return SI_OK;
}

// The long list of a built index: the threshold (the tuning's, or 64 x the median length
// rounded up to a power of two), the positions and ends of the longer intervals and the max
// tree over those ends.  Left out (len = 0) for indexes under 2^16 intervals, and when the
// chosen threshold leaves more than half of the intervals long.
template <typename S>
int build_long_list(siIndex* ix, hipStream_t st) {
    const uint32_t n = (uint32_t)ix->n;
    int64_t len = si_tune_long_len();
    const bool automatic = len == 0;
    if (len < 0 || (automatic && n < (1u << 16))) return SI_OK;
    const S* d_st = (const S*)ix->d_starts;
    const S* d_en = (const S*)ix->d_ends;
    if (automatic) {
        constexpr uint32_t M = 4096;
        int64_t* d_s = nullptr;
        HIP_TRY(hipMalloc(&d_s, M * sizeof(int64_t)));
        std::vector<int64_t> h(M);
        k_len_sample<S><<<grid_for(M, 256), 256, 0, st>>>(d_st, d_en, n, M, d_s);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_s, M * sizeof(int64_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        hipFree(d_s);
        if (e != hipSuccess) return si_set_error(SI_ERR_HIP, "long list lengths: %s", hipGetErrorString(e));
        std::nth_element(h.begin(), h.begin() + M / 2, h.end());
        const int64_t med = std::max<int64_t>(h[M / 2], 1);
        len = 1;
        while (len < 64 * med && len < ((int64_t)1 << 62)) len <<= 1;
    }
    LongList& L = ix->lng;
    uint32_t* flags = nullptr;
    uint32_t* ws = nullptr;
    HIP_TRY(hipMalloc(&flags, (size_t)n * 4));
    HIP_TRY(hipMalloc(&L.d_lrank, ((size_t)n + 1) * 4));
    HIP_TRY(hipMalloc(&ws, scan_ws_items(n) * 4));
    k_long_flags<S><<<grid_for(n, 256), 256, 0, st>>>(d_st, d_en, n, len, flags);
    LAUNCH_CHECK();
    int rc = device_excl_scan<uint32_t, uint32_t>(flags, n, L.d_lrank, L.d_lrank + n, ws, st);
    uint32_t nlong = 0;
    if (rc == SI_OK) {
        hipError_t e = hipMemcpyAsync(&nlong, L.d_lrank + n, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = si_set_error(SI_ERR_HIP, "long list count: %s", hipGetErrorString(e));
    }
    hipFree(ws);
    if (rc != SI_OK || nlong == 0 || (automatic && nlong > n / 2)) {
        hipFree(flags);
        L.release();
        return rc;
    }
    const uint32_t npad = (nlong + 63u) & ~63u;
    HIP_TRY(hipMalloc(&L.d_pos, (size_t)npad * 4));
    HIP_TRY(hipMalloc(&L.d_ends, (size_t)npad * sizeof(S)));
    if (ix->d_ends32) HIP_TRY(hipMalloc(&L.d_ends32, (size_t)npad * 4));
    k_long_fill<S><<<grid_for(std::max(n, npad), 256), 256, 0, st>>>(d_en, ix->d_ends32, flags, L.d_lrank, n, nlong,
                                                                     npad, L.d_pos, (S*)L.d_ends, L.d_ends32);
    LAUNCH_CHECK();
    // max tree over the long ends, as over the index's ends
    std::vector<uint32_t> sizes{nlong};
    while (sizes.back() > (uint32_t)FAN && (int)sizes.size() < MAX_LEVELS)
        sizes.push_back((sizes.back() + FAN - 1) / FAN);
    L.nlev = (int)sizes.size();
    size_t tot = 0;
    for (int l = 1; l < L.nlev; ++l) { L.lvl_off[l] = tot; L.lvl_size[l] = sizes[l]; tot += sizes[l]; }
    L.lvl_size[0] = nlong;
    if (tot) {
        HIP_TRY(hipMalloc(&L.d_tree, tot * sizeof(S)));
        for (int l = 1; l < L.nlev; ++l) {
            const S* in = (l == 1) ? (const S*)L.d_ends : (const S*)L.d_tree + L.lvl_off[l - 1];
            k_tree_level<S><<<grid_for((size_t)sizes[l] * FAN, 256), 256, 0, st>>>(in, sizes[l - 1],
                                                                                   (S*)L.d_tree + L.lvl_off[l], sizes[l]);
            LAUNCH_CHECK();
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    hipFree(flags);
    L.len = len;
    L.n = nlong;
    return SI_OK;
}

template <typename S>
int build_impl(siIndex* ix, const void* starts, const void* ends, const int32_t* values, size_t n,
               int start_sorted, int end_sorted, int on_device, hipStream_t st) {
    auto t0 = std::chrono::steady_clock::now();
    ix->release();
    if (n == 0) {
        ix->built = true;
        ix->last_build_ms = 0;
        return SI_OK;
    }
    if (n >= (size_t)NONE) return si_set_error(SI_ERR_ARG, "index size %zu exceeds 2^32-2", n);
    const size_t nb = n * sizeof(S);
    S *in_s = nullptr, *in_e = nullptr;
    int32_t* in_v = nullptr;
    // Device copies of the inputs (or the caller's device arrays).
    if (on_device) {
        in_s = (S*)starts;
        in_e = (S*)ends;
    } else {
        HIP_TRY(hipMalloc(&in_s, nb));
        HIP_TRY(hipMalloc(&in_e, nb));
        HIP_TRY(hipMemcpyAsync(in_s, starts, nb, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(in_e, ends, nb, hipMemcpyHostToDevice, st));
    }
    if (values) {
        if (on_device) in_v = (int32_t*)values;
        else {
            HIP_TRY(hipMalloc(&in_v, n * sizeof(int32_t)));
            HIP_TRY(hipMemcpyAsync(in_v, values, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        }
    }
    // Sorted arrays are padded to a multiple of 64 entries (see the file comment).
    const size_t npad = (n + 63) & ~(size_t)63;
    HIP_TRY(hipMalloc(&ix->d_starts, npad * sizeof(S)));
    HIP_TRY(hipMalloc(&ix->d_ends, npad * sizeof(S)));
    HIP_TRY(hipMalloc(&ix->d_branch, npad * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&ix->d_exit, npad * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&ix->d_perm, n * sizeof(uint32_t)));
    if (values) HIP_TRY(hipMalloc(&ix->d_values, npad * sizeof(int32_t)));
    if (npad > n) {
        k_pad<S><<<1, 64, 0, st>>>((S*)ix->d_starts, (S*)ix->d_ends, ix->d_branch, ix->d_values, n, npad);
        LAUNCH_CHECK();
    }

    const unsigned g = grid_for(n, 256);
    const bool presorted = start_sorted && end_sorted;
    if (presorted) {
        k_iota<<<g, 256, 0, st>>>(ix->d_perm, n);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(ix->d_starts, in_s, nb, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(ix->d_ends, in_e, nb, hipMemcpyDeviceToDevice, st));
        if (values) HIP_TRY(hipMemcpyAsync(ix->d_values, in_v, n * 4, hipMemcpyDeviceToDevice, st));
    } else {
        // one allocation: two key buffers, two payload buffers, radix temporaries
        const size_t kb = align256(n * 8), vb = align256(n * 4);
        char* tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, 2 * kb + 2 * vb + rs_temp_bytes(n)));
        uint64_t* k0 = (uint64_t*)tmp;
        uint64_t* k1 = (uint64_t*)(tmp + kb);
        uint32_t* v0 = (uint32_t*)(tmp + 2 * kb);
        uint32_t* v1 = (uint32_t*)(tmp + 2 * kb + vb);
        RsTemp rt = rs_temp_carve(tmp + 2 * kb + 2 * vb, n);
        k_iota<<<g, 256, 0, st>>>(v0, n);
        LAUNCH_CHECK();
        uint64_t* ks;
        uint32_t* vs;
        int rc;
        if (sizeof(S) == 4) {
            k_key_i32<<<g, 256, 0, st>>>((const int32_t*)in_s, (const int32_t*)in_e, k0, n);
            LAUNCH_CHECK();
            rc = radix_sort_pairs(k0, v0, k1, v1, n, 0, 8, rt, &ks, &vs, st);
        } else {
            // two stable passes: end descending, then start ascending
            k_key_end_desc_i64<<<g, 256, 0, st>>>((const int64_t*)in_e, k0, n);
            LAUNCH_CHECK();
            rc = radix_sort_pairs(k0, v0, k1, v1, n, 0, 8, rt, &ks, &vs, st);
            if (rc == SI_OK) {
                uint64_t* kf = (ks == k0) ? k1 : k0;   // free key buffer
                uint32_t* vf = (vs == v0) ? v1 : v0;
                k_key_start_i64<<<g, 256, 0, st>>>((const int64_t*)in_s, vs, kf, n);
                LAUNCH_CHECK();
                rc = radix_sort_pairs(kf, vs, ks, vf, n, 0, 8, rt, &ks, &vs, st);
            }
        }
        if (rc == SI_OK) {
            HIP_TRY(hipMemcpyAsync(ix->d_perm, vs, n * 4, hipMemcpyDeviceToDevice, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        hipFree(tmp);
        if (rc != SI_OK) return rc;
        k_gather<S><<<g, 256, 0, st>>>(in_s, in_e, in_v, ix->d_perm, (S*)ix->d_starts, (S*)ix->d_ends,
                                       ix->d_values, n);
        LAUNCH_CHECK();
    }
    // max tree over ends
    {
        std::vector<uint32_t> sizes{(uint32_t)n};
        while (sizes.back() > (uint32_t)FAN && (int)sizes.size() < MAX_LEVELS)
            sizes.push_back((sizes.back() + FAN - 1) / FAN);
        ix->nlev = (int)sizes.size();
        size_t tot = 0;
        for (int l = 1; l < ix->nlev; ++l) { ix->lvl_off[l] = tot; ix->lvl_size[l] = sizes[l]; tot += sizes[l]; }
        ix->lvl_size[0] = (uint32_t)n;
        if (tot) {
            HIP_TRY(hipMalloc(&ix->d_tree, tot * sizeof(S)));
            for (int l = 1; l < ix->nlev; ++l) {
                const S* in = (l == 1) ? (const S*)ix->d_ends : (const S*)ix->d_tree + ix->lvl_off[l - 1];
                S* out = (S*)ix->d_tree + ix->lvl_off[l];
                k_tree_level<S><<<grid_for((size_t)sizes[l] * FAN, 256), 256, 0, st>>>(in, sizes[l - 1], out, sizes[l]);
                LAUNCH_CHECK();
            }
        }
    }
    ix->n = n;
    k_branch<S><<<g, 256, 0, st>>>(ix->tree<S>(), ix->d_branch);
    LAUNCH_CHECK();
    k_exit<S><<<grid_for(npad, 256), 256, 0, st>>>((const S*)ix->d_ends, ix->d_branch, (uint32_t)n, (uint32_t)npad,
                                                 ix->d_exit);
    LAUNCH_CHECK();
    // intervals with start > end (queries on such an index take the walk, see siIndex)
    {
        unsigned long long* d_inv = nullptr;
        HIP_TRY(hipMalloc(&d_inv, sizeof(unsigned long long)));
        unsigned long long h_inv = 0;
        hipError_t e = hipMemsetAsync(d_inv, 0, sizeof(unsigned long long), st);
        if (e == hipSuccess) {
            k_count_inverted<S><<<g, 256, 0, st>>>((const S*)ix->d_starts, (const S*)ix->d_ends, n, d_inv);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&h_inv, d_inv, sizeof(h_inv), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        hipFree(d_inv);
        if (e != hipSuccess) return si_set_error(SI_ERR_HIP, "inverted count: %s", hipGetErrorString(e));
        ix->n_inverted = (size_t)h_inv;
    }
    // upper_bound search tree over the starts
    {
        const int rc = build_key_tree<S>((const S*)ix->d_starts, n, ix->ub, st);
        if (rc != SI_OK) return rc;
    }
    // ends in ascending order, for the rank form of count
    {
        const size_t kb = align256(n * 8), vb = align256(n * 4);
        char* tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, 2 * kb + 2 * vb + rs_temp_bytes(n)));
        uint64_t* k0 = (uint64_t*)tmp;
        uint64_t* k1 = (uint64_t*)(tmp + kb);
        uint32_t* v0 = (uint32_t*)(tmp + 2 * kb);
        uint32_t* v1 = (uint32_t*)(tmp + 2 * kb + vb);
        RsTemp rt = rs_temp_carve(tmp + 2 * kb + 2 * vb, n);
        k_end_keys<S><<<g, 256, 0, st>>>((const S*)ix->d_ends, k0, v0, n);
        LAUNCH_CHECK();
        uint64_t* ks;
        uint32_t* vs;
        int rc = radix_sort_pairs(k0, v0, k1, v1, n, 0, (int)sizeof(S), rt, &ks, &vs, st);
        if (rc == SI_OK) {
            hipError_t e = hipMalloc(&ix->d_ends_sorted, n * sizeof(S));
            if (e == hipSuccess) {
                k_end_unkey<S><<<g, 256, 0, st>>>(ks, (S*)ix->d_ends_sorted, n);
                e = hipGetLastError();
            }
            if (e != hipSuccess) rc = si_set_error(SI_ERR_HIP, "sorted ends: %s", hipGetErrorString(e));
        }
        hipError_t es = hipStreamSynchronize(st);
        hipFree(tmp);
        if (rc != SI_OK) return rc;
        if (es != hipSuccess) return si_set_error(SI_ERR_HIP, "build: %s", hipGetErrorString(es));
    }
    // search tree over the sorted ends (rank count)
    {
        const int rc = build_key_tree<S>((const S*)ix->d_ends_sorted, n, ix->eb, st);
        if (rc != SI_OK) return rc;
    }
    // 32-bit ends for the wave-cooperative walk of int64 indexes whose ends span < 2^32:
    // a third fewer bytes per position than ends + exitp at 8 + 4
    if (sizeof(S) == 8) {
        int64_t lohi[2];
        HIP_TRY(hipMemcpyAsync(&lohi[0], (const int64_t*)ix->d_ends_sorted, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&lohi[1], (const int64_t*)ix->d_ends_sorted + (n - 1), 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if ((uint64_t)lohi[1] - (uint64_t)lohi[0] < 0xFFFFFFFFull) {
            const size_t npad = (n + 63) & ~(size_t)63;
            HIP_TRY(hipMalloc(&ix->d_ends32, npad * sizeof(uint32_t)));
            ix->ends32_base = lohi[0];
            k_ends32<<<grid_for(npad, 256), 256, 0, st>>>((const int64_t*)ix->d_ends, n, npad, lohi[0],
                                                           ix->d_ends32);
            LAUNCH_CHECK();
        }
    }
    {
        const int rc = build_long_list<S>(ix, st);
        if (rc != SI_OK) return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (!on_device) { hipFree(in_s); hipFree(in_e); if (in_v) hipFree(in_v); }
    ix->built = true;
    ix->last_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SI_OK;
}

}  // namespace

// ============================================================================
// C entry points
// ============================================================================
SI_EXPORT siIndex* siIndexCreate(int coord_bytes) {
    if (coord_bytes != 4 && coord_bytes != 8) {
        si_set_error(SI_ERR_ARG, "coord_bytes must be 4 or 8");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        si_set_error(SI_ERR_HIP, "no HIP device available");
        return nullptr;
    }
    siIndex* ix = new siIndex();
    ix->cb = coord_bytes;
    hipGetDevice(&ix->device);
    return ix;
}

SI_EXPORT void siIndexDestroy(siIndex* ix) {
    if (!ix) return;
    for (siLearned* m : ix->models) si_learned_forget_index(m);
    ix->release();
    ix->one.release();
    delete ix;
}

SI_EXPORT int siIndexBuild(siIndex* ix, const void* starts, const void* ends, const int32_t* values,
                           size_t n, int start_sorted, int end_sorted, int on_device, void* stream) {
    if (!ix) return si_set_error(SI_ERR_ARG, "null index");
    if (n && (!starts || !ends)) return si_set_error(SI_ERR_ARG, "null starts/ends");
    hipStream_t st = (hipStream_t)stream;
    int rc = ix->cb == 4 ? build_impl<int32_t>(ix, starts, ends, values, n, start_sorted, end_sorted, on_device, st)
                         : build_impl<int64_t>(ix, starts, ends, values, n, start_sorted, end_sorted, on_device, st);
    if (rc != SI_OK) ix->release();
    return rc;
}

SI_EXPORT size_t siIndexSize(const siIndex* ix) { return ix ? ix->n : 0; }
SI_EXPORT int siIndexCoordBytes(const siIndex* ix) { return ix ? ix->cb : 0; }
SI_EXPORT double siIndexLastBuildMs(const siIndex* ix) { return ix ? ix->last_build_ms : 0.0; }
SI_EXPORT const void* siIndexDeviceStarts(const siIndex* ix) { return ix ? ix->d_starts : nullptr; }
SI_EXPORT const void* siIndexDeviceEnds(const siIndex* ix) { return ix ? ix->d_ends : nullptr; }
SI_EXPORT const uint32_t* siIndexDeviceBranch(const siIndex* ix) { return ix ? ix->d_branch : nullptr; }
SI_EXPORT const uint32_t* siIndexDevicePerm(const siIndex* ix) { return ix ? ix->d_perm : nullptr; }

SI_EXPORT int siIndexDownload(const siIndex* ix, void* starts, void* ends, int32_t* values,
                              uint32_t* perm, uint64_t* branch) {
    int rc = check_ix(ix);
    if (rc) return rc;
    const size_t n = ix->n;
    if (n == 0) return SI_OK;
    if (values && !ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
    if (starts) HIP_TRY(hipMemcpy(starts, ix->d_starts, n * ix->cb, hipMemcpyDeviceToHost));
    if (ends) HIP_TRY(hipMemcpy(ends, ix->d_ends, n * ix->cb, hipMemcpyDeviceToHost));
    if (values) HIP_TRY(hipMemcpy(values, ix->d_values, n * 4, hipMemcpyDeviceToHost));
    if (perm) HIP_TRY(hipMemcpy(perm, ix->d_perm, n * 4, hipMemcpyDeviceToHost));
    if (branch) {
        std::vector<uint32_t> b(n);
        HIP_TRY(hipMemcpy(b.data(), ix->d_branch, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) branch[i] = (b[i] == NONE) ? UINT64_MAX : (uint64_t)b[i];
    }
    return SI_OK;
}

