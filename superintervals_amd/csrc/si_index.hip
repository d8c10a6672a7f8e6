// si_index.hip -- device build and batched queries of the superset index.
//
// Build (replaces superintervals.hpp:112-130 + :1415-1439, c_superintervals.h:462-521):
//   1. optional stable LSD radix sort of (start asc, end desc) keys with the input
//      position as payload (8-bit digits, passes whose digit is constant are skipped);
//   2. gather of starts/ends/values into sorted order;
//   3. a 64-ary max tree over ends;
//   4. branch[i] = previous-greater-or-equal of ends[i], one lane per element: search
//      the element's own 64-group, then climb the max tree to the nearest group on the
//      left whose max is >= ends[i] and descend to its rightmost such entry.  The
//      previous-greater-or-equal is unique, so this equals the reference's stack.
// Queries (replace superintervals.hpp:501-516, :651-905, c_superintervals.h:537-756):
//   upper_bound + the backward walk, one query per lane; search_idxs in two passes
//   (count -> exclusive scan -> fill).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "si_common.h"
#include "si_kernels.h"

using namespace sik;

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
SI_EXPORT const char* siVersion(void) { return "superintervals-mi355x 0.1.0 (reference API 1.0.0)"; }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return si_set_error(SI_ERR_HIP, "%s failed: %s (%s:%d)", #expr,               \
                                hipGetErrorString(_e), __FILE__, __LINE__);               \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

namespace {

inline unsigned grid_for(size_t n, unsigned per_block) {
    return (unsigned)((n + per_block - 1) / per_block);
}

// ============================================================================
// Wave / block primitives
// ============================================================================
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        T t = __shfl_xor(v, d, WAVE);
        v = t > v ? t : v;
    }
    return v;
}

// Exclusive scan across a 256-thread block; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_excl_scan256(T v, T* s_w /*[4]*/, T* total) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    T inc = wave_incl_scan(v);
    if (lane == WAVE - 1) s_w[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T x = s_w[k];
        pre += (k < w) ? x : (T)0;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// ============================================================================
// Exclusive scan (reduce -> scan of tile sums -> downsweep), 4096 items per tile
// ============================================================================
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const TI* __restrict__ in, size_t n,
                                                              TO* __restrict__ tile_sums) {
    __shared__ TO s_w[4];
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    TO sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        size_t i = base + (size_t)k * SCAN_THREADS + threadIdx.x;
        if (i < n) sum += (TO)in[i];
    }
    TO tot;
    block_excl_scan256<TO>(sum, s_w, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of nt tile sums in place; *grand receives the total.
template <typename TO>
__global__ __launch_bounds__(1024) void k_scan_tiles(TO* __restrict__ ts, size_t nt, TO* grand) {
    __shared__ TO s_part[1024];
    __shared__ TO s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (size_t base = 0; base < nt; base += 1024) {
        size_t i = base + threadIdx.x;
        TO v = (i < nt) ? ts[i] : (TO)0;
        // inclusive scan over 1024 threads: wave scan then wave totals
        TO inc = wave_incl_scan(v);
        const int lane = lane_id(), w = threadIdx.x >> 6;
        if (lane == WAVE - 1) s_part[w] = inc;
        __syncthreads();
        if (threadIdx.x < 16) {
            TO x = s_part[threadIdx.x];
            TO acc = 0;
            for (int k = 0; k < 16; ++k) {
                TO y = s_part[k];
                acc += (k < (int)threadIdx.x) ? y : (TO)0;
            }
            s_part[16 + threadIdx.x] = acc;
            (void)x;
        }
        __syncthreads();
        TO excl = s_carry + s_part[16 + w] + inc - v;
        __syncthreads();
        if (i < nt) ts[i] = excl;
        if (threadIdx.x == 1023) s_carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0 && grand) *grand = s_carry;
}

// Downsweep: out[i] = tile prefix + exclusive scan within the tile.  Each thread owns
// SCAN_ITEMS consecutive items.  in and out may alias.
template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const TI* in, size_t n,
                                                            const TO* __restrict__ tile_pre,
                                                            TO* out) {
    __shared__ TO s_w[4];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    TO v[SCAN_ITEMS];
    TO sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        size_t i = base + k;
        v[k] = (i < n) ? (TO)in[i] : (TO)0;
        sum += v[k];
    }
    TO tot;
    TO run = block_excl_scan256<TO>(sum, s_w, &tot) + tile_pre[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        size_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

// Exclusive scan of n items (TI) into out (TO); out may alias in when TI == TO.
// If total != nullptr (device pointer) the grand total is stored there.
// ws must hold at least scan_ws_items(n) TO values.
inline size_t scan_ws_items(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

template <typename TI, typename TO>
int device_excl_scan(const TI* in, size_t n, TO* out, TO* total, TO* ws, hipStream_t st) {
    if (n == 0) {
        if (total) HIP_TRY(hipMemsetAsync(total, 0, sizeof(TO), st));
        return SI_OK;
    }
    size_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    k_scan_reduce<TI, TO><<<(unsigned)nt, SCAN_THREADS, 0, st>>>(in, n, ws);
    LAUNCH_CHECK();
    k_scan_tiles<TO><<<1, 1024, 0, st>>>(ws, nt, total);
    LAUNCH_CHECK();
    k_scan_down<TI, TO><<<(unsigned)nt, SCAN_THREADS, 0, st>>>(in, n, ws, out);
    LAUNCH_CHECK();
    return SI_OK;
}

// ============================================================================
// Radix sort: stable LSD over 64-bit keys with a 32-bit payload
// ============================================================================
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;   // 4096 items per tile
constexpr int RS_BINS = 256;

// Histograms of all eight digits at once (to skip constant-digit passes).
__global__ __launch_bounds__(256) void k_rs_hist8(const uint64_t* __restrict__ keys, size_t n,
                                                   uint32_t* __restrict__ hist8 /*[8][256]*/) {
    __shared__ uint32_t h[8][RS_BINS];
    for (int k = threadIdx.x; k < 8 * RS_BINS; k += blockDim.x) (&h[0][0])[k] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t key = keys[i];
#pragma unroll
        for (int d = 0; d < 8; ++d) atomicAdd(&h[d][(key >> (8 * d)) & 255], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 8 * RS_BINS; k += blockDim.x) {
        uint32_t c = (&h[0][0])[k];
        if (c) atomicAdd(&hist8[k], c);
    }
}

// Per-tile digit counts, digit-major: th[d * ntiles + tile].
__global__ __launch_bounds__(RS_THREADS) void k_rs_tile_hist(const uint64_t* __restrict__ keys,
                                                             size_t n, int shift,
                                                             uint32_t* __restrict__ th,
                                                             uint32_t ntiles) {
    __shared__ uint32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ROUNDS; ++r) {
        size_t i = base + (size_t)r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    th[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter.  Each round ranks 256 items (4 waves x 64 lanes, in input order) by
// matching digits with 8 ballots; ranks accumulate across rounds in LDS, the tile is
// staged in LDS in digit order and written out so equal digits land contiguously.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, size_t n, int shift, const uint32_t* __restrict__ th_counts,
    const uint32_t* __restrict__ th_base, uint32_t ntiles) {
    __shared__ uint64_t s_key[RS_TILE];
    __shared__ uint32_t s_val[RS_TILE];
    __shared__ uint32_t s_wcnt[4][RS_BINS];
    __shared__ uint32_t s_run[RS_BINS];
    __shared__ uint32_t s_dstart[RS_BINS];
    __shared__ uint32_t s_w[4];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * RS_TILE;
    const uint32_t tile_n = (uint32_t)min((size_t)RS_TILE, n - base);

    // digit starts within the tile = exclusive scan of this tile's digit counts
    uint32_t my_cnt = th_counts[(size_t)tid * ntiles + tile];
    uint32_t tot;
    s_dstart[tid] = block_excl_scan256<uint32_t>(my_cnt, s_w, &tot);
    s_run[tid] = 0;
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t li = (uint32_t)r * RS_THREADS + tid;
        const bool valid = li < tile_n;
        uint64_t key = valid ? kin[base + li] : 0ull;
        uint32_t val = valid ? vin[base + li] : 0u;
        uint32_t digit = (uint32_t)(key >> shift) & 255u;
        for (int k = lane; k < RS_BINS; k += WAVE) s_wcnt[w][k] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) s_wcnt[w][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        {   // per digit: prefix over the four waves, advance the running count
            uint32_t c0 = s_wcnt[0][tid], c1 = s_wcnt[1][tid], c2 = s_wcnt[2][tid], c3 = s_wcnt[3][tid];
            uint32_t run = s_run[tid];
            s_wcnt[0][tid] = run;
            s_wcnt[1][tid] = run + c0;
            s_wcnt[2][tid] = run + c0 + c1;
            s_wcnt[3][tid] = run + c0 + c1 + c2;
            s_run[tid] = run + c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (valid) {
            uint32_t slot = s_dstart[digit] + s_wcnt[w][digit] + rank;
            s_key[slot] = key;
            s_val[slot] = val;
        }
        __syncthreads();
    }
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t slot = (uint32_t)r * RS_THREADS + tid;
        if (slot < tile_n) {
            uint64_t key = s_key[slot];
            uint32_t digit = (uint32_t)(key >> shift) & 255u;
            size_t dst = (size_t)th_base[(size_t)digit * ntiles + tile] + (slot - s_dstart[digit]);
            kout[dst] = key;
            vout[dst] = s_val[slot];
        }
    }
}

// Stable sort of (keys, vals) by keys.  On return *keys_sorted / *vals_sorted point at
// whichever of the two buffers holds the result.
int radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n,
                     uint64_t** keys_sorted, uint32_t** vals_sorted, hipStream_t st) {
    *keys_sorted = k0;
    *vals_sorted = v0;
    if (n < 2) return SI_OK;
    uint32_t* d_hist8 = nullptr;
    HIP_TRY(hipMalloc(&d_hist8, 8 * RS_BINS * sizeof(uint32_t)));
    HIP_TRY(hipMemsetAsync(d_hist8, 0, 8 * RS_BINS * sizeof(uint32_t), st));
    unsigned hb = std::min<unsigned>(grid_for(n, 256 * 16), 2048);
    k_rs_hist8<<<hb, 256, 0, st>>>(k0, n, d_hist8);
    LAUNCH_CHECK();
    std::vector<uint32_t> h8(8 * RS_BINS);
    HIP_TRY(hipMemcpyAsync(h8.data(), d_hist8, h8.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipFree(d_hist8));

    const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    const size_t nth = (size_t)ntiles * RS_BINS;
    uint32_t *d_th = nullptr, *d_tb = nullptr, *d_ws = nullptr;
    HIP_TRY(hipMalloc(&d_th, nth * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&d_tb, nth * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&d_ws, scan_ws_items(nth) * sizeof(uint32_t)));
    uint64_t *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    int rc = SI_OK;
    for (int d = 0; d < 8 && rc == SI_OK; ++d) {
        bool constant = false;
        for (int b = 0; b < RS_BINS; ++b)
            if (h8[d * RS_BINS + b] == n) { constant = true; break; }
        if (constant) continue;
        const int shift = 8 * d;
        k_rs_tile_hist<<<ntiles, RS_THREADS, 0, st>>>(ka, n, shift, d_th, ntiles);
        if (hipGetLastError() != hipSuccess) { rc = si_set_error(SI_ERR_HIP, "radix histogram launch"); break; }
        rc = device_excl_scan<uint32_t, uint32_t>(d_th, nth, d_tb, nullptr, d_ws, st);
        if (rc != SI_OK) break;
        k_rs_scatter<<<ntiles, RS_THREADS, 0, st>>>(ka, va, kb, vb, n, shift, d_th, d_tb, ntiles);
        if (hipGetLastError() != hipSuccess) { rc = si_set_error(SI_ERR_HIP, "radix scatter launch"); break; }
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    hipError_t e1 = hipStreamSynchronize(st);
    hipFree(d_th);
    hipFree(d_tb);
    hipFree(d_ws);
    if (rc != SI_OK) return rc;
    if (e1 != hipSuccess) return si_set_error(SI_ERR_HIP, "radix sort: %s", hipGetErrorString(e1));
    *keys_sorted = ka;
    *vals_sorted = va;
    return SI_OK;
}

// ============================================================================
// Build kernels
// ============================================================================
__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
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

// ============================================================================
// Query kernels
// ============================================================================
constexpr int Q_THREADS = 256;

template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_upper_bound(IndexView<S> ix, const S* __restrict__ xs,
                                                           size_t n, uint32_t* __restrict__ out) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = upper_bound_pos(ix.starts, ix.n, xs[q]);
}

template <typename S>
__global__ __launch_bounds__(Q_THREADS) void k_count(IndexView<S> ix, const S* __restrict__ qs,
                                                     const S* __restrict__ qe, size_t nq,
                                                     uint32_t* __restrict__ counts,
                                                     uint64_t* __restrict__ counts64) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const S s = qs[q], e = qe[q];
    uint32_t c = walk_count(ix, upper_bound_pos(ix.starts, ix.n, e), s);
    if (counts) counts[q] = c;
    if (counts64) counts64[q] = c;
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

// Second pass.  ORDER 0: strict descending.  ORDER 1: leading run ascending, then the
// walk's remaining hits descending.  VALUES: write values[pos] instead of pos.
template <typename S, int ORDER, bool VALUES>
__global__ __launch_bounds__(Q_THREADS) void k_fill(IndexView<S> ix, const S* __restrict__ qs,
                                                    const S* __restrict__ qe, size_t nq,
                                                    const uint64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ out_idx,
                                                    int32_t* __restrict__ out_val) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const S s = qs[q], e = qe[q];
    uint64_t o = offsets[q];
    uint32_t i = upper_bound_pos(ix.starts, ix.n, e);
    if (ORDER == 1) {
        const uint32_t top = i;
        while (i != NONE && s <= ix.ends[i]) --i;
        for (uint32_t j = i + 1u; j != top + 1u; ++j) out_idx[o++] = j;   // i == NONE -> from 0
        if (i == NONE) return;
        i = ix.branch[i];
    }
    while (i != NONE) {
        if (s <= ix.ends[i]) {
            if (VALUES) out_val[o++] = vals[i];
            else out_idx[o++] = i;
            --i;
        } else {
            i = ix.branch[i];
        }
    }
}

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

}  // namespace

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
    uint32_t* d_perm = nullptr;
    int32_t* d_values = nullptr;
    void* d_tree = nullptr;   // levels 1.. of the max tree, contiguous
    int nlev = 0;
    size_t lvl_off[MAX_LEVELS] = {0};   // element offsets into d_tree (level >= 1)
    uint32_t lvl_size[MAX_LEVELS] = {0};
    double last_build_ms = 0.0;

    void release() {
        hipFree(d_starts); hipFree(d_ends); hipFree(d_branch); hipFree(d_perm);
        hipFree(d_values); hipFree(d_tree);
        d_starts = d_ends = d_tree = nullptr;
        d_branch = d_perm = nullptr;
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
        v.n = (uint32_t)n;
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

namespace {

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
    HIP_TRY(hipMalloc(&ix->d_starts, nb));
    HIP_TRY(hipMalloc(&ix->d_ends, nb));
    HIP_TRY(hipMalloc(&ix->d_branch, n * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&ix->d_perm, n * sizeof(uint32_t)));
    if (values) HIP_TRY(hipMalloc(&ix->d_values, n * sizeof(int32_t)));

    const unsigned g = grid_for(n, 256);
    const bool presorted = start_sorted && end_sorted;
    if (presorted) {
        k_iota<<<g, 256, 0, st>>>(ix->d_perm, n);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(ix->d_starts, in_s, nb, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(ix->d_ends, in_e, nb, hipMemcpyDeviceToDevice, st));
        if (values) HIP_TRY(hipMemcpyAsync(ix->d_values, in_v, n * 4, hipMemcpyDeviceToDevice, st));
    } else {
        uint64_t *k0 = nullptr, *k1 = nullptr;
        uint32_t *v0 = nullptr, *v1 = nullptr;
        HIP_TRY(hipMalloc(&k0, n * 8));
        HIP_TRY(hipMalloc(&k1, n * 8));
        HIP_TRY(hipMalloc(&v0, n * 4));
        HIP_TRY(hipMalloc(&v1, n * 4));
        k_iota<<<g, 256, 0, st>>>(v0, n);
        LAUNCH_CHECK();
        uint64_t* ks;
        uint32_t* vs;
        int rc;
        if (sizeof(S) == 4) {
            k_key_i32<<<g, 256, 0, st>>>((const int32_t*)in_s, (const int32_t*)in_e, k0, n);
            LAUNCH_CHECK();
            rc = radix_sort_pairs(k0, v0, k1, v1, n, &ks, &vs, st);
        } else {
            // two stable passes: end descending, then start ascending
            k_key_end_desc_i64<<<g, 256, 0, st>>>((const int64_t*)in_e, k0, n);
            LAUNCH_CHECK();
            rc = radix_sort_pairs(k0, v0, k1, v1, n, &ks, &vs, st);
            if (rc == SI_OK) {
                uint64_t* kf = (ks == k0) ? k1 : k0;   // free key buffer
                uint32_t* vf = (vs == v0) ? v1 : v0;
                k_key_start_i64<<<g, 256, 0, st>>>((const int64_t*)in_s, vs, kf, n);
                LAUNCH_CHECK();
                rc = radix_sort_pairs(kf, vs, ks, vf, n, &ks, &vs, st);
            }
        }
        if (rc == SI_OK) {
            HIP_TRY(hipMemcpyAsync(ix->d_perm, vs, n * 4, hipMemcpyDeviceToDevice, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        hipFree(k0); hipFree(k1); hipFree(v0); hipFree(v1);
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
    HIP_TRY(hipStreamSynchronize(st));
    if (!on_device) { hipFree(in_s); hipFree(in_e); if (in_v) hipFree(in_v); }
    ix->built = true;
    ix->last_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SI_OK;
}

inline int check_ix(const siIndex* ix) {
    if (!ix) return si_set_error(SI_ERR_ARG, "null index");
    if (!ix->built) return si_set_error(SI_ERR_NOT_BUILT, "index has not been built");
    return SI_OK;
}

#define DISPATCH_S(ix, ...)                                                  \
    ((ix)->cb == 4 ? [&] { using S = int32_t; return __VA_ARGS__; }()        \
                   : [&] { using S = int64_t; return __VA_ARGS__; }())

template <typename S>
int count_impl(const siIndex* ix, const S* qs, const S* qe, size_t nq, uint32_t* counts,
               uint64_t* counts64, hipStream_t st) {
    if (nq == 0) return SI_OK;
    k_count<S><<<grid_for(nq, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), qs, qe, nq, counts, counts64);
    LAUNCH_CHECK();
    return SI_OK;
}

template <typename S>
int fill_impl(const siIndex* ix, const S* qs, const S* qe, size_t nq, int order,
              const uint64_t* offsets, uint32_t* idx, int32_t* vals, hipStream_t st) {
    if (nq == 0) return SI_OK;
    const unsigned g = grid_for(nq, Q_THREADS);
    if (vals) {
        if (!ix->d_values) return si_set_error(SI_ERR_ARG, "index was built without values");
        k_fill<S, 0, true><<<g, Q_THREADS, 0, st>>>(ix->view<S>(), qs, qe, nq, offsets, ix->d_values, nullptr, vals);
    } else if (order == SI_ORDER_DESC) {
        k_fill<S, 0, false><<<g, Q_THREADS, 0, st>>>(ix->view<S>(), qs, qe, nq, offsets, nullptr, idx, nullptr);
    } else {
        k_fill<S, 1, false><<<g, Q_THREADS, 0, st>>>(ix->view<S>(), qs, qe, nq, offsets, nullptr, idx, nullptr);
    }
    LAUNCH_CHECK();
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
    ix->release();
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

SI_EXPORT int siUpperBound(const siIndex* ix, const void* xs, size_t n, uint32_t* out, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (n == 0) return SI_OK;
    if (!xs || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, ([&]() -> int {
        k_upper_bound<S><<<grid_for(n, Q_THREADS), Q_THREADS, 0, st>>>(ix->view<S>(), (const S*)xs, n, out);
        LAUNCH_CHECK();
        return SI_OK;
    }()));
}

SI_EXPORT int siCount(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts,
                      void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq && (!qs || !qe || !counts)) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, count_impl<S>(ix, (const S*)qs, (const S*)qe, nq, counts, nullptr, st));
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

SI_EXPORT size_t siSearchWorkspaceBytes(size_t nq) { return scan_ws_items(nq + 1) * sizeof(uint64_t); }

SI_EXPORT int siSearchOffsets(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                              uint32_t* counts, uint64_t* offsets, void* workspace,
                              size_t workspace_bytes, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (!offsets || (nq && (!qs || !qe))) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    const size_t need = siSearchWorkspaceBytes(nq);
    uint64_t* ws = (uint64_t*)workspace;
    bool own = false;
    if (!ws) {
        HIP_TRY(hipMallocAsync((void**)&ws, need, st));
        own = true;
    } else if (workspace_bytes < need) {
        return si_set_error(SI_ERR_ARG, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
    }
    rc = DISPATCH_S(ix, count_impl<S>(ix, (const S*)qs, (const S*)qe, nq, counts, offsets, st));
    if (rc == SI_OK) rc = device_excl_scan<uint64_t, uint64_t>(offsets, nq, offsets, offsets + nq, ws, st);
    if (own) {
        hipError_t e = hipFreeAsync(ws, st);
        if (rc == SI_OK && e != hipSuccess) rc = si_set_error(SI_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    }
    return rc;
}

SI_EXPORT int siSearchFill(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                           const uint64_t* offsets, uint32_t* idx, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (order != SI_ORDER_DESC && order != SI_ORDER_VECTOR) return si_set_error(SI_ERR_ARG, "bad order %d", order);
    if (nq && (!qs || !qe || !offsets || !idx)) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, fill_impl<S>(ix, (const S*)qs, (const S*)qe, nq, order, offsets, idx, nullptr, st));
}

SI_EXPORT int siSearchValuesFill(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                                 const uint64_t* offsets, int32_t* values, void* stream) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq && (!qs || !qe || !offsets || !values)) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    return DISPATCH_S(ix, fill_impl<S>(ix, (const S*)qs, (const S*)qe, nq, SI_ORDER_DESC, offsets, nullptr, values, st));
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

// ---- host-buffer forms ------------------------------------------------------
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
    int alloc(size_t bytes) {
        if (bytes == 0) bytes = 8;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) { p = nullptr; return si_set_error(SI_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e)); }
        return SI_OK;
    }
};
}  // namespace

SI_EXPORT int siCountHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (nq == 0) return SI_OK;
    if (!qs || !qe || !counts) return si_set_error(SI_ERR_ARG, "null buffer");
    const size_t qb = nq * ix->cb;
    DevBuf dq, dc;
    if ((rc = dq.alloc(2 * qb)) || (rc = dc.alloc(nq * 4))) return rc;
    HIP_TRY(hipMemcpy(dq.p, qs, qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)dq.p + qb, qe, qb, hipMemcpyHostToDevice));
    rc = siCount(ix, dq.p, (char*)dq.p + qb, nq, (uint32_t*)dc.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(counts, dc.p, nq * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

SI_EXPORT int siUpperBoundHost(const siIndex* ix, const void* xs, size_t n, uint32_t* out) {
    int rc = check_ix(ix);
    if (rc) return rc;
    if (n == 0) return SI_OK;
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
    const size_t qb = nq * ix->cb;
    DevBuf dq, dout;
    if ((rc = dq.alloc(2 * qb)) || (rc = dout.alloc(nq))) return rc;
    HIP_TRY(hipMemcpy(dq.p, qs, qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)dq.p + qb, qe, qb, hipMemcpyHostToDevice));
    rc = siHasOverlaps(ix, dq.p, (char*)dq.p + qb, nq, (uint8_t*)dout.p, nullptr);
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
    const size_t qb = nq * ix->cb;
    DevBuf dq, dc, dv;
    if ((rc = dq.alloc(2 * qb)) || (rc = dc.alloc(nq * 4)) || (rc = dv.alloc(nq * 4))) return rc;
    HIP_TRY(hipMemcpy(dq.p, qs, qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)dq.p + qb, qe, qb, hipMemcpyHostToDevice));
    rc = siCoverage(ix, dq.p, (char*)dq.p + qb, nq, (uint32_t*)dc.p, (int32_t*)dv.p, nullptr);
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
    if (values && order != SI_ORDER_DESC) return si_set_error(SI_ERR_ARG, "values are emitted in descending order only");
    if (total) *total = 0;
    if (nq == 0) { offsets[0] = 0; return SI_OK; }
    if (!qs || !qe) return si_set_error(SI_ERR_ARG, "null queries");
    const size_t qb = nq * ix->cb;
    DevBuf dq, doff, dws;
    if ((rc = dq.alloc(2 * qb)) || (rc = doff.alloc((nq + 1) * 8)) || (rc = dws.alloc(siSearchWorkspaceBytes(nq))))
        return rc;
    HIP_TRY(hipMemcpy(dq.p, qs, qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((char*)dq.p + qb, qe, qb, hipMemcpyHostToDevice));
    const void* dqs = dq.p;
    const void* dqe = (char*)dq.p + qb;
    rc = siSearchOffsets(ix, dqs, dqe, nq, nullptr, (uint64_t*)doff.p, dws.p, siSearchWorkspaceBytes(nq), nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(offsets, doff.p, (nq + 1) * 8, hipMemcpyDeviceToHost));
    const uint64_t tot = offsets[nq];
    if (total) *total = tot;
    if (tot > cap) return si_set_error(SI_ERR_CAPACITY, "%llu results exceed capacity %zu", (unsigned long long)tot, cap);
    if (tot == 0) return SI_OK;
    if (!values && !idx) return si_set_error(SI_ERR_ARG, "null idx");
    DevBuf dout;
    if ((rc = dout.alloc(tot * 4))) return rc;
    if (values) rc = siSearchValuesFill(ix, dqs, dqe, nq, (const uint64_t*)doff.p, (int32_t*)dout.p, nullptr);
    else rc = siSearchFill(ix, dqs, dqe, nq, order, (const uint64_t*)doff.p, (uint32_t*)dout.p, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(values ? (void*)values : (void*)idx, dout.p, tot * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}
