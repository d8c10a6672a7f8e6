// si_prims.h -- device primitives shared by the build and query translation units:
// wave/block scans and reductions, a three-kernel exclusive scan and a stable LSD
// radix sort of 64-bit keys with a 32-bit payload.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "si_common.h"
#include "si_kernels.h"

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return si_set_error(SI_ERR_HIP, "%s failed: %s (%s:%d)", #expr,               \
                                hipGetErrorString(_e), __FILE__, __LINE__);               \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

namespace sik {
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

// Inclusive scan of a 32-bit value across the wave in six DPP adds: shifts by 1, 2, 4, 8
// within each row of 16 lanes (a lane with no source in its row adds 0, bound_ctrl), then
// row 0's total broadcast into rows 1 and 3 (row_bcast:15, rows 0 and 2 not written) and
// lane 31's into rows 2 and 3 (row_bcast:31).  A DPP source VGPR written by the previous
// VALU instruction needs two wait states, hence the s_nop 1 before each add.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    asm volatile(
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:0\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:0\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:0\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:0\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_add_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "s_nop 1"
        : "+v"(v));
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
// SCAN_ITEMS consecutive items; the tile passes through LDS (one pad entry per 32 items)
// so that global loads and stores are coalesced across the wave.  A block reads and
// writes only its own tile, all reads before any write, so in and out may alias.
template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const TI* in, size_t n,
                                                            const TO* __restrict__ tile_pre,
                                                            TO* out) {
    __shared__ TO s_w[4];
    __shared__ TO s_v[SCAN_TILE + SCAN_TILE / 32];
    const size_t tbase = (size_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t li = (uint32_t)k * SCAN_THREADS + threadIdx.x;
        const size_t i = tbase + li;
        s_v[li + li / 32] = (i < n) ? (TO)in[i] : (TO)0;
    }
    __syncthreads();
    TO v[SCAN_ITEMS];
    TO sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t li = threadIdx.x * SCAN_ITEMS + k;
        v[k] = s_v[li + li / 32];
        sum += v[k];
    }
    TO tot;
    TO run = block_excl_scan256<TO>(sum, s_w, &tot) + tile_pre[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t li = threadIdx.x * SCAN_ITEMS + k;
        s_v[li + li / 32] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t li = (uint32_t)k * SCAN_THREADS + threadIdx.x;
        const size_t i = tbase + li;
        if (i < n) out[i] = s_v[li + li / 32];
    }
}

// Exclusive scan of n items (TI) into out (TO); out may alias in when TI == TO.
// If total != nullptr (device pointer) the grand total is stored there.
// ws must hold at least scan_ws_items(n) TO values.
constexpr size_t scan_ws_items(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

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
// 2048-item tiles: the scatter's LDS (keys, payloads, digit starts: 25 KB) lets six
// blocks share a CU; 4096-item tiles (two blocks) were 10 % slower per pass
#ifndef SI_RS_ROUNDS
#define SI_RS_ROUNDS 8
#endif
constexpr int RS_ROUNDS = SI_RS_ROUNDS;   // rounds of 256 items per tile
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;   // items per tile
constexpr int RS_BINS = 256;

// Histograms of digits [dlo, dhi) at once (to skip constant-digit passes), tile by tile:
// each thread loads its RS_ROUNDS keys of a tile before counting them.  A digit equal in
// every lane of a wave (the case this pass looks for) is added once by one lane; others
// by one LDS atomic per lane.
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist8(const uint64_t* __restrict__ keys, size_t n, int dlo,
                                                          int dhi, uint32_t* __restrict__ hist8 /*[8][256]*/) {
    __shared__ uint32_t h[8][RS_BINS];
    for (int k = threadIdx.x; k < 8 * RS_BINS; k += RS_THREADS) (&h[0][0])[k] = 0;
    __syncthreads();
    const size_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const size_t base = t * RS_TILE;
        const uint32_t tile_n = (uint32_t)min((size_t)RS_TILE, n - base);
        uint64_t key[RS_ROUNDS];
#pragma unroll
        for (int r = 0; r < RS_ROUNDS; ++r) {
            const uint32_t li = (uint32_t)r * RS_THREADS + threadIdx.x;
            key[r] = li < tile_n ? keys[base + li] : 0ull;
        }
#pragma unroll
        for (int r = 0; r < RS_ROUNDS; ++r) {
            const bool valid = (uint32_t)r * RS_THREADS + threadIdx.x < tile_n;
            const uint64_t vm = __builtin_amdgcn_ballot_w64(valid);
            if (vm == 0ull) continue;   // wave-uniform
            for (int d = dlo; d < dhi; ++d) {
                const uint32_t digit = (uint32_t)(key[r] >> (8 * d)) & 255u;
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(digit);   // lane 0's: valid whenever any is
                if (__builtin_amdgcn_ballot_w64(valid && digit == d0) == vm) {
                    if (lane_id() == 0) atomicAdd(&h[d][d0], (uint32_t)__popcll(vm));
                } else if (valid) {
                    atomicAdd(&h[d][digit], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 8 * RS_BINS; k += RS_THREADS) {
        const uint32_t c = (&h[0][0])[k];
        if (c) atomicAdd(&hist8[k], c);
    }
}

// Per-tile digit counts, digit-major: th[d * ntiles + tile].  A digit equal across a wave
// is added once, from one lane.
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
        const size_t i = base + (size_t)r * RS_THREADS + threadIdx.x;
        const bool valid = i < n;
        const uint64_t vm = __builtin_amdgcn_ballot_w64(valid);
        if (vm == 0ull) continue;   // wave-uniform
        const uint32_t digit = valid ? (uint32_t)(keys[i] >> shift) & 255u : 0u;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(digit);   // lane 0's: valid whenever any is
        if (__builtin_amdgcn_ballot_w64(valid && digit == d0) == vm) {
            if (lane_id() == 0) atomicAdd(&h[d0], (uint32_t)__popcll(vm));
        } else if (valid) {
            atomicAdd(&h[digit], 1u);
        }
    }
    __syncthreads();
    th[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Ranking of one tile of up to RS_TILE items held in registers (item r * 256 + tid in
// key[r]): off[r] receives the item's position among the tile's items of equal digit, in
// item order, and the return value is the tile's count of digit tid.  Each round of 256
// items finds, per wave, the lanes with equal digits by 8 ballots; the wave's leader of
// each digit records their number in s_c[round * 4 + wave][digit], and one column scan
// per digit over the 32 (round, wave) entries turns those numbers into each entry's start.
// s_c (RS_ROUNDS * 4 rows of RS_BINS 16-bit counts) must be free on entry; it is free
// again once every thread has passed a barrier after the call.
__device__ __forceinline__ uint32_t rs_rank_tile(const uint64_t (&key)[RS_ROUNDS], uint32_t tile_n, int shift,
                                                 uint16_t (*s_c)[RS_BINS], uint32_t (&off)[RS_ROUNDS]) {
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    uint32_t* const s_c32 = reinterpret_cast<uint32_t*>(&s_c[0][0]);
#pragma unroll
    for (int k = 0; k < RS_ROUNDS * 4 * RS_BINS / 2 / RS_THREADS; ++k) s_c32[k * RS_THREADS + tid] = 0u;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const bool valid = (uint32_t)r * RS_THREADS + tid < tile_n;
        const uint32_t digit = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __builtin_amdgcn_ballot_w64((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) s_c[r * 4 + w][digit] = (uint16_t)__popcll(peers);
        off[r] = rank;
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < RS_ROUNDS * 4; ++k) {
        const uint32_t c = s_c[k][tid];
        s_c[k][tid] = (uint16_t)run;
        run += c;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) off[r] += s_c[r * 4 + w][(uint32_t)(key[r] >> shift) & 255u];
    return run;
}

// Where a scatter writes item (key, val) at position dst: the sort's buffers, or, in a
// sort's last pass, whatever arrays its caller wants the sorted items unpacked into.
struct ArrayOut {
    uint64_t* __restrict__ k;
    uint32_t* __restrict__ v;
    __device__ __forceinline__ void put(size_t dst, uint64_t key, uint32_t val) const {
        k[dst] = key;
        v[dst] = val;
    }
};

// Stable scatter of one tile (tile-major digit offsets from the scan of the per-tile digit
// counts): the tile is ranked in registers, staged in LDS in digit order and written out so
// equal digits land contiguously.
template <class Out>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, Out out, size_t n, int shift,
    const uint32_t* __restrict__ th_base, uint32_t ntiles) {
    __shared__ union {
        uint16_t c[RS_ROUNDS * 4][RS_BINS];
        uint64_t key[RS_TILE];
    } s_u;
    __shared__ uint32_t s_val[RS_TILE];
    __shared__ uint32_t s_dstart[RS_BINS];
    __shared__ uint32_t s_w[4];

    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * RS_TILE;
    const uint32_t tile_n = (uint32_t)min((size_t)RS_TILE, n - base);
    uint64_t key[RS_ROUNDS];
    uint32_t val[RS_ROUNDS], off[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t li = (uint32_t)r * RS_THREADS + tid;
        key[r] = li < tile_n ? kin[base + li] : 0ull;
        val[r] = li < tile_n ? vin[base + li] : 0u;
    }
    const uint32_t cnt = rs_rank_tile(key, tile_n, shift, s_u.c, off);
    uint32_t tot;
    s_dstart[tid] = block_excl_scan256<uint32_t>(cnt, s_w, &tot);   // its barriers free s_u.c
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        if ((uint32_t)r * RS_THREADS + tid < tile_n) {
            const uint32_t slot = s_dstart[(uint32_t)(key[r] >> shift) & 255u] + off[r];
            s_u.key[slot] = key[r];
            s_val[slot] = val[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t slot = (uint32_t)r * RS_THREADS + tid;
        if (slot < tile_n) {
            const uint64_t k = s_u.key[slot];
            const uint32_t digit = (uint32_t)(k >> shift) & 255u;
            const size_t dst = (size_t)th_base[(size_t)digit * ntiles + tile] + (slot - s_dstart[digit]);
            out.put(dst, k, s_val[slot]);
        }
    }
}

// Temporary device memory of a radix sort of n items (see rs_temp_bytes).
struct RsTemp {
    uint32_t* hist8;   // [8][256] digit histograms
    uint32_t* th;      // per-tile digit counts, digit-major
    uint32_t* tb;      // their exclusive scan
    uint32_t* ws;      // scan workspace
};

constexpr size_t rs_ntiles(size_t n) { return (n + RS_TILE - 1) / RS_TILE; }

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t rs_temp_bytes(size_t n) {
    const size_t nth = rs_ntiles(n) * RS_BINS;
    return align256(8 * RS_BINS * 4) + 2 * align256(nth * 4) + align256(scan_ws_items(nth) * 4);
}

inline RsTemp rs_temp_carve(void* base, size_t n) {
    const size_t nth = rs_ntiles(n) * RS_BINS;
    char* p = (char*)base;
    RsTemp t;
    t.hist8 = (uint32_t*)p; p += align256(8 * RS_BINS * 4);
    t.th = (uint32_t*)p; p += align256(nth * 4);
    t.tb = (uint32_t*)p; p += align256(nth * 4);
    t.ws = (uint32_t*)p;
    return t;
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(uint64_t* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Stable sort of (keys, vals) by key digits [digit_lo, digit_hi) (8-bit digits,
// least significant first).  Digits that are equal for every key are skipped, which
// needs one host read-back of the digit histograms; skip_constant = false sorts every
// digit in [digit_lo, digit_hi) with no read-back, so the sort can be captured in a graph.
// tmp comes from rs_temp_carve.  On return *keys_sorted / *vals_sorted point at whichever
// buffer holds the result.
// With last (optional), the last pass run writes the sorted items through it instead of
// into the buffers, and *last_written says whether one ran; only if none ran are the
// items in *keys_sorted / *vals_sorted.
template <class LastOut = ArrayOut>
int radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n,
                     int digit_lo, int digit_hi, const RsTemp& tmp, uint64_t** keys_sorted,
                     uint32_t** vals_sorted, hipStream_t st, bool skip_constant = true,
                     const LastOut* last = nullptr, bool* last_written = nullptr) {
    if (last_written) *last_written = false;
    *keys_sorted = k0;
    *vals_sorted = v0;
    if (n < 2) return SI_OK;
    uint32_t h8[8 * RS_BINS] = {0};
    if (skip_constant) {
        k_zero_u32<<<grid_for(8 * RS_BINS, 256), 256, 0, st>>>(tmp.hist8, 8 * RS_BINS);
        LAUNCH_CHECK();
        const unsigned hb = (unsigned)std::min<size_t>(rs_ntiles(n), 1024);
        k_rs_hist8<<<hb, RS_THREADS, 0, st>>>(k0, n, digit_lo, digit_hi, tmp.hist8);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(h8, tmp.hist8, sizeof(h8), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }

    const uint32_t ntiles = (uint32_t)rs_ntiles(n);
    const size_t nth = (size_t)ntiles * RS_BINS;
    uint64_t *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    bool sort_digit[8] = {false};   // the digits sorted: all but those constant over the keys
    int last_d = -1;
    for (int d = digit_lo; d < digit_hi; ++d) {
        bool constant = false;
        for (int b = 0; skip_constant && b < RS_BINS; ++b)
            if (h8[d * RS_BINS + b] == n) { constant = true; break; }
        sort_digit[d] = !constant;
        if (!constant) last_d = d;
    }
    for (int d = digit_lo; d < digit_hi; ++d) {
        if (!sort_digit[d]) continue;
        const int shift = 8 * d;
        k_rs_tile_hist<<<ntiles, RS_THREADS, 0, st>>>(ka, n, shift, tmp.th, ntiles);
        LAUNCH_CHECK();
        int rc = device_excl_scan<uint32_t, uint32_t>(tmp.th, nth, tmp.tb, nullptr, tmp.ws, st);
        if (rc != SI_OK) return rc;
        if (last && d == last_d) {
            k_rs_scatter<LastOut><<<ntiles, RS_THREADS, 0, st>>>(ka, va, *last, n, shift, tmp.tb, ntiles);
            if (last_written) *last_written = true;
        } else {
            k_rs_scatter<ArrayOut><<<ntiles, RS_THREADS, 0, st>>>(ka, va, ArrayOut{kb, vb}, n, shift, tmp.tb, ntiles);
        }
        LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    *keys_sorted = ka;
    *vals_sorted = va;
    return SI_OK;
}

}  // namespace
}  // namespace sik
