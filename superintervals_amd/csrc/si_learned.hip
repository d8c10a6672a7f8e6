// si_learned.hip -- learned position index (RMI) for upper_bound.
//
// No counterpart in the reference: it replaces most of the binary search of
// upper_bound (superintervals.hpp:501-516) on a built index and keeps the result
// bit-exact by construction.
//
// Model (FP32), trained from the index's own (start -> rank) pairs:
//   root   x = (key - key_min) / (key_max - key_min);  h1 = relu(W1 x + b1) (16),
//          h2 = relu(W2 h1 + b2) (16),  y = w3 . h2 + b3  ~ rank / n;
//          predicted leaf = clamp(floor(y * L), 0, L - 1) with L = 4096 leaves.
//          Initialised as the identity CDF (h1_j = relu(x - j/16), W2 = I, y = h2_0),
//          i.e. the untrained network predicts the leaf from x alone.
//   leaf   leaf l owns positions [l n / L, (l + 1) n / L); p = base + c + a (key - key0)
//          in leaf-local integer offsets so FP32 keeps full precision, a and c fitted by
//          least squares over the leaf's keys; [elo, ehi] the exact bounds of
//          (upper_bound - p) over every key the leaf serves.
// Training: minibatch MSE on (x_i, i / n) with hand-written forward + backward kernels
// (per-block partial gradients, then a fixed-order reduction: deterministic) and Adam.
// Query: the root's leaf, corrected by a galloping search of the leaves' first keys
// (LDS) to the leaf that holds upper_bound(x); a binary search of starts in the leaf's
// window; a one-read check of the result, with a whole-leaf search as the fallback.  A
// better root means fewer correction steps; the output never depends on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "si_internal.h"
#include "si_prims.h"

namespace {

constexpr int H = 16;                          // hidden width
constexpr int NPARAM = H + H + H * H + H + H + 1;   // W1, b1, W2, b2, w3, b3 = 321
constexpr int OFF_W1 = 0, OFF_B1 = H, OFF_W2 = 2 * H, OFF_B2 = 2 * H + H * H, OFF_W3 = 3 * H + H * H,
              OFF_B3 = 4 * H + H * H;
constexpr int RMI_THREADS = 256;
constexpr int GRAD_SLOTS = NPARAM + 1;         // gradient + loss

struct Leaf {
    int64_t key0;   // the leaf's first key
    uint32_t base;  // the leaf's first position
    uint32_t last;  // the leaf's last position
    float a, c;     // p = base + c + a * (key - key0)
    int32_t elo, ehi;
};

__device__ __forceinline__ float root_forward(const float* __restrict__ P, float x) {
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) h1[j] = fmaxf(fmaf(P[OFF_W1 + j], x, P[OFF_B1 + j]), 0.f);
    float y = P[OFF_B3];
#pragma unroll
    for (int k = 0; k < H; ++k) {
        float acc = P[OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < H; ++j) acc = fmaf(P[OFF_W2 + k * H + j], h1[j], acc);
        y = fmaf(P[OFF_W3 + k], fmaxf(acc, 0.f), y);
    }
    return y;
}

__device__ __forceinline__ uint32_t leaf_of(float y, uint32_t L) {
    float f = y * (float)L;
    if (!(f > 0.f)) return 0;   // also catches NaN
    if (f >= (float)(L - 1)) return L - 1;
    return (uint32_t)f;
}

template <typename S>
__device__ __forceinline__ float norm_key(S key, int64_t kmin, float inv_span) {
    return (float)((double)((int64_t)key - kmin) * (double)inv_span);
}

__device__ __forceinline__ int64_t leaf_pred(const Leaf& lf, int64_t key) {
    const float d = (float)(key - lf.key0);
    const float f = fminf(fmaxf(floorf(fmaf(lf.a, d, lf.c)), -1e15f), 1e15f);   // keep the cast defined
    return (int64_t)lf.base + (int64_t)f;
}

// Training samples: sample j of step t is position hash(t, j) mod n.
__device__ __forceinline__ uint32_t sample_pos(uint32_t step, uint32_t j, uint32_t n) {
    uint64_t z = ((uint64_t)step << 32 | j) * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)((z >> 32) * (uint64_t)n >> 32);
}

// Forward + backward over B samples; block b writes its partial sums of dLoss/dparam and
// of the loss to part[b * GRAD_SLOTS ...].  Loss = mean over the batch of (y - t)^2.
// Per chunk of CH samples: phase 1, one thread per sample, runs the forward pass and the
// backward pass down to the pre-activation deltas and stores them in LDS; phase 2, one
// thread per parameter, sums that parameter's gradient over the chunk in sample order.
constexpr int CH = 128;

template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_rmi_grad(const S* __restrict__ starts, uint32_t n,
                                                          int64_t kmin, float inv_span,
                                                          const float* __restrict__ P, uint32_t step,
                                                          uint32_t B, const uint32_t* __restrict__ explicit_pos,
                                                          float* __restrict__ part) {
    __shared__ float s_p[NPARAM];
    __shared__ float s_x[CH], s_dy[CH], s_r2[CH];
    __shared__ float s_h1[CH][H], s_h2[CH][H], s_d1[CH][H], s_d2[CH][H];
    for (int k = threadIdx.x; k < NPARAM; k += RMI_THREADS) s_p[k] = P[k];
    const float invB = 1.f / (float)B;
    // each thread owns parameters k0 = tid and k1 = tid + RMI_THREADS (slot NPARAM = loss)
    const int k0 = threadIdx.x, k1 = threadIdx.x + RMI_THREADS;
    float acc0 = 0.f, acc1 = 0.f;
    for (uint32_t c0 = blockIdx.x * CH; c0 < B; c0 += gridDim.x * CH) {
        __syncthreads();
        if (threadIdx.x < CH) {
            const uint32_t j = c0 + threadIdx.x;
            const int t_ = threadIdx.x;
            if (j < B) {
                const uint32_t i = explicit_pos ? explicit_pos[j] : sample_pos(step, j, n);
                const float x = norm_key(starts[i], kmin, inv_span);
                const float t = (float)i / (float)n;
                float h1[H], h2[H];
#pragma unroll
                for (int a = 0; a < H; ++a) h1[a] = fmaxf(fmaf(s_p[OFF_W1 + a], x, s_p[OFF_B1 + a]), 0.f);
                float y = s_p[OFF_B3];
#pragma unroll
                for (int k = 0; k < H; ++k) {
                    float acc = s_p[OFF_B2 + k];
#pragma unroll
                    for (int a = 0; a < H; ++a) acc = fmaf(s_p[OFF_W2 + k * H + a], h1[a], acc);
                    h2[k] = fmaxf(acc, 0.f);
                    y = fmaf(s_p[OFF_W3 + k], h2[k], y);
                }
                const float r = y - t;
                const float dy = 2.f * r * invB;
                float d2[H], dh1[H];
#pragma unroll
                for (int a = 0; a < H; ++a) dh1[a] = 0.f;
#pragma unroll
                for (int k = 0; k < H; ++k) {
                    d2[k] = (h2[k] > 0.f) ? dy * s_p[OFF_W3 + k] : 0.f;
#pragma unroll
                    for (int a = 0; a < H; ++a) dh1[a] = fmaf(d2[k], s_p[OFF_W2 + k * H + a], dh1[a]);
                }
                s_x[t_] = x;
                s_dy[t_] = dy;
                s_r2[t_] = r * r * invB;
#pragma unroll
                for (int a = 0; a < H; ++a) {
                    s_h1[t_][a] = h1[a];
                    s_h2[t_][a] = h2[a];
                    s_d2[t_][a] = d2[a];
                    s_d1[t_][a] = (h1[a] > 0.f) ? dh1[a] : 0.f;
                }
            } else {
                s_x[t_] = 0.f; s_dy[t_] = 0.f; s_r2[t_] = 0.f;
#pragma unroll
                for (int a = 0; a < H; ++a) { s_h1[t_][a] = 0.f; s_h2[t_][a] = 0.f; s_d2[t_][a] = 0.f; s_d1[t_][a] = 0.f; }
            }
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)CH, B - c0);
        for (int which = 0; which < 2; ++which) {
            const int k = which ? k1 : k0;
            if (k > NPARAM) continue;
            float g = 0.f;
            for (uint32_t t_ = 0; t_ < m; ++t_) {
                float term;
                if (k < OFF_B1) term = s_d1[t_][k] * s_x[t_];                               // W1
                else if (k < OFF_W2) term = s_d1[t_][k - OFF_B1];                           // b1
                else if (k < OFF_B2) term = s_d2[t_][(k - OFF_W2) / H] * s_h1[t_][(k - OFF_W2) % H];   // W2
                else if (k < OFF_W3) term = s_d2[t_][k - OFF_B2];                           // b2
                else if (k < OFF_B3) term = s_dy[t_] * s_h2[t_][k - OFF_W3];               // w3
                else if (k == OFF_B3) term = s_dy[t_];                                      // b3
                else term = s_r2[t_];                                                       // loss
                g += term;
            }
            if (which) acc1 += g;
            else acc0 += g;
        }
    }
    if (k0 <= NPARAM) part[(size_t)blockIdx.x * GRAD_SLOTS + k0] = acc0;
    if (k1 <= NPARAM) part[(size_t)blockIdx.x * GRAD_SLOTS + k1] = acc1;
}

// grad[k] = sum over blocks of part[b][k], in block order.
__global__ void k_rmi_reduce(const float* __restrict__ part, int nblocks, float* __restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GRAD_SLOTS) return;
    float v = 0.f;
    for (int b = 0; b < nblocks; ++b) v += part[(size_t)b * GRAD_SLOTS + k];
    grad[k] = v;
}

__global__ void k_rmi_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM) return;
    const float gk = grad[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    P[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// Leaves own contiguous position ranges: leaf l holds positions [P(l), P(l + 1)) with
// P(l) = l * n / L, so a leaf never mixes keys from separate parts of the index, whatever
// the root predicts.  The root only has to name the leaf; the query corrects that guess
// by a galloping search of the leaves' first keys (kept in LDS).
__host__ __device__ __forceinline__ uint32_t leaf_begin(uint32_t l, uint32_t n, uint32_t L) {
    return (uint32_t)(((uint64_t)l * n) / L);
}

// The leaf holding position i: the largest l with P(l) <= i.
__device__ __forceinline__ uint32_t leaf_of_pos(uint32_t i, uint32_t n, uint32_t L) {
    return (uint32_t)((((uint64_t)i + 1) * L - 1) / n);
}

// The leaf's prediction, clamped to its positions.  Monotone non-decreasing in the key
// (slope >= 0, float rounding, fma, floor and clamp are all monotone), which the error
// bounds below rely on.
__device__ __forceinline__ int64_t leaf_pred_c(const Leaf& lf, int64_t key, int64_t lo, int64_t hi) {
    return std::min<int64_t>(std::max<int64_t>(leaf_pred(lf, key), lo), hi);
}

// Pass 1: least squares of (position - P(l)) on (key - key0) over the leaf's positions, in
// double; one block per leaf, so no atomics.  kb[l] receives the leaf's first key.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_leaf_fit(const S* __restrict__ starts, uint32_t n, uint32_t L,
                                                          Leaf* __restrict__ leaves, S* __restrict__ kb) {
    __shared__ double s_v[5][RMI_THREADS / WAVE];
    const uint32_t l = blockIdx.x;
    const uint32_t b = leaf_begin(l, n, L), e = leaf_begin(l + 1, n, L);
    const int64_t key0 = (int64_t)starts[b];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = b + threadIdx.x; i < e; i += RMI_THREADS) {
        const double d = (double)((int64_t)starts[i] - key0), r = (double)(i - b);
        v[0] += 1.0; v[1] += d; v[2] += r; v[3] += d * d; v[4] += d * r;
    }
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double x = v[c];
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) x += __shfl_xor(x, sh, WAVE);
        if (lane_id() == 0) s_v[c][w] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[5];
        for (int c = 0; c < 5; ++c) {
            t[c] = 0.0;
            for (int k = 0; k < RMI_THREADS / WAVE; ++k) t[c] += s_v[c][k];
        }
        const double cnt = t[0];
        const double md = t[1] / cnt, mr = t[2] / cnt;
        const double var = t[3] / cnt - md * md, cov = t[4] / cnt - md * mr;
        const double a = (var > 0.0 && cov > 0.0) ? cov / var : 0.0;
        Leaf lf;
        lf.key0 = key0;
        lf.base = b;
        lf.last = e - 1;
        lf.a = (float)a;
        lf.c = (float)(mr - a * md);
        lf.elo = INT32_MAX;
        lf.ehi = INT32_MIN;
        leaves[l] = lf;
        kb[l] = (S)key0;
    }
}

// Pass 2: exact bounds [elo, ehi] of upper_bound(x) - pred(x) over every key x the leaf
// serves (kb[l] <= x < kb[l + 1]).  upper_bound is constant between consecutive distinct
// keys and pred is monotone, so the extremes sit at the ends of those pieces: at each key's
// last position (ub = that position), just below each key's first position (ub = first - 1,
// pred <= pred(key)), and just below the next leaf's first key (ub = the leaf's last
// position; for the last leaf, keys above every start clamp to it as well).
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_leaf_err(const S* __restrict__ starts, uint32_t n, uint32_t L,
                                                          Leaf* __restrict__ leaves) {
    const uint32_t i = blockIdx.x * RMI_THREADS + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = leaf_of_pos(i, n, L);
    const uint32_t pl = leaf_begin(l, n, L), ph = leaf_begin(l + 1, n, L) - 1;
    const Leaf lf = leaves[l];
    const S k = starts[i];
    const int64_t p = leaf_pred_c(lf, (int64_t)k, pl, ph);
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (i + 1 == n || starts[i + 1] != k) lo = hi = (int64_t)i - p;
    if (i > pl && starts[i - 1] != k) {
        const int64_t e1 = (int64_t)i - 1 - p;
        lo = std::min(lo, e1);
        hi = std::max(hi, e1);
    }
    if (i == ph) {   // the top piece of the leaf: keys up to the next leaf's first key
        const int64_t ptop = (i + 1 < n) ? leaf_pred_c(lf, (int64_t)starts[i + 1] - 1, pl, ph) : (int64_t)ph;
        const int64_t e2 = (int64_t)ph - ptop;
        lo = std::min(lo, e2);
        hi = std::max(hi, e2);
    }
    if (lo > hi) return;
    lo = std::max<int64_t>(lo, INT32_MIN / 2);
    hi = std::min<int64_t>(hi, INT32_MAX / 2);
    atomicMin(&leaves[l].elo, (int32_t)lo);
    atomicMax(&leaves[l].ehi, (int32_t)hi);
}

// Last r in [lo, hi] with starts[r] <= x, or lo - 1 (branchless halving).
template <typename S>
__device__ __forceinline__ int64_t last_le(const S* __restrict__ st, int64_t lo, int64_t hi, S x) {
    const S* base = st + lo;
    uint32_t len = (uint32_t)(hi - lo + 1);
    while (len > 1) {
        const uint32_t half = len >> 1;
        base = (base[half] <= x) ? base + half : base;
        len -= half;
    }
    return (int64_t)(base - st) - ((*base <= x) ? 0 : 1);
}

// The leaf holding upper_bound(x): from the guess lg, gallop through the leaves' first keys
// (LDS) to the bracket s_kb[lo] <= x < s_kb[hi], then halve it.  -1 when x is below every
// first key.
template <typename S>
__device__ __forceinline__ int32_t correct_leaf(const S* __restrict__ s_kb, uint32_t L, int32_t lg, S x) {
    int32_t lo, hi;
    if (s_kb[lg] <= x) {
        lo = lg;
        hi = lg + 1;
        int32_t step = 1;
        while (hi < (int32_t)L && s_kb[hi] <= x) {
            lo = hi;
            step <<= 1;
            hi = lo + step;
        }
        hi = std::min(hi, (int32_t)L);
    } else {
        hi = lg;
        lo = lg - 1;
        int32_t step = 1;
        while (lo >= 0 && s_kb[lo] > x) {
            hi = lo;
            step <<= 1;
            lo = hi - step;
        }
        lo = std::max(lo, -1);
    }
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (s_kb[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// upper_bound(x) inside leaf lo (whose first key is <= x): the leaf's prediction window
// (narrowed by shrink per side, a test of the fallback), a binary search of starts in it,
// a one-read check that the result is the last start <= x, and the whole leaf as fallback.
template <typename S>
__device__ __forceinline__ uint32_t leaf_search(const IndexView<S>& ix, const Leaf* __restrict__ leaves, int32_t lo,
                                                S x, int32_t shrink, unsigned long long& f_acc,
                                                unsigned long long& w_acc) {
    if (lo < 0) return NONE;
    const Leaf lf = leaves[lo];
    const int64_t pl = lf.base, ph = lf.last;
    const int64_t p = leaf_pred_c(lf, (int64_t)x, pl, ph);
    const int64_t wl = std::max<int64_t>(p + lf.elo + shrink, pl);
    const int64_t wh = std::min<int64_t>(p + lf.ehi - shrink, ph);
    bool ok = false;
    int64_t r = pl - 1;
    if (wl <= wh) {
        w_acc += (unsigned long long)(wh - wl + 1);
        r = last_le(ix.starts, wl, wh, x);
        ok = (r >= wl) && (r < wh || wh == ph || ix.starts[wh + 1] > x);
    }
    if (!ok) {
        f_acc += 1;
        r = last_le(ix.starts, pl, ph, x);   // starts[pl] <= x, so r >= pl
    }
    return (uint32_t)r;
}

// Persistent blocks stage the root's parameters and the leaves' first keys in LDS.
template <typename S>
__device__ __forceinline__ void stage_model(float* s_p, S* s_kb, const float* __restrict__ P, const S* __restrict__ kb,
                                            uint32_t L) {
    for (int k = threadIdx.x; k < NPARAM; k += RMI_THREADS) s_p[k] = P[k];
    for (uint32_t k = threadIdx.x; k < L; k += RMI_THREADS) s_kb[k] = kb[k];
    __syncthreads();
}

template <typename S>
__device__ __forceinline__ void add_stats(unsigned long long* stats, unsigned long long f_acc,
                                          unsigned long long w_acc, unsigned long long c_acc) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        f_acc += __shfl_xor(f_acc, d, WAVE);
        w_acc += __shfl_xor(w_acc, d, WAVE);
        c_acc += __shfl_xor(c_acc, d, WAVE);
    }
    if (lane_id() == 0) {
        atomicAdd(&stats[0], f_acc);
        atomicAdd(&stats[1], w_acc);
        atomicAdd(&stats[2], c_acc);
    }
}

// Learned upper_bound.  The root names a leaf; the first keys correct it to the leaf holding
// upper_bound(x); the leaf's window, which provably contains the answer, is searched, with a
// one-read check and the whole leaf as fallback.  stats (optional): [0] fallbacks, [1] window
// widths, [2] leaf corrections.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_ub_learned(IndexView<S> ix, int64_t kmin, float inv_span,
                                                            const float* __restrict__ P,
                                                            const Leaf* __restrict__ leaves,
                                                            const S* __restrict__ kb, uint32_t L,
                                                            const S* __restrict__ xs, size_t nx,
                                                            uint32_t* __restrict__ out,
                                                            unsigned long long* __restrict__ stats, int32_t shrink) {
    extern __shared__ __align__(16) char s_mem[];
    float* s_p = reinterpret_cast<float*>(s_mem);
    S* s_kb = reinterpret_cast<S*>(s_mem + ((NPARAM * 4 + 15) & ~15));
    stage_model(s_p, s_kb, P, kb, L);
    unsigned long long f_acc = 0, w_acc = 0, c_acc = 0;
    for (size_t q0 = (size_t)blockIdx.x * RMI_THREADS; q0 < nx; q0 += (size_t)gridDim.x * RMI_THREADS) {
        const size_t q = q0 + threadIdx.x;
        if (q >= nx) continue;
        const S x = xs[q];
        const int32_t lg = (int32_t)leaf_of(root_forward(s_p, norm_key(x, kmin, inv_span)), L);
        const int32_t lo = correct_leaf(s_kb, L, lg, x);
        if (lo >= 0) c_acc += (unsigned long long)(lo > lg ? lo - lg : lg - lo);
        out[q] = leaf_search(ix, leaves, lo, x, shrink, f_acc, w_acc);
    }
    if (stats) add_stats<S>(stats, f_acc, w_acc, c_acc);
}

// The plan's two bounds through the model with one root evaluation per query: idx[k] =
// upper_bound(e[k]) from the root's leaf, and lb[k] = the first position starting at or after
// s[k] = upper_bound(s[k] - 1) + 1 (0 when s[k] is the smallest value), its leaf corrected
// from e's leaf (the query's start is close below its end).  The leaf correction is exact
// from any guess, so both results equal the plain search's.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_plan_bounds_learned(IndexView<S> ix, int64_t kmin, float inv_span,
                                                                     const float* __restrict__ P,
                                                                     const Leaf* __restrict__ leaves,
                                                                     const S* __restrict__ kb, uint32_t L,
                                                                     const S* __restrict__ qe,
                                                                     const S* __restrict__ qs, size_t nq,
                                                                     uint32_t* __restrict__ idx,
                                                                     uint32_t* __restrict__ lb, int32_t shrink) {
    constexpr S SMIN = (sizeof(S) == 4) ? (S)INT32_MIN : (S)INT64_MIN;
    extern __shared__ __align__(16) char s_mem[];
    float* s_p = reinterpret_cast<float*>(s_mem);
    S* s_kb = reinterpret_cast<S*>(s_mem + ((NPARAM * 4 + 15) & ~15));
    stage_model(s_p, s_kb, P, kb, L);
    unsigned long long f_acc = 0, w_acc = 0;
    for (size_t q0 = (size_t)blockIdx.x * RMI_THREADS; q0 < nq; q0 += (size_t)gridDim.x * RMI_THREADS) {
        const size_t q = q0 + threadIdx.x;
        if (q >= nq) continue;
        const S x = qe[q];
        const int32_t lg = (int32_t)leaf_of(root_forward(s_p, norm_key(x, kmin, inv_span)), L);
        const int32_t lo = correct_leaf(s_kb, L, lg, x);
        idx[q] = leaf_search(ix, leaves, lo, x, shrink, f_acc, w_acc);
        if (lb) {
            const S s = qs[q];
            if (s == SMIN) {
                lb[q] = 0u;
            } else {
                const S x2 = (S)(s - 1);
                const int32_t lo2 = correct_leaf(s_kb, L, lo < 0 ? 0 : lo, x2);
                lb[q] = leaf_search(ix, leaves, lo2, x2, shrink, f_acc, w_acc) + 1u;   // NONE + 1 wraps to 0
            }
        }
    }
}

}  // namespace

// ============================================================================
// The model attached to an index
// ============================================================================
struct siLearned {
    uint32_t L = 4096;
    uint64_t generation = 0;   // the index build it was trained on
    int64_t kmin = 0;
    float inv_span = 1.f;
    float* d_params = nullptr;   // NPARAM
    Leaf* d_leaves = nullptr;    // L
    void* d_kb = nullptr;        // L first keys of the leaves (coordinate type)
    bool trained = false;
    int steps_done = 0;
    float last_loss = 0.f;
    int32_t max_abs_err = 0;
    double mean_window = 0.0;
    const siIndex* ix = nullptr;
    size_t ix_n = 0;
    ~siLearned() {
        (void)hipFree(d_params);
        (void)hipFree(d_leaves);
        (void)hipFree(d_kb);
    }
};

namespace {

void init_params(float* P) {
    std::memset(P, 0, NPARAM * sizeof(float));
    for (int j = 0; j < H; ++j) {
        P[OFF_W1 + j] = 1.f;
        P[OFF_B1 + j] = -(float)j / (float)H;
    }
    for (int k = 0; k < H; ++k) P[OFF_W2 + k * H + k] = 1.f;
    P[OFF_W3 + 0] = 1.f;   // y = h2_0 = relu(x): the identity CDF
}

template <typename S>
int fit_leaves(siLearned* m, const siIndex* ix, hipStream_t st) {
    const uint32_t n = (uint32_t)ix->n, L = m->L;
    IndexView<S> v = ix->view<S>();
    k_leaf_fit<S><<<L, RMI_THREADS, 0, st>>>(v.starts, n, L, m->d_leaves, (S*)m->d_kb);
    LAUNCH_CHECK();
    k_leaf_err<S><<<grid_for(n, RMI_THREADS), RMI_THREADS, 0, st>>>(v.starts, n, L, m->d_leaves);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<Leaf> h(L);
    HIP_TRY(hipMemcpy(h.data(), m->d_leaves, (size_t)L * sizeof(Leaf), hipMemcpyDeviceToHost));
    int32_t mx = 0;
    for (const Leaf& lf : h)
        if (lf.elo <= lf.ehi) mx = std::max(mx, std::max(std::abs(lf.elo), std::abs(lf.ehi)));
    m->max_abs_err = mx;
    return SI_OK;
}

template <typename S>
int train_impl(siLearned* m, const siIndex* ix, int steps, uint32_t batch, float lr, hipStream_t st) {
    const uint32_t n = (uint32_t)ix->n;
    IndexView<S> v = ix->view<S>();
    S lo, hi;
    HIP_TRY(hipMemcpy(&lo, v.starts, sizeof(S), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&hi, v.starts + (n - 1), sizeof(S), hipMemcpyDeviceToHost));
    m->kmin = (int64_t)lo;
    m->inv_span = (hi > lo) ? (float)(1.0 / ((double)hi - (double)lo)) : 1.f;
    float P[NPARAM];
    init_params(P);
    HIP_TRY(hipMemcpy(m->d_params, P, sizeof(P), hipMemcpyHostToDevice));
    if (steps > 0) {
        const int nblocks = std::max(1, std::min<int>(256, (int)((batch + RMI_THREADS - 1) / RMI_THREADS)));
        float *d_part = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
        HIP_TRY(hipMalloc(&d_part, (size_t)nblocks * GRAD_SLOTS * 4));
        HIP_TRY(hipMalloc(&d_grad, GRAD_SLOTS * 4));
        HIP_TRY(hipMalloc(&d_m, NPARAM * 4));
        HIP_TRY(hipMalloc(&d_v, NPARAM * 4));
        HIP_TRY(hipMemsetAsync(d_m, 0, NPARAM * 4, st));
        HIP_TRY(hipMemsetAsync(d_v, 0, NPARAM * 4, st));
        const float b1 = 0.9f, b2 = 0.999f;
        double p1 = 1.0, p2 = 1.0;
        for (int t = 0; t < steps; ++t) {
            k_rmi_grad<S><<<nblocks, RMI_THREADS, 0, st>>>(v.starts, n, m->kmin, m->inv_span, m->d_params,
                                                           (uint32_t)t, batch, nullptr, d_part);
            LAUNCH_CHECK();
            k_rmi_reduce<<<grid_for(GRAD_SLOTS, 256), 256, 0, st>>>(d_part, nblocks, d_grad);
            LAUNCH_CHECK();
            p1 *= b1;
            p2 *= b2;
            k_rmi_adam<<<grid_for(NPARAM, 256), 256, 0, st>>>(m->d_params, d_m, d_v, d_grad, lr, b1, b2, 1e-8f,
                                                              (float)(1.0 - p1), (float)(1.0 - p2));
            LAUNCH_CHECK();
        }
        float loss = 0.f;
        HIP_TRY(hipMemcpyAsync(&loss, d_grad + NPARAM, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        m->last_loss = loss;
        hipFree(d_part);
        hipFree(d_grad);
        hipFree(d_m);
        hipFree(d_v);
    }
    m->steps_done = steps;
    return fit_leaves<S>(m, ix, st);
}

}  // namespace

// ============================================================================
// C entry points
// ============================================================================
SI_EXPORT siLearned* siLearnedCreate(const siIndex* ix, uint32_t leaves) {
    if (check_ix(ix) != SI_OK) return nullptr;
    if (ix->n < 2) { si_set_error(SI_ERR_ARG, "a learned index needs at least 2 intervals"); return nullptr; }
    if (leaves < 1 || leaves > 8192) { si_set_error(SI_ERR_ARG, "leaves must be in [1, 8192]"); return nullptr; }
    siLearned* m = new siLearned();
    m->L = (uint32_t)std::min<size_t>(leaves, ix->n);   // every leaf holds at least one position
    m->ix = ix;
    const_cast<siIndex*>(ix)->models.push_back(m);
    m->ix_n = ix->n;
    m->generation = ix->generation;
    if (hipMalloc(&m->d_params, NPARAM * 4) != hipSuccess ||
        hipMalloc(&m->d_leaves, (size_t)m->L * sizeof(Leaf)) != hipSuccess ||
        hipMalloc(&m->d_kb, (size_t)m->L * ix->cb) != hipSuccess) {
        si_set_error(SI_ERR_NOMEM, "learned index allocation");
        siLearnedDestroy(m);
        return nullptr;
    }
    return m;
}

SI_EXPORT void siLearnedDestroy(siLearned* m) {
    if (!m) return;
    if (m->ix) {   // the index still exists: unlink the model from it
        siIndex* ix = const_cast<siIndex*>(m->ix);
        if (ix->learned == m) ix->learned = nullptr;
        ix->models.erase(std::remove(ix->models.begin(), ix->models.end(), m), ix->models.end());
    }
    delete m;
}

void si_learned_forget_index(siLearned* m) {
    if (m) m->ix = nullptr;
}

SI_EXPORT int siLearnedParamCount(void) { return NPARAM; }

SI_EXPORT int siLearnedTrain(siLearned* m, int steps, uint32_t batch, float lr, void* stream) {
    if (!m) return si_set_error(SI_ERR_ARG, "null model");
    if (steps < 0 || batch == 0) return si_set_error(SI_ERR_ARG, "bad training arguments");
    const siIndex* ix = m->ix;
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->generation != m->generation) return si_set_error(SI_ERR_ARG, "the index was rebuilt; create a new model");
    hipStream_t st = (hipStream_t)stream;
    rc = (ix->cb == 4) ? train_impl<int32_t>(m, ix, steps, batch, lr, st)
                       : train_impl<int64_t>(m, ix, steps, batch, lr, st);
    if (rc == SI_OK) {
        m->trained = true;
        const_cast<siIndex*>(ix)->learned = m;   // the plan may use it (tuning knob ub_learned)
    }
    return rc;
}

SI_EXPORT int siLearnedGetParams(const siLearned* m, float* params) {
    if (!m || !params) return si_set_error(SI_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpy(params, m->d_params, NPARAM * 4, hipMemcpyDeviceToHost));
    return SI_OK;
}

// stats: [0] max |error| over keys, [1] final training loss (as float bits), [2] non-empty leaves
SI_EXPORT int siLearnedInfo(const siLearned* m, int64_t* max_abs_err, float* last_loss, uint32_t* used_leaves) {
    if (!m) return si_set_error(SI_ERR_ARG, "null model");
    if (max_abs_err) *max_abs_err = m->max_abs_err;
    if (last_loss) *last_loss = m->last_loss;
    if (used_leaves) {
        std::vector<Leaf> h(m->L);
        HIP_TRY(hipMemcpy(h.data(), m->d_leaves, (size_t)m->L * sizeof(Leaf), hipMemcpyDeviceToHost));
        uint32_t u = 0;
        for (const Leaf& lf : h) u += (lf.elo <= lf.ehi) ? 1u : 0u;
        *used_leaves = u;
    }
    return SI_OK;
}

// Gradient of the training loss at the given parameters over explicit sample positions
// (for checking the backward kernel).  grad receives NPARAM + 1 floats (the last is the
// loss).  params may be NULL (the model's current parameters).
SI_EXPORT int siLearnedGradient(siLearned* m, const float* params, const uint32_t* pos, uint32_t count,
                                float* grad) {
    if (!m || !pos || !grad || count == 0) return si_set_error(SI_ERR_ARG, "bad arguments");
    const siIndex* ix = m->ix;
    int rc = check_ix(ix);
    if (rc) return rc;
    for (uint32_t k = 0; k < count; ++k)
        if (pos[k] >= ix->n) return si_set_error(SI_ERR_ARG, "sample position out of range");
    float *d_p = nullptr, *d_part = nullptr, *d_grad = nullptr;
    uint32_t* d_pos = nullptr;
    const int nblocks = std::max(1, std::min<int>(256, (int)((count + RMI_THREADS - 1) / RMI_THREADS)));
    HIP_TRY(hipMalloc(&d_p, NPARAM * 4));
    HIP_TRY(hipMalloc(&d_pos, (size_t)count * 4));
    HIP_TRY(hipMalloc(&d_part, (size_t)nblocks * GRAD_SLOTS * 4));
    HIP_TRY(hipMalloc(&d_grad, GRAD_SLOTS * 4));
    if (params) HIP_TRY(hipMemcpy(d_p, params, NPARAM * 4, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemcpy(d_p, m->d_params, NPARAM * 4, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(d_pos, pos, (size_t)count * 4, hipMemcpyHostToDevice));
    if (ix->cb == 4)
        k_rmi_grad<int32_t><<<nblocks, RMI_THREADS>>>((const int32_t*)ix->d_starts, (uint32_t)ix->n, m->kmin,
                                                      m->inv_span, d_p, 0, count, d_pos, d_part);
    else
        k_rmi_grad<int64_t><<<nblocks, RMI_THREADS>>>((const int64_t*)ix->d_starts, (uint32_t)ix->n, m->kmin,
                                                      m->inv_span, d_p, 0, count, d_pos, d_part);
    LAUNCH_CHECK();
    k_rmi_reduce<<<grid_for(GRAD_SLOTS, 256), 256>>>(d_part, nblocks, d_grad);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpy(grad, d_grad, GRAD_SLOTS * 4, hipMemcpyDeviceToHost));
    hipFree(d_p);
    hipFree(d_pos);
    hipFree(d_part);
    hipFree(d_grad);
    return SI_OK;
}

// Normalisation used by the model: x = (key - kmin) * inv_span.
SI_EXPORT int siLearnedNorm(const siLearned* m, int64_t* kmin, float* inv_span) {
    if (!m) return si_set_error(SI_ERR_ARG, "null model");
    if (kmin) *kmin = m->kmin;
    if (inv_span) *inv_span = m->inv_span;
    return SI_OK;
}

namespace {
unsigned learned_grid(size_t n);
int launch_learned(const siLearned* m, const siIndex* ix, const void* xs, size_t n, uint32_t* out,
                   unsigned long long* stats, hipStream_t st);
}  // namespace

// upper_bound through the model (device buffers).  stats (device, may be NULL) receives
// [0] += fallbacks to the whole-leaf search, [1] += sum of window widths, [2] += sum of
// |leaf - root's leaf| corrections.
SI_EXPORT int siUpperBoundLearned(const siLearned* m, const void* xs, size_t n, uint32_t* out,
                                  uint64_t* stats, void* stream) {
    if (!m || !m->trained) return si_set_error(SI_ERR_ARG, "model not trained");
    const siIndex* ix = m->ix;
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->generation != m->generation) return si_set_error(SI_ERR_ARG, "the index was rebuilt; retrain the model");
    if (n == 0) return SI_OK;
    if (!xs || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    return launch_learned(m, ix, xs, n, out, (unsigned long long*)stats, (hipStream_t)stream);
}

int si_learned_plan_bounds(const siIndex* ix, const void* qe, const void* qs, size_t nq, uint32_t* idx,
                           uint32_t* lb, void* stream) {
    const siLearned* m = ix ? ix->learned : nullptr;
    if (!m || !m->trained || m->generation != ix->generation)
        return si_set_error(SI_ERR_ARG, "ub_learned: the index has no trained learned model of its current build");
    if (nq == 0) return SI_OK;
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = learned_grid(nq);
    const size_t lds = ((NPARAM * 4 + 15) & ~15) + (size_t)m->L * ix->cb;
    const int32_t shrink = (int32_t)si_tune_learned_shrink();
    if (ix->cb == 4)
        k_plan_bounds_learned<int32_t><<<g, RMI_THREADS, lds, st>>>(
            ix->view<int32_t>(), m->kmin, m->inv_span, m->d_params, m->d_leaves, (const int32_t*)m->d_kb, m->L,
            (const int32_t*)qe, (const int32_t*)qs, nq, idx, lb, shrink);
    else
        k_plan_bounds_learned<int64_t><<<g, RMI_THREADS, lds, st>>>(
            ix->view<int64_t>(), m->kmin, m->inv_span, m->d_params, m->d_leaves, (const int64_t*)m->d_kb, m->L,
            (const int64_t*)qe, (const int64_t*)qs, nq, idx, lb, shrink);
    LAUNCH_CHECK();
    return SI_OK;
}

namespace {
// Persistent grid: enough blocks to fill the GPU, each stages the model once.
unsigned learned_grid(size_t n) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cus = std::max(cus, 1);
    }
    return (unsigned)std::min<size_t>(grid_for(n, RMI_THREADS), (size_t)cus * 8);
}

int launch_learned(const siLearned* m, const siIndex* ix, const void* xs, size_t n, uint32_t* out,
                   unsigned long long* stats, hipStream_t st) {
    const unsigned g = learned_grid(n);
    const size_t lds = ((NPARAM * 4 + 15) & ~15) + (size_t)m->L * ix->cb;
    const int32_t shrink = (int32_t)si_tune_learned_shrink();
    if (ix->cb == 4)
        k_ub_learned<int32_t><<<g, RMI_THREADS, lds, st>>>(ix->view<int32_t>(), m->kmin, m->inv_span, m->d_params,
                                                           m->d_leaves, (const int32_t*)m->d_kb, m->L,
                                                           (const int32_t*)xs, n, out, stats, shrink);
    else
        k_ub_learned<int64_t><<<g, RMI_THREADS, lds, st>>>(ix->view<int64_t>(), m->kmin, m->inv_span, m->d_params,
                                                           m->d_leaves, (const int64_t*)m->d_kb, m->L,
                                                           (const int64_t*)xs, n, out, stats, shrink);
    LAUNCH_CHECK();
    return SI_OK;
}
}  // namespace
