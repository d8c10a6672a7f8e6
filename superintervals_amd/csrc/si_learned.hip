// si_learned.hip -- learned position index (RMI) for upper_bound.
//
// No counterpart in the reference: it replaces most of the binary search of
// upper_bound (superintervals.hpp:501-516) on a built index and keeps the result
// bit-exact by construction.
//
// Model (FP32), trained from the index's own (start -> rank) pairs:
//   root   x = (key - key_min) / (key_max - key_min);  h1 = relu(W1 x + b1) (16),
//          h2 = relu(W2 h1 + b2) (16),  y = w3 . h2 + b3  ~ rank / n;
//          leaf = clamp(floor(y * L), 0, L - 1) with L = 4096 leaves.
//          Initialised as the identity CDF (h1_j = relu(x - j/16), W2 = I, y = h2_0),
//          i.e. the untrained network routes by x alone.
//   leaf   p = base + c + a * (float)(key - key0), in leaf-local integer offsets so
//          FP32 keeps full precision; a, c fitted by least squares over the keys the
//          root routes to the leaf; [elo, ehi] the exact min / max of
//          (true upper_bound - p) over those keys.
// Training: minibatch MSE on (x_i, i / n) with hand-written forward + backward kernels
// (per-block partial gradients, then a fixed-order reduction: deterministic) and Adam.
// Query: p, then a binary search of starts in [p + elo - 1, p + ehi] (clamped), then a
// check that the result is the global upper_bound (the entries just outside the window
// bracket the key); if not, the plain binary search answers.
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
    int64_t key0;   // smallest key routed to the leaf (INT64_MAX: empty)
    int64_t base;   // smallest position routed to the leaf
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
    return lf.base + (int64_t)f;
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

// Leaf statistics, pass 1: smallest key and position routed to each leaf.  Keys are sorted,
// so the first key routed to a leaf is its smallest.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_leaf_first(const S* __restrict__ starts, uint32_t n,
                                                            int64_t kmin, float inv_span,
                                                            const float* __restrict__ P, uint32_t L,
                                                            uint32_t* __restrict__ leaf_of_key,
                                                            unsigned long long* __restrict__ first_pos) {
    __shared__ float s_p[NPARAM];
    for (int k = threadIdx.x; k < NPARAM; k += RMI_THREADS) s_p[k] = P[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * RMI_THREADS + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = leaf_of(root_forward(s_p, norm_key(starts[i], kmin, inv_span)), L);
    leaf_of_key[i] = l;
    atomicMin(&first_pos[l], (unsigned long long)i);
}

// Pass 2: least-squares sums over (d = key - key0, r = pos - base) per leaf, in double.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_leaf_sums(const S* __restrict__ starts, uint32_t n,
                                                           const uint32_t* __restrict__ leaf_of_key,
                                                           const unsigned long long* __restrict__ first_pos,
                                                           double* __restrict__ sums /*[L][5]*/) {
    const uint32_t i = blockIdx.x * RMI_THREADS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t l = valid ? leaf_of_key[i] : 0xFFFFFFFFu;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid) {
        const uint32_t f = (uint32_t)first_pos[l];
        const double d = (double)((int64_t)starts[i] - (int64_t)starts[f]);
        const double r = (double)(i - f);
        v[0] = 1.0; v[1] = d; v[2] = r; v[3] = d * d; v[4] = d * r;
    }
    // aggregate the wave when all its valid lanes share one leaf (keys are sorted, so
    // leaves are mostly runs); lane 0 is valid whenever any lane is
    const uint32_t l0 = __shfl(l, 0, WAVE);
    const bool uni = __all(!valid || l == l0);
    if (uni) {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            double x = v[c];
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1) x += __shfl_xor(x, s, WAVE);
            v[c] = x;
        }
        if (lane_id() == 0 && l0 != 0xFFFFFFFFu)
            for (int c = 0; c < 5; ++c) atomicAdd(&sums[(size_t)l0 * 5 + c], v[c]);
    } else if (valid) {
        for (int c = 0; c < 5; ++c) atomicAdd(&sums[(size_t)l * 5 + c], v[c]);
    }
}

template <typename S>
__global__ void k_leaf_fit(const S* __restrict__ starts, uint32_t L, const unsigned long long* __restrict__ first_pos,
                           const double* __restrict__ sums, Leaf* __restrict__ leaves) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    Leaf lf;
    const unsigned long long f = first_pos[l];
    const double cnt = sums[(size_t)l * 5 + 0];
    if (f == ~0ull || cnt <= 0.0) {
        lf.key0 = INT64_MAX; lf.base = 0; lf.a = 0.f; lf.c = 0.f; lf.elo = 1; lf.ehi = 0;   // empty
    } else {
        const double sd = sums[l * 5 + 1], sr = sums[l * 5 + 2], sdd = sums[l * 5 + 3], sdr = sums[l * 5 + 4];
        const double md = sd / cnt, mr = sr / cnt;
        const double var = sdd / cnt - md * md, cov = sdr / cnt - md * mr;
        const double a = (var > 0.0) ? cov / var : 0.0;
        lf.key0 = (int64_t)starts[f];
        lf.base = (int64_t)f;
        lf.a = (float)a;
        lf.c = (float)(mr - a * md);
        lf.elo = INT32_MAX;
        lf.ehi = INT32_MIN;
    }
    leaves[l] = lf;
}

// Pass 3: exact error bounds.  Only the last key of a run of equal starts is a distinct
// upper_bound value (ub(key) = that position).
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_leaf_err(const S* __restrict__ starts, uint32_t n,
                                                          const uint32_t* __restrict__ leaf_of_key,
                                                          Leaf* __restrict__ leaves) {
    const uint32_t i = blockIdx.x * RMI_THREADS + threadIdx.x;
    if (i >= n) return;
    if (i + 1 < n && starts[i + 1] == starts[i]) return;
    const uint32_t l = leaf_of_key[i];
    const Leaf lf = leaves[l];
    int64_t e = (int64_t)i - leaf_pred(lf, (int64_t)starts[i]);
    e = std::max<int64_t>(std::min<int64_t>(e, INT32_MAX / 2), INT32_MIN / 2);
    atomicMin(&leaves[l].elo, (int32_t)e);
    atomicMax(&leaves[l].ehi, (int32_t)e);
}

// Learned upper_bound.  stats (optional) counts fallbacks and sums window widths.
template <typename S>
__global__ __launch_bounds__(RMI_THREADS) void k_ub_learned(IndexView<S> ix, int64_t kmin, float inv_span,
                                                            const float* __restrict__ P,
                                                            const Leaf* __restrict__ leaves, uint32_t L,
                                                            const S* __restrict__ xs, size_t nx,
                                                            uint32_t* __restrict__ out,
                                                            unsigned long long* __restrict__ stats) {
    __shared__ float s_p[NPARAM];
    for (int k = threadIdx.x; k < NPARAM; k += RMI_THREADS) s_p[k] = P[k];
    __syncthreads();
    const size_t q = (size_t)blockIdx.x * RMI_THREADS + threadIdx.x;
    const bool valid = q < nx;
    const S x = valid ? xs[q] : (S)0;
    const uint32_t n = ix.n;
    uint32_t res = NONE;
    bool ok = false;
    uint32_t width = 0;
    if (!valid || n == 0) {
        ok = true;
    } else {
        const Leaf lf = leaves[leaf_of(root_forward(s_p, norm_key(x, kmin, inv_span)), L)];
        if (lf.elo <= lf.ehi) {
            const int64_t p = leaf_pred(lf, (int64_t)x);
            int64_t lo = p + lf.elo - 1, hi = p + lf.ehi;
            lo = std::max<int64_t>(lo, 0);
            hi = std::min<int64_t>(hi, (int64_t)n - 1);
            if (lo <= hi) {
                width = (uint32_t)(hi - lo + 1);
                // last r in [lo, hi] with starts[r] <= x, or lo - 1
                const S* base = ix.starts + lo;
                uint32_t len = (uint32_t)(hi - lo + 1);
                while (len > 1) {
                    const uint32_t half = len >> 1;
                    base = (base[half] <= x) ? base + half : base;
                    len -= half;
                }
                const int64_t r = (int64_t)(base - ix.starts) - ((*base <= x) ? 0 : 1);
                const bool top_ok = (r < hi) || (hi == (int64_t)n - 1) || (ix.starts[hi + 1] > x);
                const bool bot_ok = (r >= lo) || (lo == 0) || (ix.starts[lo - 1] <= x);
                if (top_ok && bot_ok) {
                    ok = true;
                    res = (r < 0) ? NONE : (uint32_t)r;
                }
            }
        }
    }
    if (!ok) res = upper_bound_pos(ix.starts, n, x);
    if (valid) out[q] = res;
    if (stats) {
        unsigned long long f = ok ? 0ull : 1ull, w = width;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            f += __shfl_xor(f, d, WAVE);
            w += __shfl_xor(w, d, WAVE);
        }
        if (lane_id() == 0) {
            atomicAdd(&stats[0], f);
            atomicAdd(&stats[1], w);
        }
    }
}

}  // namespace

// ============================================================================
// The model attached to an index
// ============================================================================
struct siLearned {
    uint32_t L = 4096;
    int64_t kmin = 0;
    float inv_span = 1.f;
    float* d_params = nullptr;   // NPARAM
    Leaf* d_leaves = nullptr;    // L
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
    uint32_t* d_leaf = nullptr;
    unsigned long long* d_first = nullptr;
    double* d_sums = nullptr;
    HIP_TRY(hipMalloc(&d_leaf, (size_t)n * 4));
    HIP_TRY(hipMalloc(&d_first, (size_t)L * 8));
    HIP_TRY(hipMalloc(&d_sums, (size_t)L * 5 * 8));
    HIP_TRY(hipMemsetAsync(d_first, 0xFF, (size_t)L * 8, st));
    HIP_TRY(hipMemsetAsync(d_sums, 0, (size_t)L * 5 * 8, st));
    const unsigned g = grid_for(n, RMI_THREADS);
    k_leaf_first<S><<<g, RMI_THREADS, 0, st>>>(v.starts, n, m->kmin, m->inv_span, m->d_params, L, d_leaf, d_first);
    LAUNCH_CHECK();
    k_leaf_sums<S><<<g, RMI_THREADS, 0, st>>>(v.starts, n, d_leaf, d_first, d_sums);
    LAUNCH_CHECK();
    k_leaf_fit<S><<<grid_for(L, 256), 256, 0, st>>>(v.starts, L, d_first, d_sums, m->d_leaves);
    LAUNCH_CHECK();
    k_leaf_err<S><<<g, RMI_THREADS, 0, st>>>(v.starts, n, d_leaf, m->d_leaves);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<Leaf> h(L);
    HIP_TRY(hipMemcpy(h.data(), m->d_leaves, (size_t)L * sizeof(Leaf), hipMemcpyDeviceToHost));
    int32_t mx = 0;
    for (const Leaf& lf : h)
        if (lf.elo <= lf.ehi) mx = std::max(mx, std::max(std::abs(lf.elo), std::abs(lf.ehi)));
    m->max_abs_err = mx;
    hipFree(d_leaf);
    hipFree(d_first);
    hipFree(d_sums);
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
    if (leaves < 1 || leaves > (1u << 24)) { si_set_error(SI_ERR_ARG, "leaves must be in [1, 2^24]"); return nullptr; }
    siLearned* m = new siLearned();
    m->L = leaves;
    m->ix = ix;
    m->ix_n = ix->n;
    if (hipMalloc(&m->d_params, NPARAM * 4) != hipSuccess ||
        hipMalloc(&m->d_leaves, (size_t)leaves * sizeof(Leaf)) != hipSuccess) {
        si_set_error(SI_ERR_NOMEM, "learned index allocation");
        delete m;
        return nullptr;
    }
    return m;
}

SI_EXPORT void siLearnedDestroy(siLearned* m) { delete m; }

SI_EXPORT int siLearnedParamCount(void) { return NPARAM; }

SI_EXPORT int siLearnedTrain(siLearned* m, int steps, uint32_t batch, float lr, void* stream) {
    if (!m) return si_set_error(SI_ERR_ARG, "null model");
    if (steps < 0 || batch == 0) return si_set_error(SI_ERR_ARG, "bad training arguments");
    const siIndex* ix = m->ix;
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->n != m->ix_n) return si_set_error(SI_ERR_ARG, "the index was rebuilt; create a new model");
    hipStream_t st = (hipStream_t)stream;
    rc = (ix->cb == 4) ? train_impl<int32_t>(m, ix, steps, batch, lr, st)
                       : train_impl<int64_t>(m, ix, steps, batch, lr, st);
    if (rc == SI_OK) m->trained = true;
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

// upper_bound through the model (device buffers).  stats (device, may be NULL) receives
// [0] += fallbacks to the plain search, [1] += sum of window widths.
SI_EXPORT int siUpperBoundLearned(const siLearned* m, const void* xs, size_t n, uint32_t* out,
                                  uint64_t* stats, void* stream) {
    if (!m || !m->trained) return si_set_error(SI_ERR_ARG, "model not trained");
    const siIndex* ix = m->ix;
    int rc = check_ix(ix);
    if (rc) return rc;
    if (ix->n != m->ix_n) return si_set_error(SI_ERR_ARG, "the index was rebuilt; retrain the model");
    if (n == 0) return SI_OK;
    if (!xs || !out) return si_set_error(SI_ERR_ARG, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(n, RMI_THREADS);
    if (ix->cb == 4)
        k_ub_learned<int32_t><<<g, RMI_THREADS, 0, st>>>(ix->view<int32_t>(), m->kmin, m->inv_span, m->d_params,
                                                         m->d_leaves, m->L, (const int32_t*)xs, n, out,
                                                         (unsigned long long*)stats);
    else
        k_ub_learned<int64_t><<<g, RMI_THREADS, 0, st>>>(ix->view<int64_t>(), m->kmin, m->inv_span, m->d_params,
                                                         m->d_leaves, m->L, (const int64_t*)xs, n, out,
                                                         (unsigned long long*)stats);
    LAUNCH_CHECK();
    return SI_OK;
}
