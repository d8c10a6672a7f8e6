// workload.cpp -- deterministic synthetic interval / query generators.
//
// Every draw is a pure function of (seed, stream, index): a SplitMix64
// finaliser over a Weyl sequence.  Generation therefore parallelises over
// threads and reproduces bit-for-bit for a given seed.  Continuous draws
// (log-normal, geometric, Gaussian hotspots) use double arithmetic from libm
// on top of the integer stream.
//
// Workloads (all end-inclusive: end = start + L - 1):
//   1  N intervals, int32, start ~ U[0, 1e7), L ~ Geometric(p = 1/500) on {1,2,..}
//      capped at 1e4.  Queries: same distribution from their own seed.
//   2  start ~ U[0, 2.5e8), L = clamp(round(LogNormal(6, 1.5)), 1, 1e6).
//      Queries: start ~ U[0, 2.5e8), L = max(1, round(LogNormal(7, 1))).
//   3  int64 on [0, 3e9): start 70 % uniform, 30 % from 200 Gaussian hotspots
//      (sd 5e4); L: 20 % log-uniform [1e5, 1e7], else max(1, round(LogNormal(5, 1))).
//      Queries: start from the same mixture; 1 % wide, L log-uniform [1e6, 1e7],
//      else L = max(1, round(LogNormal(5, 1))).
//   4  int32 on [0, 2e9): 10 segments of 2e8 with densities 100^(p_k / 9) for a
//      fixed permutation p of 0..9; L = max(1, round(LogNormal(6, 1.2))).
//      Queries: half point queries (L = 1), half L = max(1, round(LogNormal(4, 1))).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "si_common.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed * 0xD1B54A32D192ED03ULL + stream * 0x8CB92BA72F3D8DD7ULL + 1);
}
inline uint64_t draw(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * 0x9E3779B97F4A7C15ULL); }
inline double u01(uint64_t r) { return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
inline uint64_t below(uint64_t r, uint64_t m) { return (uint64_t)(((unsigned __int128)r * m) >> 64); }
inline double normal(uint64_t key, uint64_t i) {
    double a = u01(draw(key, 2 * i)), b = u01(draw(key, 2 * i + 1));
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}
inline int64_t lognormal_len(uint64_t key, uint64_t i, double mu, double sigma) {
    return (int64_t)std::llround(std::exp(mu + sigma * normal(key, i)));
}
inline int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

enum : uint64_t { ST_START = 1, ST_LEN = 2, ST_MIX = 3, ST_HOT = 4, ST_KIND = 5, ST_SEG = 6 };

const int kSegPerm[10] = {3, 7, 0, 9, 5, 1, 8, 2, 6, 4};

struct Gen {
    int cfg;
    uint64_t seed;
    bool query;
    double seg_cdf[10];
    int64_t hot[200];

    Gen(int c, uint64_t s, bool q) : cfg(c), seed(s), query(q) {
        double tot = 0;
        for (int k = 0; k < 10; ++k) tot += std::pow(100.0, kSegPerm[k] / 9.0);
        double acc = 0;
        for (int k = 0; k < 10; ++k) { acc += std::pow(100.0, kSegPerm[k] / 9.0) / tot; seg_cdf[k] = acc; }
        seg_cdf[9] = 1.0;
        // Hotspot centres depend on the interval seed's family only, so the
        // queries of workload 3 aim at the same hotspots as the intervals.
        uint64_t hk = stream_key(0x5EED0F407507ULL, ST_HOT);
        for (int k = 0; k < 200; ++k) hot[k] = (int64_t)below(draw(hk, k), 3000000000ULL);
    }

    int64_t start(uint64_t i) const {
        uint64_t ks = stream_key(seed, ST_START);
        switch (cfg) {
            case 1: return (int64_t)below(draw(ks, i), 10000000ULL);
            case 2: return (int64_t)below(draw(ks, i), 250000000ULL);
            case 3: {
                uint64_t km = stream_key(seed, ST_MIX);
                uint64_t r = draw(km, i);
                if (u01(r) < 0.7) return (int64_t)below(draw(ks, i), 3000000000ULL);
                int h = (int)below(draw(km, i + 0x100000000ULL), 200);
                double x = (double)hot[h] + 5e4 * normal(ks, i);
                return clamp64((int64_t)std::floor(x), 0, 2999999999LL);
            }
            case 4: {
                double u = u01(draw(stream_key(seed, ST_SEG), i));
                int k = 0;
                while (k < 9 && u > seg_cdf[k]) ++k;
                return (int64_t)k * 200000000LL + (int64_t)below(draw(ks, i), 200000000ULL);
            }
            default: return 0;
        }
    }

    int64_t length(uint64_t i) const {
        uint64_t kl = stream_key(seed, ST_LEN);
        switch (cfg) {
            case 1: {
                // Geometric(p) on {1,2,...} by inversion: 1 + floor(ln U / ln(1-p)).
                double u = u01(draw(kl, i));
                int64_t L = 1 + (int64_t)std::floor(std::log(u) / std::log(1.0 - 1.0 / 500.0));
                return clamp64(L, 1, 10000);
            }
            case 2:
                if (query) return std::max<int64_t>(1, lognormal_len(kl, i, 7.0, 1.0));
                return clamp64(lognormal_len(kl, i, 6.0, 1.5), 1, 1000000);
            case 3: {
                uint64_t kk = stream_key(seed, ST_KIND);
                double u = u01(draw(kk, i));
                if (query) {
                    if (u < 0.01) {
                        double v = u01(draw(kk, i + 0x100000000ULL));
                        return (int64_t)std::llround(std::exp(std::log(1e6) + v * (std::log(1e7) - std::log(1e6))));
                    }
                    return std::max<int64_t>(1, lognormal_len(kl, i, 5.0, 1.0));
                }
                if (u < 0.2) {
                    double v = u01(draw(kk, i + 0x100000000ULL));
                    return (int64_t)std::llround(std::exp(std::log(1e5) + v * (std::log(1e7) - std::log(1e5))));
                }
                return std::max<int64_t>(1, lognormal_len(kl, i, 5.0, 1.0));
            }
            case 4:
                if (query) {
                    if (draw(stream_key(seed, ST_KIND), i) & 1) return 1;
                    return std::max<int64_t>(1, lognormal_len(kl, i, 4.0, 1.0));
                }
                return std::max<int64_t>(1, lognormal_len(kl, i, 6.0, 1.2));
            default: return 1;
        }
    }
};

template <typename F>
void parallel_for(size_t n, F f) {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 16) t = 16;
    if (n < 65536) t = 1;
    std::vector<std::thread> ws;
    for (unsigned w = 0; w < t; ++w)
        ws.emplace_back([&, w] {
            size_t a = n * w / t, b = n * (w + 1) / t;
            for (size_t i = a; i < b; ++i) f(i);
        });
    for (auto& x : ws) x.join();
}

template <typename S>
int gen(int cfg, uint64_t seed, size_t n, int sorted, bool query, S* st, S* en) {
    Gen g(cfg, seed, query);
    std::vector<int64_t> s64(n);
    parallel_for(n, [&](size_t i) { s64[i] = g.start(i); });
    if (sorted) std::sort(s64.begin(), s64.end());
    parallel_for(n, [&](size_t i) {
        int64_t L = g.length(i);
        int64_t s = s64[i], e = s + L - 1;
        if (sizeof(S) == 4) e = std::min<int64_t>(e, INT32_MAX);
        st[i] = (S)s;
        en[i] = (S)e;
    });
    return SI_OK;
}

int gen_any(int cfg, uint64_t seed, size_t n, int coord_bytes, int sorted, bool query, void* st, void* en) {
    if (cfg < 1 || cfg > 4) return si_set_error(SI_ERR_ARG, "workload: unknown config %d", cfg);
    if (n && (!st || !en)) return si_set_error(SI_ERR_ARG, "workload: null output");
    if (coord_bytes == 4) {
        if (cfg == 3) return si_set_error(SI_ERR_ARG, "workload 3 needs int64 coordinates");
        return gen<int32_t>(cfg, seed, n, sorted, query, (int32_t*)st, (int32_t*)en);
    }
    if (coord_bytes == 8) return gen<int64_t>(cfg, seed, n, sorted, query, (int64_t*)st, (int64_t*)en);
    return si_set_error(SI_ERR_ARG, "workload: coord_bytes must be 4 or 8");
}

}  // namespace

SI_EXPORT int siWorkloadIntervals(int config, uint64_t seed, size_t n, int coord_bytes, int sorted,
                                  void* starts, void* ends) {
    return gen_any(config, seed, n, coord_bytes, sorted, false, starts, ends);
}

SI_EXPORT int siWorkloadQueries(int config, uint64_t seed, size_t n, int coord_bytes, void* qs,
                                void* qe) {
    return gen_any(config, seed, n, coord_bytes, 0, true, qs, qe);
}
