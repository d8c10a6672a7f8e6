/*
 * si_oracle.c -- CPU restatement of the SuperIntervals hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is never linked into, called by, or shipped with the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows (kcleal/superintervals v1.0.0):
 *   src/superintervals.hpp  (C++ template, "hpp")
 *   src/c_superintervals.h  (C implementation / C-ABI, "c.h")
 *
 * Pinning (what independently checks each function, see tests/test_oracle.py):
 *   or_sort_*          brute-force comparison sort + golden arrays from the compiled reference
 *   or_branch_*        O(n^2) brute-force previous-greater-or-equal on small inputs + golden hash
 *   or_upper_bound_*   numpy.searchsorted(side="right") - 1
 *   or_count_*         brute force #{j : starts[j] <= e && ends[j] >= s}, golden counts,
 *                      and rank difference for s <= e
 *   or_search_*        closed form [j for j = idx..0 if ends[j] >= s] (strict mode),
 *                      closed form of the vector order, golden per-query hashes
 *   or_values_*        data[] gather of the strict list
 *   or_coverage_i32    brute-force sum of min(e,ends)-max(s,starts) (c.h:758-792)
 *   or_walk_stats_*    exact visit/jump counts; checked against an independent
 *                      Python re-walk on small inputs and against h = count
 *
 * Coordinates are either int32 ("i32") or int64 ("i64"); indices are int64
 * with -1 meaning "none" (SIZE_MAX / SI_NONE in the reference).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define OR_NONE ((int64_t)-1)

/* ------------------------------------------------------------------------ */
/* Order-sensitive 64-bit list hash used by the golden fixtures and tests.   */
/* h = fold over elements of mix(h ^ (x + 0x9E37..)), seeded with the length */
/* ------------------------------------------------------------------------ */
static inline uint64_t or_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t or_hash_list(const int64_t* v, size_t n) {
    uint64_t h = or_mix64((uint64_t)n + 0x51ED270B27B1ULL);
    for (size_t k = 0; k < n; ++k) h = or_mix64(h ^ ((uint64_t)v[k] + 0x9E3779B97F4A7C15ULL));
    return h;
}

uint64_t or_hash_list_u32(const uint32_t* v, size_t n) {
    uint64_t h = or_mix64((uint64_t)n + 0x51ED270B27B1ULL);
    for (size_t k = 0; k < n; ++k) h = or_mix64(h ^ ((uint64_t)v[k] + 0x9E3779B97F4A7C15ULL));
    return h;
}

/* Per-query hashes of a CSR result (offsets[nq+1], idx as uint32 or int64). */
void or_hash_csr_u32(size_t nq, const uint64_t* off, const uint32_t* idx, uint64_t* out) {
    for (size_t q = 0; q < nq; ++q) out[q] = or_hash_list_u32(idx + off[q], (size_t)(off[q + 1] - off[q]));
}
void or_hash_csr_i64(size_t nq, const uint64_t* off, const int64_t* idx, uint64_t* out) {
    for (size_t q = 0; q < nq; ++q) out[q] = or_hash_list(idx + off[q], (size_t)(off[q + 1] - off[q]));
}
void or_hash_csr_i32(size_t nq, const uint64_t* off, const int32_t* v, uint64_t* out) {
    for (size_t q = 0; q < nq; ++q) {
        uint64_t h = or_mix64((uint64_t)(off[q + 1] - off[q]) + 0x51ED270B27B1ULL);
        for (uint64_t k = off[q]; k < off[q + 1]; ++k)
            h = or_mix64(h ^ ((uint64_t)(int64_t)v[k] + 0x9E3779B97F4A7C15ULL));
        out[q] = h;
    }
}

static int or_nthreads_clamp(int t) { return t < 1 ? 1 : (t > 256 ? 256 : t); }

#define OR_DEFINE(TY, SFX)                                                          \
                                                                                    \
typedef struct { TY start, end; int32_t data; } or_item_##SFX;                      \
                                                                                    \
/* Comparators: c.h:425-438 (start asc, end desc; end desc for tie blocks). */     \
static inline int or_less_start_##SFX(const or_item_##SFX* a, const or_item_##SFX* b) { \
    if (a->start != b->start) return a->start < b->start;                          \
    return a->end > b->end;                                                         \
}                                                                                   \
static inline int or_less_end_##SFX(const or_item_##SFX* a, const or_item_##SFX* b) {   \
    return a->end > b->end;                                                         \
}                                                                                   \
                                                                                    \
/* Stable bottom-up merge sort of an AoS block.  The C reference sorts with    */  \
/* qsort (c.h:451), which glibc 2.35 implements as a stable merge sort; stable */  \
/* order is therefore the C reference's order.  (The C++ template's std::sort, */  \
/* hpp:1403, is unstable: only the payload order among exact (start,end)       */  \
/* duplicates can differ; sorted starts/ends are identical.)                   */  \
static void or_msort_##SFX(or_item_##SFX* a, size_t n, int by_start) {             \
    if (n < 2) return;                                                              \
    or_item_##SFX* tmp = (or_item_##SFX*)malloc(n * sizeof(or_item_##SFX));         \
    or_item_##SFX* src = a; or_item_##SFX* dst = tmp;                               \
    for (size_t w = 1; w < n; w *= 2) {                                             \
        for (size_t lo = 0; lo < n; lo += 2 * w) {                                  \
            size_t mid = lo + w < n ? lo + w : n;                                   \
            size_t hi = lo + 2 * w < n ? lo + 2 * w : n;                            \
            size_t i = lo, j = mid, k = lo;                                         \
            while (i < mid && j < hi) {                                             \
                int take_right = by_start ? or_less_start_##SFX(&src[j], &src[i])   \
                                          : or_less_end_##SFX(&src[j], &src[i]);    \
                dst[k++] = take_right ? src[j++] : src[i++];                        \
            }                                                                       \
            while (i < mid) dst[k++] = src[i++];                                    \
            while (j < hi) dst[k++] = src[j++];                                     \
        }                                                                           \
        or_item_##SFX* t = src; src = dst; dst = t;                                 \
    }                                                                               \
    if (src != a) memcpy(a, src, n * sizeof(or_item_##SFX));                        \
    free(tmp);                                                                      \
}                                                                                   \
                                                                                    \
/* sortBlock, c.h:440-460 / hpp:1396-1409: AoS copy, sort, scatter back. */        \
static void or_sort_block_##SFX(TY* starts, TY* ends, int32_t* data,               \
                                size_t lo, size_t hi, int by_start) {               \
    size_t m = hi - lo;                                                             \
    or_item_##SFX* t = (or_item_##SFX*)malloc((m ? m : 1) * sizeof(or_item_##SFX)); \
    for (size_t i = 0; i < m; ++i) {                                                \
        t[i].start = starts[lo + i]; t[i].end = ends[lo + i]; t[i].data = data[lo + i]; \
    }                                                                               \
    or_msort_##SFX(t, m, by_start);                                                 \
    for (size_t i = 0; i < m; ++i) {                                                \
        starts[lo + i] = t[i].start; ends[lo + i] = t[i].end; data[lo + i] = t[i].data; \
    }                                                                               \
    free(t);                                                                        \
}                                                                                   \
                                                                                    \
/* add() sortedness flags, hpp:95-105 / c.h:408-413.  Returns bit0=start_sorted, */ \
/* bit1=end_sorted after appending the arrays in order.                        */  \
int or_flags_##SFX(size_t n, const TY* starts, const TY* ends) {                   \
    int ss = 1, es = 1;                                                             \
    for (size_t i = 1; i < n; ++i) {                                                \
        if (ss) {                                                                   \
            ss = (starts[i] < starts[i - 1]) ? 0 : 1;                               \
            if (ss && starts[i] == starts[i - 1] && ends[i] > ends[i - 1]) es = 0;  \
        }                                                                           \
    }                                                                               \
    return ss | (es << 1);                                                          \
}                                                                                   \
                                                                                    \
/* sort_intervals, hpp:1415-1439 / sortIntervals c.h:462-485. */                   \
void or_sort_##SFX(size_t n, TY* starts, TY* ends, int32_t* data,                  \
                   int start_sorted, int end_sorted) {                              \
    if (!start_sorted) {                                                            \
        or_sort_block_##SFX(starts, ends, data, 0, n, 1);                           \
    } else if (!end_sorted) {                                                       \
        size_t it = 0;                                                              \
        while (it < n) {                                                            \
            size_t be = it + 1; int need = 0;                                       \
            while (be < n && starts[be] == starts[it]) {                            \
                if (ends[be] > ends[be - 1]) need = 1;                              \
                ++be;                                                               \
            }                                                                       \
            if (need) or_sort_block_##SFX(starts, ends, data, it, be, 0);           \
            it = be;                                                                \
        }                                                                           \
    }                                                                               \
}                                                                                   \
                                                                                    \
/* Branch array, hpp:112-130 / indexSuperIntervals c.h:487-521: monotonic stack; */ \
/* the stack pops while top.end < ends[i], so branch[i] is the nearest j < i    */  \
/* with ends[j] >= ends[i], or none.                                           */  \
void or_branch_##SFX(size_t n, const TY* ends, int64_t* branch) {                  \
    if (n == 0) return;                                                             \
    int64_t* st = (int64_t*)malloc(n * sizeof(int64_t));                            \
    size_t top = 0;                                                                 \
    for (size_t i = 0; i < n; ++i) branch[i] = OR_NONE;                             \
    st[top++] = 0;                                                                  \
    for (size_t i = 1; i < n; ++i) {                                                \
        while (top > 0 && ends[st[top - 1]] < ends[i]) --top;                       \
        if (top > 0) branch[i] = st[top - 1];                                       \
        st[top++] = (int64_t)i;                                                     \
    }                                                                               \
    free(st);                                                                       \
}                                                                                   \
                                                                                    \
/* upper_bound, hpp:501-516: branchless halving; last i with starts[i] <= x,   */  \
/* else none.                                                                  */  \
int64_t or_upper_bound_##SFX(size_t n, const TY* starts, TY x) {                   \
    if (n == 0) return OR_NONE;                                                     \
    size_t length = n, idx = 0;                                                     \
    while (length > 1) {                                                            \
        size_t half = length / 2;                                                   \
        idx += (starts[idx + half] <= x) * (length - half);                         \
        length = half;                                                              \
    }                                                                               \
    if (starts[idx] > x) return idx == 0 ? OR_NONE : (int64_t)idx - 1;              \
    return (int64_t)idx;                                                            \
}                                                                                   \
                                                                                    \
/* count, hpp:651-825, restated with the scalar 16-wide block scan (hpp:688-700)*/ \
/* (the SI_NOSIMD form; the AVX2/NEON forms count the same predicate over 32). */  \
uint64_t or_count_##SFX(size_t n, const TY* starts, const TY* ends,                \
                        const int64_t* branch, TY s, TY e) {                        \
    if (n == 0) return 0;                                                           \
    int64_t ii = or_upper_bound_##SFX(n, starts, e);                                \
    if (ii == OR_NONE) return 0;                                                    \
    size_t i = (size_t)ii; uint64_t found = 0;                                      \
    const size_t block = 16;                                                        \
    while (i > 0) {                                                                 \
        if (s <= ends[i]) {                                                         \
            ++found; --i;                                                           \
            while (i > block) {                                                     \
                size_t c = 0;                                                       \
                for (size_t j = i; j > i - block; --j) c += (s <= ends[j]) ? 1 : 0; \
                found += c; i -= block;                                             \
                if (c < block && s > ends[i + 1]) break;                            \
            }                                                                       \
        } else {                                                                    \
            if (branch[i] == OR_NONE) return found;                                 \
            i = (size_t)branch[i];                                                  \
        }                                                                           \
    }                                                                               \
    if (i == 0 && s <= ends[0] && starts[0] <= e) ++found;                          \
    return found;                                                                   \
}                                                                                   \
                                                                                    \
/* Overlap indices.                                                            */  \
/* mode 0 = strict descending: searchIdxs c.h:610-643, IndexIterator hpp:153-235. */\
/* mode 1 = C++ vector form, hpp:879-905: leading run [i*+1 .. idx] ascending   */ \
/*          (CountingIterator insert, hpp:891-895), then the walk descending.  */  \
/* out may be NULL (length only).  Returns the number of hits.                 */  \
size_t or_search_##SFX(size_t n, const TY* starts, const TY* ends,                 \
                       const int64_t* branch, TY s, TY e, int mode, int64_t* out) { \
    if (n == 0) return 0;                                                           \
    int64_t idx = or_upper_bound_##SFX(n, starts, e);                               \
    if (idx == OR_NONE) return 0;                                                   \
    size_t h = 0; int64_t i = idx;                                                  \
    while (i != OR_NONE && s <= ends[i]) --i;                                       \
    if (mode == 1) {                                                                \
        for (int64_t j = i + 1; j <= idx; ++j) { if (out) out[h] = j; ++h; }        \
    } else {                                                                        \
        for (int64_t j = idx; j > i; --j) { if (out) out[h] = j; ++h; }             \
    }                                                                               \
    if (i == OR_NONE) return h;                                                     \
    i = branch[i];                                                                  \
    while (i != OR_NONE) {                                                          \
        if (s <= ends[i]) { if (out) out[h] = i; ++h; --i; }                        \
        else i = branch[i];                                                         \
    }                                                                               \
    return h;                                                                       \
}                                                                                   \
                                                                                    \
/* search_values, hpp:551-579 / searchValues c.h:575-608: data[] of the strict */  \
/* descending list.                                                            */  \
size_t or_values_##SFX(size_t n, const TY* starts, const TY* ends,                 \
                       const int64_t* branch, const int32_t* data, TY s, TY e,      \
                       int32_t* out) {                                              \
    if (n == 0) return 0;                                                           \
    int64_t idx = or_upper_bound_##SFX(n, starts, e);                               \
    if (idx == OR_NONE) return 0;                                                   \
    size_t h = 0; int64_t i = idx;                                                  \
    while (i != OR_NONE && s <= ends[i]) { if (out) out[h] = data[i]; ++h; --i; }   \
    if (i == OR_NONE) return h;                                                     \
    i = branch[i];                                                                  \
    while (i != OR_NONE) {                                                          \
        if (s <= ends[i]) { if (out) out[h] = data[i]; ++h; --i; }                  \
        else i = branch[i];                                                         \
    }                                                                               \
    return h;                                                                       \
}                                                                                   \
                                                                                    \
/* has_overlaps, hpp:865-871 / anyOverlaps c.h:570-573. */                        \
int or_has_overlaps_##SFX(size_t n, const TY* starts, const TY* ends, TY s, TY e) { \
    if (n == 0) return 0;                                                           \
    int64_t idx = or_upper_bound_##SFX(n, starts, e);                               \
    return idx != OR_NONE && s <= ends[idx];                                        \
}                                                                                   \
                                                                                    \
/* Instrumented walk of the C/iterator form (c.h:610-643): h hits, V reads of  */  \
/* ends[] in the walk, J reads of branch[].  These are the per-query counts in */  \
/* the algorithmic byte model.                                                 */  \
void or_walk_stats_##SFX(size_t n, const TY* starts, const TY* ends,               \
                         const int64_t* branch, TY s, TY e,                         \
                         uint64_t* h_out, uint64_t* v_out, uint64_t* j_out) {       \
    uint64_t h = 0, v = 0, jn = 0;                                                  \
    int64_t i = (n == 0) ? OR_NONE : or_upper_bound_##SFX(n, starts, e);            \
    if (i != OR_NONE) {                                                             \
        while (i != OR_NONE) { ++v; if (s <= ends[i]) { ++h; --i; } else break; }   \
        if (i != OR_NONE) {                                                         \
            i = branch[i]; ++jn;                                                    \
            while (i != OR_NONE) {                                                  \
                ++v;                                                                \
                if (s <= ends[i]) { ++h; --i; } else { i = branch[i]; ++jn; }       \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    *h_out = h; *v_out = v; *j_out = jn;                                            \
}                                                                                   \
                                                                                    \
/* ---- Multi-threaded batch forms (static contiguous query chunks; the built */   \
/* index is shared read-only).  Used as the ported CPU baseline and to produce */  \
/* per-query results for the parity tests.                                    */   \
typedef struct {                                                                    \
    size_t n; const TY* starts; const TY* ends; const int64_t* branch;              \
    const TY* qs; const TY* qe; size_t q0, q1; int mode;                            \
    uint32_t* counts; const uint64_t* offsets; int64_t* out; uint64_t* hashes;      \
    uint64_t* stats; uint64_t total;                                                \
} or_job_##SFX;                                                                     \
                                                                                    \
static void* or_count_worker_##SFX(void* p) {                                       \
    or_job_##SFX* j = (or_job_##SFX*)p; uint64_t t = 0;                             \
    for (size_t q = j->q0; q < j->q1; ++q) {                                        \
        uint64_t c = or_count_##SFX(j->n, j->starts, j->ends, j->branch, j->qs[q], j->qe[q]); \
        if (j->counts) j->counts[q] = (uint32_t)c;                                  \
        t += c;                                                                     \
    }                                                                               \
    j->total = t; return NULL;                                                      \
}                                                                                   \
                                                                                    \
static void* or_search_worker_##SFX(void* p) {                                      \
    or_job_##SFX* j = (or_job_##SFX*)p; uint64_t t = 0;                             \
    int64_t* scratch = NULL; size_t cap = 0;                                        \
    for (size_t q = j->q0; q < j->q1; ++q) {                                        \
        int64_t* dst;                                                               \
        if (j->out) {                                                               \
            dst = j->out + j->offsets[q];                                           \
        } else {                                                                    \
            size_t need = or_search_##SFX(j->n, j->starts, j->ends, j->branch,      \
                                          j->qs[q], j->qe[q], j->mode, NULL);       \
            if (need > cap) { cap = need * 2 + 16;                                  \
                scratch = (int64_t*)realloc(scratch, cap * sizeof(int64_t)); }      \
            dst = scratch;                                                          \
        }                                                                           \
        size_t h = or_search_##SFX(j->n, j->starts, j->ends, j->branch,             \
                                   j->qs[q], j->qe[q], j->mode, dst);               \
        if (j->counts) j->counts[q] = (uint32_t)h;                                  \
        if (j->hashes) j->hashes[q] = or_hash_list(dst, h);                         \
        t += h;                                                                     \
    }                                                                               \
    free(scratch); j->total = t; return NULL;                                       \
}                                                                                   \
                                                                                    \
static void* or_stats_worker_##SFX(void* p) {                                       \
    or_job_##SFX* j = (or_job_##SFX*)p; uint64_t H = 0, V = 0, J = 0;               \
    for (size_t q = j->q0; q < j->q1; ++q) {                                        \
        uint64_t h, v, jn;                                                          \
        or_walk_stats_##SFX(j->n, j->starts, j->ends, j->branch, j->qs[q], j->qe[q], &h, &v, &jn); \
        H += h; V += v; J += jn;                                                    \
    }                                                                               \
    j->stats[0] = H; j->stats[1] = V; j->stats[2] = J; return NULL;                 \
}                                                                                   \
                                                                                    \
static uint64_t or_run_##SFX(int nthreads, void* (*fn)(void*), or_job_##SFX* base, size_t nq) { \
    nthreads = or_nthreads_clamp(nthreads);                                         \
    pthread_t th[256]; or_job_##SFX jobs[256]; uint64_t st[256][3];                 \
    uint64_t tot = 0;                                                               \
    for (int t = 0; t < nthreads; ++t) {                                            \
        jobs[t] = *base;                                                            \
        jobs[t].q0 = nq * (size_t)t / (size_t)nthreads;                             \
        jobs[t].q1 = nq * (size_t)(t + 1) / (size_t)nthreads;                       \
        jobs[t].stats = st[t];                                                      \
        pthread_create(&th[t], NULL, fn, &jobs[t]);                                 \
    }                                                                               \
    for (int t = 0; t < nthreads; ++t) { pthread_join(th[t], NULL); tot += jobs[t].total; } \
    if (base->stats) {                                                              \
        base->stats[0] = base->stats[1] = base->stats[2] = 0;                       \
        for (int t = 0; t < nthreads; ++t) {                                        \
            base->stats[0] += st[t][0]; base->stats[1] += st[t][1]; base->stats[2] += st[t][2]; \
        }                                                                           \
    }                                                                               \
    return tot;                                                                     \
}                                                                                   \
                                                                                    \
/* counts[q] for every query (counts may be NULL); returns the total. */          \
uint64_t or_count_batch_##SFX(int nthreads, size_t n, const TY* starts, const TY* ends, \
                              const int64_t* branch, size_t nq, const TY* qs, const TY* qe, \
                              uint32_t* counts) {                                   \
    or_job_##SFX b; memset(&b, 0, sizeof(b));                                       \
    b.n = n; b.starts = starts; b.ends = ends; b.branch = branch; b.qs = qs; b.qe = qe; \
    b.counts = counts;                                                              \
    return or_run_##SFX(nthreads, or_count_worker_##SFX, &b, nq);                   \
}                                                                                   \
                                                                                    \
/* Per-query lists.  With offsets/out given, writes each list at out+offsets[q]; */ \
/* otherwise uses a per-thread scratch (lengths, hashes only).  counts and      */ \
/* hashes may be NULL.  Returns the total number of hits.                      */  \
uint64_t or_search_batch_##SFX(int nthreads, size_t n, const TY* starts, const TY* ends, \
                               const int64_t* branch, size_t nq, const TY* qs, const TY* qe, \
                               int mode, uint32_t* counts, const uint64_t* offsets, \
                               int64_t* out, uint64_t* hashes) {                    \
    or_job_##SFX b; memset(&b, 0, sizeof(b));                                       \
    b.n = n; b.starts = starts; b.ends = ends; b.branch = branch; b.qs = qs; b.qe = qe; \
    b.mode = mode; b.counts = counts; b.offsets = offsets; b.out = out; b.hashes = hashes; \
    return or_run_##SFX(nthreads, or_search_worker_##SFX, &b, nq);                  \
}                                                                                   \
                                                                                    \
/* Totals of (h, V, J) over all queries into stats[3]. */                          \
void or_walk_stats_batch_##SFX(int nthreads, size_t n, const TY* starts, const TY* ends, \
                               const int64_t* branch, size_t nq, const TY* qs, const TY* qe, \
                               uint64_t* stats) {                                   \
    or_job_##SFX b; memset(&b, 0, sizeof(b));                                       \
    b.n = n; b.starts = starts; b.ends = ends; b.branch = branch; b.qs = qs; b.qe = qe; \
    b.stats = stats;                                                                \
    or_run_##SFX(nthreads, or_stats_worker_##SFX, &b, nq);                          \
}

OR_DEFINE(int32_t, i32)
OR_DEFINE(int64_t, i64)

/* coverage, c.h:758-792 (int32 C form): count and sum of hi - lo, where      */
/* hi = min(ends[i], e) and lo = max(starts[i], s), over the strict walk; the  */
/* sum is int32 and has no +1 (c.h:771-774).                                   */
void or_coverage_i32(size_t n, const int32_t* starts, const int32_t* ends,
                     const int64_t* branch, int32_t s, int32_t e,
                     uint64_t* count_out, int32_t* cov_out) {
    uint64_t c = 0; uint32_t cov = 0;   /* unsigned accumulation = int32 wraparound */
    *count_out = 0; *cov_out = 0;
    if (n == 0) return;
    int64_t i = or_upper_bound_i32(n, starts, e);
    if (i == OR_NONE) return;
    while (i != OR_NONE && s <= ends[i]) {
        int32_t hi = ends[i] < e ? ends[i] : e, lo = starts[i] > s ? starts[i] : s;
        ++c; cov += (uint32_t)hi - (uint32_t)lo; --i;
    }
    if (i != OR_NONE) {
        i = branch[i];
        while (i != OR_NONE) {
            if (s <= ends[i]) {
                int32_t hi = ends[i] < e ? ends[i] : e, lo = starts[i] > s ? starts[i] : s;
                ++c; cov += (uint32_t)hi - (uint32_t)lo; --i;
            } else {
                i = branch[i];
            }
        }
    }
    *count_out = c; *cov_out = (int32_t)cov;
}
