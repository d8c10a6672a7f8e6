/*
 * c_superintervals.h -- drop-in C-ABI of libsuperintervals_hip.so.
 *
 * Declares the query-path functions of the reference's C implementation
 * (kcleal/superintervals v1.0.0, src/c_superintervals.h) with identical names,
 * signatures and struct layouts, so a C caller (or an FFI binding) links against
 * the HIP library instead of compiling the reference header.  Unlike the
 * reference this header holds declarations only; the definitions live in the
 * shared library.
 *
 * Each declaration names the reference function it replaces (file:line in
 * src/c_superintervals.h).  The index is built on the GPU by
 * indexSuperIntervals; every query runs on the GPU.  Results are identical to
 * the reference: same counts, same int32 indices in the same (descending)
 * order, same payload order for stable input (the reference's qsort on
 * glibc 2.35).
 *
 * Differences a caller can observe:
 *   - cSuperIntervals carries one extra trailing field (dev).  The reference's
 *     fields keep their order, types and meaning; after indexSuperIntervals the
 *     host arrays hold the sorted intervals and branch holds the superset index
 *     exactly as in the reference.
 *   - Single-query functions are GPU round trips (tens of microseconds each).
 *     Use the si*Batch functions below for throughput.
 *   - Device failures on the reference's void functions are reported through
 *     siGetLastError() (si_device.h) and a zero/none result.
 *   - Of the set operations (c_superintervals.h:264-332) this library provides
 *     intersection and difference, the two that query an index (batched device
 *     searches).  mergeOverlaps, intervalGaps, unionWith, symmetricDifference,
 *     intervalSpan, expandIntervals, flankIntervals and uniqueIntervals are
 *     host-side sorts and merges of the arrays and are not part of it.
 *   - addInterval and clearSuperIntervals release the device index: queries on a
 *     handle changed since its last indexSuperIntervals report SI_ERR_NOT_BUILT
 *     through siGetLastError() and return no result (the reference leaves them
 *     undefined).
 */
#ifndef SUPERINTERVALS_HIP_C_HEADER
#define SUPERINTERVALS_HIP_C_HEADER 1

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "si_device.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SUPERINTERVALS_C_VERSION "1.0.0"

/* c_superintervals.h:56 */
#define SI_NONE ((size_t)-1)

/* c_superintervals.h:61-65 */
typedef struct {
    int32_t start;
    int32_t end;
    int32_t data;
} Interval;

/* c_superintervals.h:70-73 */
typedef struct {
    int32_t start;
    int32_t end;
} KeyPair;

/* c_superintervals.h:81-91, plus the trailing device handle. */
typedef struct {
    int32_t* starts;
    int32_t* ends;
    int32_t* data;
    size_t* branch;
    size_t size;
    size_t capacity;
    size_t idx;
    bool startSorted;
    bool endSorted;
    siIndex* dev; /* device-resident index (owned); not in the reference */
} cSuperIntervals;

/* c_superintervals.h:98-102 */
typedef struct {
    int32_t* data;
    size_t size;
    size_t capacity;
} cIndexResult;

/* c_superintervals.h:107-111 */
typedef struct {
    KeyPair* data;
    size_t size;
    size_t capacity;
} cKeyResult;

/* c_superintervals.h:116-120 */
typedef struct {
    Interval* data;
    size_t size;
    size_t capacity;
} cItemResult;

/* c_superintervals.h:128 */
typedef int32_t (*cCombineFn)(int32_t accumulated, int32_t next);

/* ---- Lifecycle (c_superintervals.h:135-156) ---- */
cSuperIntervals* createSuperIntervals(void);                       /* :135, :363-375 */
void destroySuperIntervals(cSuperIntervals* si);                   /* :138, :377-384 */
void clearSuperIntervals(cSuperIntervals* si);                     /* :141, :386-391 */
void reserveSuperIntervals(cSuperIntervals* si, size_t n);         /* :144, :393-400 */
void addInterval(cSuperIntervals* si, int32_t start, int32_t end, int32_t value); /* :153, :402-419 */
size_t sizeSuperIntervals(const cSuperIntervals* si);              /* :156, :421-423 */

/* ---- Indexing (c_superintervals.h:163-169) ---- */
void sortIntervals(cSuperIntervals* si);                           /* :163, :462-485 (device sort) */
void indexSuperIntervals(cSuperIntervals* si);                     /* :169, :487-521 (device build) */

/* ---- Element access (c_superintervals.h:176-185) ---- */
bool intervalAt(const cSuperIntervals* si, size_t index, Interval* out); /* :176, :523-531 */
int32_t startAt(const cSuperIntervals* si, size_t index);          /* :179, :533 */
int32_t endAt(const cSuperIntervals* si, size_t index);            /* :182, :534 */
int32_t dataAt(const cSuperIntervals* si, size_t index);           /* :185, :535 */

/* ---- Cursor (c_superintervals.h:195) ---- */
size_t upperBound(cSuperIntervals* si, int32_t value);             /* :195, :537-568 */

/* ---- Queries (c_superintervals.h:202-240) ---- */
bool anyOverlaps(cSuperIntervals* si, int32_t start, int32_t end); /* :202, :570-573 */
size_t countOverlaps(cSuperIntervals* si, int32_t start, int32_t end);            /* :205, :729-756 */
void searchValues(cSuperIntervals* si, int32_t start, int32_t end, cIndexResult* found); /* :211, :575-608 */
void searchIdxs(cSuperIntervals* si, int32_t start, int32_t end, cIndexResult* found);   /* :214, :610-643 */
void searchKeys(cSuperIntervals* si, int32_t start, int32_t end, cKeyResult* found);     /* :217, :645-682 */
void searchItems(cSuperIntervals* si, int32_t start, int32_t end, cItemResult* found);   /* :220, :684-723 */
void searchPoint(cSuperIntervals* si, int32_t point, cIndexResult* found);               /* :223, :725-727 */
void coverage(cSuperIntervals* si, int32_t start, int32_t end,
              size_t* count_out, int32_t* coverage_out);                                  /* :230, :758-792 */
void findOverlaps(cSuperIntervals* si, int32_t start, int32_t end,
                  int32_t* found, size_t* found_size);                                    /* :239, :794-821 */

/* ---- Set operations (c_superintervals.h:286-292) ----
 * Both return a new, unindexed handle (caller owns; destroySuperIntervals).  other must be
 * indexed; si is read in its current order.  The searches of si's intervals in other run
 * as one batch on the GPU. */
/* :286, :921-940 -- every non-empty overlap piece, per si interval in searchIdxs (descending)
 * order of other's hits; combine(si data, other data), NULL keeps si's data. */
cSuperIntervals* intersection(const cSuperIntervals* si, cSuperIntervals* other, cCombineFn combine);
/* :292, :950-975 -- the parts of each si interval that no interval of other covers, with
 * its data.  A hit ending at INT32_MAX covers the rest of the interval (the reference's
 * ce + 1 overflows there). */
cSuperIntervals* difference(const cSuperIntervals* si, cSuperIntervals* other);

/* ---- Result helpers (c_superintervals.h:339-357) ---- */
cIndexResult createIndexResult(void);     /* :339, :1070-1073 */
void clearIndexResult(cIndexResult* r);   /* :341, :1074 */
void destroyIndexResult(cIndexResult* r); /* :343, :1075 */
cKeyResult createKeyResult(void);         /* :346, :1077-1080 */
void clearKeyResult(cKeyResult* r);       /* :348, :1081 */
void destroyKeyResult(cKeyResult* r);     /* :350, :1082 */
cItemResult createItemResult(void);       /* :353, :1084-1087 */
void clearItemResult(cItemResult* r);     /* :355, :1088 */
void destroyItemResult(cItemResult* r);   /* :357, :1089 */

/* ---- Batch extensions (no single-call counterpart in the reference; shaped after
 *      intervalmap.pyx:363-494 count_batch / search_idxs_batch / search_values_batch).
 *      Host buffers; the handle must be indexed.  Each per-query result equals the
 *      single-query function's.  Return SI_OK or an SI_ERR_* code. ---- */

/* counts[q] = countOverlaps(si, qs[q], qe[q]) */
int siCountBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                 uint32_t* counts);

/* CSR of searchIdxs (descending) results: offsets[n + 1], idx[cap].  *total receives the
 * number of indices; SI_ERR_CAPACITY if it exceeds cap (offsets are still written, so a
 * caller may size idx from offsets[n] and call again). */
int siSearchIdxsBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                      uint64_t* offsets, int32_t* idx, size_t cap, uint64_t* total);

/* CSR of searchValues results (data payloads, descending position order). */
int siSearchValuesBatch(cSuperIntervals* si, const int32_t* qs, const int32_t* qe, size_t n,
                        uint64_t* offsets, int32_t* values, size_t cap, uint64_t* total);

#ifdef __cplusplus
}
#endif

#endif /* SUPERINTERVALS_HIP_C_HEADER */
