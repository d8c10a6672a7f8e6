/*
 * si_device.h -- device-resident superset index and batched overlap queries
 * on AMD Instinct MI355X (gfx950).
 *
 * This is the batch boundary of libsuperintervals_hip.so.  The reference
 * (kcleal/superintervals v1.0.0) answers one query per call; its only batch
 * surface is the Python loop in src/superintervals/intervalmap.pyx:363-494
 * (count_batch / search_idxs_batch / search_values_batch).  The entry points
 * below take whole batches of queries, produce per-query results identical to
 * the reference's single-query functions, and are the calls the C++ template
 * (superintervals.hpp), the C-ABI (c_superintervals.h) and the Python package
 * are built on.
 *
 * Conventions
 *  - Plain pointers and sizes only.  "stream" is a hipStream_t passed as void*
 *    (NULL = the default stream).  Every query entry point is stream-ordered
 *    and re-entrant: concurrent calls on one built index from several streams
 *    or threads are safe (the index is read-only after build).
 *  - Coordinates are int32_t or int64_t, chosen by coord_bytes (4 or 8) when
 *    the index is created; query arrays use the same type.
 *  - Positions ("idx") are uint32_t positions in the sorted order; the maximum
 *    index size is 2^32 - 2 intervals.  SI_INDEX_NONE marks "no interval".
 *  - Functions named *Host take host buffers and synchronise; all others take
 *    device buffers (hipMalloc'd or torch tensors) and do not synchronise.
 *  - Every function returns SI_OK (0) or an SI_ERR_* code; siGetLastError()
 *    returns a message for the calling thread's last failure.
 */
#ifndef SI_DEVICE_H
#define SI_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SI_OK 0
#define SI_ERR_ARG 1          /* invalid argument (null pointer, bad size, bad mode) */
#define SI_ERR_HIP 2          /* a HIP runtime call or kernel launch failed */
#define SI_ERR_NOT_BUILT 3    /* query on an index that has not been built */
#define SI_ERR_CAPACITY 4     /* output capacity smaller than the number of results */
#define SI_ERR_NOMEM 5        /* device or host allocation failed */
#define SI_ERR_UNSUPPORTED 6  /* operation not available for this coordinate type */
#define SI_ERR_MISMATCH 7     /* a fill pass found a different number of hits than its offsets */

/* Output order of index lists. */
#define SI_ORDER_DESC 0   /* strict descending position: c_superintervals.h:610-643 searchIdxs,
                             superintervals.hpp:153-235 IndexIterator, search_values order */
#define SI_ORDER_VECTOR 1 /* superintervals.hpp:879-905 search_idxs(s, e, std::vector&):
                             the leading contiguous run [i*+1 .. idx] ascending, then the
                             branch walk's hits descending */

#define SI_INDEX_NONE 0xFFFFFFFFu

typedef struct siIndex siIndex;

/* Message of the calling thread's last failed call ("" if none). */
const char* siGetLastError(void);

/* Library version string. */
const char* siVersion(void);

/* ---- Index lifecycle ------------------------------------------------------ */

/* Create an empty index for coordinates of coord_bytes (4 = int32, 8 = int64).
 * Returns NULL on failure. */
siIndex* siIndexCreate(int coord_bytes);

/* Free the index and its device memory. */
void siIndexDestroy(siIndex* ix);

/* Build the index from n intervals (end-inclusive), replacing any previous
 * contents.  Replaces superintervals.hpp:112-130 (build: sort_intervals
 * :1415-1439 then the branch stack) and c_superintervals.h:487-521
 * (indexSuperIntervals).
 *   starts, ends   n coordinates each, host memory if on_device == 0, else device.
 *   values         optional int32 payload (NULL: none); kept in sorted order on
 *                  the device for siSearchValuesFill.
 *   start_sorted, end_sorted  the reference's add() flags (superintervals.hpp:95-105):
 *                  start_sorted != 0 promises non-decreasing starts; end_sorted != 0
 *                  additionally promises non-increasing ends within equal starts.
 *                  When both hold the sort is skipped; otherwise a stable device sort
 *                  by (start ascending, end descending) runs, so equal (start, end)
 *                  pairs keep their input order (the C reference's qsort order).
 * Synchronises the stream. */
int siIndexBuild(siIndex* ix, const void* starts, const void* ends, const int32_t* values,
                 size_t n, int start_sorted, int end_sorted, int on_device, void* stream);

/* Number of intervals in the built index (0 before the first build). */
size_t siIndexSize(const siIndex* ix);

/* Coordinate width in bytes (4 or 8). */
int siIndexCoordBytes(const siIndex* ix);

/* Copy the built arrays to host buffers (each may be NULL):
 *   starts, ends  sorted coordinates (n each)
 *   values        int32 payload in sorted order (index built with values only)
 *   perm          perm[k] = input position of the interval at sorted position k
 *   branch        superset index as 64-bit values, UINT64_MAX = none (the
 *                 reference's SIZE_MAX), ready to store into a size_t array. */
int siIndexDownload(const siIndex* ix, void* starts, void* ends, int32_t* values, uint32_t* perm,
                    uint64_t* branch);

/* Device pointers of the built arrays (valid until the next build/destroy). */
const void* siIndexDeviceStarts(const siIndex* ix);
const void* siIndexDeviceEnds(const siIndex* ix);
const uint32_t* siIndexDeviceBranch(const siIndex* ix);
const uint32_t* siIndexDevicePerm(const siIndex* ix);

/* Milliseconds spent in the last siIndexBuild on the device (sort, branch, trees). */
double siIndexLastBuildMs(const siIndex* ix);

/* ---- Launch tuning ----------------------------------------------------------
 * Process-wide launch choices of the walk kernels (results never depend on them):
 *   "tile"       sorted queries per workgroup tile (64 .. 2^20, default 1024)
 *   "count_wgs"  resident workgroups per CU in the count pass (0 = occupancy limit)
 *   "fill_wgs"   resident workgroups per CU in the fill passes (0 = occupancy limit)
 *   "count_rank"   1 (default): siCount* use the rank form (upper_bound + 1 - #{ends < s},
 *                  exact for s <= e; inverted queries walk); 0: the walk
 *   "offsets_rank" 1 (default): siSearchOffsets* count with the rank form; 0: the walk
 *   "lane_steps"   line steps a lane spends on one query before the walk is parked for
 *                  the wave-cooperative kernel (1 .. 2^30, default 256)
 *   "ub_tree"      1 (default): siUpperBound, the plan's upper_bound and the rank-form
 *                  count descend 64-ary trees of first keys (top level of at most 4096
 *                  keys in Eytzinger order in LDS, then one 64-key node per level) over
 *                  the starts / the sorted ends; 0: plain halving
 *   "plan_sync"    1 (default): siPlanQueries reads the key-digit histograms back once to
 *                  skip constant digits (one stream synchronisation); 0: it sorts every
 *                  digit without a read-back, so the planned calls can be graph-captured
 *   "heavy_ends32" 1 (default): the wave-cooperative walk of an int64 index whose ends
 *                  span less than 2^32 reads a 32-bit copy of them (8 B per position with
 *                  exitp instead of 12); 0: the int64 ends
 *   "plan_skip_low" 0 (default) .. 2: lowest 8-bit digits of the query end that
 *                  siPlanQueries leaves unsorted (one radix pass fewer each; the walk
 *                  allows the coarser order at some cost in LDS window hits)
 *   "range_skip"   1 (default): the plan also finds each query's lower_bound and the walks
 *                  take [lower_bound(start), upper_bound(end)] as hits without reading ends
 *                  (exact unless the index holds an interval with start > end, where it is
 *                  off); 0: they walk that run too
 *   "ub_learned"   0 (default); 1: siPlanQueries finds upper_bound / lower_bound through the
 *                  learned model trained on the index (siLearnedTrain attaches it), when the
 *                  index has one; results are identical either way
 *   "learned_shrink" 0 (default); a test knob: the learned windows are narrowed by this many
 *                  positions per side, forcing the model's whole-leaf fallback
 * superintervals_amd/autotune.py measures candidates and stores the best per workload. */
int siTuneSet(const char* name, int64_t value);
int64_t siTuneGet(const char* name);   /* -1 for an unknown name */

/* ---- Planned batches (device buffers) --------------------------------------
 * A plan holds one batch of queries sorted by end together with each query's
 * upper_bound position, so the count and fill kernels walk neighbouring queries in
 * neighbouring lanes.  A plan is a caller-owned device buffer of siPlanBytes bytes;
 * it is valid for the index and queries it was made from until the index is rebuilt.
 * Calls that share one plan must be ordered on one stream (they reuse its scratch).
 * siPlanQueries synchronises the stream once (radix pass selection) unless the tuning
 * knob "plan_sync" is 0. */
size_t siPlanBytes(const siIndex* ix, size_t nq);

int siPlanQueries(const siIndex* ix, const void* qs, const void* qe, size_t nq, void* plan,
                  size_t plan_bytes, void* stream);

/* counts[q] for the planned queries, in input order (see siCount). */
int siCountPlanned(const siIndex* ix, void* plan, size_t nq, uint32_t* counts, void* stream);

/* offsets[0..nq] for the planned queries (see siSearchOffsets); counts may be NULL. */
int siSearchOffsetsPlanned(const siIndex* ix, void* plan, size_t nq, uint32_t* counts,
                           uint64_t* offsets, void* stream);

/* Fill pass for the planned queries (see siSearchFill). */
int siSearchFillPlanned(const siIndex* ix, void* plan, size_t nq, int order,
                        const uint64_t* offsets, uint32_t* idx, void* stream);

/* Checks the last fill pass run with this plan (synchronises the stream): SI_OK when every
 * query's fill wrote exactly the hits its offsets give, SI_ERR_MISMATCH otherwise.  A fill
 * never writes outside a query's segment [offsets[q], offsets[q+1]), so offsets that
 * disagree with the index (the wrong index, or changed offsets) cut or leave gaps in the
 * lists; this call reports it.  The host-buffer forms (siSearchHost and the C-ABI) run it
 * after every fill.  No reference counterpart (the reference fills growing vectors). */
int siPlanCheck(const siIndex* ix, const void* plan, void* stream);

/* Payload fill pass for the planned queries (see siSearchValuesFill). */
int siSearchValuesFillPlanned(const siIndex* ix, void* plan, size_t nq, const uint64_t* offsets,
                              int32_t* values, void* stream);

/* Bounded fill passes: as above, but only output slots [0, capacity) are written (each
 * query's segment is cut at capacity).  For callers that size the output before the
 * counts are known, e.g. a captured graph of plan -> offsets -> fill with fixed buffers
 * (set the tuning knob "plan_sync" to 0 first): after a replay, offsets[nq] <= capacity
 * means the output is complete.  No reference counterpart (the reference fills growing
 * std::vectors, superintervals.hpp:879-905). */
int siSearchFillPlannedBounded(const siIndex* ix, void* plan, size_t nq, int order,
                               const uint64_t* offsets, uint32_t* idx, uint64_t capacity, void* stream);
int siSearchValuesFillPlannedBounded(const siIndex* ix, void* plan, size_t nq, const uint64_t* offsets,
                                     int32_t* values, uint64_t capacity, void* stream);

/* ---- Search without allocation ---------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * two-pass search of siSearchOffsets + siSearchFillPlannedBounded in one call, with the
 * batch's plan kept in a static device array of the library instead of allocated memory
 * (56 MB for the 2^20-query maximum, held on every device the library runs on until the
 * process ends).  Calls no hipMalloc/hipFree and no allocating library routine, and reads
 * nothing back to the host, so it can be captured in a graph.
 *   nq <= SI_NOALLOC_MAX_QUERIES, else SI_ERR_ARG.
 *   counts (may be NULL) and offsets[nq + 1] as siSearchOffsets; with capacity > 0 the
 *   positions in the given order go to idx[0 .. capacity) as siSearchFillPlannedBounded
 *   (offsets[nq] > capacity means the lists were cut); capacity 0 skips the fill.
 * Every call on one device shares that device's array and leaves its contents behind:
 * calls that overlap on one device (concurrent streams or threads) race and must be
 * serialised by the caller.  Results equal siSearchOffsets + siSearchFill. */
#define SI_NOALLOC_MAX_QUERIES (1u << 20)
int siSearchNoAlloc(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                    uint32_t* counts, uint64_t* offsets, uint32_t* idx, uint64_t capacity, void* stream);

/* ---- Set tables: batches over many resident indexes --------------------------
 * No counterpart in the reference (its only multi-index structure is a map from sequence
 * name to IntervalMap, examples/bed-intersect-si.rs:100-123).  A set table holds up to
 * max_sets built int32 indexes in numbered slots; one batch may hold queries of any mix of
 * slots (set_ids[q] = the slot of query q; an id >= max_sets is an empty set).  Results are
 * each slot's own: every query's count and position list equal what siSearchOffsets /
 * siSearchFill give on that slot's index alone.  Putting or evicting a slot touches no other
 * slot.  The table holds device pointers of the indexes: put a slot again after rebuilding
 * its index (calls on a table whose index was rebuilt since it was put return
 * SI_ERR_NOT_BUILT), and do not destroy an index while batches that use it are in flight.
 * Both calls are stream-ordered with no host read-back, so they can be captured in a graph;
 * calls that share one plan must be ordered on one stream. */
typedef struct siSetTable siSetTable;

/* A table of max_sets slots (1 .. 4095) for indexes of coord_bytes (4: int32 only). */
siSetTable* siSetTableCreate(int coord_bytes, uint32_t max_sets);
void siSetTableDestroy(siSetTable* t);

/* Puts a built index into slot (replacing what was there); ix = NULL empties the slot. */
int siSetTablePut(siSetTable* t, uint32_t slot, const siIndex* ix);

/* Device bytes of a plan for nq queries of this table. */
size_t siSetTablePlanBytes(const siSetTable* t, size_t nq);

/* Plans the batch (sorted by slot, then end) into plan and writes offsets[nq + 1] (and
 * counts[nq] when not NULL) in input order, as siSearchOffsets per slot. */
int siSetTableSearchOffsets(const siSetTable* t, const uint32_t* set_ids, const void* qs, const void* qe,
                            size_t nq, uint32_t* counts, uint64_t* offsets, void* plan, size_t plan_bytes,
                            void* stream);

/* Fill pass of a planned batch: each query's positions in its slot's index, in the given
 * order, at idx[offsets[q] ..), writing only slots below capacity (see
 * siSearchFillPlannedBounded). */
int siSetTableSearchFill(const siSetTable* t, void* plan, size_t nq, int order, const uint64_t* offsets,
                         uint32_t* idx, uint64_t capacity, void* stream);

/* ---- Batched queries on device buffers ------------------------------------
 * These plan the batch internally (stream-ordered scratch) and then run. */

/* upper_bound for n values: out[k] = last position with starts[pos] <= xs[k], or
 * SI_INDEX_NONE.  Replaces superintervals.hpp:501-516 / c_superintervals.h:537-568. */
int siUpperBound(const siIndex* ix, const void* xs, size_t n, uint32_t* out, void* stream);

/* Overlap counts: counts[q] = #{j : starts[j] <= qe[q] && ends[j] >= qs[q]}, computed by
 * the superset walk.  Replaces superintervals.hpp:651-825 (count) and
 * c_superintervals.h:729-756 (countOverlaps). */
int siCount(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts,
            void* stream);

/* has_overlaps / anyOverlaps (superintervals.hpp:865-871, c_superintervals.h:570-573). */
int siHasOverlaps(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint8_t* out,
                  void* stream);

/* Workspace bytes siSearchOffsets needs for nq queries (either coordinate width). */
size_t siSearchWorkspaceBytes(size_t nq);

/* First pass of search_idxs: offsets[0..nq] (nq + 1 entries) = exclusive prefix sum of
 * the per-query overlap counts; offsets[nq] is the total.  workspace may be NULL (the
 * library then allocates stream-ordered memory) or point to siSearchWorkspaceBytes(nq)
 * device bytes; it then holds the batch's plan, usable with siSearchFillPlanned.
 * counts may be NULL or receive the per-query counts. */
int siSearchOffsets(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                    uint32_t* counts, uint64_t* offsets, void* workspace, size_t workspace_bytes,
                    void* stream);

/* Second pass: write each query's overlap positions at idx[offsets[q] .. offsets[q+1]) in
 * the given order (SI_ORDER_DESC or SI_ORDER_VECTOR).  Replaces superintervals.hpp:879-905
 * (search_idxs), :153-235 (IndexIterator) and c_superintervals.h:610-643 (searchIdxs). */
int siSearchFill(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                 const uint64_t* offsets, uint32_t* idx, void* stream);

/* As siSearchFill but writes the int32 payload given at build time, in strict
 * descending order.  Replaces superintervals.hpp:551-579 (search_values) and
 * c_superintervals.h:575-608 (searchValues). */
int siSearchValuesFill(const siIndex* ix, const void* qs, const void* qe, size_t nq,
                       const uint64_t* offsets, int32_t* values, void* stream);

/* coverage for int32 indexes: count_out[q] and cov_out[q] = sum over overlaps of
 * min(ends, qe) - max(starts, qs) in int32 arithmetic (no +1), as
 * c_superintervals.h:758-792. */
int siCoverage(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* count_out,
               int32_t* cov_out, void* stream);

/* Keys (and payload) of result positions, for batched search_keys / search_items: after a
 * descending fill (SI_ORDER_DESC), out_starts[k] = starts[idx[k]], out_ends[k] =
 * ends[idx[k]], out_values[k] = values[idx[k]] (coordinate type for starts/ends; any output
 * may be NULL).  Together they give superintervals.hpp:913-971 (search_keys, search_items:
 * the same descending walk) for a whole batch. */
int siGatherKeys(const siIndex* ix, const uint32_t* idx, size_t n, void* out_starts, void* out_ends,
                 int32_t* out_values, void* stream);

/* ---- Set operations over a batched search ---------------------------------------
 * A's intervals k (as, ae, optional data ad, in the caller's order, which is the order of
 * the output) are queries against the index ix (the second set B).
 *
 * siIntersectPieces replaces superintervals.hpp:1158-1174 (intersection): with offsets /
 * idx from a search of A's intervals in ix in the C++ vector order (siSearchFill with
 * SI_ORDER_VECTOR), output j (j in [offsets[k], offsets[k+1])) is the overlap
 * [max(as[k], starts[h]), min(ae[k], ends[h])] of A's interval k with B's interval h =
 * idx[j]; out_adata[j] = ad[k] (the default combine keeps A's data), out_bdata[j] =
 * values[h].  Any output may be NULL.
 *
 * siDifferenceOffsets / siDifferenceFill replace superintervals.hpp:1188-1211
 * (difference): with offsets / idx from a descending search (SI_ORDER_DESC), the first
 * writes the CSR piece_offsets[na + 1] of the uncovered pieces of each A interval
 * (workspace of siSearchWorkspaceBytes(na) bytes), the second writes the pieces
 * [out_starts, out_ends] with A's data, in the reference's order. */
int siIntersectPieces(const siIndex* ix, const void* as, const void* ae, const int32_t* ad, size_t na,
                      const uint64_t* offsets, const uint32_t* idx, void* out_starts, void* out_ends,
                      int32_t* out_adata, int32_t* out_bdata, void* stream);
int siDifferenceOffsets(const siIndex* ix, const void* as, const void* ae, size_t na, const uint64_t* offsets,
                        const uint32_t* idx, uint64_t* piece_offsets, void* workspace, size_t workspace_bytes,
                        void* stream);
int siDifferenceFill(const siIndex* ix, const void* as, const void* ae, const int32_t* ad, size_t na,
                     const uint64_t* offsets, const uint32_t* idx, const uint64_t* piece_offsets,
                     void* out_starts, void* out_ends, int32_t* out_data, void* stream);

/* Reference-walk statistics summed over the batch: totals[0] = hits, totals[1] = reads of
 * ends[] by the walk, totals[2] = reads of branch[] (jumps), totals[3] = queries whose
 * upper_bound found a position.  The walk is c_superintervals.h:610-643. */
int siWalkStats(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint64_t* totals,
                void* stream);

/* ---- Learned position index (no counterpart in the reference) ---------------
 * A two-stage FP32 model of rank(start): a root MLP 1 -> 16 -> 16 -> 1 predicts which of
 * `leaves` linear models serves a key; leaf l owns the positions [l n / L, (l + 1) n / L)
 * and carries the exact bounds of its prediction error over every key it serves.
 * siUpperBoundLearned corrects the predicted leaf by a search of the leaves' first keys,
 * searches only that leaf's window, and checks the result (falling back to the whole
 * leaf), so its output always equals siUpperBound's. */
typedef struct siLearned siLearned;

/* Attach an untrained model to a built index (the index must outlive the model).
 * leaves in [1, 8192] (at most the index size). */
siLearned* siLearnedCreate(const siIndex* ix, uint32_t leaves);
void siLearnedDestroy(siLearned* m);

/* Train on the GPU: `steps` Adam steps of minibatch MSE (batch samples of the index's own
 * (start, position / n) pairs, drawn deterministically), then fit the leaves and their
 * error bounds.  steps = 0 keeps the identity-CDF initialisation (routing by key alone).
 * Attaches the model to its index, so that with the tuning knob "ub_learned" = 1 the plan
 * of every batch (siPlanQueries, and the unplanned calls) finds its bounds through it;
 * siLearnedDestroy and rebuilding the index detach it.  Synchronises the stream. */
int siLearnedTrain(siLearned* m, int steps, uint32_t batch, float lr, void* stream);

/* Number of root parameters (W1[16], b1[16], W2[16][16], b2[16], w3[16], b3). */
int siLearnedParamCount(void);
int siLearnedGetParams(const siLearned* m, float* params);

/* Normalisation of keys: x = (key - kmin) * inv_span. */
int siLearnedNorm(const siLearned* m, int64_t* kmin, float* inv_span);

/* max_abs_err: largest |error bound| of any leaf; last_loss: final minibatch loss;
 * used_leaves: number of leaves.  Each may be NULL. */
int siLearnedInfo(const siLearned* m, int64_t* max_abs_err, float* last_loss, uint32_t* used_leaves);

/* Training-loss gradient at params (host, NULL = the model's) over the given sample
 * positions, computed by the training kernels; grad receives siLearnedParamCount() + 1
 * floats, the last being the loss. */
int siLearnedGradient(siLearned* m, const float* params, const uint32_t* pos, uint32_t count, float* grad);

/* upper_bound through the model, device buffers; identical to siUpperBound.  stats (device,
 * may be NULL, 3 entries) accumulates [0] fallbacks to the whole-leaf search, [1] total
 * window width, [2] total |served leaf - predicted leaf|. */
int siUpperBoundLearned(const siLearned* m, const void* xs, size_t n, uint32_t* out, uint64_t* stats,
                        void* stream);

/* ---- Host-buffer convenience forms (copy in, compute, copy out, synchronise) -- */

int siCountHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* counts);

/* Two-pass search into host CSR buffers.  offsets has nq + 1 entries; idx has capacity
 * cap.  *total receives the number of results; SI_ERR_CAPACITY if it exceeds cap (offsets
 * are still written).  values != NULL selects payload output instead of positions
 * (then order must be SI_ORDER_DESC and idx may be NULL). */
int siSearchHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, int order,
                 uint64_t* offsets, uint32_t* idx, int32_t* values, size_t cap, uint64_t* total);

int siUpperBoundHost(const siIndex* ix, const void* xs, size_t n, uint32_t* out);

int siHasOverlapsHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint8_t* out);

int siCoverageHost(const siIndex* ix, const void* qs, const void* qe, size_t nq, uint32_t* count_out,
                   int32_t* cov_out);

/* ---- Synthetic workloads --------------------------------------------------- */

/* Deterministic generators for the benchmark workloads 1-4 (see workload.cpp).
 * starts/ends are host buffers of n coordinates of coord_bytes each. */
int siWorkloadIntervals(int workload, uint64_t seed, size_t n, int coord_bytes, int sorted,
                        void* starts, void* ends);
int siWorkloadQueries(int workload, uint64_t seed, size_t n, int coord_bytes, void* qs, void* qe);

#ifdef __cplusplus
}
#endif

#endif /* SI_DEVICE_H */
