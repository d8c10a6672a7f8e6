// si_common.h -- internal helpers shared by the library's translation units:
// export macro, status codes and the per-thread error message.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/si_device.h"

#define SI_EXPORT extern "C" __attribute__((visibility("default")))

// Records a formatted message for siGetLastError() and returns code.
int si_set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Set operations for the C-ABI on host buffers (int32 indexes), run on the device: A's
// intervals are searched in ix in descending hit order.
//   si_intersect_host: off[na + 1] CSR over A's intervals, per hit its position pos and the
//                      piece [max(starts), min(ends)] (pieces with start > end included).
//   si_difference_host: the uncovered pieces of every A interval in order, with A's data.
int si_intersect_host(const siIndex* ix, const int32_t* as, const int32_t* ae, size_t na,
                      std::vector<uint64_t>& off, std::vector<uint32_t>& pos, std::vector<int32_t>& ps,
                      std::vector<int32_t>& pe);
int si_difference_host(const siIndex* ix, const int32_t* as, const int32_t* ae, const int32_t* ad, size_t na,
                       std::vector<int32_t>& ps, std::vector<int32_t>& pe, std::vector<int32_t>& pd);

// Learned position model (si_learned.hip) for the plan: idx[k] = upper_bound(qe[k]) and, with
// lb, lb[k] = the first position whose start is >= qs[k], through the model attached to ix.
// SI_ERR_ARG when ix has no trained model of its current build.
int si_learned_plan_bounds(const siIndex* ix, const void* qe, const void* qs, size_t nq, uint32_t* idx,
                           uint32_t* lb, void* stream);
// Test-only tuning: every learned leaf window narrowed by this many positions on each side
// (forces the whole-leaf fallback); 0 in normal use.
int si_tune_learned_shrink();
// The index a model was created on is being destroyed: the model forgets it.
void si_learned_forget_index(siLearned* m);
// Build-time tuning of the long list: > 0 the length threshold, 0 chosen by the build from
// the lengths, < 0 no list.
int64_t si_tune_long_len();
