// Host-side group-by driver (C++), called by the C ABI in capi.cpp.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace mig {

// A column owned by a result: every buffer came from `alloc`.
struct owned_column {
  mig_column col{};
  size_t data_bytes    = 0;
  size_t mask_bytes    = 0;
  size_t offsets_bytes = 0;
};

struct device_allocator {
  mig_memory_resource mr{};
  bool has_mr   = false;
  bool borrowed = false;  // columns are the caller's buffers: allocate fails, deallocate does nothing
  hipStream_t stream = nullptr;
  void* allocate(size_t bytes);
  void deallocate(void* p, size_t bytes);
};


}  // namespace mig

struct mig_groupby_result {
  int32_t num_groups = 0;
  std::vector<mig::owned_column> keys;
  std::vector<std::vector<mig::owned_column>> results;
  mig::device_allocator alloc;
  ~mig_groupby_result();
};

namespace mig {

// Statistics of the last call on this thread (exposed by mig_last_stats).
enum stat_index : int {
  STAT_STRATEGY = 0,   // 0 = global table, 1 = LDS pre-aggregation
  STAT_KEY_MODE,       // key_mode
  STAT_CAPACITY,       // table slots
  STAT_EST_GROUPS,     // estimated groups
  STAT_GROUPS,         // actual groups
  STAT_RETRIES,        // table rebuilds after overflow
  STAT_STATE_BYTES,    // bytes of table + accumulators
  STAT_T_ESTIMATE,     // phase times (ms) when phase timing is on
  STAT_T_INIT,
  STAT_T_ACCUMULATE,
  STAT_T_SECOND_PASS,
  STAT_T_COMPACT,
  STAT_T_OUTPUTS,
  STAT_T_RPA_HIST,       // partitioned path kernels (ms)
  STAT_T_RPA_SCATTER1,
  STAT_T_RPA_SCATTER2,
  STAT_T_RPA_AGGREGATE,
  STAT_LDS_PREAGG,       // 1 when the global path pre-aggregated in LDS
  STAT_RPA_SPECULATIVE,  // 1 when the partitioned path's bins were sized without a histogram
  STAT_VERIFIED_PAIRS,   // general keys, hashed form: rows compared with their group's representative
  STAT_COUNT
};
double* thread_stats();
void set_phase_timing(int level);
void set_strategy(int s);

std::unique_ptr<mig_groupby_result> groupby_aggregate(mig_column_view const* keys, int32_t num_keys,
                                                      int32_t null_policy,
                                                      mig_aggregation_request const* requests,
                                                      int32_t num_requests, hipStream_t stream,
                                                      mig_memory_resource const* mr, bool keys_sorted = false);

// Memory of an allocation-free call: device scratch carved from a fixed
// arena, and the caller's output buffers (num_keys key columns, then one
// column per aggregation in request order; see mig_groupby_aggregate_into).
struct call_memory {
  uint8_t* arena     = nullptr;
  size_t arena_bytes = 0;
  mig_column* out_keys    = nullptr;
  mig_column* out_results = nullptr;
};
// The group-by on the device-wide table with no allocation: returns the
// number of groups and fills the caller's column structs.
int32_t groupby_aggregate_into(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                               mig_aggregation_request const* requests, int32_t num_requests, hipStream_t stream,
                               call_memory const& mem);

// mig_groupby_aggregate_into (groupby_noalloc.hip): the same on the
// library's fixed per-device scratch array of noalloc_scratch_bytes().
size_t noalloc_scratch_bytes();
int32_t groupby_aggregate_noalloc(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                                  mig_aggregation_request const* requests, int32_t num_requests,
                                  mig_column* out_keys, mig_column* out_results, hipStream_t stream);

// The sort-based group-by (sort_groupby.cpp): MEDIAN and QUANTILE, and
// NUNIQUE / NTH_ELEMENT of unsorted keys, with the other requested kinds.
std::unique_ptr<mig_groupby_result> sort_groupby(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                                                 mig_aggregation_request const* requests, int32_t num_requests,
                                                 hipStream_t stream, mig_memory_resource const* mr, bool keys_sorted);
int num_cus_of_device();

void hash_rows(mig_column_view const* cols, int32_t ncols, uint32_t seed, uint32_t* out,
               hipStream_t stream);

}  // namespace mig
