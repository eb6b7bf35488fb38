// extern "C" boundary of the library (declared in include/mig_groupby.h).
// Exceptions never cross it: each entry point converts them to a status code
// and keeps the message for mig_last_error(), the way the reference's JNI
// layer converts cudf exceptions with CATCH_STD (java/src/main/native/src/
// TableJni.cpp:3716).
#include <cstring>
#include <new>
#include <string>

#include "batch_host.hpp"
#include "groupby.hpp"
#include "partitioning.hpp"

namespace {

std::string& last_error()
{
  static thread_local std::string msg;
  return msg;
}

template <typename F>
mig_status guarded(F&& f)
{
  try {
    last_error().clear();
    f();
    return MIG_OK;
  } catch (mig::mig_error const& e) {
    last_error() = e.what();
    return e.status;
  } catch (std::bad_alloc const&) {
    last_error() = "host allocation failed";
    return MIG_ERROR_ALLOC;
  } catch (std::exception const& e) {
    last_error() = e.what();
    return MIG_ERROR_LOGIC;
  } catch (...) {
    last_error() = "unknown error";
    return MIG_ERROR_LOGIC;
  }
}

mig::owned_column* result_col(mig_groupby_result const* r, int32_t request, int32_t agg)
{
  if (!r) mig::fail(MIG_ERROR_LOGIC, "null result handle");
  if (request < 0 || request >= int32_t(r->results.size()))
    mig::fail(MIG_ERROR_LOGIC, "request index out of range");
  auto& v = const_cast<mig_groupby_result*>(r)->results[request];
  if (agg < 0 || agg >= int32_t(v.size())) mig::fail(MIG_ERROR_LOGIC, "aggregation index out of range");
  return &v[agg];
}

// The C view of an owned column, with the allocation sizes of its buffers.
mig_column exported(mig::owned_column const& c)
{
  mig_column out    = c.col;
  out.data_bytes    = int64_t(c.data_bytes);
  out.mask_bytes    = int64_t(c.mask_bytes);
  out.offsets_bytes = int64_t(c.offsets_bytes);
  return out;
}

mig::owned_column* key_col(mig_groupby_result const* r, int32_t i)
{
  if (!r) mig::fail(MIG_ERROR_LOGIC, "null result handle");
  if (i < 0 || i >= int32_t(r->keys.size())) mig::fail(MIG_ERROR_LOGIC, "key index out of range");
  return &const_cast<mig_groupby_result*>(r)->keys[i];
}

mig::owned_column* table_col(mig_table const* t, int32_t i)
{
  if (!t) mig::fail(MIG_ERROR_LOGIC, "null table handle");
  if (i < 0 || i >= int32_t(t->columns.size())) mig::fail(MIG_ERROR_OUT_OF_RANGE, "column index out of range");
  return &const_cast<mig_table*>(t)->columns[i];
}

}  // namespace

extern "C" {

const char* mig_last_error(void) { return last_error().c_str(); }

int32_t mig_abi_version(void) { return MIG_ABI_VERSION; }

void* mig_device_allocate(size_t bytes, void* stream)
{
  void* p = nullptr;
  if (hipMallocAsync(&p, bytes == 0 ? 1 : bytes, static_cast<hipStream_t>(stream)) != hipSuccess) {
    last_error() = "hipMallocAsync failed";
    return nullptr;
  }
  return p;
}

void mig_device_deallocate(void* ptr, size_t, void* stream)
{
  if (ptr) (void)hipFreeAsync(ptr, static_cast<hipStream_t>(stream));
}

mig_status mig_hash_murmurhash3_x86_32(const mig_column_view* columns, int32_t num_columns,
                                       uint32_t seed, uint32_t* out, void* stream)
{
  return guarded([&] {
    if (num_columns > 0 && !columns) mig::fail(MIG_ERROR_LOGIC, "null column array");
    mig::hash_rows(columns, num_columns, seed, out, static_cast<hipStream_t>(stream));
  });
}

mig_status mig_groupby_aggregate(const mig_column_view* keys, int32_t num_keys, int32_t null_policy,
                                 const mig_aggregation_request* requests, int32_t num_requests,
                                 void* stream, const mig_memory_resource* mr,
                                 mig_groupby_result** out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    if (num_keys > 0 && !keys) mig::fail(MIG_ERROR_LOGIC, "null key array");
    if (num_requests > 0 && !requests) mig::fail(MIG_ERROR_LOGIC, "null request array");
    auto r = mig::groupby_aggregate(keys, num_keys, null_policy, requests, num_requests,
                                    static_cast<hipStream_t>(stream), mr);
    *out = r.release();
  });
}

mig_status mig_groupby_aggregate_into(const mig_column_view* keys, int32_t num_keys, int32_t null_policy,
                                      const mig_aggregation_request* requests, int32_t num_requests,
                                      mig_column* out_keys, mig_column* out_results, int32_t* num_groups,
                                      void* stream)
{
  return guarded([&] {
    if (!num_groups) mig::fail(MIG_ERROR_LOGIC, "null group count");
    *num_groups = 0;
    if (num_keys > 0 && (!keys || !out_keys)) mig::fail(MIG_ERROR_LOGIC, "null key or key output array");
    if (num_requests > 0 && !requests) mig::fail(MIG_ERROR_LOGIC, "null request array");
    int64_t outs = 0;
    for (int32_t r = 0; r < num_requests; r++) outs += requests[r].num_aggregations > 0 ? requests[r].num_aggregations : 0;
    if (outs > 0 && !out_results) mig::fail(MIG_ERROR_LOGIC, "null result column array");
    *num_groups = mig::groupby_aggregate_noalloc(keys, num_keys, null_policy, requests, num_requests, out_keys,
                                                 out_results, static_cast<hipStream_t>(stream));
  });
}

size_t mig_groupby_into_scratch_bytes(void) { return mig::noalloc_scratch_bytes(); }

mig_status mig_groupby_aggregate_sorted(const mig_column_view* keys, int32_t num_keys, int32_t null_policy,
                                        const mig_aggregation_request* requests, int32_t num_requests,
                                        void* stream, const mig_memory_resource* mr,
                                        mig_groupby_result** out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    if (num_keys > 0 && !keys) mig::fail(MIG_ERROR_LOGIC, "null key array");
    if (num_requests > 0 && !requests) mig::fail(MIG_ERROR_LOGIC, "null request array");
    auto r = mig::groupby_aggregate(keys, num_keys, null_policy, requests, num_requests,
                                    static_cast<hipStream_t>(stream), mr, true);
    *out = r.release();
  });
}

int32_t mig_result_num_groups(const mig_groupby_result* r) { return r ? r->num_groups : 0; }
int32_t mig_result_num_keys(const mig_groupby_result* r) { return r ? int32_t(r->keys.size()) : 0; }
int32_t mig_result_num_requests(const mig_groupby_result* r)
{
  return r ? int32_t(r->results.size()) : 0;
}
int32_t mig_result_num_aggregations(const mig_groupby_result* r, int32_t request)
{
  if (!r || request < 0 || request >= int32_t(r->results.size())) return 0;
  return int32_t(r->results[request].size());
}

mig_status mig_result_key(const mig_groupby_result* r, int32_t i, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    *out = exported(*key_col(r, i));
  });
}

mig_status mig_result_column(const mig_groupby_result* r, int32_t request, int32_t agg, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    *out = exported(*result_col(r, request, agg));
  });
}

mig_status mig_result_release_key(mig_groupby_result* r, int32_t i, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    auto* c = key_col(r, i);
    *out    = exported(*c);
    *c      = mig::owned_column{};
  });
}

mig_status mig_result_release_column(mig_groupby_result* r, int32_t request, int32_t agg, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    auto* c = result_col(r, request, agg);
    *out    = exported(*c);
    *c      = mig::owned_column{};
  });
}

void mig_result_free(mig_groupby_result* r) { delete r; }

int32_t mig_last_stats(double* values, int32_t n)
{
  double const* s = mig::thread_stats();
  for (int32_t i = 0; i < n && i < mig::STAT_COUNT; i++) values[i] = s[i];
  return mig::STAT_COUNT;
}

void mig_set_phase_timing(int32_t enable) { mig::set_phase_timing(enable == 2 ? 2 : (enable != 0 ? 1 : 0)); }

void mig_set_strategy(int32_t strategy)
{
  if (strategy >= MIG_STRATEGY_AUTO && strategy <= MIG_STRATEGY_PARTITIONED_EXACT) mig::set_strategy(strategy);
}

mig_status mig_batch_plan_create(const mig_batch_instance* instances, int32_t num_instances,
                                 int32_t num_keys, int32_t null_policy,
                                 const mig_aggregation* aggregations, const int32_t* num_aggregations,
                                 int32_t num_requests, const mig_batch_options* options, void* stream,
                                 mig_batch_plan** out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    *out = mig::batch_plan_create(instances, num_instances, num_keys, null_policy, aggregations,
                                  num_aggregations, num_requests, options, static_cast<hipStream_t>(stream));
  });
}

mig_status mig_batch_execute(mig_batch_plan* plan, void* stream)
{
  return guarded([&] { mig::batch_execute(plan, static_cast<hipStream_t>(stream)); });
}

mig_status mig_batch_synchronize(mig_batch_plan* plan, void* stream)
{
  return guarded([&] { mig::batch_synchronize(plan, static_cast<hipStream_t>(stream)); });
}

int32_t mig_batch_num_groups(const mig_batch_plan* plan, int32_t instance)
{
  return mig::batch_num_groups(plan, instance);
}

mig_status mig_batch_result_key(const mig_batch_plan* plan, int32_t instance, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    *out = mig::batch_result_key(plan, instance);
  });
}

mig_status mig_batch_result_column(const mig_batch_plan* plan, int32_t instance, int32_t request,
                                   int32_t agg, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    *out = mig::batch_result_column(plan, instance, request, agg);
  });
}

void mig_batch_plan_free(mig_batch_plan* plan) { mig::batch_plan_free(plan); }

int32_t mig_table_num_rows(const mig_table* t) { return t ? t->num_rows : 0; }

int32_t mig_table_num_columns(const mig_table* t) { return t ? int32_t(t->columns.size()) : 0; }

mig_status mig_table_column(const mig_table* t, int32_t i, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    *out = exported(*table_col(t, i));
  });
}

mig_status mig_table_release_column(mig_table* t, int32_t i, mig_column* out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output column");
    auto* c = table_col(t, i);
    *out    = exported(*c);
    *c      = mig::owned_column{};
  });
}

void mig_table_free(mig_table* t) { delete t; }

mig_status mig_hash_partition(const mig_column_view* input, int32_t num_columns, const int32_t* columns_to_hash,
                              int32_t num_columns_to_hash, int32_t num_partitions, int32_t hash_function,
                              uint32_t seed, void* stream, const mig_memory_resource* mr, mig_table** out,
                              int32_t* partition_offsets)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    if (num_partitions > 0 && !partition_offsets) mig::fail(MIG_ERROR_LOGIC, "null offsets array");
    std::vector<int32_t> offsets;
    auto t = mig::hash_partition(input, num_columns, columns_to_hash, num_columns_to_hash, num_partitions,
                                 hash_function, seed, static_cast<hipStream_t>(stream), mr, offsets);
    if (!offsets.empty()) std::memcpy(partition_offsets, offsets.data(), sizeof(int32_t) * offsets.size());
    *out = t.release();
  });
}

mig_status mig_partition(const mig_column_view* input, int32_t num_columns, const mig_column_view* partition_map,
                         int32_t num_partitions, void* stream, const mig_memory_resource* mr, mig_table** out,
                         int32_t* partition_offsets)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    if (!partition_map) mig::fail(MIG_ERROR_LOGIC, "null partition map");
    if (!partition_offsets) mig::fail(MIG_ERROR_LOGIC, "null offsets array");
    std::vector<int32_t> offsets;
    auto t = mig::partition_by_map(input, num_columns, *partition_map, num_partitions,
                                   static_cast<hipStream_t>(stream), mr, offsets);
    std::memcpy(partition_offsets, offsets.data(), sizeof(int32_t) * offsets.size());
    *out = t.release();
  });
}

mig_status mig_concatenate(const mig_column_view* tables, int32_t num_tables, int32_t num_columns, void* stream,
                           const mig_memory_resource* mr, mig_table** out)
{
  return guarded([&] {
    if (!out) mig::fail(MIG_ERROR_LOGIC, "null output handle");
    *out = nullptr;
    *out = mig::concatenate(tables, num_tables, num_columns, static_cast<hipStream_t>(stream), mr).release();
  });
}

}  // extern "C"
