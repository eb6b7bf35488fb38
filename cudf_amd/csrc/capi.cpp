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
  if (strategy >= MIG_STRATEGY_AUTO && strategy <= MIG_STRATEGY_PARTITIONED) mig::set_strategy(strategy);
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

}  // extern "C"
