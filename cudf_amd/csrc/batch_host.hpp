// Host driver of the batched group-by (batch.cpp), called by capi.cpp.
#pragma once

#include "common.hpp"

namespace mig {

mig_batch_plan* batch_plan_create(mig_batch_instance const* instances, int32_t num_instances,
                                  int32_t num_keys, int32_t null_policy,
                                  mig_aggregation const* aggregations, int32_t const* num_aggregations,
                                  int32_t num_requests, mig_batch_options const* options,
                                  hipStream_t stream);
void batch_execute(mig_batch_plan* plan, hipStream_t stream);
void batch_synchronize(mig_batch_plan* plan, hipStream_t stream);
int32_t batch_num_groups(mig_batch_plan const* plan, int32_t instance);
mig_column batch_result_key(mig_batch_plan const* plan, int32_t instance);
mig_column batch_result_column(mig_batch_plan const* plan, int32_t instance, int32_t request,
                               int32_t agg);
void batch_plan_free(mig_batch_plan* plan);

}  // namespace mig
