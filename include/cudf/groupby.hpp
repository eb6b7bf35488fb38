// cudf::groupby::groupby on the MI355X hash path: drop-in for
// cpp/include/cudf/groupby.hpp:61-199 (aggregation_request,
// aggregation_result, groupby(keys, null_policy, ...).aggregate(requests,
// stream, mr)).  Header-only C++ over the C ABI of mig_groupby.h; all device
// work is in libmig_groupby.so.
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "aggregation.hpp"
#include "column/column.hpp"
#include "table/table_view.hpp"
#include "utilities/error.hpp"
#include "utilities/memory_resource.hpp"
#include "utilities/span.hpp"

namespace cudf {
namespace groupby {

/// groupby.hpp:61-64
struct aggregation_request {
  column_view values;
  std::vector<std::unique_ptr<groupby_aggregation>> aggregations;
};

/// groupby.hpp:88-91
struct aggregation_result {
  std::vector<std::unique_ptr<column>> results;
};

namespace detail {

/// Wrap a column released from a result handle; its buffers are freed
/// through `mr` with `stream`.
inline std::unique_ptr<column> adopt(mig_column const& c, device_memory_resource* mr, stream_view stream)
{
  auto const type = data_type{static_cast<type_id>(c.type_id), c.scale};
  std::vector<std::unique_ptr<column>> children;
  if (type.id() == type_id::STRING) {
    children.push_back(std::make_unique<column>(data_type{type_id::INT32}, c.size + 1,
                                                device_buffer{c.offsets, size_t(c.offsets_bytes), mr, stream},
                                                device_buffer{}, 0));
  }
  return std::make_unique<column>(type, c.size, device_buffer{c.data, size_t(c.data_bytes), mr, stream},
                                  device_buffer{c.null_mask, size_t(c.mask_bytes), mr, stream}, c.null_count,
                                  std::move(children));
}

}  // namespace detail

/// Group-by over `keys` (groupby.hpp:109-200).  Keys are not copied; the
/// caller keeps them alive.  Unsorted keys take the hash path (groups in
/// unspecified order); with `keys_are_sorted == sorted::YES` the groups are
/// the runs of equal adjacent key rows, in input order (the sort path with
/// pre-sorted keys).
class groupby {
 public:
  groupby() = delete;
  explicit groupby(table_view const& keys, null_policy null_handling = null_policy::EXCLUDE,
                   sorted keys_are_sorted = sorted::NO, std::vector<order> const& column_order = {},
                   std::vector<null_order> const& null_precedence = {})
    : _keys{keys}, _include_null_keys{null_handling}, _keys_are_sorted{keys_are_sorted}
  {
    (void)column_order;
    (void)null_precedence;
  }

  /// groupby.hpp:187-199.  Result buffers come from `mr` (the library's own
  /// resource when null); `mr` must outlive the returned columns.
  std::pair<std::unique_ptr<table>, std::vector<aggregation_result>> aggregate(
    host_span<aggregation_request const> requests, stream_view stream = get_default_stream(),
    device_memory_resource* mr = get_current_device_resource())
  {
    std::vector<mig_column_view> keys;
    for (auto const& c : _keys) keys.push_back(c.c_view());
    std::vector<std::vector<mig_aggregation>> aggs(requests.size());
    std::vector<mig_aggregation_request> reqs(requests.size());
    for (size_t r = 0; r < requests.size(); r++) {
      for (auto const& a : requests[r].aggregations)
        aggs[r].push_back(mig_aggregation{static_cast<int32_t>(a->kind), a->_ddof,
                                          static_cast<int32_t>(a->_null_handling), a->_n,
                                          static_cast<int32_t>(a->_interpolation),
                                          static_cast<int32_t>(a->_quantiles.size()), a->_quantiles.data()});
      reqs[r].values           = requests[r].values.c_view();
      reqs[r].aggregations     = aggs[r].data();
      reqs[r].num_aggregations = int32_t(aggs[r].size());
    }
    mig_memory_resource cmr = mr->c_resource();
    mig_groupby_result* handle = nullptr;
    auto const entry = _keys_are_sorted == sorted::YES ? mig_groupby_aggregate_sorted : mig_groupby_aggregate;
    cudf::detail::throw_on_error(entry(
      keys.data(), int32_t(keys.size()), _include_null_keys == null_policy::INCLUDE ? MIG_NULL_INCLUDE : MIG_NULL_EXCLUDE,
      reqs.data(), int32_t(reqs.size()), stream.value(), &cmr, &handle));
    struct closer {
      mig_groupby_result* h;
      ~closer() { mig_result_free(h); }
    } guard{handle};

    std::vector<std::unique_ptr<column>> key_cols;
    for (int32_t i = 0; i < mig_result_num_keys(handle); i++) {
      mig_column c{};
      cudf::detail::throw_on_error(mig_result_release_key(handle, i, &c));
      key_cols.push_back(detail::adopt(c, mr, stream));
    }
    std::vector<aggregation_result> results(requests.size());
    for (int32_t r = 0; r < mig_result_num_requests(handle); r++) {
      for (int32_t a = 0; a < mig_result_num_aggregations(handle, r); a++) {
        mig_column c{};
        cudf::detail::throw_on_error(mig_result_release_column(handle, r, a, &c));
        results[r].results.push_back(detail::adopt(c, mr, stream));
      }
    }
    return {std::make_unique<table>(std::move(key_cols)), std::move(results)};
  }

 private:
  table_view _keys;
  null_policy _include_null_keys;
  sorted _keys_are_sorted;
};

}  // namespace groupby
}  // namespace cudf
