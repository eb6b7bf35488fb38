// cudf::concatenate of tables (drop-in for cpp/include/cudf/concatenate.hpp:95-98),
// over the C ABI.
#pragma once

#include <memory>
#include <vector>

#include "partitioning.hpp"

namespace cudf {

inline std::unique_ptr<table> concatenate(std::vector<table_view> const& tables_to_concat,
                                          stream_view stream         = get_default_stream(),
                                          device_memory_resource* mr = get_current_device_resource())
{
  int32_t const nc = tables_to_concat.empty() ? 0 : tables_to_concat.front().num_columns();
  std::vector<mig_column_view> cols;
  for (auto const& t : tables_to_concat) {
    if (t.num_columns() != nc) throw logic_error("Mismatch in table columns to concatenate.");
    for (auto const& c : t) cols.push_back(c.c_view());
  }
  mig_memory_resource cmr = mr->c_resource();
  mig_table* handle       = nullptr;
  detail::throw_on_error(mig_concatenate(cols.data(), int32_t(tables_to_concat.size()), nc, stream.value(), &cmr,
                                         &handle));
  return detail::adopt_table(handle, mr, stream);
}

}  // namespace cudf
