// cudf::hash_partition and cudf::partition (drop-in for
// cpp/include/cudf/partitioning.hpp:41-112), over the C ABI.  Rows inside a
// partition keep their input order.
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "groupby.hpp"
#include "hashing.hpp"

namespace cudf {

enum class hash_id { HASH_IDENTITY = 0, HASH_MURMUR3 };  // partitioning.hpp:41-44

namespace detail {

inline std::unique_ptr<table> adopt_table(mig_table* handle, device_memory_resource* mr, stream_view stream)
{
  struct closer {
    mig_table* h;
    ~closer() { mig_table_free(h); }
  } guard{handle};
  std::vector<std::unique_ptr<column>> cols;
  for (int32_t i = 0; i < mig_table_num_columns(handle); i++) {
    mig_column c{};
    throw_on_error(mig_table_release_column(handle, i, &c));
    cols.push_back(groupby::detail::adopt(c, mr, stream));
  }
  return std::make_unique<table>(std::move(cols));
}

inline std::vector<mig_column_view> c_views(table_view const& t)
{
  std::vector<mig_column_view> v;
  for (auto const& c : t) v.push_back(c.c_view());
  return v;
}

}  // namespace detail

inline std::pair<std::unique_ptr<table>, std::vector<size_type>> hash_partition(
  table_view const& input,
  std::vector<size_type> const& columns_to_hash,
  int num_partitions,
  hash_id hash_function          = hash_id::HASH_MURMUR3,
  uint32_t seed                  = hashing::DEFAULT_HASH_SEED,
  stream_view stream             = get_default_stream(),
  device_memory_resource* mr     = get_current_device_resource())
{
  auto cols = detail::c_views(input);
  std::vector<size_type> offsets(num_partitions > 0 ? size_t(num_partitions) : 0);
  mig_memory_resource cmr = mr->c_resource();
  mig_table* handle       = nullptr;
  detail::throw_on_error(mig_hash_partition(cols.data(), int32_t(cols.size()), columns_to_hash.data(),
                                            int32_t(columns_to_hash.size()), num_partitions,
                                            static_cast<int32_t>(hash_function), seed, stream.value(), &cmr,
                                            &handle, offsets.data()));
  return {detail::adopt_table(handle, mr, stream), std::move(offsets)};
}

inline std::pair<std::unique_ptr<table>, std::vector<size_type>> partition(
  table_view const& t,
  column_view const& partition_map,
  size_type num_partitions,
  device_memory_resource* mr = get_current_device_resource())
{
  auto cols = detail::c_views(t);
  auto map  = partition_map.c_view();
  std::vector<size_type> offsets(size_t(num_partitions > 0 ? num_partitions : 0) + 1);
  mig_memory_resource cmr = mr->c_resource();
  mig_table* handle       = nullptr;
  stream_view stream      = get_default_stream();
  detail::throw_on_error(mig_partition(cols.data(), int32_t(cols.size()), &map, num_partitions, stream.value(),
                                       &cmr, &handle, offsets.data()));
  return {detail::adopt_table(handle, mr, stream), std::move(offsets)};
}

}  // namespace cudf
