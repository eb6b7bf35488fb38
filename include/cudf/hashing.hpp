// cudf::hashing::murmurhash3_x86_32 (drop-in for cpp/include/cudf/hashing.hpp:61):
// MurmurHash3_x86_32 row hashes of a table, over the C ABI.
#pragma once

#include <memory>

#include "column/column.hpp"
#include "table/table_view.hpp"
#include "utilities/error.hpp"
#include "utilities/memory_resource.hpp"

namespace cudf {
namespace hashing {

static constexpr uint32_t DEFAULT_HASH_SEED = 0;  // hashing.hpp:42

inline std::unique_ptr<column> murmurhash3_x86_32(table_view const& input,
                                                  uint32_t seed                  = DEFAULT_HASH_SEED,
                                                  stream_view stream             = get_default_stream(),
                                                  device_memory_resource* mr     = get_current_device_resource())
{
  size_type const n = input.num_rows();
  size_t const bytes = sizeof(uint32_t) * size_t(n > 0 ? n : 1);
  void* out          = mr->allocate(bytes, stream);
  if (!out) throw std::bad_alloc();
  device_buffer buf{out, bytes, mr, stream};
  std::vector<mig_column_view> cols;
  for (auto const& c : input) cols.push_back(c.c_view());
  cudf::detail::throw_on_error(mig_hash_murmurhash3_x86_32(cols.data(), int32_t(cols.size()), seed,
                                                           static_cast<uint32_t*>(out), stream.value()));
  return std::make_unique<column>(data_type{type_id::UINT32}, n, std::move(buf), device_buffer{}, 0);
}

}  // namespace hashing
}  // namespace cudf
