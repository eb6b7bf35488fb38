// cudf data model subset used by the hash group-by (drop-in for
// cpp/include/cudf/types.hpp:95-235).  Header-only; the compute path is the
// C ABI of mig_groupby.h.
#pragma once

#include <cstddef>
#include <cstdint>

namespace cudf {

using size_type    = int32_t;   // types.hpp:95
using bitmask_type = uint32_t;  // types.hpp:96

/// Null count not yet computed (types.hpp UNKNOWN_NULL_COUNT).
constexpr size_type UNKNOWN_NULL_COUNT{-1};

/// Same numbering as the reference (types.hpp:203-235) and MIG_TYPE_*.
enum class type_id : int32_t {
  EMPTY,
  INT8,
  INT16,
  INT32,
  INT64,
  UINT8,
  UINT16,
  UINT32,
  UINT64,
  FLOAT32,
  FLOAT64,
  BOOL8,
  TIMESTAMP_DAYS,
  TIMESTAMP_SECONDS,
  TIMESTAMP_MILLISECONDS,
  TIMESTAMP_MICROSECONDS,
  TIMESTAMP_NANOSECONDS,
  DURATION_DAYS,
  DURATION_SECONDS,
  DURATION_MILLISECONDS,
  DURATION_MICROSECONDS,
  DURATION_NANOSECONDS,
  DICTIONARY32,
  STRING,
  LIST,
  DECIMAL32,
  DECIMAL64,
  DECIMAL128,
  STRUCT,
  NUM_TYPE_IDS
};

/// Logical type of a column element (types.hpp:240-303).
class data_type {
 public:
  data_type() = default;
  explicit constexpr data_type(type_id id) : _id{id} {}
  constexpr data_type(type_id id, int32_t scale) : _id{id}, _fixed_point_scale{scale} {}
  [[nodiscard]] constexpr type_id id() const noexcept { return _id; }
  [[nodiscard]] constexpr int32_t scale() const noexcept { return _fixed_point_scale; }
  constexpr bool operator==(data_type const& o) const { return _id == o._id && _fixed_point_scale == o._fixed_point_scale; }
  constexpr bool operator!=(data_type const& o) const { return !(*this == o); }

 private:
  type_id _id{type_id::EMPTY};
  int32_t _fixed_point_scale{};
};

enum class null_policy : bool { EXCLUDE, INCLUDE };  // types.hpp:126-129
enum class null_equality : bool { EQUAL, UNEQUAL };  // types.hpp:151-154
enum class sorted : bool { NO, YES };                // types.hpp:167
enum class order : bool { ASCENDING, DESCENDING };
enum class null_order : bool { AFTER, BEFORE };
// types.hpp:192-198
enum class interpolation : int32_t { LINEAR, LOWER, HIGHER, MIDPOINT, NEAREST };

/// Element width in bytes of a fixed-width type, 0 otherwise.
constexpr std::size_t size_of(data_type t)
{
  switch (t.id()) {
    case type_id::INT8:
    case type_id::UINT8:
    case type_id::BOOL8: return 1;
    case type_id::INT16:
    case type_id::UINT16: return 2;
    case type_id::INT32:
    case type_id::UINT32:
    case type_id::FLOAT32:
    case type_id::TIMESTAMP_DAYS:
    case type_id::DURATION_DAYS:
    case type_id::DECIMAL32: return 4;
    case type_id::INT64:
    case type_id::UINT64:
    case type_id::FLOAT64:
    case type_id::TIMESTAMP_SECONDS:
    case type_id::TIMESTAMP_MILLISECONDS:
    case type_id::TIMESTAMP_MICROSECONDS:
    case type_id::TIMESTAMP_NANOSECONDS:
    case type_id::DURATION_SECONDS:
    case type_id::DURATION_MILLISECONDS:
    case type_id::DURATION_MICROSECONDS:
    case type_id::DURATION_NANOSECONDS:
    case type_id::DECIMAL64: return 8;
    case type_id::DECIMAL128: return 16;
    default: return 0;
  }
}

}  // namespace cudf
