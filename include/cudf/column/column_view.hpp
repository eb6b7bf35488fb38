// Non-owning column view (drop-in for the fields and accessors of
// cpp/include/cudf/column/column_view.hpp:240-247).  Strings follow the
// reference layout: child 0 is the offsets column, the chars are the
// parent's data (column_device_view.cuh:444-453).
#pragma once

#include <vector>

#include "../types.hpp"

namespace cudf {

class column_view {
 public:
  column_view() = default;
  column_view(data_type type, size_type size, void const* data, bitmask_type const* null_mask = nullptr,
              size_type null_count = UNKNOWN_NULL_COUNT, size_type offset = 0,
              std::vector<column_view> children = {})
    : _type{type},
      _size{size},
      _data{data},
      _null_mask{null_mask},
      _null_count{null_mask == nullptr ? 0 : null_count},
      _offset{offset},
      _children{std::move(children)}
  {
  }

  [[nodiscard]] data_type type() const noexcept { return _type; }
  [[nodiscard]] size_type size() const noexcept { return _size; }
  [[nodiscard]] bool is_empty() const noexcept { return _size == 0; }
  [[nodiscard]] void const* head() const noexcept { return _data; }
  template <typename T>
  [[nodiscard]] T const* data() const noexcept
  {
    return static_cast<T const*>(_data) + _offset;
  }
  [[nodiscard]] bitmask_type const* null_mask() const noexcept { return _null_mask; }
  [[nodiscard]] bool nullable() const noexcept { return _null_mask != nullptr; }
  /// Null count as given (UNKNOWN_NULL_COUNT when not known; the library then
  /// counts on the device).
  [[nodiscard]] size_type null_count() const noexcept { return _null_count; }
  [[nodiscard]] bool has_nulls() const noexcept { return _null_count != 0 && _null_mask != nullptr; }
  [[nodiscard]] size_type offset() const noexcept { return _offset; }
  [[nodiscard]] size_type num_children() const noexcept { return size_type(_children.size()); }
  [[nodiscard]] column_view const& child(size_type i) const { return _children.at(i); }

  /// The C ABI descriptor of this view.
  [[nodiscard]] mig_column_view c_view() const
  {
    mig_column_view v{};
    v.type_id    = static_cast<int32_t>(_type.id());
    v.scale      = _type.scale();
    v.size       = _size;
    v.offset     = _offset;
    v.data       = _data;
    v.null_mask  = _null_mask;
    v.null_count = _null_count;
    if (_type.id() == type_id::STRING && !_children.empty()) {
      v.offsets_type_id = static_cast<int32_t>(_children[0].type().id());
      v.offsets         = _children[0].head();
    }
    return v;
  }

 private:
  data_type _type{type_id::EMPTY};
  size_type _size{0};
  void const* _data{nullptr};
  bitmask_type const* _null_mask{nullptr};
  size_type _null_count{0};
  size_type _offset{0};
  std::vector<column_view> _children;
};

}  // namespace cudf
