// Owning column (drop-in for the parts of cpp/include/cudf/column/column.hpp
// that group-by results use).  Buffers are released through the memory
// resource they came from.
#pragma once

#include <memory>
#include <vector>

#include "../utilities/memory_resource.hpp"
#include "column_view.hpp"

namespace cudf {

/// A device buffer owned through a memory resource.
class device_buffer {
 public:
  device_buffer() = default;
  device_buffer(void* p, std::size_t bytes, device_memory_resource* mr, stream_view s)
    : _p{p}, _bytes{bytes}, _mr{mr}, _stream{s}
  {
  }
  device_buffer(device_buffer const&)            = delete;
  device_buffer& operator=(device_buffer const&) = delete;
  device_buffer(device_buffer&& o) noexcept { swap(o); }
  device_buffer& operator=(device_buffer&& o) noexcept
  {
    device_buffer tmp{std::move(o)};
    swap(tmp);
    return *this;
  }
  ~device_buffer()
  {
    if (_p && _mr) _mr->deallocate(_p, _bytes, _stream);
  }
  [[nodiscard]] void* data() const noexcept { return _p; }
  [[nodiscard]] std::size_t size() const noexcept { return _bytes; }

 private:
  void swap(device_buffer& o) noexcept
  {
    std::swap(_p, o._p);
    std::swap(_bytes, o._bytes);
    std::swap(_mr, o._mr);
    std::swap(_stream, o._stream);
  }
  void* _p{nullptr};
  std::size_t _bytes{0};
  device_memory_resource* _mr{nullptr};
  stream_view _stream{};
};

class column {
 public:
  column(data_type type, size_type size, device_buffer data, device_buffer null_mask, size_type null_count,
         std::vector<std::unique_ptr<column>> children = {})
    : _type{type},
      _size{size},
      _data{std::move(data)},
      _null_mask{std::move(null_mask)},
      _null_count{null_count},
      _children{std::move(children)}
  {
  }

  [[nodiscard]] data_type type() const noexcept { return _type; }
  [[nodiscard]] size_type size() const noexcept { return _size; }
  [[nodiscard]] size_type null_count() const noexcept { return _null_count; }
  [[nodiscard]] bool nullable() const noexcept { return _null_mask.data() != nullptr; }
  [[nodiscard]] bool has_nulls() const noexcept { return _null_count > 0; }
  [[nodiscard]] size_type num_children() const noexcept { return size_type(_children.size()); }
  [[nodiscard]] column& child(size_type i) { return *_children.at(i); }

  [[nodiscard]] column_view view() const
  {
    std::vector<column_view> kids;
    for (auto const& c : _children) kids.push_back(c->view());
    return column_view{_type,
                       _size,
                       _data.data(),
                       static_cast<bitmask_type const*>(_null_mask.data()),
                       _null_count,
                       0,
                       std::move(kids)};
  }
  operator column_view() const { return view(); }

 private:
  data_type _type;
  size_type _size;
  device_buffer _data;
  device_buffer _null_mask;
  size_type _null_count;
  std::vector<std::unique_ptr<column>> _children;
};

}  // namespace cudf
