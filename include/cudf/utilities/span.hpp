// Minimal host_span (drop-in for cpp/include/cudf/utilities/span.hpp): a
// non-owning view of contiguous host elements.
#pragma once

#include <cstddef>
#include <vector>

namespace cudf {

template <typename T>
class host_span {
 public:
  constexpr host_span() = default;
  constexpr host_span(T* data, std::size_t size) : _data{data}, _size{size} {}
  template <typename U, typename A>
  host_span(std::vector<U, A>& v) : _data{v.data()}, _size{v.size()}
  {
  }
  template <typename U, typename A>
  host_span(std::vector<U, A> const& v) : _data{v.data()}, _size{v.size()}
  {
  }
  [[nodiscard]] constexpr T* data() const noexcept { return _data; }
  [[nodiscard]] constexpr std::size_t size() const noexcept { return _size; }
  [[nodiscard]] constexpr bool empty() const noexcept { return _size == 0; }
  constexpr T& operator[](std::size_t i) const { return _data[i]; }
  constexpr T* begin() const noexcept { return _data; }
  constexpr T* end() const noexcept { return _data + _size; }

 private:
  T* _data{nullptr};
  std::size_t _size{0};
};

}  // namespace cudf
