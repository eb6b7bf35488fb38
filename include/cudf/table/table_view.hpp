// Non-owning table view (drop-in for cpp/include/cudf/table/table_view.hpp:200)
// and the owning table returned as unique keys.
#pragma once

#include <memory>
#include <vector>

#include "../column/column.hpp"

namespace cudf {

class table_view {
 public:
  table_view() = default;
  table_view(std::vector<column_view> const& cols) : _columns{cols} {}
  table_view(std::initializer_list<column_view> cols) : _columns{cols} {}

  [[nodiscard]] size_type num_columns() const noexcept { return size_type(_columns.size()); }
  [[nodiscard]] size_type num_rows() const noexcept { return _columns.empty() ? 0 : _columns[0].size(); }
  [[nodiscard]] column_view const& column(size_type i) const { return _columns.at(i); }
  [[nodiscard]] auto begin() const noexcept { return _columns.begin(); }
  [[nodiscard]] auto end() const noexcept { return _columns.end(); }

 private:
  std::vector<column_view> _columns;
};

class table {
 public:
  explicit table(std::vector<std::unique_ptr<cudf::column>> columns) : _columns{std::move(columns)} {}
  [[nodiscard]] size_type num_columns() const noexcept { return size_type(_columns.size()); }
  [[nodiscard]] size_type num_rows() const noexcept { return _columns.empty() ? 0 : _columns[0]->size(); }
  [[nodiscard]] cudf::column& get_column(size_type i) { return *_columns.at(i); }
  [[nodiscard]] table_view view() const
  {
    std::vector<column_view> v;
    for (auto const& c : _columns) v.push_back(c->view());
    return table_view{v};
  }
  std::vector<std::unique_ptr<cudf::column>> release() { return std::move(_columns); }

 private:
  std::vector<std::unique_ptr<cudf::column>> _columns;
};

}  // namespace cudf
