/*
 * LogicalTable: ordered columns plus a name -> index map.
 *
 * Mirrors legate::dataframe::LogicalTable (cpp/include/legate_dataframe/core/table.hpp:49-341 in
 * the reference): same constructors, get_column(index|name), get_columns(),
 * get_column_names() (name -> index map), get_column_name_vector(), select(indices|names).
 */
#pragma once

#include <map>
#include <string>
#include <vector>

#include <legate_dataframe/core/column.hpp>

namespace legate::dataframe {

class LogicalTable {
 public:
  LogicalTable() = default;
  LogicalTable(std::vector<LogicalColumn> columns, std::map<std::string, size_t> column_names);
  LogicalTable(std::vector<LogicalColumn> columns, const std::vector<std::string>& column_names);

  static LogicalTable empty_like(const LogicalTable& other);

  [[nodiscard]] int32_t num_columns() const { return (int32_t)columns_.size(); }
  [[nodiscard]] size_t num_rows() const;
  [[nodiscard]] LogicalColumn& get_column(size_t column_index) { return columns_.at(column_index); }
  [[nodiscard]] const LogicalColumn& get_column(size_t column_index) const { return columns_.at(column_index); }
  [[nodiscard]] LogicalColumn& get_column(const std::string& name) { return get_column(column_names_.at(name)); }
  [[nodiscard]] const LogicalColumn& get_column(const std::string& name) const
  {
    return get_column(column_names_.at(name));
  }
  [[nodiscard]] const std::vector<LogicalColumn>& get_columns() const { return columns_; }
  [[nodiscard]] const std::map<std::string, size_t>& get_column_names() const noexcept { return column_names_; }
  [[nodiscard]] std::vector<std::string> get_column_name_vector() const;
  [[nodiscard]] LogicalTable select(const std::vector<size_t>& column_indices) const;
  [[nodiscard]] LogicalTable select(const std::vector<std::string>& column_names) const;
  [[nodiscard]] bool unbound() const { return false; }
  [[nodiscard]] std::string repr(size_t max_rows = 30) const;

  /* C ABI view; `storage` keeps the column views alive. */
  ldf_table_view view(std::vector<ldf_column_view>& storage) const;

 private:
  std::vector<LogicalColumn> columns_;
  std::map<std::string, size_t> column_names_;
};

}  // namespace legate::dataframe
