/*
 * legate::dataframe::groupby_aggregation -- same signature and validation as the reference
 * (cpp/include/legate_dataframe/groupby_aggregation.hpp:64-70, cpp/src/groupby_aggregation.cpp:135-232),
 * executed by the HIP groupby behind ldf_groupby (include/ldf_hip.h).
 *
 * FunctionOptions / CountOptions stand in for arrow::compute's classes of the same names; only
 * CountOptions is accepted, by "count" and "count_distinct", as in the reference's GPU variant
 * (cpp/src/groupby_aggregation.cu:25-88).
 */
#pragma once

#include <memory>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include <legate_dataframe/core/table.hpp>

namespace legate::dataframe {

struct FunctionOptions {
  virtual ~FunctionOptions() = default;
  [[nodiscard]] virtual std::string type_name() const = 0;
};

struct CountOptions : FunctionOptions {
  enum CountMode : int32_t { ONLY_VALID = 0, ONLY_NULL = 1, ALL = 2 };
  CountMode mode = ONLY_VALID;
  explicit CountOptions(CountMode m = ONLY_VALID) : mode(m) {}
  [[nodiscard]] std::string type_name() const override { return "CountOptions"; }
};

using ColumnAggregation =
  std::tuple<std::string, std::string, std::optional<std::shared_ptr<FunctionOptions>>, std::string>;

/**
 * @brief Perform a groupby and aggregation in a single operation.
 *
 * `column_aggregations` holds (input column, kind, options, output column).  Kinds: "sum",
 * "product", "min", "max", "count", "mean", "variance", "stddev", "count_distinct", "count_all"
 * ("count_all" takes "" as input column).  The output holds the key columns, then one column per
 * aggregation in order; group order is unspecified.
 */
LogicalTable groupby_aggregation(const LogicalTable& table,
                                 const std::vector<std::string>& keys,
                                 const std::vector<ColumnAggregation>& column_aggregations);

}  // namespace legate::dataframe
