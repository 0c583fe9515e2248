/*
 * legate::dataframe::join -- same enums, overloads, defaults and errors as the reference
 * (cpp/include/legate_dataframe/join.hpp:25-145, cpp/src/join.cpp:328-504), executed by the HIP
 * hash join behind ldf_join (include/ldf_hip.h).
 */
#pragma once

#include <string>
#include <vector>

#include <legate_dataframe/core/table.hpp>

namespace legate::dataframe {

enum class JoinType : int32_t { INNER = 0, LEFT, FULL, SEMI, ANTI };
enum class BroadcastInput : int32_t { AUTO = 0, LEFT, RIGHT };

/**
 * @brief Perform a join between the specified tables.
 *
 * @throw std::invalid_argument if the number of elements in `lhs_keys` and `rhs_keys` differ,
 * if SEMI/ANTI get rhs output columns, or if the output column names are not unique.
 * @throw std::runtime_error if a broadcast is forced where a repartition would be needed.
 */
LogicalTable join(const LogicalTable& lhs,
                  const LogicalTable& rhs,
                  const std::vector<size_t>& lhs_keys,
                  const std::vector<size_t>& rhs_keys,
                  JoinType join_type,
                  const std::vector<size_t>& lhs_out_columns,
                  const std::vector<size_t>& rhs_out_columns,
                  bool nulls_equal         = true,
                  BroadcastInput broadcast = BroadcastInput::AUTO);

LogicalTable join(const LogicalTable& lhs,
                  const LogicalTable& rhs,
                  const std::vector<std::string>& lhs_keys,
                  const std::vector<std::string>& rhs_keys,
                  JoinType join_type,
                  const std::vector<std::string>& lhs_out_columns,
                  const std::vector<std::string>& rhs_out_columns,
                  bool nulls_equal         = true,
                  BroadcastInput broadcast = BroadcastInput::AUTO);

/* Outputs every lhs column and (except for SEMI/ANTI) every rhs column. */
LogicalTable join(const LogicalTable& lhs,
                  const LogicalTable& rhs,
                  const std::vector<size_t>& lhs_keys,
                  const std::vector<size_t>& rhs_keys,
                  JoinType join_type,
                  bool nulls_equal         = true,
                  BroadcastInput broadcast = BroadcastInput::AUTO);

}  // namespace legate::dataframe
