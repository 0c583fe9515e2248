/*
 * Local hash partition: the one-GPU step of task::repartition_by_hash
 * (cpp/src/core/repartition_by_hash.cu:296-310 in the reference: cudf::hash_partition with
 * HASH_MURMUR3 and DEFAULT_HASH_SEED = 0, then a split at the partition offsets).
 */
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include <legate_dataframe/core/table.hpp>

namespace legate::dataframe {

/* Rows grouped by murmur3(row of `columns`, seed) mod num_partitions, keeping input order inside a
 * partition; returns the table and num_partitions+1 offsets. */
std::pair<LogicalTable, std::vector<int64_t>> hash_partition(const LogicalTable& table,
                                                             const std::vector<std::string>& columns,
                                                             int32_t num_partitions,
                                                             uint32_t seed = 0);

}  // namespace legate::dataframe
