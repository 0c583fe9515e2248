/*
 * distinct (reference: cpp/include/legate_dataframe/stream_compaction.hpp:62-76 and the client in
 * cpp/src/stream_compaction.cpp:120-146).
 */
#pragma once

#include <string>
#include <vector>

#include <legate_dataframe/core/table.hpp>

namespace legate::dataframe {

/* One row per distinct tuple of the `keys` columns, with every column of `tbl`; which row of a key
 * is kept is unspecified in the reference (this build keeps the first, in input order).
 * `high_cardinality` only decides, in the reference, whether a local distinct runs before the
 * shuffle between GPUs; on one GPU it changes nothing. */
LogicalTable distinct(const LogicalTable& tbl, const std::vector<std::string>& keys, bool high_cardinality = false);

}  // namespace legate::dataframe
