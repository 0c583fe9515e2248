// join / groupby_aggregation / hash_partition / distinct: the client half of the reference operators
// (validation, output schema, name -> index mapping) on top of the C ABI.
#include <algorithm>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>

#include <legate_dataframe/core/runtime.hpp>
#include <legate_dataframe/groupby_aggregation.hpp>
#include <legate_dataframe/stream_compaction.hpp>
#include <legate_dataframe/hash_partition.hpp>
#include <legate_dataframe/join.hpp>

namespace legate::dataframe {

namespace {

std::vector<LogicalColumn> adopt_all(const std::vector<ldf_column>& cols, const std::vector<DataType>& types, int n)
{
  std::vector<LogicalColumn> out;
  for (int i = 0; i < n; ++i) out.push_back(LogicalColumn::adopt(cols[i], types[i]));
  return out;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// join (reference: cpp/src/join.cpp:328-504)
// ---------------------------------------------------------------------------------------------
LogicalTable join(const LogicalTable& lhs, const LogicalTable& rhs, const std::vector<size_t>& lhs_keys,
                  const std::vector<size_t>& rhs_keys, JoinType join_type, const std::vector<size_t>& lhs_out_columns,
                  const std::vector<size_t>& rhs_out_columns, bool nulls_equal, BroadcastInput broadcast)
{
  if (lhs_keys.size() != rhs_keys.size())
    throw std::invalid_argument("The size of `lhs_keys` and `rhs_keys` must be equal");
  auto lhs_out = lhs.select(lhs_out_columns);
  auto rhs_out = rhs.select(rhs_out_columns);
  switch (join_type) {
    case JoinType::INNER:
    case JoinType::LEFT:
    case JoinType::FULL: break;
    case JoinType::SEMI:
    case JoinType::ANTI:
      if (rhs_out.num_columns() > 0)
        throw std::invalid_argument("rhs must have no output columns for SEMI and ANTI join");
      break;
    default: throw std::invalid_argument("Unknown JoinType");
  }
  // Broadcast hints only select the communication pattern across GPUs; the same forced
  // broadcasts the reference rejects are rejected here (cpp/src/join.cpp:387-396).
  if (broadcast != BroadcastInput::AUTO &&
      (join_type == JoinType::FULL || (broadcast == BroadcastInput::LEFT && join_type != JoinType::INNER))) {
    throw std::runtime_error(
      "Force broadcast was indicated, but repartitioning is required. "
      "FULL joins do not support broadcasting and LEFT/SEMI/ANTI joins only for the "
      "right hand side argument.");
  }
  auto ret_names = lhs_out.get_column_name_vector();
  for (auto& n : rhs_out.get_column_name_vector()) ret_names.push_back(n);
  {
    std::set<std::string> uniq(ret_names.begin(), ret_names.end());
    if (uniq.size() != ret_names.size()) throw std::invalid_argument("column names must be unique");
  }

  std::vector<ldf_column_view> lv, rv;
  ldf_table_view lt = lhs.view(lv);
  ldf_table_view rt = rhs.view(rv);
  std::vector<int32_t> lk(lhs_keys.begin(), lhs_keys.end()), rk(rhs_keys.begin(), rhs_keys.end());
  std::vector<int32_t> lo(lhs_out_columns.begin(), lhs_out_columns.end()), ro(rhs_out_columns.begin(), rhs_out_columns.end());
  for (auto k : lk)
    if (k < 0 || k >= lhs.num_columns()) throw std::out_of_range("lhs key index out of range");
  for (auto k : rk)
    if (k < 0 || k >= rhs.num_columns()) throw std::out_of_range("rhs key index out of range");
  int ncols = (int)(lo.size() + ro.size());
  std::vector<ldf_column> cols(std::max(1, ncols));
  ldf_table_out out{ncols, cols.data()};
  detail::check_status(ldf_join(&lt, &rt, lk.data(), rk.data(), (int32_t)lk.size(), (int32_t)join_type, nulls_equal ? 1 : 0,
                                lo.data(), (int32_t)lo.size(), ro.data(), (int32_t)ro.size(), nullptr, get_stream(), &out));
  std::vector<DataType> types;
  for (auto i : lhs_out_columns) types.push_back(lhs.get_column(i).type());
  for (auto i : rhs_out_columns) types.push_back(rhs.get_column(i).type());
  return LogicalTable(adopt_all(cols, types, ncols), ret_names);
}

LogicalTable join(const LogicalTable& lhs, const LogicalTable& rhs, const std::vector<std::string>& lhs_keys,
                  const std::vector<std::string>& rhs_keys, JoinType join_type,
                  const std::vector<std::string>& lhs_out_columns, const std::vector<std::string>& rhs_out_columns,
                  bool nulls_equal, BroadcastInput broadcast)
{
  std::vector<size_t> lk, rk, lo, ro;
  const auto& ln = lhs.get_column_names();
  const auto& rn = rhs.get_column_names();
  for (auto& n : lhs_keys) lk.push_back(ln.at(n));
  for (auto& n : rhs_keys) rk.push_back(rn.at(n));
  for (auto& n : lhs_out_columns) lo.push_back(ln.at(n));
  for (auto& n : rhs_out_columns) ro.push_back(rn.at(n));
  return join(lhs, rhs, lk, rk, join_type, lo, ro, nulls_equal, broadcast);
}

LogicalTable join(const LogicalTable& lhs, const LogicalTable& rhs, const std::vector<size_t>& lhs_keys,
                  const std::vector<size_t>& rhs_keys, JoinType join_type, bool nulls_equal, BroadcastInput broadcast)
{
  std::vector<size_t> lo(lhs.num_columns());
  std::iota(lo.begin(), lo.end(), 0);
  size_t nr = (join_type == JoinType::SEMI || join_type == JoinType::ANTI) ? 0 : rhs.num_columns();
  std::vector<size_t> ro(nr);
  std::iota(ro.begin(), ro.end(), 0);
  return join(lhs, rhs, lhs_keys, rhs_keys, join_type, lo, ro, nulls_equal, broadcast);
}

// ---------------------------------------------------------------------------------------------
// groupby_aggregation (reference: cpp/src/groupby_aggregation.cpp:100-232 and the kind map of
// cpp/src/groupby_aggregation.cu:46-88)
// ---------------------------------------------------------------------------------------------
namespace {
const std::map<std::string, int32_t>& agg_kinds()
{
  static const std::map<std::string, int32_t> m = {
    {"sum", LDF_AGG_SUM},           {"product", LDF_AGG_PRODUCT},
    {"min", LDF_AGG_MIN},           {"max", LDF_AGG_MAX},
    {"count", LDF_AGG_COUNT},       {"mean", LDF_AGG_MEAN},
    {"variance", LDF_AGG_VARIANCE}, {"stddev", LDF_AGG_STDDEV},
    {"approximate_median", LDF_AGG_APPROX_MEDIAN},
    {"count_distinct", LDF_AGG_COUNT_DISTINCT},
    {"count_all", LDF_AGG_COUNT_ALL}};
  return m;
}

std::string strip_hash_prefix(const std::string& k) { return k.rfind("hash_", 0) == 0 ? k.substr(5) : k; }

DataType agg_output_type(int32_t kind, DataType in)
{
  auto is_float = in.id == TypeId::FLOAT || in.id == TypeId::DOUBLE;
  auto is_uint  = in.id == TypeId::BOOL || in.id == TypeId::UINT8 || in.id == TypeId::UINT16 ||
                 in.id == TypeId::UINT32 || in.id == TypeId::UINT64;
  switch (kind) {
    case LDF_AGG_COUNT:
    case LDF_AGG_COUNT_ALL:
    case LDF_AGG_COUNT_DISTINCT: return DataType(TypeId::INT64);
    case LDF_AGG_MEAN:
    case LDF_AGG_VARIANCE:
    case LDF_AGG_STDDEV:
    case LDF_AGG_APPROX_MEDIAN: return DataType(TypeId::DOUBLE);
    case LDF_AGG_SUM:
    case LDF_AGG_PRODUCT:
      return is_float ? DataType(TypeId::DOUBLE) : (is_uint ? DataType(TypeId::UINT64) : DataType(TypeId::INT64));
    default: return in;
  }
}
}  // namespace

LogicalTable groupby_aggregation(const LogicalTable& table, const std::vector<std::string>& keys,
                                 const std::vector<ColumnAggregation>& column_aggregations)
{
  if (keys.empty()) throw std::invalid_argument("must pass at least one key column");
  std::vector<int32_t> key_idx;
  std::vector<std::string> out_names;
  std::vector<DataType> out_types;
  for (auto& k : keys) {
    key_idx.push_back((int32_t)table.get_column_names().at(k));
    out_names.push_back(k);
    out_types.push_back(table.get_column(k).type());
  }
  std::vector<ldf_agg> aggs;
  for (const auto& [in_col_name, kind_in, options, out_col_name] : column_aggregations) {
    std::string kind = strip_hash_prefix(kind_in);
    auto it          = agg_kinds().find(kind);
    if (it == agg_kinds().end()) throw std::invalid_argument("unknown aggregation function: " + kind_in);
    ldf_agg a{};
    a.kind       = it->second;
    a.count_mode = LDF_COUNT_ONLY_VALID;
    if (options.has_value()) {
      auto* co = dynamic_cast<CountOptions*>(options.value().get());
      if (kind != "count" && kind != "count_distinct")
        throw std::invalid_argument("Options are not supported for this aggregation");
      if (co == nullptr)
        throw std::invalid_argument("options type '" + options.value()->type_name() +
                                    "' does not match expected type 'CountOptions' for function '" + kind + "'");
      a.count_mode = co->mode;
    }
    if (kind == "count_all") {
      if (!in_col_name.empty()) throw std::invalid_argument("count_all must use the empty string as column name.");
      a.in_col = key_idx.at(0);
    } else {
      a.in_col = (int32_t)table.get_column_names().at(in_col_name);
    }
    if (std::find(out_names.begin(), out_names.end(), out_col_name) != out_names.end())
      throw std::invalid_argument("name conflict in the output columns: " + out_col_name);
    out_names.push_back(out_col_name);
    out_types.push_back(agg_output_type(a.kind, table.get_column((size_t)a.in_col).type()));
    aggs.push_back(a);
  }
  std::vector<ldf_column_view> tv;
  ldf_table_view t = table.view(tv);
  int ncols        = (int)(key_idx.size() + aggs.size());
  std::vector<ldf_column> cols(ncols);
  ldf_table_out out{ncols, cols.data()};
  detail::check_status(ldf_groupby(&t, key_idx.data(), (int32_t)key_idx.size(), aggs.data(), (int32_t)aggs.size(),
                                   nullptr, get_stream(), &out));
  return LogicalTable(adopt_all(cols, out_types, ncols), out_names);
}

// ---------------------------------------------------------------------------------------------
// hash partition (reference: cpp/src/core/repartition_by_hash.cu:264-316, local step)
// ---------------------------------------------------------------------------------------------
std::pair<LogicalTable, std::vector<int64_t>> hash_partition(const LogicalTable& table,
                                                             const std::vector<std::string>& columns,
                                                             int32_t num_partitions, uint32_t seed)
{
  std::vector<int32_t> cols;
  for (auto& c : columns) cols.push_back((int32_t)table.get_column_names().at(c));
  std::vector<ldf_column_view> tv;
  ldf_table_view t = table.view(tv);
  int ncols        = table.num_columns();
  std::vector<ldf_column> out_cols(std::max(1, ncols));
  ldf_table_out out{ncols, out_cols.data()};
  std::vector<int64_t> offsets((size_t)std::max(0, num_partitions) + 1);
  detail::check_status(ldf_hash_partition(&t, cols.data(), (int32_t)cols.size(), num_partitions, seed, nullptr,
                                          get_stream(), &out, offsets.data()));
  std::vector<DataType> types;
  for (auto& c : table.get_columns()) types.push_back(c.type());
  return {LogicalTable(adopt_all(out_cols, types, ncols), table.get_column_name_vector()), offsets};
}

// ---------------------------------------------------------------------------------------------
// distinct (reference: cpp/src/stream_compaction.cpp:120-146 and cpp/src/stream_compaction.cu:38-72)
// ---------------------------------------------------------------------------------------------
LogicalTable distinct(const LogicalTable& tbl, const std::vector<std::string>& keys, bool high_cardinality)
{
  (void)high_cardinality;  // only selects a pre-shuffle local distinct in the reference
  std::vector<int32_t> key_idx;
  for (const std::string& key : keys) key_idx.push_back((int32_t)tbl.get_column_names().at(key));
  std::vector<ldf_column_view> tv;
  ldf_table_view t = tbl.view(tv);
  int ncols        = tbl.num_columns();
  std::vector<ldf_column> out_cols(std::max(1, ncols));
  ldf_table_out out{ncols, out_cols.data()};
  detail::check_status(ldf_distinct(&t, key_idx.data(), (int32_t)key_idx.size(), nullptr, get_stream(), &out));
  std::vector<DataType> types;
  for (auto& c : tbl.get_columns()) types.push_back(c.type());
  return LogicalTable(adopt_all(out_cols, types, ncols), tbl.get_column_name_vector());
}

}  // namespace legate::dataframe
