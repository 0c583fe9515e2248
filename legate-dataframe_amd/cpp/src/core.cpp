// LogicalColumn / LogicalTable storage, host <-> device copies, stream and status handling.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <set>
#include <sstream>

#include <legate_dataframe/core/column.hpp>
#include <legate_dataframe/core/runtime.hpp>
#include <legate_dataframe/core/table.hpp>

namespace legate::dataframe {

namespace {
thread_local void* g_stream = nullptr;

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw HipError(LDF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

void set_stream(void* stream) { g_stream = stream; }
void* get_stream() { return g_stream; }

namespace detail {

void check_status(int rc)
{
  if (rc == LDF_OK) return;
  std::string msg = ldf_last_error();
  if (rc == LDF_E_INVALID) throw std::invalid_argument(msg);
  throw HipError(rc, msg);
}

// Column buffers come from the library's default (caching) allocator, the same one that makes
// the buffers of the operators' outputs (called with a NULL ldf_allocator), and go back to it:
// one owner per buffer, as the reference's allocator hands its buffers to the column it binds
// (cpp/src/core/column.cu:64-70, cpp/src/core/allocator.cu:60-81).
void* device_alloc(size_t bytes)
{
  void* p = ldf_pool_alloc(bytes ? bytes : 1, g_stream);
  if (p == nullptr) throw HipError(LDF_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  return p;
}

void device_free(void* p)
{
  if (p) ldf_pool_free(p, g_stream);
}

void copy_to_device(void* dst, const void* src, size_t bytes)
{
  if (!bytes) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)g_stream), "hipMemcpyAsync H2D");
  hip_check(hipStreamSynchronize((hipStream_t)g_stream), "hipStreamSynchronize");
}

void copy_to_host(void* dst, const void* src, size_t bytes)
{
  if (!bytes) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)g_stream), "hipMemcpyAsync D2H");
  hip_check(hipStreamSynchronize((hipStream_t)g_stream), "hipStreamSynchronize");
}

struct ColumnStorage {
  DataType type;
  int64_t length = 0;
  std::shared_ptr<void> data, validity, offsets;
  int64_t data_bytes     = 0;  // chars for strings
  int64_t validity_bytes = 0;
  int64_t null_count     = 0;
};

static std::shared_ptr<void> owned(void* p)
{
  return std::shared_ptr<void>(p, [](void* q) { device_free(q); });
}

}  // namespace detail

using detail::ColumnStorage;

// ---------------------------------------------------------------------------------------------
// DataType
// ---------------------------------------------------------------------------------------------
int DataType::width() const
{
  switch (id) {
    case TypeId::BOOL:
    case TypeId::UINT8:
    case TypeId::INT8: return 1;
    case TypeId::UINT16:
    case TypeId::INT16: return 2;
    case TypeId::UINT32:
    case TypeId::INT32:
    case TypeId::FLOAT:
    case TypeId::DATE32: return 4;
    case TypeId::LARGE_STRING: return 0;
    default: return 8;
  }
}

std::string DataType::ToString() const
{
  switch (id) {
    case TypeId::BOOL: return "bool";
    case TypeId::UINT8: return "uint8";
    case TypeId::INT8: return "int8";
    case TypeId::UINT16: return "uint16";
    case TypeId::INT16: return "int16";
    case TypeId::UINT32: return "uint32";
    case TypeId::INT32: return "int32";
    case TypeId::UINT64: return "uint64";
    case TypeId::INT64: return "int64";
    case TypeId::FLOAT: return "float";
    case TypeId::DOUBLE: return "double";
    case TypeId::DATE32: return "date32[day]";
    case TypeId::TIMESTAMP: {
      static const char* u[] = {"s", "ms", "us", "ns"};
      return std::string("timestamp[") + u[unit & 3] + "]";
    }
    case TypeId::LARGE_STRING: return "large_string";
  }
  return "unknown";
}

// ---------------------------------------------------------------------------------------------
// LogicalColumn
// ---------------------------------------------------------------------------------------------
static std::vector<uint8_t> pack_mask(const std::vector<bool>& m)
{
  std::vector<uint8_t> bits((m.size() + 63) / 64 * 8, 0);
  for (size_t i = 0; i < m.size(); ++i)
    if (m[i]) bits[i >> 3] |= (uint8_t)(1u << (i & 7));
  return bits;
}

LogicalColumn LogicalColumn::from_host_buffers(DataType type, int64_t length, const void* data, size_t data_bytes,
                                               const uint8_t* validity_bits, const int64_t* offsets)
{
  if (length < 0) throw std::invalid_argument("negative column length");
  auto s        = std::make_shared<ColumnStorage>();
  s->type       = type;
  s->length     = length;
  s->data_bytes = (int64_t)data_bytes;
  void* d       = detail::device_alloc(data_bytes);
  s->data       = detail::owned(d);
  detail::copy_to_device(d, data, data_bytes);
  if (type.id == TypeId::LARGE_STRING) {
    if (offsets == nullptr) throw std::invalid_argument("string column without offsets");
    void* o    = detail::device_alloc((size_t)(length + 1) * 8);
    s->offsets = detail::owned(o);
    detail::copy_to_device(o, offsets, (size_t)(length + 1) * 8);
  }
  if (validity_bits != nullptr) {
    s->validity_bytes = (length + 63) / 64 * 8;
    std::vector<uint8_t> bits((size_t)s->validity_bytes, 0);
    std::memcpy(bits.data(), validity_bits, (size_t)(length + 7) / 8);
    int64_t nulls = 0;
    for (int64_t i = 0; i < length; ++i) nulls += !((bits[i >> 3] >> (i & 7)) & 1);
    s->null_count = nulls;
    void* v       = detail::device_alloc(bits.size());
    s->validity   = detail::owned(v);
    detail::copy_to_device(v, bits.data(), bits.size());
  }
  LogicalColumn c;
  c.s_ = s;
  return c;
}

LogicalColumn LogicalColumn::from_host(DataType type, int64_t length, const void* data, size_t data_bytes,
                                       const std::vector<bool>& null_mask)
{
  if (!null_mask.empty() && (int64_t)null_mask.size() != length)
    throw std::invalid_argument("null mask length differs from the data length");
  std::vector<uint8_t> bits;
  if (!null_mask.empty()) bits = pack_mask(null_mask);
  return from_host_buffers(type, length, data, data_bytes, null_mask.empty() ? nullptr : bits.data(), nullptr);
}

LogicalColumn::LogicalColumn(const std::vector<std::string>& data, const std::vector<bool>& null_mask)
{
  std::vector<int64_t> offs(1, 0);
  std::string chars;
  for (const auto& s : data) {
    chars += s;
    offs.push_back((int64_t)chars.size());
  }
  if (!null_mask.empty() && null_mask.size() != data.size())
    throw std::invalid_argument("null mask length differs from the data length");
  std::vector<uint8_t> bits;
  if (!null_mask.empty()) bits = pack_mask(null_mask);
  *this = from_host_buffers(DataType(TypeId::LARGE_STRING), (int64_t)data.size(), chars.data(), chars.size(),
                            null_mask.empty() ? nullptr : bits.data(), offs.data());
}

LogicalColumn LogicalColumn::from_device(DataType type, int64_t length, void* data, uint8_t* validity, int64_t* offsets,
                                         int64_t chars_bytes, std::shared_ptr<void> keepalive, int64_t null_count)
{
  auto s        = std::make_shared<ColumnStorage>();
  s->type       = type;
  s->length     = length;
  s->data       = std::shared_ptr<void>(keepalive, data);
  s->data_bytes = type.id == TypeId::LARGE_STRING ? chars_bytes : length * type.width();
  if (validity) {
    s->validity       = std::shared_ptr<void>(keepalive, validity);
    s->validity_bytes = (length + 7) / 8;
  }
  if (offsets) s->offsets = std::shared_ptr<void>(keepalive, offsets);
  s->null_count = null_count;
  if (validity && null_count < 0) {
    std::vector<uint8_t> bits((size_t)(length + 7) / 8);
    detail::copy_to_host(bits.data(), validity, bits.size());
    int64_t nulls = 0;
    for (int64_t i = 0; i < length; ++i) nulls += !((bits[i >> 3] >> (i & 7)) & 1);
    s->null_count = nulls;
  }
  LogicalColumn c;
  c.s_ = s;
  return c;
}

LogicalColumn LogicalColumn::adopt(const ldf_column& c, DataType type)
{
  auto s            = std::make_shared<ColumnStorage>();
  s->type           = type;
  s->length         = c.length;
  s->data           = detail::owned(c.data);
  s->data_bytes     = c.data_bytes;
  s->validity       = c.validity ? detail::owned(c.validity) : nullptr;
  s->validity_bytes = c.validity_bytes;
  s->offsets        = (type.id == TypeId::LARGE_STRING && c.offsets) ? detail::owned(c.offsets) : nullptr;
  s->null_count     = c.validity ? c.null_count : 0;
  LogicalColumn col;
  col.s_ = s;
  return col;
}

LogicalColumn LogicalColumn::empty_like(DataType type, bool nullable)
{
  std::vector<int64_t> offs(1, 0);
  uint8_t zero = 0;
  return from_host_buffers(type, 0, &zero, 0, nullable ? &zero : nullptr, offs.data());
}

DataType LogicalColumn::type() const
{
  if (!s_) throw std::runtime_error("unbound column");
  return s_->type;
}
bool LogicalColumn::nullable() const { return s_ && s_->validity != nullptr; }
size_t LogicalColumn::num_rows() const { return s_ ? (size_t)s_->length : 0; }
int64_t LogicalColumn::null_count() const { return s_ ? s_->null_count : 0; }

ldf_column_view LogicalColumn::view() const
{
  if (!s_) throw std::runtime_error("unbound column");
  ldf_column_view v;
  v.type_id    = (int32_t)s_->type.id;
  v.reserved   = 0;
  v.length     = s_->length;
  v.data       = s_->data.get();
  v.validity   = static_cast<const uint8_t*>(s_->validity.get());
  v.offsets    = static_cast<const int64_t*>(s_->offsets.get());
  v.null_count = s_->null_count;
  return v;
}

std::vector<uint8_t> LogicalColumn::data_to_host() const
{
  std::vector<uint8_t> out((size_t)s_->data_bytes);
  detail::copy_to_host(out.data(), s_->data.get(), out.size());
  return out;
}

std::vector<uint8_t> LogicalColumn::validity_to_host() const
{
  if (!s_->validity) return {};
  std::vector<uint8_t> out((size_t)(s_->length + 7) / 8);
  detail::copy_to_host(out.data(), s_->validity.get(), out.size());
  return out;
}

std::vector<int64_t> LogicalColumn::offsets_to_host() const
{
  if (!s_->offsets) return {};
  std::vector<int64_t> out((size_t)s_->length + 1);
  detail::copy_to_host(out.data(), s_->offsets.get(), out.size() * 8);
  return out;
}

std::vector<bool> LogicalColumn::null_mask() const
{
  std::vector<bool> m(num_rows(), true);
  auto bits = validity_to_host();
  if (!bits.empty())
    for (size_t i = 0; i < m.size(); ++i) m[i] = (bits[i >> 3] >> (i & 7)) & 1;
  return m;
}

std::vector<std::string> LogicalColumn::to_strings() const
{
  if (type().id != TypeId::LARGE_STRING) throw std::invalid_argument("not a string column");
  auto chars = data_to_host();
  auto offs  = offsets_to_host();
  std::vector<std::string> out(num_rows());
  for (size_t i = 0; i < out.size(); ++i)
    out[i] = std::string(reinterpret_cast<const char*>(chars.data()) + offs[i], (size_t)(offs[i + 1] - offs[i]));
  return out;
}

std::string LogicalColumn::repr(size_t max_rows) const
{
  std::ostringstream os;
  os << "LogicalColumn(" << type().ToString() << ", rows=" << num_rows() << ", nullable=" << nullable() << ")";
  (void)max_rows;
  return os.str();
}

// ---------------------------------------------------------------------------------------------
// LogicalTable
// ---------------------------------------------------------------------------------------------
LogicalTable::LogicalTable(std::vector<LogicalColumn> columns, std::map<std::string, size_t> column_names)
  : columns_(std::move(columns)), column_names_(std::move(column_names))
{
  if (columns_.size() != column_names_.size())
    throw std::invalid_argument("LogicalTable: number of columns and column names differ (names must be unique)");
  for (auto& c : columns_)
    if (c.num_rows() != columns_.at(0).num_rows())
      throw std::invalid_argument("LogicalTable: all columns must have the same number of rows");
}

static std::map<std::string, size_t> name_map(const std::vector<std::string>& names)
{
  std::map<std::string, size_t> m;
  for (size_t i = 0; i < names.size(); ++i)
    if (!m.emplace(names[i], i).second) throw std::invalid_argument("column names must be unique: " + names[i]);
  return m;
}

LogicalTable::LogicalTable(std::vector<LogicalColumn> columns, const std::vector<std::string>& column_names)
  : LogicalTable(std::move(columns), name_map(column_names))
{
}

LogicalTable LogicalTable::empty_like(const LogicalTable& other)
{
  std::vector<LogicalColumn> cols;
  for (auto& c : other.columns_) cols.push_back(LogicalColumn::empty_like(c));
  return LogicalTable(std::move(cols), other.column_names_);
}

size_t LogicalTable::num_rows() const { return columns_.empty() ? 0 : columns_.at(0).num_rows(); }

std::vector<std::string> LogicalTable::get_column_name_vector() const
{
  std::vector<std::string> ret(column_names_.size());
  for (const auto& [name, idx] : column_names_) ret.at(idx) = name;
  return ret;
}

LogicalTable LogicalTable::select(const std::vector<size_t>& column_indices) const
{
  auto names = get_column_name_vector();
  std::vector<LogicalColumn> cols;
  std::vector<std::string> sel;
  for (auto i : column_indices) {
    cols.push_back(columns_.at(i));
    sel.push_back(names.at(i));
  }
  return LogicalTable(std::move(cols), sel);
}

LogicalTable LogicalTable::select(const std::vector<std::string>& column_names) const
{
  std::vector<size_t> idx;
  for (auto& n : column_names) idx.push_back(column_names_.at(n));
  return select(idx);
}

ldf_table_view LogicalTable::view(std::vector<ldf_column_view>& storage) const
{
  storage.clear();
  for (auto& c : columns_) storage.push_back(c.view());
  ldf_table_view v;
  v.ncols = (int32_t)storage.size();
  v.cols  = storage.data();
  return v;
}

std::string LogicalTable::repr(size_t max_rows) const
{
  std::ostringstream os;
  os << "LogicalTable(rows=" << num_rows() << ")\n";
  auto names = get_column_name_vector();
  for (size_t i = 0; i < columns_.size(); ++i) os << "  " << names[i] << ": " << columns_[i].repr(max_rows) << "\n";
  return os.str();
}

}  // namespace legate::dataframe
