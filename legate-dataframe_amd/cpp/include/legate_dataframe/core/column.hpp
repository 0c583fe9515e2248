/*
 * LogicalColumn: a column resident in the HBM of the current HIP device.
 *
 * Mirrors legate::dataframe::LogicalColumn (cpp/include/legate_dataframe/core/column.hpp:58-447
 * in the reference) without Legate: the column owns (or borrows, with a keep-alive) device
 * buffers in the Arrow layout of include/ldf_hip.h -- values, an LSB validity bitmap, and
 * int64 offsets + chars for strings (utf8 is normalised to large_utf8 as the reference's
 * normalize_arrow_type does, cpp/src/utils.cpp:141-146).
 */
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include <ldf_hip.h>

namespace legate::dataframe {

/* Arrow type ids (arrow::Type::type values) of the supported column types. */
enum class TypeId : int32_t {
  BOOL         = LDF_TYPE_BOOL,
  UINT8        = LDF_TYPE_UINT8,
  INT8         = LDF_TYPE_INT8,
  UINT16       = LDF_TYPE_UINT16,
  INT16        = LDF_TYPE_INT16,
  UINT32       = LDF_TYPE_UINT32,
  INT32        = LDF_TYPE_INT32,
  UINT64       = LDF_TYPE_UINT64,
  INT64        = LDF_TYPE_INT64,
  FLOAT        = LDF_TYPE_FLOAT,
  DOUBLE       = LDF_TYPE_DOUBLE,
  DATE32       = LDF_TYPE_DATE32,
  TIMESTAMP    = LDF_TYPE_TIMESTAMP,
  LARGE_STRING = LDF_TYPE_LARGE_STRING
};

/* Stand-in for arrow::DataType: the id plus the timestamp unit (0 s, 1 ms, 2 us, 3 ns). */
struct DataType {
  TypeId id    = TypeId::INT64;
  int32_t unit = 0;
  DataType() = default;
  DataType(TypeId i, int32_t u = 0) : id(i), unit(u) {}
  [[nodiscard]] int width() const;  // bytes per value, 0 for strings
  [[nodiscard]] std::string ToString() const;
  bool operator==(const DataType& o) const { return id == o.id && unit == o.unit; }
  bool operator!=(const DataType& o) const { return !(*this == o); }
};

template <typename T>
constexpr TypeId type_id_of()
{
  if constexpr (std::is_same_v<T, bool>) return TypeId::BOOL;
  else if constexpr (std::is_same_v<T, uint8_t>) return TypeId::UINT8;
  else if constexpr (std::is_same_v<T, int8_t>) return TypeId::INT8;
  else if constexpr (std::is_same_v<T, uint16_t>) return TypeId::UINT16;
  else if constexpr (std::is_same_v<T, int16_t>) return TypeId::INT16;
  else if constexpr (std::is_same_v<T, uint32_t>) return TypeId::UINT32;
  else if constexpr (std::is_same_v<T, int32_t>) return TypeId::INT32;
  else if constexpr (std::is_same_v<T, uint64_t>) return TypeId::UINT64;
  else if constexpr (std::is_same_v<T, int64_t>) return TypeId::INT64;
  else if constexpr (std::is_same_v<T, float>) return TypeId::FLOAT;
  else if constexpr (std::is_same_v<T, double>) return TypeId::DOUBLE;
  else static_assert(sizeof(T) == 0, "unsupported element type");
}

namespace detail {
struct ColumnStorage;
}

class LogicalColumn {
 public:
  LogicalColumn() = default;

  /* Convenience constructors for tests (the reference's column.hpp:112-190): host values and an
   * optional validity mask (true = valid). */
  template <typename T, typename = std::enable_if_t<std::is_arithmetic_v<T>>>
  LogicalColumn(const std::vector<T>& data, const std::vector<bool>& null_mask = {})
  {
    if constexpr (std::is_same_v<T, bool>) {
      std::vector<uint8_t> bytes(data.begin(), data.end());
      *this = from_host(DataType(TypeId::BOOL), (int64_t)data.size(), bytes.data(), bytes.size(), null_mask);
    } else {
      *this = from_host(DataType(type_id_of<T>()), (int64_t)data.size(), data.data(), data.size() * sizeof(T),
                        null_mask);
    }
  }
  LogicalColumn(const std::vector<std::string>& data, const std::vector<bool>& null_mask = {});

  /* Copies host buffers to the device.  `validity_bits` is an LSB bitmap or null. */
  static LogicalColumn from_host_buffers(DataType type, int64_t length, const void* data, size_t data_bytes,
                                         const uint8_t* validity_bits, const int64_t* offsets);
  static LogicalColumn from_host(DataType type, int64_t length, const void* data, size_t data_bytes,
                                 const std::vector<bool>& null_mask);
  /* Wraps device buffers without copying; `keepalive` owns them. */
  static LogicalColumn from_device(DataType type, int64_t length, void* data, uint8_t* validity, int64_t* offsets,
                                   int64_t chars_bytes, std::shared_ptr<void> keepalive, int64_t null_count = -1);
  /* Adopts a column the C ABI returned (buffers from the default allocator). */
  static LogicalColumn adopt(const ldf_column& c, DataType type);
  /* An empty (zero-row) column; the reference's empty_like (column.hpp:246). */
  static LogicalColumn empty_like(DataType type, bool nullable);
  static LogicalColumn empty_like(const LogicalColumn& other) { return empty_like(other.type(), other.nullable()); }

  [[nodiscard]] DataType type() const;
  [[nodiscard]] DataType arrow_type() const { return type(); }
  [[nodiscard]] bool nullable() const;
  [[nodiscard]] size_t num_rows() const;
  [[nodiscard]] int64_t null_count() const;
  [[nodiscard]] bool unbound() const { return s_ == nullptr; }

  /* C ABI view of the device buffers. */
  [[nodiscard]] ldf_column_view view() const;

  /* Host copies. */
  [[nodiscard]] std::vector<uint8_t> data_to_host() const;
  [[nodiscard]] std::vector<uint8_t> validity_to_host() const;  // empty when not nullable
  [[nodiscard]] std::vector<int64_t> offsets_to_host() const;   // strings only
  [[nodiscard]] std::vector<bool> null_mask() const;            // true = valid
  template <typename T>
  [[nodiscard]] std::vector<T> to_vector() const
  {
    auto raw = data_to_host();
    std::vector<T> out(num_rows());
    if constexpr (std::is_same_v<T, bool>) {
      for (size_t i = 0; i < out.size(); ++i) out[i] = raw[i] != 0;
    } else {
      if (sizeof(T) != (size_t)type().width()) throw std::invalid_argument("to_vector: element size mismatch");
      std::copy(reinterpret_cast<const T*>(raw.data()), reinterpret_cast<const T*>(raw.data()) + out.size(),
                out.begin());
    }
    return out;
  }
  [[nodiscard]] std::vector<std::string> to_strings() const;
  [[nodiscard]] std::string repr(size_t max_rows = 30) const;

 private:
  std::shared_ptr<const detail::ColumnStorage> s_;
};

}  // namespace legate::dataframe
