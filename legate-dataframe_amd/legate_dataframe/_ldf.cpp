// pybind11 bindings of the C++ host API (the role the reference's Cython modules play:
// python/legate_dataframe/lib/{join,groupby_aggregation}.pyx, lib/core/{table,column}.pyx).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <legate_dataframe/core/runtime.hpp>
#include <legate_dataframe/groupby_aggregation.hpp>
#include <legate_dataframe/hash_partition.hpp>
#include <legate_dataframe/join.hpp>
#include <legate_dataframe/stream_compaction.hpp>

namespace py = pybind11;
using namespace legate::dataframe;

namespace {

LogicalColumn column_from_host(int32_t type_id, int32_t unit, int64_t length, py::array_t<uint8_t> data,
                               py::object validity, py::object offsets)
{
  auto d = data.request();
  std::vector<uint8_t> vbits;
  const uint8_t* vp = nullptr;
  if (!validity.is_none()) {
    auto va = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>(validity);
    vbits.assign(va.data(), va.data() + va.size());
    vbits.resize((size_t)(length + 7) / 8, 0);
    vp = vbits.data();
  }
  std::vector<int64_t> offs;
  if (!offsets.is_none()) {
    auto oa = py::array_t<int64_t, py::array::c_style | py::array::forcecast>(offsets);
    offs.assign(oa.data(), oa.data() + oa.size());
  }
  py::gil_scoped_release nogil;
  return LogicalColumn::from_host_buffers(DataType((TypeId)type_id, unit), length, d.ptr, (size_t)d.size, vp,
                                          offs.empty() ? nullptr : offs.data());
}

LogicalColumn column_from_device(int32_t type_id, int32_t unit, int64_t length, uintptr_t data, uintptr_t validity,
                                 uintptr_t offsets, int64_t chars_bytes, py::object keepalive, int64_t null_count)
{
  // the Python owner(s) of the device buffers stay alive while the column does
  auto* holder = new py::object(std::move(keepalive));
  std::shared_ptr<void> keep(holder, [](void* p) {
    py::gil_scoped_acquire gil;
    delete static_cast<py::object*>(p);
  });
  return LogicalColumn::from_device(DataType((TypeId)type_id, unit), length, reinterpret_cast<void*>(data),
                                    reinterpret_cast<uint8_t*>(validity), reinterpret_cast<int64_t*>(offsets),
                                    chars_bytes, keep, null_count);
}

py::tuple column_to_host(const LogicalColumn& c)
{
  std::vector<uint8_t> data, valid;
  std::vector<int64_t> offs;
  {
    py::gil_scoped_release nogil;
    data  = c.data_to_host();
    valid = c.validity_to_host();
    if (c.type().id == TypeId::LARGE_STRING) offs = c.offsets_to_host();
  }
  py::array_t<uint8_t> d(data.size());
  std::copy(data.begin(), data.end(), d.mutable_data());
  py::object v = py::none();
  if (c.nullable()) {
    py::array_t<uint8_t> va(valid.size());
    std::copy(valid.begin(), valid.end(), va.mutable_data());
    v = va;
  }
  py::object o = py::none();
  if (!offs.empty()) {
    py::array_t<int64_t> oa(offs.size());
    std::copy(offs.begin(), offs.end(), oa.mutable_data());
    o = oa;
  }
  return py::make_tuple(d, v, o);
}

std::vector<ColumnAggregation> to_aggs(const std::vector<std::tuple<std::string, std::string, int32_t, std::string>>& in)
{
  std::vector<ColumnAggregation> out;
  for (auto& [col, kind, mode, name] : in) {
    std::optional<std::shared_ptr<FunctionOptions>> opt;
    if (mode >= 0) opt = std::make_shared<CountOptions>((CountOptions::CountMode)mode);
    out.emplace_back(col, kind, opt, name);
  }
  return out;
}

}  // namespace

PYBIND11_MODULE(_ldf, m)
{
  m.doc() = "MI355X hash join / groupby / hash partition (C++ host API bindings)";

  py::register_exception<HipError>(m, "HipError", PyExc_RuntimeError);

  py::class_<LogicalColumn>(m, "LogicalColumn")
    .def_static("from_host", &column_from_host, py::arg("type_id"), py::arg("unit"), py::arg("length"), py::arg("data"),
                py::arg("validity"), py::arg("offsets"))
    .def_static("from_device", &column_from_device, py::arg("type_id"), py::arg("unit"), py::arg("length"),
                py::arg("data"), py::arg("validity"), py::arg("offsets"), py::arg("chars_bytes"), py::arg("keepalive"),
                py::arg("null_count") = -1)
    .def("to_host", &column_to_host)
    .def("num_rows", &LogicalColumn::num_rows)
    .def("nullable", &LogicalColumn::nullable)
    .def("null_count", &LogicalColumn::null_count)
    .def("type_id", [](const LogicalColumn& c) { return (int32_t)c.type().id; })
    .def("type_unit", [](const LogicalColumn& c) { return c.type().unit; })
    .def("type_name", [](const LogicalColumn& c) { return c.type().ToString(); })
    .def("device_pointers",
         [](const LogicalColumn& c) {
           auto v = c.view();
           return py::make_tuple((uintptr_t)v.data, (uintptr_t)v.validity, (uintptr_t)v.offsets);
         })
    .def("__repr__", [](const LogicalColumn& c) { return c.repr(); });

  py::class_<LogicalTable>(m, "LogicalTable")
    .def(py::init<std::vector<LogicalColumn>, const std::vector<std::string>&>())
    .def("num_columns", &LogicalTable::num_columns)
    .def("num_rows", &LogicalTable::num_rows)
    .def("get_column", py::overload_cast<size_t>(&LogicalTable::get_column, py::const_))
    .def("get_column_by_name", py::overload_cast<const std::string&>(&LogicalTable::get_column, py::const_))
    .def("get_columns", &LogicalTable::get_columns)
    .def("get_column_names", &LogicalTable::get_column_name_vector)
    .def("select", py::overload_cast<const std::vector<std::string>&>(&LogicalTable::select, py::const_))
    .def("__repr__", [](const LogicalTable& t) { return t.repr(); });

  m.def(
    "join",
    [](const LogicalTable& lhs, const LogicalTable& rhs, const std::vector<std::string>& lhs_keys,
       const std::vector<std::string>& rhs_keys, int32_t join_type, const std::vector<std::string>& lhs_out,
       const std::vector<std::string>& rhs_out, bool nulls_equal, int32_t broadcast) {
      py::gil_scoped_release nogil;
      return join(lhs, rhs, lhs_keys, rhs_keys, (JoinType)join_type, lhs_out, rhs_out, nulls_equal,
                  (BroadcastInput)broadcast);
    },
    py::arg("lhs"), py::arg("rhs"), py::arg("lhs_keys"), py::arg("rhs_keys"), py::arg("join_type"), py::arg("lhs_out"),
    py::arg("rhs_out"), py::arg("nulls_equal"), py::arg("broadcast"));

  m.def(
    "groupby_aggregation",
    [](const LogicalTable& t, const std::vector<std::string>& keys,
       const std::vector<std::tuple<std::string, std::string, int32_t, std::string>>& aggs) {
      auto a = to_aggs(aggs);
      py::gil_scoped_release nogil;
      return groupby_aggregation(t, keys, a);
    },
    py::arg("table"), py::arg("keys"), py::arg("column_aggregations"));

  m.def(
    "hash_partition",
    [](const LogicalTable& t, const std::vector<std::string>& cols, int32_t nparts, uint32_t seed) {
      py::gil_scoped_release nogil;
      return hash_partition(t, cols, nparts, seed);
    },
    py::arg("table"), py::arg("columns"), py::arg("num_partitions"), py::arg("seed") = 0);

  m.def(
    "distinct",
    [](const LogicalTable& t, const std::vector<std::string>& keys, bool high_cardinality) {
      py::gil_scoped_release nogil;
      return distinct(t, keys, high_cardinality);
    },
    py::arg("tbl"), py::arg("keys"), py::arg("high_cardinality") = false);

  m.def("set_stream", [](uintptr_t s) { set_stream(reinterpret_cast<void*>(s)); });
  m.def("get_stream", []() { return reinterpret_cast<uintptr_t>(get_stream()); });
  m.def("version", []() { return std::string(ldf_version()); });
}
