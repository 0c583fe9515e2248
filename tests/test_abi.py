"""The C ABI library loads and exports every entry point include/ldf_hip.h declares; argument
validation answers with status codes before any device work (no GPU needed)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "ldf_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ldf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from legate_dataframe import _capi

    lib = _capi.load()
    names = header_functions()
    assert "ldf_join" in names and "ldf_groupby" in names and "ldf_hash_partition" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_capi.EXPORTED)
    assert lib.ldf_version().decode().startswith("ldf-hip")


def test_invalid_arguments_return_status_without_device():
    from legate_dataframe import _capi as C

    L = C.load()
    # NULL tables
    rc = L.ldf_join(None, None, None, None, 1, 0, 1, None, 0, None, 0, None, None, None)
    assert rc == -1 and "NULL" in L.ldf_last_error().decode()
    # bad join type / SEMI with rhs outputs
    col = C.ColumnView(C.INT64, 0, 0, None, None, None, 0)
    lv, _a = C.make_table_view([col])
    rv, _b = C.make_table_view([col])
    outc = (C.Column * 2)()
    out = C.TableOut(2, outc)
    rc = L.ldf_join(ctypes.byref(lv), ctypes.byref(rv), C.int32_array([0]), C.int32_array([0]), 1, 9, 1,
                    C.int32_array([0]), 1, C.int32_array([0]), 1, None, None, ctypes.byref(out))
    assert rc == -1 and "JoinType" in L.ldf_last_error().decode()
    rc = L.ldf_join(ctypes.byref(lv), ctypes.byref(rv), C.int32_array([0]), C.int32_array([0]), 1, 3, 1,
                    C.int32_array([0]), 1, C.int32_array([0]), 1, None, None, ctypes.byref(out))
    assert rc == -1 and "SEMI" in L.ldf_last_error().decode()
    # key column types must match
    col32 = C.ColumnView(C.INT32, 0, 0, None, None, None, 0)
    rv2, _c = C.make_table_view([col32])
    rc = L.ldf_join(ctypes.byref(lv), ctypes.byref(rv2), C.int32_array([0]), C.int32_array([0]), 1, 0, 1,
                    C.int32_array([0]), 1, C.int32_array([0]), 1, None, None, ctypes.byref(out))
    assert rc == -1 and "types differ" in L.ldf_last_error().decode()
    # groupby without keys
    tv, _d = C.make_table_view([col])
    gout = C.TableOut(2, outc)
    rc = L.ldf_groupby(ctypes.byref(tv), C.int32_array([]), 0, None, 0, None, None, ctypes.byref(gout))
    assert rc == -1 and "key" in L.ldf_last_error().decode()
    # too many partitions
    offs = (ctypes.c_int64 * 5000)()
    rc = L.ldf_hash_partition(ctypes.byref(tv), C.int32_array([0]), 1, 4096, 0, None, None, ctypes.byref(gout), offs)
    assert rc == -4
    # distinct: no keys, a key index out of range, too few output slots
    rc = L.ldf_distinct(ctypes.byref(tv), C.int32_array([]), 0, None, None, ctypes.byref(gout))
    assert rc == -1 and "key" in L.ldf_last_error().decode()
    rc = L.ldf_distinct(ctypes.byref(tv), C.int32_array([3]), 1, None, None, ctypes.byref(gout))
    assert rc == -1 and "out of range" in L.ldf_last_error().decode()
    rc = L.ldf_distinct(ctypes.byref(tv), C.int32_array([0]), 1, None, None, ctypes.byref(C.TableOut(0, outc)))
    assert rc == -1 and "slots" in L.ldf_last_error().decode()


def test_python_api_mirrors_reference_signatures():
    import inspect

    from legate_dataframe.lib.groupby_aggregation import groupby_aggregation
    from legate_dataframe.lib.join import BroadcastInput, JoinType, join

    assert [m.name for m in JoinType] == ["INNER", "LEFT", "FULL", "SEMI", "ANTI"]
    assert [int(m) for m in JoinType] == [0, 1, 2, 3, 4]
    assert [m.name for m in BroadcastInput] == ["AUTO", "LEFT", "RIGHT"]
    p = inspect.signature(join).parameters
    assert list(p) == ["lhs", "rhs", "lhs_keys", "rhs_keys", "join_type", "lhs_out_columns", "rhs_out_columns",
                       "nulls_equal", "broadcast"]
    assert p["nulls_equal"].default is True
    assert list(inspect.signature(groupby_aggregation).parameters) == ["table", "keys", "column_aggregations"]
    from legate_dataframe.lib.stream_compaction import distinct

    p = inspect.signature(distinct).parameters
    assert list(p) == ["tbl", "keys", "high_cardinality"]
    assert p["keys"].default is None and p["high_cardinality"].default is False
    assert p["high_cardinality"].kind is inspect.Parameter.KEYWORD_ONLY


def test_groupby_python_validation_without_device():
    import pyarrow.compute as pc

    from legate_dataframe.lib.groupby_aggregation import _count_mode, groupby_aggregation

    assert _count_mode(pc.CountOptions(mode="all")) == 2
    assert _count_mode(pc.CountOptions(mode="only_null")) == 1
    assert _count_mode(pc.CountOptions()) == 0
    with pytest.raises(ValueError):
        groupby_aggregation(None, ["k"], [("x", "count_all", "n")])
    with pytest.raises(TypeError):
        groupby_aggregation(None, ["k"], [("x", ("count", "not-options"), "n")])
