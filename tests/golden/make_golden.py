"""Writes the golden fixtures under tests/golden/ (run once; the JSON files are committed).

Expected results come from Apache Arrow's Acero engine (pyarrow 25.0 in this image; the
reference pins Arrow 19.0, conda/recipes/legate-dataframe/conda_build_config.yaml:37-38), which
is what the reference's CPU task variants call:
  * joins: Acero hashjoin (cpp/src/join.cpp:214-227), with the reference's nulls_equal=true
    emulation (coalesce the key with 0/"" and add a validity "_mask" key column,
    cpp/src/join.cpp:117-209);
  * groupby: Acero hash aggregate (cpp/src/groupby_aggregation.cpp:87-91) via
    pyarrow.TableGroupBy, the same call python/tests/test_groupby_aggregation.py:28-53 uses.
Input tables follow the reference tests: python/tests/test_join.py:30-67 (make_param),
cpp/tests/test_join.cpp:269-304 (LeftJoinOnNulls, whose expected table is hard-coded there),
python/tests/test_groupby_aggregation.py:56-161.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc

HERE = Path(__file__).resolve().parent
JT = {"inner": "inner", "left": "left outer", "full": "full outer", "semi": "left semi", "anti": "left anti"}


def col_json(arr: pa.ChunkedArray | pa.Array):
    arr = arr.combine_chunks() if isinstance(arr, pa.ChunkedArray) else arr
    t = arr.type
    vals = arr.to_pylist()
    if pa.types.is_floating(t):
        vals = [None if v is None else (float(v) if not np.isnan(v) else "nan") for v in vals]
    return {"type": str(t), "values": vals}


def table_json(t: pa.Table):
    return {"names": t.column_names, "columns": [col_json(t.column(i)) for i in range(t.num_columns)]}


def arrow_join(lhs: pa.Table, rhs: pa.Table, lk, rk, how: str, nulls_equal: bool) -> pa.Table:
    """The reference's CPU join (Acero), including its null-key emulation."""
    lt, rt = lhs, rhs
    lkeys, rkeys = list(lk), list(rk)
    if nulls_equal:
        lkeys, rkeys = [], []
        for a, b in zip(lk, rk):
            la, rb = lhs.column(a), rhs.column(b)
            need = la.null_count > 0 or rb.null_count > 0
            for tab_name, key, arr in (("l", a, la), ("r", b, rb)):
                t = lt if tab_name == "l" else rt
                keys = lkeys if tab_name == "l" else rkeys
                if arr.null_count > 0:
                    fill = pa.scalar("" if pa.types.is_string(arr.type) or pa.types.is_large_string(arr.type) else 0,
                                     type=arr.type)
                    t = t.append_column(key + "_replaced", pc.coalesce(arr, fill))
                    keys.append(key + "_replaced")
                else:
                    keys.append(key)
                if need:
                    t = t.append_column(key + "_mask", pc.is_valid(arr))
                    keys.append(key + "_mask")
                if tab_name == "l":
                    lt = t
                else:
                    rt = t
    # keep only the original columns on output, positional like the reference's output table
    lcols = lhs.column_names
    rcols = [] if how in ("semi", "anti") else rhs.column_names
    # rename to avoid collisions, as the reference's suffixes do
    lt2 = lt.rename_columns([f"L.{c}" for c in lt.column_names])
    rt2 = rt.rename_columns([f"R.{c}" for c in rt.column_names])
    res = lt2.join(rt2, keys=[f"L.{k}" for k in lkeys], right_keys=[f"R.{k}" for k in rkeys], join_type=JT[how],
                   use_threads=False, coalesce_keys=False)
    out = res.select([f"L.{c}" for c in lcols] + [f"R.{c}" for c in rcols])
    return out.rename_columns(lcols + rcols)


def join_cases():
    rng = np.random.RandomState(7)
    cases = []
    for a_rows in (60, 80):
        for b_rows in (60, 80):
            a = np.arange(a_rows, dtype="int64")
            b = np.arange(b_rows, dtype="int64")
            rng.shuffle(a)
            rng.shuffle(b)
            payload_a = pa.array(np.arange(a_rows), mask=np.array([i % 5 == 3 for i in range(a_rows)]))
            cases.append((f"perm_{a_rows}_{b_rows}", pa.table({"a": a, "payload_a": payload_a}),
                          pa.table({"b": b, "payload_b": np.arange(b.size) * -1}), ["a"], ["b"]))
            cases.append((f"perm_str_{a_rows}_{b_rows}",
                          pa.table({"a": a, "payload_a": [str(i) for i in range(a.size)]}),
                          pa.table({"b": b, "payload_b": [str(i * -1) for i in range(b.size)]}), ["a"], ["b"]))
    cases.append(("dup_keys", pa.table({"a": [1, 2, 3, 4, 5], "payload_a": np.arange(5)}),
                  pa.table({"b": [1, 1, 2, 2, 5, 6], "payload_b": np.arange(6) * -1}), ["a"], ["b"]))
    cases.append(("two_keys_order", pa.table({"a_a": [1, 2, 3, 4, 5], "b_a": [5, 4, 3, 2, 1]}),
                  pa.table({"a_b": [5, 4, 3, 2, 1], "b_b": [1, 2, 3, 4, 5]}), ["a_a", "b_a"], ["b_b", "a_b"]))
    # nullable keys (int32 + string), cpp/tests/test_join.cpp:269-304 tables
    cases.append(("left_join_on_nulls",
                  pa.table({"a": pa.array([3, 1, 0], pa.int32(), mask=np.array([False, True, False])),
                            "b": ["s0", "s1", "s2"], "c": pa.array([0, 1, 2], pa.int32())}),
                  pa.table({"d": pa.array([0, 5, 3, 7], pa.int32(), mask=np.array([False, False, False, True])),
                            "e": ["s1", "s0", "s0", "s1"], "f": pa.array([1, 4, 2, 8], pa.int32())}),
                  ["a", "b"], ["d", "e"]))
    # random nullable int64 keys with duplicates on both sides
    n1, n2 = 300, 200
    k1 = rng.randint(0, 50, n1)
    k2 = rng.randint(0, 50, n2)
    cases.append(("random_nulls",
                  pa.table({"k": pa.array(k1, pa.int64(), mask=rng.rand(n1) < 0.1),
                            "x": pa.array(rng.rand(n1), mask=rng.rand(n1) < 0.2)}),
                  pa.table({"k": pa.array(k2, pa.int64(), mask=rng.rand(n2) < 0.1),
                            "y": pa.array(rng.randint(-1000, 1000, n2), pa.int64())}), ["k"], ["k"]))
    return cases


def groupby_cases():
    cases = []
    basic = [
        (["k1"], pa.table({"k1": ["x", "x", "y", "y", "z"], "d1": [1, 2, 0, 4, 1], "d2": [3, 2, 4, 5, 1]})),
        (["k1", "k2"], pa.table({"k1": ["x", "y", "y", "y", "x"], "d1": [1, 2, 0, 4, 1],
                                 "k2": ["y", "x", "y", "x", "y"], "d2": [3, 2, 4, 5, 1]})),
    ]
    basic_aggs = [[("d1", "sum", "sum")], [("d1", "min", "min"), ("d1", "max", "max")],
                  [("d2", "mean", "mean"), ("d1", "product", "prod")]]
    for i, (keys, t) in enumerate(basic):
        for j, aggs in enumerate(basic_aggs):
            cases.append((f"basic_{i}_{j}", t, keys, aggs))
    kinds = ["sum", "product", "min", "max", "count", "count_all", "mean", "variance", "stddev", "count_distinct",
             ("count_distinct", "all")]
    for value_type in ("int64", "float64", "uint8"):
        for key_type in ("int64", "string", "float32"):
            rng = np.random.RandomState(42)
            n_keys, n = 10, 1000
            if key_type == "int64":
                key_a = pa.array(rng.randint(0, n_keys, n), type=pa.int64())
            elif key_type == "string":
                key_a = pa.array([f"key_{i % n_keys}" for i in range(n)], type=pa.string())
            else:
                key_a = pa.array(rng.randint(0, n_keys, n).astype(np.float32), type=pa.float32())
            key_b = pa.array(rng.randint(0, n_keys, n), type=pa.int64())
            value_a = pa.array(rng.random(n).astype(value_type), type=value_type)
            value_b = pa.array(rng.random(n), type=pa.float64())
            t = pa.Table.from_arrays([key_a, value_a, key_b, value_b], names=["key_a", "value_a", "key_b", "value_b"])
            for kind in kinds:
                kname = kind if isinstance(kind, str) else f"{kind[0]}_{kind[1]}"
                aggs = [("value_a" if kind != "count_all" else None, kind, f"value_a_{kname}"),
                        ("value_b" if kind != "count_all" else None, kind, f"value_b_{kname}")]
                cases.append((f"numeric_{value_type}_{key_type}_{kname}", t, ["key_a", "key_b"], aggs))
    # nulls in keys and values, all count modes
    rng = np.random.RandomState(3)
    n = 500
    t = pa.table({"k": pa.array(rng.randint(0, 20, n), pa.int32(), mask=rng.rand(n) < 0.05),
                  "v": pa.array(rng.randint(-(2**40), 2**40, n), pa.int64(), mask=rng.rand(n) < 0.1),
                  "f": pa.array(rng.uniform(-1, 1, n), mask=rng.rand(n) < 0.1),
                  "u": pa.array(rng.randint(0, 30, n), pa.int32(), mask=rng.rand(n) < 0.1)})
    aggs = [("v", "sum", "s"), ("v", "min", "mn"), ("v", "max", "mx"), ("v", ("count", "only_valid"), "c"),
            ("v", ("count", "only_null"), "cn"), ("v", ("count", "all"), "ca"), ("f", "sum", "fs"), ("f", "mean", "fm"),
            ("u", "count_distinct", "nu"), ("u", ("count_distinct", "all"), "nua"),
            ("u", ("count_distinct", "only_null"), "nun"), (None, "count_all", "n")]
    cases.append(("nulls_all_modes", t, ["k"], aggs))
    return cases


def arrow_groupby(t: pa.Table, keys, aggs) -> pa.Table:
    spec = []
    for col, kind, _ in aggs:
        if kind == "count_all":
            spec.append(([], "count_all"))
        elif isinstance(kind, tuple):
            spec.append((col, kind[0], pc.CountOptions(mode=kind[1])))
        else:
            spec.append((col, kind))
    res = pa.TableGroupBy(t, keys, use_threads=False).aggregate(spec)
    return res.rename_columns(list(keys) + [o for _, _, o in aggs])


def main():
    joins = []
    for name, lhs, rhs, lk, rk in join_cases():
        for how in ("inner", "left", "full", "semi", "anti"):
            for ne in (True, False):
                res = arrow_join(lhs, rhs, lk, rk, how, ne)
                joins.append({"name": name, "how": how, "nulls_equal": ne, "lhs": table_json(lhs),
                              "rhs": table_json(rhs), "lhs_keys": lk, "rhs_keys": rk, "expected": table_json(res)})
    # the reference's hard-coded expectation (cpp/tests/test_join.cpp:283-290): LEFT join on
    # (a,b)=(d,e) with lhs_out a,b,c and rhs_out f
    left_on_nulls = {
        "names": ["a", "b", "c", "f"],
        "columns": [{"type": "int32", "values": [None, 0, 3]}, {"type": "string", "values": ["s1", "s2", "s0"]},
                    {"type": "int32", "values": [1, 2, 0]}, {"type": "int32", "values": [8, None, 2]}],
    }
    groupbys = []
    tables = {}
    for name, t, keys, aggs in groupby_cases():
        res = arrow_groupby(t, keys, aggs)
        tj = table_json(t)
        tid = next((k for k, v in tables.items() if v == tj), None)
        if tid is None:
            tid = f"t{len(tables)}"
            tables[tid] = tj
        groupbys.append({"name": name, "table": tid, "keys": keys,
                         "aggs": [[c, k if isinstance(k, str) else list(k), o] for c, k, o in aggs],
                         "expected": table_json(res)})
    (HERE / "join_cases.json").write_text(json.dumps({"cases": joins, "left_join_on_nulls_expected": left_on_nulls}))
    (HERE / "groupby_cases.json").write_text(json.dumps({"tables": tables, "cases": groupbys}))
    print(len(joins), "join cases,", len(groupbys), "groupby cases")


if __name__ == "__main__":
    main()
