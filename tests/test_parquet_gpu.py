"""Parquet ingest (reference API: python/legate_dataframe/lib/parquet.pyx:44-140): files read on
the host with pyarrow, columns moved to HBM.  Round trips are exact; a join and a groupby over the
read tables equal the oracle; the served form's tables load from Parquet and answer queries exactly
as the plain entry points do on the same tables."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from ldf_testing import assert_groupby_equal, groupby_via_oracle, join_via_oracle, sorted_rows, table_fingerprint

pytestmark = pytest.mark.gpu


def _table(rng, n, tag):
    return pa.table({"k": pa.array(rng.integers(0, n // 3, n), mask=rng.random(n) < 0.05),
                     "x": rng.random(n), "i": rng.integers(-(2**40), 2**40, n).astype(np.int32),
                     "s": pa.array([f"{tag}{v}" for v in rng.integers(0, 99, n)], mask=rng.random(n) < 0.02),
                     "t": pa.array(rng.integers(946684800000, 1893456000000, n), pa.timestamp("ms"))})


def test_parquet_round_trip_glob_and_columns(gpu, tmp_path):
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.parquet import parquet_read, parquet_write

    rng = np.random.default_rng(1)
    a, b = _table(rng, 5000, "a"), _table(rng, 3000, "b")
    pq.write_table(a, tmp_path / "part.0.parquet")
    pq.write_table(b, tmp_path / "part.1.parquet")
    got = parquet_read(str(tmp_path / "*.parquet")).to_arrow()
    exp = pa.concat_tables([a, b])
    assert got.num_rows == exp.num_rows and got.column_names == exp.column_names
    assert sorted_rows(got) == sorted_rows(exp)
    assert got.column("s").type == pa.large_string()
    sub = parquet_read([tmp_path / "part.1.parquet"], columns=["s", "k"]).to_arrow()
    assert sub.column_names == ["s", "k"] and sub.column("k").to_pylist() == b.column("k").to_pylist()
    out = tmp_path / "out"
    parquet_write(LogicalTable.from_arrow(a), out)
    back = parquet_read(out / "part.0.parquet").to_arrow()
    assert back.column("x").to_pylist() == a.column("x").to_pylist()
    assert back.column("s").to_pylist() == a.column("s").to_pylist()


def test_join_and_groupby_over_parquet_tables(gpu, tmp_path):
    from legate_dataframe.lib.groupby_aggregation import groupby_aggregation
    from legate_dataframe.lib.join import JoinType, join
    from legate_dataframe.lib.parquet import parquet_read

    rng = np.random.default_rng(2)
    lhs, rhs = _table(rng, 20000, "l"), _table(rng, 4000, "r")
    pq.write_table(lhs, tmp_path / "lhs.parquet")
    pq.write_table(rhs.rename_columns(["k2", "x2", "i2", "s2", "t2"]), tmp_path / "rhs.parquet")
    L = parquet_read(tmp_path / "lhs.parquet")
    R = parquet_read(tmp_path / "rhs.parquet")
    got = join(L, R, lhs_keys=["k"], rhs_keys=["k2"], join_type=JoinType.LEFT).to_arrow()
    exp = join_via_oracle(L.to_arrow(), R.to_arrow(), ["k"], ["k2"], "left", True)
    assert sorted_rows(got) == sorted_rows(exp)
    aggs = [("x", "sum", "xs"), ("i", "max", "im"), ("s", "count_distinct", "sd")]
    g = groupby_aggregation(L, ["k"], aggs).to_arrow()
    assert_groupby_equal(g, groupby_via_oracle(L.to_arrow(), ["k"], aggs), ["k"], rel=1e-12)


def test_served_tables_from_parquet(gpu, tmp_path):
    from legate_dataframe import _capi as C

    rng = np.random.default_rng(3)
    paths = []
    for j, n in enumerate((300_000, 1_200_000)):
        t = pa.table({"k": rng.integers(0, n, n), "a": rng.integers(-(2**40), 2**40, n), "f": rng.random(n)})
        p = tmp_path / f"t{j}.parquet"
        pq.write_table(t, p, row_group_size=100_000)
        paths.append(p)
    tables = [C.columns_from_parquet(p) for p in paths]
    assert [c.type_id for c in tables[0]] == [C.INT64, C.INT64, C.DOUBLE]
    srv = C.Server(hbm_budget_bytes=20e9, slots=2, max_batch=16)
    for tb in tables:
        srv.add_table(tb)
    aggs = [(1, "sum", None), (2, "sum", None), (None, "count_all", None)]
    queries = []
    for j in range(12):
        t = j % 2
        n = tables[t][0].length
        if j % 3:
            s = int(rng.integers(1000, 200_000))
            queries.append(("join", t, [C.column_from_numpy(rng.integers(0, n, s), C.INT64),
                                        C.column_from_numpy(rng.random(s), C.DOUBLE)]))
        else:
            queries.append(("groupby", t, (64 * int(rng.integers(0, 100)), n)))
    ids = [srv.submit_join(t, a, 0) if k == "join" else srv.submit_groupby(t, a[0], a[1], aggs) for k, t, a in queries]
    for qid, (kind, t, arg) in zip(ids, queries):
        res, _ = srv.wait(qid)
        if kind == "join":
            plain = C.join(arg, tables[t], [0], [0], 0)
        else:
            b, e = arg
            plain = C.groupby([C.DeviceColumn(c.type_id, e - b, c.data[b * C.WIDTH[c.type_id]:e * C.WIDTH[c.type_id]])
                               for c in tables[t]], [0], aggs)
        assert np.array_equal(table_fingerprint(res), table_fingerprint(plain)), kind
        srv.release(qid)
    srv.close()
