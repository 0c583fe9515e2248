"""GPU distinct parity: ldf_distinct (through the C++ API and the C ABI) keeps one row per distinct
key tuple -- the key's first row, rows in input order -- with every column.  Bar: bit-exact against
the reference's own test case (python/tests/test_stream_compaction.py:97-121, compared sorted as
the reference does and also in input order) and against a first-occurrence scan in Python (a dict
over key tuples) on random tables with nulls, strings, composite keys and floating-point keys
(as cudf::distinct with nan_equality::ALL_EQUAL: every NaN one key, -0.0 equal to 0.0; NaNs with
other payload bits included); at 50M rows, against torch.unique's first index of each key."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pytest

from ldf_testing import bits_of as _bits
from ldf_testing import first_rows

pytestmark = pytest.mark.gpu


def gpu_distinct(t: pa.Table, keys, high_cardinality=False) -> pa.Table:
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.stream_compaction import distinct

    return distinct(LogicalTable.from_arrow(t), keys, high_cardinality=high_cardinality).to_arrow()


def assert_rows_equal(got: pa.Table, exp: pa.Table):
    assert got.column_names == exp.column_names
    assert got.num_rows == exp.num_rows
    for name in exp.column_names:
        g = [_bits(v) for v in got.column(name).to_pylist()]
        e = [_bits(v) for v in exp.column(name).to_pylist()]
        assert g == e, name


@pytest.mark.parametrize("keys, high_cardinality",
                         [(["b", "a"], True), (["a", "b"], False), (["b", "c"], False), (None, False)])
def test_distinct_reference_case(gpu, keys, high_cardinality):
    df = pa.table({"a": [1, 1, 2, 3, 4], "b": [2, 2, 4, 4, None], "c": [10, 10, 20, 30, 40]})
    res = gpu_distinct(df, keys, high_cardinality)
    expect = pa.table({"a": [1, 2, 3, 4], "b": [2, 4, 4, None], "c": [10, 20, 30, 40]})
    sort_by = [("a", "ascending"), ("b", "ascending"), ("c", "ascending")]
    assert_rows_equal(res.sort_by(sort_by), expect.sort_by(sort_by))
    assert_rows_equal(res, expect)  # first rows, in input order


def _random_table(kind: str, n: int, seed: int) -> tuple[pa.Table, list]:
    rng = np.random.default_rng(seed)
    payload = {"p_f": pa.array(rng.uniform(-1, 1, n), mask=rng.random(n) < 0.05),
               "p_s": pa.array([f"s{x}" for x in rng.integers(0, 1000, n)], mask=rng.random(n) < 0.05)}
    if kind == "int64_nulls":
        t = pa.table({"k": pa.array(rng.integers(0, 50_000, n), mask=rng.random(n) < 0.02), **payload})
        return t, ["k"]
    if kind == "int32_string":
        words = np.array(["", "a", "bb", "ccc", "a much longer string value", "dd"])
        t = pa.table({"k1": pa.array(rng.integers(0, 300, n).astype(np.int32), mask=rng.random(n) < 0.03),
                      "k2": pa.array(list(words[rng.integers(0, len(words), n)]), mask=rng.random(n) < 0.03),
                      **payload})
        return t, ["k1", "k2"]
    if kind in ("float64_special", "float32_special"):
        vals = np.array([0.0, -0.0, np.nan, -np.nan, 1.5, -2.25, np.inf, -np.inf, 3.0])
        dt = np.float64 if kind == "float64_special" else np.float32
        kv = vals[rng.integers(0, len(vals), n)].astype(dt)
        if dt == np.float64:  # a NaN with other payload bits
            kv[rng.integers(0, n, 50)] = np.array([0x7FF0000000000123], np.uint64).view(np.float64)[0]
        t = pa.table({"k": pa.array(kv, mask=rng.random(n) < 0.05), **payload})
        return t, ["k"]
    if kind == "int8_all_columns":
        t = pa.table({"k": pa.array(rng.integers(-128, 128, n).astype(np.int8)),
                      "q": pa.array(rng.integers(0, 3, n).astype(np.int16))})
        return t, ["k", "q"]
    if kind == "all_distinct":
        t = pa.table({"k": pa.array(rng.permutation(n).astype(np.int64)), **payload})
        return t, ["k"]
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["int64_nulls", "int32_string", "float64_special", "float32_special",
                                  "int8_all_columns", "all_distinct"])
def test_distinct_matches_first_occurrence(gpu, kind):
    n = 200_000
    t, keys = _random_table(kind, n, seed=5)
    got = gpu_distinct(t, keys)
    exp = t.take(pa.array(first_rows(t, keys), pa.int64()))
    assert_rows_equal(got, exp.cast(got.schema))


def test_distinct_empty_and_single_row(gpu):
    t = pa.table({"k": pa.array([], pa.int64()), "s": pa.array([], pa.large_utf8())})
    got = gpu_distinct(t, ["k"])
    assert got.num_rows == 0 and got.column_names == ["k", "s"]
    t = pa.table({"k": pa.array([None], pa.int64()), "s": pa.array(["x"], pa.large_utf8())})
    assert_rows_equal(gpu_distinct(t, ["k"]), t)


def test_distinct_unknown_key_raises(gpu):
    t = pa.table({"k": pa.array([1, 2], pa.int64())})
    with pytest.raises(Exception):
        gpu_distinct(t, ["nope"])


def test_distinct_50M_rows_against_torch_unique(gpu):
    """50M int64 keys over 10M values with a float payload, through the C ABI on device memory:
    the kept rows are each key's first row (scatter-min of the row index over torch.unique's
    inverse), in row order."""
    import torch

    from legate_dataframe import _capi as C

    n = 50_000_000
    g = torch.Generator(device="cuda").manual_seed(11)
    k = torch.randint(0, 10_000_000, (n,), device="cuda", generator=g, dtype=torch.int64)
    v = torch.rand(n, device="cuda", generator=g, dtype=torch.float64)
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    out = C.distinct([dc(k, C.INT64), dc(v, C.DOUBLE)], [0])
    m = out[0].length
    uniq, inv = torch.unique(k, return_inverse=True)
    first = torch.full((uniq.numel(),), n, device="cuda", dtype=torch.int64)
    first.scatter_reduce_(0, inv, torch.arange(n, device="cuda", dtype=torch.int64), reduce="amin")
    first, _ = torch.sort(first)
    assert m == uniq.numel()
    assert torch.equal(out[0].data.view(torch.int64)[:m], k[first])
    assert torch.equal(out[1].data.view(torch.float64)[:m], v[first])


@pytest.mark.parametrize("fooled", [False, True])
def test_distinct_small_set_overflow_falls_back(gpu, fooled):
    """2^21 rows, every key distinct: the cardinality sample sees no repeat and the call uses the
    full-size set at once.  fooled: the sampled rows (every 8th) all hold one key, so the sample
    predicts few keys; the first set (n/4 slots) cannot hold the ~1.8M keys, rows hit the probe limit
    and the call runs again with the full-size set.  Either way each key's first row is kept, in
    order."""
    import torch

    from legate_dataframe import _capi as C

    n = 1 << 21
    g = torch.Generator(device="cuda").manual_seed(13)
    k = torch.randperm(n, device="cuda", generator=g).to(torch.int64) * 7 - 3
    if fooled:
        k[::8] = -1
    out = C.distinct([C.DeviceColumn(C.INT64, n, k.view(torch.uint8))], [0])
    keep = torch.ones(n, dtype=torch.bool, device="cuda")
    if fooled:
        keep[8::8] = False  # only the first -1 (row 0) is kept
    exp = k[keep]
    assert out[0].length == exp.numel()
    assert torch.equal(out[0].data.view(torch.int64)[:exp.numel()], exp)


def test_distinct_many_key_columns_and_repeated_keys(gpu):
    """keys=None on a 12-column table (every column a key: more than fit in a kernel's argument
    block, so the key descriptors live in device memory) and a key list naming a column twice."""
    rng = np.random.default_rng(21)
    n = 30_000
    cols = {f"c{i}": rng.integers(0, 3 if i < 6 else 2, n).astype(np.int8 if i % 2 else np.int32) for i in range(12)}
    cols["c3"] = pa.array(cols["c3"], mask=rng.random(n) < 0.1)
    cols["s"] = [f"s{v}" for v in rng.integers(0, 2, n)]
    t = pa.table(cols)
    got = gpu_distinct(t, None)
    exp = t.take(pa.array(first_rows(t, t.column_names), pa.int64()))
    assert_rows_equal(got, exp.cast(got.schema))
    got = gpu_distinct(t, ["c1", "c0", "c1", "s"])
    exp = t.take(pa.array(first_rows(t, ["c1", "c0", "s"]), pa.int64()))
    assert_rows_equal(got, exp.cast(got.schema))
