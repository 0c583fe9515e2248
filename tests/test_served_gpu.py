"""Served form (ldf_server_*): each query's result is the plain entry point's result on the same
query (ldf_join of the probe with the table; ldf_groupby of the table rows), whether the query ran
alone or batched with others in a mix of concurrent batches.  Compared as sorted row multisets,
values bit for bit."""
import threading

import numpy as np
import pytest

from legate_dataframe import _capi as C

pytestmark = pytest.mark.gpu

AGGS = [(1, "sum", None), (2, "sum", None), (None, "count_all", None), (2, "mean", None), (1, "min", None),
        (1, "max", None)]


def _col(vals, tid, valid=None):
    return C.column_from_numpy(np.asarray(vals), tid, valid)


def _rows(cols):
    data = []
    for c in cols:
        v, ok = C.column_to_numpy(c)
        if ok is None:
            ok = np.ones(len(v), bool)
        if v.dtype.kind == "f":
            v = v.view(np.int64 if v.dtype.itemsize == 8 else np.int32)
        data.append([int(x) if o else None for x, o in zip(v, ok)])
    return sorted(zip(*data), key=lambda r: tuple((0, 0) if x is None else (1, x) for x in r))


def _tables(rng):
    out = []
    for n, zipf in ((200_000, False), (1_000_000, True), (3_000_000, False)):
        k = (rng.zipf(1.1, n) - 1) % n if zipf else rng.integers(0, n, n)
        out.append([_col(k, C.INT64), _col(rng.integers(-(2**40), 2**40, n), C.INT64),
                    _col(rng.random(n) * 2 - 1, C.DOUBLE)])
    return out


def _queries(rng, tables):
    qs = []
    for j in range(24):
        t = int(rng.integers(0, len(tables)))
        n = tables[t][0].length
        size = int(np.exp(rng.uniform(np.log(1e3), np.log(3e5))))
        if j % 2 == 0:
            pk = rng.integers(0, n, size)
            qs.append(("join", t, [_col(pk, C.INT64), _col(rng.random(size), C.DOUBLE)]))
        else:
            size = min(size, n)
            b = int(rng.integers(0, (n - size) // 64 + 1)) * 64
            qs.append(("groupby", t, (b, min(n, b + size))))
    qs.append(("groupby", 2, (64, 2_500_000)))  # over 2^21 rows: the plain path
    return qs


def _plain(tables, q):
    kind, t, arg = q
    if kind == "join":
        return C.join(arg, tables[t], [0], [0], 0)
    b, e = arg
    sl = [C.DeviceColumn(c.type_id, e - b, c.data[b * C.WIDTH[c.type_id]:e * C.WIDTH[c.type_id]]) for c in tables[t]]
    return C.groupby(sl, [0], AGGS)


def _submit(srv, q):
    kind, t, arg = q
    if kind == "join":
        return srv.submit_join(t, arg, 0)
    return srv.submit_groupby(t, arg[0], arg[1], AGGS)


@pytest.fixture(scope="module")
def served(gpu):
    rng = np.random.default_rng(2024)
    tables = _tables(rng)
    qs = _queries(rng, tables)
    return tables, qs, [_rows(_plain(tables, q)) for q in qs]


def test_served_mix_equals_plain_and_alone(served):
    tables, qs, expected = served
    srv = C.Server(hbm_budget_bytes=64e9, slots=4, max_batch=8)
    tids = [srv.add_table(t) for t in tables]
    assert tids == [0, 1, 2]
    # the mix: two client threads submitting concurrently, results waited in any order
    ids = [None] * len(qs)

    def client(part):
        for j in part:
            ids[j] = _submit(srv, qs[j])

    th = [threading.Thread(target=client, args=(list(range(k, len(qs), 2)),)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    batched = 0
    for j in reversed(range(len(qs))):
        cols, st = srv.wait(ids[j])
        assert _rows(cols) == expected[j], j
        assert st["end_ns"] >= st["start_ns"] >= st["submit_ns"]
        batched = max(batched, st["batch_queries"])
        assert st["plain"] == (1 if j == len(qs) - 1 else 0)
        srv.release(ids[j])
    assert batched > 1  # queries that arrived together ran together
    srv.close()
    # alone: one query at a time
    srv = C.Server(hbm_budget_bytes=64e9, slots=1, max_batch=1)
    for t in tables:
        srv.add_table(t)
    for j, q in enumerate(qs):
        qid = _submit(srv, q)
        cols, st = srv.wait(qid)
        assert st["batch_queries"] == 1
        assert _rows(cols) == expected[j], j
        srv.release(qid)
    srv.close()


def test_served_tight_budget_still_completes(served):
    """A budget that holds the indexes and barely one query: queries run one at a time (the oldest
    runs even when it alone exceeds the budget) and all complete with the right results."""
    tables, qs, expected = served
    srv = C.Server(hbm_budget_bytes=0.4e9, slots=2, max_batch=16)
    for t in tables:
        srv.add_table(t)
    sub = list(range(0, 8))
    ids = [_submit(srv, qs[j]) for j in sub]
    for j, qid in zip(sub, ids):
        cols, st = srv.wait(qid)
        assert _rows(cols) == expected[j], j
        srv.release(qid)
    srv.close()


def test_served_errors(gpu):
    rng = np.random.default_rng(1)
    srv = C.Server(hbm_budget_bytes=8e9, slots=1, max_batch=4)
    t = [_col(rng.integers(0, 100, 1000), C.INT64), _col(rng.random(1000), C.DOUBLE)]
    srv.add_table(t)
    with pytest.raises(ValueError):  # unknown table
        srv.submit_groupby(5, 0, 10, [(1, "sum", None)])
    with pytest.raises(ValueError):  # rows outside the table
        srv.submit_groupby(0, 10, 1001, [(1, "sum", None)])
    qid = srv.submit_groupby(0, 0, 100, [(1, "count_distinct", None)])  # a plain-only kind: alone
    cols, st = srv.wait(qid)
    assert st["plain"] == 1 and cols[0].length == len(np.unique(C.column_to_numpy(t[0])[0][:100]))
    srv.release(qid)
    with pytest.raises(ValueError):  # schema differs from the first table
        srv.add_table([_col(rng.integers(0, 100, 10), C.INT32), _col(rng.random(10), C.DOUBLE)])
    qid = srv.submit_groupby(0, 0, 1000, [(1, "sum", None), (None, "count_all", None)])
    cols, st = srv.wait(qid)
    assert cols[0].length == len(np.unique(C.column_to_numpy(t[0])[0]))
    srv.release(qid)
    with pytest.raises(ValueError):  # released
        srv.wait(qid)
    srv.close()


def test_served_groupby_unaligned_row_ranges_with_nulls(gpu):
    """Groupby row ranges at any start row over a table with null keys and values: the validity of
    a range that does not start on a 64-row word is a bit-shifted copy; every result equals the
    plain groupby of the same rows (sliced on the host), batched and alone."""
    rng = np.random.default_rng(17)
    n = 300_000
    k, kv = rng.integers(0, 5000, n), rng.random(n) > 0.05
    a, av = rng.integers(-(2**40), 2**40, n), rng.random(n) > 0.1
    f = rng.random(n)
    table = [_col(k, C.INT64, kv), _col(a, C.INT64, av), _col(f, C.DOUBLE)]
    srv = C.Server(hbm_budget_bytes=8e9, slots=2, max_batch=8)
    srv.add_table(table)
    ranges = [(1, 1000), (7, 200_003), (63, 64), (65, 299_999), (12345, 12346), (0, n), (99_999, 100_063), (8, 16)]
    ids = [srv.submit_groupby(0, b, e, AGGS) for b, e in ranges]
    for qid, (b, e) in zip(ids, ranges):
        cols, _ = srv.wait(qid)
        plain = C.groupby([_col(k[b:e], C.INT64, kv[b:e]), _col(a[b:e], C.INT64, av[b:e]), _col(f[b:e], C.DOUBLE)],
                          [0], AGGS)
        assert _rows(cols) == _rows(plain), (b, e)
        srv.release(qid)
    srv.close()


def _vals(cols):
    """Rows as a sorted multiset; strings as str, floats as raw bits, None for nulls."""
    data = []
    for c in cols:
        v, ok = C.column_to_numpy(c)
        if isinstance(v, list):
            data.append(list(v))
            continue
        if ok is None:
            ok = np.ones(len(v), bool)
        if v.dtype.kind == "f":
            v = v.view(np.int64 if v.dtype.itemsize == 8 else np.int32)
        data.append([int(x) if o else None for x, o in zip(v, ok)])
    return sorted(zip(*data), key=lambda r: tuple((0, "") if x is None else (1, x) for x in r))


def test_served_string_columns_and_plain_kinds(gpu):
    """A table with a string key and a string-only schema for the probes: such tables get no key
    index and their queries run alone through the plain calls -- joins of the probe with the
    whole table, groupbys over row ranges that start anywhere (string offsets of a range point into
    the table's characters) with COUNT_DISTINCT and MEAN.  Each result equals the plain call's."""
    rng = np.random.default_rng(23)
    n = 120_000
    words = np.array([f"k{i:06d}" * (1 + i % 3) for i in range(6000)], dtype=object)
    keys = list(words[rng.integers(0, 6000, n)])
    a, av = rng.integers(-(2**40), 2**40, n), rng.random(n) > 0.1
    f = rng.random(n)
    table = [C.column_from_strings(keys), _col(a, C.INT64, av), _col(f, C.DOUBLE)]
    srv = C.Server(hbm_budget_bytes=8e9, slots=2, max_batch=8)
    srv.add_table(table)
    aggs = [(1, "sum", None), (2, "mean", None), (1, "count_distinct", None), (None, "count_all", None)]
    ranges = [(0, n), (1, 999), (77, 100_077), (64, 128)]
    probes = []
    for size in (500, 20_000):
        pk = list(words[rng.integers(0, 8000, size) % 6000])
        probes.append([C.column_from_strings(pk), _col(rng.integers(0, 9, size), C.INT32)])
    ids = [srv.submit_groupby(0, b, e, aggs) for b, e in ranges] + [srv.submit_join(0, pr, 0) for pr in probes]
    for qid, (b, e) in zip(ids, ranges):
        cols, st = srv.wait(qid)
        assert st["plain"] == 1
        plain = C.groupby([C.column_from_strings(keys[b:e]), _col(a[b:e], C.INT64, av[b:e]), _col(f[b:e], C.DOUBLE)],
                          [0], aggs)
        got, exp = _vals(cols), _vals(plain)
        assert [(r[0], r[1], r[3], r[4]) for r in got] == [(r[0], r[1], r[3], r[4]) for r in exp], (b, e)
        gm = np.array([np.int64(r[2]).view(np.float64) for r in got])
        em = np.array([np.int64(r[2]).view(np.float64) for r in exp])
        np.testing.assert_allclose(gm, em, rtol=1e-12)
        srv.release(qid)
    for qid, pr in zip(ids[len(ranges):], probes):
        cols, st = srv.wait(qid)
        assert st["plain"] == 1
        assert _vals(cols) == _vals(C.join(pr, table, [0], [0], 0))
        srv.release(qid)
    srv.close()
