"""Batched form (ldf_batch_*) against the plain entry points, instance by instance.

Each instance of a batch must give exactly what ldf_join / ldf_groupby give on that instance
alone (rows in another order): join rows compared as sorted multisets, groupby rows sorted by
key with every value compared bit for bit (FP64 SUM / MEAN included: both paths sum exactly).
Checked for direct launches and for hipGraph replays, on instances of uneven sizes including
empty ones, null keys and null values, and on the fallbacks a batch plans for (partitions whose
keys overflow the LDS table, staging re-plans)."""
import numpy as np
import pytest

from legate_dataframe import _capi as C

pytestmark = pytest.mark.gpu

CFGS = [{"load_factor": 0.5, "tile_rows": 1024, "lds_slots": 512, "use_graph": 0},
        {"load_factor": 0.7, "tile_rows": 2048, "lds_slots": 1024, "use_graph": 1}]


def _col(vals, tid, valid=None):
    return C.column_from_numpy(np.asarray(vals), tid, valid)


def _host(c):
    v, ok = C.column_to_numpy(c)
    if ok is None:
        ok = np.ones(len(v), bool)
    return v, ok


def _rows(cols, lo=0, hi=None):
    """Rows [lo, hi) of the columns as sortable tuples (None for nulls, raw bits for floats)."""
    data = []
    for c in cols:
        v, ok = _host(c)
        if v.dtype.kind == "f":
            v = v.view(np.int64 if v.dtype.itemsize == 8 else np.int32)
        data.append([int(x) if o else None for x, o in zip(v[lo:hi], ok[lo:hi])])
    rows = list(zip(*data))
    return sorted(rows, key=lambda r: tuple((0, 0) if x is None else (1, x) for x in r))


def _join_instances(rng, sizes, nullable):
    L, R = [], []
    for npr, nb in sizes:
        kr = max(4, 2 * nb)
        bk = rng.integers(0, kr, nb)
        pk = rng.integers(0, kr, npr)
        bv = (rng.random(nb) > 0.1) if nullable else None
        pv = (rng.random(npr) > 0.1) if nullable else None
        L.append([_col(pk, C.INT64, pv), _col(rng.random(npr), C.DOUBLE), _col(rng.integers(0, 100, npr), C.INT32)])
        R.append([_col(bk, C.INT64, bv), _col(rng.integers(-(2**60), 2**60, nb), C.INT64,
                                              (rng.random(nb) > 0.2) if nullable else None)])
    return L, R


@pytest.mark.parametrize("cfg", range(len(CFGS)))
@pytest.mark.parametrize("jt", [0, 1, 2, 3, 4], ids=["inner", "left", "full", "semi", "anti"])
def test_batch_join_matches_plain(gpu, jt, cfg):
    rng = np.random.default_rng(10 + jt)
    sizes = [(0, 0), (1, 0), (0, 5), (3000, 300), (1, 1), (17, 900), (50000, 5000), (1234, 123), (256, 2048),
             (10, 10)]
    L, R = _join_instances(rng, sizes, nullable=(jt != 3))
    b = C.batch_join(L, R, [0], [0], jt, config=CFGS[cfg])
    for rep in range(2):  # the second run replays the graph
        b.run()
        cols, offs = b.result()
        assert offs[0] == 0 and len(offs) == len(sizes) + 1
        for i in range(len(sizes)):
            plain = C.join(L[i], R[i], [0], [0], jt)
            assert offs[i + 1] - offs[i] == plain[0].length, (rep, i)
            assert _rows(cols, offs[i], offs[i + 1]) == _rows(plain), (rep, i)
    b.close()


def test_batch_join_nulls_equal_false_and_selected_columns(gpu):
    rng = np.random.default_rng(3)
    L, R = _join_instances(rng, [(2000, 200), (500, 700), (0, 3)], nullable=True)
    b = C.batch_join(L, R, [0], [0], 1, nulls_equal=False, lhs_out=[2, 0], rhs_out=[1], config=CFGS[0])
    b.run()
    cols, offs = b.result()
    for i in range(3):
        plain = C.join(L[i], R[i], [0], [0], 1, nulls_equal=False, lhs_out=[2, 0], rhs_out=[1])
        assert _rows(cols, offs[i], offs[i + 1]) == _rows(plain)
    # output nullability follows the plain path: LEFT makes the rhs columns nullable
    assert cols[0].validity is None and cols[1].validity is not None and cols[2].validity is not None
    # FULL with nulls not equal: null build keys match nothing and come out as unmatched rhs rows
    b = C.batch_join(L, R, [0], [0], 2, nulls_equal=False, config=CFGS[1])
    for rep in range(2):
        b.run()
        cols, offs = b.result()
        for i in range(3):
            plain = C.join(L[i], R[i], [0], [0], 2, nulls_equal=False)
            assert _rows(cols, offs[i], offs[i + 1]) == _rows(plain), (rep, i)
    assert all(c.validity is not None for c in cols)  # FULL: every column nullable
    b.close()


def _gb_instances(rng, specs, knull=False, vnull=False):
    tabs = []
    for n, ng in specs:
        k = rng.integers(0, max(1, ng), n).astype(np.int32)
        i64 = rng.integers(-(2**40), 2**40, n)
        f64 = rng.random(n) * 2 - 1
        f32 = rng.standard_normal(n).astype(np.float32)
        f32[rng.random(n) < 0.01] = np.nan
        kv = (rng.random(n) > 0.05) if knull else None
        tabs.append([_col(k, C.INT32, kv), _col(i64, C.INT64, (rng.random(n) > 0.1) if vnull else None),
                     _col(f64, C.DOUBLE, (rng.random(n) > 0.1) if vnull else None), _col(f32, C.FLOAT)])
    return tabs


GB_AGGS = [(1, "sum", None), (2, "sum", None), (2, "mean", None), (1, "count", None), (1, "count", "all"),
           (2, "count", "only_null"), (None, "count_all", None), (1, "min", None), (1, "max", None), (3, "min", None),
           (3, "max", None), (3, "count", None)]


def _check_groupby(b, tabs, aggs):
    cols, offs = b.result()
    for i, t in enumerate(tabs):
        plain = C.groupby(t, [0], aggs)
        assert offs[i + 1] - offs[i] == plain[0].length, i
        assert _rows(cols, offs[i], offs[i + 1]) == _rows(plain), i


@pytest.mark.parametrize("cfg", range(len(CFGS)))
@pytest.mark.parametrize("nulls", [False, True], ids=["no_nulls", "nulls"])
def test_batch_groupby_matches_plain(gpu, cfg, nulls):
    rng = np.random.default_rng(7)
    specs = [(0, 1), (1, 1), (5000, 3), (20000, 1000), (100000, 20000), (777, 700), (3000, 3000), (40000, 50)]
    tabs = _gb_instances(rng, specs, knull=nulls, vnull=nulls)
    b = C.batch_groupby(tabs, [0], GB_AGGS, config=CFGS[cfg])
    for _ in range(2):
        b.run()
        _check_groupby(b, tabs, GB_AGGS)
    b.close()


def test_batch_groupby_partition_overflow_paths(gpu):
    """Keys the sampled estimate misses: every sampled row (row index a multiple of n/2048) holds
    key 0 and the other rows are all distinct, so the planned partitions receive far more keys
    than their LDS tables hold -- exercising the hash sub-passes and the staging re-plan."""
    rng = np.random.default_rng(11)
    tabs = []
    for n in (20000, 300000):
        k = np.arange(1, n + 1, dtype=np.int32)
        k[:: max(1, n // 2048)] = 0
        tabs.append([_col(k, C.INT32), _col(rng.integers(0, 1000, n), C.INT64), _col(rng.random(n), C.DOUBLE),
                     _col(rng.random(n).astype(np.float32), C.FLOAT)])
    aggs = [(1, "sum", None), (2, "sum", None), (None, "count_all", None)]
    b = C.batch_groupby(tabs, [0], aggs, config={"tile_rows": 1024, "lds_slots": 256, "use_graph": 1})
    b.run()
    _check_groupby(b, tabs, aggs)


def test_batch_autotune_and_identical_results(gpu):
    rng = np.random.default_rng(5)
    tabs = _gb_instances(rng, [(30000, 2000), (500, 100), (60000, 30000)])
    aggs = [(1, "sum", None), (2, "mean", None), (None, "count_all", None)]
    b = C.batch_groupby(tabs, [0], aggs)  # every knob autotuned
    used = b.config_used()
    assert used["tile_rows"] in (1024, 2048, 4096) and used["lds_slots"] in (256, 512, 1024, 2048)
    b.run()
    _check_groupby(b, tabs, aggs)
    L, R = _join_instances(rng, [(4000, 400), (100, 10)], nullable=False)
    bj = C.batch_join(L, R, [0], [0], 0)
    used = bj.config_used()
    assert 0 < used["load_factor"] <= 0.9 and used["tile_rows"] in (1024, 2048, 4096)
    bj.run()
    cols, offs = bj.result()
    for i in range(2):
        assert _rows(cols, offs[i], offs[i + 1]) == _rows(C.join(L[i], R[i], [0], [0], 0))


def _values(cols, lo=0, hi=None):
    """Columns' rows [lo, hi) as lists of values (None for nulls, raw bits for floats)."""
    out = []
    for c in cols:
        v, ok = C.column_to_numpy(c)
        if isinstance(v, list):  # strings
            out.append(list(v[lo:hi]))
            continue
        if ok is None:
            ok = np.ones(len(v), bool)
        if v.dtype.kind == "f":
            v = v.view(np.int64 if v.dtype.itemsize == 8 else np.int32)
        out.append([int(x) if o else None for x, o in zip(v[lo:hi], ok[lo:hi])])
    return out


def _rowset(cols, lo=0, hi=None):
    """The rows as a sorted multiset (a probe row's matches come in no fixed order)."""
    rows = list(zip(*_values(cols, lo, hi)))
    return sorted(rows, key=lambda r: tuple((0, "") if x is None else (1, x) for x in r))


def _strings(rng, n, nkinds):
    words = [f"w{i:05d}" * (1 + i % 3) for i in range(nkinds)]
    return [words[j] for j in rng.integers(0, nkinds, n)]


@pytest.mark.parametrize("jt", [0, 1, 2, 3, 4], ids=["inner", "left", "full", "semi", "anti"])
def test_batch_join_instance_by_instance_schemas(gpu, jt):
    """Schemas the work-queue kernels do not take -- two key columns, string payloads -- run
    instance by instance through the plain join: each instance's rows equal ldf_join's on that
    instance alone (as multisets: a probe row's matches come in no fixed order), strings
    included."""
    rng = np.random.default_rng(40 + jt)
    sizes = [(0, 0), (200, 0), (0, 50), (3000, 300), (17, 900), (1234, 123)]
    L, R = [], []
    for npr, nb in sizes:
        kr = max(4, nb)
        pv = rng.random(npr) > 0.1
        L.append([_col(rng.integers(0, kr, npr), C.INT64, pv), _col(rng.integers(0, 3, npr), C.INT32),
                  C.column_from_strings(_strings(rng, npr, 50))])
        R.append([_col(rng.integers(0, kr, nb), C.INT64), _col(rng.integers(0, 3, nb), C.INT32),
                  C.column_from_strings([None if x else s for s, x in zip(_strings(rng, nb, 70), rng.random(nb) < 0.2)]),
                  _col(rng.random(nb), C.DOUBLE)])
    rhs_out = [] if jt in (3, 4) else [2, 3]
    b = C.batch_join(L, R, [0, 1], [0, 1], jt, lhs_out=[0, 2], rhs_out=rhs_out)
    assert b.config_used()["use_graph"] == 0
    for rep in range(2):
        b.run()
        cols, offs = b.result()
        assert len(cols) == 2 + len(rhs_out) and offs[0] == 0
        for i in range(len(sizes)):
            plain = C.join(L[i], R[i], [0, 1], [0, 1], jt, lhs_out=[0, 2], rhs_out=rhs_out)
            assert offs[i + 1] - offs[i] == plain[0].length, (rep, i)
            assert _rowset(cols, offs[i], offs[i + 1]) == _rowset(plain), (rep, i)
    b.close()


def test_batch_groupby_instance_by_instance_schemas(gpu):
    """A string key with COUNT_DISTINCT and VARIANCE (kinds the work-queue kernels do not take):
    each instance's groups equal ldf_groupby's on that instance alone (as sorted rows; VARIANCE
    compared within 1e-12, its summation order being free)."""
    rng = np.random.default_rng(77)
    tabs = []
    for n, nk in [(0, 1), (1, 1), (5000, 40), (20000, 900), (777, 700)]:
        keys = _strings(rng, n, nk)
        tabs.append([C.column_from_strings(keys), _col(rng.integers(0, 50, n), C.INT32, rng.random(n) > 0.1),
                     _col(rng.random(n), C.DOUBLE)])
    aggs = [(1, "count_distinct", None), (2, "variance", None), (2, "sum", None), (None, "count_all", None)]
    b = C.batch_groupby(tabs, [0], aggs)
    for rep in range(2):
        b.run()
        cols, offs = b.result()
        for i, t in enumerate(tabs):
            plain = C.groupby(t, [0], aggs)
            assert offs[i + 1] - offs[i] == plain[0].length, (rep, i)
            got, exp = _rowset(cols, offs[i], offs[i + 1]), _rowset(plain)
            assert [(r[0], r[1], r[3], r[4]) for r in got] == [(r[0], r[1], r[3], r[4]) for r in exp], (rep, i)
            gv = np.array([np.nan if r[2] is None else np.int64(r[2]).view(np.float64) for r in got])
            ev = np.array([np.nan if r[2] is None else np.int64(r[2]).view(np.float64) for r in exp])
            np.testing.assert_allclose(gv, ev, rtol=1e-12, equal_nan=True)
    b.close()


def test_batch_errors(gpu):
    rng = np.random.default_rng(1)
    L, R = _join_instances(rng, [(10, 10)], nullable=False)
    with pytest.raises(ValueError):  # not a JoinType
        C.batch_join(L, R, [0], [0], 7, config=CFGS[0])
    S = [[C.column_from_strings(["a", "b"])]]
    b = C.batch_join(S, S, [0], [0], 0, config=CFGS[0])  # a string key: instance by instance
    b.run()
    cols, offs = b.result()
    assert list(offs) == [0, 2] and C.column_to_numpy(cols[0])[0] == ["a", "b"]
    b.close()
    with pytest.raises(ValueError):
        C.batch_join(L, R, [0], [0], 0, config={"load_factor": 1.5, "tile_rows": 1024, "lds_slots": 512})
    with pytest.raises(ValueError):  # SEMI with rhs columns
        C.batch_join(L, R, [0], [0], 3, rhs_out=[1], config=CFGS[0])
    tabs = _gb_instances(rng, [(100, 10)])
    C.batch_groupby(tabs, [0], [(1, "count_distinct", None)], config=CFGS[0]).close()  # instance by instance
    big = [[_col(np.zeros(4, np.int32), C.INT32), _col(np.array([1e300, 1.0, 2.0, 3.0]), C.DOUBLE)]]
    b = C.batch_groupby(big, [0], [(1, "sum", None)], config=CFGS[0])
    b.run()
    with pytest.raises(C.LdfError):  # outside the exact FP64 range
        b.result()
