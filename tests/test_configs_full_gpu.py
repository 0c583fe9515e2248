"""The benchmark's configurations at their full sizes, on the exact inputs bench.py times, checked
against the CPU oracle (the reference tests' pattern: the GPU result against the CPU engine,
python/tests/test_groupby_aggregation.py:28-53, cpp/tests/test_join.cpp:63-113).

* 100M-row groupby (7 aggregations, 1% nulls): every group against the oracle -- keys, integer
  SUM/MIN/MAX/COUNT and NUNIQUE bit-exact, FP64 SUM/MEAN within 1e-12 relative.
* Batched form, 4096 join + groupby instances: every instance's fingerprint (row count and two
  wrap-around sums of per-row hashes: order-independent) equals the plain entry point's on that
  instance, for the captured run and a graph replay; 64 instances spanning 1K..4M rows are also
  compared with the oracle (join rows by fingerprint of the oracle's rows; groupby exactly, FP64
  within 1e-12).
* Served form, 8 tables of 10M..200M rows and 200 queries: every query's result fingerprint equals
  the plain entry point's, alone and batched; joins against the two smallest tables and groupbys
  over up to 2M rows are compared with the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from ldf_testing import device_columns_from_numpy, table_fingerprint

pytestmark = pytest.mark.gpu


def _np(c):
    from legate_dataframe import _capi as C

    return C.column_to_numpy(c)


def _by_key(keys, *cols):
    o = np.argsort(keys, kind="stable")
    return (keys[o],) + tuple(c[o] if c is not None else None for c in cols)


def _close(a, b, rel):
    return np.all((np.abs(a - b) <= rel * np.abs(b)) | (a == b))


# ---------------------------------------------------------------------------------------------
# 100M-row groupby
# ---------------------------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_groupby_100m_rows_all_aggregations_vs_oracle(gpu):
    import torch

    import bench
    from legate_dataframe import _capi as C

    n = 100_000_000
    cols, host = bench.make_groupby_inputs(torch, C, n, seed=2)
    out = C.groupby(cols, [0], bench.GB_AGGS)
    torch.cuda.synchronize()
    k = host["k"].cpu().numpy()
    f = host["f"].cpu().numpy()
    i = host["i"].cpu().numpy()
    u = host["u"].cpu().numpy()
    m = [x.cpu().numpy() for x in host["masks"]]
    del cols, host
    oc = lambda t, v, ok: oracle.OCol(t, v, ok)  # noqa: E731
    fcol, icol, ucol = oc(oracle.T_DOUBLE, f, m[0]), oc(oracle.T_INT64, i, m[1]), oc(oracle.T_INT32, u, m[2])
    reps, res = oracle.groupby([oracle.OCol(oracle.T_INT32, k)],
                               [(fcol, "sum", None), (fcol, "mean", None), (icol, "sum", None), (icol, "min", None),
                                (icol, "max", None), (icol, "count", None), (ucol, "count_distinct", None)])
    print(f"[groupby 100M] oracle: {len(reps)} groups", flush=True)
    gk = _np(out[0])[0]
    assert len(gk) == len(reps)
    got = [_np(c) for c in out[1:]]
    o = np.argsort(gk)
    eo = np.argsort(k[reps])
    assert np.array_equal(gk[o], k[reps][eo])
    # FP64 SUM, MEAN: validity exact, values within 1e-12 relative of the compensated sums
    for j in (0, 1):
        gv, gok = got[j]
        bits, f64, ok = res[j]
        assert np.array_equal(gok[o], ok[eo])
        sel = ok[eo]
        assert _close(gv[o][sel], f64[eo][sel], 1e-12), j
    # int64 SUM / MIN / MAX bit-exact (wrap-around sums), COUNT and NUNIQUE exact
    for j in (2, 3, 4):
        gv, gok = got[j]
        bits, f64, ok = res[j]
        assert np.array_equal(gok[o], ok[eo])
        sel = ok[eo]
        assert np.array_equal(gv[o][sel], bits[eo][sel].view(np.int64)), j
    for j in (5, 6):
        gv, _ = got[j]
        assert np.array_equal(gv[o], res[j][0][eo].view(np.int64)), j
    # the totals the groups add up to
    assert int(got[5][0].sum()) == int(m[1].sum())


# ---------------------------------------------------------------------------------------------
# batched form, 4096 instances
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def batched_inputs(gpu):
    import torch

    import bench
    from legate_dataframe import _capi as C

    L, R, G, keep, info = bench.make_batched_inputs(torch, C, 4096)
    torch.cuda.synchronize()
    yield L, R, G, info
    del L, R, G, keep
    torch.cuda.empty_cache()
    C.pool_release()


def _oracle_pick(sizes, count=64):
    order = np.argsort(sizes, kind="stable")
    return order[np.linspace(0, len(sizes) - 1, count).round().astype(int)]


@pytest.mark.timeout(900)
def test_batched_4096_joins_vs_plain_and_oracle(gpu, batched_inputs):
    import torch

    from legate_dataframe import _capi as C

    L, R, G, info = batched_inputs
    ninst = len(L)
    plain = np.zeros((ninst, 3), np.int64)
    for i in range(ninst):
        plain[i] = table_fingerprint(C.join(L[i], R[i], [0], [0], 0))[0]
    print(f"[batched join] plain fingerprints done, {int(plain[:, 0].sum())} rows", flush=True)
    bj = C.batch_join(L, R, [0], [0], 0, config={"use_graph": 1})
    for run in range(2):  # the first run captures the graph, the second replays it
        bj.run()
        cols, offs = bj.result()
        fp = table_fingerprint(cols, offs)
        bad = np.nonzero(np.any(fp != plain, axis=1))[0]
        assert len(bad) == 0, (run, bad[:10])
        del cols
    bj.close()
    # 64 instances spanning the size range against the oracle
    pick = _oracle_pick(np.array([c[0].length for c in L]))
    assert L[pick[0]][0].length < 2000 and L[pick[-1]][0].length > 3_000_000
    for i in pick:
        pk, pp = _np(L[i][0])[0], _np(L[i][1])[0]
        bk, bp = _np(R[i][0])[0], _np(R[i][1])[0]
        lm, rm = oracle.join([oracle.OCol(oracle.T_INT64, pk)], [oracle.OCol(oracle.T_INT64, bk)], "inner")
        exp = device_columns_from_numpy([(pk[lm], C.INT64, None), (pp[lm], C.DOUBLE, None),
                                         (bk[rm], C.INT64, None), (bp[rm], C.INT64, None)])
        assert np.array_equal(table_fingerprint(exp)[0], plain[i]), i
    torch.cuda.synchronize()


@pytest.mark.timeout(900)
def test_batched_4096_groupbys_vs_plain_and_oracle(gpu, batched_inputs):
    import bench
    from legate_dataframe import _capi as C

    L, R, G, info = batched_inputs
    ninst = len(G)
    plain = np.zeros((ninst, 3), np.int64)
    for i in range(ninst):
        plain[i] = table_fingerprint(C.groupby(G[i], [0], bench.BATCH_GB_AGGS))[0]
    print(f"[batched groupby] plain fingerprints done, {int(plain[:, 0].sum())} groups", flush=True)
    bg = C.batch_groupby(G, [0], bench.BATCH_GB_AGGS, config={"use_graph": 1})
    for run in range(2):
        bg.run()
        cols, offs = bg.result()
        fp = table_fingerprint(cols, offs)
        bad = np.nonzero(np.any(fp != plain, axis=1))[0]
        assert len(bad) == 0, (run, bad[:10])
        del cols
    bg.close()
    pick = _oracle_pick(np.array([c[0].length for c in G]))
    for i in pick:
        k, a, f = (_np(c)[0] for c in G[i])
        reps, res = oracle.groupby([oracle.OCol(oracle.T_INT32, k)],
                                   [(oracle.OCol(oracle.T_INT64, a), "sum", None),
                                    (oracle.OCol(oracle.T_DOUBLE, f), "sum", None),
                                    (oracle.OCol(oracle.T_INT64, a), "count", None)])
        out = C.groupby(G[i], [0], bench.BATCH_GB_AGGS)
        gk, ga, gf, gc = (_np(c)[0] for c in out)
        gk, ga, gf, gc = _by_key(gk, ga, gf, gc)
        ek, ea, ef, ec = _by_key(k[reps], res[0][0].view(np.int64), res[1][1], res[2][0].view(np.int64))
        assert np.array_equal(gk, ek) and np.array_equal(ga, ea) and np.array_equal(gc, ec), i
        assert _close(gf, ef, 1e-12), i


# ---------------------------------------------------------------------------------------------
# served form, 8 tables of 10M..200M rows, 200 queries
# ---------------------------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_served_200_queries_vs_plain_and_oracle(gpu):
    import torch

    import bench
    from legate_dataframe import _capi as C

    tables, queries, sizes = bench.make_served_inputs(torch, C, 200)
    torch.cuda.synchronize()
    aggs = bench.SERVED_AGGS

    def plain(q):
        kind, t, arg = q
        if kind == "join":
            return C.join(arg, tables[t], [0], [0], 0)
        return C.groupby(bench._table_slice(C, tables[t], *arg), [0], aggs)

    exp = np.stack([table_fingerprint(plain(q))[0] for q in queries])
    print(f"[served] plain fingerprints done ({int(exp[:, 0].sum())} result rows)", flush=True)
    srv = C.Server(hbm_budget_bytes=200e9, slots=4, max_batch=64)
    for tb in tables:
        srv.add_table(tb)
    for wave in range(2):  # a burst of every query, then the same queries one at a time
        ids = []
        for kind, t, arg in queries if wave == 0 else queries[:40]:
            ids.append(srv.submit_join(t, arg, 0) if kind == "join" else srv.submit_groupby(t, arg[0], arg[1], aggs))
            if wave == 1:
                res, st = srv.wait(ids[-1])
                assert np.array_equal(table_fingerprint(res)[0], exp[len(ids) - 1]), ("alone", len(ids) - 1)
                srv.release(ids[-1])
        if wave == 0:
            batched = 0
            for j, qid in enumerate(ids):
                res, st = srv.wait(qid)
                batched += st["batch_queries"] > 1
                assert np.array_equal(table_fingerprint(res)[0], exp[j]), ("burst", j, queries[j][0])
                srv.release(qid)
            assert batched > 0
    srv.close()
    # against the oracle: joins with the two smallest tables, groupbys over up to 2M rows
    host_tables = {}
    checked = 0
    for j, (kind, t, arg) in enumerate(queries):
        if kind == "join" and t in (0, 1) and arg[0].length <= 2_000_000:
            if t not in host_tables:
                host_tables[t] = [_np(c)[0] for c in tables[t]]
            bk, ba, bf = host_tables[t]
            pk, pp = _np(arg[0])[0], _np(arg[1])[0]
            lm, rm = oracle.join([oracle.OCol(oracle.T_INT64, pk)], [oracle.OCol(oracle.T_INT64, bk)], "inner")
            e = device_columns_from_numpy([(pk[lm], C.INT64, None), (pp[lm], C.DOUBLE, None), (bk[rm], C.INT64, None),
                                           (ba[rm], C.INT64, None), (bf[rm], C.DOUBLE, None)])
            assert np.array_equal(table_fingerprint(e)[0], exp[j]), j
            checked += 1
        elif kind == "groupby" and arg[1] - arg[0] <= 2_000_000:
            b, e_ = arg
            k, a, f = (_np(c)[0] for c in bench._table_slice(C, tables[t], b, e_))
            reps, res = oracle.groupby([oracle.OCol(oracle.T_INT64, k)],
                                       [(oracle.OCol(oracle.T_INT64, a), "sum", None),
                                        (oracle.OCol(oracle.T_DOUBLE, f), "sum", None), (None, "count_all", None)])
            out = plain(queries[j])
            gk, ga, gf, gc = _by_key(*(_np(c)[0] for c in out))
            ek, ea, ef, ec = _by_key(k[reps], res[0][0].view(np.int64), res[1][1], res[2][0].view(np.int64))
            assert np.array_equal(gk, ek) and np.array_equal(ga, ea) and np.array_equal(gc, ec), j
            assert _close(gf, ef, 1e-12), j
            checked += 1
    print(f"[served] {checked} queries checked against the oracle", flush=True)
    assert checked >= 20
    del tables, queries
    torch.cuda.empty_cache()
    C.pool_release()
