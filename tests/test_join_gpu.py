"""GPU join parity: the HIP hash join (through the C++ API and the C ABI) against the Acero
golden fixtures, the reference's hard-coded LeftJoinOnNulls table and the CPU oracle.
Bar: bit-exact, as sorted multisets of output rows."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pytest

import oracle
from ldf_testing import fixture_table, join_fixture, join_via_oracle, ocol, sorted_rows

pytestmark = pytest.mark.gpu

HOW = {"inner": 0, "left": 1, "full": 2, "semi": 3, "anti": 4}


def gpu_join(lhs: pa.Table, rhs: pa.Table, lk, rk, how, nulls_equal=True, lhs_out=None, rhs_out=None) -> pa.Table:
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.join import JoinType, join

    louts = lhs.column_names if lhs_out is None else lhs_out
    routs = ([] if how in ("semi", "anti") else rhs.column_names) if rhs_out is None else rhs_out
    if set(louts) & set(routs):  # output names must be unique
        rhs = rhs.rename_columns(["r." + c for c in rhs.column_names])
        rk = ["r." + c for c in rk]
        rhs_out = None if rhs_out is None else ["r." + c for c in rhs_out]
    res = join(LogicalTable.from_arrow(lhs), LogicalTable.from_arrow(rhs), lhs_keys=lk, rhs_keys=rk,
               join_type=JoinType(HOW[how]), lhs_out_columns=lhs_out, rhs_out_columns=rhs_out, nulls_equal=nulls_equal)
    return res.to_arrow()


def _norm(t: pa.Table) -> pa.Table:
    # output strings are large_string; the fixtures use string
    cols = [c.cast(pa.string()) if pa.types.is_large_string(c.type) else c for c in t.columns]
    return pa.Table.from_arrays(cols, names=t.column_names)


@pytest.mark.parametrize("case", range(len(join_fixture()["cases"])))
def test_join_matches_acero_fixture(gpu, case):
    c = join_fixture()["cases"][case]
    lhs, rhs = fixture_table(c["lhs"]), fixture_table(c["rhs"])
    got = _norm(gpu_join(lhs, rhs, c["lhs_keys"], c["rhs_keys"], c["how"], c["nulls_equal"]))
    exp = fixture_table(c["expected"])
    assert got.num_rows == exp.num_rows
    assert sorted_rows(got) == sorted_rows(exp)


def test_left_join_on_nulls(gpu):
    """cpp/tests/test_join.cpp:269-304 (the reference's hand-written expectation)."""
    c = [x for x in join_fixture()["cases"] if x["name"] == "left_join_on_nulls"][0]
    lhs, rhs = fixture_table(c["lhs"]), fixture_table(c["rhs"])
    got = _norm(gpu_join(lhs, rhs, ["a", "b"], ["d", "e"], "left", True, ["a", "b", "c"], ["f"]))
    exp = fixture_table(join_fixture()["left_join_on_nulls_expected"])
    assert sorted_rows(got) == sorted_rows(exp)
    # nullability rules (cpp/src/join.cpp:348-383): LEFT forces the rhs outputs nullable
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.join import JoinType, join

    r = join(LogicalTable.from_arrow(pa.table({"k": [1, 2]})), LogicalTable.from_arrow(pa.table({"j": [1], "v": [5]})),
             lhs_keys=["k"], rhs_keys=["j"], join_type=JoinType.LEFT)
    assert not r["k"].nullable() and r["v"].nullable()
    r = join(LogicalTable.from_arrow(pa.table({"k": [1, 2]})), LogicalTable.from_arrow(pa.table({"j": [1], "v": [5]})),
             lhs_keys=["k"], rhs_keys=["j"], join_type=JoinType.FULL)
    assert r["k"].nullable() and r["v"].nullable()
    r = join(LogicalTable.from_arrow(pa.table({"k": [1, 2]})), LogicalTable.from_arrow(pa.table({"j": [1], "v": [5]})),
             lhs_keys=["k"], rhs_keys=["j"], join_type=JoinType.INNER)
    assert not r["k"].nullable() and not r["v"].nullable()


def test_column_names_and_strings(gpu):
    """python/tests/test_join.py:156-178 with string keys and selected out-columns."""
    lhs = pa.table({"key": ["1", "2", "3"], "data0": [1, 2, 3]})
    rhs = pa.table({"key": ["3", "2", "1"], "data1": ["1", "2", "3"]})
    got = gpu_join(lhs, rhs, ["key"], ["key"], "inner", True, ["data0", "key"], ["data1"])
    assert got.column_names == ["data0", "key", "data1"]
    rows = sorted(zip(*[c.to_pylist() for c in got.columns]))
    assert rows == [(1, "1", "3"), (2, "2", "2"), (3, "3", "1")]  # test_join.py:174's expected table


@pytest.mark.parametrize("how", list(HOW))
@pytest.mark.parametrize("nulls_equal", [True, False])
@pytest.mark.parametrize("seed", [0, 1])
def test_join_random_vs_oracle(gpu, how, nulls_equal, seed):
    """int64 keys with duplicates on both sides, null keys and null payloads."""
    rng = np.random.default_rng(seed)
    nl, nr = 30000, 7000
    lhs = pa.table({"k": pa.array(rng.integers(0, 5000, nl), pa.int64(), mask=rng.random(nl) < 0.05),
                    "x": pa.array(rng.random(nl), mask=rng.random(nl) < 0.1),
                    "s": pa.array([f"l{v}" for v in rng.integers(0, 100, nl)])})
    rhs = pa.table({"k": pa.array(rng.integers(0, 5000, nr), pa.int64(), mask=rng.random(nr) < 0.05),
                    "y": pa.array(rng.integers(-(2**60), 2**60, nr), pa.int64()),
                    "t": pa.array(rng.integers(0, 2**31, nr).astype(np.int32))})
    got = gpu_join(lhs, rhs, ["k"], ["k"], how, nulls_equal)
    exp = join_via_oracle(lhs, rhs, ["k"], ["k"], how, nulls_equal)
    assert got.num_rows == exp.num_rows
    assert sorted_rows(_norm(got)) == sorted_rows(_norm(exp))


@pytest.mark.parametrize("how", ["inner", "left", "full"])
def test_join_composite_keys_vs_oracle(gpu, how):
    """(int32, string, float64) keys with nulls: the generic (fingerprint + row compare) path."""
    rng = np.random.default_rng(9)
    nl, nr = 5000, 3000

    def t(n, name):
        return pa.table({"a": pa.array(rng.integers(0, 8, n).astype(np.int32), mask=rng.random(n) < 0.05),
                         "b": pa.array([f"s{v}" for v in rng.integers(0, 5, n)], mask=rng.random(n) < 0.05),
                         "c": pa.array(rng.integers(0, 3, n).astype(np.float64)),
                         name: pa.array(np.arange(n, dtype=np.int64))})

    lhs, rhs = t(nl, "p"), t(nr, "q")
    for ne in (True, False):
        got = gpu_join(lhs, rhs, ["a", "b", "c"], ["a", "b", "c"], how, ne, ["p"], ["q"])
        exp = join_via_oracle(lhs, rhs, ["a", "b", "c"], ["a", "b", "c"], how, ne, ["p"], ["q"])
        assert sorted_rows(got) == sorted_rows(exp)


def test_join_skewed_build_many_to_many(gpu):
    """A build key repeated 20000 times (the skew of the Zipf build side): the emit pass
    spreads one probe row's matches over the whole block."""
    rng = np.random.default_rng(3)
    rk = np.concatenate([np.zeros(20000, np.int64), rng.integers(1, 1000, 20000)])
    lk = rng.integers(0, 1000, 50000)
    lhs = pa.table({"k": lk, "i": np.arange(len(lk))})
    rhs = pa.table({"k": rk, "j": np.arange(len(rk))})
    got = gpu_join(lhs, rhs, ["k"], ["k"], "inner", True, ["i"], ["j"])
    lm, rm = oracle.join([ocol(lhs["k"])], [ocol(rhs["k"])], "inner")
    assert got.num_rows == len(lm)
    g = np.stack([got["i"].to_numpy(), got["j"].to_numpy()], 1)
    e = np.stack([lm, rm], 1).astype(np.int64)
    assert np.array_equal(g[np.lexsort(g.T[::-1])], e[np.lexsort(e.T[::-1])])


def test_join_edge_cases(gpu):
    # key bits equal to the table's empty marker (-1 = all ones), empty sides, all-null keys
    lhs = pa.table({"k": pa.array([-1, -1, 0, None], pa.int64()), "i": [0, 1, 2, 3]})
    rhs = pa.table({"k": pa.array([-1, None, 5], pa.int64()), "j": [10, 11, 12]})
    for how in HOW:
        for ne in (True, False):
            got = gpu_join(lhs, rhs, ["k"], ["k"], how, ne)
            exp = join_via_oracle(lhs, rhs, ["k"], ["k"], how, ne)
            assert sorted_rows(got) == sorted_rows(exp), (how, ne)
    empty = pa.table({"k": pa.array([], pa.int64()), "j": pa.array([], pa.int64())})
    for how in HOW:
        got = gpu_join(lhs, empty, ["k"], ["k"], how, True)
        exp = join_via_oracle(lhs, empty, ["k"], ["k"], how, True)
        assert sorted_rows(got) == sorted_rows(exp), how
        got = gpu_join(empty, lhs, ["k"], ["k"], how, True)
        exp = join_via_oracle(empty, lhs, ["k"], ["k"], how, True)
        assert sorted_rows(got) == sorted_rows(exp), how
    nulls = pa.table({"k": pa.array([None] * 5, pa.int64()), "i": range(5)})
    got = gpu_join(nulls, nulls, ["k"], ["k"], "inner", True, ["i"], [])
    assert got.num_rows == 25


def test_join_errors(gpu):
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.join import BroadcastInput, JoinType, join

    a = LogicalTable.from_arrow(pa.table({"k": [1, 2], "v": [1, 2]}))
    b = LogicalTable.from_arrow(pa.table({"k": [1, 2], "w": [1, 2]}))
    with pytest.raises(ValueError):
        join(a, b, lhs_keys=["k", "v"], rhs_keys=["k"], join_type=JoinType.INNER)
    with pytest.raises(ValueError):
        join(a, b, lhs_keys=["k"], rhs_keys=["k"], join_type=JoinType.SEMI, rhs_out_columns=["w"])
    with pytest.raises(ValueError):  # duplicate output names
        join(a, b, lhs_keys=["k"], rhs_keys=["k"], join_type=JoinType.INNER)
    with pytest.raises(RuntimeError):
        join(a, b, lhs_keys=["k"], rhs_keys=["k"], join_type=JoinType.FULL, rhs_out_columns=["w"],
             broadcast=BroadcastInput.RIGHT)


def test_inner_join_1m_x_10m_closed_form(gpu):
    """The 1M-build x 10M-probe inner join at full size: build keys are a permutation of
    0..999999, probe keys uniform in [0, 2M); probe row i matches iff key_i < 1M, with the
    build row holding that key.  Checked exactly through the C ABI."""
    import torch

    from legate_dataframe import _capi as C

    m, n = 1_000_000, 10_000_000
    g = torch.Generator(device="cuda").manual_seed(1)
    bkeys = torch.randperm(m, device="cuda", generator=g).to(torch.int64)
    bpay = torch.randint(-(2**62), 2**62, (m,), device="cuda", generator=g, dtype=torch.int64)
    pkeys = torch.randint(0, 2 * m, (n,), device="cuda", generator=g, dtype=torch.int64)
    ppay = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    out = C.join([dc(pkeys, C.INT64), dc(ppay, C.DOUBLE)], [dc(bkeys, C.INT64), dc(bpay, C.INT64)], [0], [0], 0)
    lk = out[0].data.view(torch.int64)
    lp = out[1].data.view(torch.float64)
    rk = out[2].data.view(torch.int64)
    rp = out[3].data.view(torch.int64)
    match = pkeys < m
    assert out[0].length == int(match.sum())
    assert torch.equal(lk, rk)  # joined keys agree
    # the multiset of (probe key, probe payload) equals the matching probe rows
    o = torch.argsort(lp)
    e = torch.argsort(ppay[match])
    assert torch.equal(lp[o], ppay[match][e]) and torch.equal(lk[o], pkeys[match][e])
    # each rhs payload is the one stored with that key
    where = torch.empty(m, dtype=torch.int64, device="cuda")
    where[bkeys] = torch.arange(m, device="cuda")
    assert torch.equal(rp, bpay[where[rk]])


@pytest.mark.parametrize("how", ["inner", "left", "full"])
def test_join_many_output_columns_vs_oracle(gpu, how):
    """Eleven output columns per side of every width (1/2/4/8 bytes, strings), with nulls: the
    columns are gathered in batches that share one read of the gather map."""
    rng = np.random.default_rng(21)

    def t(n, p):
        cols = {f"{p}k": pa.array(rng.integers(0, 400, n), pa.int64(), mask=rng.random(n) < 0.05)}
        for j, ty in enumerate([pa.int8(), pa.int16(), pa.int32(), pa.int64(), pa.uint8(), pa.float32(), pa.float64()]):
            v = rng.integers(0, 100, n).astype(ty.to_pandas_dtype())
            cols[f"{p}{j}"] = pa.array(v, ty, mask=rng.random(n) < 0.1)
        for j in range(3):
            strs = ["".join(chr(65 + x) for x in rng.integers(0, 26, rng.integers(0, 40))) for _ in range(n)]
            cols[f"{p}s{j}"] = pa.array(strs, mask=rng.random(n) < 0.1 * j)
        return pa.table(cols)

    lhs, rhs = t(3000, "l"), t(1500, "r")
    got = gpu_join(lhs, rhs, ["lk"], ["rk"], how)
    exp = join_via_oracle(lhs, rhs, ["lk"], ["rk"], how, True)
    assert got.num_rows == exp.num_rows
    assert sorted_rows(_norm(got)) == sorted_rows(_norm(exp))


@pytest.mark.parametrize("how", ["inner", "left"])
def test_join_string_lengths_vs_oracle(gpu, how):
    """String payloads of every length class: empty, short (the row-per-lane copy through LDS
    windows, rows crossing window edges), up to 256 bytes, and multi-kilobyte rows (the
    per-word copy), on both sides (probe side in order, build side through packed records)."""
    rng = np.random.default_rng(33)

    def strs(n):
        lens = rng.choice([0, 1, 3, 7, 17, 31, 64, 200, 256, 257, 1000, 4099], n,
                          p=[.08, .1, .1, .1, .15, .15, .1, .08, .04, .04, .04, .02])
        return [bytes(rng.integers(33, 127, ln, dtype=np.uint8)).decode() for ln in lens]

    def t(n, p):
        return pa.table({f"{p}k": pa.array(rng.integers(0, 500, n), pa.int64()),
                         f"{p}s": pa.array(strs(n), mask=rng.random(n) < 0.05),
                         f"{p}v": pa.array(rng.integers(0, 1 << 40, n), pa.int64())})

    lhs, rhs = t(6000, "l"), t(2500, "r")
    got = gpu_join(lhs, rhs, ["lk"], ["rk"], how)
    exp = join_via_oracle(lhs, rhs, ["lk"], ["rk"], how, True)
    assert got.num_rows == exp.num_rows
    assert sorted_rows(_norm(got)) == sorted_rows(_norm(exp))


def _fmix64(k: np.ndarray) -> np.ndarray:
    """MurmurHash3's 64-bit finaliser (the join table's slot hash), in wrapping uint64."""
    k = k.astype(np.uint64)
    with np.errstate(over="ignore"):
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xFF51AFD7ED558CCD)
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xC4CEB9FE1A85EC53)
        k ^= k >> np.uint64(33)
    return k


@pytest.mark.parametrize("how", ["inner", "left", "full", "semi", "anti"])
def test_join_build_tile_overflow_falls_back(gpu, how):
    """5000 distinct build keys chosen to share one 4096-slot tile of the 16384-slot table: the
    tiled LDS build cannot hold them, which is found when the output size is read back; the
    join then rebuilds over the whole table and probes again (same result for every join type)."""
    rng = np.random.default_rng(5)
    cand = rng.integers(0, 1 << 40, 200_000, dtype=np.int64)
    tile0 = cand[((_fmix64(cand) & np.uint64(16383)) >> np.uint64(12)) == 0]
    keys = np.unique(tile0)[:5000]
    assert keys.size == 5000
    rhs = pa.table({"rk": pa.array(rng.permutation(keys)), "rv": pa.array(rng.integers(0, 1000, 5000))})
    pk = np.concatenate([rng.choice(keys, 6000), rng.integers(0, 1 << 40, 2000)])
    lhs = pa.table({"lk": pa.array(pk), "lv": pa.array(rng.integers(0, 1000, pk.size))})
    got = gpu_join(lhs, rhs, ["lk"], ["rk"], how)
    exp = join_via_oracle(lhs, rhs, ["lk"], ["rk"], how, True)
    assert got.num_rows == exp.num_rows
    assert sorted_rows(_norm(got)) == sorted_rows(_norm(exp))
