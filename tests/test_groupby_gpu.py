"""GPU groupby parity: the HIP groupby (through the C++ API and the C ABI) against the Acero
golden fixtures and the CPU oracle.  Bar: keys, integer SUM/MIN/MAX/COUNT and COUNT_DISTINCT
bit-exact; float64 SUM/MEAN within 1e-12 relative of the oracle's compensated (Neumaier) sums;
VARIANCE/STDDEV/PRODUCT of floats within 1e-10 (summation order differs)."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc
import pytest

from ldf_testing import assert_groupby_equal, fixture_table, groupby_fixture, groupby_via_oracle, rows_of

pytestmark = pytest.mark.gpu


def gpu_groupby(t: pa.Table, keys, aggs) -> pa.Table:
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.groupby_aggregation import groupby_aggregation

    spec = []
    for col, kind, out in aggs:
        if isinstance(kind, (list, tuple)):
            kind = (kind[0], pc.CountOptions(mode=kind[1]))
        spec.append((col, kind, out))
    return groupby_aggregation(LogicalTable.from_arrow(t), keys, spec).to_arrow()


def _norm(t: pa.Table) -> pa.Table:
    cols = [c.cast(pa.string()) if pa.types.is_large_string(c.type) else c for c in t.columns]
    return pa.Table.from_arrays(cols, names=t.column_names)


def _tol(aggs):
    kinds = {(k[0] if isinstance(k, (list, tuple)) else k) for _, k, _ in aggs}
    return 1e-10 if kinds & {"variance", "stddev", "product"} else 1e-12


@pytest.mark.parametrize("case", range(len(groupby_fixture()["cases"])))
def test_groupby_matches_fixture_and_oracle(gpu, case):
    fx = groupby_fixture()
    c = fx["cases"][case]
    t = fixture_table(fx["tables"][c["table"]])
    aggs = [(a, k, o) for a, k, o in c["aggs"]]
    got = _norm(gpu_groupby(t, c["keys"], aggs))
    # schema: key types kept, Arrow result types
    exp = fixture_table(c["expected"])
    assert got.schema.types == exp.schema.types, (got.schema, exp.schema)
    assert_groupby_equal(got, exp, c["keys"], rel=1e-9)  # Acero: naive float summation
    assert_groupby_equal(got, _norm(groupby_via_oracle(t, c["keys"], aggs)), c["keys"], rel=_tol(aggs))


def _c2_like(n, seed, n_keys=100_000, s=1.1, null_frac=0.01):
    """The 100M-row groupby's column mix at a size the oracle checks in seconds."""
    rng = np.random.default_rng(seed)
    ranks = np.arange(1, n_keys + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -s)
    cdf /= cdf[-1]
    k = np.searchsorted(cdf, rng.random(n)).astype(np.int32)
    m = lambda: rng.random(n) < null_frac  # noqa: E731
    return pa.table({
        "k": k,
        "f": pa.array(rng.uniform(-1, 1, n), mask=m()),
        "i": pa.array(rng.integers(-(2**40), 2**40, n, endpoint=True), mask=m()),
        "u": pa.array(rng.integers(0, 10_000_000, n).astype(np.int32), mask=m()),
    })


C2_AGGS = [("f", "sum", "f_sum"), ("f", "mean", "f_mean"), ("i", "sum", "i_sum"), ("i", "min", "i_min"),
           ("i", "max", "i_max"), ("i", "count", "i_count"), ("u", "count_distinct", "u_nunique")]


@pytest.mark.parametrize("seed", [0, 1])
def test_groupby_c2_mix_vs_oracle(gpu, seed):
    t = _c2_like(2_000_000, seed)
    got = gpu_groupby(t, ["k"], C2_AGGS)
    exp = groupby_via_oracle(t, ["k"], C2_AGGS)
    assert_groupby_equal(got, exp, ["k"], rel=1e-12)


@pytest.mark.parametrize("keys_kind", ["uniform_300k", "zipf_1M", "few_16", "nullkeys_int64", "int64_wide",
                                       "uniform_1200k", "uniform_wide_4M", "int16_signed", "int8_signed",
                                       "int32_outliers_nulls"])
def test_groupby_partitioned_path_matches_global_path(gpu, keys_kind, monkeypatch):
    """Inputs of >= 1M rows take the partitioned path (hot keys in LDS, the rest hash-partitioned
    and aggregated per partition in LDS); LDF_GROUPBY_PARTITIONED=0 forces the global-table path.
    Both must equal the oracle (integers exactly, FP64 sums within 1e-12).  The key counts span
    2^10, 2^11 and 2^12 partitions, and (uniform_wide_4M, ~2.9M keys) sub-passes per partition;
    1- and 2-byte signed keys (negative values zero-extend the same way in the hot key table
    and in the rows) take the packed hot-key lookup of the row classification.  Keys of at most
    4 bytes whose sampled range spans at most n / 4 values are classified through a table over
    that range; int32_outliers_nulls adds null keys and rare keys far outside the sampled range,
    which the table does not cover and the hash probe classifies."""
    rng = np.random.default_rng(21)
    n = 4_500_000 if keys_kind == "uniform_wide_4M" else 1_500_000
    if keys_kind == "uniform_1200k":
        k = pa.array(rng.integers(0, 1_200_000, n).astype(np.int32))
    elif keys_kind == "uniform_wide_4M":
        k = pa.array(rng.integers(0, 4_000_000, n).astype(np.int32))
    elif keys_kind == "uniform_300k":
        k = pa.array(rng.integers(0, 300_000, n).astype(np.int32))
    elif keys_kind == "zipf_1M":
        k = _c2_like(n, 3, n_keys=1_000_000)["k"]
    elif keys_kind == "few_16":
        k = pa.array(rng.integers(0, 16, n).astype(np.int32))
    elif keys_kind == "int16_signed":
        k = pa.array(rng.integers(-20_000, 20_000, n).astype(np.int16))
    elif keys_kind == "int8_signed":
        k = pa.array(rng.integers(-128, 128, n).astype(np.int8))
    elif keys_kind == "int32_outliers_nulls":
        kk = rng.integers(0, 100_000, n).astype(np.int32)
        far = rng.choice(n, 40, replace=False)
        kk[far] = rng.choice(np.array([-7, -(2**31), 2**31 - 1, 5_000_000, 100_000, 100_001], np.int32), 40)
        k = pa.array(kk, mask=rng.random(n) < 0.02)
    elif keys_kind == "nullkeys_int64":
        k = pa.array(rng.integers(0, 50_000, n), pa.int64(), mask=rng.random(n) < 0.02)
    else:
        k = pa.array(rng.integers(-(2**62), 2**62, n // 3).repeat(3)[:n])
    t = pa.table({"k": k,
                  "f": pa.array(rng.uniform(-1, 1, n), mask=rng.random(n) < 0.01),
                  "i": pa.array(rng.integers(-(2**40), 2**40, n), mask=rng.random(n) < 0.01),
                  "u": pa.array(rng.integers(0, 10_000, n).astype(np.int32), mask=rng.random(n) < 0.01)})
    aggs = [("f", "sum", "fs"), ("f", "mean", "fm"), ("i", "sum", "is"), ("i", "min", "imin"), ("i", "max", "imax"),
            ("i", "count", "ic"), ("f", "max", "fmax"), (None, "count_all", "n")]
    if keys_kind in ("uniform_300k", "zipf_1M", "few_16", "uniform_1200k", "uniform_wide_4M", "int16_signed",
                     "int8_signed", "int32_outliers_nulls"):
        aggs.append(("u", "count_distinct", "nu"))
    got = gpu_groupby(t, ["k"], aggs)
    monkeypatch.setenv("LDF_GROUPBY_PARTITIONED", "0")
    ref = gpu_groupby(t, ["k"], aggs)
    assert_groupby_equal(got, ref, ["k"], rel=0.0)  # both paths are exact and order-independent
    assert_groupby_equal(got, groupby_via_oracle(t, ["k"], aggs), ["k"], rel=1e-12)


def test_groupby_float_sum_exact_under_cancellation(gpu):
    """Values spanning 24 orders of magnitude with heavy cancellation: the 192-bit fixed-point
    accumulation returns the correctly rounded exact sum (checked with rational arithmetic)."""
    from fractions import Fraction

    rng = np.random.default_rng(4)
    n = 20000
    k = rng.integers(0, 10, n).astype(np.int64)
    x = rng.uniform(-1, 1, n) * 10.0 ** rng.integers(-12, 12, n)
    x = np.concatenate([x, -x[: n // 2]])  # cancel half of the values exactly
    k = np.concatenate([k, k[: n // 2]])
    t = pa.table({"k": k, "x": x})
    got = gpu_groupby(t, ["k"], [("x", "sum", "s")])
    for key, s in zip(got["k"].to_pylist(), got["s"].to_pylist()):
        exact = sum((Fraction(v) for v in x[k == key]), Fraction(0))
        assert s == float(exact), (key, s, float(exact))


def _fixed96_sums(k, x):
    """Per key: sum of trunc(x * 2^96) over the key's values as a Python integer (the fixed-point
    semantics of the GPU's exact sums), then correctly rounded to float64 by true division."""
    m, e = np.frexp(x)
    mant = (m * 2.0**53).astype(np.int64)  # exact: x = mant * 2^(e - 53)
    sh = (e.astype(np.int64) - 53 + 96).tolist()
    acc = {}
    for key, mt, s in zip(k.tolist(), mant.tolist(), sh):
        v = mt << s if s >= 0 else (abs(mt) >> -s) * (1 if mt >= 0 else -1)
        acc[key] = acc.get(key, 0) + v
    return {key: v / (1 << 96) for key, v in acc.items()}


def test_groupby_float_sum_exact_partitioned_path(gpu):
    """The exact sums of the partitioned path (over 2^20 rows: hot keys in LDS, other keys through
    partitions): mixed signs, magnitudes from 1e-20 to 1e17 and exact cancellation; every group's
    sum is bit for bit the correctly rounded sum of trunc(x * 2^96)."""
    rng = np.random.default_rng(11)
    n = 1_500_000
    hot = rng.integers(0, 8, n // 3)
    cold = rng.integers(8, 300_000, n - n // 3)
    k = rng.permutation(np.concatenate([hot, cold])).astype(np.int32)
    x = rng.choice([-1.0, 1.0], n) * rng.uniform(1, 2, n) * 10.0 ** rng.uniform(-20, 17, n)
    x[: n // 4] = -x[n // 4 : n // 2]  # exact cancellation between rows of different keys
    t = pa.table({"k": k, "x": x})
    got = gpu_groupby(t, ["k"], [("x", "sum", "s"), ("x", "count", "c")])
    ref = _fixed96_sums(k, x)
    gk, gs = got["k"].to_pylist(), got["s"].to_pylist()
    assert len(gk) == len(ref)
    bad = [(key, s, ref[key]) for key, s in zip(gk, gs) if s != ref[key]]
    assert not bad, bad[:5]


def test_groupby_out_of_range_values_fall_back(gpu):
    """|x| >= 2^62, inf and NaN leave the fixed-point range: the FP64 fallback keeps Arrow's
    IEEE results (inf, nan)."""
    t = pa.table({"k": [1, 1, 2, 2, 3, 3], "x": [1e300, 1e300, np.inf, 1.0, np.nan, 2.0]})
    got = gpu_groupby(t, ["k"], [("x", "sum", "s"), ("x", "mean", "m")])
    d = dict(zip(got["k"].to_pylist(), zip(got["s"].to_pylist(), got["m"].to_pylist())))
    assert d[1] == (2e300, 1e300) and d[2][0] == np.inf and np.isnan(d[3][0])


def test_groupby_nulls_and_float_keys(gpu):
    rng = np.random.default_rng(8)
    n = 50000
    t = pa.table({"k": pa.array(rng.integers(0, 50, n).astype(np.float32), mask=rng.random(n) < 0.02),
                  "s": pa.array([f"g{v}" for v in rng.integers(0, 7, n)], mask=rng.random(n) < 0.02),
                  "v": pa.array(rng.integers(-100, 100, n).astype(np.int16), mask=rng.random(n) < 0.1),
                  "w": pa.array(rng.random(n).astype(np.float32), mask=rng.random(n) < 0.1)})
    aggs = [("v", "sum", "vs"), ("v", "min", "vmin"), ("v", "max", "vmax"), ("w", "min", "wmin"),
            ("w", "max", "wmax"), ("w", "mean", "wm"), ("v", ["count", "all"], "ca"), ("v", ["count", "only_null"], "cn"),
            ("w", ["count_distinct", "all"], "nd"), (None, "count_all", "n")]
    for keys in (["k"], ["k", "s"], ["s"]):
        got = _norm(gpu_groupby(t, keys, aggs))
        exp = _norm(groupby_via_oracle(t, keys, aggs))
        assert_groupby_equal(got, exp, keys, rel=1e-12)


def test_groupby_c2_full_size_properties(gpu):
    """The 100M-row, 1M-key Zipf(1.1) groupby at full size (integer columns only; the bench's own
    100M-row input with nulls, float sums and NUNIQUE is checked against the oracle in
    test_configs_full_gpu.py), through size-independent properties: group count = distinct keys,
    per-key counts = torch.unique's counts, integer sums / min / max against a torch reduction by
    sorted key."""
    import torch

    from legate_dataframe import _capi as C

    n, nk = 100_000_000, 1_000_000
    g = torch.Generator(device="cuda").manual_seed(2)
    ranks = torch.arange(1, nk + 1, device="cuda", dtype=torch.float64)
    cdf = torch.cumsum(ranks.pow(-1.1), 0)
    cdf /= cdf[-1].clone()
    k = torch.searchsorted(cdf, torch.rand(n, device="cuda", generator=g, dtype=torch.float64)).to(torch.int32)
    k.clamp_(max=nk - 1)
    vi = torch.randint(-(2**40), 2**40 + 1, (n,), device="cuda", generator=g, dtype=torch.int64)
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    out = C.groupby([dc(k, C.INT32), dc(vi, C.INT64)], [0],
                    [(1, "sum", None), (1, "min", None), (1, "max", None), (1, "count", None)])
    keys = out[0].data.view(torch.int32)[: out[0].length]
    uk, counts = torch.unique(k, return_counts=True)
    assert out[0].length == uk.numel()
    order = torch.argsort(keys)
    assert torch.equal(keys[order], uk)
    assert torch.equal(out[4].data.view(torch.int64)[order], counts)
    # exact int64 reference by sorting on the key
    so = torch.argsort(k, stable=True)
    ks, vs = k[so], vi[so]
    seg = torch.repeat_interleave(torch.arange(uk.numel(), device="cuda"), counts)
    ref_sum = torch.zeros(uk.numel(), dtype=torch.int64, device="cuda").index_add_(0, seg, vs)
    ref_min = torch.full((uk.numel(),), 2**62, dtype=torch.int64, device="cuda").scatter_reduce_(0, seg, vs, "amin")
    ref_max = torch.full((uk.numel(),), -(2**62), dtype=torch.int64, device="cuda").scatter_reduce_(0, seg, vs, "amax")
    assert torch.equal(out[1].data.view(torch.int64)[order], ref_sum)
    assert torch.equal(out[2].data.view(torch.int64)[order], ref_min)
    assert torch.equal(out[3].data.view(torch.int64)[order], ref_max)
    del ks, vs, so, seg


@pytest.mark.parametrize("keys_kind", ["zipf_1M", "uniform_2M"])
def test_groupby_c2_full_size_count_distinct(gpu, keys_kind):
    """COUNT_DISTINCT at 100M rows: per key, the number of distinct (key, value) pairs from
    torch.unique.  zipf_1M (1M keys, Zipf(1.1)): hot keys' pairs stored in row order and split by
    pair partition in many tiles per block, the other keys' ~7.8K pairs per partition deduplicated
    in two hash-filtered passes.  uniform_2M (2M keys, none hot): ~24K pairs per partition, split
    into five groups through global memory before deduplication."""
    import torch

    from legate_dataframe import _capi as C

    n = 100_000_000
    nk = 1_000_000 if keys_kind == "zipf_1M" else 2_000_000
    g = torch.Generator(device="cuda").manual_seed(7)
    if keys_kind == "zipf_1M":
        ranks = torch.arange(1, nk + 1, device="cuda", dtype=torch.float64)
        cdf = torch.cumsum(ranks.pow(-1.1), 0)
        cdf /= cdf[-1].clone()
        k = torch.searchsorted(cdf, torch.rand(n, device="cuda", generator=g, dtype=torch.float64)).to(torch.int32)
        k.clamp_(max=nk - 1)
        del cdf
    else:
        k = torch.randint(0, nk, (n,), device="cuda", generator=g, dtype=torch.int32)
    u = torch.randint(0, 10_000_000, (n,), device="cuda", generator=g, dtype=torch.int32)
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    out = C.groupby([dc(k, C.INT32), dc(u, C.INT32)], [0], [(1, "count_distinct", None)])
    keys = out[0].data.view(torch.int32)[: out[0].length].to(torch.int64)
    got = out[1].data.view(torch.int64)[: out[0].length]
    pairs = torch.unique((k.to(torch.int64) << 32) | u.to(torch.int64))
    ref = torch.bincount(pairs >> 32, minlength=nk)
    del pairs
    assert torch.equal(got, ref[keys])
    assert int(got.sum()) == int(ref.sum())


def test_groupby_count_distinct_group_overflow(gpu):
    """COUNT_DISTINCT where the keys outside the hot set have ~8K rows each and nearly all of a
    key's rows share one value: ~2 such keys per partition, so a partition's ~16K pairs split by
    pair hash into three dedupe groups very unevenly (the repeated pairs all land in one group,
    far past a third of the partition).  Expected counts from torch.unique over the (key, value)
    pairs."""
    import torch

    from legate_dataframe import _capi as C

    n, nk = 24_000_000, 3_000
    g = torch.Generator(device="cuda").manual_seed(11)
    k = torch.randint(0, nk, (n,), device="cuda", generator=g, dtype=torch.int32)
    few = torch.randint(0, 100, (n,), device="cuda", generator=g, dtype=torch.int32)
    u = torch.where(torch.rand(n, device="cuda", generator=g) < 0.9, torch.full_like(few, 5), few)
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    out = C.groupby([dc(k, C.INT32), dc(u, C.INT32)], [0], [(1, "count_distinct", None)])
    keys = out[0].data.view(torch.int32)[: out[0].length].to(torch.int64)
    got = out[1].data.view(torch.int64)[: out[0].length]
    pairs = torch.unique((k.to(torch.int64) << 32) | u.to(torch.int64))
    ref = torch.bincount(pairs >> 32, minlength=nk)
    assert out[0].length == nk
    assert torch.equal(got, ref[keys])


@pytest.mark.parametrize("n", [5_000, 1_500_000])
def test_groupby_approximate_median_vs_oracle(gpu, n):
    """approximate_median is the exact median (cudf's, which the reference's GPU variant uses):
    against the oracle's sort-based median (pinned by numpy.median on the CPU), bit for bit, with
    nulls, NaN, all-null groups, integer and float inputs, null and string keys, and beside other
    aggregations (1.5M rows: the groups come from the partitioned path)."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, n)
    x[rng.random(n) < 0.01] = np.nan
    k = rng.integers(0, max(10, n // 50), n)
    xm = rng.random(n) < 0.1
    xm[k == 3] = True  # a group without valid values
    t = pa.table({"k": pa.array(k, mask=rng.random(n) < 0.01), "s": [f"g{v}" for v in rng.integers(0, 5, n)],
                  "x": pa.array(x, mask=xm), "i": rng.integers(-(2**40), 2**40, n).astype(np.int64),
                  "u": rng.integers(0, 1000, n).astype(np.uint16), "f": rng.random(n).astype(np.float32)})
    aggs = [("x", "approximate_median", "xm"), ("i", "sum", "is"), ("i", "approximate_median", "im"),
            ("u", "approximate_median", "um"), ("f", "approximate_median", "fm"), ("x", "count", "xc")]
    for keys in (["k"], ["k", "s"]):
        got = _norm(gpu_groupby(t, keys, aggs))
        assert got.schema.field("xm").type == pa.float64()
        exp = _norm(groupby_via_oracle(t, keys, aggs))
        assert_groupby_equal(got, exp, keys, rel=0.0)


def test_groupby_median_only(gpu):
    t = pa.table({"k": [1, 1, 1, 2, 2, 3], "v": [5, 1, 3, 2, 8, 7]})
    got = gpu_groupby(t, ["k"], [("v", "approximate_median", "m")])
    assert dict(zip(got["k"].to_pylist(), got["m"].to_pylist())) == {1: 3.0, 2: 5.0, 3: 7.0}


def test_groupby_errors(gpu):
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.groupby_aggregation import groupby_aggregation

    t = LogicalTable.from_arrow(pa.table({"k": [1, 2], "x": [1.0, 2.0]}))
    with pytest.raises(ValueError):
        groupby_aggregation(t, [], [("x", "sum", "s")])
    with pytest.raises(ValueError):
        groupby_aggregation(t, ["k"], [("x", "sum", "k")])  # name conflict with a key
    with pytest.raises(ValueError):
        groupby_aggregation(t, ["k"], [("x", "no_such_kind", "s")])
    with pytest.raises(ValueError):
        groupby_aggregation(t, ["k"], [("x", ("sum", pc.CountOptions()), "s")])
