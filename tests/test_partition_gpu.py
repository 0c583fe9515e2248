"""GPU hash partition parity: row hashes equal the oracle's murmur3 restatement bit for bit, and
the partitioned table equals the oracle's stable partition exactly (rows keep input order inside a
partition), for fixed-width, nullable, multi-column and string tables."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pytest

import oracle
from ldf_testing import ocol

pytestmark = pytest.mark.gpu


def gpu_partition(t: pa.Table, cols, nparts, seed=0):
    from legate_dataframe import LogicalTable
    from legate_dataframe.lib.hash_partition import hash_partition

    res, offs = hash_partition(LogicalTable.from_arrow(t), cols, nparts, seed)
    return res.to_arrow(), offs


def _expected(t: pa.Table, cols, nparts, seed):
    perm, offs = oracle.partition([ocol(t.column(c)) for c in cols], nparts, seed)
    return t.take(pa.array(perm)), offs


@pytest.mark.parametrize("nparts", [1, 7, 256, 2048])
@pytest.mark.parametrize("seed", [0, 12345])
def test_partition_int64_exact(gpu, nparts, seed):
    rng = np.random.default_rng(nparts)
    t = pa.table({"k": rng.integers(-(2**63), 2**63 - 1, 300_001, dtype=np.int64),
                  "v": rng.integers(0, 100, 300_001).astype(np.int16)})
    got, offs = gpu_partition(t, ["k"], nparts, seed)
    exp, eoffs = _expected(t, ["k"], nparts, seed)
    assert list(offs) == list(eoffs)
    assert got.equals(exp)


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.uint16, np.int8])
@pytest.mark.parametrize("nparts", [3, 64, 100, 256, 512, 1000, 2048])
def test_partition_single_column_exact(gpu, dtype, nparts):
    """One fixed-width column hashed on itself (the value-staged kernels): every width, power-of-two
    and other partition counts, partition counts whose per-wave lane masks fit the staging area and
    ones that do not; the partitioned column equals the oracle's stable partition row for row."""
    rng = np.random.default_rng(nparts * 7 + np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    n = 250_007
    t = pa.table({"k": rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)})
    got, offs = gpu_partition(t, ["k"], nparts, 0)
    exp, eoffs = _expected(t, ["k"], nparts, 0)
    assert list(offs) == list(eoffs)
    assert got.equals(exp)


def test_partition_nullable_multicolumn_strings(gpu):
    rng = np.random.default_rng(5)
    n = 100_000
    t = pa.table({"a": pa.array(rng.integers(0, 1000, n).astype(np.int32), mask=rng.random(n) < 0.1),
                  "s": pa.array([f"key{v}" for v in rng.integers(0, 5000, n)], mask=rng.random(n) < 0.05),
                  "x": pa.array(rng.random(n), mask=rng.random(n) < 0.2)})
    for cols in (["a"], ["s"], ["a", "s"], ["x", "a"]):
        got, offs = gpu_partition(t, cols, 64, 0)
        exp, eoffs = _expected(t, cols, 64, 0)
        assert list(offs) == list(eoffs)
        got = pa.Table.from_arrays([c.cast(pa.string()) if pa.types.is_large_string(c.type) else c
                                    for c in got.columns], names=got.column_names)
        assert got.equals(exp), cols


@pytest.mark.parametrize("nullable", [False, True])
def test_partition_float_keys_nan_and_signed_zero(gpu, nullable):
    """Float keys as libcudf hashes them: NaNs of any payload share a partition, -0.0 shares 0.0's
    (both the single-column value-staged path and the generic path), exact against the oracle."""
    rng = np.random.default_rng(9)
    n = 120_000
    nan_bits = np.array([0x7FF8000000000000, 0xFFF8000000000000, 0x7FF0000000000001, 0x7FF8DEADBEEF0001],
                        np.uint64)
    d = rng.normal(size=n)
    pick = rng.integers(0, 6, n)
    d[pick == 0] = nan_bits[rng.integers(0, 4, n)].view(np.float64)[pick == 0]
    d[pick == 1] = 0.0
    d[pick == 2] = -0.0
    f = d.astype(np.float32)
    f32_nans = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0x7FC12345], np.uint32).view(np.float32)
    f[pick == 0] = f32_nans[rng.integers(0, 4, n)][pick == 0]
    mask = (rng.random(n) < 0.05) if nullable else None
    t = pa.table({"d": pa.array(d, mask=mask), "f": pa.array(f, mask=mask)})
    for cols, tab in ((["d"], t.select(["d"])), (["f"], t.select(["f"])), (["d"], t), (["f", "d"], t)):
        got, offs = gpu_partition(tab, cols, 64, 0)
        exp, eoffs = _expected(tab, cols, 64, 0)
        assert list(offs) == list(eoffs), cols
        for c in tab.column_names:  # compare bits: NaN payloads must travel unchanged
            gb = got[c].combine_chunks()
            eb = exp[c].combine_chunks()
            assert gb.is_valid().equals(eb.is_valid())
            w = np.uint64 if c == "d" else np.uint32
            assert np.array_equal(gb.fill_null(0).to_numpy().view(w), eb.fill_null(0).to_numpy().view(w)), (cols, c)
        # rows equal under libcudf's comparison share a partition
        part = np.repeat(np.arange(64), np.diff(offs))
        key = got[cols[0]].combine_chunks()
        vals = key.fill_null(0).to_numpy()
        valid = np.asarray(key.is_valid())
        if len(cols) == 1:
            for sel in (np.isnan(vals) & valid, (vals == 0) & valid):
                assert len(set(part[sel].tolist())) == 1


def test_row_hash_matches_oracle(gpu):
    import torch

    from legate_dataframe import _capi as C

    rng = np.random.default_rng(2)
    n = 200_000
    x = rng.integers(-(2**63), 2**63 - 1, n, dtype=np.int64)
    y = rng.integers(0, 2**31, n).astype(np.int32)
    yv = rng.random(n) > 0.1
    cols = [C.column_from_numpy(x, C.INT64), C.column_from_numpy(y, C.INT32, valid=yv)]
    for seed in (0, 7):
        h = C.murmur3_hash(cols, [0, 1], seed).cpu().numpy().view(np.uint32)
        e = oracle.row_hashes([oracle.OCol(oracle.T_INT64, x), oracle.OCol(oracle.T_INT32, y, yv)], seed)
        assert np.array_equal(h, e)
    s = [f"str{i}" * (i % 5) for i in range(1000)]
    h = C.murmur3_hash([C.column_from_strings(s)], [0]).cpu().numpy().view(np.uint32)
    assert np.array_equal(h, oracle.row_hashes([oracle.OCol(oracle.T_LARGE_STRING, s)], 0))


def test_partition_empty_and_tiny(gpu):
    t = pa.table({"k": pa.array([], pa.int64())})
    got, offs = gpu_partition(t, ["k"], 4)
    assert got.num_rows == 0 and list(offs) == [0] * 5
    t = pa.table({"k": pa.array([42], pa.int64())})
    got, offs = gpu_partition(t, ["k"], 4)
    exp, eoffs = _expected(t, ["k"], 4, 0)
    assert list(offs) == list(eoffs) and got.equals(exp)


@pytest.mark.parametrize("n", [200_000_000, 1_000_000_000])
def test_partition_keys_invariants_at_scale(gpu, n):
    """200M and 1B int64 keys (one GPU's whole share of the shuffle) into 256 partitions: offsets
    are monotone and sum to n, every key lands in its murmur3 partition (the device row hash is
    checked bit for bit against the oracle's on a 1M-key slice first), the multiset of keys is
    preserved (wrap-around sum and a sum of a mixed form of every key), and partition 17 equals
    the input rows of that partition in input order."""
    import torch

    from legate_dataframe import _capi as C

    P = 256
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randint(-(2**63), 2**63 - 1, (n,), device="cuda", generator=g, dtype=torch.int64)
    col = C.DeviceColumn(C.INT64, n, x.view(torch.uint8))
    out, offs = C.hash_partition([col], [0], P)
    y = out[0].data.view(torch.int64)
    xs = x[:1_000_000].cpu().numpy()
    hs = C.murmur3_hash([C.DeviceColumn(C.INT64, len(xs), x[:1_000_000].view(torch.uint8))], [0])
    assert np.array_equal(hs.cpu().numpy().view(np.uint32), oracle.row_hashes([oracle.OCol(oracle.T_INT64, xs)], 0))
    assert offs[0] == 0 and offs[-1] == n and all(offs[i] <= offs[i + 1] for i in range(P))
    h = C.murmur3_hash([C.DeviceColumn(C.INT64, n, y.view(torch.uint8))], [0]).to(torch.int64) & 0xFFFFFFFF
    part = torch.repeat_interleave(torch.arange(P, device="cuda"), torch.tensor(np.diff(offs), device="cuda"))
    assert torch.equal(h % P, part)
    del h, part
    assert int(x.sum()) == int(y.sum())  # wrap-around sums: the key multiset is preserved
    assert int((x * 0x1E3779B97F4A7C15 >> 7).sum()) == int((y * 0x1E3779B97F4A7C15 >> 7).sum())
    # partition 17 exactly: its rows in input order
    hx = C.murmur3_hash([col], [0]).to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(y[offs[17]:offs[18]], x[hx % P == 17])
