"""Helpers shared by the tests: golden fixtures, oracle columns, comparisons."""

from __future__ import annotations

import json
import math
import struct
from functools import lru_cache
from pathlib import Path

import numpy as np
import pyarrow as pa

GOLDEN = Path(__file__).resolve().parent / "golden"

ARROW_TYPE = {"int8": pa.int8(), "int16": pa.int16(), "int32": pa.int32(), "int64": pa.int64(),
              "uint8": pa.uint8(), "uint16": pa.uint16(), "uint32": pa.uint32(), "uint64": pa.uint64(),
              "float": pa.float32(), "double": pa.float64(), "bool": pa.bool_(), "string": pa.string(),
              "large_string": pa.large_string()}
TYPE_ID = {"bool": 1, "uint8": 2, "int8": 3, "uint16": 4, "int16": 5, "uint32": 6, "int32": 7, "uint64": 8,
           "int64": 9, "float": 11, "double": 12, "string": 34, "large_string": 34}


@lru_cache(maxsize=None)
def join_fixture():
    return json.loads((GOLDEN / "join_cases.json").read_text())


@lru_cache(maxsize=None)
def groupby_fixture():
    return json.loads((GOLDEN / "groupby_cases.json").read_text())


def _vals(c):
    out = []
    for v in c["values"]:
        out.append(float("nan") if v == "nan" else v)
    return out


def fixture_table(tj) -> pa.Table:
    arrays = [pa.array(_vals(c), type=ARROW_TYPE[c["type"]]) for c in tj["columns"]]
    return pa.Table.from_arrays(arrays, names=tj["names"])


def ocol(arr: pa.Array):
    """pyarrow array -> oracle column"""
    from oracle import OCol

    arr = arr.combine_chunks() if isinstance(arr, pa.ChunkedArray) else arr
    tname = str(arr.type)
    tid = TYPE_ID[tname]
    valid = None if arr.null_count == 0 else np.asarray(arr.is_valid())
    if tid == 34:
        return OCol(tid, arr.to_pylist(), valid)
    if tid == 1:
        vals = np.asarray(arr.fill_null(False), dtype=np.uint8)
    else:
        vals = arr.fill_null(0).to_numpy(zero_copy_only=False) if arr.null_count else arr.to_numpy(zero_copy_only=False)
    return OCol(tid, vals, valid)


def rows_of(t: pa.Table):
    """Rows as tuples with NaN made comparable, for sorted multiset comparison."""
    cols = [t.column(i).to_pylist() for i in range(t.num_columns)]
    out = []
    for r in zip(*cols):
        out.append(tuple(("<nan>" if isinstance(v, float) and math.isnan(v) else v) for v in r))
    return out


def sort_key(row):
    return tuple((0, "") if v is None else (1, str(type(v)), v) for v in row)


def sorted_rows(t: pa.Table):
    return sorted(rows_of(t), key=sort_key)


def join_via_oracle(lhs: pa.Table, rhs: pa.Table, lk, rk, how, nulls_equal, lhs_out=None, rhs_out=None) -> pa.Table:
    """Materialises the oracle's pair list into the reference's output table."""
    import oracle

    lm, rm = oracle.join([ocol(lhs.column(k)) for k in lk], [ocol(rhs.column(k)) for k in rk], how, nulls_equal)
    lhs_out = lhs.column_names if lhs_out is None else lhs_out
    rhs_out = ([] if how in ("semi", "anti") else rhs.column_names) if rhs_out is None else rhs_out
    cols, names = [], []
    for side, table, m, outs in (("l", lhs, lm, lhs_out), ("r", rhs, rm, rhs_out)):
        idx = pa.array(np.where(m < 0, 0, m).astype(np.int64))
        mask = pa.array(m < 0)
        for c in outs:
            src = table.column(c).combine_chunks()
            if table.num_rows == 0:
                g = pa.nulls(len(m), type=src.type)
            else:
                g = src.take(idx)
                g = pa.compute.if_else(mask, pa.scalar(None, type=src.type), g)
            cols.append(g)
            names.append(c)
    return pa.Table.from_arrays(cols, names=names)


def assert_groupby_equal(got: pa.Table, exp: pa.Table, keys, rel=1e-12):
    """Groups matched by key; integers/strings exactly, floats within `rel` relative error."""
    assert got.num_rows == exp.num_rows, (got.num_rows, exp.num_rows)
    assert got.column_names == exp.column_names
    g = {tuple(r[: len(keys)]): r for r in rows_of(got)}
    e = {tuple(r[: len(keys)]): r for r in rows_of(exp)}
    assert set(g) == set(e)
    for k, er in e.items():
        gr = g[k]
        for name, a, b in zip(exp.column_names, gr, er):
            if isinstance(b, float) and isinstance(a, float):
                assert abs(a - b) <= rel * max(abs(b), 1e-300) or a == b, (name, k, a, b)
            else:
                assert a == b, (name, k, a, b)


def groupby_via_oracle(t: pa.Table, keys, aggs) -> pa.Table:
    """Materialises the oracle's groupby into the reference's output table (Arrow types)."""
    import oracle

    kcols = [ocol(t.column(k)) for k in keys]
    spec = []
    for col, kind, _ in aggs:
        mode = None
        if isinstance(kind, (list, tuple)):
            kind, mode = kind
        spec.append((ocol(t.column(col)) if col is not None else None, kind, mode))
    reps, res = oracle.groupby(kcols, spec)
    cols = [t.column(k).combine_chunks().take(pa.array(reps)) for k in keys]
    for (col, kind, _), (bits, f64, valid) in zip(aggs, res):
        kind = kind[0] if isinstance(kind, (list, tuple)) else kind
        vt = t.column(col).type if col is not None else pa.int64()
        if kind in ("count", "count_all", "count_distinct"):
            arr = pa.array(bits.view(np.int64))
        elif kind in ("mean", "variance", "stddev", "approximate_median") or \
                (kind in ("sum", "product") and pa.types.is_floating(vt)):
            arr = pa.array(f64, mask=~valid)
        elif kind in ("sum", "product"):
            arr = pa.array(bits.view(np.uint64 if pa.types.is_unsigned_integer(vt) or vt == pa.bool_() else np.int64),
                           mask=~valid)
        else:  # min / max: raw bits of the input type
            npt = {pa.float32(): np.float32, pa.float64(): np.float64}.get(vt, None)
            if vt == pa.bool_():
                arr = pa.array(bits.astype(bool), mask=~valid)
            else:
                width = vt.bit_width // 8
                raw = bits.astype(np.uint64).view(np.uint8).reshape(-1, 8)[:, :width].copy().reshape(-1)
                arr = pa.array(raw.view(npt or vt.to_pandas_dtype()), type=vt, mask=~valid)
        cols.append(arr)
    return pa.Table.from_arrays(cols, names=list(keys) + [o for _, _, o in aggs])


def bits_of(v):
    """A value for exact comparison: floats by their bits (so NaN == NaN, -0.0 != 0.0)."""
    return ("f", struct.pack("<d", v)) if isinstance(v, float) else v


def first_rows(t: pa.Table, keys) -> list:
    """Row index of each distinct key tuple's first occurrence, in row order (the oracle's)."""
    from oracle import distinct_first_rows

    return distinct_first_rows([t.column(k).to_pylist() for k in keys])


# ---------------------------------------------------------------------------------------------
# order-independent fingerprints of device tables (for results too large to sort on the host)
# ---------------------------------------------------------------------------------------------
_M1 = 0xBF58476D1CE4E5B9 - (1 << 64)
_M2 = 0x94D049BB133111EB - (1 << 64)
_G = 0x9E3779B97F4A7C15 - (1 << 64)


def _s64(x):
    """A Python int wrapped to a signed 64-bit value."""
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >= 1 << 63 else x


def _lsr(torch, x, s):
    return (x >> s) & ((1 << (64 - s)) - 1)


def _mix(torch, x):
    """splitmix64's finaliser on int64 tensors (wrap-around arithmetic)."""
    x = x ^ _lsr(torch, x, 30)
    x = x * _M1
    x = x ^ _lsr(torch, x, 27)
    x = x * _M2
    return x ^ _lsr(torch, x, 31)


def column_values_i64(torch, col, lo, hi):
    """Rows [lo, hi) of a fixed-width DeviceColumn as int64 (raw bits, zero-extended) and a bool
    validity tensor."""
    w = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[_width(col)]
    v = col.data[lo * _width(col): hi * _width(col)].view(w).to(torch.int64)
    if _width(col) == 4:
        v = v & 0xFFFFFFFF
    elif _width(col) == 2:
        v = v & 0xFFFF
    if col.validity is None:
        ok = torch.ones(hi - lo, dtype=torch.bool, device=v.device)
    else:
        idx = torch.arange(lo, hi, device=v.device)
        ok = ((col.validity[idx >> 3].to(torch.int64) >> (idx & 7)) & 1).bool()
    return v, ok


def _width(col):
    from legate_dataframe import _capi as C

    return C.WIDTH[col.type_id]


def table_fingerprint(cols, offsets=None, chunk=1 << 26):
    """Per segment of rows (one segment, or offsets[i]:offsets[i+1]): the row count and the
    wrap-around sums of two independent 64-bit hashes of each row's values (null-aware).  Equal
    fingerprints mean equal row multisets up to a hash collision.  Returns an int64 array
    [nseg, 3]."""
    import torch

    n = cols[0].length
    if offsets is None:
        offsets = np.array([0, n], np.int64)
    nseg = len(offsets) - 1
    dev = cols[0].data.device
    out = torch.zeros((nseg, 3), dtype=torch.int64, device=dev)
    offs_t = torch.from_numpy(np.asarray(offsets, np.int64)).to(dev)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        h = torch.zeros(hi - lo, dtype=torch.int64, device=dev)
        for j, c in enumerate(cols):
            v, ok = column_values_i64(torch, c, lo, hi)
            v = torch.where(ok, v, torch.full_like(v, 0x5BD1E995 + j))
            h = _mix(torch, h * 31 + v + _s64(_G * (j + 1)) + ok.to(torch.int64))
        h2 = _mix(torch, h ^ 0x2545F4914F6CDD1D)
        seg = torch.searchsorted(offs_t, torch.arange(lo, hi, device=dev), right=True) - 1
        out[:, 0].index_add_(0, seg, torch.ones_like(h))
        out[:, 1].index_add_(0, seg, h)
        out[:, 2].index_add_(0, seg, h2)
    return out.cpu().numpy()


def device_columns_from_numpy(arrays):
    """[(values, type_id, valid or None)] -> DeviceColumns (for fingerprints of host tables)."""
    from legate_dataframe import _capi as C

    return [C.column_from_numpy(v, t, ok) for v, t, ok in arrays]
