"""ORACLE -- test infrastructure only (parity checker and CPU baseline).

ctypes wrapper of liboracle.so, the single-threaded C restatement of the reference's CPU task
variants for the join / groupby / hash-partition path (see oracle.c for the file:line each
function follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package; the product never imports it.  `distinct_first_rows` (pure Python, small inputs) is
the restatement for `distinct`; it is pinned by the reference's own test case and by
pandas.DataFrame.duplicated(keep="first") in tests/test_distinct_checker.py.

Columns are given as :class:`OCol` (numpy values or a list of str, an optional numpy bool
validity array, and the Arrow type id).
"""

from __future__ import annotations

import ctypes
import struct
import subprocess
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "liboracle.so"

T_BOOL, T_UINT8, T_INT8, T_UINT16, T_INT16, T_UINT32, T_INT32, T_UINT64, T_INT64 = 1, 2, 3, 4, 5, 6, 7, 8, 9
T_FLOAT, T_DOUBLE, T_DATE32, T_TIMESTAMP, T_LARGE_STRING = 11, 12, 16, 18, 34
_NP = {T_BOOL: np.uint8, T_UINT8: np.uint8, T_INT8: np.int8, T_UINT16: np.uint16, T_INT16: np.int16,
       T_UINT32: np.uint32, T_INT32: np.int32, T_UINT64: np.uint64, T_INT64: np.int64, T_FLOAT: np.float32,
       T_DOUBLE: np.float64, T_DATE32: np.int32, T_TIMESTAMP: np.int64}
AGG = {"sum": 0, "product": 1, "min": 2, "max": 3, "count": 4, "mean": 5, "variance": 6, "stddev": 7,
       "approximate_median": 8, "count_distinct": 9, "count_all": 10}
MODE = {"only_valid": 0, "only_null": 1, "all": 2}
JOIN = {"inner": 0, "left": 1, "full": 2, "semi": 3, "anti": 4}


class _Col(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("width", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("valid", ctypes.c_void_p), ("offsets", ctypes.c_void_p)]


class _Agg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mode", ctypes.c_int32), ("input", _Col)]


class _AggOut(ctypes.Structure):
    _fields_ = [("bits", ctypes.POINTER(ctypes.c_uint64)), ("f64", ctypes.POINTER(ctypes.c_double)),
                ("valid", ctypes.POINTER(ctypes.c_uint8))]


class OCol:
    """A host column for the oracle."""

    def __init__(self, type_id: int, values, valid=None):
        self.type_id = type_id
        self.valid = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.uint8))
        if type_id == T_LARGE_STRING:
            enc = [("" if s is None else s).encode() for s in values]
            self.offsets = np.zeros(len(enc) + 1, np.int64)
            self.offsets[1:] = np.cumsum([len(e) for e in enc])
            self.values = np.frombuffer(b"".join(enc) + b"\0", np.uint8).copy()
            self.n = len(enc)
            self.width = 0
        else:
            self.values = np.ascontiguousarray(np.asarray(values).astype(_NP[type_id], copy=False))
            self.offsets = None
            self.n = len(self.values)
            self.width = self.values.dtype.itemsize

    def c(self) -> _Col:
        return _Col(self.type_id, self.width, self.values.ctypes.data,
                    self.valid.ctypes.data if self.valid is not None else None,
                    self.offsets.ctypes.data if self.offsets is not None else None)


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB


def load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build()
        L = ctypes.CDLL(str(_LIB))
        P = ctypes.POINTER
        L.orc_join.restype = ctypes.c_int64
        L.orc_join.argtypes = [P(_Col), P(_Col), ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                               ctypes.c_int, P(P(ctypes.c_int32)), P(P(ctypes.c_int32))]
        L.orc_groupby.restype = ctypes.c_int64
        L.orc_groupby.argtypes = [P(_Col), ctypes.c_int, ctypes.c_int64, P(_Agg), ctypes.c_int, P(P(ctypes.c_int64)),
                                  P(_AggOut)]
        L.orc_free_agg_out.argtypes = [P(_AggOut), ctypes.c_int]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_murmur3_32.restype = ctypes.c_uint32
        L.orc_murmur3_32.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_uint32]
        L.orc_row_hashes.argtypes = [P(_Col), ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_void_p]
        L.orc_partition.restype = ctypes.c_int
        L.orc_partition.argtypes = [P(_Col), ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _cols(cols: Sequence[OCol]):
    return (_Col * max(1, len(cols)))(*[c.c() for c in cols])


def join(lkeys: Sequence[OCol], rkeys: Sequence[OCol], how: str, nulls_equal: bool = True
         ) -> Tuple[np.ndarray, np.ndarray]:
    """Row pairs (lhs row, rhs row) of the join; -1 marks the null-filled side."""
    L = load()
    nl = lkeys[0].n
    nr = rkeys[0].n
    lp = ctypes.POINTER(ctypes.c_int32)()
    rp = ctypes.POINTER(ctypes.c_int32)()
    n = L.orc_join(_cols(lkeys), _cols(rkeys), len(lkeys), nl, nr, JOIN[how], int(nulls_equal), ctypes.byref(lp),
                   ctypes.byref(rp))
    if n < 0:
        raise MemoryError("oracle join failed")
    lm = np.ctypeslib.as_array(lp, shape=(n,)).copy() if n else np.zeros(0, np.int32)
    rm = np.ctypeslib.as_array(rp, shape=(n,)).copy() if n else np.zeros(0, np.int32)
    L.orc_free(ctypes.cast(lp, ctypes.c_void_p))
    L.orc_free(ctypes.cast(rp, ctypes.c_void_p))
    return lm, rm


def groupby(keys: Sequence[OCol], aggs: Sequence[Tuple[Optional[OCol], str, Optional[str]]]):
    """Returns (rep_rows, [(bits uint64, f64, valid bool) per aggregation]); groups are in order
    of first appearance and rep_rows[g] is the first row of group g."""
    L = load()
    n = keys[0].n
    arr = (_Agg * max(1, len(aggs)))()
    for i, (col, kind, mode) in enumerate(aggs):
        arr[i].kind = AGG[kind]
        arr[i].mode = MODE[mode or "only_valid"]
        if col is not None:
            arr[i].input = col.c()
    outs = (_AggOut * max(1, len(aggs)))()
    rep = ctypes.POINTER(ctypes.c_int64)()
    ng = L.orc_groupby(_cols(keys), len(keys), n, arr, len(aggs), ctypes.byref(rep), outs)
    if ng < 0:
        raise MemoryError("oracle groupby failed")
    reps = np.ctypeslib.as_array(rep, shape=(ng,)).copy() if ng else np.zeros(0, np.int64)
    res = []
    for i in range(len(aggs)):
        if ng:
            res.append((np.ctypeslib.as_array(outs[i].bits, shape=(ng,)).copy(),
                        np.ctypeslib.as_array(outs[i].f64, shape=(ng,)).copy(),
                        np.ctypeslib.as_array(outs[i].valid, shape=(ng,)).astype(bool)))
        else:
            res.append((np.zeros(0, np.uint64), np.zeros(0), np.zeros(0, bool)))
    L.orc_free_agg_out(outs, len(aggs))
    L.orc_free(ctypes.cast(rep, ctypes.c_void_p))
    return reps, res


def murmur3_32(data: bytes, seed: int = 0) -> int:
    return int(load().orc_murmur3_32(data, len(data), seed))


def row_hashes(cols: Sequence[OCol], seed: int = 0) -> np.ndarray:
    n = cols[0].n
    out = np.zeros(max(1, n), np.uint32)
    load().orc_row_hashes(_cols(cols), len(cols), n, seed, out.ctypes.data)
    return out[:n]


def partition(cols: Sequence[OCol], nparts: int, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """(perm, offsets): output position i holds input row perm[i]; partition p spans
    offsets[p]:offsets[p+1]."""
    n = cols[0].n
    perm = np.zeros(max(1, n), np.int64)
    offs = np.zeros(nparts + 1, np.int64)
    if load().orc_partition(_cols(cols), len(cols), n, nparts, seed, perm.ctypes.data, offs.ctypes.data):
        raise MemoryError("oracle partition failed")
    return perm[:n], offs


def _distinct_key(v):
    """A key value as cudf::distinct compares it with null_equality::EQUAL and
    nan_equality::ALL_EQUAL (cpp/src/stream_compaction.cu:63-71 in the reference): nulls (None)
    equal, every NaN one key, -0.0 equal to 0.0, other floats by their bits."""
    if isinstance(v, float):
        if v != v:
            return ("nan",)
        return ("f", struct.pack("<d", v + 0.0))  # -0.0 + 0.0 == +0.0
    return v


def distinct_first_rows(key_columns: Sequence[Sequence]) -> List[int]:
    """Row index of each distinct key tuple's first occurrence, in row order: the rows
    DistinctTask keeps (cpp/src/stream_compaction.cu:38-72; KEEP_ANY there, the first row of each
    key here).  `key_columns` are Python lists (pyarrow to_pylist())."""
    seen, out = set(), []
    for i, key in enumerate(zip(*[[_distinct_key(v) for v in c] for c in key_columns])):
        if key not in seen:
            seen.add(key)
            out.append(i)
    return out
