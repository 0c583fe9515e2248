"""ctypes binding of the C ABI in include/ldf_hip.h (libldf_hip.so).

Device memory is plumbing here: buffers are torch tensors on the current HIP device, the
library allocates its outputs and temporaries through callbacks that hand out torch tensors
(so the caching allocator serves them), and calls run on torch's current stream.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence

_PKG = Path(__file__).resolve().parent.parent
_LIB_PATH = _PKG / "libldf_hip.so"

# type ids (arrow::Type::type values)
BOOL, UINT8, INT8, UINT16, INT16, UINT32, INT32, UINT64, INT64 = 1, 2, 3, 4, 5, 6, 7, 8, 9
FLOAT, DOUBLE, DATE32, TIMESTAMP, LARGE_STRING = 11, 12, 16, 18, 34

WIDTH = {BOOL: 1, UINT8: 1, INT8: 1, UINT16: 2, INT16: 2, UINT32: 4, INT32: 4, UINT64: 8, INT64: 8,
         FLOAT: 4, DOUBLE: 8, DATE32: 4, TIMESTAMP: 8, LARGE_STRING: 0}

AGG = {"sum": 0, "product": 1, "min": 2, "max": 3, "count": 4, "mean": 5, "variance": 6,
       "stddev": 7, "approximate_median": 8, "count_distinct": 9, "count_all": 10}
COUNT_MODE = {"only_valid": 0, "only_null": 1, "all": 2}

EXPORTED = [
    "ldf_join", "ldf_groupby", "ldf_hash_partition", "ldf_hash_partition_noalloc", "ldf_murmur3_hash", "ldf_distinct",
    "ldf_bools_to_bitmask", "ldf_bitmask_to_bools", "ldf_ranges_to_offsets",
    "ldf_offsets_to_ranges", "ldf_free_table", "ldf_last_error", "ldf_version",
    "ldf_profile_enable", "ldf_profile_reset", "ldf_profile_get", "ldf_pool_release", "ldf_pool_alloc",
    "ldf_pool_free", "ldf_pool_stats", "ldf_hbm_probe",
    "ldf_batch_join_create", "ldf_batch_groupby_create", "ldf_batch_run", "ldf_batch_result",
    "ldf_batch_config_used", "ldf_batch_destroy", "ldf_server_create", "ldf_server_add_table",
    "ldf_server_submit", "ldf_server_wait", "ldf_server_release", "ldf_server_destroy",
]


class ColumnView(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("reserved", ctypes.c_int32), ("length", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("validity", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("null_count", ctypes.c_int64)]


class TableView(ctypes.Structure):
    _fields_ = [("ncols", ctypes.c_int32), ("cols", ctypes.POINTER(ColumnView))]


class Column(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("reserved", ctypes.c_int32), ("length", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("validity", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("null_count", ctypes.c_int64), ("data_bytes", ctypes.c_int64), ("validity_bytes", ctypes.c_int64)]


class TableOut(ctypes.Structure):
    _fields_ = [("ncols", ctypes.c_int32), ("cols", ctypes.POINTER(Column))]


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


class Allocator(ctypes.Structure):
    _fields_ = [("alloc", ALLOC_FN), ("free", FREE_FN), ("ctx", ctypes.c_void_p)]


class Agg(ctypes.Structure):
    _fields_ = [("in_col", ctypes.c_int32), ("kind", ctypes.c_int32), ("count_mode", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("load_factor", ctypes.c_double), ("tile_rows", ctypes.c_int32), ("lds_slots", ctypes.c_int32),
                ("use_graph", ctypes.c_int32)]


class LdfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib_path() -> Path:
    return _LIB_PATH


def load(require: bool = True):
    """Loads libldf_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            if not require:
                return None
            raise ImportError(f"{_LIB_PATH} is missing: run `python legate-dataframe_amd/build.py` first")
        import torch  # noqa: F401  (its HIP runtime must be the one the library binds to)

        L = ctypes.CDLL(str(_LIB_PATH))
        P = ctypes.POINTER
        L.ldf_join.argtypes = [P(TableView), P(TableView), P(ctypes.c_int32), P(ctypes.c_int32), ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32, P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_int32),
                               ctypes.c_int32, P(Allocator), ctypes.c_void_p, P(TableOut)]
        L.ldf_groupby.argtypes = [P(TableView), P(ctypes.c_int32), ctypes.c_int32, P(Agg), ctypes.c_int32,
                                  P(Allocator), ctypes.c_void_p, P(TableOut)]
        L.ldf_hash_partition.argtypes = [P(TableView), P(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_uint32, P(Allocator), ctypes.c_void_p, P(TableOut),
                                         P(ctypes.c_int64)]
        L.ldf_distinct.argtypes = [P(TableView), P(ctypes.c_int32), ctypes.c_int32, P(Allocator), ctypes.c_void_p,
                                   P(TableOut)]
        L.ldf_hash_partition_noalloc.argtypes = [P(TableView), P(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_uint32, ctypes.c_void_p, P(TableOut), P(ctypes.c_int64)]
        L.ldf_murmur3_hash.argtypes = [P(TableView), P(ctypes.c_int32), ctypes.c_int32, ctypes.c_uint32,
                                       ctypes.c_void_p, ctypes.c_void_p]
        L.ldf_bools_to_bitmask.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.ldf_bitmask_to_bools.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.ldf_ranges_to_offsets.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.ldf_offsets_to_ranges.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p]
        L.ldf_free_table.argtypes = [P(TableOut), P(Allocator), ctypes.c_void_p]
        L.ldf_pool_alloc.argtypes = [ctypes.c_size_t, ctypes.c_void_p]
        L.ldf_pool_alloc.restype = ctypes.c_void_p
        L.ldf_pool_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ldf_pool_free.restype = None
        L.ldf_pool_stats.argtypes = [P(ctypes.c_int64)] * 4
        L.ldf_pool_stats.restype = None
        L.ldf_hbm_probe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    P(ctypes.c_double), ctypes.c_void_p]
        L.ldf_hbm_probe.restype = ctypes.c_int
        L.ldf_profile_enable.argtypes = [ctypes.c_int]
        L.ldf_profile_get.argtypes = [ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_double)]
        L.ldf_profile_get.restype = ctypes.c_int
        L.ldf_batch_join_create.argtypes = [ctypes.c_int32, P(TableView), P(TableView), P(ctypes.c_int32),
                                            P(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_int32), ctypes.c_int32,
                                            P(BatchConfig), ctypes.c_void_p, P(ctypes.c_void_p)]
        L.ldf_batch_groupby_create.argtypes = [ctypes.c_int32, P(TableView), P(ctypes.c_int32), ctypes.c_int32,
                                               P(Agg), ctypes.c_int32, P(BatchConfig), ctypes.c_void_p,
                                               P(ctypes.c_void_p)]
        L.ldf_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ldf_batch_result.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(TableOut), P(ctypes.c_int64)]
        L.ldf_batch_config_used.argtypes = [ctypes.c_void_p, P(BatchConfig)]
        L.ldf_batch_destroy.argtypes = [ctypes.c_void_p]
        L.ldf_batch_destroy.restype = None
        for f in ("ldf_batch_join_create", "ldf_batch_groupby_create", "ldf_batch_run", "ldf_batch_result",
                  "ldf_batch_config_used"):
            getattr(L, f).restype = ctypes.c_int
        L.ldf_last_error.restype = ctypes.c_char_p
        L.ldf_version.restype = ctypes.c_char_p
        for f in ("ldf_join", "ldf_groupby", "ldf_hash_partition", "ldf_hash_partition_noalloc", "ldf_murmur3_hash",
                  "ldf_distinct",
                  "ldf_bools_to_bitmask",
                  "ldf_bitmask_to_bools", "ldf_ranges_to_offsets", "ldf_offsets_to_ranges"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ldf_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        raise LdfError(rc, msg)


class TorchAllocator:
    """ldf_allocator whose buffers are torch uint8 tensors (kept alive until freed / adopted)."""

    def __init__(self, device=None):
        import torch

        self.torch = torch
        self.device = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.live: Dict[int, "torch.Tensor"] = {}
        self._alloc = ALLOC_FN(self._do_alloc)
        self._free = FREE_FN(self._do_free)
        self.struct = Allocator(self._alloc, self._free, None)

    def _do_alloc(self, ctx, nbytes, stream):
        try:
            t = self.torch.empty(int(nbytes), dtype=self.torch.uint8, device=self.device)
        except Exception:  # out of memory -> NULL -> LDF_E_OOM
            return None
        p = t.data_ptr()
        self.live[p] = t
        return p

    def _do_free(self, ctx, ptr, nbytes, stream):
        self.live.pop(ptr, None)

    def adopt(self, ptr: int) -> "torch.Tensor":
        """Takes ownership of a buffer the library returned (removes it from the live set)."""
        return self.live.pop(ptr)


def stream_ptr() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream


def make_table_view(cols: Sequence[ColumnView]):
    arr = (ColumnView * max(1, len(cols)))(*cols)
    return TableView(len(cols), arr), arr


def int32_array(xs: Sequence[int]):
    xs = list(xs)
    return (ctypes.c_int32 * max(1, len(xs)))(*xs)


# ---------------------------------------------------------------------------------------------
# device columns as torch tensors (used by the bindings' tests and the benchmark)
# ---------------------------------------------------------------------------------------------
class DeviceColumn:
    """A column resident in HBM: torch uint8/typed tensors for data, validity and offsets."""

    def __init__(self, type_id: int, length: int, data, validity=None, offsets=None, null_count: int = -1):
        self.type_id = type_id
        self.length = length
        self.data = data
        self.validity = validity
        self.offsets = offsets
        self.null_count = null_count

    def view(self) -> ColumnView:
        return ColumnView(self.type_id, 0, self.length,
                          self.data.data_ptr() if self.data is not None else None,
                          self.validity.data_ptr() if self.validity is not None else None,
                          self.offsets.data_ptr() if self.offsets is not None else None,
                          self.null_count)

    @property
    def nullable(self) -> bool:
        return self.validity is not None


_NP_OF = {BOOL: "uint8", UINT8: "uint8", INT8: "int8", UINT16: "uint16", INT16: "int16", UINT32: "uint32",
          INT32: "int32", UINT64: "uint64", INT64: "int64", FLOAT: "float32", DOUBLE: "float64",
          DATE32: "int32", TIMESTAMP: "int64"}


def np_dtype(type_id: int):
    import numpy as np

    return np.dtype(_NP_OF[type_id])


def column_from_numpy(values, type_id: int, valid=None, device=None) -> DeviceColumn:
    """Fixed-width column from a numpy array (+ optional bool validity array)."""
    import numpy as np
    import torch

    dev = device or torch.device("cuda", torch.cuda.current_device())
    v = np.ascontiguousarray(np.asarray(values).astype(np_dtype(type_id), copy=False))
    data = torch.from_numpy(v.view(np.uint8).reshape(-1).copy()).to(dev)
    validity = None
    nc = 0
    if valid is not None:
        valid = np.asarray(valid, dtype=bool)
        bits = np.packbits(valid, bitorder="little")
        pad = (-len(bits)) % 8
        bits = np.concatenate([bits, np.zeros(pad + 8, np.uint8)])
        validity = torch.from_numpy(bits).to(dev)
        nc = int((~valid).sum())
    return DeviceColumn(type_id, len(v), data, validity, None, nc)


def column_from_strings(strs, valid=None, device=None) -> DeviceColumn:
    import numpy as np
    import torch

    dev = device or torch.device("cuda", torch.cuda.current_device())
    enc = [("" if s is None else s).encode() for s in strs]
    offs = np.zeros(len(enc) + 1, np.int64)
    offs[1:] = np.cumsum([len(e) for e in enc])
    chars = np.frombuffer(b"".join(enc) + b"\0", np.uint8).copy()
    if valid is None and any(s is None for s in strs):
        valid = [s is not None for s in strs]
    col = DeviceColumn(LARGE_STRING, len(enc), torch.from_numpy(chars).to(dev), None,
                       torch.from_numpy(offs).to(dev), 0)
    if valid is not None:
        valid = np.asarray(valid, dtype=bool)
        bits = np.packbits(valid, bitorder="little")
        col.validity = torch.from_numpy(np.concatenate([bits, np.zeros(8, np.uint8)])).to(dev)
        col.null_count = int((~valid).sum())
    return col


def column_from_arrow(arr, device=None) -> DeviceColumn:
    """A pyarrow array (fixed width, timestamp or string) as a DeviceColumn in HBM."""
    import numpy as np
    import pyarrow as pa

    if isinstance(arr, pa.ChunkedArray):
        arr = arr.combine_chunks()
    valid = np.asarray(arr.is_valid()) if arr.null_count else None
    if pa.types.is_string(arr.type) or pa.types.is_large_string(arr.type):
        return column_from_strings(arr.to_pylist(), valid, device)
    t = arr.type
    if pa.types.is_timestamp(t):
        tid, vals = TIMESTAMP, arr.cast(pa.int64()).fill_null(0).to_numpy()
    elif t == pa.bool_():
        tid, vals = BOOL, np.asarray(arr.fill_null(False)).astype(np.uint8)
    else:
        tid = {pa.uint8(): UINT8, pa.int8(): INT8, pa.uint16(): UINT16, pa.int16(): INT16, pa.uint32(): UINT32,
               pa.int32(): INT32, pa.uint64(): UINT64, pa.int64(): INT64, pa.float32(): FLOAT, pa.float64(): DOUBLE,
               pa.date32(): DATE32}[t]
        vals = arr.fill_null(0).to_numpy(zero_copy_only=False) if arr.null_count else arr.to_numpy(zero_copy_only=False)
    return column_from_numpy(vals, tid, valid, device)


def columns_from_parquet(files, columns=None, device=None) -> List[DeviceColumn]:
    """Parquet files (read on the host with pyarrow) as DeviceColumns in HBM, e.g. the tables a
    served-form Server registers with add_table."""
    from legate_dataframe.lib.parquet import read_arrow

    t = read_arrow(files, columns=columns)
    return [column_from_arrow(t.column(i), device) for i in range(t.num_columns)]


def column_to_numpy(c: DeviceColumn):
    """(values, valid) on the host; strings come back as a list of str (None for null)."""
    import numpy as np

    n = c.length
    valid = None
    if c.validity is not None:
        bits = c.validity.cpu().numpy()
        valid = np.unpackbits(bits, bitorder="little")[:n].astype(bool)
    if c.type_id == LARGE_STRING:
        offs = c.offsets.cpu().numpy()
        chars = c.data.cpu().numpy().tobytes()
        vals = [chars[offs[i]:offs[i + 1]].decode() for i in range(n)]
        if valid is not None:
            vals = [v if ok else None for v, ok in zip(vals, valid)]
        return vals, valid
    raw = c.data.cpu().numpy()[: n * WIDTH[c.type_id]]
    return raw.view(np_dtype(c.type_id)).copy(), valid


def _adopt_columns(alloc: TorchAllocator, out: TableOut) -> List[DeviceColumn]:
    import torch

    cols = []
    for i in range(out.ncols):
        c = out.cols[i]
        data = alloc.adopt(c.data) if c.data else None
        if data is not None:  # allocations are rounded up; expose the exact byte length
            data = data[: c.data_bytes if c.type_id == LARGE_STRING else c.length * WIDTH[c.type_id]]
        validity = alloc.adopt(c.validity) if c.validity else None
        offsets = None
        if c.type_id == LARGE_STRING and c.offsets:
            offsets = alloc.adopt(c.offsets)[: (c.length + 1) * 8].view(torch.int64)
        cols.append(DeviceColumn(c.type_id, c.length, data, validity, offsets, c.null_count))
    return cols


def join(lhs: Sequence[DeviceColumn], rhs: Sequence[DeviceColumn], lhs_keys, rhs_keys, join_type: int,
         nulls_equal: bool = True, lhs_out=None, rhs_out=None, alloc: Optional[TorchAllocator] = None):
    L = load()
    alloc = alloc or TorchAllocator()
    lhs_out = list(range(len(lhs))) if lhs_out is None else list(lhs_out)
    if rhs_out is None:
        rhs_out = [] if join_type in (3, 4) else list(range(len(rhs)))
    lv, _la = make_table_view([c.view() for c in lhs])
    rv, _ra = make_table_view([c.view() for c in rhs])
    ncols = len(lhs_out) + len(rhs_out)
    cols = (Column * max(1, ncols))()
    out = TableOut(ncols, cols)
    check(L.ldf_join(ctypes.byref(lv), ctypes.byref(rv), int32_array(lhs_keys), int32_array(rhs_keys), len(lhs_keys),
                     join_type, int(bool(nulls_equal)), int32_array(lhs_out), len(lhs_out), int32_array(rhs_out),
                     len(rhs_out), ctypes.byref(alloc.struct), stream_ptr(), ctypes.byref(out)))
    return _adopt_columns(alloc, out)


def groupby(table: Sequence[DeviceColumn], keys, aggs, alloc: Optional[TorchAllocator] = None):
    """aggs: list of (in_col, kind_name, count_mode_name)."""
    L = load()
    alloc = alloc or TorchAllocator()
    tv, _ta = make_table_view([c.view() for c in table])
    arr = (Agg * max(1, len(aggs)))(*[Agg(-1 if c is None else c, AGG[k], COUNT_MODE[m or "only_valid"], 0)
                                      for c, k, m in aggs])
    ncols = len(keys) + len(aggs)
    cols = (Column * ncols)()
    out = TableOut(ncols, cols)
    check(L.ldf_groupby(ctypes.byref(tv), int32_array(keys), len(keys), arr, len(aggs), ctypes.byref(alloc.struct),
                        stream_ptr(), ctypes.byref(out)))
    return _adopt_columns(alloc, out)


def distinct(table: Sequence[DeviceColumn], keys, alloc: Optional[TorchAllocator] = None):
    """Distinct rows by the `keys` columns: every column, the first row of each key tuple."""
    L = load()
    alloc = alloc or TorchAllocator()
    tv, _ta = make_table_view([c.view() for c in table])
    outc = (Column * max(1, len(table)))()
    out = TableOut(len(table), outc)
    check(L.ldf_distinct(ctypes.byref(tv), int32_array(keys), len(keys), ctypes.byref(alloc.struct), stream_ptr(),
                         ctypes.byref(out)))
    return _adopt_columns(alloc, out)


def hash_partition(table: Sequence[DeviceColumn], cols, nparts: int, seed: int = 0,
                   alloc: Optional[TorchAllocator] = None):
    L = load()
    alloc = alloc or TorchAllocator()
    tv, _ta = make_table_view([c.view() for c in table])
    outc = (Column * max(1, len(table)))()
    out = TableOut(len(table), outc)
    offs = (ctypes.c_int64 * (nparts + 1))()
    check(L.ldf_hash_partition(ctypes.byref(tv), int32_array(cols), len(cols), nparts, seed,
                               ctypes.byref(alloc.struct), stream_ptr(), ctypes.byref(out), offs))
    return _adopt_columns(alloc, out), list(offs)


def murmur3_hash(table: Sequence[DeviceColumn], cols, seed: int = 0):
    import torch

    L = load()
    n = table[0].length if table else 0
    tv, _ta = make_table_view([c.view() for c in table])
    out = torch.empty(max(1, n), dtype=torch.int32, device=table[0].data.device)
    check(L.ldf_murmur3_hash(ctypes.byref(tv), int32_array(cols), len(cols), seed, out.data_ptr(), stream_ptr()))
    return out[:n]


# ---------------------------------------------------------------------------------------------
# library-default allocator (stream-ordered pool, no Python in the allocation path)
# ---------------------------------------------------------------------------------------------
class RawTable:
    """Output table whose buffers belong to the library's default pool; free() returns them."""

    def __init__(self, ncols: int):
        self.cols = (Column * max(1, ncols))()
        self.out = TableOut(ncols, self.cols)

    def free(self):
        load().ldf_free_table(ctypes.byref(self.out), None, stream_ptr())
        self.out.ncols = 0


def groupby_raw(table: Sequence[DeviceColumn], keys, aggs) -> RawTable:
    L = load()
    tv, _ta = make_table_view([c.view() for c in table])
    arr = (Agg * max(1, len(aggs)))(*[Agg(-1 if c is None else c, AGG[k], COUNT_MODE[m or "only_valid"], 0)
                                      for c, k, m in aggs])
    rt = RawTable(len(keys) + len(aggs))
    check(L.ldf_groupby(ctypes.byref(tv), int32_array(keys), len(keys), arr, len(aggs), None, stream_ptr(),
                        ctypes.byref(rt.out)))
    return rt


def join_raw(lhs, rhs, lhs_keys, rhs_keys, join_type: int, nulls_equal=True, lhs_out=None, rhs_out=None) -> RawTable:
    L = load()
    lhs_out = list(range(len(lhs))) if lhs_out is None else list(lhs_out)
    if rhs_out is None:
        rhs_out = [] if join_type in (3, 4) else list(range(len(rhs)))
    lv, _la = make_table_view([c.view() for c in lhs])
    rv, _ra = make_table_view([c.view() for c in rhs])
    rt = RawTable(len(lhs_out) + len(rhs_out))
    check(L.ldf_join(ctypes.byref(lv), ctypes.byref(rv), int32_array(lhs_keys), int32_array(rhs_keys), len(lhs_keys),
                     join_type, int(bool(nulls_equal)), int32_array(lhs_out), len(lhs_out), int32_array(rhs_out),
                     len(rhs_out), None, stream_ptr(), ctypes.byref(rt.out)))
    return rt


def hash_partition_raw(table, cols, nparts: int, seed: int = 0):
    L = load()
    tv, _ta = make_table_view([c.view() for c in table])
    rt = RawTable(len(table))
    offs = (ctypes.c_int64 * (nparts + 1))()
    check(L.ldf_hash_partition(ctypes.byref(tv), int32_array(cols), len(cols), nparts, seed, None, stream_ptr(),
                               ctypes.byref(rt.out), offs))
    return rt, list(offs)


def hash_partition_noalloc(table, cols, nparts: int, out_data, seed: int = 0):
    """ldf_hash_partition_noalloc: `out_data` holds one caller-allocated uint8 tensor per column
    (at least rows * width bytes); returns the partitioned columns (views of those tensors) and the
    nparts + 1 offsets."""
    L = load()
    tv, _ta = make_table_view([c.view() for c in table])
    cols_out = (Column * max(1, len(table)))()
    for i, t in enumerate(out_data):
        cols_out[i].data = t.data_ptr()
        cols_out[i].data_bytes = t.numel()
    out = TableOut(len(table), cols_out)
    offs = (ctypes.c_int64 * (nparts + 1))()
    check(L.ldf_hash_partition_noalloc(ctypes.byref(tv), int32_array(cols), len(cols), nparts, seed, stream_ptr(),
                                       ctypes.byref(out), offs))
    res = [DeviceColumn(cols_out[i].type_id, cols_out[i].length, out_data[i][: cols_out[i].length * WIDTH[cols_out[i].type_id]])
           for i in range(out.ncols)]
    return res, list(offs)


def pool_release() -> None:
    """Returns the default allocator's cached device blocks to the device."""
    load().ldf_pool_release()


def pool_stats() -> dict:
    """Blocks in use and cached blocks (with bytes) of the default allocator."""
    v = [ctypes.c_int64() for _ in range(4)]
    load().ldf_pool_stats(*[ctypes.byref(x) for x in v])
    return dict(live_blocks=v[0].value, live_bytes=v[1].value, cached_blocks=v[2].value, cached_bytes=v[3].value)


def hbm_probe(src, dst, nbytes: int, copy: bool, nblocks: int) -> float:
    """Milliseconds of one read (copy=False) or copy (copy=True) of `nbytes` at device tensor
    `src` (into `dst`) by the library's HBM probe kernels, on the current torch stream."""
    ms = ctypes.c_double()
    check(load().ldf_hbm_probe(src.data_ptr(), dst.data_ptr(), nbytes, int(copy), nblocks, ctypes.byref(ms),
                               stream_ptr()))
    return ms.value


def profile_enable(on: bool = True) -> None:
    load().ldf_profile_enable(int(on))


def profile_reset() -> None:
    load().ldf_profile_reset()


def profile_get(kernel: str):
    """(launches, total ms) recorded with HIP events for `kernel`."""
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    load().ldf_profile_get(kernel.encode(), ctypes.byref(n), ctypes.byref(ms))
    return n.value, ms.value


# ---------------------------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------------------------
class _DevBuf:
    """A device buffer owned by the library, exposed to torch without a copy."""

    def __init__(self, ptr: int, nbytes: int):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False), "version": 3}


def _view_bytes(ptr: int, nbytes: int):
    import torch

    if not ptr or nbytes <= 0:
        return torch.empty(0, dtype=torch.uint8, device="cuda")
    return torch.as_tensor(_DevBuf(ptr, nbytes), device="cuda")


def _out_view(c) -> DeviceColumn:
    """A DeviceColumn viewing an output column whose buffers the library keeps owning."""
    import torch

    offsets = None
    if c.type_id == LARGE_STRING:
        data = _view_bytes(c.data, c.data_bytes)
        offsets = _view_bytes(c.offsets, (c.length + 1) * 8).view(torch.int64)
    else:
        data = _view_bytes(c.data, c.length * WIDTH[c.type_id])
    valid = _view_bytes(c.validity, c.validity_bytes) if c.validity else None
    return DeviceColumn(c.type_id, c.length, data, valid, offsets, c.null_count)


def _config(config) -> Optional[BatchConfig]:
    if config is None:
        return None
    c = config if isinstance(config, dict) else dict(config)
    return BatchConfig(float(c.get("load_factor", 0.0)), int(c.get("tile_rows", 0)), int(c.get("lds_slots", 0)),
                       int(c.get("use_graph", 1)))


class Batch:
    """A planned batch of joins or groupbys (ldf_batch_*): run() enqueues it, result() waits and
    returns the concatenated output columns (views of the batch's buffers) and the per-instance
    row offsets."""

    def __init__(self, handle, ncols: int, ninst: int, keep):
        self.h = handle
        self.ncols = ncols
        self.ninst = ninst
        self._keep = keep

    def run(self) -> None:
        check(load().ldf_batch_run(self.h, stream_ptr()))

    def result(self):
        import numpy as np

        cols = (Column * max(1, self.ncols))()
        out = TableOut(self.ncols, cols)
        offs = (ctypes.c_int64 * (self.ninst + 1))()
        check(load().ldf_batch_result(self.h, stream_ptr(), ctypes.byref(out), offs))
        res = [_out_view(out.cols[i]) for i in range(out.ncols)]
        return res, np.frombuffer(offs, dtype=np.int64).copy()

    def config_used(self) -> dict:
        c = BatchConfig()
        check(load().ldf_batch_config_used(self.h, ctypes.byref(c)))
        return {"load_factor": c.load_factor, "tile_rows": c.tile_rows, "lds_slots": c.lds_slots,
                "use_graph": c.use_graph}

    def close(self) -> None:
        if self.h:
            load().ldf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _table_views(tables):
    keep = []
    arr = (TableView * len(tables))()
    for i, t in enumerate(tables):
        tv, a = make_table_view([c.view() for c in t])
        arr[i] = tv
        keep.append(a)
    return arr, keep


def batch_join(lhs_tables, rhs_tables, lhs_keys, rhs_keys, join_type: int, nulls_equal: bool = True,
               lhs_out=None, rhs_out=None, config=None) -> Batch:
    L = load()
    assert len(lhs_tables) == len(rhs_tables) and len(lhs_tables) > 0
    lhs_out = list(range(len(lhs_tables[0]))) if lhs_out is None else list(lhs_out)
    if rhs_out is None:
        rhs_out = [] if join_type in (3, 4) else list(range(len(rhs_tables[0])))
    la, lk = _table_views(lhs_tables)
    ra, rk = _table_views(rhs_tables)
    cfg = _config(config)
    h = ctypes.c_void_p()
    check(L.ldf_batch_join_create(len(lhs_tables), la, ra, int32_array(lhs_keys), int32_array(rhs_keys),
                                  len(lhs_keys), join_type, int(bool(nulls_equal)), int32_array(lhs_out),
                                  len(lhs_out), int32_array(rhs_out), len(rhs_out),
                                  ctypes.byref(cfg) if cfg is not None else None, stream_ptr(), ctypes.byref(h)))
    return Batch(h, len(lhs_out) + len(rhs_out), len(lhs_tables), (lhs_tables, rhs_tables, la, ra, lk, rk))


def batch_groupby(tables, keys, aggs, config=None) -> Batch:
    """aggs: list of (in_col, kind_name, count_mode_name) as for groupby()."""
    L = load()
    ta, tk = _table_views(tables)
    arr = (Agg * max(1, len(aggs)))(*[Agg(-1 if c is None else c, AGG[k], COUNT_MODE[m or "only_valid"], 0)
                                      for c, k, m in aggs])
    cfg = _config(config)
    h = ctypes.c_void_p()
    check(L.ldf_batch_groupby_create(len(tables), ta, int32_array(keys), len(keys), arr, len(aggs),
                                     ctypes.byref(cfg) if cfg is not None else None, stream_ptr(), ctypes.byref(h)))
    return Batch(h, len(keys) + len(aggs), len(tables), (tables, ta, tk))


# ---------------------------------------------------------------------------------------------
# served form
# ---------------------------------------------------------------------------------------------
QUERY_JOIN, QUERY_GROUPBY = 0, 1


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("slots", ctypes.c_int32), ("max_batch", ctypes.c_int32)]


class Query(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("table", ctypes.c_int32), ("probe", TableView),
                ("probe_key", ctypes.c_int32), ("naggs", ctypes.c_int32), ("aggs", ctypes.POINTER(Agg)),
                ("row_begin", ctypes.c_int64), ("row_end", ctypes.c_int64)]


class QueryStats(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_int64), ("start_ns", ctypes.c_int64), ("end_ns", ctypes.c_int64),
                ("batch_queries", ctypes.c_int32), ("plain", ctypes.c_int32)]


def _server_lib():
    L = load()
    if not getattr(L, "_server_types", False):
        P = ctypes.POINTER
        L.ldf_server_create.argtypes = [P(ServerConfig), P(ctypes.c_void_p)]
        L.ldf_server_add_table.argtypes = [ctypes.c_void_p, P(TableView), ctypes.c_int32, P(ctypes.c_int32)]
        L.ldf_server_submit.argtypes = [ctypes.c_void_p, P(Query), P(ctypes.c_int64)]
        L.ldf_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64, P(TableOut), P(QueryStats)]
        L.ldf_server_release.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ldf_server_destroy.argtypes = [ctypes.c_void_p]
        L.ldf_server_destroy.restype = None
        for f in ("ldf_server_create", "ldf_server_add_table", "ldf_server_submit", "ldf_server_wait",
                  "ldf_server_release"):
            getattr(L, f).restype = ctypes.c_int
        L._server_types = True
    return L


class Server:
    """Query server over device-resident tables (ldf_server_*): submit_* return a query id at once;
    wait() blocks for the result (views of server-owned buffers, valid until release())."""

    def __init__(self, hbm_budget_bytes: float = 200e9, slots: int = 4, max_batch: int = 64):
        L = _server_lib()
        self.h = ctypes.c_void_p()
        check(L.ldf_server_create(ctypes.byref(ServerConfig(float(hbm_budget_bytes), slots, max_batch)),
                                  ctypes.byref(self.h)))
        self._keep = {}
        self._tables = []

    def add_table(self, cols, key_col: int = 0) -> int:
        tv, arr = make_table_view([c.view() for c in cols])
        tid = ctypes.c_int32()
        check(_server_lib().ldf_server_add_table(self.h, ctypes.byref(tv), key_col, ctypes.byref(tid)))
        self._tables.append((cols, arr))
        return tid.value

    def _submit(self, q: Query, keep) -> int:
        qid = ctypes.c_int64()
        check(_server_lib().ldf_server_submit(self.h, ctypes.byref(q), ctypes.byref(qid)))
        self._keep[qid.value] = keep
        return qid.value

    def submit_join(self, table: int, probe_cols, probe_key: int = 0) -> int:
        tv, arr = make_table_view([c.view() for c in probe_cols])
        q = Query(QUERY_JOIN, table, tv, probe_key, 0, None, 0, 0)
        return self._submit(q, (probe_cols, arr))

    def submit_groupby(self, table: int, row_begin: int, row_end: int, aggs) -> int:
        arr = (Agg * max(1, len(aggs)))(*[Agg(-1 if c is None else c, AGG[k], COUNT_MODE[m or "only_valid"], 0)
                                          for c, k, m in aggs])
        q = Query(QUERY_GROUPBY, table, TableView(0, None), 0, len(aggs), arr, row_begin, row_end)
        return self._submit(q, arr)

    def wait(self, qid: int):
        cols = (Column * 40)()
        out = TableOut(40, cols)
        st = QueryStats()
        check(_server_lib().ldf_server_wait(self.h, qid, ctypes.byref(out), ctypes.byref(st)))
        res = [_out_view(out.cols[i]) for i in range(out.ncols)]
        stats = {"submit_ns": st.submit_ns, "start_ns": st.start_ns, "end_ns": st.end_ns,
                 "batch_queries": st.batch_queries, "plain": st.plain}
        return res, stats

    def release(self, qid: int) -> None:
        check(_server_lib().ldf_server_release(self.h, qid))
        self._keep.pop(qid, None)

    def close(self) -> None:
        if self.h:
            _server_lib().ldf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
