"""LogicalColumn: an HBM-resident column (reference: python/legate_dataframe/lib/core/column.pyx).

Conversion to/from pyarrow keeps Arrow semantics: utf8 is stored as large_utf8 (the reference
normalises the same way), booleans are stored one byte per row (Legate's layout), timestamps
keep their unit.
"""

from __future__ import annotations

from typing import Any, Optional

import numpy as np
import pyarrow as pa

from legate_dataframe._native import lib

# arrow::Type::type ids of the supported types
BOOL, UINT8, INT8, UINT16, INT16, UINT32, INT32, UINT64, INT64 = 1, 2, 3, 4, 5, 6, 7, 8, 9
FLOAT, DOUBLE, DATE32, TIMESTAMP, LARGE_STRING = 11, 12, 16, 18, 34

_FIXED = {
    pa.bool_(): BOOL, pa.uint8(): UINT8, pa.int8(): INT8, pa.uint16(): UINT16, pa.int16(): INT16,
    pa.uint32(): UINT32, pa.int32(): INT32, pa.uint64(): UINT64, pa.int64(): INT64, pa.float32(): FLOAT,
    pa.float64(): DOUBLE, pa.date32(): DATE32,
}
_UNITS = ["s", "ms", "us", "ns"]
_NP = {BOOL: np.uint8, UINT8: np.uint8, INT8: np.int8, UINT16: np.uint16, INT16: np.int16, UINT32: np.uint32,
       INT32: np.int32, UINT64: np.uint64, INT64: np.int64, FLOAT: np.float32, DOUBLE: np.float64,
       DATE32: np.int32, TIMESTAMP: np.int64}


def arrow_to_type_id(t: pa.DataType):
    """(type id, timestamp unit) of a pyarrow type."""
    if t in _FIXED:
        return _FIXED[t], 0
    if pa.types.is_timestamp(t):
        return TIMESTAMP, _UNITS.index(t.unit)
    if pa.types.is_string(t) or pa.types.is_large_string(t):
        return LARGE_STRING, 0
    raise TypeError(f"unsupported column type: {t}")


def type_id_to_arrow(type_id: int, unit: int = 0) -> pa.DataType:
    if type_id == TIMESTAMP:
        return pa.timestamp(_UNITS[unit])
    if type_id == LARGE_STRING:
        return pa.large_string()
    for k, v in _FIXED.items():
        if v == type_id:
            return k
    raise TypeError(f"unknown type id {type_id}")


class LogicalColumn:
    """A column resident in the HBM of the current device."""

    def __init__(self, obj: Any = None, *, _handle=None):
        if _handle is not None:
            self._handle = _handle
        elif isinstance(obj, LogicalColumn):
            self._handle = obj._handle
        elif isinstance(obj, (pa.Array, pa.ChunkedArray)):
            self._handle = LogicalColumn.from_arrow(obj)._handle
        else:
            raise TypeError("LogicalColumn() takes a pyarrow array or a LogicalColumn")

    # ---- construction ----------------------------------------------------------------------
    @staticmethod
    def from_arrow(array) -> "LogicalColumn":
        if isinstance(array, pa.ChunkedArray):
            array = array.combine_chunks() if array.num_chunks != 1 else array.chunk(0)
        tid, unit = arrow_to_type_id(array.type)
        n = len(array)
        validity = None
        if array.null_count > 0:
            validity = np.packbits(np.asarray(array.is_valid()), bitorder="little")
        offsets = None
        if tid == LARGE_STRING:
            array = array.cast(pa.large_string())
            bufs = array.buffers()
            offs = np.frombuffer(bufs[1], np.int64)[array.offset: array.offset + n + 1]
            chars = np.frombuffer(bufs[2], np.uint8)[offs[0]: offs[-1]] if bufs[2] is not None and n else \
                np.zeros(0, np.uint8)
            offsets = (offs - offs[0]).astype(np.int64)
            data = np.ascontiguousarray(chars)
        elif tid == BOOL:
            data = np.asarray(array.fill_null(False), dtype=np.uint8) if n else np.zeros(0, np.uint8)
        else:
            width = np.dtype(_NP[tid]).itemsize
            buf = array.buffers()[1]
            vals = np.frombuffer(buf, np.uint8)[array.offset * width: (array.offset + n) * width] if buf is not None \
                else np.zeros(0, np.uint8)
            data = np.ascontiguousarray(vals)
        return LogicalColumn(_handle=lib.LogicalColumn.from_host(tid, unit, n, data, validity, offsets))

    @staticmethod
    def from_torch(values, valid=None, *, type_id: Optional[int] = None) -> "LogicalColumn":
        """Wraps a device tensor without copying (the tensor is kept alive by the column).
        `valid` is an optional uint8 tensor holding the LSB validity bitmap."""
        import torch

        dt = {torch.int8: INT8, torch.uint8: UINT8, torch.int16: INT16, torch.int32: INT32, torch.int64: INT64,
              torch.float32: FLOAT, torch.float64: DOUBLE, torch.bool: BOOL}
        tid = type_id if type_id is not None else dt[values.dtype]
        values = values.contiguous()
        keep = (values, valid)
        return LogicalColumn(_handle=lib.LogicalColumn.from_device(
            tid, 0, values.numel() if tid != LARGE_STRING else 0, values.data_ptr(),
            valid.data_ptr() if valid is not None else 0, 0, 0, keep, -1))

    @staticmethod
    def empty_like_logical_column(col: "LogicalColumn") -> "LogicalColumn":
        return LogicalColumn.from_arrow(pa.array([], type=col.dtype()))

    # ---- properties ------------------------------------------------------------------------
    def num_rows(self) -> int:
        return self._handle.num_rows()

    def dtype(self) -> pa.DataType:
        return type_id_to_arrow(self._handle.type_id(), self._handle.type_unit())

    def nullable(self) -> bool:
        return self._handle.nullable()

    def null_count(self) -> int:
        return self._handle.null_count()

    def is_scalar(self) -> bool:
        return False

    # ---- conversion ------------------------------------------------------------------------
    def to_arrow(self) -> pa.Array:
        tid = self._handle.type_id()
        t = self.dtype()
        n = self.num_rows()
        data, valid, offsets = self._handle.to_host()
        vbuf = pa.py_buffer(valid.tobytes()) if valid is not None else None
        if tid == LARGE_STRING:
            return pa.Array.from_buffers(t, n, [vbuf, pa.py_buffer(offsets.tobytes()), pa.py_buffer(data.tobytes())])
        if tid == BOOL:
            vals = data[:n].astype(bool)
            mask = None
            if valid is not None:
                mask = ~np.unpackbits(valid, bitorder="little")[:n].astype(bool)
            return pa.array(vals, type=pa.bool_(), mask=mask)
        return pa.Array.from_buffers(t, n, [vbuf, pa.py_buffer(data.tobytes())])

    def to_array(self):
        """Values as a numpy array (nulls not allowed)."""
        a = self.to_arrow()
        if a.null_count:
            raise ValueError("column has nulls")
        return a.to_numpy(zero_copy_only=False)

    def __repr__(self) -> str:
        return f"LogicalColumn(dtype={self.dtype()}, num_rows={self.num_rows()}, nullable={self.nullable()})"
