"""Parquet ingest and output (reference: python/legate_dataframe/lib/parquet.pyx:44-140,
cpp/src/parquet.cpp).

The files are read and written on the host with pyarrow's Parquet reader and writer; the columns
then move to (or from) HBM as one LogicalTable.  This feeds the device-resident tables of the join,
the groupby and the served form; decoding Parquet on the GPU is not part of this build.
"""

from __future__ import annotations

import glob
import pathlib
from typing import Iterable, List, Optional

import pyarrow as pa
import pyarrow.parquet as pq

from legate_dataframe.lib.core.table import LogicalTable


def _file_list(files) -> List[str]:
    if isinstance(files, str):
        out = sorted(glob.glob(files))
    elif isinstance(files, pathlib.Path):
        out = [str(files)]
    else:
        out = [str(f) for f in files]
    if not out:
        raise ValueError("parquet_read: no files match")
    return out


def read_arrow(files, *, columns: Optional[Iterable[str]] = None) -> pa.Table:
    """The files' rows in file order as one pyarrow table (utf8 made large_utf8, the storage type
    of string columns here, as the reference normalises it)."""
    cols = None if columns is None else list(columns)
    parts = [pq.read_table(f, columns=cols) for f in _file_list(files)]
    t = pa.concat_tables(parts) if len(parts) > 1 else parts[0]
    fields = [pa.field(f.name, pa.large_string() if pa.types.is_string(f.type) else f.type) for f in t.schema]
    return t.cast(pa.schema(fields)).combine_chunks()


def parquet_read(files, *, columns=None, ignore_row_groups=None) -> LogicalTable:
    """Read Parquet files into a logical table.

    ``files``: a glob string (sorted matches), a path, or an iterable of paths; ``columns``: the
    subset of columns to read, in that order.  ``ignore_row_groups`` only changes how the
    reference chunks its distributed read and is accepted for compatibility."""
    del ignore_row_groups
    return LogicalTable.from_arrow(read_arrow(files, columns=columns))


def parquet_write(tbl: LogicalTable, path) -> None:
    """Write a logical table to ``path/part.0.parquet`` (one partition: one GPU's table)."""
    d = pathlib.Path(path)
    d.mkdir(parents=True, exist_ok=True)
    pq.write_table(tbl.to_arrow(), d / "part.0.parquet")
