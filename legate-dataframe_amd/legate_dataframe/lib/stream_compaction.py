"""distinct (reference: python/legate_dataframe/lib/stream_compaction.pyx:65-94)."""

from __future__ import annotations

from typing import Iterable, Optional

from legate_dataframe._native import lib, sync_stream
from legate_dataframe.lib.core.table import LogicalTable


def distinct(tbl: LogicalTable, keys: Optional[Iterable[str]] = None, *, high_cardinality: bool = False
             ) -> LogicalTable:
    """Distinct rows of ``tbl`` by the ``keys`` columns (all columns when None): one row per
    distinct key tuple, with every column.  For columns that are not keys any row's value of the
    key is possible (this build returns the key's first row, so rows keep their input order).
    ``high_cardinality`` only matters across several GPUs, where it skips a local distinct before
    the shuffle."""
    if keys is None:
        keys = tbl.get_column_names()
    sync_stream()
    return LogicalTable.from_handle(lib.distinct(tbl._handle, list(keys), bool(high_cardinality)))
