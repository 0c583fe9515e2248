"""groupby_aggregation (reference: python/legate_dataframe/lib/groupby_aggregation.pyx:48-117)."""

from __future__ import annotations

import re
from typing import Iterable, Tuple

from legate_dataframe._native import lib, sync_stream
from legate_dataframe.lib.core.table import LogicalTable

# pyarrow prints CountOptions' mode as NON_NULL / NULLS / ALL
_MODES = {"NON_NULL": 0, "ONLY_VALID": 0, "NULLS": 1, "ONLY_NULL": 1, "ALL": 2}


def _count_mode(options) -> int:
    """Mode of a pyarrow.compute.CountOptions (from its repr, e.g. 'CountOptions(mode=ALL)')."""
    import pyarrow.compute as pc

    if not isinstance(options, pc.FunctionOptions):
        raise TypeError("options must be an arrow `FunctionOptions` object.")
    m = re.match(r"CountOptions\(mode=(\w+)\)", str(options))
    if m is None:
        raise ValueError(f"options type '{type(options).__name__}' is not supported; only CountOptions is")
    return _MODES[m.group(1)]


def groupby_aggregation(
    table: LogicalTable, keys: Iterable[str], column_aggregations: Iterable[Tuple[str, str, str]]
) -> LogicalTable:
    """Perform a groupby and aggregation in a single operation.

    Available aggregations: "sum", "product", "min", "max", "count", "mean", "variance",
    "stddev", "approximate_median", "count_distinct", "count_all".  "approximate_median" is the
    exact median (float64), as the reference's GPU variant computes it.  Each column aggregation is
    ``(<input-column-name>, <aggregation-kind>, <output-column-name>)``; the kind may be
    ``(kind, pyarrow.compute.CountOptions(...))`` for "count" and "count_distinct".
    "count_all" takes ``None`` as its input column.

    Returns a table with the key columns and then the aggregated columns in the given order.
    """
    aggs = []
    for in_col_name, kind, out_col_name in column_aggregations:
        mode = -1
        if kind == "count_all":
            if in_col_name is not None:
                raise ValueError("input column name must be `None` for 'count_all'")
            in_col_name = ""
        if isinstance(kind, tuple):
            kind, py_options = kind
            mode = _count_mode(py_options)
        if not isinstance(kind, str):
            raise TypeError("aggregation kind must be string or tuple of string and options.")
        aggs.append((in_col_name, kind, mode, out_col_name))
    sync_stream()
    return LogicalTable.from_handle(lib.groupby_aggregation(table._handle, list(keys), aggs))
