"""LogicalTable (reference: python/legate_dataframe/lib/core/table.pyx)."""

from __future__ import annotations

from typing import Iterable, List

import pyarrow as pa

from legate_dataframe._native import lib
from legate_dataframe.lib.core.column import LogicalColumn


class LogicalTable:
    """Ordered HBM-resident columns with unique names."""

    def __init__(self, columns: Iterable[LogicalColumn], column_names: Iterable[str]):
        cols = [c._handle if isinstance(c, LogicalColumn) else LogicalColumn(c)._handle for c in columns]
        self._handle = lib.LogicalTable(cols, list(column_names))

    @staticmethod
    def from_handle(handle) -> "LogicalTable":
        t = LogicalTable.__new__(LogicalTable)
        t._handle = handle
        return t

    @staticmethod
    def from_arrow(table: pa.Table) -> "LogicalTable":
        cols = [LogicalColumn.from_arrow(table.column(i)) for i in range(table.num_columns)]
        return LogicalTable(cols, table.column_names)

    def num_columns(self) -> int:
        return self._handle.num_columns()

    def num_rows(self) -> int:
        return self._handle.num_rows()

    def get_column(self, column: int | str) -> LogicalColumn:
        if isinstance(column, str):
            names = self.get_column_names()
            if column not in names:
                raise KeyError(column)
            return LogicalColumn(_handle=self._handle.get_column_by_name(column))
        return LogicalColumn(_handle=self._handle.get_column(column))

    def __getitem__(self, column):
        if isinstance(column, (list, tuple)):
            return self.select(column)
        return self.get_column(column)

    def select(self, columns) -> "LogicalTable":
        names = self.get_column_names()
        sel = [names[c] if isinstance(c, int) else c for c in columns]
        return LogicalTable.from_handle(self._handle.select(sel))

    def get_column_names(self) -> List[str]:
        return list(self._handle.get_column_names())

    def to_arrow(self) -> pa.Table:
        names = self.get_column_names()
        return pa.Table.from_arrays([self.get_column(i).to_arrow() for i in range(len(names))], names=names)

    def __repr__(self) -> str:
        return self._handle.__repr__()
