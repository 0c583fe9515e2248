"""join (reference: python/legate_dataframe/lib/join.pyx:21-146)."""

from __future__ import annotations

from enum import IntEnum
from typing import Iterable, Optional

from legate_dataframe._native import lib, sync_stream
from legate_dataframe.lib.core.table import LogicalTable


class JoinType(IntEnum):
    """Options for the join type (``INNER``, ``LEFT``, ``FULL``, ``SEMI``, ``ANTI``)."""

    INNER = 0
    LEFT = 1
    FULL = 2
    SEMI = 3
    ANTI = 4


class BroadcastInput(IntEnum):
    """Options for input broadcasting (``AUTO``, ``LEFT``, and ``RIGHT``)."""

    AUTO = 0
    LEFT = 1
    RIGHT = 2


def join(
    lhs: LogicalTable,
    rhs: LogicalTable,
    *,
    lhs_keys: Iterable[str],
    rhs_keys: Iterable[str],
    join_type: JoinType,
    lhs_out_columns: Optional[Iterable[str]] = None,
    rhs_out_columns: Optional[Iterable[str]] = None,
    nulls_equal: bool = True,
    broadcast: BroadcastInput = BroadcastInput.AUTO,
) -> LogicalTable:
    """Perform a join between the specified tables.

    By default the result includes the columns of `lhs` and of `rhs` (none of `rhs` for SEMI and
    ANTI).  `lhs` is the probe side and `rhs` the build side of the hash join.

    Raises
    ------
    ValueError
        If the number of elements in `lhs_keys` and `rhs_keys` differ or if the column names of
        `lhs_out_columns` and `rhs_out_columns` are not unique.
    RuntimeError
        If a broadcast is forced where a repartition would be needed.
    """
    if lhs_out_columns is None:
        lhs_out_columns = lhs.get_column_names()
    if rhs_out_columns is None:
        if join_type == JoinType.ANTI or join_type == JoinType.SEMI:
            rhs_out_columns = []  # cannot select any columns with these joins
        else:
            rhs_out_columns = rhs.get_column_names()
    sync_stream()
    return LogicalTable.from_handle(
        lib.join(lhs._handle, rhs._handle, list(lhs_keys), list(rhs_keys), int(join_type), list(lhs_out_columns),
                 list(rhs_out_columns), bool(nulls_equal), int(broadcast)))
