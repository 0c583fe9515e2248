"""Local murmur3 hash partition: the one-GPU step of the reference's repartition_by_hash
(cpp/src/core/repartition_by_hash.cu:296-310: cudf::hash_partition(HASH_MURMUR3, seed 0) + split)."""

from __future__ import annotations

from typing import Iterable, List, Tuple

from legate_dataframe._native import lib, sync_stream
from legate_dataframe.lib.core.table import LogicalTable


def hash_partition(table: LogicalTable, columns: Iterable[str], num_partitions: int, seed: int = 0
                   ) -> Tuple[LogicalTable, List[int]]:
    """Rows grouped by murmur3(row of `columns`) mod `num_partitions` (input order kept inside a
    partition) and the ``num_partitions + 1`` partition offsets."""
    sync_stream()
    t, offsets = lib.hash_partition(table._handle, list(columns), int(num_partitions), int(seed))
    return LogicalTable.from_handle(t), list(offsets)
