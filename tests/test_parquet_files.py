"""Host side of the Parquet ingest (no GPU): file lists from a glob, a path or a list; utf8 read
as large_utf8; columns selected in the order asked."""

from __future__ import annotations

import pyarrow as pa
import pyarrow.parquet as pq
import pytest


def test_read_arrow_file_lists_and_types(tmp_path):
    from legate_dataframe.lib.parquet import read_arrow

    pq.write_table(pa.table({"a": [1, 2], "s": ["x", "y"]}), tmp_path / "p1.parquet")
    pq.write_table(pa.table({"a": [3], "s": ["z"]}), tmp_path / "p0.parquet")
    t = read_arrow(str(tmp_path / "p*.parquet"))  # sorted glob: p0 then p1
    assert t.column("a").to_pylist() == [3, 1, 2] and t.schema.field("s").type == pa.large_string()
    assert read_arrow([tmp_path / "p1.parquet"], columns=["s", "a"]).column_names == ["s", "a"]
    assert read_arrow(tmp_path / "p0.parquet").num_rows == 1
    with pytest.raises(ValueError):
        read_arrow(str(tmp_path / "nothing*.parquet"))
