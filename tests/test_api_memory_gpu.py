"""The C++/Python API hands every output buffer back to the allocator that made it: after many
LogicalTable joins, groupbys, partitions and distincts the default allocator's blocks in use and
the device's free memory stay flat (the reference binds each buffer to exactly one owner,
cpp/src/core/column.cu:64-70, cpp/src/core/allocator.cu:60-81)."""

from __future__ import annotations

import gc

import numpy as np
import pyarrow as pa
import pytest

pytestmark = pytest.mark.gpu


def test_api_outputs_return_to_the_pool(gpu):
    import torch

    from legate_dataframe import LogicalTable
    from legate_dataframe import _capi as C
    from legate_dataframe.lib.groupby_aggregation import groupby_aggregation
    from legate_dataframe.lib.hash_partition import hash_partition
    from legate_dataframe.lib.join import JoinType, join
    from legate_dataframe.lib.stream_compaction import distinct

    rng = np.random.default_rng(0)
    n = 50_000
    lhs = LogicalTable.from_arrow(pa.table({"k": rng.integers(0, 20_000, n),
                                            "x": pa.array(rng.random(n), mask=rng.random(n) < 0.1),
                                            "s": [f"s{v}" for v in rng.integers(0, 999, n)]}))
    rhs = LogicalTable.from_arrow(pa.table({"k2": rng.integers(0, 20_000, n // 4), "y": rng.integers(0, 9, n // 4)}))
    gtab = LogicalTable.from_arrow(pa.table({"g": rng.integers(0, 500, n).astype(np.int32),
                                             "v": rng.integers(-100, 100, n)}))

    def work(i):
        jt = (JoinType.INNER, JoinType.LEFT, JoinType.FULL)[i % 3]
        out = join(lhs, rhs, lhs_keys=["k"], rhs_keys=["k2"], join_type=jt)
        rows = out.num_rows()
        g = groupby_aggregation(gtab, ["g"], [("v", "sum", "s"), ("v", "count", "c")])
        assert g.num_rows() == 500
        if i % 10 == 0:
            p, offs = hash_partition(lhs, ["k"], 8)
            assert offs[-1] == n
            d = distinct(gtab, ["g"])
            assert d.num_rows() == 500
            del p, d
        del out, g
        return rows

    expected = [work(i) for i in range(3)]  # warm the cache
    gc.collect()
    torch.cuda.synchronize()
    base = C.pool_stats()
    free_base = torch.cuda.mem_get_info()[0]
    for i in range(1000):
        assert work(i) == expected[i % 3]
        if i % 250 == 249:
            gc.collect()
            torch.cuda.synchronize()
            st = C.pool_stats()
            print(f"[api-memory] iteration {i + 1}: {st}", flush=True)
            # blocks in use: the three input tables' buffers only
            assert st["live_blocks"] == base["live_blocks"], (base, st)
            assert st["cached_bytes"] <= base["cached_bytes"] + (64 << 20), (base, st)
            assert torch.cuda.mem_get_info()[0] >= free_base - (64 << 20)
