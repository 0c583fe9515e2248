"""ldf_hash_partition_noalloc (scratch in a fixed device array, caller-provided outputs) against
ldf_hash_partition through the C ABI: identical columns and offsets, bit for bit, on two calls in a
row (the second with a smaller input, so stale scratch from the first would show), for a single
int64 key column, a float64 key with NaN and -0.0, a multi-column table and 2048 partitions; the
default allocator's blocks do not change across the calls; invalid inputs are refused."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cols(C, torch, n, seed, kind):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if kind == "int64":
        x = torch.randint(-(2**63), 2**63 - 1, (n,), device="cuda", generator=g, dtype=torch.int64)
        return [C.DeviceColumn(C.INT64, n, x.view(torch.uint8))]
    if kind == "float64":
        x = torch.randn(n, device="cuda", generator=g, dtype=torch.float64)
        x[::7] = float("nan")
        x[1::11] = -0.0
        x[2::13] = 0.0
        return [C.DeviceColumn(C.DOUBLE, n, x.view(torch.uint8))]
    a = torch.randint(0, 1000, (n,), device="cuda", generator=g, dtype=torch.int32)
    b = torch.randint(-(2**62), 2**62, (n,), device="cuda", generator=g, dtype=torch.int64)
    c = torch.randint(0, 255, (n,), device="cuda", generator=g, dtype=torch.uint8)
    return [C.DeviceColumn(C.INT32, n, a.view(torch.uint8)), C.DeviceColumn(C.INT64, n, b.view(torch.uint8)),
            C.DeviceColumn(C.UINT8, n, c)]


@pytest.mark.parametrize("kind,hash_cols,nparts", [("int64", [0], 256), ("float64", [0], 64),
                                                    ("multi", [0, 1], 100), ("int64", [0], 2048)])
def test_noalloc_partition_matches_allocating_entry_point(gpu, kind, hash_cols, nparts):
    import torch

    from legate_dataframe import _capi as C

    sizes = (3_000_017, 1_000_003)  # the second call is smaller: scratch left by the first is stale
    widths = {C.INT64: 8, C.DOUBLE: 8, C.INT32: 4, C.UINT8: 1}
    cap = max(sizes)
    outs = None
    for call, n in enumerate(sizes):
        cols = _cols(C, torch, n, 100 + call, kind)
        if outs is None:  # caller-owned outputs, allocated once, large enough for both calls
            outs = [torch.full((cap * widths[c.type_id],), 0xCD, dtype=torch.uint8, device="cuda") for c in cols]
        torch.cuda.synchronize()
        before = C.pool_stats()
        got, goffs = C.hash_partition_noalloc(cols, hash_cols, nparts, outs, seed=7)
        torch.cuda.synchronize()
        assert C.pool_stats() == before, "the no-alloc entry point used the default allocator"
        print(f"[noalloc] {kind} call {call}: n={n} nparts={nparts}", flush=True)
        rt, eoffs = C.hash_partition_raw(cols, hash_cols, nparts, seed=7)
        assert goffs == eoffs
        assert goffs[0] == 0 and goffs[-1] == n
        for i, c in enumerate(got):
            exp = C._view_bytes(rt.cols[i].data, n * widths[cols[i].type_id])
            assert c.length == n and c.type_id == cols[i].type_id
            assert torch.equal(c.data, exp), (call, i)
        rt.free()


def test_noalloc_partition_rejects_invalid_inputs(gpu):
    import torch

    from legate_dataframe import _capi as C

    n = 1000
    x = torch.arange(n, device="cuda", dtype=torch.int64)
    small = [torch.empty(8 * n - 8, dtype=torch.uint8, device="cuda")]
    with pytest.raises(ValueError, match="data_bytes"):
        C.hash_partition_noalloc([C.DeviceColumn(C.INT64, n, x.view(torch.uint8))], [0], 16, small)
    valid = torch.full(((n + 63) // 64 * 8,), 0xFF, dtype=torch.uint8, device="cuda")
    out = [torch.empty(8 * n, dtype=torch.uint8, device="cuda")]
    with pytest.raises(ValueError, match="non-nullable"):
        C.hash_partition_noalloc([C.DeviceColumn(C.INT64, n, x.view(torch.uint8), valid)], [0], 16, out)
    with pytest.raises(C.LdfError, match="2048"):
        C.hash_partition_noalloc([C.DeviceColumn(C.INT64, n, x.view(torch.uint8))], [0], 4096, out)
    # empty input
    got, offs = C.hash_partition_noalloc([C.DeviceColumn(C.INT64, 0, x[:0].view(torch.uint8))], [0], 4, out)
    assert offs == [0] * 5 and got[0].length == 0
