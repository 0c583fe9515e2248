"""The order-independent table fingerprint used by the full-size GPU tests (CPU tensors here):
invariant under row permutation, sensitive to one changed value, a changed validity bit, a moved
value between columns and a dropped row."""

from __future__ import annotations

import numpy as np

from ldf_testing import table_fingerprint


def _cols(torch, C, arrays):
    out = []
    for v, tid, ok in arrays:
        v = np.ascontiguousarray(v)
        valid = None
        if ok is not None:
            valid = torch.from_numpy(np.concatenate([np.packbits(ok, bitorder="little"), np.zeros(8, np.uint8)]))
        out.append(C.DeviceColumn(tid, len(v), torch.from_numpy(v.view(np.uint8).copy()), valid))
    return out


def test_fingerprint_properties():
    import torch

    from legate_dataframe import _capi as C

    rng = np.random.default_rng(0)
    n = 5000
    a = rng.integers(-(2**62), 2**62, n)
    b = rng.random(n)
    c = rng.integers(0, 9, n).astype(np.int32)
    ok = rng.random(n) > 0.2
    base = table_fingerprint(_cols(torch, C, [(a, C.INT64, None), (b, C.DOUBLE, None), (c, C.INT32, ok)]))
    p = rng.permutation(n)
    assert np.array_equal(base, table_fingerprint(_cols(torch, C, [(a[p], C.INT64, None), (b[p], C.DOUBLE, None),
                                                                   (c[p], C.INT32, ok[p])])))
    assert base[0, 0] == n
    a2 = a.copy()
    a2[17] ^= 1
    assert not np.array_equal(base, table_fingerprint(_cols(torch, C, [(a2, C.INT64, None), (b, C.DOUBLE, None),
                                                                       (c, C.INT32, ok)])))
    ok2 = ok.copy()
    ok2[3] = not ok2[3]
    assert not np.array_equal(base, table_fingerprint(_cols(torch, C, [(a, C.INT64, None), (b, C.DOUBLE, None),
                                                                       (c, C.INT32, ok2)])))
    c2 = c.copy()
    c2[[5, 6]] = c2[[6, 5]] if c2[5] != c2[6] else c2[[5, 6]] + 1  # values swapped between rows
    assert not np.array_equal(base, table_fingerprint(_cols(torch, C, [(a, C.INT64, None), (b, C.DOUBLE, None),
                                                                       (c2, C.INT32, ok)])))
    assert not np.array_equal(base, table_fingerprint(_cols(torch, C, [(a[1:], C.INT64, None), (b[1:], C.DOUBLE, None),
                                                                       (c[1:], C.INT32, ok[1:])])))
    # segments: the fingerprint of each segment equals the fingerprint of that slice alone
    offs = np.array([0, 100, 100, 2500, n])
    seg = table_fingerprint(_cols(torch, C, [(a, C.INT64, None), (c, C.INT32, ok)]), offs)
    for s in range(4):
        lo, hi = offs[s], offs[s + 1]
        one = table_fingerprint(_cols(torch, C, [(a[lo:hi], C.INT64, None), (c[lo:hi], C.INT32, ok[lo:hi])]))
        assert np.array_equal(seg[s], one[0])
