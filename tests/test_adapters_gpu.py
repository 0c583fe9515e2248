"""Legate-layout adapters on the GPU (reference: cpp/src/core/null_mask.cu:38-114,
cpp/src/core/ranges.cu:33-138): one byte per row validity <-> LSB bitmap with the null count, and
Rect<1> string ranges (global, inclusive hi) <-> local int64 offsets, checked against numpy; then
one ldf_join on inputs held in Legate's layout, converted and bound back exactly as the task
variant would do it, against the oracle."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _lib():
    from legate_dataframe import _capi as C

    return C, C.load()


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 63, 64, 65, 1000, 100_003])
def test_bools_bitmask_round_trip(gpu, n):
    import torch

    C, L = _lib()
    rng = np.random.default_rng(n)
    bools = (rng.random(n) > 0.3).astype(np.uint8)
    if n > 5:
        bools[rng.integers(0, n, 3)] = 7  # any nonzero byte is true
    db = torch.from_numpy(np.concatenate([bools, np.zeros(1, np.uint8)])).cuda()
    nbytes = (n + 7) // 8
    mask = torch.full((nbytes + 16,), 0xAB, dtype=torch.uint8, device="cuda")  # guard bytes past the bitmap
    nc = torch.full((1,), 12345, dtype=torch.int64, device="cuda")  # overwritten, not accumulated
    C.check(L.ldf_bools_to_bitmask(db.data_ptr(), n, mask.data_ptr(), nc.data_ptr(), C.stream_ptr()))
    got = mask.cpu().numpy()
    exp = np.packbits(bools != 0, bitorder="little")
    assert np.array_equal(got[:nbytes], exp)
    assert np.all(got[nbytes:] == 0xAB)
    assert int(nc.item()) == int((bools == 0).sum())
    back = torch.full((n + 1,), 9, dtype=torch.uint8, device="cuda")
    C.check(L.ldf_bitmask_to_bools(mask.data_ptr(), n, back.data_ptr(), C.stream_ptr()))
    b = back.cpu().numpy()
    assert np.array_equal(b[:n], (bools != 0).astype(np.uint8)) and b[n] == 9


@pytest.mark.parametrize("n", [0, 1, 2, 257, 50_001])
@pytest.mark.parametrize("base", [0, 12_345_678_901])
def test_ranges_offsets_round_trip(gpu, n, base):
    import torch

    C, L = _lib()
    rng = np.random.default_rng(n + 1)
    lens = rng.integers(0, 6, n)  # zero-length strings: hi = lo - 1
    lo = base + np.concatenate([[0], np.cumsum(lens)[:-1]]) if n else np.zeros(0, np.int64)
    ranges = np.stack([lo, lo + lens - 1], 1).astype(np.int64).reshape(-1) if n else np.zeros(0, np.int64)
    dr = torch.from_numpy(np.concatenate([ranges, [0, 0]]).astype(np.int64)).cuda()
    offs = torch.full((n + 2,), -7, dtype=torch.int64, device="cuda")
    C.check(L.ldf_ranges_to_offsets(dr.data_ptr(), n, offs.data_ptr(), C.stream_ptr()))
    o = offs.cpu().numpy()
    exp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    assert np.array_equal(o[:n + 1], exp) and o[n + 1] == -7
    back = torch.full((2 * n + 2,), -7, dtype=torch.int64, device="cuda")
    for b in (0, base):
        C.check(L.ldf_offsets_to_ranges(offs.data_ptr(), n, b, back.data_ptr(), C.stream_ptr()))
        r = back.cpu().numpy()
        assert np.array_equal(r[:2 * n], (ranges - base + b) if n else ranges) and np.all(r[2 * n:] == -7)


def test_join_on_legate_layout_inputs(gpu):
    """Inputs as a Legate PhysicalTable holds them -- one byte per row validity, strings as
    global Rect<1> ranges into a chars store that starts at a nonzero global index -- converted
    with the adapters, joined, and the output bound back to bools and local ranges (the task
    variant's move_into); compared with the oracle's join rows."""
    import torch

    C, L = _lib()
    rng = np.random.default_rng(42)
    n, m = 20_000, 3_000
    lk = rng.integers(0, 4000, n).astype(np.int64)
    lvalid = (rng.random(n) > 0.1).astype(np.uint8)
    lstr = [f"L{i}:" + "x" * int(rng.integers(0, 9)) for i in range(n)]
    rk = rng.integers(0, 4000, m).astype(np.int64)
    rpay = rng.integers(-(2**40), 2**40, m).astype(np.int64)

    def legate_strings(strs, global_base):
        enc = [s.encode() for s in strs]
        lens = np.array([len(e) for e in enc], np.int64)
        lo = global_base + np.concatenate([[0], np.cumsum(lens)[:-1]])
        ranges = np.stack([lo, lo + lens - 1], 1).reshape(-1).astype(np.int64)
        chars = np.frombuffer(b"".join(enc) + b"\0", np.uint8)
        return torch.from_numpy(ranges.copy()).cuda(), torch.from_numpy(chars.copy()).cuda()

    stream = C.stream_ptr()
    # lhs key: Legate bools -> bitmap (+ null count), the view the task variant builds
    lk_d = torch.from_numpy(lk).cuda()
    lb_d = torch.from_numpy(lvalid).cuda()
    lmask = torch.zeros(((n + 63) // 64) * 8, dtype=torch.uint8, device="cuda")
    nc = torch.zeros(1, dtype=torch.int64, device="cuda")
    C.check(L.ldf_bools_to_bitmask(lb_d.data_ptr(), n, lmask.data_ptr(), nc.data_ptr(), stream))
    # lhs strings: global ranges (this rank's chars start at global index 777) -> local offsets
    ranges_d, chars_d = legate_strings(lstr, 777)
    loffs = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    C.check(L.ldf_ranges_to_offsets(ranges_d.data_ptr(), n, loffs.data_ptr(), stream))
    lhs = [C.DeviceColumn(C.INT64, n, lk_d.view(torch.uint8), lmask, None, int(nc.item())),
           C.DeviceColumn(C.LARGE_STRING, n, chars_d, None, loffs, 0)]
    rhs = [C.column_from_numpy(rk, C.INT64), C.column_from_numpy(rpay, C.INT64)]
    out = C.join(lhs, rhs, [0], [0], 1)  # LEFT, nulls_equal: null lhs keys match nothing in rhs
    # bind the output back in Legate's layout
    key, s, rkey, rp = out
    nout = key.length
    kb = torch.empty(nout, dtype=torch.uint8, device="cuda")
    C.check(L.ldf_bitmask_to_bools(key.validity.data_ptr(), nout, kb.data_ptr(), stream))
    rb = torch.empty(nout, dtype=torch.uint8, device="cuda")
    C.check(L.ldf_bitmask_to_bools(rp.validity.data_ptr(), nout, rb.data_ptr(), stream))
    sr = torch.empty(2 * nout, dtype=torch.int64, device="cuda")
    C.check(L.ldf_offsets_to_ranges(s.offsets.data_ptr(), nout, 0, sr.data_ptr(), stream))
    torch.cuda.synchronize()
    chars = s.data.cpu().numpy().tobytes()
    sr = sr.cpu().numpy().reshape(-1, 2)
    got_str = [chars[a:b + 1].decode() for a, b in sr]
    got_key = [int(v) if ok else None for v, ok in zip(key.data.view(torch.int64).cpu().numpy(), kb.cpu().numpy())]
    got_rp = [int(v) if ok else None for v, ok in zip(rp.data.view(torch.int64).cpu().numpy(), rb.cpu().numpy())]
    got = sorted(zip(got_str, [k if k is not None else -1 for k in got_key], [v if v is not None else 0 for v in got_rp]))
    lm, rm = oracle.join([oracle.OCol(oracle.T_INT64, lk, lvalid.astype(bool))], [oracle.OCol(oracle.T_INT64, rk)],
                         "left")
    exp = sorted((lstr[a], int(lk[a]) if lvalid[a] else -1, int(rpay[b]) if b >= 0 else 0) for a, b in zip(lm, rm))
    assert got == exp
    assert sum(1 for v in got_key if v is None) == int((lvalid == 0).sum())
