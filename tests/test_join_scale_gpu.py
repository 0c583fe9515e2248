"""Left and full join at the 50M-build x 500M-probe scale (and a 10x smaller copy), checked
through properties that hold at any size.

Workload (bench.make_join_c3_inputs): lhs (probe) int64 key U[0, 1.6 n_build) with 10% null
keys, rhs (build) int64 key Zipf(0.8) over 0.8 n_build values (the hottest key repeats ~3e5
times at full size), int32 / f64 / timestamp[ms] / large_utf8 payloads on both sides, all 10
columns out, nulls_equal=True (build keys are never null, so a null probe key never matches).

Every input string starts with its row number in base 64, so each output row names the lhs row
rL and rhs row rR it was gathered from.  Checked exactly (no tolerance: this is integer, byte and
index work), following the pairing rule of cudf_join / the reference's null-fill
(cpp/src/join.cu:32-155):
  * row count = sum over probe rows of max(1, matches) (+ unmatched build rows for FULL);
  * each probe row appears max(1, matches) times; its partners have its key, and their count,
    sum and sum of squares of build-row ids equal those of the build rows holding that key;
  * every output column equals its source row's value bit for bit (strings byte for byte),
    and the unmatched side is null in all five of its columns;
  * FULL: each build row whose key no probe row has appears exactly once, with a null lhs side.
"""
import pytest

import bench
from legate_dataframe import _capi as C

pytestmark = pytest.mark.gpu

SIZES = {"c3_10x_smaller": (50_000_000, 5_000_000), "c3_full": (500_000_000, 50_000_000)}


def _bits(torch, bm, n):
    sh = torch.arange(8, device=bm.device, dtype=torch.uint8)
    return ((bm[: (n + 7) // 8].unsqueeze(1) >> sh) & 1).view(-1)[:n].bool()


def _valid(torch, col):
    return _bits(torch, col.validity, col.length) if col.validity is not None else \
        torch.ones(col.length, dtype=torch.bool, device="cuda")


def _decode_tags(torch, col, lut):
    n, total = col.length, col.data.numel()
    out = torch.zeros(n, dtype=torch.int64, device="cuda")
    chunk = 1 << 27
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        o = col.offsets[a:b]
        acc = out[a:b]
        for j in range(8):
            idx = (o + j).clamp_(max=max(0, total - 1))
            acc |= lut[col.data[idx].long()] << (6 * j)
    return out


def _check_strings(torch, col, rows, src_offs, src_chars):
    """Output string i (rows where it is non-null) equals source string rows[i], byte for byte."""
    n = col.length
    lens = col.offsets[1:] - col.offsets[:-1]
    valid = _valid(torch, col)
    src_len = src_offs[rows + 1] - src_offs[rows]
    assert torch.equal(lens[valid], src_len[valid])
    assert int(lens[~valid].abs().sum()) == 0
    chunk = 1 << 25
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        lo, hi = int(col.offsets[a]), int(col.offsets[b])
        if hi == lo:
            continue
        row_of = torch.repeat_interleave(torch.arange(a, b, device="cuda"), lens[a:b])
        pos = torch.arange(lo, hi, device="cuda") - col.offsets[row_of]
        src = src_offs[rows[row_of]] + pos
        assert torch.equal(col.data[lo:hi], src_chars[src]), f"string bytes differ in rows [{a}, {b})"


def _check_fixed(torch, col, rows, sel, src, dtype):
    got = col.data.view(dtype)[: col.length]
    assert torch.equal(got[sel], src[rows[sel]])


@pytest.fixture(scope="module", params=list(SIZES))
def c3(request, gpu):
    import torch

    n_probe, n_build = SIZES[request.param]
    print(f"[c3] generating {request.param}: probe {n_probe} x build {n_build}", flush=True)
    L, R, hl, hr, n_vals = bench.make_join_c3_inputs(torch, C, n_probe, n_build)
    bkey, pkey, pvalid = hr["key"], hl["key"], hl["key_valid"]
    cb = torch.bincount(bkey, minlength=2 * n_vals)  # build rows per key value
    cnt = torch.where(pvalid, cb[pkey], torch.zeros_like(pkey))  # matches of each probe row
    ids = torch.arange(n_build, device="cuda", dtype=torch.int64)
    s1 = torch.zeros(2 * n_vals, dtype=torch.int64, device="cuda").index_add_(0, bkey, ids)
    s2 = torch.zeros(2 * n_vals, dtype=torch.int64, device="cuda").index_add_(0, bkey, ids * ids)
    pc = torch.bincount(pkey[pvalid], minlength=2 * n_vals)  # valid probe rows per key value
    lut = torch.zeros(256, dtype=torch.int64, device="cuda")
    lut[torch.tensor(list(bench.TAG_ALPHABET), device="cuda").long()] = torch.arange(64, device="cuda")
    ctx = dict(L=L, R=R, hl=hl, hr=hr, cnt=cnt, s1=s1, s2=s2, pc=pc, lut=lut, n_probe=n_probe, n_build=n_build)
    torch.cuda.synchronize()
    yield ctx
    ctx.clear()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("jt", [1, 2], ids=["left", "full"])
def test_join_c3_properties(c3, jt):
    import torch

    L, R, hl, hr, cnt, lut = c3["L"], c3["R"], c3["hl"], c3["hr"], c3["cnt"], c3["lut"]
    n_probe, n_build = c3["n_probe"], c3["n_build"]
    print(f"[c3] join type {jt}, probe {n_probe} x build {n_build}", flush=True)
    out = C.join(L, R, [0], [0], jt)
    torch.cuda.synchronize()
    assert len(out) == 10
    n_out = out[0].length
    unmatched_build = c3["pc"][hr["key"]] == 0
    exp_rows = int(cnt.clamp(min=1).sum()) + (int(unmatched_build.sum()) if jt == 2 else 0)
    assert n_out == exp_rows

    # which side each output row has; all five columns of a side share that side's validity
    vL = _valid(torch, out[4])
    vR = _valid(torch, out[9])
    # (lhs payloads and all rhs columns are non-null in the inputs; build keys are never null)
    for i in range(1, 4):
        assert torch.equal(_valid(torch, out[i]), vL)
    for i in range(5):
        assert torch.equal(_valid(torch, out[5 + i]), vR)
    if jt == 1:
        assert bool(vL.all()) and all(out[i].validity is None for i in range(1, 5))
    rL = _decode_tags(torch, out[4], lut)
    rR = _decode_tags(torch, out[9], lut)
    rLs = torch.where(vL, rL, torch.zeros_like(rL))
    rRs = torch.where(vR, rR, torch.zeros_like(rR))
    assert bool((rLs < n_probe).all()) and bool((rRs < n_build).all())
    assert torch.equal(_valid(torch, out[0]), vL & hl["key_valid"][rLs])

    # lhs side rows: probe row multiplicities, matches and partners
    occ = torch.bincount(rL[vL], minlength=n_probe)
    assert torch.equal(occ, cnt.clamp(min=1))
    matched = cnt[rLs] > 0
    assert torch.equal(vR & vL, matched & vL)
    both = vL & vR
    assert torch.equal(hr["key"][rRs[both]], hl["key"][rLs[both]])
    sum1 = torch.zeros(n_probe, dtype=torch.int64, device="cuda").index_add_(0, rL[both], rR[both])
    sum2 = torch.zeros(n_probe, dtype=torch.int64, device="cuda").index_add_(0, rL[both], rR[both] * rR[both])
    pk = hl["key"]
    zero = torch.zeros_like(pk)
    assert torch.equal(sum1, torch.where(cnt > 0, c3["s1"][pk], zero))
    assert torch.equal(sum2, torch.where(cnt > 0, c3["s2"][pk], zero))

    # FULL: unmatched build rows, once each, with a null lhs side
    if jt == 2:
        only_r = vR & ~vL
        assert torch.equal(torch.bincount(rR[only_r], minlength=n_build), unmatched_build.long())
        for i in range(5):
            assert not bool(_valid(torch, out[i])[only_r].any())
    else:
        assert not bool((vR & ~vL).any())

    # every value equals its source row's
    kv = _valid(torch, out[0])
    assert torch.equal(kv[vL], hl["key_valid"][rL[vL]])
    _check_fixed(torch, out[0], rLs, vL & kv, hl["key"], torch.int64)
    _check_fixed(torch, out[1], rLs, vL, hl["i32"], torch.int32)
    _check_fixed(torch, out[2], rLs, vL, hl["f64"].view(torch.int64), torch.int64)
    _check_fixed(torch, out[3], rLs, vL, hl["ts"], torch.int64)
    _check_fixed(torch, out[5], rRs, vR, hr["key"], torch.int64)
    _check_fixed(torch, out[6], rRs, vR, hr["i32"], torch.int32)
    _check_fixed(torch, out[7], rRs, vR, hr["f64"].view(torch.int64), torch.int64)
    _check_fixed(torch, out[8], rRs, vR, hr["ts"], torch.int64)
    _check_strings(torch, out[4], rLs, hl["offs"], hl["chars"])
    _check_strings(torch, out[9], rRs, hr["offs"], hr["chars"])
    # null counts reported with the columns
    for c in out:
        assert c.null_count == c.length - int(_valid(torch, c).sum())
    del out
    torch.cuda.empty_cache()
