"""Pins the CPU oracle (no GPU): published MurmurHash3 vectors, the reference CPU engine's
(Acero) results on the reference tests' tables, brute force and exact rational arithmetic."""

from __future__ import annotations

import itertools
import struct
from fractions import Fraction

import numpy as np
import pyarrow as pa
import pytest

import oracle
from oracle import OCol, T_INT32, T_INT64, T_DOUBLE, T_LARGE_STRING, T_UINT8
from ldf_testing import (assert_groupby_equal, fixture_table, groupby_fixture, groupby_via_oracle, join_fixture,
                         join_via_oracle, ocol, sorted_rows)


# ---------------------------------------------------------------------------------------------
# murmur3
# ---------------------------------------------------------------------------------------------
# Published MurmurHash3_x86_32 test vectors (SMHasher test suite / reference implementations).
MURMUR_VECTORS = [
    (b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\xff\xff\xff\xff", 0, 0x76293B50), (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE), (b"\x21\x43\x65", 0, 0x7E4A8634),
    (b"\x21\x43", 0, 0xA0F7B07A), (b"\x21", 0, 0x72661CF4), (b"\x00\x00\x00\x00", 0, 0x2362F9DE),
    (b"\x00\x00\x00", 0, 0x85F0B427), (b"\x00\x00", 0, 0x30F4C306), (b"\x00", 0, 0x514E28B7),
    (b"aaaa", 0x9747B28C, 0x5A97808A), (b"aaa", 0x9747B28C, 0x283E0130), (b"aa", 0x9747B28C, 0x5D211726),
    (b"a", 0x9747B28C, 0x7FA09EA6), (b"abcd", 0x9747B28C, 0xF0478627), (b"abc", 0x9747B28C, 0xC84A62DD),
    (b"ab", 0x9747B28C, 0x74875592), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
]


@pytest.mark.parametrize("data,seed,expected", MURMUR_VECTORS)
def test_murmur3_vectors(data, seed, expected):
    assert oracle.murmur3_32(data, seed) == expected


def test_murmur3_smhasher_verification():
    # SMHasher's VerificationTest: hash {}, {0}, {0,1}, ... {0..254} with seed 256-i, then hash
    # the 256 little-endian results with seed 0; MurmurHash3_x86_32's value is 0xB0F57EE3.
    key = bytes(range(256))
    hashes = b"".join(struct.pack("<I", oracle.murmur3_32(key[:i], 256 - i)) for i in range(256))
    assert oracle.murmur3_32(hashes, 0) == 0xB0F57EE3


def _combine(lhs, rhs):
    return (lhs ^ ((rhs + 0x9E3779B9 + (lhs << 6) + (lhs >> 2)) & 0xFFFFFFFF)) & 0xFFFFFFFF


def test_row_hash_composition():
    """row hash = fold of hash_combine(h, murmur3(element bytes, seed)) from h = seed; null -> ~0."""
    x = np.array([0, 1, -1, 2**40, -(2**62)], np.int64)
    y = np.array([7, 8, 9, 10, 11], np.int32)
    yv = np.array([1, 0, 1, 1, 1], bool)
    for seed in (0, 42):
        got = oracle.row_hashes([OCol(T_INT64, x), OCol(T_INT32, y, yv)], seed)
        for i in range(len(x)):
            h = _combine(seed, oracle.murmur3_32(struct.pack("<q", int(x[i])), seed))
            e = oracle.murmur3_32(struct.pack("<i", int(y[i])), seed) if yv[i] else 0xFFFFFFFF
            assert got[i] == _combine(h, e)
    s = ["", "a", "hello"]
    got = oracle.row_hashes([OCol(T_LARGE_STRING, s)], 0)
    assert [int(v) for v in got] == [_combine(0, oracle.murmur3_32(t.encode(), 0)) for t in s]


def test_row_hash_float_normalisation_published_vectors():
    """libcudf's MurmurHash3_x86_32<float/double> hashes normalize_nans_and_zeros(value): -0.0
    hashes as the bytes of 0.0, whose murmur3 is a published vector (4 zero bytes, seed 0:
    0x2362F9DE), and every NaN hashes as the canonical quiet NaN's bytes."""
    from oracle import T_FLOAT
    f32 = np.array([0.0, -0.0], np.float32)
    h = oracle.row_hashes([OCol(T_FLOAT, f32)], 0)
    assert [int(v) for v in h] == [_combine(0, 0x2362F9DE)] * 2
    nans32 = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0x7FC12345, 0xFFFFFFFF], np.uint32).view(np.float32)
    h = oracle.row_hashes([OCol(T_FLOAT, nans32)], 0)
    assert len({int(v) for v in h}) == 1
    assert int(h[0]) == _combine(0, oracle.murmur3_32(struct.pack("<I", 0x7FC00000), 0))
    nans64 = np.array([0x7FF8000000000000, 0xFFF8000000000000, 0x7FF0000000000001, 0x7FF8DEADBEEF0001],
                      np.uint64).view(np.float64)
    d = np.concatenate([nans64, np.array([0.0, -0.0, 1.5, -1.5])])
    h = oracle.row_hashes([OCol(T_DOUBLE, d)], 0)
    assert len({int(v) for v in h[:4]}) == 1 and h[4] == h[5] and h[6] != h[7]
    assert int(h[0]) == _combine(0, oracle.murmur3_32(struct.pack("<Q", 0x7FF8000000000000), 0))
    assert int(h[4]) == _combine(0, oracle.murmur3_32(b"\0" * 8, 0))
    # other values hash their own bytes
    assert int(h[6]) == _combine(0, oracle.murmur3_32(struct.pack("<d", 1.5), 0))


def test_partition_contract():
    rng = np.random.default_rng(1)
    x = rng.integers(-(2**63), 2**63 - 1, 5000, dtype=np.int64)
    for nparts in (1, 3, 16, 256):
        perm, offs = oracle.partition([OCol(T_INT64, x)], nparts, 0)
        assert offs[0] == 0 and offs[-1] == len(x) and np.all(np.diff(offs) >= 0)
        assert sorted(perm.tolist()) == list(range(len(x)))  # every row exactly once
        h = oracle.row_hashes([OCol(T_INT64, x)], 0)
        for p in range(nparts):
            rows = perm[offs[p]:offs[p + 1]]
            assert np.all(h[rows] % nparts == p)
            assert np.all(np.diff(rows) > 0)  # input order kept inside a partition
        if nparts & (nparts - 1) == 0:
            assert np.array_equal(h % nparts, h & (nparts - 1))


# ---------------------------------------------------------------------------------------------
# join
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("case", range(len(join_fixture()["cases"])))
def test_join_oracle_matches_acero(case):
    c = join_fixture()["cases"][case]
    lhs, rhs = fixture_table(c["lhs"]), fixture_table(c["rhs"])
    got = join_via_oracle(lhs, rhs, c["lhs_keys"], c["rhs_keys"], c["how"], c["nulls_equal"])
    exp = fixture_table(c["expected"])
    assert sorted_rows(got) == sorted_rows(exp)


def test_join_oracle_left_join_on_nulls_hardcoded():
    """cpp/tests/test_join.cpp:269-304: the reference's hand-written expectation."""
    c = [x for x in join_fixture()["cases"] if x["name"] == "left_join_on_nulls" and x["how"] == "left"
         and x["nulls_equal"]][0]
    lhs, rhs = fixture_table(c["lhs"]), fixture_table(c["rhs"])
    got = join_via_oracle(lhs, rhs, ["a", "b"], ["d", "e"], "left", True, ["a", "b", "c"], ["f"])
    exp = fixture_table(join_fixture()["left_join_on_nulls_expected"])
    assert sorted_rows(got) == sorted_rows(exp)


def _brute_join(lk, rk, lv, rv, how, nulls_equal):
    def eq(i, j):
        for a, b, va, vb in zip(lk, rk, lv, rv):
            if not va[i] or not vb[j]:
                if not nulls_equal or va[i] != vb[j]:
                    return False
                continue
            if a[i] != b[j]:
                return False
        return True

    nl, nr = len(lk[0]), len(rk[0])
    out, matched = [], set()
    for i in range(nl):
        ms = [j for j in range(nr) if eq(i, j)]
        if how in ("inner", "left", "full"):
            out += [(i, j) for j in ms]
            matched.update(ms)
            if not ms and how in ("left", "full"):
                out.append((i, -1))
        elif how == "semi" and ms:
            out.append((i, -1))
        elif how == "anti" and not ms:
            out.append((i, -1))
    if how == "full":
        out += [(-1, j) for j in range(nr) if j not in matched]
    return sorted(out)


@pytest.mark.parametrize("how", ["inner", "left", "full", "semi", "anti"])
@pytest.mark.parametrize("nulls_equal", [True, False])
@pytest.mark.parametrize("seed", range(3))
def test_join_oracle_brute_force(how, nulls_equal, seed):
    rng = np.random.default_rng(seed)
    nl, nr = rng.integers(0, 40), rng.integers(0, 40)
    lk = [rng.integers(0, 6, nl), np.array([f"s{v}" for v in rng.integers(0, 3, nl)])]
    rk = [rng.integers(0, 6, nr), np.array([f"s{v}" for v in rng.integers(0, 3, nr)])]
    lv = [rng.random(nl) > 0.2, rng.random(nl) > 0.1]
    rv = [rng.random(nr) > 0.2, rng.random(nr) > 0.1]
    L = [OCol(T_INT64, lk[0], lv[0]), OCol(T_LARGE_STRING, list(lk[1]), lv[1])]
    R = [OCol(T_INT64, rk[0], rv[0]), OCol(T_LARGE_STRING, list(rk[1]), rv[1])]
    lm, rm = oracle.join(L, R, how, nulls_equal)
    got = sorted(zip(lm.tolist(), rm.tolist()))
    assert got == _brute_join(lk, rk, lv, rv, how, nulls_equal)


def test_join_oracle_closed_form_permutation_build():
    """Build keys = a permutation of 0..m-1 (the 1M x 10M inner-join shape, scaled down):
    probe row i matches exactly when key_i < m, with the build row holding that key."""
    rng = np.random.default_rng(5)
    m, n = 1000, 10000
    bkeys = rng.permutation(m).astype(np.int64)
    pkeys = rng.integers(0, 2 * m, n)
    lm, rm = oracle.join([OCol(T_INT64, pkeys)], [OCol(T_INT64, bkeys)], "inner")
    where = np.empty(m, np.int64)
    where[bkeys] = np.arange(m)
    exp_l = np.nonzero(pkeys < m)[0]
    assert sorted(zip(lm.tolist(), rm.tolist())) == sorted(zip(exp_l.tolist(), where[pkeys[exp_l]].tolist()))


# ---------------------------------------------------------------------------------------------
# groupby
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("case", range(len(groupby_fixture()["cases"])))
def test_groupby_oracle_matches_acero(case):
    fx = groupby_fixture()
    c = fx["cases"][case]
    t = fixture_table(fx["tables"][c["table"]])
    aggs = [(a, k, o) for a, k, o in c["aggs"]]
    got = groupby_via_oracle(t, c["keys"], aggs)
    exp = fixture_table(c["expected"])
    # float results differ from Acero's in rounding only: summation order differs
    assert_groupby_equal(got, exp, c["keys"], rel=1e-9)


def test_groupby_oracle_float_sum_exact_rational():
    rng = np.random.default_rng(11)
    n = 4000
    k = rng.integers(0, 7, n)
    x = rng.uniform(-1, 1, n) * 10.0 ** rng.integers(-8, 8, n)
    reps, [(bits, f64, valid)] = oracle.groupby([OCol(T_INT64, k)], [(OCol(T_DOUBLE, x), "sum", None)])
    for g, r in enumerate(reps):
        exact = sum((Fraction(v) for v in x[k == k[r]]), Fraction(0))
        assert abs(Fraction(f64[g]) - exact) <= Fraction(1, 10**12) * abs(exact)


def test_groupby_oracle_brute_force():
    rng = np.random.default_rng(12)
    n = 3000
    k1 = rng.integers(0, 5, n)
    k1v = rng.random(n) > 0.1
    k2 = [f"k{v}" for v in rng.integers(0, 4, n)]
    v = rng.integers(-(2**62), 2**62, n)
    vv = rng.random(n) > 0.2
    keys = [OCol(T_INT64, k1, k1v), OCol(T_LARGE_STRING, k2)]
    aggs = [(OCol(T_INT64, v, vv), "sum", None), (OCol(T_INT64, v, vv), "min", None),
            (OCol(T_INT64, v, vv), "max", None), (OCol(T_INT64, v, vv), "count", None),
            (OCol(T_UINT8, v % 7, vv), "count_distinct", None), (None, "count_all", None)]
    reps, res = oracle.groupby(keys, aggs)
    groups = {}
    for i in range(n):
        groups.setdefault((k1[i] if k1v[i] else None, k2[i]), []).append(i)
    assert len(reps) == len(groups)
    for g, r in enumerate(reps):
        rows = groups[(k1[r] if k1v[r] else None, k2[r])]
        vals = [int(v[i]) for i in rows if vv[i]]
        wrap = (sum(vals) + 2**63) % 2**64 - 2**63
        assert res[0][2][g] == bool(vals)
        if vals:
            assert int(res[0][0][g].view(np.int64)) == wrap
            assert int(res[1][0][g].view(np.int64)) == min(vals)
            assert int(res[2][0][g].view(np.int64)) == max(vals)
        assert int(res[3][0][g]) == len(vals)
        assert int(res[4][0][g]) == len({int(v[i]) % 7 for i in rows if vv[i]})
        assert int(res[5][0][g]) == len(rows)


def test_groupby_oracle_median_vs_numpy():
    """approximate_median as the GPU variant computes it (cudf's exact median): per group the
    median of the valid values, numpy.median of the group's values; no valid value -> null;
    NaN sorts after every number; integer inputs give the double of the middle value(s)."""
    rng = np.random.default_rng(13)
    n = 6000
    k = rng.integers(0, 40, n)
    x = rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, n)
    xv = rng.random(n) > 0.15
    xv[k == 7] = False  # a group without valid values
    i = rng.integers(-(2**40), 2**40, n)
    reps, res = oracle.groupby([OCol(T_INT64, k)], [(OCol(T_DOUBLE, x, xv), "approximate_median", None),
                                                    (OCol(T_INT64, i), "approximate_median", None)])
    for g, r in enumerate(reps):
        sel = k == k[r]
        vals = x[sel & xv]
        assert res[0][2][g] == (len(vals) > 0)
        if len(vals):
            assert res[0][1][g] == np.median(vals)
        assert res[1][1][g] == np.median(i[sel].astype(np.float64)) or len(set(i[sel])) < 2
        srt = sorted(int(v) for v in i[sel])
        m = len(srt)
        assert res[1][1][g] == (float(srt[m // 2]) if m % 2 else 0.5 * float(srt[m // 2 - 1]) + 0.5 * float(srt[m // 2]))
    # NaN after every number: a group {1, 2, nan} has median 2, {1, nan} has 0.5 + 0.5 nan = nan
    reps, res = oracle.groupby([OCol(T_INT64, np.array([0, 0, 0, 1, 1]))],
                               [(OCol(T_DOUBLE, np.array([np.nan, 1.0, 2.0, np.nan, 1.0])), "approximate_median", None)])
    assert res[0][1][0] == 2.0 and np.isnan(res[0][1][1])
