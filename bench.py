#!/usr/bin/env python
"""Benchmark of the MI355X hash join / groupby / hash-partition path.

Default (the line the driver records): the 100M-row groupby on one GPU -- int32 key Zipf(1.1)
over 1M values; float64 U[-1,1) (SUM, MEAN); int64 U[-2^40, 2^40] (SUM, MIN, MAX, COUNT); int32
U[0, 10^7) (NUNIQUE); 1% nulls in each value column -- one plain groupby call per step, inputs
resident in HBM, outputs allocated inside the timed region.  `value` = rows/s.

roofline: the dominant kernel (gb_aggregate) timed with HIP events on its own stream during the
timed steps; achieved = its algorithmic bytes (input columns read once) / its mean duration.
cpu_baseline: Arrow Acero (pyarrow, use_threads=False: the engine and options of the reference
CPU variant) on the whole 100M-row workload on 1 core (`value`), the same on N single-threaded
processes over hash partitions of the rows (`n_cores`), and the oracle's C restatement on 1 core
(`port`); the 1M x 10M inner join carries the same three figures.

Other workloads (also reported as `secondary` entries of the default line): --workload join
(1M-build x 10M-probe inner join), --workload join_left / join_full (50M-build x 500M-probe with
string payloads), --workload partition (1B int64 keys -> 256 partitions), --copy-calibration (a
plain copy of known size, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE).  --workload distinct
(100M rows by an int64 key over 10^7 values, with a float64 column) is measured on request only.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "legate-dataframe_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "join/groupby rows/s + % HBM roofline; batched-vs-plain speedup; p99 ms"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------
# synthetic inputs (device resident)
# ---------------------------------------------------------------------------------------------
def zipf_keys(torch, n, n_keys, s, gen, dtype):
    ranks = torch.arange(1, n_keys + 1, device="cuda", dtype=torch.float64)
    cdf = torch.cumsum(ranks.pow(-s), 0)
    cdf /= cdf[-1].clone()
    out = torch.empty(n, dtype=dtype, device="cuda")
    chunk = 1 << 26
    for i in range(0, n, chunk):
        m = min(chunk, n - i)
        u = torch.rand(m, device="cuda", generator=gen, dtype=torch.float64)
        out[i:i + m] = torch.searchsorted(cdf, u).clamp_(max=n_keys - 1).to(dtype)
    return out


def bitmap(torch, valid_bool):
    """LSB-first validity bitmap (uint8) of a bool tensor, padded to whole 64-bit words."""
    n = valid_bool.numel()
    pad = (-n) % 64
    v = torch.cat([valid_bool, torch.zeros(pad, dtype=torch.bool, device=valid_bool.device)]) if pad else valid_bool
    w = torch.tensor([1, 2, 4, 8, 16, 32, 64, 128], dtype=torch.uint8, device=v.device)
    return (v.view(-1, 8).to(torch.uint8) * w).sum(1, dtype=torch.uint8)


def make_groupby_inputs(torch, C, n, seed=2):
    g = torch.Generator(device="cuda").manual_seed(seed)
    k = zipf_keys(torch, n, 1_000_000, 1.1, g, torch.int32)
    f = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    i = torch.randint(-(2**40), 2**40 + 1, (n,), device="cuda", generator=g, dtype=torch.int64)
    u = torch.randint(0, 10_000_000, (n,), device="cuda", generator=g, dtype=torch.int32)
    masks = [torch.rand(n, device="cuda", generator=g) >= 0.01 for _ in range(3)]
    cols = [
        C.DeviceColumn(C.INT32, n, k.view(torch.uint8)),
        C.DeviceColumn(C.DOUBLE, n, f.view(torch.uint8), bitmap(torch, masks[0])),
        C.DeviceColumn(C.INT64, n, i.view(torch.uint8), bitmap(torch, masks[1])),
        C.DeviceColumn(C.INT32, n, u.view(torch.uint8), bitmap(torch, masks[2])),
    ]
    host = dict(k=k, f=f, i=i, u=u, masks=masks)
    return cols, host


TAG_ALPHABET = b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"


def tagged_strings(torch, n, g):
    """n ASCII strings of length U{8..32} (large_utf8: int64 offsets + chars).  The first 8
    characters spell the row number in base 64 (low digit first), the rest are random printable
    characters, so a gathered string names the input row it came from."""
    lens = torch.randint(8, 33, (n,), device="cuda", generator=g, dtype=torch.int64)
    offs = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    torch.cumsum(lens, 0, out=offs[1:])
    del lens
    chars = torch.randint(33, 127, (int(offs[-1]),), device="cuda", generator=g, dtype=torch.uint8)
    alpha = torch.tensor(list(TAG_ALPHABET), dtype=torch.uint8, device="cuda")
    chunk = 1 << 27
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        rows = torch.arange(a, b, device="cuda", dtype=torch.int64)
        for j in range(8):
            chars[offs[a:b] + j] = alpha[(rows >> (6 * j)) & 63]
    return offs, chars


def make_join_c3_inputs(torch, C, n_probe=500_000_000, n_build=50_000_000, seed=3):
    """Left/full join workload at 50M x 500M: lhs (probe) key int64 U[0, 1.6 n_build) with 10%
    null keys; rhs (build) key int64 = Zipf(0.8) rank - 1 over 0.8 n_build values; both sides
    carry int32 U, f64 U[-1,1), timestamp[ms] U[2000-01-01, 2030-01-01) and large_utf8 payloads."""
    g = torch.Generator(device="cuda").manual_seed(seed)

    def side(n, key, key_valid):
        i32 = torch.randint(-(2**31), 2**31, (n,), device="cuda", generator=g, dtype=torch.int32)
        f64 = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
        ts = torch.randint(946684800000, 1893456000000, (n,), device="cuda", generator=g, dtype=torch.int64)
        offs, chars = tagged_strings(torch, n, g)
        kv = bitmap(torch, key_valid) if key_valid is not None else None
        cols = [C.DeviceColumn(C.INT64, n, key.view(torch.uint8), kv), C.DeviceColumn(C.INT32, n, i32.view(torch.uint8)),
                C.DeviceColumn(C.DOUBLE, n, f64.view(torch.uint8)), C.DeviceColumn(C.TIMESTAMP, n, ts.view(torch.uint8)),
                C.DeviceColumn(C.LARGE_STRING, n, chars, None, offs)]
        return cols, dict(key=key, key_valid=key_valid, i32=i32, f64=f64, ts=ts, offs=offs, chars=chars)

    n_vals = max(1, n_build * 4 // 5)
    bkey = zipf_keys(torch, n_build, n_vals, 0.8, g, torch.int64)
    R, hr = side(n_build, bkey, None)
    pkey = torch.randint(0, 2 * n_vals, (n_probe,), device="cuda", generator=g, dtype=torch.int64)
    pvalid = torch.rand(n_probe, device="cuda", generator=g) >= 0.10
    L, hl = side(n_probe, pkey, pvalid)
    return L, R, hl, hr, n_vals


def column_bytes(C, type_id, n, char_bytes, nullable):
    """Bytes of a column in the minimal layout: values (8 B offset per row + chars for strings)
    plus one validity bit per row when nullable."""
    b = 8 * n + char_bytes if type_id == C.LARGE_STRING else n * C.WIDTH[type_id]
    return b + ((n + 7) // 8 if nullable else 0)


def input_bytes(C, cols):
    return sum(column_bytes(C, c.type_id, c.length, int(c.data.numel()) if c.type_id == C.LARGE_STRING else 0,
                            c.validity is not None) for c in cols)


def output_bytes(C, rt):
    return sum(column_bytes(C, c.type_id, c.length, c.data_bytes if c.type_id == C.LARGE_STRING else 0,
                            bool(c.validity)) for c in rt.cols[:rt.out.ncols])


GB_AGGS = [(1, "sum", None), (1, "mean", None), (2, "sum", None), (2, "min", None), (2, "max", None),
           (2, "count", None), (3, "count_distinct", None)]


def groupby_alg_bytes(n, ngroups):
    """Inputs read once (values + 1 bit/row validity of the 3 nullable columns) and outputs
    written once (key i32 + 2 f64 + 4 i64 + 1 i64 per group + 6 validity bitmaps)."""
    inp = n * (4 + 8 + 8 + 4) + 3 * (n // 8)
    out = ngroups * (4 + 8 * 7) + 6 * ((ngroups + 7) // 8)
    return inp, out


# ---------------------------------------------------------------------------------------------
# CPU baseline: the reference CPU variant's engine (Arrow Acero, single-threaded) on the host
# ---------------------------------------------------------------------------------------------
# GroupByAggregationTask::cpu_variant and JoinTask::cpu_variant run Acero with use_threads=false
# (cpp/src/groupby_aggregation.cpp:87-91, cpp/src/join.cpp:214-227), one task per Legate CPU;
# with N CPUs the rows are first hash-partitioned so each task sees whole groups / key matches.
# The same Acero calls run here through pyarrow: once on 1 core over the whole workload, and as N
# single-threaded worker processes, each on its hash partition of the rows (N = the host's cores,
# at most 16, the GPU box's CPU share).  Workers are spawned before the GPU is initialised and
# read the rows from memory-mapped files; only the Acero call itself is timed.
CPU_WORKERS_MAX = 16
CPU_POOL = None  # (multiprocessing pool, workers) when the CPU baseline runs


def cpu_worker_count():
    return max(1, min(CPU_WORKERS_MAX, os.cpu_count() or 1))


def start_cpu_pool():
    import multiprocessing as mp

    n = cpu_worker_count()
    pool = mp.get_context("spawn").Pool(n)
    pool.map(_w_ready, range(n))  # every worker is up before the parent touches the GPU
    return pool, n


def _w_ready(_):
    return os.getpid()


def _mm(path, lo, hi):
    import numpy as np

    return np.ascontiguousarray(np.load(path, mmap_mode="r")[lo:hi])


def _w_acero_groupby(job):
    """One GroupByAggregationTask::cpu_variant on rows [lo, hi): the 7 aggregations."""
    import pyarrow as pa

    paths, lo, hi = job
    m = [_mm(paths[f"m{j}"], lo, hi) for j in range(3)]
    tab = pa.table({"k": _mm(paths["k"], lo, hi), "f": pa.array(_mm(paths["f"], lo, hi), mask=~m[0]),
                    "i": pa.array(_mm(paths["i"], lo, hi), mask=~m[1]), "u": pa.array(_mm(paths["u"], lo, hi), mask=~m[2])})
    t0 = time.perf_counter()
    out = pa.TableGroupBy(tab, ["k"], use_threads=False).aggregate(
        [("f", "sum"), ("f", "mean"), ("i", "sum"), ("i", "min"), ("i", "max"), ("i", "count"), ("u", "count_distinct")])
    return time.perf_counter() - t0, out.num_rows


def _w_oracle_groupby(job):
    """The oracle's single-threaded C restatement of the same groupby (a side figure)."""
    import oracle

    paths, lo, hi = job
    m = [_mm(paths[f"m{j}"], lo, hi) for j in range(3)]
    f, i, u = _mm(paths["f"], lo, hi), _mm(paths["i"], lo, hi), _mm(paths["u"], lo, hi)
    fc, ic = oracle.OCol(oracle.T_DOUBLE, f, m[0]), oracle.OCol(oracle.T_INT64, i, m[1])
    aggs = [(fc, "sum", None), (fc, "mean", None), (ic, "sum", None), (ic, "min", None), (ic, "max", None),
            (ic, "count", None), (oracle.OCol(oracle.T_INT32, u, m[2]), "count_distinct", None)]
    keys = [oracle.OCol(oracle.T_INT32, _mm(paths["k"], lo, hi))]
    t0 = time.perf_counter()
    reps, _ = oracle.groupby(keys, aggs)
    return time.perf_counter() - t0, len(reps)


def _w_acero_join(job):
    """One JoinTask::cpu_variant: INNER hashjoin of lhs rows [llo, lhi) (probe) with rhs rows
    [rlo, rhi) (build); no nulls, so the nulls_equal emulation adds no mask columns."""
    import pyarrow as pa

    paths, llo, lhi, rlo, rhi = job
    lhs = pa.table({"k": _mm(paths["pk"], llo, lhi), "x": _mm(paths["pp"], llo, lhi)})
    rhs = pa.table({"kr": _mm(paths["bk"], rlo, rhi), "y": _mm(paths["bp"], rlo, rhi)})
    t0 = time.perf_counter()
    out = lhs.join(rhs, keys="k", right_keys="kr", join_type="inner", use_threads=False)
    return time.perf_counter() - t0, out.num_rows


def _w_oracle_join(job):
    import oracle

    paths, llo, lhi, rlo, rhi = job
    pk, bk = _mm(paths["pk"], llo, lhi), _mm(paths["bk"], rlo, rhi)
    t0 = time.perf_counter()
    lm, _ = oracle.join([oracle.OCol(oracle.T_INT64, pk)], [oracle.OCol(oracle.T_INT64, bk)], "inner")
    return time.perf_counter() - t0, len(lm)


class HostFiles:
    """Host copies of device columns as .npy files in a temporary directory (workers map them)."""

    def __init__(self):
        import tempfile

        self.dir = tempfile.mkdtemp(prefix="ldf_cpu_baseline_")
        self.paths = {}

    def put(self, name, tensor_or_array):
        import numpy as np

        a = tensor_or_array.cpu().numpy() if hasattr(tensor_or_array, "cpu") else np.asarray(tensor_or_array)
        path = os.path.join(self.dir, name + ".npy")
        np.save(path, a)
        self.paths[name] = path

    def close(self):
        import shutil

        shutil.rmtree(self.dir, ignore_errors=True)


def hash_slices(torch, key, nparts):
    """Stable hash partition of the rows by `key` (a device tensor): (row order, offsets)."""
    import numpy as np

    h = key.to(torch.int64) * -7046029254386353131  # 0x9E3779B97F4A7C15 as int64
    part = ((h >> 32) & 0x7FFFFFFF) % nparts
    order = torch.argsort(part, stable=True)
    counts = torch.bincount(part, minlength=nparts).cpu().numpy()
    return order, np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)


def cpu_baseline_groupby(torch, pool, nw, host, n):
    files = HostFiles()
    try:
        for name in ("k", "f", "i", "u"):
            files.put(name, host[name])
        for j, mk in enumerate(host["masks"]):
            files.put(f"m{j}", mk)
        whole = dict(files.paths)
        order, offs = hash_slices(torch, host["k"], nw)
        for name in ("k", "f", "i", "u"):
            files.put(name + "_p", host[name][order])
        for j, mk in enumerate(host["masks"]):
            files.put(f"m{j}_p", mk[order])
        del order
        parts = {key: files.paths[key + "_p"] for key in whole}
        log(f"[cpu] groupby: 1 core, Acero on {n} rows")
        t1, g1 = pool.apply(_w_acero_groupby, ((whole, 0, n),))
        log(f"[cpu] groupby: {nw} workers, Acero on hash slices")
        res = pool.map(_w_acero_groupby, [(parts, int(offs[j]), int(offs[j + 1])) for j in range(nw)])
        tn = max(r[0] for r in res)
        gn = sum(r[1] for r in res)
        log("[cpu] groupby: 1 core, the oracle (C restatement)")
        to, go = pool.apply(_w_oracle_groupby, ((whole, 0, n),))
    finally:
        files.close()
    assert g1 == gn == go, (g1, gn, go)
    return {"value": n / t1, "unit": "rows/s", "cores": 1, "kind": "reference",
            "engine": "pyarrow 25 Acero TableGroupBy(use_threads=False): the engine and options of "
                      "GroupByAggregationTask::cpu_variant (cpp/src/groupby_aggregation.cpp:87-91); the Legate task "
                      "wrapper around it is not built here",
            "sample": f"the whole workload: all {n} rows, 7 aggregations, {g1} groups, {t1:.1f} s on 1 core",
            "n_cores": {"value": n / tn, "cores": nw,
                        "note": f"{nw} single-threaded processes, each Acero on one hash partition of the rows "
                                "(whole groups per process, as Legate's CPU tasks after the shuffle; the shuffle "
                                "itself is not timed); value = rows / slowest process"},
            "port": {"value": n / to, "cores": 1, "kind": "port",
                     "note": "the oracle's single-threaded C restatement, same rows"}}


def cpu_baseline_join(torch, pool, nw, cols):
    pk, pp, bk, bp = cols
    n, m = pk.numel(), bk.numel()
    files = HostFiles()
    try:
        for name, t in (("pk", pk), ("pp", pp), ("bk", bk), ("bp", bp)):
            files.put(name, t)
        whole = dict(files.paths)
        lo_, loffs = hash_slices(torch, pk, nw)
        ro_, roffs = hash_slices(torch, bk, nw)
        files.put("pk_p", pk[lo_])
        files.put("pp_p", pp[lo_])
        files.put("bk_p", bk[ro_])
        files.put("bp_p", bp[ro_])
        parts = {key: files.paths[key + "_p"] for key in whole}
        t1, r1 = pool.apply(_w_acero_join, ((whole, 0, n, 0, m),))
        res = pool.map(_w_acero_join, [(parts, int(loffs[j]), int(loffs[j + 1]), int(roffs[j]), int(roffs[j + 1]))
                                       for j in range(nw)])
        to, ro = pool.apply(_w_oracle_join, ((whole, 0, n, 0, m),))
    finally:
        files.close()
    tn = max(r[0] for r in res)
    assert r1 == sum(r[1] for r in res) == ro, (r1, ro)
    return {"value": (n + m) / t1, "unit": "rows/s", "cores": 1, "kind": "reference",
            "engine": "pyarrow 25 Acero hashjoin (Table.join, use_threads=False; build = rhs): the engine and options "
                      "of JoinTask::cpu_variant (cpp/src/join.cpp:214-227)",
            "sample": f"the whole workload: {n} probe x {m} build rows, {r1} output rows, {t1:.2f} s on 1 core",
            "n_cores": {"value": (n + m) / tn, "cores": nw,
                        "note": f"{nw} single-threaded processes, each joining one hash partition of both sides"},
            "port": {"value": (n + m) / to, "cores": 1, "kind": "port", "note": "the oracle's C restatement"}}


def profiles_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, calibrated
    against a copy kernel of known size), recorded under profiles/ for this kernel/workload."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


# ---------------------------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------------------------
def run_groupby(args, torch, C):
    n = args.rows or 100_000_000
    cols, host = make_groupby_inputs(torch, C, n)
    torch.cuda.synchronize()
    ngroups = None
    for _ in range(args.warmup):
        rt = C.groupby_raw(cols, [0], GB_AGGS)
        ngroups = rt.cols[0].length
        rt.free()
    torch.cuda.synchronize()
    C.profile_reset()
    C.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rt = C.groupby_raw(cols, [0], GB_AGGS)
        ngroups = rt.cols[0].length
        rt.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    C.profile_enable(False)
    launches, kms = C.profile_get("gb_aggregate")
    ms_step = (t1 - t0) * 1e3 / args.steps
    inp, outb = groupby_alg_bytes(n, ngroups)
    k_avg_ms = kms / max(1, launches)
    achieved = inp / (k_avg_ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": n / (ms_step * 1e-3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64+f64", "data": "synthetic (seeded torch generator, device resident)",
        "config": {"workload": "groupby 100M rows: int32 key Zipf(1.1) over 1M; f64 U[-1,1) SUM,MEAN; int64 "
                               "U[-2^40,2^40] SUM,MIN,MAX,COUNT; int32 U[0,1e7) NUNIQUE; 1% nulls per value column",
                   "rows": n, "groups": ngroups, "form": "plain, one call per step"},
        "roofline": {"bound": "hbm", "kernel": "gb_aggregate", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": profiles_traffic("gb_aggregate_groupby100M"),
                     "traffic_source": "profiles/gb_aggregate_groupby100M_traffic.json (rocprofv3 --pmc FETCH_SIZE and "
                                       "WRITE_SIZE runs of this workload, copy-calibrated; not measured in this run)",
                     "kernel_ms": k_avg_ms, "launches": launches, "alg_bytes_per_launch": inp,
                     "whole_step_frac": (inp + outb) / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if CPU_POOL is not None:
        res["cpu_baseline"] = cpu_baseline_groupby(torch, CPU_POOL[0], CPU_POOL[1], host, n)
    return res


def run_join(args, torch, C):
    m, n = 1_000_000, args.rows or 10_000_000
    g = torch.Generator(device="cuda").manual_seed(1)
    bkeys = torch.randperm(m, device="cuda", generator=g).to(torch.int64)
    bpay = torch.randint(-(2**62), 2**62, (m,), device="cuda", generator=g, dtype=torch.int64)
    pkeys = torch.randint(0, 2 * m, (n,), device="cuda", generator=g, dtype=torch.int64)
    ppay = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    L = [dc(pkeys, C.INT64), dc(ppay, C.DOUBLE)]
    R = [dc(bkeys, C.INT64), dc(bpay, C.INT64)]
    for _ in range(args.warmup):
        C.join_raw(L, R, [0], [0], 0).free()
    torch.cuda.synchronize()
    C.profile_reset()
    C.profile_enable(True)
    t0 = time.perf_counter()
    nout = 0
    for _ in range(args.steps):
        rt = C.join_raw(L, R, [0], [0], 0)
        nout = rt.cols[0].length
        rt.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    C.profile_enable(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    alg = (n + m) * 16 + nout * 32
    kern = {k: C.profile_get(k) for k in JOIN_KERNELS}
    res = {"workload": "inner join build 1M (permutation keys) x probe 10M (uniform [0,2M)), i64 keys, "
                       "i64/f64 payloads", "rows_per_s": (n + m) / (ms_step * 1e-3), "ms_per_step": ms_step,
           "out_rows": nout, "alg_bytes": alg, "whole_step_frac": alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernels_ms": {k: v[1] / max(1, v[0]) for k, v in kern.items()}}
    if CPU_POOL is not None:
        log("[cpu] inner join 1M x 10M: Acero on 1 core and on hash slices")
        res["cpu_baseline"] = cpu_baseline_join(torch, CPU_POOL[0], CPU_POOL[1], (pkeys, ppay, bkeys, bpay))
    return res


JOIN_KERNELS = ("join_build", "join_count", "join_emit", "gather", "gather_str")


def run_join_c3(args, torch, C, join_type):
    n_probe = args.rows or 500_000_000
    n_build = max(1, n_probe // 10)
    L, R, _, _, n_vals = make_join_c3_inputs(torch, C, n_probe, n_build)
    torch.cuda.synchronize()
    inb = input_bytes(C, L) + input_bytes(C, R)
    for _ in range(args.warmup):
        C.join_raw(L, R, [0], [0], join_type).free()
    torch.cuda.synchronize()
    C.profile_reset()
    C.profile_enable(True)
    t0 = time.perf_counter()
    nout, outb = 0, 0
    for _ in range(args.steps):
        rt = C.join_raw(L, R, [0], [0], join_type)
        nout, outb = rt.cols[0].length, output_bytes(C, rt)
        rt.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    C.profile_enable(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern = {k: C.profile_get(k) for k in JOIN_KERNELS}
    name = {1: "left", 2: "full"}[join_type]
    del L, R
    torch.cuda.empty_cache()
    C.pool_release()
    return {"workload": f"{name} join build {n_build} (int64 key Zipf(0.8) over {n_vals}) x probe {n_probe} "
                        f"(int64 key U[0,{2 * n_vals}), 10% null), payloads int32/f64/timestamp[ms]/large_utf8 "
                        "(8-32 chars) on both sides, all 10 columns out",
            "rows_per_s": (n_probe + n_build) / (ms_step * 1e-3), "ms_per_step": ms_step, "out_rows": nout,
            "alg_bytes": inb + outb, "whole_step_frac": (inb + outb) / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernels_ms": {k: v[1] / max(1, v[0]) for k, v in kern.items()}}


BATCH_GB_AGGS = [(1, "sum", None), (2, "sum", None), (1, "count", None)]


def make_batched_inputs(torch, C, ninst=4096, seed=4):
    """4096 independent join + groupby instances: probe rows log-uniform in [1e3, 4e6], build =
    probe / 10, int64 keys U[0, 2 build) on both sides, build payload int64, probe payload f64;
    the groupby table of an instance has the probe's row count, an int32 key U[0, g) with g
    log-uniform in [1e3, 1e5], an int64 and a f64 value (SUM, SUM, COUNT)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    p = np.exp(rng.uniform(np.log(1e3), np.log(4e6), ninst)).astype(np.int64)
    b = np.maximum(1, p // 10)
    ng = np.exp(rng.uniform(np.log(1e3), np.log(1e5), ninst)).astype(np.int64)
    P, B = int(p.sum()), int(b.sum())
    g = torch.Generator(device="cuda").manual_seed(seed)
    pk = torch.empty(P, dtype=torch.int64, device="cuda")
    bk = torch.empty(B, dtype=torch.int64, device="cuda")
    gk = torch.empty(P, dtype=torch.int32, device="cuda")
    pp = torch.rand(P, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
    bp = torch.randint(-(2**62), 2**62, (B,), device="cuda", generator=g, dtype=torch.int64)
    gi = torch.randint(-(2**40), 2**40, (P,), device="cuda", generator=g, dtype=torch.int64)
    po = np.concatenate([[0], np.cumsum(p)])
    bo = np.concatenate([[0], np.cumsum(b)])
    for i in range(ninst):
        hi = int(2 * b[i])
        pk[po[i]:po[i + 1]].random_(0, hi, generator=g)
        bk[bo[i]:bo[i + 1]].random_(0, hi, generator=g)
        gk[po[i]:po[i + 1]].random_(0, int(ng[i]), generator=g)
    dc = lambda t, tid: C.DeviceColumn(tid, t.numel(), t.view(torch.uint8))  # noqa: E731
    L, R, G = [], [], []
    for i in range(ninst):
        ps, bs = slice(int(po[i]), int(po[i + 1])), slice(int(bo[i]), int(bo[i + 1]))
        L.append([dc(pk[ps], C.INT64), dc(pp[ps], C.DOUBLE)])
        R.append([dc(bk[bs], C.INT64), dc(bp[bs], C.INT64)])
        G.append([dc(gk[ps], C.INT32), dc(gi[ps], C.INT64), dc(pp[ps], C.DOUBLE)])
    keep = (pk, bk, gk, pp, bp, gi)
    return L, R, G, keep, dict(probe_rows=P, build_rows=B, instances=ninst)


def run_batched(args, torch, C):
    ninst = args.instances
    L, R, G, keep, info = make_batched_inputs(torch, C, ninst)
    torch.cuda.synchronize()
    log(f"[batched] {ninst} instances ready, {info['probe_rows']} probe rows")

    def plain_step():
        for i in range(ninst):
            C.join_raw(L[i], R[i], [0], [0], 0).free()
            C.groupby_raw(G[i], [0], BATCH_GB_AGGS).free()

    plain_step()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_step()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    C.pool_release()
    log(f"[batched] plain launches: {plain_ms:.0f} ms")

    t0 = time.perf_counter()
    bj = C.batch_join(L, R, [0], [0], 0, config={"use_graph": 1})
    bg = C.batch_groupby(G, [0], BATCH_GB_AGGS, config={"use_graph": 1})
    torch.cuda.synchronize()
    create_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(2):  # warm-up; the first run captures the graphs
        bj.run()
        bg.run()
    torch.cuda.synchronize()
    steps = max(3, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        bj.run()
        bg.run()
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / steps
    jc, joff = bj.result()
    gc, goff = bg.result()
    res = {"workload": f"{ninst} join + groupby instances: probe rows log-uniform [1e3, 4e6], build = probe/10, "
                       "int64 keys U[0, 2 build), int64/f64 payloads; groupby int32 key with 1e3..1e5 distinct, "
                       "SUM(int64), SUM(f64), COUNT",
           **info, "join_out_rows": int(joff[-1]), "groups": int(goff[-1]),
           "plain_ms": plain_ms, "batched_ms": batched_ms, "speedup": plain_ms / batched_ms,
           "batched_create_ms": create_ms, "speedup_incl_create": plain_ms / (batched_ms + create_ms),
           "rows_per_s_batched": (2 * info["probe_rows"] + info["build_rows"]) / (batched_ms * 1e-3),
           "config_used": {"join": bj.config_used(), "groupby": bg.config_used()}}
    bj.close()
    bg.close()
    del L, R, G, keep
    torch.cuda.empty_cache()
    C.pool_release()
    return res


SERVED_AGGS = [(1, "sum", None), (2, "sum", None), (None, "count_all", None)]


def make_served_inputs(torch, C, nq, seed=5):
    """8 device-resident tables of 10M..200M rows (log-spaced; int64 key uniform over the row count
    for even tables, Zipf(1.1) over it for odd ones; int64 and f64 payloads) and nq queries: half
    inner joins of a probe (int64 key uniform over the table's key range, f64 payload) with a table,
    half groupby SUM(int64), SUM(f64), COUNT(*) over a row range; sizes log-uniform [1e4, 5e7]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    sizes = [int(round(10e6 * 20 ** (i / 7))) for i in range(8)]
    tables = []
    for i, n in enumerate(sizes):
        k = zipf_keys(torch, n, n, 1.1, g, torch.int64) if i % 2 else \
            torch.randint(0, n, (n,), device="cuda", generator=g, dtype=torch.int64)
        a = torch.randint(-(2**40), 2**40, (n,), device="cuda", generator=g, dtype=torch.int64)
        f = torch.rand(n, device="cuda", generator=g, dtype=torch.float64) * 2 - 1
        tables.append([C.DeviceColumn(C.INT64, n, k.view(torch.uint8)), C.DeviceColumn(C.INT64, n, a.view(torch.uint8)),
                       C.DeviceColumn(C.DOUBLE, n, f.view(torch.uint8))])
    queries = []
    for j in range(nq):
        t = int(rng.integers(0, 8))
        n = sizes[t]
        s = int(np.exp(rng.uniform(np.log(1e4), np.log(5e7))))
        if j % 2 == 0:
            pk = torch.randint(0, n, (s,), device="cuda", generator=g, dtype=torch.int64)
            pp = torch.rand(s, device="cuda", generator=g, dtype=torch.float64)
            queries.append(("join", t, [C.DeviceColumn(C.INT64, s, pk.view(torch.uint8)),
                                        C.DeviceColumn(C.DOUBLE, s, pp.view(torch.uint8))]))
        else:
            s = min(s, n)
            b = int(rng.integers(0, (n - s) // 64 + 1)) * 64
            queries.append(("groupby", t, (b, min(n, b + s))))
    return tables, queries, sizes


def _table_slice(C, cols, b, e):
    return [C.DeviceColumn(c.type_id, e - b, c.data[b * C.WIDTH[c.type_id]:e * C.WIDTH[c.type_id]]) for c in cols]


def run_served(args, torch, C):
    import threading

    import numpy as np

    nq = args.queries
    t_start = time.perf_counter()
    tables, queries, sizes = make_served_inputs(torch, C, nq)
    torch.cuda.synchronize()
    log(f"[served] inputs ready: 8 tables, {nq} queries ({time.perf_counter() - t_start:.1f} s)")

    # plain sequential: one plain call per query (the join builds the table's hash table each time)
    svc = []
    for j, (kind, t, arg) in enumerate(queries):
        if j % 20 == 0:
            log(f"[served] plain sequential: query {j}/{nq} ({time.perf_counter() - t_start:.1f} s)")
        t0 = time.perf_counter()
        if kind == "join":
            C.join_raw(arg, tables[t], [0], [0], 0).free()
        else:
            C.groupby_raw(_table_slice(C, tables[t], *arg), [0], SERVED_AGGS).free()
        torch.cuda.synchronize()
        svc.append(time.perf_counter() - t0)
    svc = np.array(svc)
    C.pool_release()

    srv = C.Server(hbm_budget_bytes=200e9, slots=4, max_batch=64)
    t0 = time.perf_counter()
    for tb in tables:
        srv.add_table(tb)
    index_ms = (time.perf_counter() - t0) * 1e3
    log(f"[served] server up, indexes built in {index_ms:.0f} ms")

    def serve(arrivals):
        ids = [None] * nq
        start = time.perf_counter()

        def submitter():
            for j, (kind, t, arg) in enumerate(queries):
                d = start + arrivals[j] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[j] = srv.submit_join(t, arg, 0) if kind == "join" else srv.submit_groupby(t, arg[0], arg[1],
                                                                                              SERVED_AGGS)

        th = threading.Thread(target=submitter)
        th.start()
        lat, ends, firsts, batch = [], [], [], []
        for j in range(nq):
            while ids[j] is None:
                time.sleep(0.0002)
            if j % 50 == 0:
                log(f"[served] waiting for query {j}/{nq}")
            _, st = srv.wait(ids[j])
            srv.release(ids[j])
            lat.append((st["end_ns"] - st["submit_ns"]) * 1e-6)
            ends.append(st["end_ns"])
            firsts.append(st["submit_ns"])
            batch.append(st["batch_queries"])
        th.join()
        lat = np.array(lat)
        wall = (max(ends) - min(firsts)) * 1e-9
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "queries_per_s": nq / wall, "mean_batch": float(np.mean(batch))}

    burst = serve(np.zeros(nq))  # every query at once: the server's capacity
    log(f"[served] burst done: {burst}")
    rate = 0.7 * burst["queries_per_s"]
    rng = np.random.default_rng(55)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nq))
    poisson = serve(arrivals)
    log(f"[served] poisson done: {poisson}")
    # one query at a time through the same server: persistent join indexes, no batching or overlap
    seq = []
    for kind, t, arg in queries:
        t0 = time.perf_counter()
        qid = srv.submit_join(t, arg, 0) if kind == "join" else srv.submit_groupby(t, arg[0], arg[1], SERVED_AGGS)
        srv.wait(qid)
        srv.release(qid)
        seq.append(time.perf_counter() - t0)
    seq = np.array(seq)
    log(f"[served] sequential with persistent indexes: {nq / seq.sum():.1f} q/s")
    srv.close()
    # the plain path under the same arrivals: one query at a time, first come first served
    t_free, lat_plain = 0.0, []
    for a, s_ in zip(arrivals, svc):
        t_free = max(t_free, a) + s_
        lat_plain.append((t_free - a) * 1e3)
    lat_plain = np.array(lat_plain)
    del tables, queries
    torch.cuda.empty_cache()
    C.pool_release()
    return {"workload": f"{nq} queries (half inner joins of a probe with a table, half groupby SUM,SUM,COUNT over "
                        "a row range; sizes log-uniform [1e4, 5e7]) over 8 resident tables of "
                        f"{sizes[0] // 10**6}M..{sizes[-1] // 10**6}M rows (int64 keys uniform / Zipf(1.1)); "
                        "HBM budget 200 GB, 4 slots",
            "served_burst": burst, "served_poisson": {"arrival_rate_per_s": rate, **poisson},
            "plain_sequential": {"queries_per_s": nq / float(svc.sum()), "mean_service_ms": float(svc.mean() * 1e3),
                                 "p50_ms_at_same_arrivals": float(np.percentile(lat_plain, 50)),
                                 "p99_ms_at_same_arrivals": float(np.percentile(lat_plain, 99))},
            "sequential_with_indexes": {"queries_per_s": nq / float(seq.sum()),
                                        "mean_service_ms": float(seq.mean() * 1e3),
                                        "note": "each query alone through the server (its persistent join "
                                                "indexes), no batching or overlap"},
            "throughput_vs_sequential": burst["queries_per_s"] / (nq / float(svc.sum())),
            "throughput_vs_sequential_with_indexes": burst["queries_per_s"] / (nq / float(seq.sum())),
            "index_build_ms": index_ms}


def run_partition(args, torch, C):
    n, P = args.rows or 1_000_000_000, 256
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randint(-(2**63), 2**63 - 1, (n,), device="cuda", generator=g, dtype=torch.int64)
    col = [C.DeviceColumn(C.INT64, n, x.view(torch.uint8))]
    for _ in range(args.warmup):
        rt, _ = C.hash_partition_raw(col, [0], P)
        rt.free()
    torch.cuda.synchronize()
    C.profile_reset()
    C.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rt, _ = C.hash_partition_raw(col, [0], P)
        rt.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    C.profile_enable(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern = {k: C.profile_get(k) for k in ("hp_hist", "hp_scatter")}
    alg = 16 * n
    return {"workload": f"hash_partition {n} int64 keys -> {P} partitions (murmur3, seed 0)",
            "keys_per_s": n / (ms_step * 1e-3), "ms_per_step": ms_step, "alg_bytes": alg,
            "whole_step_frac": alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernels_ms": {k: v[1] / max(1, v[0]) for k, v in kern.items()}}


def run_distinct(args, torch, C):
    n = args.rows or 100_000_000
    g = torch.Generator(device="cuda").manual_seed(6)
    nk = args.distinct_keys or 10_000_000
    if nk >= n:  # every key distinct: a permutation of the rows
        k = torch.randperm(n, device="cuda", generator=g).to(torch.int64)
    else:
        k = torch.randint(0, nk, (n,), device="cuda", generator=g, dtype=torch.int64)
    v = torch.rand(n, device="cuda", generator=g, dtype=torch.float64)
    cols = [C.DeviceColumn(C.INT64, n, k.view(torch.uint8)), C.DeviceColumn(C.DOUBLE, n, v.view(torch.uint8))]
    for _ in range(args.warmup):
        out = C.distinct(cols, [0])
        del out
    torch.cuda.synchronize()
    C.profile_reset()
    C.profile_enable(True)
    m = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = C.distinct(cols, [0])
        m = out[0].length
        del out
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    C.profile_enable(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern = {k: C.profile_get(k) for k in ("distinct_insert",)}
    alg = 8 * n + 32 * m  # the key column once; the kept rows' two columns read and written once
    keys = "a permutation of the rows (every key distinct)" if nk >= n else f"U[0, {nk})"
    return {"workload": f"distinct {n} rows by an int64 key {keys} with a float64 column (first row per key)",
            "rows_per_s": n / (ms_step * 1e-3), "ms_per_step": ms_step, "out_rows": m, "alg_bytes": alg,
            "whole_step_frac": alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernels_ms": {k: v[1] / max(1, v[0]) for k, v in kern.items()}}


def copy_calibration(torch, nbytes=4 << 30, reps=5):
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def hbm_ceiling(torch, C, nbytes=8_000_000_000, reps=3):
    """What HBM gives on this GPU: the library's read and copy probe kernels (16-byte loads, four in
    flight per thread) over 8 GB, best of `reps` per grid size; TB/s counts read plus write bytes."""
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    out = {"bytes": nbytes, "read": {}, "copy": {}}
    for grid in (1024, 2048, 4096, 8192):
        r = min(C.hbm_probe(src, dst, nbytes, False, grid) for _ in range(reps))
        c = min(C.hbm_probe(src, dst, nbytes, True, grid) for _ in range(reps))
        out["read"][grid] = {"ms": r, "tb_s": nbytes / (r * 1e-3) / 1e12}
        out["copy"][grid] = {"ms": c, "tb_s": 2 * nbytes / (c * 1e-3) / 1e12}
    out["read_best_tb_s"] = max(v["tb_s"] for v in out["read"].values())
    out["copy_best_tb_s"] = max(v["tb_s"] for v in out["copy"].values())
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--workload",
                    choices=["groupby", "join", "join_left", "join_full", "partition", "batched", "served",
                             "distinct"],
                    default="groupby")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--queries", type=int, default=200)
    ap.add_argument("--distinct-keys", type=int, default=0,
                    help="distinct workload: key values U[0, K) (default 1e7); K >= rows gives every key distinct")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="time the library's HBM read and copy probe kernels over 8 GB and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the path on a single MI355X (no collectives); "
              "multi-GPU runs are not supported.", file=sys.stderr)
        sys.exit(2)

    global CPU_POOL
    if not args.no_cpu and args.workload in ("groupby", "join") and not args.copy_calibration \
            and not args.hbm_ceiling:
        CPU_POOL = start_cpu_pool()  # before the GPU is initialised
    try:
        _main(args)
    finally:
        if CPU_POOL is not None:
            CPU_POOL[0].close()
            CPU_POOL[0].join()


def _main(args):
    import torch

    from legate_dataframe import _capi as C

    C.load()
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.hbm_ceiling:
        hbm_ceiling(torch, C)
        return
    if args.workload == "groupby":
        res = run_groupby(args, torch, C)
        if not args.no_secondary:
            sa = argparse.Namespace(**vars(args))
            sa.rows, sa.steps, sa.warmup = 0, max(3, args.steps // 2), 2
            log("[bench] secondary: inner join 1M x 10M")
            res["secondary"] = {"join_1Mx10M": run_join(sa, torch, C)}
            sa.rows = 0
            log("[bench] secondary: hash partition 1B -> 256")
            res["secondary"]["partition_1B_256"] = run_partition(sa, torch, C)
            C.pool_release()
            sa.rows, sa.steps, sa.warmup = 0, 3, 1
            log("[bench] secondary: left join 50M x 500M")
            res["secondary"]["left_join_50Mx500M"] = run_join_c3(sa, torch, C, 1)
            sa.steps = 3
            log("[bench] secondary: batched 4096 instances")
            res["secondary"]["batched_4096"] = run_batched(sa, torch, C)
            log("[bench] secondary: served 200 queries")
            res["secondary"]["served_200"] = run_served(sa, torch, C)
    elif args.workload == "join":
        res = run_join(args, torch, C)
    elif args.workload == "batched":
        res = run_batched(args, torch, C)
    elif args.workload == "served":
        res = run_served(args, torch, C)
    elif args.workload == "distinct":
        res = run_distinct(args, torch, C)
    elif args.workload in ("join_left", "join_full"):
        res = run_join_c3(args, torch, C, 1 if args.workload == "join_left" else 2)
    else:
        res = run_partition(args, torch, C)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
