"""Benchmark of the GROUP BY operator on one MI355X.

Workload (the headline high-cardinality aggregate):
  SELECT k, COUNT(*), SUM(v1), MIN(v2), MAX(v2), AVG(v3) FROM t GROUP BY k
  100M rows; k int64 = sparse ids of Zipf(1.1) ranks over 10M values;
  v1 int64 uniform [-1e9, 1e9]; v2 int32 uniform; v3 double Normal(100, 25)
  (seeded synthetic data, slothdb_amd/csrc/datagen.cpp).

A step is one full pass of the operator over the resident input: Init
(reset), consume_device, finalize (HLL sizing, table init, aggregate,
compaction). Inputs sit in HBM before the timed region; results stay in HBM
(the host copy is not part of the step).

Prints one JSON line. `roofline` is for the dominant kernel of the radix
strategy, k_rp_scatter (the one pass that reads every input row): algorithmic
bytes = 28 B/row of input read once, over the kernel's average duration
measured with HIP events on the operator's stream (plan_info ms_aggregate).
`traffic` is the HBM bytes one launch moves, from rocprofv3 FETCH_SIZE and
WRITE_SIZE, read from profiles/k_rp_scatter_highcard_100m_traffic.json.
`cpu_baseline` times the reference's own engine (kind "reference"): its
sources compiled into oracle/_ref/ref_query (oracle/Makefile, target cli),
which runs the same query through the engine's C API on the first 20M rows of
the workload written as Parquet (the engine's fused Parquet scan + per-thread
aggregation + merge + result materialization, its default 8 threads; median
of 3 runs after a warm-up run). The restated CPU path (oracle/oracle.c
orc_port_groupby, the per-thread-map path of physical_planner.cpp:7613-7931)
on the same rows in memory is reported beside it. Without oracle/_ref/ref_query
the restatement is the baseline (kind "port").

`--ref-calibration` times the reference engine and the restatement on the
three aggregate configurations (plain 1M, high-cardinality 20M of 100M,
composite 5M of 50M), median of 5 runs each, then exits.

`--copy-calibration` instead copies 1 GiB device to device three times with
hipMemcpy: the reference point that scales the rocprofv3 byte counters.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Grouped rows/sec + HBM GB/s % of peak; learned/batched vs plain GPU speedup"
HBM_PEAK_GBS = 8000.0
ROW_BYTES = 8 + 8 + 4 + 8     # k, v1, v2, v3
GROUP_BYTES = 8 + 8 + 8 + 4 + 4 + 8  # key, count, sum, min, max, avg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--keys", type=int, default=10_000_000)
    ap.add_argument("--strategy", type=int, default=0)
    ap.add_argument("--cpu-sample-rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--ref-sample-rows", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-queries", type=int, default=512)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched mix (no plain runs), print its JSON, then exit")
    ap.add_argument("--batch-tune", default=None, metavar="TILE,SLOTS,BLOCK,WPC",
                    help="with --batch-only: this launch shape instead of autotuning")
    ap.add_argument("--no-strings", action="store_true")
    ap.add_argument("--no-parquet", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the plain and composite configurations")
    ap.add_argument("--cpu-calibration", action="store_true",
                    help="time the CPU restatement against the reference's own RadixCountAgg "
                         "(oracle/_ref, built from the reference sources), then exit")
    ap.add_argument("--ref-calibration", action="store_true",
                    help="time the reference engine (oracle/_ref/ref_query, built from the reference "
                         "sources) and the CPU restatement on the three aggregate configurations, then exit")
    ap.add_argument("--batch-split", action="store_true",
                    help="time the batched mix split by the queries' distinct keys, then exit")
    ap.add_argument("--param", action="append", default=[], metavar="NAME=VALUE",
                    help="operator tuning knob (gbagg_set_param), e.g. part_bits=8; repeatable")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_calibration:
        cpu_calibration(args.rows, args.keys, args.cpu_threads)
        return
    if args.batch_split:
        batch_split()
        return
    if args.ref_calibration:
        ref_calibration(args.cpu_threads)
        return
    if args.batch_only:
        tune = [int(x) for x in args.batch_tune.split(",")] if args.batch_tune else None
        print(json.dumps({"batched": bench_batched(args.batch_queries, plain=False, tune=tune)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from slothdb_amd import datagen, gbagg as G

    t0 = time.time()
    q, cols = datagen.highcard_workload(n=args.rows, nkeys=args.keys)
    q.strategy = args.strategy
    print(f"[bench] generated {args.rows} rows in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dcols = [G.DeviceColumn(c) for c in cols]
    knobs = {k: int(v) for k, v in (p.split("=", 1) for p in args.param)}
    op = G.GpuHashAggregate(q, knobs or None)
    op.Init()

    def step():
        op.Init()
        op.consume_device(dcols)
        op.finalize()
        return op.plan_info()

    for _ in range(args.warmup):
        step()
    _check = G.lib().gbagg_device_synchronize
    _check()
    agg_ms, tot_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info = step()
        agg_ms.append(info["ms_aggregate"])
        tot_ms.append(info["ms_total"])
    _check()
    elapsed = time.perf_counter() - t0
    groups = info["groups"]
    ms_per_step = elapsed * 1e3 / args.steps
    rows_per_s = args.rows * args.steps / elapsed
    agg_avg = sum(agg_ms) / len(agg_ms)
    tot_avg = sum(tot_ms) / len(tot_ms)
    agg_bytes = args.rows * ROW_BYTES
    achieved = agg_bytes / (agg_avg * 1e-3) / 1e9
    pipe_bytes = agg_bytes + groups * GROUP_BYTES
    pipe_gbs = pipe_bytes / (tot_avg * 1e-3) / 1e9
    print(f"[bench] step {ms_per_step:.3f} ms (device {tot_avg:.3f} ms, aggregate kernel {agg_avg:.3f} ms), "
          f"groups {groups}, plan {info}", file=sys.stderr, flush=True)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_rp_scatter_highcard_100m_traffic.json")
    if os.path.exists(tpath) and info["strategy"] == 4 and args.rows == 100_000_000:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        m = min(args.cpu_sample_rows, args.rows)
        sample = [G.Column(c.type, c.values[:m]) for c in cols]
        O.port_groupby(q, [G.Column(c.type, c.values[:100_000]) for c in cols], args.cpu_threads)
        t1 = time.perf_counter()
        O.port_groupby(q, sample, args.cpu_threads)
        cpu_s = time.perf_counter() - t1
        port = {"value": m / cpu_s, "unit": "rows/s", "cores": args.cpu_threads, "kind": "port",
                "sample": f"first {m} rows of the same workload, restated per-thread-map path "
                          f"(physical_planner.cpp:7613-7931), in memory, {cpu_s:.2f} s wall"}
        print(f"[bench] cpu port {m} rows in {cpu_s:.2f}s", file=sys.stderr, flush=True)
        ref = reference_time("highcard", q, cols, min(args.ref_sample_rows, args.rows), reps=3)
        if ref:
            cpu = {"value": ref["rows_per_s"], "unit": "rows/s", "cores": ref["threads"],
                   "host_cores": len(os.sched_getaffinity(0)), "kind": "reference",
                   "sample": ref["sample"], "scan_aggregate_s": ref["scan_aggregate_s"], "port": port}
            print(f"[bench] cpu reference {ref['rows']} rows: {ref['median_s']:.2f}s", file=sys.stderr, flush=True)
        else:
            cpu = dict(port, host_cores=len(os.sched_getaffinity(0)))

    configs = None if args.no_configs else bench_configs()
    batched = None if args.no_batch else bench_batched(args.batch_queries)
    strings = None if args.no_strings else bench_strings()
    parquet = None if args.no_parquet else bench_parquet()

    line = {
        "metric": METRIC, "value": rows_per_s, "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "highcard_groupby_100M", "rows": args.rows, "distinct_ids": args.keys,
                   "key_dist": "zipf(1.1) over sparse int64 ids", "groups": groups,
                   "aggs": "COUNT(*), SUM(v1 int64), MIN/MAX(v2 int32), AVG(v3 double)",
                   "strategy": info["strategy"], "selector_used": info["selector_used"],
                   "selector_path": info["selector_path"], "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_rp_scatter" if info["strategy"] == 4 else "k_aggregate",
                     "bytes_per_launch": agg_bytes, "avg_ms": agg_avg},
        "pipeline": {"algorithmic_bytes": pipe_bytes, "device_ms": tot_avg, "GB/s": pipe_gbs,
                     "frac": pipe_gbs / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "configs": configs,
        "batched": batched,
        "strings": strings,
        "parquet": parquet,
    }
    print(json.dumps(line), flush=True)
    op.close()


def bench_parquet(n=20_000_000, runs=5):
    """The high-cardinality columns (k, v1, v2, v3) for n rows written by
    pyarrow (uncompressed, 122,880-row row groups, its default dictionary
    encoding, which falls back to PLAIN pages for these columns), decoded on
    the GPU from an HBM-resident file image: the whole file in one call, then
    decoded and aggregated (the high-cardinality GROUP BY) as one batch.
    Checked against the same query over the generated columns."""
    try:
        import io
        import pyarrow as pa
        import pyarrow.parquet as pq
    except ImportError:
        return None
    import numpy as np
    from slothdb_amd import datagen, gbagg as G
    t0 = time.time()
    q, cols = datagen.highcard_workload(n=n)
    schema = pa.schema([pa.field(nm, t, nullable=False) for nm, t in
                        [("k", pa.int64()), ("v1", pa.int64()), ("v2", pa.int32()), ("v3", pa.float64())]])
    tab = pa.table({"k": cols[0].values, "v1": cols[1].values, "v2": cols[2].values, "v3": cols[3].values},
                   schema=schema)
    buf = io.BytesIO()
    pq.write_table(tab, buf, compression="NONE", row_group_size=122_880)
    image = buf.getvalue()
    del tab, buf
    print(f"[bench] parquet file {len(image)} bytes in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    sync = G.lib().gbagg_device_synchronize
    f = G.ParquetFile(image)
    f.upload()
    nrg = f.num_row_groups
    dec = []
    for i in range(runs + 1):
        f.reset()
        sync()
        t = time.perf_counter()
        f.decode(0, last=nrg)
        sync()
        if i:
            dec.append((time.perf_counter() - t) * 1e3)
    per_rg = []  # one decode call per row group (the RGConsumerFn shape), each waiting
    for i in range(3):
        f.reset()
        sync()
        t = time.perf_counter()
        for rg in range(nrg):
            f.decode(rg)
        sync()
        if i:
            per_rg.append((time.perf_counter() - t) * 1e3)
    # one queued call per row group (gbagg_parquet_decode_async), one wait at
    # the end; the C calls are made with prepared arguments, so the loop
    # measures the entry point rather than Python's wrapper objects
    import ctypes as C
    dec_async = G.lib().gbagg_parquet_decode_async
    ncols = f.num_columns
    idx = np.arange(ncols, dtype=np.uint32)
    views = (G.ColumnView * ncols)()
    nout = C.c_uint64()
    a_idx, a_views, a_n = idx.ctypes.data, C.addressof(views), C.byref(nout)
    per_rg_async = []
    for i in range(runs + 1):
        f.reset()
        sync()
        t = time.perf_counter()
        for rg in range(nrg):
            st = dec_async(f._h, rg, a_idx, ncols, a_views, a_n)
            if st:
                G._check(st)
        f.sync()
        if i:
            per_rg_async.append((time.perf_counter() - t) * 1e3)
    op = G.GpuHashAggregate(q)
    full = []
    rows = None
    for i in range(runs + 1):
        f.reset()
        sync()
        t = time.perf_counter()
        dcols = f.decode(0, last=nrg)
        op.Init()
        op.consume_device(dcols)
        op.finalize()
        info = op.plan_info()
        sync()
        if i:
            full.append((time.perf_counter() - t) * 1e3)
        rows = info["groups"]
    got = sorted(op.result().rows())
    # the engine's row-group consumer: each row group queued and its views
    # handed to the operator at once, one wait before finalize
    rg_full = []
    for i in range(runs + 1):
        f.reset()
        sync()
        t = time.perf_counter()
        op.Init()
        for rg in range(nrg):
            op.consume_device(f.decode_async(rg))
        f.sync()
        op.finalize()
        sync()
        if i:
            rg_full.append((time.perf_counter() - t) * 1e3)
    got_rg = sorted(op.result().rows())
    op.close()
    f.close()
    # the same row groups as host batches (gbagg_consume: buffers valid only
    # during the call), from the generated columns
    bounds = [(a, min(n, a + 122_880)) for a in range(0, n, 122_880)]
    host_rg = []
    op = G.GpuHashAggregate(q)
    for i in range(3):
        sync()
        t = time.perf_counter()
        op.Init()
        for a, b in bounds:
            op.consume([G.Column(c.type, c.values[a:b], None) for c in cols])
        op.finalize()
        sync()
        if i:
            host_rg.append((time.perf_counter() - t) * 1e3)
    got_host = sorted(op.result().rows())
    op.close()
    ref = sorted(G.run(q, cols, device=True)[0].rows())
    same = got == ref and got_rg == ref and got_host == ref  # exact, AVG included
    ms = min(dec)
    print(f"[bench] parquet: {n} rows, {len(image)} bytes: decode {ms:.2f} ms ({min(per_rg):.2f} ms as {nrg} row "
          f"group calls that wait, {min(per_rg_async):.2f} ms queued), decode + aggregate {min(full):.2f} ms "
          f"({min(rg_full):.2f} ms queued per row group; {min(host_rg):.2f} ms as {len(bounds)} host batches), "
          f"groups {rows}, equal to the in-memory run {same}", file=sys.stderr, flush=True)
    return {"workload": f"high-cardinality columns, {n} rows, pyarrow Parquet (uncompressed, 122880-row row groups)",
            "file_bytes": len(image), "row_groups": nrg, "decode_ms": ms,
            "decode_rows_per_s": n / (ms * 1e-3), "decode_file_GB/s": len(image) / (ms * 1e-3) / 1e9,
            "decode_per_row_group_ms": min(per_rg), "decode_per_row_group_queued_ms": min(per_rg_async),
            "decode_aggregate_ms": min(full), "decode_aggregate_per_row_group_queued_ms": min(rg_full),
            "host_batches_aggregate_ms": min(host_rg), "groups": rows, "equal_to_in_memory": same}


def bench_strings(runs=5):
    """The composite workload (50M rows, WHERE c < 30 GROUP BY a, b ORDER BY
    COUNT(*) DESC LIMIT 10) with b given as its strings (offsets + bytes, 1M
    distinct strings of 4-32 bytes) rather than dictionary codes, HBM-resident:
    the GPU interns b, then aggregates as over codes. Best of `runs` per form."""
    from slothdb_amd import datagen, gbagg as G
    t0 = time.time()
    q, cols = datagen.composite_workload()
    scol = datagen.strings_from_codes(cols[1])
    print(f"[bench] strings workload in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dcodes = [G.DeviceColumn(c) for c in cols]
    dstr = G.DeviceColumn(scol)
    sync = G.lib().gbagg_device_synchronize

    def best(dcols):
        op = G.GpuHashAggregate(q)
        out = None
        for i in range(runs + 1):
            sync()
            t = time.perf_counter()
            op.Init()
            op.consume_device(dcols)
            op.finalize()
            rows = op.result().rows()
            sync()
            t = (time.perf_counter() - t) * 1e3
            info = op.plan_info()
            if i and (out is None or t < out[0]):
                out = (t, info, rows)
        op.close()
        return out

    t_codes, info_c, rows_c = best(dcodes)
    t_str, info_s, rows_s = best([dcodes[0], dstr, dcodes[2]])
    same = sorted(r[2] for r in rows_c) == sorted(r[2] for r in rows_s)
    n = len(scol)
    nbytes = int(scol.values[-1])
    ms = info_s["ms_strings"]
    # interning reads the offsets and the bytes once and writes a code per row
    io_bytes = n * 8 + nbytes
    for c in dcodes + [dstr]:
        c.free()
    print(f"[bench] strings: {n} rows, {nbytes} bytes, {info_s['strings_distinct']} distinct: intern {ms:.2f} ms, "
          f"query {t_str:.2f} ms with strings vs {t_codes:.2f} ms with codes, same top counts {same}",
          file=sys.stderr, flush=True)
    return {"workload": "composite 50M rows, WHERE c < 30 GROUP BY a, b ORDER BY COUNT(*) DESC LIMIT 10, "
                        "b as strings (1M distinct, 4-32 B)",
            "rows": n, "string_bytes": nbytes, "distinct": info_s["strings_distinct"],
            "intern_ms": ms, "intern_rows_per_s": n / (ms * 1e-3), "intern_GB/s": io_bytes / (ms * 1e-3) / 1e9,
            "query_ms_strings": t_str, "query_ms_codes": t_codes, "top_counts_equal": same}


def bench_configs(steps=10):
    """The other two aggregate configurations, HBM-resident, AUTO plan, the
    same step as the headline (Init + consume_device + finalize, results left
    in HBM): the plain aggregate (1M rows, 1000 int32 keys; launch-bound, its
    HBM fraction is reported but not meaningful) and the composite key with
    filter and TopN (50M rows, int32 a + dictionary-coded VARCHAR b, WHERE
    c < 30, ORDER BY COUNT(*) DESC LIMIT 10). Algorithmic bytes: the input
    columns read once plus the result written once."""
    from slothdb_amd import datagen, gbagg as G
    sync = G.lib().gbagg_device_synchronize
    out = {}
    for name, (q, cols), in_bytes, out_bytes in (
            ("plain_1M", datagen.plain_workload(n=1_000_000), 1_000_000 * 12, 1000 * 36),
            ("composite_filter_topn_50M", datagen.composite_workload(), 50_000_000 * 12, 10 * 20)):
        dcols = [G.DeviceColumn(c) for c in cols]
        op = G.GpuHashAggregate(q)
        for _ in range(3):
            op.Init(); op.consume_device(dcols); op.finalize()
        sync()
        dev = []
        t = time.perf_counter()
        for _ in range(steps):
            op.Init(); op.consume_device(dcols); op.finalize()
            dev.append(op.plan_info()["ms_total"])
        sync()
        ms = (time.perf_counter() - t) * 1e3 / steps
        info = op.plan_info()
        dms = sum(dev) / len(dev)
        n = len(cols[0].values)
        gbs = (in_bytes + out_bytes) / (dms * 1e-3) / 1e9
        out[name] = {"rows": n, "ms_per_step": ms, "rows_per_s": n / (ms * 1e-3), "device_ms": dms,
                     "algorithmic_bytes": in_bytes + out_bytes, "GB/s": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                     "strategy": info["strategy"], "selector_used": info["selector_used"],
                     "groups": info["groups"]}
        print(f"[bench] {name}: {ms:.3f} ms per step (device {dms:.3f} ms), {n / (ms * 1e-3):.3g} rows/s, "
              f"strategy {info['strategy']}", file=sys.stderr, flush=True)
        op.close()
        for c in dcols:
            c.free()
    return out


def cpu_calibration(n, nkeys, threads):
    """COUNT(*) GROUP BY k over the high-cardinality keys: the oracle's
    restatement of RadixCountAgg (radix_count_agg.cpp:17-106) against the
    reference's own class compiled into oracle/_ref, same threads, same keys;
    best of 3 each. Their outputs must be the same (key, count) set."""
    import ctypes as C
    import numpy as np
    from oracle import oracle as O
    from slothdb_amd import datagen
    L = O.ref()
    if L is None:
        print(json.dumps({"cpu_calibration": None, "reason": "oracle/_ref not built"}), flush=True)
        return
    keys = np.ascontiguousarray(datagen.highcard(n=n, nkeys=nkeys, columns=("k",))["k"], np.int64)
    cap = min(n, nkeys) + 16
    ok, oc, m = np.empty(cap, np.int64), np.empty(cap, np.int64), C.c_uint64()
    t_ref, t_port = [], []
    for _ in range(3):
        t = time.perf_counter()
        L.ref_radix_count(keys.ctypes.data, len(keys), threads, 0, ok.ctypes.data, oc.ctypes.data, cap, C.byref(m))
        t_ref.append(time.perf_counter() - t)
        t = time.perf_counter()
        pk, pc = O.port_radix_count(keys, threads)
        t_port.append(time.perf_counter() - t)
    ref = dict(zip(ok[:m.value].tolist(), oc[:m.value].tolist()))
    same = ref == dict(zip(pk.tolist(), pc.tolist()))
    print(json.dumps({"cpu_calibration": {
        "workload": f"COUNT(*) GROUP BY k, {n} rows, high-cardinality keys", "threads": threads,
        "reference_radix_count_s": min(t_ref), "port_radix_count_s": min(t_port),
        "port_over_reference": min(t_port) / min(t_ref), "same_result": same}}), flush=True)


REF_QUERY = os.path.join(ROOT, "oracle", "_ref", "ref_query")
# Parquet column names and the query, per configuration (the columns of the
# generated workloads, in order)
REF_CONFIGS = {
    "plain": (("k", "v"), "SELECT k, COUNT(*), SUM(v), MIN(v), MAX(v) FROM '{f}' GROUP BY k"),
    "highcard": (("k", "v1", "v2", "v3"),
                 "SELECT k, COUNT(*), SUM(v1), MIN(v2), MAX(v2), AVG(v3) FROM '{f}' GROUP BY k"),
    "composite": (("a", "b", "c"),
                  "SELECT a, b, COUNT(*), SUM(c) FROM '{f}' WHERE c < 30 GROUP BY a, b "
                  "ORDER BY COUNT(*) DESC LIMIT 10"),
}


def _arrow(col, n):
    """The first n rows of a workload column as a pyarrow array (VARCHAR codes
    with their dictionary as a dictionary-encoded string column)."""
    import numpy as np
    import pyarrow as pa
    v = col.values[:n]
    if col.dictionary is not None:
        d = col.dictionary
        dic = pa.StringArray.from_buffers(len(d), pa.py_buffer(d.offsets.astype(np.int32)), pa.py_buffer(d.data))
        return pa.DictionaryArray.from_arrays(pa.array(v.astype(np.int32)), dic)
    return pa.array(v)


def reference_time(config, q, cols, n, reps):
    """The reference engine's own time for configuration `config` on the first
    n rows: the rows written as Parquet (uncompressed, 122,880-row row groups)
    to a temporary directory, then oracle/_ref/ref_query runs the query
    reps + 1 times through the engine's C API (SLOTH_PROFILE=1; the first run
    warms the page cache and is dropped). None when ref_query or pyarrow is
    missing."""
    import shutil
    import statistics
    import subprocess
    import tempfile
    if not os.path.exists(REF_QUERY):
        return None
    try:
        import pyarrow as pa
        import pyarrow.parquet as pq
    except ImportError:
        return None
    names, sql = REF_CONFIGS[config]
    tmp = tempfile.mkdtemp(prefix="slothdb_ref_")
    try:
        path = os.path.join(tmp, config + ".parquet")
        pq.write_table(pa.table({nm: _arrow(c, n) for nm, c in zip(names, cols)}), path,
                       compression="NONE", row_group_size=122_880)
        r = subprocess.run([REF_QUERY, str(reps + 1), sql.format(f=path)], env=dict(os.environ, SLOTH_PROFILE="1"),
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900)
        if r.returncode != 0:
            print(f"[bench] reference engine failed: {r.stderr.strip()[-300:]}", file=sys.stderr, flush=True)
            return None
        runs = [l.split() for l in r.stdout.splitlines() if l.startswith("run ")]
        secs = [float(x[2]) for x in runs][1:]
        walls = [float(l.split("wall=")[1].split("ms")[0]) * 1e-3 for l in r.stderr.splitlines()
                 if "RunParallelRGs wall=" in l][1:]
        threads = [int(l.split(" nt=")[1].split()[0]) for l in r.stderr.splitlines() if " nt=" in l]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    med = statistics.median(secs)
    return {"rows": n, "median_s": med, "runs_s": secs, "rows_per_s": n / med,
            "scan_aggregate_s": statistics.median(walls) if walls else None,
            "threads": threads[0] if threads else None, "result_rows": int(runs[-1][3]),
            "sample": f"first {n} rows of the same workload as Parquet (uncompressed, 122880-row row groups); "
                      f"the reference engine (oracle/_ref/ref_query, built from its sources) runs the query "
                      f"through its C API: fused Parquet scan, per-thread aggregation, merge, result "
                      f"materialized; median of {reps} runs after a warm-up run"}


def ref_calibration(threads):
    """The reference engine against the CPU restatement (orc_port_groupby) on
    the three aggregate configurations, median of 5 runs each. The reference
    reads the rows from Parquet and materializes its result; the restatement
    reads them from memory. Prints one JSON line."""
    import statistics
    from oracle import oracle as O
    from slothdb_amd import datagen, gbagg as G
    out = {"host_cores": len(os.sched_getaffinity(0)), "port_threads": threads}
    for config, (q, cols), n in (("plain", datagen.plain_workload(n=1_000_000), 1_000_000),
                                 ("highcard", datagen.highcard_workload(n=20_000_000), 20_000_000),
                                 ("composite", datagen.composite_workload(n=5_000_000), 5_000_000)):
        ref = reference_time(config, q, cols, n, reps=5)
        pm = None
        if config != "composite":  # the restatement covers one integer group column
            sample = [G.Column(c.type, c.values[:n]) for c in cols]
            port = []
            for _ in range(5):
                t = time.perf_counter()
                O.port_groupby(q, sample, threads)
                port.append(time.perf_counter() - t)
            pm = statistics.median(port)
        out[config] = {"rows": n, "reference": ref, "port_median_s": pm, "port_rows_per_s": n / pm if pm else None,
                       "port_over_reference": (pm / ref["median_s"]) if ref and pm else None}
        print(f"[bench] {config}: reference {ref['median_s'] if ref else float('nan'):.3f}s, "
              f"port {pm if pm else float('nan'):.3f}s", file=sys.stderr, flush=True)
    print(json.dumps({"ref_calibration": out}), flush=True)


def batch_split(nq=512):
    """The batched mix, and its queries split by distinct keys, each as a
    batch of its own (autotuned, best of 5): where the batched time goes."""
    from slothdb_amd import datagen, gbagg as G
    specs = datagen.batch_specs(nq)
    host = datagen.batch_workload(nq=nq)
    sync = G.lib().gbagg_device_synchronize
    out = []
    for name, sel in [("all", lambda d: True), ("distinct<=10000", lambda d: d <= 10_000),
                      ("distinct>10000", lambda d: d > 10_000), ("distinct>100000", lambda d: d > 100_000)]:
        dev = [(q, [G.DeviceColumn(c) for c in cols]) for (q, cols), sp in zip(host, specs) if sel(sp[1])]
        rows = sum(c[0].n for _, c in dev)
        b = G.GpuBatch()
        b.set_queries(dev)
        b.autotune()
        best = 1e9
        for _ in range(5):
            sync()
            t = time.perf_counter()
            b.run()
            sync()
            best = min(best, time.perf_counter() - t)
        out.append({"queries": name, "count": len(dev), "rows": rows, "batched_ms": best * 1e3,
                    "rows_per_s": rows / best, "tune": b.info()["tune"]})
        b.close()
        for _, cs in dev:
            for c in cs:
                c.free()
    print(json.dumps({"batch_split": out}), flush=True)


def bench_batched(nq, plain=True, tune=None):
    """The mixed-query workload: nq COUNT/SUM GROUP BY queries (rows
    log-uniform 1e3..4e6, distinct log-uniform 1..1e6, Zipf s in [0, 1.2]),
    HBM-resident. Plain = the single-query operator once per query; batched =
    the captured graph after autotuning (tuning time excluded, reported)."""
    from slothdb_amd import datagen, gbagg as G
    host = datagen.batch_workload(nq=nq)
    rows = sum(len(c[0]) for _, c in host)
    dev = [(q, [G.DeviceColumn(c) for c in cols]) for q, cols in host]
    sync = G.lib().gbagg_device_synchronize
    ops = [G.GpuHashAggregate(q) for q, _ in dev] if plain else []
    plain_s = []
    for rep in range(3 if plain else 0):
        sync()
        t = time.perf_counter()
        for op, (_, dcols) in zip(ops, dev):
            op.Init()
            op.consume_device(dcols)
            op.finalize()
        sync()
        plain_s.append(time.perf_counter() - t)
    b = G.GpuBatch()
    b.set_queries(dev)
    t = time.perf_counter()
    if tune:
        b.set_tune(*tune)
    else:
        b.autotune()
    tune_s = time.perf_counter() - t
    walls, devs = [], []
    for rep in range(5):
        sync()
        t = time.perf_counter()
        b.run()
        sync()
        walls.append(time.perf_counter() - t)
        devs.append(b.info()["ms_device"])
    info = b.info()
    # exactness spot check against the single-query operator
    mism = 0
    for i in range(0, nq, max(1, nq // 16)) if plain else ():
        mism += sorted(b.result(i).rows()) != sorted(ops[i].result().rows())
    # the same tuned batch with the selector off (modes from the sketch rule alone)
    b.set_param("selector", 0)
    b.set_tune(**info["tune"])
    walls0 = []
    for rep in range(5):
        sync()
        t = time.perf_counter()
        b.run()
        sync()
        walls0.append(time.perf_counter() - t)
    for op in ops:
        op.close()
    b.close()
    plain_ms = min(plain_s) * 1e3 if plain_s else None
    batch_ms = min(walls) * 1e3
    in_bytes = rows * 12
    print(f"[bench] batched: {nq} queries, {rows} rows: plain {plain_ms or float('nan'):.1f} ms, batched {batch_ms:.2f} ms "
          f"(device {min(devs):.2f} ms; selector planned {info['selector_queries']} queries; without it "
          f"{min(walls0) * 1e3:.2f} ms), tune {tune_s:.1f}s, mismatches {mism}", file=sys.stderr, flush=True)
    return {"workload": f"{nq} mixed COUNT/SUM queries", "rows": rows, "plain_ms": plain_ms,
            "batched_ms": batch_ms, "batched_device_ms": min(devs), "speedup": plain_ms / batch_ms if plain_ms else None,
            "batched_rows_per_s": rows / (batch_ms * 1e-3), "batched_hbm_frac": in_bytes / (batch_ms * 1e-3) / 8e12,
            "selector_queries": info["selector_queries"], "batched_ms_without_selector": min(walls0) * 1e3,
            "tune": info["tune"], "autotune_s": tune_s, "spot_check_mismatches": mism}


def copy_calibration():
    """1 GiB device-to-device copies through the HIP runtime (known bytes)."""
    import ctypes as C
    from slothdb_amd import gbagg as G
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    G._check(G.lib().gbagg_device_alloc(n, C.byref(a)))
    G._check(G.lib().gbagg_device_alloc(n, C.byref(b)))
    hip = C.CDLL("libamdhip64.so")
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    G._check(G.lib().gbagg_device_free(a))
    G._check(G.lib().gbagg_device_free(b))
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


if __name__ == "__main__":
    main()
