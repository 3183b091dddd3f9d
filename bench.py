"""Benchmark: hash join + group-by on one MI355X (the headline workload), with
the as-of join (binary search vs learned index), the 1M-row group-by, the
trade-table filter, the full tick pipeline and the batched group-by reported
beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--quick]

Headline workload: build 2^24 rows (unique I64 keys = seeded permutation, I32
group column in [0,10000)); probe 2^28 rows (Zipf(1.0) keys over the build
domain, 10% misses, F64 payload); SUM(payload), COUNT(*) per group. Inputs
are generated in HBM before the timed region. One step = the whole query
(build index + pre-grouping, streaming probe with per-block LDS partials,
merge, emission of the ~10^4 result rows). value = probe rows per second.

The dominant kernel (the streaming probe) is timed with HIP events on the
stream it runs on; its algorithmic bytes are 16 per probe row (8-byte key +
8-byte payload), against MI355X's 8 TB/s HBM3E peak.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def load_pkg():
    import importlib.util
    name = "query_engine_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, ROOT / "query-engine_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_tuning(args):
    """The persisted autotuner choices: {workload signature: entry}."""
    try:
        return json.loads(Path(args.tuning_file).read_text())
    except (OSError, ValueError):
        return {}


def save_tuning(args, key, entry):
    data = load_tuning(args)
    data[key] = entry
    path = Path(args.tuning_file)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


def tuning_for(args, key, tune_fn):
    """The entry for `key`: from the tuning file when present (unless
    --retune), else from the autotuner, which is then written to the file."""
    entry = None if args.retune else load_tuning(args).get(key)
    if entry is not None:
        return dict(entry, source="file", file=str(args.tuning_file))
    t0 = time.perf_counter()
    entry = tune_fn()
    entry["autotune_s"] = time.perf_counter() - t0
    save_tuning(args, key, entry)
    return dict(entry, source="autotune", file=str(args.tuning_file))


def join_groupby_workload(qe, ctx, args):
    build_n = 1 << args.build_log2
    probe_n = 1 << args.probe_log2
    groups = 10000
    log(f"generating join workload: build {build_n} probe {probe_n}")
    (pk, pp), (bk, bg) = ctx.gen_join(20260615, build_n, probe_n, groups, args.zipf, args.miss)
    aggs = [(qe.SUM, 1), (qe.COUNT_STAR, 0)]
    # Launch shape: the persisted autotuner choice for this workload on this
    # GPU, or a fresh autotune (timed on this query) written back to the file.
    key = (f"join_groupby/{ctx.device_info()['arch'].split(':')[0]}/build2^{args.build_log2}/"
           f"probe2^{args.probe_log2}/zipf{args.zipf:g}/miss{args.miss:g}/groups{groups}/sum,count_star")
    tuning = tuning_for(args, key, lambda: ctx.join_autotune([pk, pp], [bk, bg], [0], [0], qe.INNER, [3], aggs))
    pairs = args.join_pairs or tuning["pairs"]
    ctx.set_tuning(qe.TUNE_JOIN_PAIRS, pairs)
    ctx.set_tuning(qe.TUNE_JOIN_BLOCKS_PER_CU, tuning["blocks_per_cu"])
    tuning["applied"] = {"pairs": pairs, "blocks_per_cu": tuning["blocks_per_cu"]}

    # The C ABI call a host engine makes per query; its arguments are
    # marshalled once, so the step times the engine, not Python list handling.
    run = ctx.prepare_join_aggregate([pk, pp], [bk, bg], [0], [0], qe.INNER, [3], aggs)

    def step():
        t = run()
        rows = t.rows
        t.free()
        return rows

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.set_timing(True)
    ctx.timing("join_agg_probe")  # clear
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nrows = step()
    ctx.sync()
    t1 = time.perf_counter()
    kms, launches = ctx.timing("join_agg_probe")
    ctx.set_timing(False)
    # The event timing above sits inside the timed loop (it synchronises per
    # launch); re-time the steps without events for the headline figure.
    ctx.sync()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    t3 = time.perf_counter()
    ms_step = (t3 - t2) * 1e3 / args.steps
    # Step-time distribution over separate repetitions (each step ends with the
    # result on the host, so each is timed on its own).
    reps = []
    for _ in range(max(20, args.steps)):
        ta = time.perf_counter()
        step()
        reps.append((time.perf_counter() - ta) * 1e3)
    kernel_ms = kms / max(launches, 1)
    # The floor of the probe: its map lookups alone (the probe keys streamed,
    # one 16-bit entry gathered per key from a map of the build's key range),
    # best of 6.
    gmap, sink = qe.DevBuffer(ctx, 2 * build_n), qe.DevBuffer(ctx, 1024)
    gather_ms = min(ctx.map_gather(pk, 0, build_n, gmap, sink) for _ in range(6))
    gmap.free()
    sink.free()
    probe_bytes = probe_n * 16
    step_bytes = probe_n * 16 + build_n * 12 + groups * 20
    res = {
        "ms_per_step": ms_step,
        "ms_per_step_with_events": (t1 - t0) * 1e3 / args.steps,
        "step_ms_p50": float(np.percentile(reps, 50)), "step_ms_p99": float(np.percentile(reps, 99)),
        "step_reps": len(reps),
        "probe_rows": probe_n, "build_rows": build_n, "groups_out": int(nrows),
        "kernel": "join_agg_stream (direct-map fused probe)",
        "kernel_ms": kernel_ms,
        "kernel_launches": launches,
        "probe_bytes": probe_bytes, "step_bytes": step_bytes,
        "value": probe_n / (ms_step / 1e3),
        "step_achieved_gbs": step_bytes / (ms_step / 1e3) / 1e9,
        "kernel_achieved_gbs": probe_bytes / (kernel_ms / 1e3) / 1e9 if kernel_ms > 0 else None,
        "map_gather_floor_ms": gather_ms,
        "kernel_over_gather_floor": kernel_ms / gather_ms if gather_ms > 0 else None,
        "tuning": tuning,
    }
    cols = {"pk": pk, "pp": pp, "bk": bk, "bg": bg}
    return res, cols


def cpu_baseline_join(qe, ctx, cols, args):
    """Morsel-parallel CPU restatement timed on a bounded sample: the full
    2^24-row build and the first 2^24 probe rows (1/16 of the probe)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    L = O.oracle()
    sample = min(1 << args.cpu_probe_log2, cols["pk"].len)
    log(f"cpu baseline: downloading sample ({sample} probe rows)")
    bk, _ = cols["bk"].numpy()
    bg, _ = cols["bg"].numpy()
    pk_s = ctx.slice(cols["pk"], 0, sample)
    pp_s = ctx.slice(cols["pp"], 0, sample)
    pkn, _ = pk_s.numpy()
    ppn, _ = pp_s.numpy()
    threads = int(L.orc_bl_hw_threads())
    maxg = 1 << 16
    og = np.empty(maxg, np.int32)
    osum = np.empty(maxg, np.float64)
    ocnt = np.empty(maxg, np.int64)
    t0 = time.perf_counter()
    g = L.orc_bl_join_groupby(pkn.ctypes.data, ppn.ctypes.data, sample, bk.ctypes.data, bg.ctypes.data,
                              len(bk), threads, 1024, maxg, og.ctypes.data, osum.ctypes.data, ocnt.ctypes.data)
    t1 = time.perf_counter()
    # Parity spot check: the GPU on the same sample.
    t = ctx.join_aggregate([pk_s, pp_s], [cols["bk"], cols["bg"]], [0], [0], qe.INNER, [3],
                           [(qe.SUM, 1), (qe.COUNT_STAR, 0)])
    gv = t.numpy()
    t.free()
    pk_s.free()
    pp_s.free()
    # Parity on the sample: the oracle's join -> aggregate with Neumaier sums;
    # keys and COUNTs exact, SUM within 1e-12 relative.
    from _compare import diff, table_rows
    t4 = time.perf_counter()
    want = O.join_aggregate([(pkn, None, O.I64), (ppn, None, O.F64)], [(bk, None, O.I64), (bg, None, O.I32)],
                            [0], [0], 0, [3], [(O.SUM, 1), (O.COUNT_STAR, 0)], sum_mode=1)
    t5 = time.perf_counter()
    mismatch = diff(table_rows(gv), want, 1e-12, 1e-12, False)
    cpu = {int(og[i]): (float(osum[i]), int(ocnt[i])) for i in range(g)}
    port_vs_oracle = all(cpu.get(r[0], (None, -1))[1] == r[2] for r in want) and len(cpu) == len(want)
    res = {
        "value": sample / (t1 - t0), "unit": "probe rows/s", "cores": threads, "kind": "port",
        "sample": f"full 2^{args.build_log2}-row build + first {sample} probe rows; build single-threaded "
                  f"then morsel-parallel probe (1024-row morsels) with per-thread partial aggregates "
                  f"(oracle/qe_baseline.cpp)",
        "seconds": t1 - t0, "gpu_parity_on_sample": mismatch is None,
        "parity_check": "GPU vs oracle (single-thread join -> aggregate, Neumaier sums) on the sample: "
                        "keys and COUNTs exact, SUM within 1e-12 rel",
        "parity_mismatch": mismatch, "oracle_single_thread_s": t5 - t4,
        "port_counts_equal_oracle": bool(port_vs_oracle),
    }
    # The reference's own operators (compiled from its sources into oracle/_ref
    # where this tree was built), single-threaded: Aggregate(HashJoin(Scan, Scan)).
    if O.ref_available():
        rs = min(1 << args.cpu_ref_log2, sample)
        pc = [(pkn[:rs], None, O.I64), (ppn[:rs], None, O.F64)]
        bc = [(bk, None, O.I64), (bg, None, O.I32)]
        t2 = time.perf_counter()
        rows = O.ref_join_aggregate(pc, bc, [0], [0], 0, [3], [(O.SUM, 1), (O.COUNT_STAR, 0)])
        t3 = time.perf_counter()
        res["reference_single_thread"] = {
            "value": rs / (t3 - t2), "unit": "probe rows/s", "cores": 1, "kind": "reference",
            "sample": f"full 2^{args.build_log2}-row build + first {rs} probe rows through the reference's "
                      f"HashJoin and Aggregate (2048-row batches), table construction included",
            "seconds": t3 - t2, "groups": len(rows)}
    return res


def asof_workload(qe, ctx, args):
    """Right 400*2^20 rows (sym Zipf(1.1) over 5000, bursty per-sym microsecond
    times, price), left 100*2^20 probes; device sort + learned-index training
    (build phase, reported separately); probe phase timed per method."""
    rn, ln = args.asof_right, args.asof_left
    log(f"generating as-of workload: right {rn} left {ln}")
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(20260616, rn, ln, 5000, 1.1, sessions=args.asof_sessions)
    t0 = time.perf_counter()
    idx = ctx.asof_build([rs], rt)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, args.bucket_knots)
    ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, args.data_lines)
    idx.train(args.piece_rows, args.train_steps, args.train_lr)
    ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, 1)
    ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, 1)
    ctx.sync()
    t2 = time.perf_counter()
    out = qe.DevBuffer(ctx, 4 * ln)
    res = {"right_rows": rn, "left_rows": ln, "build_sort_s": t1 - t0, "train_s": t2 - t1,
           "piece_rows": args.piece_rows, "train_steps": args.train_steps, "bucket_knots": args.bucket_knots,
           "data_lines": args.data_lines}
    results = {}
    srt = qe.ASOF_OUT_SORTED
    for name, mode, timer in (("binary", qe.ASOF_BINARY, "asof_probe_binary"),
                              ("learned", qe.ASOF_LEARNED, "asof_probe_learned"),
                              ("learned_no_mlp", qe.ASOF_LEARNED_NO_MLP, "asof_probe_learned_no_mlp"),
                              ("binary_sorted", qe.ASOF_BINARY | srt, "asof_probe_binary"),
                              ("learned_sorted", qe.ASOF_LEARNED | srt, "asof_probe_learned"),
                              ("learned_no_mlp_sorted", qe.ASOF_LEARNED_NO_MLP | srt, "asof_probe_learned_no_mlp")):
        for _ in range(2):
            idx.probe_into([ls], lt, out, mode=mode)
        ctx.set_timing(True)
        ctx.timing(timer)
        for _ in range(args.asof_reps):
            idx.probe_into([ls], lt, out, mode=mode)
        ms, n = ctx.timing(timer)
        ctx.set_timing(False)
        results[name] = out.numpy(np.uint32, ln)
        res[f"{name}_ms"] = ms / max(n, 1)
        res[f"{name}_probes_per_s"] = ln / (ms / max(n, 1) / 1e3)
    st = idx.stats()
    res["learned_equals_binary"] = bool((results["binary"] == results["learned"]).all()
                                        and (results["binary"] == results["learned_no_mlp"]).all()
                                        and (results["binary_sorted"] == results["learned_sorted"]).all()
                                        and (results["binary_sorted"] == results["learned_no_mlp_sorted"]).all())
    res["speedup_learned_vs_binary"] = res["binary_ms"] / res["learned_ms"]
    res["speedup_mlp_vs_no_mlp"] = res["learned_no_mlp_ms"] / res["learned_ms"]
    res["speedup_learned_vs_binary_sorted_output"] = res["binary_sorted_ms"] / res["learned_sorted_ms"]
    res["speedup_mlp_vs_no_mlp_sorted_output"] = res["learned_no_mlp_sorted_ms"] / res["learned_sorted_ms"]
    res["matched_frac"] = float((results["binary"] != qe.NO_ROW).mean())
    probe_bytes = ln * 16  # left sym 4 + ts 8 + match index 4
    res["learned_achieved_gbs"] = probe_bytes / (res["learned_ms"] / 1e3) / 1e9
    res["binary_achieved_gbs"] = probe_bytes / (res["binary_ms"] / 1e3) / 1e9
    res["learned_sorted_achieved_gbs"] = probe_bytes / (res["learned_sorted_ms"] / 1e3) / 1e9
    res["stats"] = {k: (float(v) if isinstance(v, float) else int(v)) for k, v in st.items()}
    if not args.skip_cpu:
        res["cpu_baseline"] = cpu_baseline_asof(qe, ctx, idx, ls, lt, results["binary"], st, args)
    out.free()
    idx.free()
    for c in (rs, rt, rp, ls, lt):
        c.free()
    return res


def cpu_baseline_asof(qe, ctx, idx, ls, lt, gpu_match, st, args):
    """Parallel CPU upper_bound probe over the same sorted right side (the
    reference has no parallel as-of); bounded probe sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    L = O.oracle()
    n = int(st["build_rows"])
    parts = int(st["partitions"])
    ts, rows, offs = idx.sorted(n, parts)
    sample = min(1 << args.cpu_asof_log2, ls.len)
    lsn, _ = ls.numpy()
    ltn, _ = lt.numpy()
    lsn, ltn = np.ascontiguousarray(lsn[:sample]), np.ascontiguousarray(ltn[:sample])
    out = np.empty(sample, np.uint32)
    threads = int(L.orc_bl_hw_threads())
    t0 = time.perf_counter()
    L.orc_bl_asof_upper_bound(ts.ctypes.data, offs.ctypes.data, parts, lsn.ctypes.data, ltn.ctypes.data, sample,
                              threads, out.ctypes.data)
    t1 = time.perf_counter()
    mapped = np.where(out == qe.NO_ROW, qe.NO_ROW, rows[np.minimum(out, max(n - 1, 0))])
    res = {"value": sample / (t1 - t0), "unit": "probe rows/s", "cores": threads, "kind": "port",
           "sample": f"first {sample} probes against the full sorted right side; std::upper_bound per probe "
                     f"(oracle/qe_baseline.cpp)",
           "gpu_parity_on_sample": bool((mapped == gpu_match[:sample]).all())}
    res["single_thread_asof_join"] = cpu_asof_single_thread(qe, ctx, O, args)
    return res


def cpu_asof_single_thread(qe, ctx, O, args):
    """The whole as-of join single-threaded on a sample of the same generator
    (right 2^22 rows, left 2^20, 5000 Zipf(1.1) symbols): the oracle's
    restatement and, where this tree was built with it, the reference's
    AsofJoin (sort of both sides + cursor merge, tsx/asof.cpp:234-308, 2048-row
    batches), each timed over the operator run alone. The GPU binary probe on
    the same sample is checked against the reference's pairs."""
    rn, ln = (1 << 22, 1 << 20) if not args.quick else (1 << 18, 1 << 16)
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(20260616, rn, ln, 5000, 1.1)
    idx = ctx.asof_build([rs], rt)
    gpu = idx.probe([ls], lt, mode=qe.ASOF_BINARY)
    idx.free()
    cols = [c.numpy()[0] for c in (rs, rt, rp, ls, lt)]
    for c in (rs, rt, rp, ls, lt):
        c.free()
    bc = [(cols[0], None, O.I32), (cols[1], None, O.I64), (cols[2], None, O.F64)]
    pc = [(cols[3], None, O.I32), (cols[4], None, O.I64)]
    t0 = time.perf_counter()
    p, b = O.asof_pairs(pc, bc, [0], [0], 1, 1, join_type=0)
    t1 = time.perf_counter()
    want = np.full(ln, qe.NO_ROW, dtype=np.uint32)
    want[p] = b
    res = {"sample": f"gen_asof(seed 20260616) at right {rn} / left {ln} rows, inner as-of on sym, ts <=",
           "cores": 1, "unit": "left rows/s", "port_value": ln / (t1 - t0),
           "gpu_equals_port": bool((want == gpu).all())}
    if O.ref_available():
        R = O.ref()
        bcs, pcs = O.Cols(bc), O.Cols(pc)
        t2 = time.perf_counter()
        h = R.ref_asof_run(pcs.arr, 2, pcs.n, bcs.arr, 3, bcs.n, O._u32([0]), O._u32([0]), 1, 1, 1, 0, 0, 0, 2048)
        t3 = time.perf_counter()
        rows = O._ref_rows(h)  # frees h; each row is probe cols, probe row id, build cols, build row id
        rwant = np.full(ln, qe.NO_ROW, dtype=np.uint32)
        rwant[np.array([r[2] for r in rows], dtype=np.int64)] = [r[6] for r in rows]
        res.update({"value": ln / (t3 - t2), "kind": "reference", "reference_seconds": t3 - t2,
                    "gpu_equals_reference": bool((rwant == gpu).all()),
                    "restatement_vs_reference": (t1 - t0) / (t3 - t2)})
    return res


def batched_workload(qe, ctx, args):
    """64 independent GROUP BY queries over 4096 morsels (64 per query), morsel
    rows log-uniform [2^11, 2^21], groups per query log-uniform [16, 10^7]
    (capped by rows), I32/I64 keys, F64/I64 values with 2% NULLs;
    COUNT(*), SUM, AVG per group."""
    r = np.random.default_rng(20260617)
    nq, per = args.batch_queries, args.batch_morsels // args.batch_queries
    lo, hi = np.log(2 ** 11), np.log(2 ** args.batch_max_log2)
    queries, total_rows = [], 0
    log(f"generating batched workload: {nq} queries x {per} morsels")
    for q in range(nq):
        sizes = np.exp(r.uniform(lo, hi, per)).astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        n = int(offs[-1])
        groups = int(min(np.exp(r.uniform(np.log(16), np.log(1e7))), n))
        kt = qe.I32 if r.random() < 0.5 else qe.I64
        vt = qe.F64 if r.random() < 0.5 else qe.I64
        k = ctx.gen_uniform(20260617, 2 * q, kt, n, 0, groups)
        v = ctx.gen_uniform(20260617, 2 * q + 1, vt, n, 0, 1000, 0.02)
        queries.append((k, v, offs, [(qe.COUNT_STAR, 0), (qe.SUM, 1), (qe.AVG, 1)], groups))
        total_rows += n
    plan = ctx.batch(queries)
    res = {"queries": nq, "morsels": nq * per, "rows": total_rows}
    strat = {}
    for (name, _), (_k, _v, offs, *_rest) in zip(plan.strategies(), queries):
        q_, r_ = strat.get(name, (0, 0))
        strat[name] = (q_ + 1, r_ + int(offs[-1]))
    res["strategies"] = {name: {"queries": q_, "rows": r_} for name, (q_, r_) in sorted(strat.items())}
    key = (f"batched_groupby/{ctx.device_info()['arch'].split(':')[0]}/queries{nq}/morsels{nq * per}/"
           f"max2^{args.batch_max_log2}")

    def autotune():
        t = plan.autotune()
        return {"tile_rows": t[0], "blocks_per_cu": t[1], "lds_slots": t[2]}

    res["tuning"] = tuning_for(args, key, autotune)
    tune = (res["tuning"]["tile_rows"], res["tuning"]["blocks_per_cu"], res["tuning"]["lds_slots"])

    def timed(mode, reps):
        plan.run(mode, tuning=tune, collect=False)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            plan.run(mode, tuning=tune, collect=False)
        ctx.sync()
        return (time.perf_counter() - t) / reps

    res["persistent_s"] = timed(qe.BATCH_PERSISTENT, 3)
    res["graph_s"] = timed(qe.BATCH_GRAPH, 3)
    if not args.no_per_batch:
        res["per_batch_s"] = timed(qe.BATCH_PER_BATCH, 1)
        res["speedup_persistent_vs_per_batch"] = res["per_batch_s"] / res["persistent_s"]
        res["speedup_graph_vs_per_batch"] = res["per_batch_s"] / res["graph_s"]
    res["rows_per_s_graph"] = total_rows / res["graph_s"]
    # Parity between the modes: every query's groups, keys and aggregates
    # (COUNT, SUM, AVG) identical in bits once rows are ordered by key.
    outs = {m: plan.run(m, tuning=tune) for m in (qe.BATCH_PERSISTENT, qe.BATCH_GRAPH)}
    same = True
    for a, b in zip(outs[qe.BATCH_PERSISTENT], outs[qe.BATCH_GRAPH]):
        ca, cb = a.numpy(), b.numpy()
        same &= a.rows == b.rows and len(ca) == len(cb)
        if same:
            oa, ob = np.argsort(ca[0][0], kind="stable"), np.argsort(cb[0][0], kind="stable")
            for (va, ma), (vb, mb) in zip(ca, cb):
                same &= np.array_equal(va[oa].view(np.uint8), vb[ob].view(np.uint8))
                same &= (ma is None and mb is None) or (ma is not None and mb is not None
                                                        and np.array_equal(ma[oa], mb[ob]))
        a.free()
        b.free()
    res["graph_equals_persistent"] = bool(same)
    res["graph_equals_persistent_check"] = "all result columns, bitwise, rows ordered by key"
    plan.free()
    for k, v, *_ in queries:
        k.free()
        v.free()
    return res


PMC_SUMMARY = ROOT / "profiles" / "join_agg_stream_pmc_build2e24_probe2e28.md"


def measured_traffic(args):
    """HBM bytes per launch of the probe kernel, from the committed rocprofv3
    PMC summary of this workload (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy kernel of known size); None for
    other workload sizes or without the summary."""
    import re
    if (args.build_log2, args.probe_log2, args.zipf, args.miss) != (24, 28, 1.0, 0.1) or not PMC_SUMMARY.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", PMC_SUMMARY.read_text())
    return float(m.group(1)) if m else None


def copy_ceiling(qe, ctx, nbytes=1 << 30, reps=5):
    """Measured HBM ceiling: a plain 16-byte streaming copy (read + write)."""
    src, dst = qe.DevBuffer(ctx, nbytes), qe.DevBuffer(ctx, nbytes)
    best = 1e30
    for _ in range(reps + 1):
        best = min(best, ctx.stream_copy(dst, src, nbytes))
    src.free()
    dst.free()
    return {"bytes_moved": 2 * nbytes, "best_ms": best, "gbs": 2 * nbytes / (best / 1e3) / 1e9}


def groupby_1m(qe, ctx, args):
    """Plain group-by: 2^20 rows, I64 key uniform in [0,1024), F64 value in
    [0,1000) with 5% NULLs; COUNT(*), SUM, AVG per key. The GPU query is timed
    end to end (result on the host); beside it the CPU paths on the same rows."""
    n = 1 << 20
    k, v = ctx.gen_groupby(20260615, n, 1024, 0.05)
    aggs = [(qe.COUNT_STAR, 0), (qe.SUM, 1), (qe.AVG, 1)]
    for _ in range(3):
        ctx.aggregate([k, v], [0], aggs).free()
    ctx.sync()
    t = time.perf_counter()
    reps = 20
    for _ in range(reps):
        ctx.aggregate([k, v], [0], aggs).free()
    ctx.sync()
    s = (time.perf_counter() - t) / reps
    res = {"rows": n, "ms_per_query": s * 1e3, "rows_per_s": n / s}
    if not args.skip_cpu:
        tab = ctx.aggregate([k, v], [0], aggs)
        gv = tab.numpy()
        tab.free()
        res["cpu_baseline"] = cpu_baseline_groupby(qe, k.numpy(), v.numpy(), gv)
    k.free()
    v.free()
    return res


def _best_of(fn, reps):
    best = 1e30
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t)
    return best


def cpu_baseline_groupby(qe, kcol, vcol, gpu_cols):
    """The plain group-by on the host cores, whole workload (2^20 rows), best
    of 5: the port single-threaded and morsel-parallel (oracle/qe_baseline.cpp,
    1024-row morsels, per-thread hash maps merged at the end), the oracle's
    restatement (ops/aggregate.cpp order, Neumaier sums) and, where this tree
    was built with it, the reference's own Aggregate over 2048-row batches
    (compiled from its sources into oracle/_ref). restatement_vs_reference is
    the restatement's time over the reference's."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    from _compare import diff, table_rows
    L = O.oracle()
    key, _ = kcol
    val, valid = vcol
    n = len(key)
    words = qe.pack_validity(valid) if valid is not None else None
    wptr = words.ctypes.data if words is not None else None
    maxg = 1 << 12
    ok_, oc, osum, onn = (np.empty(maxg, np.int64), np.empty(maxg, np.int64), np.empty(maxg, np.float64),
                          np.empty(maxg, np.int64))
    threads = int(L.orc_bl_hw_threads())

    def port(th):
        return lambda: L.orc_bl_groupby_i64_f64(key.ctypes.data, val.ctypes.data, wptr, n, th, 1024, maxg,
                                                ok_.ctypes.data, oc.ctypes.data, osum.ctypes.data, onn.ctypes.data)

    cols = [(key, None, O.I64), (val, valid, O.F64)]
    oaggs = [(O.COUNT_STAR, 0), (O.SUM, 1), (O.AVG, 1)]
    res = {"unit": "rows/s", "kind": "port", "cores": threads,
           "sample": f"the whole workload ({n} rows), best of 5"}
    t1 = _best_of(port(1), 5)
    tm = _best_of(port(threads), 5)
    res["value"] = n / tm
    res["port_single_thread_rows_per_s"] = n / t1
    res["port_morsel_rows_per_s"] = n / tm
    want = []
    to = _best_of(lambda: want.append(O.aggregate(cols, [0], oaggs, sum_mode=1)), 5)
    res["oracle_restatement_single_thread_rows_per_s"] = n / to
    mismatch = diff(table_rows(gpu_cols), want[-1], 1e-12, 1e-12, False)
    res["gpu_parity"] = mismatch is None
    res["parity_mismatch"] = mismatch
    if O.ref_available():
        tr = _best_of(lambda: O.ref_aggregate(cols, [0], oaggs), 5)
        res["reference_single_thread"] = {"value": n / tr, "unit": "rows/s", "cores": 1, "kind": "reference",
                                          "sample": f"{n} rows through the reference's Aggregate (2048-row "
                                                    f"batches), best of 5"}
        res["restatement_vs_reference"] = to / tr
    return res


def filter_workload(qe, ctx, args):
    """Trade-table filter: 2^30 rows (sym I32 in [0,5000), ts I64, px F64
    uniform [0,200) with 2% NULLs), WHERE px > 100, all three columns
    compacted. One step = qe_gpu_filter (predicate bitmap, tile scan, fused
    compaction, result table)."""
    n = 1 << args.filter_log2
    log(f"generating filter workload: {n} rows")
    sym = ctx.gen_uniform(20260617, 0, qe.I32, n, 0, 5000)
    ts = ctx.gen_uniform(20260617, 1, qe.I64, n, 0, 1 << 40)
    px = ctx.gen_uniform(20260617, 2, qe.F64, n, 0.0, 200.0, null_frac=args.filter_null_frac)
    cols = [sym, ts, px]
    terms = [(2, qe.CMP_GT, qe.F64, 100.0)]
    for _ in range(args.warmup):
        ctx.filter(cols, terms).free()
    ctx.sync()
    t0 = time.perf_counter()
    m = 0
    for _ in range(args.steps):
        t = ctx.filter(cols, terms)
        m = t.rows
        t.free()
    ctx.sync()
    wall = (time.perf_counter() - t0) / args.steps
    # Per-kernel times in a separate pass (the event brackets synchronize).
    ctx.set_timing(True)
    ctx.timing("filter_pred_mask")
    ctx.timing("filter_compact")
    for _ in range(args.steps):
        ctx.filter(cols, terms).free()
    mask_ms, k1 = ctx.timing("filter_pred_mask")
    comp_ms, k2 = ctx.timing("filter_compact")
    ctx.set_timing(False)
    mask_ms /= max(k1, 1)
    comp_ms /= max(k2, 1)
    row_bytes = 4 + 8 + 8
    # Algorithmic bytes: the bitmap pass reads px (8 B/row) and its validity
    # (1/8 B/row) and writes 1/8 B/row; the compaction reads every column and the
    # bitmap once and writes the m surviving rows.
    mask_bytes = n * 8 + n / 8 + n / 8
    comp_bytes = n * row_bytes + n / 8 + n / 8 + m * row_bytes + m / 8
    res = {"rows": n, "passing": m, "selectivity": m / n, "px_null_frac": args.filter_null_frac, "ms_per_step": wall * 1e3,
           "rows_per_s": n / wall, "pred_mask_ms": mask_ms, "compact_ms": comp_ms,
           "pred_mask_gbs": mask_bytes / (mask_ms / 1e3) / 1e9,
           "compact_gbs": comp_bytes / (comp_ms / 1e3) / 1e9,
           "step_gbs": (n * row_bytes + n / 8 + m * row_bytes) / wall / 1e9}
    if not args.skip_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import _oracle as O
        sample = min(1 << 24, n)
        head = ctx.slice(px, 0, sample)
        pv, pvalid = head.numpy()
        c = [(pv, pvalid, O.F64)]
        t0 = time.perf_counter()
        sel = O.filter_select(c, [(0, O.GT, O.F64, 100.0)])
        t1 = time.perf_counter()
        gsel = ctx.filter_select([head], [(0, qe.CMP_GT, qe.F64, 100.0)])
        head.free()
        res["cpu_baseline"] = {"value": sample / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
                               "sample": f"first {sample} rows of px; oracle filter_select (selection only)",
                               "gpu_parity_on_sample": bool(np.array_equal(sel, gsel))}
    for c in cols:
        c.free()
    return res


def tick_cpu_baseline(qe, ctx, args, d, sym_id, sym_name, names, order, code):
    """The whole pipeline on a bounded sample (2^23 trades against 2^25 quotes
    from the same generator) through the CPU oracle (filter, as-of, join,
    group-by; single thread), timed, and compared with the GPU pipeline on the
    same sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    nq, nt, nsym = 1 << 25, 1 << 23, len(names)
    (qs, qt, qp), (ts_, tt) = ctx.gen_asof(20260619, nq, nt, nsym, 1.1)
    tp = ctx.gen_uniform(20260619, 5, qe.F64, nt, 0.0, 200.0)
    # GPU pipeline on the sample (binary-search as-of: an index per run)
    f = ctx.filter([ts_, tt, tp], [(2, qe.CMP_GT, qe.F64, 100.0)])
    idx = ctx.asof_build([qs], qt)
    match = qe.DevBuffer(ctx, 4 * max(f.rows, 1))
    idx.probe_into([f.col(0)], f.col(1), match, mode=qe.ASOF_BINARY)
    pairs = ctx.asof_pairs(match, f.rows, qe.INNER)
    g = pairs.gather([f.col(0), f.col(2)], [])
    agg = ctx.join_aggregate([g.col(0), g.col(1)], [sym_id, sym_name], [0], [0], qe.INNER, [3],
                             [(qe.COUNT_STAR, 0), (qe.AVG, 1)])
    gpu_rows = list(zip(agg.col(0).strings(), agg.col(1).numpy()[0].tolist(), agg.col(2).numpy()[0].tolist()))
    for x in (agg, g, pairs, match, idx, f):
        x.free()
    host = [c.numpy()[0] for c in (qs, qt, ts_, tt, tp)]
    for c in (qs, qt, qp, ts_, tt, tp):
        c.free()
    qsym, qts, tsym, tts, tpx = host
    sym_tab = [(np.arange(nsym, dtype=np.int32), None, O.I32),
               (np.array([code[names[i]] for i in order], dtype=np.int32), None, O.STR, d.strings)]
    t0 = time.perf_counter()
    keep = O.filter_select([(tpx, None, O.F64)], [(0, O.GT, O.F64, 100.0)])
    fs, ft, fp = tsym[keep], tts[keep], tpx[keep]
    p_idx, _ = O.asof_pairs([(fs, None, O.I32), (ft, None, O.I64)], [(qsym, None, O.I32), (qts, None, O.I64)],
                            [0], [0], 1, 1, 0)
    cpu_rows = O.join_aggregate([(fs[p_idx], None, O.I32), (fp[p_idx], None, O.F64)], sym_tab, [0], [0], 0, [3],
                                [(O.COUNT_STAR, 0), (O.AVG, 1)])
    t1 = time.perf_counter()
    from _compare import diff
    return {"value": nt / (t1 - t0), "unit": "trade rows/s", "cores": 1, "kind": "port",
            "sample": f"{nt} trades against {nq} quotes (same generator, seed 20260619), 50,000-name symbol "
                      f"table; oracle filter_select, asof (sort + per-probe search), join_aggregate",
            "seconds": t1 - t0, "groups": len(cpu_rows),
            "gpu_parity_on_sample": diff(gpu_rows, cpu_rows, rel=1e-12, abs_=1e-12) is None}


def tick_workload(qe, ctx, args):
    """The full tick pipeline: trades (2^30 rows: sym I32, ts I64 us, px F64
    uniform [0,200)) WHERE px > 100, as-of joined (backward, learned index) to
    quotes (2^31 rows: sym Zipf(1.1) over 50,000 symbols, bursty per-symbol
    microsecond times), joined on the symbol id to a 50,000-row symbol table
    whose names are dictionary-encoded VARCHAR (3-12 characters), GROUP BY
    name COUNT(*) / AVG(px), ORDER BY the count descending. The quote index
    (device sort + learned router) is built once, outside the timed steps; one
    step is the whole query over the trades."""
    nq, nt, nsym = 1 << args.tick_quotes_log2, 1 << args.tick_trades_log2, 50000
    log(f"generating tick workload: quotes {nq} trades {nt}")
    (qs, qt, qp), (ts_, tt) = ctx.gen_asof(20260618, nq, nt, nsym, 1.1)
    tp = ctx.gen_uniform(20260618, 5, qe.F64, nt, 0.0, 200.0)
    r = np.random.default_rng(20260618)
    letters = np.array(list("ABCDEFGHIJKLMNOPQRSTUVWXYZ"))
    names = set()
    while len(names) < nsym:
        k = int(r.integers(3, 13))
        names.add("".join(r.choice(letters, k)))
    names = sorted(names)
    order = r.permutation(nsym)
    d = ctx.dict([names[i] for i in r.permutation(nsym)])  # dictionary in its own order
    code = {x: i for i, x in enumerate(d.strings)}
    sym_id = ctx.upload(np.arange(nsym, dtype=np.int32), qe.I32)
    sym_name = ctx.upload(np.array([code[names[i]] for i in order], dtype=np.int32), qe.STR, dict=d)
    t0 = time.perf_counter()
    idx = ctx.asof_build([qs], qt)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, args.bucket_knots)
    ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, args.data_lines)
    idx.train(args.tick_piece_rows, args.train_steps, args.train_lr)
    ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, 1)
    ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, 1)
    ctx.sync()
    t2 = time.perf_counter()
    match = qe.DevBuffer(ctx, 4 * nt)
    perm = qe.DevBuffer(ctx, 4 * nsym)
    stages = {}

    def stamp(name, t):
        ctx.sync()
        now = time.perf_counter()
        stages[name] = stages.get(name, 0.0) + (now - t)
        return now

    def step(mode, timed_stages=False):
        t = time.perf_counter()
        f = ctx.filter([ts_, tt, tp], [(2, qe.CMP_GT, qe.F64, 100.0)])
        if timed_stages:
            t = stamp("filter", t)
        m = f.rows
        if args.tick_cluster:
            # Probes grouped by the index region they read (the as-of join's
            # probe-side ordering); the later stages are order-insensitive.
            c = idx.cluster([f.col(0)], f.col(1), [f.col(0), f.col(1), f.col(2)])
            f.free()
            f = c
            if timed_stages:
                t = stamp("asof_cluster", t)
        # Only which trades matched is used downstream (the gather takes
        # trade columns), so the probe reports sorted positions.
        idx.probe_into([f.col(0)], f.col(1), match, mode=mode | qe.ASOF_OUT_SORTED)
        pairs = ctx.asof_pairs(match, m, qe.INNER)
        if timed_stages:
            t = stamp("asof", t)
        g = pairs.gather([f.col(0), f.col(2)], [])
        if timed_stages:
            t = stamp("asof_gather", t)
        agg = ctx.join_aggregate([g.col(0), g.col(1)], [sym_id, sym_name], [0], [0], qe.INNER, [3],
                                 [(qe.COUNT_STAR, 0), (qe.AVG, 1)])
        if timed_stages:
            t = stamp("join_groupby", t)
        ng = agg.rows
        cnt = agg.col(1)
        ctx.sort_perm_into([cnt], perm, desc=[1])
        out = ctx.take([agg.col(0), agg.col(1), agg.col(2)], perm, ng)
        if timed_stages:
            t = stamp("sort", t)
        res = (m, pairs.count, ng, out)
        pairs.free()
        g.free()
        agg.free()
        f.free()
        return res

    for _ in range(max(args.warmup, 1)):
        step(qe.ASOF_LEARNED)[3].free()
    ctx.sync()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        last = step(qe.ASOF_LEARNED)
        if _ + 1 < args.steps:
            last[3].free()
    ctx.sync()
    wall = (time.perf_counter() - t3) / args.steps
    m, matched, ng, out = last
    top = [(n, c, a) for n, c, a in zip(out.col(0).strings()[:5], out.col(1).numpy()[0][:5].tolist(),
                                         out.col(2).numpy()[0][:5].tolist())]
    counts = out.col(1).numpy()[0]
    # Identical result with the binary-search as-of (the learned index is exact).
    ref = step(qe.ASOF_BINARY)
    same = (ref[0] == m and ref[1] == matched and ref[2] == ng and
            ref[3].col(0).strings() == out.col(0).strings() and
            np.array_equal(ref[3].col(1).numpy()[0], counts))
    ref[3].free()
    step(qe.ASOF_LEARNED, timed_stages=True)[3].free()
    cpu = None if args.skip_cpu else tick_cpu_baseline(qe, ctx, args, d, sym_id, sym_name, names, order, code)
    st = idx.stats()
    res = {"quotes": nq, "trades": nt, "symbols": nsym, "passing_filter": m, "asof_matched": matched,
           "groups": ng, "ms_per_step": wall * 1e3, "trade_rows_per_s": nt / wall,
           "index_build_s": t1 - t0, "index_train_s": t2 - t1, "piece_rows": args.tick_piece_rows,
           "probe_clustering": bool(args.tick_cluster),
           "router_buckets": int(st["router_buckets"]), "data_line_buckets": int(st["data_line_buckets"]),
           "stage_ms": {k: v * 1e3 for k, v in stages.items()},
           "counts_sorted_desc": bool(np.all(np.diff(counts.astype(np.int64)) <= 0)),
           "top5": top, "learned_equals_binary": bool(same), "cpu_baseline": cpu}
    out.free()
    match.free()
    perm.free()
    idx.free()
    for c in (qs, qt, qp, ts_, tt, tp, sym_id, sym_name):
        c.free()
    d.free()
    return res


def read_ceiling(qe, ctx, nbytes=1 << 30, reps=5):
    """Measured read-only HBM ceiling: 16-byte nontemporal loads, four in
    flight per thread, nothing written."""
    src, sink = qe.DevBuffer(ctx, nbytes), qe.DevBuffer(ctx, 1024)
    best = 1e30
    for _ in range(reps + 1):
        best = min(best, ctx.stream_read(src, sink, nbytes))
    src.free()
    sink.free()
    return {"bytes_read": nbytes, "best_ms": best, "gbs": nbytes / (best / 1e3) / 1e9}


def _read(path):
    try:
        return Path(path).read_text()
    except OSError:
        return ""


def result_tags(qe, ctx, args):
    """The machine state that travels with the numbers, after the reference's
    bench/bench_report.h:22-46 (CPU model, cores, RAM, governors, kernel, load,
    build provenance), plus the GPU and ROCm."""
    import glob
    import hashlib
    import platform
    cpu_model = next((ln.split(":", 1)[1].strip() for ln in _read("/proc/cpuinfo").splitlines()
                      if ln.startswith("model name")), platform.processor() or "n/a")
    mem = next((int(ln.split()[1]) * 1024 for ln in _read("/proc/meminfo").splitlines()
                if ln.startswith("MemTotal:")), 0)
    govs = sorted({_read(g).strip() for g in glob.glob("/sys/devices/system/cpu/cpu*/cpufreq/scaling_governor")})
    src = hashlib.sha256()
    for f in sorted((ROOT / "query-engine_amd" / "csrc").glob("*")) + [ROOT / "bench.py"]:
        if f.is_file():
            src.update(f.read_bytes())
    try:
        load1 = os.getloadavg()[0]
    except OSError:
        load1 = None
    return {
        "seeds": {"join_groupby": 20260615, "groupby_1m": 20260615, "asof": 20260616, "batched": 20260617,
                  "filter": 20260617, "tick_pipeline": 20260618, "tick_cpu_sample": 20260619},
        "host": {"cpu_model": cpu_model, "logical_cpus": os.cpu_count(),
                 "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
                 "ram_bytes": mem, "governors": ",".join(g for g in govs if g) or "n/a",
                 "kernel": platform.release(), "loadavg1": load1},
        "gpu": ctx.device_info(),
        "rocm": _read("/opt/rocm/.info/version").strip() or "n/a",
        "build": {"compiler": "hipcc --offload-arch=gfx950 -O3 -std=c++17, packed-fp32 ops off",
                  "source_sha16": src.hexdigest()[:16]},
        "args": {k: v for k, v in vars(args).items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-log2", type=int, default=24)
    ap.add_argument("--probe-log2", type=int, default=28)
    ap.add_argument("--cpu-probe-log2", type=int, default=24)
    ap.add_argument("--cpu-ref-log2", type=int, default=22, help="probe rows through the compiled reference")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--zipf", type=float, default=1.0, help="Zipf exponent of the probe keys")
    ap.add_argument("--miss", type=float, default=0.1, help="fraction of probe keys outside the build domain")
    ap.add_argument("--tuning-file", default=str(ROOT / "query-engine_amd" / "tuning" / "autotune.json"),
                    help="persisted autotuner choices (read, and written after a fresh autotune)")
    ap.add_argument("--retune", action="store_true", help="autotune even where the tuning file has an entry")
    ap.add_argument("--join-pairs", type=int, default=0, choices=[0, 2, 4],
                    help="row pairs per thread per tile in the streaming probe (0: the tuned value)")
    ap.add_argument("--asof-right", type=int, default=400 << 20)
    ap.add_argument("--asof-left", type=int, default=100 << 20)
    ap.add_argument("--asof-reps", type=int, default=5)
    ap.add_argument("--piece-rows", type=int, default=32,
                    help="rows per spline piece (with one knot per router bucket, the rows a bucket holds on average)")
    ap.add_argument("--tick-piece-rows", type=int, default=64,
                    help="rows per spline piece of the tick pipeline's 2^31-row quote index")
    ap.add_argument("--tick-cluster", type=int, default=1, choices=[0, 1],
                    help="tick pipeline: group the filtered trades by the index region their probes read")
    ap.add_argument("--train-steps", type=int, default=300)
    ap.add_argument("--train-lr", type=float, default=0.003)
    ap.add_argument("--bucket-knots", type=int, default=1, help="spline knots per learned-router bucket")
    ap.add_argument("--data-lines", type=int, default=1, choices=[0, 1],
                    help="learned index: bucket rows inline where they fit (1) or knot lines only (0)")
    ap.add_argument("--asof-sessions", type=int, default=0,
                    help="0: bursty exponential gaps (the tick shape); >0: trading days with U-shaped activity")
    ap.add_argument("--cpu-asof-log2", type=int, default=22)
    ap.add_argument("--batch-queries", type=int, default=64)
    ap.add_argument("--batch-morsels", type=int, default=4096)
    ap.add_argument("--batch-max-log2", type=int, default=21)
    ap.add_argument("--no-per-batch", action="store_true", help="skip the per-batch baseline (profiling runs)")
    ap.add_argument("--filter-log2", type=int, default=30)
    ap.add_argument("--tick-trades-log2", type=int, default=30)
    ap.add_argument("--tick-quotes-log2", type=int, default=31)
    ap.add_argument("--filter-null-frac", type=float, default=0.02)
    ap.add_argument("--only", choices=["join", "asof", "batched", "filter", "tick", "all"], default="all",
                    help="join: headline workload only; asof / batched: that secondary workload alone")
    ap.add_argument("--quick", action="store_true", help="small sizes (smoke)")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run one 1 GiB stream copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              f"run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.quick:
        args.build_log2, args.probe_log2, args.cpu_probe_log2, args.cpu_ref_log2 = 16, 20, 18, 16
        args.asof_right, args.asof_left, args.cpu_asof_log2 = 1 << 22, 1 << 20, 18
        args.batch_queries, args.batch_morsels, args.batch_max_log2 = 8, 64, 16
        args.filter_log2 = 20
        args.tick_trades_log2, args.tick_quotes_log2 = 18, 19
    qe = load_pkg()
    ctx = qe.Context(0)
    if args.calibrate_copy:
        # PMC calibration run: one copy of known size (1 GiB read + 1 GiB write).
        print(json.dumps({"copy": copy_ceiling(qe, ctx, reps=0)}), flush=True)
        return
    if args.only in ("asof", "batched", "filter", "tick"):
        # One secondary workload on its own (diagnostics; not the headline line).
        fn = {"asof": asof_workload, "batched": batched_workload, "filter": filter_workload,
              "tick": tick_workload}[args.only]
        print(json.dumps({args.only: fn(qe, ctx, args)}), flush=True)
        return
    ceiling = copy_ceiling(qe, ctx)
    rceil = read_ceiling(qe, ctx)
    log(f"stream copy ceiling {ceiling['gbs']:.0f} GB/s, read-only {rceil['gbs']:.0f} GB/s")
    res, cols = join_groupby_workload(qe, ctx, args)
    log(f"join+groupby: {res['ms_per_step']:.3f} ms/step, kernel {res['kernel_ms']:.3f} ms")
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline_join(qe, ctx, cols, args)
        log(f"cpu baseline {cpu['value']:.3e} rows/s on {cpu['cores']} threads, parity {cpu['gpu_parity_on_sample']}")
    for c in cols.values():
        c.free()
    secondary = {}
    if args.only == "all":
        secondary["groupby_1m"] = groupby_1m(qe, ctx, args)
        log(f"group-by 1M: {secondary['groupby_1m']['ms_per_query']:.3f} ms")
        secondary["asof"] = asof_workload(qe, ctx, args)
        a = secondary["asof"]
        log(f"as-of: binary {a['binary_ms']:.3f} ms, learned {a['learned_ms']:.3f} ms "
            f"(x{a['speedup_learned_vs_binary']:.2f}), identical={a['learned_equals_binary']}; sorted-position "
            f"output: binary {a['binary_sorted_ms']:.3f} ms, learned {a['learned_sorted_ms']:.3f} ms")
        secondary["filter"] = filter_workload(qe, ctx, args)
        f = secondary["filter"]
        log(f"filter: {f['ms_per_step']:.3f} ms/step ({f['rows_per_s']:.3e} rows/s), mask {f['pred_mask_gbs']:.0f} GB/s, "
            f"compact {f['compact_gbs']:.0f} GB/s")
        secondary["tick_pipeline"] = tick_workload(qe, ctx, args)
        tk = secondary["tick_pipeline"]
        log(f"tick pipeline: {tk['ms_per_step']:.1f} ms/step ({tk['trade_rows_per_s']:.3e} trade rows/s), "
            f"stages {', '.join(f'{k} {v:.1f}' for k, v in tk['stage_ms'].items())} ms")
        secondary["batched"] = batched_workload(qe, ctx, args)
        b = secondary["batched"]
        log(f"batched: graph {b['graph_s'] * 1e3:.2f} ms, per-batch {b['per_batch_s'] * 1e3:.1f} ms "
            f"(x{b['speedup_graph_vs_per_batch']:.1f})")
    kach = res["kernel_achieved_gbs"]
    line = {
        "metric": "hash join + group-by probe throughput (rows/s) and fraction of HBM peak",
        "value": res["value"],
        "unit": "probe rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64,f64",
        "data": "synthetic (seeded, generated in HBM)",
        "config": {
            "workload": "inner hash join + group-by: build 2^%d rows (unique I64 key permutation, I32 group in "
                        "[0,10000)), probe 2^%d rows (Zipf(%g) keys, %g%% misses, F64 payload), SUM/COUNT per "
                        "group" % (args.build_log2, args.probe_log2, args.zipf, 100 * args.miss),
            "build_rows": res["build_rows"], "probe_rows": res["probe_rows"], "groups": 10000,
            "parallelism": "single GPU", "tuning": res["tuning"],
        },
        "roofline": {
            "bound": "hbm", "achieved": kach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (kach / HBM_PEAK_GBS) if kach else None, "traffic": measured_traffic(args),
            "kernel": res["kernel"], "kernel_ms": res["kernel_ms"],
            "algorithmic_bytes_per_launch": res["probe_bytes"],
            "measured_copy_ceiling_gbs": ceiling["gbs"],
            "measured_read_ceiling_gbs": rceil["gbs"],
            "frac_of_measured_read_ceiling": (kach / rceil["gbs"]) if kach else None,
            "map_gather_floor_ms": res["map_gather_floor_ms"],
            "kernel_over_gather_floor": res["kernel_over_gather_floor"],
        },
        "step_latency_ms": {"p50": res["step_ms_p50"], "p99": res["step_ms_p99"], "reps": res["step_reps"]},
        "step_roofline": {"achieved": res["step_achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": res["step_achieved_gbs"] / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_step": res["step_bytes"]},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "tags": result_tags(qe, ctx, args),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
