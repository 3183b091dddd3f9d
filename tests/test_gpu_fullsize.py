"""Parity at the full sizes of the two headline workloads.

* Join -> group-by: 2^24 unique build keys with an I32 group column (10,000
  groups), 2^28 Zipf(1.0) probe keys with 10% misses and an F64 value; SUM and
  COUNT(*) per group (Aggregate over HashJoin, ops/join.cpp:108-236 and
  ops/aggregate.cpp:128-131).  The GPU table is compared with the oracle's
  single-threaded join -> aggregate with Neumaier-compensated sums: keys and
  COUNTs exact, SUM within 1e-12 relative.  The hot group folds about 1.4e7
  addends.  Two prepared runs must agree bit for bit.
* As-of: 400*2^20 right rows (5000 Zipf(1.1) symbols, bursty timestamps),
  100*2^20 left probes (tsx/asof.cpp:234-308 semantics: greatest right ts <=
  left ts within the symbol).  The device sort is checked by invariants that
  pin it uniquely as the (sym, ts, row) order: CSR offsets, rows a
  permutation, (sym, ts) of rows[i] equal to the sorted arrays, ts
  non-decreasing and rows increasing among equal ts inside each partition.
  Then the binary, learned and linear-router probes, each with original-row
  and sorted-position output, are compared with a multi-threaded CPU
  upper_bound over the verified sorted arrays on every probe.
"""
import time

import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu
NO = 0xFFFFFFFF


def _say(msg):
    print(f"[fullsize] {msg}", flush=True)


def test_join_groupby_full_size_vs_neumaier_oracle(ctx, qe):
    build_n, probe_n, groups = 1 << 24, 1 << 28, 10000
    (pk, pp), (bk, bg) = ctx.gen_join(20260615, build_n, probe_n, groups, 1.0, 0.1)
    aggs = [(qe.SUM, 1), (qe.COUNT_STAR, 0)]
    try:
        run = ctx.prepare_join_aggregate([pk, pp], [bk, bg], [0], [0], qe.INNER, [3], aggs)
        tabs = []
        for _ in range(2):
            t = run()
            tabs.append(t.numpy())
            t.free()
        cols_p = [pk.numpy()[0], pp.numpy()[0]]
        cols_b = [bk.numpy()[0], bg.numpy()[0]]
    finally:
        for d in (pk, pp, bk, bg):
            d.free()
    # Run to run: identical rows in identical order, sums identical in bits.
    a, b = tabs
    assert len(a) == len(b) == 3
    for (va, ma), (vb, mb) in zip(a, b):
        assert va.dtype == vb.dtype and len(va) == len(vb)
        assert np.array_equal(va.view(np.uint8), vb.view(np.uint8))
        assert (ma is None) == (mb is None)
    got = table_rows(a)
    t0 = time.perf_counter()
    pcols = [(cols_p[0], None, O.I64), (cols_p[1], None, O.F64)]
    bcols = [(cols_b[0], None, O.I64), (cols_b[1], None, O.I32)]
    want = O.join_aggregate(pcols, bcols, [0], [0], 0, [3],
                            [(O.SUM, 1), (O.COUNT_STAR, 0)], sum_mode=1)
    _say(f"oracle join_aggregate over {probe_n} probe rows: {time.perf_counter() - t0:.1f} s")
    assert len(want) == len(got) and len(got) <= groups
    assert_same(got, want)
    matched = sum(r[2] for r in got)
    assert 0.85 * probe_n < matched < 0.95 * probe_n
    assert max(r[2] for r in got) > 10_000_000  # the hot group


def _partition_ids(offs, n):
    return np.repeat(np.arange(len(offs) - 1, dtype=np.int32), np.diff(offs).astype(np.int64))


def test_asof_full_size_every_probe(ctx, qe):
    right_n, left_n, nsym = 400 << 20, 100 << 20, 5000
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(20260616, right_n, left_n, nsym, 1.1)
    try:
        idx = ctx.asof_build([rs], rt)
        try:
            idx.train(32, 300, 0.003)  # the bench's pieces: most buckets are data lines
            res = {}
            for name, mode in (("binary", qe.ASOF_BINARY), ("learned", qe.ASOF_LEARNED),
                               ("learned_no_mlp", qe.ASOF_LEARNED_NO_MLP)):
                res[name] = idx.probe([ls], lt, mode=mode)
                res[name + "_sorted"] = idx.probe([ls], lt, mode=mode | qe.ASOF_OUT_SORTED)
            # The probes clustered by index region, each carrying its row id.
            rid = ctx.upload(np.arange(left_n, dtype=np.int64))
            try:
                cl = idx.cluster([ls], lt, [ls, lt, rid])
                try:
                    clustered = idx.probe([cl.col(0)], cl.col(1), mode=qe.ASOF_LEARNED | qe.ASOF_OUT_SORTED)
                    perm = cl.col(2).numpy()[0]
                finally:
                    cl.free()
            finally:
                rid.free()
            st = idx.stats()
            nparts = int(st["partitions"])
            sts, rows, offs = idx.sorted(right_n, nparts)
        finally:
            idx.free()
        rsy, _ = rs.numpy()
        rts, _ = rt.numpy()
        lsy, _ = ls.numpy()
        lts, _ = lt.numpy()
    finally:
        for c in (rs, rt, rp, ls, lt):
            c.free()
    _say(f"as-of: {nparts} partitions, fallbacks {st['fallbacks']}, data lines {st['data_line_buckets']} of "
         f"{st['router_buckets']} buckets")
    assert st["fallbacks"] == 0
    assert st["data_line_buckets"] > 0.5 * st["router_buckets"]
    smin = int(rsy.min())
    assert nparts == int(rsy.max()) - smin + 1  # one direct partition per symbol value

    # CSR offsets.
    assert offs[0] == 0 and offs[-1] == right_n and (np.diff(offs.astype(np.int64)) >= 0).all()
    # rows is a permutation of [0, n).
    seen = np.zeros(right_n, dtype=bool)
    seen[rows] = True
    assert seen.all()
    del seen
    # Sorted timestamps are the right side's, in rows order; each row sits in
    # its symbol's partition.
    assert np.array_equal(rts[rows], sts)
    part = _partition_ids(offs, right_n)
    assert np.array_equal(rsy[rows].astype(np.int64) - smin, part)
    # Inside a partition: ts non-decreasing, and among equal ts rows increase
    # (a stable sort), which fixes the order completely.
    same = part[1:] == part[:-1]
    del part
    d = np.diff(sts)
    assert (d[same] >= 0).all()
    tie = same & (d == 0)
    del d, same
    r = rows.astype(np.int64)
    assert (r[1:][tie] > r[:-1][tie]).all()
    del r, tie
    _say("sort invariants hold")

    # Every probe against a CPU upper_bound over the verified arrays.
    psym = lsy.astype(np.int64) - smin
    psym = np.where((psym >= 0) & (psym < nparts), psym, -1).astype(np.int32)
    L = O.oracle()
    cpu = np.empty(left_n, dtype=np.uint32)
    t0 = time.perf_counter()
    L.orc_bl_asof_upper_bound(sts.ctypes.data, offs.ctypes.data, nparts, psym.ctypes.data, lts.ctypes.data,
                              left_n, 16, cpu.ctypes.data)
    _say(f"cpu upper_bound over {left_n} probes: {time.perf_counter() - t0:.1f} s")
    want = np.full(left_n, NO, dtype=np.uint32)
    hit = cpu != NO
    want[hit] = rows[cpu[hit]]
    assert 0.5 * left_n < hit.sum() <= left_n
    for name, got in res.items():
        # Sorted-position outputs are the CPU search's own indices.
        bad = np.flatnonzero(got != (cpu if name.endswith("_sorted") else want))
        assert bad.size == 0, f"{name}: {bad.size} probes differ, first at {bad[:5].tolist()}"
    # Clustered: a permutation of the probes, each with its own row's match.
    assert np.array_equal(np.bincount(perm, minlength=left_n), np.ones(left_n, dtype=np.int64))
    bad = np.flatnonzero(clustered != cpu[perm])
    assert bad.size == 0, f"clustered: {bad.size} probes differ, first at {bad[:5].tolist()}"
