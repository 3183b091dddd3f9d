"""Batched group-by (qe_gpu_batch_*): independent queries over uneven morsels
through the persistent kernel, the one-launch-per-2048-row-batch baseline and
the hipGraph-captured pipeline, each checked against the oracle per query;
graph replays are checked against the eager result every time, with eager
work in between."""
import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu
I32, I64, F64 = O.I32, O.I64, O.F64
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG


def make_batch(ctx, qe, seed, nq=6, morsels=8, max_rows=40000):
    r = np.random.default_rng(seed)
    queries, hostcols = [], []
    for q in range(nq):
        sizes = np.exp(r.uniform(np.log(100), np.log(max_rows), morsels)).astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        n = int(offs[-1])
        groups = int(np.exp(r.uniform(np.log(4), np.log(max(5, n // 2)))))
        ktype = I32 if q % 2 == 0 else I64
        vtype = F64 if q % 3 != 2 else I64
        k = ctx.gen_uniform(seed, 2 * q, ktype, n, 0, groups, 0.0)
        v = ctx.gen_uniform(seed, 2 * q + 1, vtype, n, -1000, 1000, 0.02)
        aggs = [(CS, 0), (SUM, 1), (AVG, 1), (CNT, 1), (MIN, 1), (MAX, 1)]
        queries.append((k, v, offs, aggs, groups))
        kv, _ = k.numpy()
        vv, vvalid = v.numpy()
        hostcols.append(([(kv, None, ktype), (vv, vvalid, vtype)], aggs))
    return queries, hostcols


def check_outputs(tables, hostcols):
    for t, (cols, aggs) in zip(tables, hostcols):
        got = table_rows(t.numpy())
        t.free()
        assert_same(got, O.aggregate(cols, [0], aggs))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_modes_match_oracle(ctx, qe, mode):
    queries, hostcols = make_batch(ctx, qe, 10 + mode)
    plan = ctx.batch(queries)
    try:
        check_outputs(plan.run(mode, tuning=(4096, 2, 256)), hostcols)
        check_outputs(plan.run(mode, tuning=(1024, 1, 0)), hostcols)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_graph_replays_match_eager(ctx, qe):
    queries, hostcols = make_batch(ctx, qe, 21, nq=4)
    plan = ctx.batch(queries)
    try:
        eager = [table_rows(t.numpy()) for t in plan.run(qe.BATCH_PERSISTENT, tuning=(4096, 2, 256))]
        for rep in range(3):
            got = plan.run(qe.BATCH_GRAPH, tuning=(4096, 2, 256))
            for t, e in zip(got, eager):
                assert_same(table_rows(t.numpy()), e)
                t.free()
            # eager work and a synchronisation between replays
            plan.run(qe.BATCH_PERSISTENT, tuning=(2048, 1, 0), collect=False)
            ctx.sync()
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_autotune_returns_a_valid_shape(ctx, qe):
    queries, hostcols = make_batch(ctx, qe, 31, nq=3, morsels=4, max_rows=20000)
    plan = ctx.batch(queries)
    try:
        tile, bpc, C = plan.autotune()
        assert tile >= 256 and bpc >= 1 and (C == 0 or C & (C - 1) == 0)
        check_outputs(plan.run(qe.BATCH_PERSISTENT, tuning=(tile, bpc, C)), hostcols)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_strategies_small_partitioned_hash(ctx, qe):
    # One batch mixing every strategy: small dense tables, partitioned dense
    # ranges (with NULL keys), hash tables for sparse keys, F64 keys and
    # aggregates over the key column; each mode against the oracle.
    r = np.random.default_rng(77)
    specs = [
        ("small", I32, lambda n: r.integers(-5, 60, n), False),
        ("part", I64, lambda n: r.integers(0, 150000, n), False),
        ("part-nullkeys", I32, lambda n: r.integers(1000, 90000, n), True),
        ("hash-sparse", I64, lambda n: r.integers(0, 3000, n) * 1000003, False),
        ("hash-f64", F64, lambda n: r.integers(0, 500, n) * 0.5, False),
    ]
    queries, hostcols = [], []
    for i, (name, kt, gen, knull) in enumerate(specs):
        n = int(r.integers(50000, 400000))
        sizes = np.diff(np.sort(np.concatenate([[0, n], r.integers(0, n, 5)])))
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        kv = np.asarray(gen(n), dtype=O.NP[kt])
        kvalid = (r.random(n) >= 0.01) if knull else None
        vv = r.normal(size=n) * 100
        vvalid = r.random(n) >= 0.02
        aggs = [(CS, 0), (SUM, 1), (AVG, 1), (MAX, 1)]
        if name == "hash-sparse":
            aggs = aggs + [(SUM, 0), (MIN, 0)]  # aggregates over the key column
        k = ctx.upload(kv, kt, kvalid)
        v = ctx.upload(vv, F64, vvalid)
        queries.append((k, v, offs, aggs, 0))
        hostcols.append(([(kv, kvalid, kt), (vv, vvalid, F64)], aggs))
    plan = ctx.batch(queries)
    try:
        for mode in (qe.BATCH_PERSISTENT, qe.BATCH_GRAPH, qe.BATCH_PER_BATCH):
            for tuning in ((16384, 2, 0), (4096, 4, 256)):
                check_outputs(plan.run(mode, tuning=tuning), hostcols)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_key_and_value_widths(ctx, qe):
    # The streaming phases are specialised on key width (4: I32/STR codes,
    # 8: I64/TS) and value width (1: BOOL, 4: I32, 8: I64/TS/F64); every pair
    # through both direct strategies (small LDS tables and partitioned ranges),
    # nullable keys and values, against the oracle.
    r = np.random.default_rng(91)
    BOOL, TS = O.BOOL, O.TS
    queries, hostcols = [], []
    for i, (kt, vt) in enumerate([(k, v) for k in (I32, TS) for v in (BOOL, I32, TS, F64)]):
        for span in (300, 50000):  # small table / partitioned range (span <= 2 x rows)
            n = int(r.integers(30000, 250000))
            sizes = np.diff(np.sort(np.concatenate([[0, n], r.integers(0, n, 3)])))
            offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
            kv = np.asarray(r.integers(-span // 2, span // 2, n), dtype=O.NP[kt])
            kvalid = r.random(n) >= 0.01
            if vt == BOOL:
                vv = r.integers(0, 2, n).astype(np.uint8)
            elif vt == F64:
                vv = r.normal(size=n) * 1e3
            else:
                vv = np.asarray(r.integers(-(1 << 30), 1 << 30, n), dtype=O.NP[vt])
            vvalid = r.random(n) >= 0.03
            aggs = [(CS, 0), (SUM, 1), (AVG, 1), (CNT, 1), (MIN, 1), (MAX, 1)]
            k = ctx.upload(kv, kt, kvalid)
            v = ctx.upload(vv, vt, vvalid)
            queries.append((k, v, offs, aggs, 0))
            hostcols.append(([(kv, kvalid, kt), (vv, vvalid, vt)], aggs))
    plan = ctx.batch(queries)
    try:
        for mode in (qe.BATCH_PERSISTENT, qe.BATCH_GRAPH):
            check_outputs(plan.run(mode, tuning=(4096, 2, 0)), hostcols)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_partitioned_staged_and_direct_scatter(ctx, qe):
    # Partitioned queries with up to 1024 partitions are counting-sorted per
    # step in LDS before they are written; those with more (here 4.3M slots in
    # 4096-slot partitions: 1050) write each entry straight to its range. Both
    # in one batch (513, 30 and 3 partitions staged), with NULL keys and values,
    # tiles that end inside a step.
    r = np.random.default_rng(93)
    queries, hostcols = [], []
    for n, span in ((2_200_000, 4_300_000), (1_300_000, 2_100_000), (70_001, 120_000), (5_000, 9_000)):
        sizes = np.diff(np.sort(np.concatenate([[0, n], r.integers(0, n, 4)])))
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        kv = r.integers(0, span, n).astype(np.int64)
        kvalid = r.random(n) >= 0.005
        vv = r.normal(size=n) * 50
        vvalid = r.random(n) >= 0.02
        aggs = [(CS, 0), (SUM, 1), (MAX, 1)] + ([(MIN, 1)] if n > 2_000_000 else [])
        k = ctx.upload(kv, I64, kvalid)
        v = ctx.upload(vv, F64, vvalid)
        queries.append((k, v, offs, aggs, 0))
        hostcols.append(([(kv, kvalid, I64), (vv, vvalid, F64)], aggs))
    expect = [O.aggregate(cols, [0], aggs) for cols, aggs in hostcols]
    plan = ctx.batch(queries)
    try:
        for tuning in ((4096, 2, 0), (3000, 1, 0), (65536, 4, 0)):
            for t, e in zip(plan.run(qe.BATCH_PERSISTENT, tuning=tuning), expect):
                got = table_rows(t.numpy())
                t.free()
                assert_same(got, e)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def test_partitioned_pass_without_scatter(ctx, qe):
    # Dense ranges of one or two LDS partitions (4096 slots each at these
    # aggregates) are folded straight from the key and value columns, one
    # block per (row slice, partition): no histogram, no scatter. Slices of
    # 2^19 rows: the 1.5M-row queries take 3 slices whose tables fold in with
    # atomics, the short ones one slice written with plain stores. A
    # three-partition query beside them keeps the scatter. Nullable keys and
    # values; I32, I64 and TS keys; F64, I64, I32 and BOOL values (every value
    # width of the row loads); every mode, two tile shapes.
    r = np.random.default_rng(95)
    cases = [  # rows, key span, key type, value type, expected strategy
        (1_500_000, 7000, I32, F64, ("part_pass", 2)),
        (300_000, 3000, I64, I64, ("part_pass", 1)),
        (1_100_000, 2500, I32, F64, ("part_pass", 1)),
        (200_000, 6000, I64, F64, ("part_pass", 2)),
        (400_000, 11000, I64, F64, ("part", 3)),
        (600_000, 3000, O.TS, O.BOOL, ("part_pass", 1)),
        (900_000, 7000, I32, I32, ("part_pass", 2)),
    ]
    queries, hostcols = [], []
    for n, span, kt, vt, _ in cases:
        sizes = np.diff(np.sort(np.concatenate([[0, n], r.integers(0, n, 4)])))
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        kv = r.integers(0, span, n).astype(O.NP[kt])
        kvalid = r.random(n) >= 0.005
        if vt == F64:
            vv = r.normal(size=n) * 50
        elif vt == O.BOOL:
            vv = r.integers(0, 2, n).astype(np.uint8)
        else:
            vv = r.integers(-(1 << 30), 1 << 30, n).astype(O.NP[vt]) * (1 << 10 if vt == I64 else 1)
        vvalid = r.random(n) >= 0.02
        aggs = [(CS, 0), (SUM, 1), (AVG, 1), (MAX, 1)]
        k = ctx.upload(kv, kt, kvalid)
        v = ctx.upload(vv, vt, vvalid)
        queries.append((k, v, offs, aggs, 0))
        hostcols.append(([(kv, kvalid, kt), (vv, vvalid, vt)], aggs))
    expect = [O.aggregate(cols, [0], aggs) for cols, aggs in hostcols]
    plan = ctx.batch(queries)
    try:
        assert plan.strategies() == [c[4] for c in cases]
        for mode in (qe.BATCH_PERSISTENT, qe.BATCH_GRAPH, qe.BATCH_PER_BATCH):
            for tuning in ((4096, 2, 0), (65536, 4, 0)):
                for t, e in zip(plan.run(mode, tuning=tuning), expect):
                    got = table_rows(t.numpy())
                    t.free()
                    assert_same(got, e)
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()
