"""Exact, bit-reproducible F64 SUM/AVG on the device.

Every finite summed F64 value is taken as an integer at a per-column scale and
added with integer atomics, so the result cannot depend on the order in which
rows, waves or blocks add (the reference folds in row order,
ops/aggregate.cpp:128-131). Checked here:
  * the same query three times gives the same bits (the fused join ->
    group-by workload scaled to 2^22 probe rows, the group-by, the batched
    group-by), and every run is within 1e-12 of the Neumaier oracle;
  * NaN and infinities follow the row-order fold (NaN if any NaN or both
    infinities, else the infinity), also when a finite sum overflows;
  * a value beyond the sampled scale (a row the 2^16-row sample skips) makes
    the query rerun with bounds from every row, and the result is the exact
    sum rounded once;
  * sums whose terms the scale represents exactly equal math.fsum bit for bit.
"""
import math

import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu
I32, I64, F64 = O.I32, O.I64, O.F64
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG


def _bits(t):
    """Result rows as raw bit patterns, ordered by the first column."""
    cols = t.numpy()
    order = np.argsort(cols[0][0], kind="stable")
    return [(c[0][order].tobytes(), None if c[1] is None else c[1][order].tobytes()) for c in cols]


def test_join_groupby_bit_reproducible(ctx):
    (pk, pp), (bk, bg) = ctx.gen_join(31, 1 << 16, 1 << 22, 10000)
    try:
        runs = []
        for _ in range(3):
            t = ctx.join_aggregate([pk, pp], [bk, bg], [0], [0], 0, [3], [(SUM, 1), (AVG, 1), (CS, 0)])
            runs.append((_bits(t), table_rows(t.numpy())))
            t.free()
        pcols = [(pk.numpy()[0], None, I64), (pp.numpy()[0], None, F64)]
        bcols = [(bk.numpy()[0], None, I64), (bg.numpy()[0], None, I32)]
    finally:
        for d in (pk, pp, bk, bg):
            d.free()
    assert runs[0][0] == runs[1][0] == runs[2][0]
    want = O.join_aggregate(pcols, bcols, [0], [0], 0, [3], [(SUM, 1), (AVG, 1), (CS, 0)])
    assert_same(runs[0][1], want)


def test_groupby_bit_reproducible(ctx, rng):
    n = 1 << 21
    k = rng.integers(0, 3000, n)
    v = rng.normal(size=n) * 1e3
    valid = rng.random(n) > 0.05
    K, V = ctx.upload(k.astype(np.int64)), ctx.upload(v, F64, valid)
    try:
        runs = []
        for _ in range(3):
            t = ctx.aggregate([K, V], [0], [(SUM, 1), (AVG, 1), (CNT, 1), (CS, 0)])
            runs.append((_bits(t), table_rows(t.numpy())))
            t.free()
        g = ctx.aggregate([V], [], [(SUM, 0), (AVG, 0)])  # global: one row
        gb = [_bits(g)]
        g.free()
        g = ctx.aggregate([V], [], [(SUM, 0), (AVG, 0)])
        gb.append(_bits(g))
        g.free()
    finally:
        K.free()
        V.free()
    assert runs[0][0] == runs[1][0] == runs[2][0]
    assert gb[0] == gb[1]
    cols = [(k, None, I64), (v, valid, F64)]
    assert_same(runs[0][1], O.aggregate(cols, [0], [(SUM, 1), (AVG, 1), (CNT, 1), (CS, 0)]))


def test_batched_bit_reproducible(ctx, qe):
    r = np.random.default_rng(77)
    queries, hostcols = [], []
    for q in range(4):
        sizes = r.integers(500, 60000, 6)
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        n = int(offs[-1])
        groups = [8, 5000, 40000, 200][q]
        key = ctx.gen_uniform(91, 2 * q, I32 if q % 2 else I64, n, 0, groups, 0.0)
        val = ctx.gen_uniform(91, 2 * q + 1, F64, n, -1e6, 1e6, 0.02)
        aggs = [(CS, 0), (SUM, 1), (AVG, 1)]
        queries.append((key, val, offs, aggs, groups))
        kv, _ = key.numpy()
        vv, vvalid = val.numpy()
        hostcols.append(([(kv, None, I32 if q % 2 else I64), (vv, vvalid, F64)], aggs))
    plan = ctx.batch(queries)
    try:
        first = None
        for mode, tuning in ((0, (4096, 2, 0)), (2, (4096, 2, 0)), (0, (16384, 4, 256))):
            tabs = plan.run(mode, tuning=tuning)
            bits = []
            for t, (cols, aggs) in zip(tabs, hostcols):
                bits.append(_bits(t))
                assert_same(table_rows(t.numpy()), O.aggregate(cols, [0], aggs))
                t.free()
            if first is None:
                first = bits
            assert bits == first  # every mode and launch shape: the same bits
    finally:
        plan.free()
        for k, v, *_ in queries:
            k.free()
            v.free()


def _seq_sum(xs):
    s = 0.0
    for x in xs:
        s += x
    return s


def _sum_avg_by_group(ctx, groups):
    k = np.array([g for g, xs in groups.items() for _ in xs], dtype=np.int64)
    v = np.array([x for xs in groups.values() for x in xs])
    K, V = ctx.upload(k), ctx.upload(v)
    try:
        t = ctx.aggregate([K, V], [0], [(SUM, 1), (AVG, 1)])
        rows = {r[0]: r[1:] for r in table_rows(t.numpy())}
        t.free()
    finally:
        K.free()
        V.free()
    for g, xs in groups.items():
        want = _seq_sum(xs)
        s, a = rows[g]
        if math.isnan(want):
            assert math.isnan(s) and math.isnan(a), g
        else:
            assert s == want, (g, s, want)
            assert a == want / len(xs), g


@pytest.mark.parametrize("case", ["nonfinite", "overflow"])
def test_nonfinite_and_overflow_follow_the_row_order_fold(ctx, case):
    inf, nan = math.inf, math.nan
    if case == "nonfinite":
        groups = {0: [1.0, nan, 2.0], 1: [inf, 2.0, 3.0], 2: [-inf, 5.0], 3: [inf, 1.0, -inf],
                  6: [0.5, 0.25, -0.125], 7: [nan, inf]}
    else:  # finite sums past DBL_MAX become infinities
        groups = {4: [1e308, 1e308], 5: [-1e308, -1e308, 1e300], 8: [1.5e308, -1e308]}
    _sum_avg_by_group(ctx, groups)


def _rows_and_launches(ctx, k, v, timer):
    K, V = ctx.upload(k), ctx.upload(v)
    try:
        ctx.set_timing(True)
        t = ctx.aggregate([K, V], [0], [(SUM, 1), (CS, 0)])
        launches = ctx.timing(timer)[1]
        ctx.set_timing(False)
        rows = {r[0]: r[1:] for r in table_rows(t.numpy())}
        t.free()
    finally:
        K.free()
        V.free()
    return rows, launches


def test_outliers_past_the_sample_keep_every_group_exact(ctx):
    # 2^20 rows in 7 groups: the scale sample reads every 16th row; rows 12345
    # and 54321 (not sampled) are 1e30 and -3e25, far past the sampled bound.
    # They go to the outlier list (summed in a fixed order at emission), so
    # the other groups keep the scale of the sampled values and stay exact.
    n = 1 << 20
    v = np.full(n, 0.1)
    v[12345] = 1e30
    v[54321] = -3e25
    k = np.arange(n, dtype=np.int64) % 7
    rows, launches = _rows_and_launches(ctx, k, v, "agg_direct")
    assert launches == 1  # no rerun
    for g in range(7):
        xs = v[k == g]
        want = math.fsum(xs)
        if np.abs(xs).max() > 1e20:  # exact sum of the ordinary rows + the outliers: two roundings
            assert abs(rows[g][0] - want) <= 2 * math.ulp(want), g
        else:
            assert rows[g][0] == want, g


def test_outlier_list_overflow_reruns_with_exact_bounds(ctx):
    # 5000 unsampled rows (never a multiple of the 16-row sample stride) at
    # 2^20 x the sampled values: more outliers than the list holds, so the
    # query reruns with bounds from every row.
    n = 1 << 20
    r = np.random.default_rng(5)
    v = r.uniform(0, 1, n)
    big = (np.arange(1, 5001) * 16 + 7)[:5000]
    v[big] = r.uniform(1 << 20, 1 << 21, 5000)
    k = np.arange(n, dtype=np.int64) % 11
    rows, launches = _rows_and_launches(ctx, k, v, "agg_direct")
    assert launches == 2  # the sampled run and the exact rerun
    got = [(g, rows[g][0], rows[g][1]) for g in range(11)]
    want = O.aggregate([(k, None, I64), (v, None, F64)], [0], [(SUM, 1), (CS, 0)])
    assert_same(got, want)


def test_exactly_representable_terms_sum_exactly(ctx, rng):
    # Terms with at most 40 significant bits above 2^-20 and a column maximum
    # near 2^20: every term is an integer multiple of the scale's quantum, so
    # the device total is the exact sum rounded once (math.fsum), whatever
    # the order; a row-order double fold would round at every step.
    n = 1 << 20
    v = rng.integers(-(1 << 40), 1 << 40, n).astype(np.float64) / (1 << 20)
    k = rng.integers(0, 50, n).astype(np.int64)
    K, V = ctx.upload(k), ctx.upload(v)
    try:
        t = ctx.aggregate([K, V], [0], [(SUM, 1)])
        rows = {r[0]: r[1] for r in table_rows(t.numpy())}
        t.free()
    finally:
        K.free()
        V.free()
    for g in range(50):
        assert rows[g] == math.fsum(v[k == g]), g
