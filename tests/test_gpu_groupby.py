"""GPU hash group-by (qe_gpu_aggregate) against the CPU oracle.

Cases follow the reference suites: the edge table and edge shapes of
tests/aggregate_differential_test.cpp:84-237 (NULL-key group, composite key,
global, all-NULL group, skew, empty input, NaN MIN/MAX), seeded random
queries, and the 1M-row group-by shape. Integers/COUNTs/keys are bit-exact;
F64 SUM/AVG are within 1e-12 relative of the Neumaier-summed oracle.
"""
import math

import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu

I32, I64, F64, BOOL, TS = O.I32, O.I64, O.F64, O.BOOL, O.TS
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def gpu_agg(ctx, cols, keys, aggs):
    dev = [ctx.upload(v, t, valid) for v, valid, t in cols]
    try:
        t = ctx.aggregate(dev, keys, aggs)
        try:
            return table_rows(t.numpy())
        finally:
            t.free()
    finally:
        for d in dev:
            d.free()


def check(ctx, cols, keys, aggs):
    got = gpu_agg(ctx, cols, keys, aggs)
    want = O.aggregate(cols, keys, aggs)
    assert_same(got, want)
    return got


def edge_table():
    # aggregate_differential_test.cpp:84-97
    return [col([1, 1, 2, 2, 3, 1, 2, 3, 0, 0], I32, nulls=(8, 9)),
            col([10, 20, 30, 40, 50, 60, 70, 80, 90, 100], I32, nulls=(0,)),
            col([1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5], F64, nulls=(2,))]


FULL = [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 1), (AVG, 1), (SUM, 2), (AVG, 2)]


def test_edge_single_key_with_null_group(ctx):
    got = check(ctx, edge_table(), [0], FULL)
    # Hand-derived: key 1 rows {0,1,5}: c1 {NULL,20,60}; c2 {1.5,2.5,6.5}.
    row1 = [r for r in got if r[0] == 1][0]
    assert row1[1:] == (3, 2, 80, 20, 60, 40.0, 10.5, 3.5)
    nullrow = [r for r in got if r[0] is None][0]
    assert nullrow[1] == 2 and nullrow[3] == 190


def test_edge_composite_key_i32_f64(ctx):
    check(ctx, edge_table(), [0, 2], [(CS, 0), (SUM, 1)])


def test_edge_global(ctx):
    got = check(ctx, edge_table(), [], FULL)
    assert len(got) == 1 and got[0][0] == 10


def test_edge_all_null_value_group(ctx):
    cols = [col([1, 1, 2, 2], I32), col([0, 0, 5, 6], I32, nulls=(0, 1))]
    got = check(ctx, cols, [0], [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 1), (AVG, 1)])
    g1 = [r for r in got if r[0] == 1][0]
    assert g1 == (1, 2, 0, None, None, None, None)


def test_edge_key_skew_crosses_batches(ctx):
    k = np.arange(6000) % 5
    v = (np.arange(6000) % 100) - 50
    cols = [col(k, I32), col(v, I64)]
    check(ctx, cols, [1], [(CS, 0), (SUM, 0)])
    check(ctx, cols, [0], [(CS, 0), (SUM, 1), (MIN, 1), (MAX, 1)])


def test_edge_empty_input(ctx):
    cols = [col([], I32), col([], F64)]
    got = check(ctx, cols, [], [(CS, 0), (CNT, 0), (SUM, 0), (MIN, 0)])
    assert got == [(0, 0, None, None)]  # global over empty: exactly one row
    assert check(ctx, cols, [0], [(CS, 0), (SUM, 1)]) == []  # keyed: zero rows


def test_edge_nan_total_order(ctx):
    nan = float("nan")
    cols = [col([1, 1, 2, 2, 3, 3, 4, 4], I32), col([nan, 1.5, nan, nan, 2.5, 3.5, nan, 0.0], F64, nulls=(7,))]
    aggs = [(MIN, 1), (MAX, 1), (SUM, 1), (AVG, 1), (CNT, 1)]
    got = {r[0]: r for r in check(ctx, cols, [0], aggs)}
    assert got[1][1] == 1.5 and math.isnan(got[1][2])
    assert math.isnan(got[2][1]) and math.isnan(got[2][2])
    assert got[3][1:3] == (2.5, 3.5)
    assert math.isnan(got[4][1]) and got[4][5] == 1
    check(ctx, cols, [], aggs)


def test_f64_key_canonicalisation(ctx):
    # -0.0 == +0.0 and NaN == NaN as group keys (ops/hashtable.h:25-32).
    nan = float("nan")
    cols = [col([0.0, -0.0, nan, float.fromhex("0x1.8p+1"), nan], F64), col([1, 2, 3, 4, 5], I64)]
    got = check(ctx, cols, [0], [(CS, 0), (SUM, 1)])
    assert len(got) == 3


def test_bool_and_ts_keys(ctx, rng):
    n = 5000
    cols = [col(rng.integers(0, 2, n), BOOL, nulls=tuple(range(0, n, 97))),
            col(rng.integers(-10**15, 10**15, n), TS),
            col(rng.normal(size=n), F64)]
    check(ctx, cols, [0], [(CS, 0), (SUM, 2), (MIN, 1), (MAX, 1)])
    check(ctx, cols, [0, 1], [(CS, 0), (AVG, 2)])


@pytest.mark.parametrize("case", range(40))
def test_random_queries(ctx, case):
    r = np.random.default_rng(1000 + case)
    n = int(r.choice([0, 1, 63, 64, 65, 2047, 2048, 2049, 10000]))
    card = int(r.choice([1, 3, 50, 1000, 100000]))
    ktype = int(r.choice([I32, I64]))
    null_p = float(r.choice([0.0, 0.05, 0.5]))

    def nulls(p):
        return tuple(np.nonzero(r.random(n) < p)[0]) if p > 0 else ()
    cols = [col(r.integers(-card, card, n), ktype, nulls(null_p)),
            col(r.integers(-1000, 1000, n), I64, nulls(null_p)),
            col(r.uniform(-1000, 1000, n), F64, nulls(null_p)),
            col(r.integers(-50, 50, n), I32)]
    nk = int(r.integers(0, 3))
    keys = [0, 3][:nk]
    aggs = [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 2), (AVG, 2), (SUM, 2), (MIN, 2), (AVG, 3)]
    check(ctx, cols, keys, aggs)


def test_hash_path_sparse_i64_keys(ctx, rng):
    n = 200000
    keys = rng.integers(-2**62, 2**62, 3000)
    cols = [col(keys[rng.integers(0, 3000, n)], I64), col(rng.uniform(0, 1000, n), F64)]
    got = check(ctx, cols, [0], [(CS, 0), (SUM, 1), (AVG, 1)])
    assert len(got) == len(np.unique(keys))


@pytest.mark.parametrize("span", [32766, 32767, 32768])
def test_emission_at_the_one_block_limit(ctx, rng, span):
    # Direct slots = key range + the NULL slot: 32767, 32768 (the largest slot
    # count emitted by the one-block group listing) and 32769 (the scan path);
    # some keys absent, NULL keys, an all-NULL value group, both emissions
    # against the oracle.
    n = 150000
    k = rng.integers(0, span - 1, n)
    k[0] = 0
    k[2:50] = span - 1  # the top key: present only here, every value NULL
    kn = tuple(range(100, 160))
    v = rng.normal(size=n) * 100
    vn = tuple(range(2, 50)) + tuple(rng.integers(200, n, 500).tolist())
    got = check(ctx, [col(k, I32, kn), col(v, F64, vn)], [0], [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (AVG, 1)])
    assert len(got) == len(np.unique(np.delete(k, kn))) + 1  # + the NULL-key group


def test_config1_shape_1m_rows(ctx, qe):
    # 2^20 rows, key I64 uniform [0,1024), F64 [0,1000) with 5% NULLs,
    # COUNT(*)/SUM/AVG per key.
    n = 1 << 20
    k, v = ctx.gen_groupby(20260615, n, 1024, 0.05)
    try:
        kv, _ = k.numpy()
        vv, valid = v.numpy()
        assert valid is not None and 0.94 < valid.mean() < 0.96
        t = ctx.aggregate([k, v], [0], [(CS, 0), (SUM, 1), (AVG, 1)])
        got = table_rows(t.numpy())
        t.free()
    finally:
        k.free()
        v.free()
    cols = [(kv, None, I64), (vv, valid, F64)]
    want = O.aggregate(cols, [0], [(CS, 0), (SUM, 1), (AVG, 1)])
    assert len(got) == 1024
    assert_same(got, want)
    # Against the reference's naive row-order fold the deviation is summation
    # order only; it stays far inside the reference's own 1e-9 tolerance.
    naive = O.aggregate(cols, [0], [(CS, 0), (SUM, 1), (AVG, 1)], sum_mode=0)
    assert_same(got, naive, rel=1e-9, abs_=1e-9)


def test_sum_of_str_is_invalid_argument(ctx, qe):
    c = ctx.upload(np.array([1, 2, 3], dtype=np.int32), qe.STR)
    try:
        with pytest.raises(ValueError, match="SUM is not defined on STR"):
            ctx.aggregate([c], [], [(SUM, 0)])
    finally:
        c.free()
