"""GPU Filter (qe_gpu_filter / qe_gpu_filter_select) against the CPU oracle.

Cases follow the reference's Filter tests (tests/ops_test.cpp:99-146: all-pass,
all-fail, NULL rows excluded, c0 >= 100 over the 300-row sequence table), the
comparison truth table (tests/expr_null_truthtable_test.cpp: null-if-any,
three-valued AND), IEEE NaN and numeric promotion (expr/expr.h:22-40,
expr/expr.cpp:133-151), empty and ragged sizes around the 4096-row tile, and
seeded random conjunctions. Selections are bit-exact; compacted columns carry
the same values and validity as the rows the selection names.
"""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

I32, I64, F64, BOOL, TS = O.I32, O.I64, O.F64, O.BOOL, O.TS
LT, LE, GT, GE, EQ, NE = O.LT, O.LE, O.GT, O.GE, O.EQ, O.NE


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def upload(ctx, cols):
    return [ctx.upload(v, t, valid) for v, valid, t in cols]


def gpu_select(ctx, cols, terms):
    dev = upload(ctx, cols)
    try:
        return ctx.filter_select(dev, terms)
    finally:
        for d in dev:
            d.free()


def check(ctx, cols, terms):
    got = gpu_select(ctx, cols, terms)
    want = O.filter_select(cols, terms)
    assert got.dtype == np.uint32 and got.tolist() == want.tolist()
    return got


def check_compact(ctx, cols, terms):
    """qe_gpu_filter: every column gathered through the oracle's selection."""
    want = O.filter_select(cols, terms)
    dev = upload(ctx, cols)
    try:
        t = ctx.filter(dev, terms)
        try:
            out = t.numpy()
        finally:
            t.free()
    finally:
        for d in dev:
            d.free()
    assert len(out) == len(cols)
    for (vals, valid, _), (gv, gvalid) in zip(cols, out):
        assert len(gv) == len(want)
        ev = np.ones(len(want), dtype=bool) if valid is None else valid[want]
        gm = np.ones(len(want), dtype=bool) if gvalid is None else gvalid
        assert (gm == ev).all()
        a, b = vals[want][ev], gv[ev]
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))  # bitwise, NaN included


def seq_table(n):
    i = np.arange(n)
    return [col(i, I32, tuple(np.nonzero(i % 7 == 0)[0])), col(i * 1.5, F64)]


def test_reference_filter_cases(ctx):
    t = seq_table(300)
    got = check(ctx, t, [(0, GE, I32, 100)])
    assert got.tolist() == [i for i in range(100, 300) if i % 7]
    assert len(check(ctx, t, [(1, GE, F64, 0.0)])) == 300
    assert len(check(ctx, t, [(1, LT, F64, 0.0)])) == 0
    check(ctx, [col([10, 20, 30, 40, 50], I32)], [(0, GT, I32, 15)])
    check_compact(ctx, t, [(0, GE, I32, 100)])


def test_nulls_nan_and_promotion(ctx):
    a = col([10, 20, 30, 40], I32, (1, 3))
    assert check(ctx, [a], [(0, GT, I32, 0)]).tolist() == [0, 2]
    assert check(ctx, [a], [(0, GT, I32, None)]).tolist() == []
    x = col([np.nan, 1.0, -0.0, 0.0, np.inf, -np.inf], F64)
    for op in range(6):
        for lit in (1.0, 0.0, -0.0, float("nan"), float("inf")):
            check(ctx, [x], [(0, op, F64, lit)])
    big = col([2**53, 2**53 + 1, 2**53 + 2, -(2**63), 2**63 - 1], I64)
    for op in range(6):
        check(ctx, [big], [(0, op, F64, float(2**53))])
        check(ctx, [big], [(0, op, I64, 2**53 + 1)])
        check(ctx, [big], [(0, op, I32, -1)])
        check(ctx, [big], [(0, op, F64, 9.3e18)])
    check(ctx, [col([1, 2, 3], I32)], [(0, LT, F64, 2.5)])
    check(ctx, [col([0, 1, 1, 0], BOOL)], [(0, EQ, BOOL, 1)])
    check(ctx, [col([5, -5, 7], TS)], [(0, LE, TS, 5)])


def test_three_valued_and(ctx):
    L = col([0, 0, 0, 1, 1, 1, 0, 0, 0], I32, (6, 7, 8))
    R = col([0, 1, 0, 0, 1, 0, 0, 1, 0], I32, (2, 5, 8))
    assert check(ctx, [L, R], [(0, EQ, I32, 1), (1, EQ, I32, 1)]).tolist() == [4]


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 255, 256, 4095, 4096, 4097, 8191, 12289, 100003])
def test_ragged_sizes_keep_order(ctx, n):
    r = np.random.default_rng(n)
    v = r.integers(-100, 100, n)
    valid = r.random(n) >= 0.2
    cols = [(v, valid, I64), (r.normal(size=n), None, F64)]
    check(ctx, cols, [(0, GE, I32, -20)])
    check(ctx, cols, [(0, NE, I64, 3), (1, LT, F64, 0.5)])
    if n:
        check_compact(ctx, cols, [(0, GE, I32, -20)])


def _rand_case(r, n):
    from test_oracle_golden import _rand_filter_case
    return _rand_filter_case(r, n)


@pytest.mark.parametrize("case", range(16))
def test_random_conjunctions(ctx, case):
    r = np.random.default_rng(1400 + case)
    n = int(r.integers(0, 20000))
    cols, terms = _rand_case(r, n)
    check(ctx, cols, terms)
    if case % 4 == 0:
        check_compact(ctx, cols, terms)


def test_all_and_none_pass_large(ctx):
    n = 3_000_017
    v = np.arange(n, dtype=np.int64)
    cols = [(v, None, I64)]
    assert len(check(ctx, cols, [(0, GE, I64, 0)])) == n
    assert len(check(ctx, cols, [(0, LT, I64, 0)])) == 0
    got = check(ctx, cols, [(0, GE, I64, n // 3), (0, LT, I64, 2 * n // 3)])
    assert got.tolist() == list(range(n // 3, 2 * n // 3))


def test_rejects_bad_terms(ctx, qe):
    dev = upload(ctx, [col([1, 2, 3], I32), col([1, 0, 1], BOOL)])
    try:
        for terms in ([(2, LT, I32, 1)],          # column out of range
                      [(0, 9, I32, 1)],           # unknown operator
                      [(0, LT, BOOL, 1)],         # I32 vs BOOL
                      [(1, LT, I32, 1)],          # BOOL vs I32
                      [(0, LT, I32, 2**40)],      # I32 literal out of range
                      [(1, EQ, BOOL, 2)],         # BOOL literal not 0/1
                      []):
            with pytest.raises(qe.InvalidArgument):
                ctx.filter_select(dev, terms)
        with pytest.raises(qe.Unsupported):
            ctx.filter_select(dev, [(0, LT, I32, 1)] * 9)
    finally:
        for d in dev:
            d.free()


def test_many_columns_all_widths_and_tile_edges(ctx):
    # 11 columns (more than one compaction launch), every value width, NULLs in
    # several columns, over 300 tiles so output validity words straddle tiles.
    r = np.random.default_rng(77)
    n = 1_234_567
    types = [I32, I64, F64, BOOL, TS, I32, F64, I64, BOOL, F64, I32]
    cols = []
    for j, t in enumerate(types):
        if t == F64:
            v = r.normal(size=n)
        elif t == BOOL:
            v = r.integers(0, 2, n)
        else:
            v = r.integers(-1000, 1000, n)
        valid = (r.random(n) >= 0.05) if j % 3 == 0 else None
        cols.append((np.asarray(v, dtype=O.NP[t]), valid, t))
    for terms in ([(2, GT, F64, 0.0)], [(0, LT, I32, 500), (4, GE, TS, -900)], [(3, EQ, BOOL, 1)],
                  [(6, GT, F64, 3.5)]):
        check_compact(ctx, cols, terms)
        check(ctx, cols, terms)
