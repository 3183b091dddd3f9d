"""Mutation catalog against the GPU operators. Every mutant is planted in the
CPU oracle (never in device code): the GPU result must equal the clean oracle
and differ from each mutant oracle on the case designed to expose it. This
shows the parity checks bite for every hazard class of the reference catalog
(MUTATION_CATALOG.md: hash probe/growth, aggregation NULL/empty/NaN/tail,
join fan-out/drop/composite, as-of boundary/direction/key)."""
import numpy as np
import pytest

import _oracle as O
from _compare import diff, table_rows

pytestmark = pytest.mark.gpu
I32, I64, F64, TS = O.I32, O.I64, O.F64, O.TS
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG
NO = O.NO_ROW


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def up(ctx, cols):
    return [ctx.upload(v, t, valid) for v, valid, t in cols]


def free(ds):
    for d in ds:
        d.free()


def test_hashtable_mutants(ctx):
    keys = [col(np.arange(4000), I64)]
    t = ctx.hash_table([I64], 0, 8, 0.9)
    d = up(ctx, keys)
    got = t.insert_or_find(d)
    free(d)
    t.free()
    clean = O.OracleHT([I64], init_cap=8, lf=0.9).insert_or_find(keys)
    assert (got == clean).all()
    for m in (1, 2):
        mt = O.OracleHT([I64], init_cap=8, lf=0.9, mutant=m)
        mg = mt.insert_or_find(keys)
        mf = mt.find(keys)
        assert not ((mg == got).all() and (mf == got).all() and mt.num_groups == 4000), m


def test_aggregate_mutants(ctx):
    nan = float("nan")
    cases = {
        1: ([col([1, 1, 2], I32), col([10, 0, 5], I32, nulls=(1,))], [(SUM, 1), (AVG, 1)]),
        2: ([col([1, 1, 2], I32), col([0, 0, 5], I32, nulls=(0, 1))], [(SUM, 1), (MIN, 1)]),
        3: ([col([1, 1, 2], I32), col([nan, 1.0, 2.0], F64)], [(MAX, 1)]),
        4: ([col([1, 2, 3, 3], I32), col([1, 2, 3, 4], I64)], [(CS, 0), (SUM, 1)]),
    }
    for m, (cols, aggs) in cases.items():
        d = up(ctx, cols)
        t = ctx.aggregate(d, [0], aggs)
        got = table_rows(t.numpy())
        t.free()
        free(d)
        assert diff(got, O.aggregate(cols, [0], aggs)) is None
        assert diff(got, O.aggregate(cols, [0], aggs, mutant=m)) is not None, m


def _pairs(p, b):
    return sorted(zip(p.tolist(), b.tolist()))


def test_join_mutants(ctx):
    cases = {
        1: ([col([1, 2, 3], I32)], [col([1, 1, 2], I32)]),
        2: ([col([1, 2, 3], I32)], [col([1, 1, 2], I32)]),
        3: ([col([1, 1], I32), col([2.0, 9.0], F64)], [col([1, 1], I32), col([1.0, 2.0], F64)]),
    }
    for m, (pk, bk) in cases.items():
        P, B = up(ctx, pk), up(ctx, bk)
        pr = ctx.hash_join(P, B, 0)
        got = _pairs(*pr.numpy())
        pr.free()
        free(P + B)
        assert got == _pairs(*O.join_pairs(pk, bk, 0))
        assert got != _pairs(*O.join_pairs(pk, bk, 0, mutant=m)), m


def test_asof_mutants(ctx, qe):
    kb = [col([1, 1, 2], I32), col([10, 20, 15], I64)]
    kp = [col([1, 1, 2], I32), col([20, 25, 15], I64)]
    cb = [col([1, 1, 1], I32), col([7, 7, 9], I32), col([10, 20, 10], TS)]
    cp = [col([1, 1], I32), col([7, 9], I32), col([25, 25], TS)]
    cases = {1: (kp, kb, [0], 1), 2: (kp, kb, [0], 1), 3: (cp, cb, [0, 1], 2)}
    for m, (pc, bc, keys, tcol) in cases.items():
        P, B = up(ctx, pc), up(ctx, bc)
        idx = ctx.asof_build([B[k] for k in keys], B[tcol])
        got = idx.probe([P[k] for k in keys], P[tcol])
        idx.free()
        free(P + B)
        gp = _pairs(np.arange(len(got))[got != NO], got[got != NO])
        assert gp == _pairs(*O.asof_pairs(pc, bc, keys, keys, tcol, tcol, 0))
        assert gp != _pairs(*O.asof_pairs(pc, bc, keys, keys, tcol, tcol, 0, mutant=m)), m
    # LEFT (kLeftWrongNull): unmatched probe rows keep the no-row index.
    lp = [col([1, 3, 1], I32), col([20, 20, 5], I64)]
    P, B = up(ctx, lp), up(ctx, kb)
    idx = ctx.asof_build([B[0]], B[1])
    got = idx.probe([P[0]], P[1])
    idx.free()
    free(P + B)
    gl = _pairs(np.arange(len(got)), got)
    assert gl == _pairs(*O.asof_pairs(lp, kb, [0], [0], 1, 1, 1))
    assert gl != _pairs(*O.asof_pairs(lp, kb, [0], [0], 1, 1, 1, mutant=4))
