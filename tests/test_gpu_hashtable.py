"""Device HashTable (qe_gpu_ht_*) against the oracle's first-seen ids: the
literal cases of tests/hashtable_test.cpp:32-189 and
tests/catalog_checks_hash.cpp:35-53, a multi-call fuzz with growth under both
NULL policies, and the golden fixtures from the compiled reference."""
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu
I32, I64, F64 = O.I32, O.I64, O.F64
NO = O.NO_ROW
GOLDEN = Path(__file__).resolve().parent / "golden"


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


class Pair:
    """The device table and the oracle table driven with the same calls."""

    def __init__(self, ctx, types, policy=0, init_cap=1024, lf=0.7):
        self.ctx = ctx
        self.dev = ctx.hash_table(types, policy, init_cap, lf)
        self.orc = O.OracleHT(types, policy=policy, init_cap=init_cap, lf=lf)

    def _up(self, cols):
        return [self.ctx.upload(v, t, valid) for v, valid, t in cols]

    def insert_or_find(self, cols, sel=None):
        d = self._up(cols)
        try:
            got = self.dev.insert_or_find(d, sel=sel)
        finally:
            for c in d:
                c.free()
        want = self.orc.insert_or_find(cols, sel)
        assert got.tolist() == want.tolist()
        assert self.dev.num_groups == self.orc.num_groups
        return got

    def find(self, cols, sel=None):
        d = self._up(cols)
        try:
            got = self.dev.find(d, sel=sel)
        finally:
            for c in d:
                c.free()
        assert got.tolist() == self.orc.find(cols, sel).tolist()
        return got


def test_literal_dense_ids_idempotence_find(ctx):
    t = Pair(ctx, [I32])
    assert t.insert_or_find([col([7, 7, 3, 7, 3, 100], I32)]).tolist() == [0, 0, 1, 0, 1, 2]
    assert t.insert_or_find([col([7, 100], I32)]).tolist() == [0, 2]
    assert t.find([col([3, 99, 7], I32)]).tolist() == [1, NO, 0]
    assert t.dev.num_groups == 3


def test_literal_nulls_both_policies(ctx):
    t = Pair(ctx, [I32], policy=0)
    assert t.insert_or_find([col([0, 5, 0, 5, 0], I32, nulls=(0, 2, 4))]).tolist() == [0, 1, 0, 1, 0]
    w, nl = t.dev.key_words(0)
    assert nl.tolist() == [True, False] and int(w[1]) == 5
    t2 = Pair(ctx, [I32], policy=1)
    assert t2.insert_or_find([col([0, 5, 0, 5], I32, nulls=(0, 2))]).tolist() == [0, 1, 2, 1]
    assert t2.find([col([0, 5], I32, nulls=(0,))]).tolist() == [NO, 1]


def test_literal_f64_and_composite_and_selection(ctx):
    t = Pair(ctx, [F64])
    assert t.insert_or_find([col([0.0, -0.0, float("nan"), -float("nan"), 1.5], F64)]).tolist() == [0, 0, 1, 1, 2]
    c = Pair(ctx, [I32, I64])
    g = c.insert_or_find([col([1, 1, 1, 2], I32), col([10, 10, 20, 10], I64)]).tolist()
    assert g == [0, 0, 1, 2]
    s = Pair(ctx, [I32])
    sel = np.array([0, 3], dtype=np.uint32)
    assert s.insert_or_find([col([9, 9, 9, 8, 8], I32)], sel=sel).tolist() == [0, 1]


def test_catalog_collision_storm_4000_sequential(ctx):
    t = Pair(ctx, [I64], init_cap=8, lf=0.9)
    keys = [col(np.arange(4000), I64)]
    assert (t.insert_or_find(keys) == np.arange(4000)).all()
    assert (t.find(keys) == np.arange(4000)).all()


@pytest.mark.parametrize("case", range(8))
def test_fuzz_multi_call_growth(ctx, case):
    r = np.random.default_rng(3000 + case)
    policy = case % 2
    types = [[I64], [I32, I64], [F64]][case % 3]
    t = Pair(ctx, types, policy=policy, init_cap=8, lf=float(r.choice([0.5, 0.9])))
    for _ in range(6):
        n = int(r.integers(0, 20000))
        cols = []
        for ty in types:
            v = r.integers(0, 3000, n) if ty != F64 else r.choice([0.0, -0.0, 1.5, float("nan"), 2.0, 3.0], n)
            cols.append(col(v, ty, tuple(np.nonzero(r.random(n) < 0.02)[0])))
        t.insert_or_find(cols)
        t.find(cols)


def test_golden_fixtures(ctx):
    for f in sorted(GOLDEN.glob("hashtable_*.json")):
        fx = json.loads(f.read_text())
        cols = [col(c["values"], c["type"], tuple(c.get("nulls", []))) for c in fx["cols"]]
        t = ctx.hash_table(fx["types"], fx["policy"], fx["init_cap"], fx["lf"])
        d = [ctx.upload(v, ty, valid) for v, valid, ty in cols]
        got = t.insert_or_find(d)
        for x in d:
            x.free()
        t.free()
        assert got.tolist() == fx["expected_ids"], f.name
