"""Pins the CPU oracle (oracle/qe_oracle.cpp) before it is trusted as the GPU
checker:
  1. the reference suites' literal known-answer cases, restated as data
     (tests/hashtable_test.cpp:32-189, tests/catalog_checks_hash.cpp:35-53,
     tests/aggregate_differential_test.cpp:84-237,
     tests/join_differential_test.cpp:103-296,
     tests/catalog_checks_join_sort.cpp:34-70,
     tests/asof_differential_test.cpp:118-318,
     tests/catalog_checks_asof.cpp:38-92);
  2. independent brute-force models written here (dict first-seen ids, dict
     group-by, nested-loop join, O(n*m) as-of, numpy stable lexsort);
  3. the compiled reference operators (oracle/_ref) on random inputs, where
     that library was built (this container; skipped elsewhere);
  4. the committed golden fixtures under tests/golden/.
Each planted oracle mutant must be caught by the same checks.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, canon, diff

I32, I64, F64, BOOL, TS = O.I32, O.I64, O.F64, O.BOOL, O.TS
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG
NO = O.NO_ROW
GOLDEN = Path(__file__).resolve().parent / "golden"

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference harness not built here")


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


# ---- splitmix64 closed form -----------------------------------------------------------
def py_mix64(x):
    M = (1 << 64) - 1
    x ^= x >> 30
    x = (x * 0xBF58476D1CE4E5B9) & M
    x ^= x >> 27
    x = (x * 0x94D049BB133111EB) & M
    x ^= x >> 31
    return x


def test_mix64_known_answers():
    L = O.oracle()
    assert L.orc_mix64(0) == 0  # the finalizer fixes 0
    for x in [1, 2, 0xDEADBEEF, (1 << 63) + 12345, (1 << 64) - 1]:
        assert L.orc_mix64(x) == py_mix64(x)
    # A bijection on a sample: no collisions.
    xs = np.arange(1 << 14, dtype=np.uint64)
    assert len({L.orc_mix64(int(v)) for v in xs}) == len(xs)


# ---- hash table literals (tests/hashtable_test.cpp) -------------------------------------
def test_ht_first_seen_dense_ids():
    t = O.OracleHT([I32])
    g = t.insert_or_find([col([7, 7, 3, 7, 3, 100], I32)])
    assert g.tolist() == [0, 0, 1, 0, 1, 2] and t.num_groups == 3


def test_ht_idempotent_across_calls():
    t = O.OracleHT([I64])
    c = [col([10, 20, 30], I64)]
    assert t.insert_or_find(c).tolist() == [0, 1, 2]
    assert t.insert_or_find(c).tolist() == [0, 1, 2] and t.num_groups == 3


def test_ht_find_misses_without_mutating():
    t = O.OracleHT([I32])
    t.insert_or_find([col([1, 2, 3], I32)])
    assert t.find([col([2, 99, 1, 12345], I32)]).tolist() == [1, NO, 0, NO]
    assert t.num_groups == 3


def test_ht_composite_keys():
    t = O.OracleHT([I32, I64])
    g = t.insert_or_find([col([1, 1, 1, 2], I32), col([10, 10, 20, 10], I64)]).tolist()
    assert g[0] == g[1] and g[0] != g[2] and g[0] != g[3] and t.num_groups == 3


def test_ht_kequal_nulls_group_together():
    t = O.OracleHT([I32], policy=0)
    g = t.insert_or_find([col([0, 5, 0, 5, 0], I32, nulls=(0, 2, 4))]).tolist()
    assert g == [0, 1, 0, 1, 0] and t.num_groups == 2
    assert t.key_is_null(0, 0) and not t.key_is_null(1, 0)
    t2 = O.OracleHT([I32, I32], policy=0)
    g2 = t2.insert_or_find([col([0, 0, 7], I32, nulls=(0, 1)), col([5, 5, 5], I32)]).tolist()
    assert g2[0] == g2[1] != g2[2] and t2.num_groups == 2


def test_ht_never_match_dead_ids():
    t = O.OracleHT([I32], policy=1)
    g = t.insert_or_find([col([0, 5, 0, 5], I32, nulls=(0, 2))]).tolist()
    assert g == [0, 1, 2, 1]  # dead ids interleave in first-seen order
    assert t.find([col([0, 5], I32, nulls=(0,))]).tolist() == [NO, 1]


def test_ht_f64_canonical_keys():
    t = O.OracleHT([F64])
    nan2 = -float("nan")
    g = t.insert_or_find([col([0.0, -0.0, float("nan"), nan2, 1.5], F64)]).tolist()
    assert g == [0, 0, 1, 1, 2]


def test_ht_key_readback():
    t = O.OracleHT([I32, F64])
    g = t.insert_or_find([col([42], I32), col([3.25], F64)]).tolist()
    assert np.uint64(t.key_word(g[0], 0)).astype(np.uint32).view(np.int32) == 42
    assert np.array([t.key_word(g[0], 1)], dtype=np.uint64).view(np.float64)[0] == 3.25


def test_ht_selection_vector():
    t = O.OracleHT([I32])
    g = t.insert_or_find([col([9, 9, 9, 8, 8], I32)], sel=[0, 3]).tolist()
    assert g == [0, 1] and t.num_groups == 2


def _sequential_4000(t):
    keys = [col(np.arange(4000), I64)]
    g = t.insert_or_find(keys)
    f = t.find(keys)
    return bool((g == np.arange(4000)).all() and (f == np.arange(4000)).all() and t.num_groups == 4000)


def test_ht_catalog_collision_storm_and_mutants():
    # tests/catalog_checks_hash.cpp:35-53: 4000 sequential keys, cap 8, LF 0.9.
    seed = 0x9E3779B97F4A7C15
    assert _sequential_4000(O.OracleHT([I64], init_cap=8, lf=0.9, seed=seed))
    assert not _sequential_4000(O.OracleHT([I64], init_cap=8, lf=0.9, seed=seed, mutant=1))
    assert not _sequential_4000(O.OracleHT([I64], init_cap=8, lf=0.9, seed=seed, mutant=2))


def _dict_model(rows_keys, nulls, policy, state):
    """First-seen ids with a Python dict (tests/hashtable_fuzz_test.cpp:54-94)."""
    out = []
    for k, nl in zip(rows_keys, nulls):
        if policy == 1 and any(nl):
            out.append(state["next"])
            state["next"] += 1
            continue
        key = tuple(None if n else v for v, n in zip(k, nl))
        if key not in state["map"]:
            state["map"][key] = state["next"]
            state["next"] += 1
        out.append(state["map"][key])
    return out


@pytest.mark.parametrize("case", range(20))
def test_ht_fuzz_vs_dict_model(case):
    r = np.random.default_rng(case)
    policy = case % 2
    ncols = 1 + case % 3
    t = O.OracleHT([I64] * ncols, policy=policy, init_cap=8, lf=float(r.choice([0.5, 0.7, 0.9])))
    state = {"map": {}, "next": 0}
    for _ in range(5):
        n = int(r.integers(0, 3000))
        cols = [col(r.integers(0, 40, n), I64, tuple(np.nonzero(r.random(n) < 0.05)[0])) for _ in range(ncols)]
        got = t.insert_or_find(cols).tolist()
        keys = list(zip(*[c[0].tolist() for c in cols])) if n else []
        nl = list(zip(*[(~c[1]).tolist() if c[1] is not None else [False] * n for c in cols])) if n else []
        assert got == _dict_model(keys, nl, policy, state)
    assert t.num_groups == state["next"]


@needs_ref
@pytest.mark.parametrize("case", range(10))
def test_ht_vs_compiled_reference(case):
    r = np.random.default_rng(100 + case)
    policy = case % 2
    types = [I32, F64, I64][: 1 + case % 3]
    a = O.OracleHT(types, policy=policy, init_cap=8, lf=0.9)
    b = O.OracleHT(types, policy=policy, init_cap=8, lf=0.9, use_ref=True)
    for _ in range(4):
        n = int(r.integers(1, 4000))
        cols = []
        for t in types:
            vals = r.integers(0, 50, n) if t != F64 else r.choice([0.0, -0.0, 1.5, float("nan"), 2.5], n)
            cols.append(col(vals, t, tuple(np.nonzero(r.random(n) < 0.05)[0])))
        sel = np.sort(r.choice(n, size=n // 2, replace=False)).astype(np.uint32) if case % 3 == 0 else None
        assert (a.insert_or_find(cols, sel) == b.insert_or_find(cols, sel)).all()
        assert (a.find(cols) == b.find(cols)).all()
    assert a.num_groups == b.num_groups
    for g in range(a.num_groups):
        for j in range(len(types)):
            assert a.key_is_null(g, j) == b.key_is_null(g, j)
            if not a.key_is_null(g, j):
                assert a.key_word(g, j) == b.key_word(g, j)


# ---- aggregate ---------------------------------------------------------------------------
def edge_table():
    return [col([1, 1, 2, 2, 3, 1, 2, 3, 0, 0], I32, nulls=(8, 9)),
            col([10, 20, 30, 40, 50, 60, 70, 80, 90, 100], I32, nulls=(0,)),
            col([1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5], F64, nulls=(2,))]


FULL = [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 1), (AVG, 1), (SUM, 2), (AVG, 2)]


def test_agg_edge_table_hand_derived():
    want = [(1, 3, 2, 80, 20, 60, 40.0, 10.5, 3.5),
            (2, 3, 3, 140, 30, 70, 140 / 3, 12.0, 6.0),
            (3, 2, 2, 130, 50, 80, 65.0, 14.0, 7.0),
            (None, 2, 2, 190, 90, 100, 95.0, 20.0, 10.0)]
    got = O.aggregate(edge_table(), [0], FULL)
    assert [r[0] for r in got] == [1, 2, 3, None]  # first-seen group order
    assert_same(got, want, rel=0, abs_=0)
    g = O.aggregate(edge_table(), [], FULL)
    assert_same(g, [(10, 9, 540, 20, 100, 60.0, 56.5, 56.5 / 9)], rel=1e-15, abs_=0)


def test_agg_all_null_group_empty_and_nan():
    got = O.aggregate([col([1, 1, 2, 2], I32), col([0, 0, 5, 6], I32, nulls=(0, 1))], [0],
                      [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 1), (AVG, 1)])
    assert got[0] == (1, 2, 0, None, None, None, None)
    empty = [col([], I32), col([], F64)]
    assert O.aggregate(empty, [], [(CS, 0), (CNT, 0), (SUM, 0), (MIN, 0)]) == [(0, 0, None, None)]
    assert O.aggregate(empty, [0], [(CS, 0), (SUM, 1)]) == []
    nan = float("nan")
    rows = O.aggregate([col([1, 1, 2, 2, 3, 3, 4, 4], I32), col([nan, 1.5, nan, nan, 2.5, 3.5, nan, 0.0], F64,
                                                                 nulls=(7,))], [0], [(MIN, 1), (MAX, 1), (CNT, 1)])
    d = {r[0]: r for r in rows}
    assert d[1][1] == 1.5 and math.isnan(d[1][2])
    assert math.isnan(d[2][1]) and math.isnan(d[2][2])
    assert d[3][1:3] == (2.5, 3.5)
    assert math.isnan(d[4][1]) and math.isnan(d[4][2]) and d[4][3] == 1


def test_agg_sum_nonfinite_both_modes():
    # The row-order fold (ops/aggregate.cpp:128-131) with IEEE addition: a lone
    # +inf or -inf stays, both infinities or any NaN give NaN. The compensated
    # mode must agree (its correction term is NaN once the sum is infinite),
    # and stays exact where the naive fold cancels (1e16 + 1 - 1e16).
    inf, nan = float("inf"), float("nan")
    g = col([1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 5], I32)
    v = col([inf, 1.0, -inf, 2.0, inf, -inf, nan, 3.0, 1e16, 1.0, -1e16], F64)
    for mode in (0, 1):
        d = {r[0]: r for r in O.aggregate([g, v], [0], [(SUM, 1), (AVG, 1)], sum_mode=mode)}
        assert d[1][1] == inf and d[1][2] == inf, mode
        assert d[2][1] == -inf and d[2][2] == -inf, mode
        assert math.isnan(d[3][1]) and math.isnan(d[4][1]), mode
        assert d[5][1] == (1.0 if mode == 1 else 0.0), mode


def _py_groupby(cols, keys, aggs):
    groups = {}
    order = []
    n = len(cols[0][0]) if cols else 0

    def val(c, i):
        v, valid, t = c
        if valid is not None and not valid[i]:
            return None
        x = v[i]
        return float(x) if t == F64 else int(x)

    def kval(c, i):
        x = val(c, i)
        if isinstance(x, float):
            if math.isnan(x):
                return ("nan",)
            if x == 0.0:
                return 0.0
        return x
    for i in range(n):
        k = tuple(kval(cols[j], i) for j in keys)
        if k not in groups:
            groups[k] = []
            order.append(k)
        groups[k].append(i)
    if not keys:
        order = [()]
        groups[()] = list(range(n))
    out = []
    for k in order:
        rows = groups[k]
        rec = [None if kk is None else (float("nan") if kk == ("nan",) else kk) for kk in k]
        for f, c in aggs:
            if f == CS:
                rec.append(len(rows))
                continue
            vs = [val(cols[c], i) for i in rows]
            vs = [v for v in vs if v is not None]
            if f == CNT:
                rec.append(len(vs))
            elif not vs:
                rec.append(None)
            elif f == SUM:
                rec.append(math.fsum(vs) if cols[c][2] == F64 else sum(vs))
            elif f == AVG:
                rec.append((math.fsum(vs) if cols[c][2] == F64 else float(sum(vs))) / len(vs))
            elif f == MIN:
                if cols[c][2] == F64:
                    real = [v for v in vs if not math.isnan(v)]
                    rec.append(min(real) if real else float("nan"))
                else:
                    rec.append(min(vs))
            else:
                if cols[c][2] == F64:
                    rec.append(float("nan") if any(math.isnan(v) for v in vs) else max(vs))
                else:
                    rec.append(max(vs))
        out.append(tuple(rec))
    return out


@pytest.mark.parametrize("case", range(20))
def test_agg_vs_bruteforce(case):
    r = np.random.default_rng(200 + case)
    n = int(r.choice([0, 1, 5, 100, 3000]))
    p = float(r.choice([0.0, 0.1, 0.6]))

    def nulls():
        return tuple(np.nonzero(r.random(n) < p)[0]) if p else ()
    cols = [col(r.integers(-5, 5, n), I32, nulls()), col(r.integers(-10**6, 10**6, n), I64, nulls()),
            col(r.normal(0, 100, n), F64, nulls()), col(r.choice([0.0, -0.0, 1.0, 2.0], n), F64)]
    keys = [[], [0], [0, 3], [3]][case % 4]
    aggs = [(CS, 0), (CNT, 1), (SUM, 1), (MIN, 1), (MAX, 2), (AVG, 2), (SUM, 2), (MIN, 2)]
    got = O.aggregate(cols, keys, aggs)
    want = _py_groupby(cols, keys, aggs)
    assert_same(got, want, rel=1e-13, abs_=1e-9)


def test_agg_mutants_are_caught():
    t = edge_table()
    clean = O.aggregate(t, [0], FULL)
    for m in (1, 4):
        assert diff(O.aggregate(t, [0], FULL, mutant=m), clean) is not None, m
    allnull = [col([1, 1, 2], I32), col([0, 0, 5], I32, nulls=(0, 1))]
    aggs = [(SUM, 1), (MIN, 1)]
    assert diff(O.aggregate(allnull, [0], aggs, mutant=2), O.aggregate(allnull, [0], aggs)) is not None
    nan = float("nan")
    nt = [col([1, 1], I32), col([nan, 1.0], F64)]
    assert diff(O.aggregate(nt, [0], [(MAX, 1)], mutant=3), O.aggregate(nt, [0], [(MAX, 1)])) is not None


@needs_ref
@pytest.mark.parametrize("case", range(12))
def test_agg_vs_compiled_reference_bit_exact(case):
    # sum_mode=0 folds F64 in row order exactly like the reference, so even the
    # F64 sums match bit for bit.
    r = np.random.default_rng(300 + case)
    n = int(r.choice([0, 1, 64, 2047, 2049, 9000]))
    p = float(r.choice([0.0, 0.2]))

    def nulls():
        return tuple(np.nonzero(r.random(n) < p)[0]) if p else ()
    cols = [col(r.integers(-20, 20, n), I32, nulls()), col(r.integers(-10**9, 10**9, n), I64, nulls()),
            col(r.normal(0, 1e3, n), F64, nulls())]
    keys = [[], [0], [0, 2]][case % 3]
    aggs = [(CS, 0), (CNT, 2), (SUM, 1), (MIN, 1), (MAX, 1), (AVG, 1), (SUM, 2), (AVG, 2), (MIN, 2), (MAX, 2)]
    got = O.aggregate(cols, keys, aggs, sum_mode=0)
    want = O.ref_aggregate(cols, keys, aggs, batch_size=int(r.choice([64, 256, 2048])))
    if keys:
        assert diff(got, want, rel=0, abs_=0) is None
    else:
        # The zero-key path reduces F64 in SIMD lane order (ops/agg_kernels.cpp:26-35),
        # which depends on the host vector width: order-only differences.
        assert diff(got, want, rel=1e-12, abs_=0) is None


# ---- join ----------------------------------------------------------------------------------
def pairs(p, b):
    return sorted(zip(p.tolist(), b.tolist()))


def test_join_literals():
    bk = [col([1, 1, 2, 5, 5, 5], I32)]
    pk = [col([1, 2, 3, 5, 0], I32, nulls=(4,))]
    inner = [(0, 0), (0, 1), (1, 2), (3, 3), (3, 4), (3, 5)]
    assert pairs(*O.join_pairs(pk, bk, 0)) == inner
    assert pairs(*O.join_pairs(pk, bk, 1)) == sorted(inner + [(2, NO), (4, NO)])
    comp_b = [col([1, 1, 2, 2], I32), col([1.0, 2.0, 1.0, 2.0], F64)]
    comp_p = [col([1, 1, 2, 2], I32), col([2.0, 9.0, 1.0, 2.0], F64)]
    assert pairs(*O.join_pairs(comp_p, comp_b, 0)) == [(0, 1), (2, 2), (3, 3)]
    # catalog_checks_join_sort.cpp:34-70: probe {1,2,3} x build {1:10,11; 2:20} = 3 rows
    cat = O.join_pairs([col([1, 2, 3], I32)], [col([1, 1, 2], I32)], 0)
    assert pairs(*cat) == [(0, 0), (0, 1), (1, 2)]
    assert len(O.join_pairs([col([7, 9, 7], I32)], [col([7] * 3000, I32)], 0)[0]) == 6000
    nn = O.join_pairs([col([1, 0, 0], I32, nulls=(1, 2))], [col([1, 0, 2], I32, nulls=(1,))], 1)
    assert pairs(*nn) == [(0, 0), (1, NO), (2, NO)]
    assert pairs(*O.join_pairs([col([1, 2, 3], I32)], [col([], I32)], 1)) == [(0, NO), (1, NO), (2, NO)]


def test_join_mutants_are_caught():
    pk, bk = [col([1, 2, 3], I32)], [col([1, 1, 2], I32)]
    clean = pairs(*O.join_pairs(pk, bk, 0))
    assert pairs(*O.join_pairs(pk, bk, 0, mutant=1)) != clean
    assert pairs(*O.join_pairs(pk, bk, 0, mutant=2)) != clean
    cp = [col([1, 1], I32), col([2.0, 9.0], F64)]
    cb = [col([1, 1], I32), col([1.0, 2.0], F64)]
    assert pairs(*O.join_pairs(cp, cb, 0, mutant=3)) != pairs(*O.join_pairs(cp, cb, 0))


@pytest.mark.parametrize("case", range(15))
def test_join_vs_nested_loop(case):
    r = np.random.default_rng(400 + case)
    nb, np_ = int(r.integers(0, 300)), int(r.integers(0, 300))
    nk = 1 + case % 2
    bcols = [col(r.integers(0, 12, nb), I64, tuple(np.nonzero(r.random(nb) < 0.1)[0])) for _ in range(nk)]
    pcols = [col(r.integers(0, 12, np_), I64, tuple(np.nonzero(r.random(np_) < 0.1)[0])) for _ in range(nk)]
    jt = case % 2

    def key(cols, i):
        out = []
        for v, valid, _ in cols:
            if valid is not None and not valid[i]:
                return None
            out.append(int(v[i]))
        return tuple(out)
    want = []
    for i in range(np_):
        ki = key(pcols, i)
        m = [j for j in range(nb) if ki is not None and key(bcols, j) == ki]
        want += [(i, j) for j in m] if m else ([(i, NO)] if jt else [])
    assert pairs(*O.join_pairs(pcols, bcols, jt)) == sorted(want)


@needs_ref
@pytest.mark.parametrize("case", range(10))
def test_join_vs_compiled_reference(case):
    r = np.random.default_rng(500 + case)
    nb, np_ = int(r.integers(0, 3000)), int(r.integers(0, 5000))
    types = [I32, F64][: 1 + case % 2]

    def mk(n):
        out = []
        for t in types:
            v = r.integers(0, 300, n) if t == I32 else r.choice([0.0, -0.0, 1.5, float("nan")], n)
            out.append(col(v, t, tuple(np.nonzero(r.random(n) < 0.05)[0])))
        return out
    pc, bc = mk(np_), mk(nb)
    jt = case % 2
    keys = list(range(len(types)))
    got = pairs(*O.join_pairs(pc, bc, jt))
    want = pairs(*O.ref_join_pairs(pc, bc, keys, keys, jt, batch_size=int(r.choice([64, 2048]))))
    assert got == want


@needs_ref
@pytest.mark.parametrize("case", range(6))
def test_join_aggregate_vs_compiled_reference(case):
    r = np.random.default_rng(600 + case)
    nb, np_ = int(r.integers(1, 2000)), int(r.integers(0, 20000))
    pc = [col(r.integers(0, 2500, np_), I64), col(r.uniform(0, 1000, np_), F64)]
    bc = [col(r.permutation(2500)[:nb] if case % 2 else r.integers(0, 2500, nb), I64),
          col(r.integers(0, 100, nb), I32)]
    aggs = [(SUM, 1), (CS, 0), (AVG, 1), (MAX, 1)]
    got = O.join_aggregate(pc, bc, [0], [0], 0, [3], aggs, sum_mode=0)
    want = O.ref_join_aggregate(pc, bc, [0], [0], 0, [3], aggs)
    assert_same(got, want, rel=0, abs_=0)


# ---- as-of ---------------------------------------------------------------------------------
def asof_build():
    return [col([1, 1, 1, 2], I32), col([10, 20, 30, 15], TS), col([100, 200, 300, 150], I64)]


def asof_probe():
    return [col([1, 1, 1, 2, 3, 1], I32), col([5, 10, 25, 15, 99, 50], TS)]


def test_asof_literals():
    b, p = asof_build(), asof_probe()
    # (1,5) none; (1,10)->row0; (1,25)->row1; (2,15)->row3; (3,99) none; (1,50)->row2
    assert pairs(*O.asof_pairs(p, b, [0], [0], 1, 1, 0)) == [(1, 0), (2, 1), (3, 3), (5, 2)]
    assert pairs(*O.asof_pairs(p, b, [0], [0], 1, 1, 1)) == [(0, NO), (1, 0), (2, 1), (3, 3), (4, NO), (5, 2)]
    assert pairs(*O.asof_pairs(p, b, [0], [0], 1, 1, 0, tolerance=10)) == [(1, 0), (2, 1), (3, 3)]
    zb, zp = [col([10, 20, 30], TS)], [col([5, 25, 35], TS)]
    assert pairs(*O.asof_pairs(zp, zb, [], [], 0, 0, 1)) == [(0, NO), (1, 1), (2, 2)]
    nb = [col([1, 1, 2, 7], I32, nulls=(3,)), col([10, 20, 0, 5], TS, nulls=(2,))]
    npb = [col([1, 2, 2, 9], I32, nulls=(3,)), col([25, 30, 0, 50], TS, nulls=(2,))]
    assert pairs(*O.asof_pairs(npb, nb, [0], [0], 1, 1, 1)) == [(0, 1), (1, NO), (2, NO), (3, NO)]
    cb = [col([1, 1, 1], I32), col([7, 7, 9], I32), col([10, 20, 10], TS)]
    cp = [col([1, 1], I32), col([7, 9], I32), col([25, 25], TS)]
    assert pairs(*O.asof_pairs(cp, cb, [0, 1], [0, 1], 2, 2, 0)) == [(0, 1), (1, 2)]
    # catalog_checks_asof.cpp:38-92: exact boundary ties match (the `<=` rule).
    kb = [col([1, 1, 2], I32), col([10, 20, 15], I64)]
    kp = [col([1, 1, 2], I32), col([20, 25, 15], I64)]
    assert pairs(*O.asof_pairs(kp, kb, [0], [0], 1, 1, 0)) == [(0, 1), (1, 1), (2, 2)]
    for m in (1, 2):
        assert pairs(*O.asof_pairs(kp, kb, [0], [0], 1, 1, 0, mutant=m)) != [(0, 1), (1, 1), (2, 2)]
    assert pairs(*O.asof_pairs(cp, cb, [0, 1], [0, 1], 2, 2, 0, mutant=3)) != [(0, 1), (1, 2)]
    # LEFT: an unmatched probe row (key 3, and key 1 before the first tick)
    # carries NULL build columns; kLeftWrongNull (tsx/asof_mutants.cpp:233-239)
    # gives it build row 0 instead.
    lp = [col([1, 3, 1], I32), col([20, 20, 5], I64)]
    left = pairs(*O.asof_pairs(lp, kb, [0], [0], 1, 1, 1))
    assert left == sorted([(0, 1), (1, NO), (2, NO)])
    assert pairs(*O.asof_pairs(lp, kb, [0], [0], 1, 1, 1, mutant=4)) != left


def brute_asof(pk, pt, bk, bt):
    """O(n*m) scan; ts unique per key in the generated data (tsx/asof.h:28-39)."""
    out = []
    for i in range(len(pt)):
        best, bts = NO, None
        for j in range(len(bt)):
            if bk[j] == pk[i] and bt[j] <= pt[i] and (bts is None or bt[j] > bts):
                best, bts = j, bt[j]
        out.append(best)
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("case", range(12))
def test_asof_vs_bruteforce(case):
    r = np.random.default_rng(700 + case)
    nb, np_ = int(r.integers(0, 400)), int(r.integers(0, 400))
    bk = r.integers(0, 6, nb)
    bt = r.permutation(np.arange(nb) * 7 + 3)  # globally distinct build times
    pk = r.integers(0, 7, np_)
    pt = r.integers(0, nb * 7 + 10, np_)
    want = brute_asof(pk, pt, bk, bt)
    p, b = O.asof_pairs([col(pk, I32), col(pt, TS)], [col(bk, I32), col(bt, TS)], [0], [0], 1, 1, 1)
    got = np.full(np_, NO, dtype=np.uint32)
    got[p] = b
    assert (got == want).all()


@needs_ref
@pytest.mark.parametrize("case", range(8))
def test_asof_vs_compiled_reference(case):
    r = np.random.default_rng(800 + case)
    nb, np_ = int(r.integers(0, 4000)), int(r.integers(0, 4000))
    bk = r.integers(0, 30, nb)
    bt = r.permutation(nb) * 3
    bc = [col(bk, I32, tuple(np.nonzero(r.random(nb) < 0.03)[0])), col(bt, I64)]
    pc = [col(r.integers(0, 33, np_), I32), col(r.integers(-5, nb * 3 + 5, np_), I64,
                                                 tuple(np.nonzero(r.random(np_) < 0.03)[0]))]
    tol = None if case % 2 else int(r.integers(0, 50))
    jt = case % 3 == 0
    got = pairs(*O.asof_pairs(pc, bc, [0], [0], 1, 1, int(jt), tolerance=tol))
    want = pairs(*O.ref_asof_pairs(pc, bc, [0], [0], 1, 1, int(jt), tolerance=tol, batch_size=256))
    assert got == want


# ---- sort ----------------------------------------------------------------------------------
@pytest.mark.parametrize("case", range(6))
def test_sort_perm_vs_numpy_lexsort(case):
    r = np.random.default_rng(900 + case)
    n = int(r.integers(0, 5000))
    a = r.integers(-50, 50, n)
    b = r.integers(-10**12, 10**12, n)
    perm = O.sort_perm([col(a, I32), col(b, I64)], [0, 1])
    assert (perm == np.lexsort((b, a), axis=0).astype(np.uint32)).all()  # stable, a primary


# ---- committed golden fixtures ---------------------------------------------------------------
def test_golden_fixtures_match_oracle():
    files = sorted(GOLDEN.glob("*.json"))
    if not files:
        pytest.skip("no golden fixtures")
    for f in files:
        fx = json.loads(f.read_text())
        kind = fx["kind"]
        cols = lambda spec: [col(np.array(c["values"], dtype=O.NP[c["type"]]), c["type"],
                                 tuple(c.get("nulls", []))) for c in spec]
        if kind == "groupby":
            got = O.aggregate(cols(fx["cols"]), fx["keys"], [tuple(a) for a in fx["aggs"]], sum_mode=0)
            want = [tuple(r) for r in fx["expected"]]
            assert_same(got, want, rel=0, abs_=0), f.name
        elif kind == "join_groupby":
            got = O.join_aggregate(cols(fx["probe"]), cols(fx["build"]), fx["pkeys"], fx["bkeys"], 0,
                                   fx["gkeys"], [tuple(a) for a in fx["aggs"]], sum_mode=0)
            assert_same(got, [tuple(r) for r in fx["expected"]], rel=0, abs_=0), f.name
        elif kind == "asof":
            p, b = O.asof_pairs(cols(fx["probe"]), cols(fx["build"]), fx["pkeys"], fx["bkeys"], fx["ptime"],
                                fx["btime"], fx["join_type"])
            assert pairs(p, b) == sorted(tuple(x) for x in fx["expected_pairs"]), f.name
        elif kind == "hashtable":
            t = O.OracleHT(fx["types"], policy=fx["policy"], init_cap=fx["init_cap"], lf=fx["lf"])
            assert t.insert_or_find(cols(fx["cols"])).tolist() == fx["expected_ids"], f.name


# ---- filter ------------------------------------------------------------------------------------
LT, LE, GT, GE, EQ, NE = O.LT, O.LE, O.GT, O.GE, O.EQ, O.NE


def seq_table(n):
    """tests/ops_test.cpp:42-56: c0 = 0..n-1 (NULL where i % 7 == 0), c1 = 1.5 * c0."""
    i = np.arange(n)
    return [col(i, I32, tuple(np.nonzero(i % 7 == 0)[0])), col(i * 1.5, F64)]


def test_filter_literals():
    # tests/ops_test.cpp:99-124: c0 >= 100 over 300 rows keeps 100..299 minus the NULL rows.
    t = seq_table(300)
    got = O.filter_select(t, [(0, GE, I32, 100)])
    assert got.tolist() == [i for i in range(100, 300) if i % 7]
    assert len(O.filter_select(t, [(1, GE, F64, 0.0)])) == 300   # TRUE everywhere keeps all
    assert len(O.filter_select(t, [(1, LT, F64, 0.0)])) == 0     # FALSE everywhere keeps none
    # tests/ops_test.cpp:126-146 (c0 > 15 keeps {30, 50}), on the physical rows.
    assert O.filter_select([col([10, 20, 30, 40, 50], I32)], [(0, GT, I32, 15)]).tolist() == [1, 2, 3, 4]
    # null-if-any comparison (tests/expr_null_truthtable_test.cpp:74-107): NULL rows never pass.
    a = col([10, 20, 30, 40], I32, (1, 3))
    assert O.filter_select([a], [(0, GT, I32, 0)]).tolist() == [0, 2]
    assert O.filter_select([a], [(0, GT, I32, None)]).tolist() == []   # a NULL literal
    assert O.filter_select([a], [(0, NE, I32, None)]).tolist() == []
    # IEEE NaN (expr/expr.h:36-40): every comparison with NaN is FALSE except !=.
    x = col([np.nan, 1.0, -0.0, 0.0, np.inf], F64)
    for op, want in [(LT, [2, 3]), (LE, [1, 2, 3]), (GT, [4]), (GE, [1, 4]), (EQ, [1]), (NE, [0, 2, 3, 4])]:
        assert O.filter_select([x], [(0, op, F64, 1.0)]).tolist() == want, op
    assert O.filter_select([x], [(0, EQ, F64, 0.0)]).tolist() == [2, 3]         # -0 == +0
    assert O.filter_select([x], [(0, NE, F64, float("nan"))]).tolist() == [0, 1, 2, 3, 4]
    assert O.filter_select([x], [(0, EQ, F64, float("nan"))]).tolist() == []
    # promotion (expr/expr.cpp:133-151): I64 vs F64 literal compares as double;
    # 2^53 + 1 rounds to 2^53 as a double, so it equals the literal 2^53.
    big = col([2**53, 2**53 + 1, 2**53 + 2, -(2**63)], I64)
    assert O.filter_select([big], [(0, EQ, F64, float(2**53))]).tolist() == [0, 1]
    assert O.filter_select([big], [(0, EQ, I64, 2**53)]).tolist() == [0]
    assert O.filter_select([big], [(0, LT, F64, -9.2e18)]).tolist() == [3]
    # I32 vs F64 literal: 2 < 2.5 as a double (not 2 < 2 on truncated integers).
    assert O.filter_select([col([1, 2, 3], I32)], [(0, LT, F64, 2.5)]).tolist() == [0, 1]
    # BOOL compares BOOL, TS compares TS.
    assert O.filter_select([col([0, 1, 1, 0], BOOL)], [(0, EQ, BOOL, 1)]).tolist() == [1, 2]
    assert O.filter_select([col([5, -5, 7], TS)], [(0, LE, TS, 5)]).tolist() == [0, 1]


def test_filter_three_valued_and():
    # tests/expr_null_truthtable_test.cpp:1-72: AND is TRUE only when both sides
    # are TRUE; F AND N = F, N AND T = N, and a Filter keeps TRUE rows only.
    # Rows enumerate (left, right) over {F, T, N} x {F, T, N}.
    L = col([0, 0, 0, 1, 1, 1, 0, 0, 0], I32, (6, 7, 8))
    R = col([0, 1, 0, 0, 1, 0, 0, 1, 0], I32, (2, 5, 8))
    got = O.filter_select([L, R], [(0, EQ, I32, 1), (1, EQ, I32, 1)])
    assert got.tolist() == [4]


def _numpy_filter(cols, terms):
    """Independent model: numpy comparisons after promotion, ANDed with validity."""
    n = len(cols[0][0])
    keep = np.ones(n, dtype=bool)
    ops = [np.less, np.less_equal, np.greater, np.greater_equal, np.equal, np.not_equal]
    for c, op, lt, v in terms:
        vals, valid, t = cols[c]
        if v is None:
            return np.zeros(0, dtype=np.uint32)
        if t == F64 or lt == F64:
            res = ops[op](np.asarray(vals).astype(np.float64), np.float64(v))
        else:
            res = ops[op](np.asarray(vals).astype(np.int64), np.int64(v))
        keep &= res if valid is None else (res & valid)
    return np.nonzero(keep)[0].astype(np.uint32)


def _rand_filter_case(r, n):
    def nulls():
        return np.ones(n, dtype=bool) if r.random() < 0.3 else r.random(n) >= 0.1
    f = r.normal(0, 10, n)
    f[r.random(n) < 0.05] = np.nan
    cols = [(r.integers(-50, 50, n).astype(np.int32), nulls(), I32),
            (r.integers(-10**12, 10**12, n), nulls(), I64),
            (np.round(f, 1), nulls(), F64),
            (r.integers(0, 2, n).astype(np.uint8), nulls(), BOOL),
            (r.integers(0, 1000, n), nulls(), TS)]
    cols = [(v, None if m.all() else m, t) for v, m, t in cols]
    terms = []
    for _ in range(int(r.integers(1, 5))):
        c = int(r.integers(0, 5))
        op = int(r.integers(0, 6))
        t = cols[c][2]
        if t == I32:
            lt = [I32, I64, F64][int(r.integers(0, 3))]
            v = float(np.round(r.normal(0, 30), 1)) if lt == F64 else int(r.integers(-60, 60))
        elif t == I64:
            lt = [I32, I64, F64][int(r.integers(0, 3))]
            v = (float(r.integers(-10**12, 10**12)) if lt == F64 else
                 int(r.integers(-2**31, 2**31)) if lt == I32 else int(r.integers(-10**12, 10**12)))
        elif t == F64:
            lt = [I32, I64, F64][int(r.integers(0, 3))]
            v = float(np.round(r.normal(0, 10), 1)) if lt == F64 else int(r.integers(-20, 20))
            if lt == F64 and r.random() < 0.1:
                v = float("nan")
        elif t == BOOL:
            lt, v = BOOL, int(r.integers(0, 2))
        else:
            lt, v = TS, int(r.integers(0, 1000))
        if r.random() < 0.03:
            v = None
        terms.append((c, op, lt, v))
    return cols, terms


@pytest.mark.parametrize("case", range(12))
def test_filter_vs_numpy_model(case):
    r = np.random.default_rng(1200 + case)
    n = int(r.integers(0, 5000))
    cols, terms = _rand_filter_case(r, n)
    assert (O.filter_select(cols, terms) == _numpy_filter(cols, terms)).all()


@needs_ref
@pytest.mark.parametrize("case", range(12))
def test_filter_vs_compiled_reference(case):
    r = np.random.default_rng(1300 + case)
    n = int(r.integers(1, 5000))
    cols, terms = _rand_filter_case(r, n)
    if any(cols[c][2] == F64 and op == NE for c, op, _, _ in terms):
        # The reference's vector compare evaluates NaN != x as FALSE in full SIMD
        # lanes (Highway's ordered not-equal) but TRUE in the scalar tail and the
        # scalar twin (expr/compare_kernels.cpp:55 vs :34, compare_scalar.cpp:5);
        # the documented rule (expr/expr.h:36-40, IEEE) is pinned by
        # test_filter_literals, so NaN is kept out of != terms here.
        v, m, t = cols[2]
        cols[2] = (np.where(np.isnan(v), 0.5, v), m, t)
        terms = [(c, op, lt, 0.25 if (lt == F64 and v is not None and v != v) else v) for c, op, lt, v in terms]
    got = O.filter_select(cols, terms)
    want = O.ref_filter_select(cols, terms, batch_size=[64, 256, 2048][case % 3])
    assert got.tolist() == want.tolist()


def test_filter_mutants_are_caught():
    t = seq_table(300)
    clean = O.filter_select(t, [(0, GE, I32, 100)])
    assert not np.array_equal(O.filter_select(t, [(0, GE, I32, 100)], mutant=1), clean)   # NULL passes
    assert not np.array_equal(O.filter_select(t, [(0, GE, I32, 100)], mutant=2), clean)   # >= made strict
    assert not np.array_equal(O.filter_select(t, [(0, GE, I32, 100)], mutant=4), clean)   # inverted
    x = [col([1, 2, 3], I32)]
    assert O.filter_select(x, [(0, LT, F64, 2.5)], mutant=3).tolist() != [0, 1]          # no promotion


# ---- dictionary-encoded VARCHAR (STR) ----------------------------------------------------------
STR = O.STR
ALPHA = ["AAPL", "MSFT", "GOOG", "AMZN", "TSLA", "NVDA", "META", "x"]  # wp7b_strings_differential_test.cpp:69-72


def str_col(vals, nulls_at=(), dict_order=None):
    """A STR column from values; each column gets its own dictionary (in
    `dict_order`, default first-seen), so equal strings carry different codes
    across columns (the cross-dictionary case)."""
    d = list(dict_order) if dict_order is not None else list(dict.fromkeys(v for v in vals if v is not None))
    idx = {s: i for i, s in enumerate(d)}
    codes = np.array([idx[v] if v is not None else 0 for v in vals], dtype=np.int32)
    valid = None
    if nulls_at or any(v is None for v in vals):
        valid = np.array([v is not None for v in vals], dtype=bool)
        valid[list(nulls_at)] = False
    return (codes, valid, STR, d)


def rand_str_col(r, n, null_pct):
    """gen_string_column's shape: values from ALPHA, NULLs, a shuffled dictionary."""
    d = list(r.permutation(ALPHA))
    vals = [None if r.random() * 100 < null_pct else str(r.choice(ALPHA)) for _ in range(n)]
    return str_col(vals, dict_order=d)


def test_str_join_literals_cross_dictionary():
    # wp7b_strings_differential_test.cpp:249-268: build "AAPL" is code 0, probe's is code 1.
    b = [str_col(["AAPL", "MSFT"]), col([10, 20], I64)]
    p = [str_col(["ZZZ", "AAPL"])]
    assert pairs(*O.join_pairs([p[0]], [b[0]])) == [(1, 0)]
    # catalog_checks_strings.cpp:34-70: the same values interned in other orders.
    b = [str_col(["AAPL", "MSFT", "GOOG"])]
    p = [str_col(["GOOG", "MSFT", "AAPL", "ZZZ"])]
    assert pairs(*O.join_pairs(p, b)) == [(0, 2), (1, 1), (2, 0)]
    # the by-code mutant (string_key_mutants kHashRawCode) pairs codes, not values
    assert pairs(*O.join_pairs(p, b, mutant=4)) != [(0, 2), (1, 1), (2, 0)]
    # NULL STR keys never join
    b = [str_col(["A", None, "B"])]
    p = [str_col([None, "B", "A"])]
    assert pairs(*O.join_pairs(p, b)) == [(1, 2), (2, 0)]


def _py_str_groupby(vals, payload):
    """Independent model: dict keyed by the string value (None groups)."""
    out = {}
    for k, v in zip(vals, payload):
        g = out.setdefault(k, [0, 0, None, None])
        g[0] += 1
        if v is not None:
            g[1] += 1
            g[2] = v if g[2] is None or v < g[2] else g[2]
            g[3] = v if g[3] is None or v > g[3] else g[3]
    return [(k, *g) for k, g in out.items()]


def _strs(c):
    codes, valid, _, d = c
    return [None if (valid is not None and not valid[i]) else d[codes[i]] for i in range(len(codes))]


@pytest.mark.parametrize("case", range(8))
def test_str_groupby_vs_python_model(case):
    r = np.random.default_rng(1500 + case)
    n = int(r.integers(1, 1200))
    k, v = rand_str_col(r, n, 15), rand_str_col(r, n, 15)
    got = O.aggregate([k, v], [0], [(CS, 0), (CNT, 1), (MIN, 1), (MAX, 1)])
    assert canon(got) == canon(_py_str_groupby(_strs(k), _strs(v)))


def test_str_groupby_duplicate_dictionary_values_and_mutant():
    # A dictionary holding one value under two codes still groups by value.
    c = (np.array([0, 1, 2, 0, 2], dtype=np.int32), None, STR, ["A", "B", "A"])
    got = O.aggregate([c], [0], [(CS, 0)])
    assert canon(got) == [("A", 4), ("B", 1)]
    assert canon(O.aggregate([c], [0], [(CS, 0)], mutant=5)) != [("A", 4), ("B", 1)]


@needs_ref
@pytest.mark.parametrize("case", range(10))
def test_str_groupby_vs_compiled_reference(case):
    # wp7b_strings_differential_test.cpp:132-165: GROUP BY a STR key with
    # COUNT/MIN/MAX over a STR payload; composite STR + I64 keys.
    r = np.random.default_rng(1600 + case)
    n = int(r.integers(1, 1200))
    cols = [rand_str_col(r, n, 15), rand_str_col(r, n, 15), col(r.integers(0, 1000, n), I64)]
    if case % 2:
        keys, aggs = [0, 2], [(CS, 0), (MAX, 0)]
    else:
        keys, aggs = [0], [(CS, 0), (CNT, 1), (MIN, 1), (MAX, 1)]
    got = O.aggregate(cols, keys, aggs)
    want = O.ref_aggregate(cols, keys, aggs, batch_size=[64, 256, 2048][case % 3])
    assert diff(got, want, rel=0, abs_=0) is None


@needs_ref
@pytest.mark.parametrize("case", range(10))
def test_str_join_vs_compiled_reference(case):
    # wp7b_strings_differential_test.cpp:167-200: independent dictionaries on
    # each side, inner and left.
    r = np.random.default_rng(1700 + case)
    nb, np_ = int(r.integers(0, 300)), int(r.integers(1, 300))
    b = [rand_str_col(r, nb, 12), col(np.arange(nb), I64)]
    p = [rand_str_col(r, np_, 12), col(1000 + np.arange(np_), I64)]
    jt = case % 2
    got = pairs(*O.join_pairs([p[0]], [b[0]], jt))
    want = pairs(*O.ref_join_pairs(p, b, [0], [0], jt, batch_size=[64, 2048][case % 2]))
    assert got == want


@needs_ref
def test_str_join_aggregate_vs_compiled_reference():
    # The tick pipeline's tail: trades join a symbol table on an I32 id, grouped
    # by the symbol's dictionary-encoded name.
    r = np.random.default_rng(1800)
    nsym = 300
    names = [f"S{i:03d}{'X' * (i % 7)}" for i in range(nsym)]
    order = list(r.permutation(nsym))
    sym_tab = [col(np.arange(nsym), I32), str_col([names[i] for i in order], dict_order=list(r.permutation(names)))]
    n = 20000
    trades = [col(r.integers(0, nsym + 20, n), I32), col(np.round(r.uniform(0, 200, n), 2), F64)]
    aggs = [(CS, 0), (AVG, 1), (MIN, 3), (MAX, 3)]
    got = O.join_aggregate(trades, sym_tab, [0], [0], 0, [3], aggs, sum_mode=0)
    want = O.ref_join_aggregate(trades, sym_tab, [0], [0], 0, [3], aggs)
    assert diff(got, want, rel=1e-12, abs_=0) is None


# ---- sort: comparison path (F64, STR, BOOL and mixed keys) --------------------------------------
def _rand_sort_cols(r, n):
    f = np.round(r.normal(size=n), 1)
    f[r.random(n) < 0.05] = np.nan
    f[r.random(n) < 0.05] = -0.0
    f[r.random(n) < 0.05] = 0.0
    return [col(f, F64, tuple(np.nonzero(r.random(n) < 0.05)[0])),
            rand_str_col(r, n, 8),
            col(r.integers(0, 2, n), BOOL, tuple(np.nonzero(r.random(n) < 0.05)[0])),
            col(r.integers(-3, 3, n), I64)]


@needs_ref
@pytest.mark.parametrize("case", range(12))
def test_sort_comparison_path_vs_compiled_reference(case):
    # wp7b_strings_differential_test.cpp:217-246 (STR ORDER BY, desc and NULL
    # placement) and mixed F64 / BOOL / integer keys through the reference Sort.
    r = np.random.default_rng(1900 + case)
    n = int(r.integers(1, 3000))
    cols = _rand_sort_cols(r, n)
    nk = 1 + case % 3
    keys = [int(k) for k in r.permutation(4)[:nk]]
    desc = [int(r.integers(0, 2)) for _ in range(nk)]
    nf = [int(r.integers(0, 2)) for _ in range(nk)]
    got = O.sort_perm(cols, keys, desc, nf)
    want = O.ref_sort_perm(cols, keys, desc, nf, batch_size=[64, 256, 2048][case % 3])
    assert got.tolist() == want.tolist()


def test_sort_f64_total_order_literals():
    # ops/sort_internal.h:100-104: NaN greatest and tied, -0 == +0 (stable)
    c = col([1.0, np.nan, -0.0, 0.0, -np.inf, np.nan, 0.0], F64)
    assert O.sort_perm([c], [0]).tolist() == [4, 2, 3, 6, 0, 1, 5]
    assert O.sort_perm([c], [0], [1]).tolist() == [1, 5, 0, 2, 3, 6, 4]
    s = str_col(["b", None, "a", "B", "b"])
    assert O.sort_perm([s], [0]).tolist() == [3, 2, 0, 4, 1]       # byte order: "B" < "a" < "b"
    assert O.sort_perm([s], [0], [1], [1]).tolist() == [1, 0, 4, 2, 3]


# ---- the morsel-parallel CPU baseline (oracle/qe_baseline.cpp) ------------------------------
def test_morsel_baseline_equals_oracle(rng):
    """The bench's CPU baseline (exec/parallel.cpp:431-599 restated: atomic
    morsel cursor, per-worker partials, merge after join) gives the oracle's
    groups, COUNTs exactly and F64 sums within 1e-12, for several thread
    counts and morsel sizes; the as-of leg equals numpy searchsorted."""
    L = O.oracle()
    # the 1M-row group-by's shape (scaled): I64 key in [0, 1024), F64 values with NULLs
    n = 50_000
    key = rng.integers(0, 1024, n).astype(np.int64)
    val = rng.uniform(0, 1000, n)
    valid = rng.random(n) > 0.05
    words = np.zeros((n + 63) // 64, dtype=np.uint64)
    for i in np.flatnonzero(valid):
        words[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    want = {r[0]: r[1:] for r in O.aggregate([(key, None, I64), (val, valid, F64)], [0],
                                             [(CS, 0), (SUM, 1), (CNT, 1)])}
    for threads, morsel in ((1, 1024), (4, 1024), (7, 333)):
        ok, oc, osum, onn = (np.empty(2048, np.int64), np.empty(2048, np.int64), np.empty(2048), np.empty(2048, np.int64))
        g = L.orc_bl_groupby_i64_f64(key.ctypes.data, val.ctypes.data, words.ctypes.data, n, threads, morsel, 2048,
                                     ok.ctypes.data, oc.ctypes.data, osum.ctypes.data, onn.ctypes.data)
        assert g == len(want)
        for i in range(g):
            c, s, nn = want[int(ok[i])]
            assert oc[i] == c and onn[i] == nn
            assert abs(osum[i] - s) <= 1e-12 * abs(s) + 1e-12
    # the join -> group-by workload's shape (scaled): duplicate build keys fan out, 10% probe misses
    nb, npr = 3000, 40_000
    bkey = rng.permutation(nb).astype(np.int64) % 2500  # 500 duplicated keys
    bgrp = rng.integers(0, 100, nb).astype(np.int32)
    pkey = np.where(rng.random(npr) < 0.9, rng.integers(0, 2500, npr), rng.integers(5000, 6000, npr)).astype(np.int64)
    ppay = rng.uniform(0, 1000, npr)
    wj = {r[0]: r[1:] for r in O.join_aggregate([(pkey, None, I64), (ppay, None, F64)],
                                                [(bkey, None, I64), (bgrp, None, I32)], [0], [0], 0, [3],
                                                [(SUM, 1), (CS, 0)])}
    for threads in (1, 5):
        og, osum, ocnt = np.empty(256, np.int32), np.empty(256), np.empty(256, np.int64)
        g = L.orc_bl_join_groupby(pkey.ctypes.data, ppay.ctypes.data, npr, bkey.ctypes.data, bgrp.ctypes.data, nb,
                                  threads, 1024, 256, og.ctypes.data, osum.ctypes.data, ocnt.ctypes.data)
        assert g == len(wj)
        for i in range(g):
            s, c = wj[int(og[i])]
            assert ocnt[i] == c
            assert abs(osum[i] - s) <= 1e-12 * abs(s) + 1e-12
    # as-of leg: upper_bound per probe over CSR segments of a sorted right side
    nsym = 50
    sym = np.sort(rng.integers(0, nsym, 5000))
    ts = np.concatenate([np.sort(rng.choice(10**6, np.count_nonzero(sym == s), replace=False)) for s in range(nsym)])
    ts = ts.astype(np.int64)
    seg = np.searchsorted(sym, np.arange(nsym + 1)).astype(np.uint64)
    psym = rng.integers(-1, nsym + 1, 20_000).astype(np.int32)
    pts = rng.integers(-10, 10**6 + 10, 20_000).astype(np.int64)
    out = np.empty(20_000, np.uint32)
    L.orc_bl_asof_upper_bound(ts.ctypes.data, seg.ctypes.data, nsym, psym.ctypes.data, pts.ctypes.data, 20_000, 3,
                              out.ctypes.data)
    for i in range(20_000):
        s = int(psym[i])
        if s < 0 or s >= nsym:
            assert out[i] == NO
            continue
        b, e = int(seg[s]), int(seg[s + 1])
        u = b + int(np.searchsorted(ts[b:e], pts[i], side="right"))
        assert out[i] == (NO if u == b else u - 1)
