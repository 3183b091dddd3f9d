"""Dictionary-encoded VARCHAR (STR) on the GPU against the CPU oracle.

Cases follow the reference's string suites
(tests/wp7b_strings_differential_test.cpp: GROUP BY a STR key with
COUNT/MIN/MAX over a STR payload, composite STR + I64 keys, inner and left
joins on STR keys whose sides have independent dictionaries, the explicit
cross-dictionary match; tests/catalog_checks_strings.cpp: the by-code mutant's
case) plus the tick pipeline's tail (trades joined to a symbol table on an I32
id, grouped by the dictionary-encoded name), a dictionary holding a value under
two codes, STR and F64 sort keys, STR as-of keys, and dictionaries carried by
filter / gather / slice outputs. STR values compare as strings; everything is
exact except F64 AVG (1e-12 relative).
"""
import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, canon
from test_oracle_golden import ALPHA, rand_str_col, str_col

pytestmark = pytest.mark.gpu

I32, I64, F64, STR = O.I32, O.I64, O.F64, O.STR
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


class Uploaded:
    """Oracle-style columns on the device (STR columns with their dictionary)."""

    def __init__(self, ctx, cols):
        self.ctx, self.dev, self.dicts = ctx, [], []
        for c in cols:
            d = None
            if c[2] == STR:
                d = ctx.dict(c[3])
                self.dicts.append(d)
            self.dev.append(ctx.upload(c[0], c[2], c[1], dict=d))

    def free(self):
        for x in self.dev:
            x.free()
        for d in self.dicts:
            d.free()


def table_rows(t):
    cols = []
    for i in range(t.ncols):
        c = t.col(i)
        if c.type == STR:
            cols.append(c.strings())
        else:
            v, valid = c.numpy()
            lst = v.tolist()
            cols.append(lst if valid is None else [x if ok else None for x, ok in zip(lst, valid.tolist())])
    return [tuple(c[r] for c in cols) for r in range(t.rows)]


def gpu_agg(ctx, cols, keys, aggs):
    u = Uploaded(ctx, cols)
    try:
        t = ctx.aggregate(u.dev, keys, aggs)
        try:
            return table_rows(t)
        finally:
            t.free()
    finally:
        u.free()


@pytest.mark.parametrize("case", range(12))
def test_str_groupby_min_max_count(ctx, case):
    r = np.random.default_rng(2100 + case)
    n = int(r.integers(1, 3000)) if case < 10 else 200_000
    cols = [rand_str_col(r, n, 15), rand_str_col(r, n, 15), col(r.integers(0, 1000, n), I64)]
    if case % 2:
        keys, aggs = [0, 2], [(CS, 0), (MAX, 0), (MIN, 1)]
    else:
        keys, aggs = [0], [(CS, 0), (CNT, 1), (MIN, 1), (MAX, 1)]
    assert_same(gpu_agg(ctx, cols, keys, aggs), O.aggregate(cols, keys, aggs), rel=0, abs_=0)


def test_str_global_min_max_and_null_group(ctx):
    c = str_col(["pear", None, "apple", "fig", "apple", "zucchini", None])
    got = gpu_agg(ctx, [c], [], [(MIN, 0), (MAX, 0), (CNT, 0), (CS, 0)])
    assert got == [("apple", "zucchini", 5, 7)]
    # one group per value, NULL keys grouped together (row order is not
    # contractual: the reference's comparator canonicalises unordered results)
    got = gpu_agg(ctx, [c], [0], [(CS, 0)])
    assert canon(got) == canon([("pear", 1), (None, 2), ("apple", 2), ("fig", 1), ("zucchini", 1)])


def test_str_dictionary_with_repeated_values(ctx):
    # one value under codes 0 and 2: grouping, MIN/MAX and joins use the value
    c = (np.array([0, 1, 2, 0, 2, 1], dtype=np.int32), None, STR, ["A", "B", "A"])
    assert canon(gpu_agg(ctx, [c], [0], [(CS, 0), (MIN, 0)])) == [("A", 4, "A"), ("B", 2, "B")]
    u = Uploaded(ctx, [c, str_col(["B", "A"])])
    try:
        p, b = ctx.hash_join([u.dev[1]], [u.dev[0]]).numpy()
        assert sorted(zip(p.tolist(), b.tolist())) == [(0, 1), (0, 5), (1, 0), (1, 2), (1, 3), (1, 4)]
    finally:
        u.free()


def test_cross_dictionary_join_literals(ctx):
    cases = [([str_col(["ZZZ", "AAPL"])], [str_col(["AAPL", "MSFT"])], [(1, 0)]),
             ([str_col(["GOOG", "MSFT", "AAPL", "ZZZ"])], [str_col(["AAPL", "MSFT", "GOOG"])],
              [(0, 2), (1, 1), (2, 0)]),
             ([str_col([None, "B", "A"])], [str_col(["A", None, "B"])], [(1, 2), (2, 0)])]
    for pc, bc, want in cases:
        u = Uploaded(ctx, pc + bc)
        try:
            pairs = ctx.hash_join([u.dev[0]], [u.dev[1]])
            p, b = pairs.numpy()
            pairs.free()
            assert sorted(zip(p.tolist(), b.tolist())) == want
            assert sorted(zip(*[x.tolist() for x in O.join_pairs(pc, bc)])) == want
        finally:
            u.free()


@pytest.mark.parametrize("case", range(10))
def test_str_join_pairs_random(ctx, case):
    r = np.random.default_rng(2200 + case)
    nb, np_ = int(r.integers(0, 2000)), int(r.integers(1, 2000))
    bc = rand_str_col(r, nb, 12) if nb else str_col([], dict_order=ALPHA)
    pc = rand_str_col(r, np_, 12)
    if case % 3 == 2:  # both sides on one dictionary
        pc = (pc[0], pc[1], STR, bc[3])
        pc = (np.minimum(pc[0], max(len(bc[3]) - 1, 0)).astype(np.int32), pc[1], STR, bc[3])
    jt = case % 2
    u = Uploaded(ctx, [pc, bc]) if case % 3 != 2 else None
    if u is None:  # one shared Dict object on both sides
        d = ctx.dict(bc[3])
        dev = [ctx.upload(pc[0], STR, pc[1], dict=d), ctx.upload(bc[0], STR, bc[1], dict=d)]
    else:
        dev = u.dev
    try:
        pairs = ctx.hash_join([dev[0]], [dev[1]], jt)
        p, b = pairs.numpy()
        pairs.free()
        wp, wb = O.join_pairs([pc], [bc], jt)
        assert sorted(zip(p.tolist(), b.tolist())) == sorted(zip(wp.tolist(), wb.tolist()))
    finally:
        if u is None:
            for x in dev:
                x.free()
            d.free()
        else:
            u.free()


@pytest.mark.parametrize("case", range(6))
def test_join_aggregate_group_by_name(ctx, case):
    # trades (sym id, px) join a symbol table (id, name) on the id; GROUP BY the
    # dictionary-encoded name: COUNT(*), AVG(px), MIN/MAX(name), COUNT(name)
    r = np.random.default_rng(2300 + case)
    nsym = int(r.choice([50, 3000]))
    names = [f"S{i:04d}{'X' * (i % 9)}" for i in range(nsym)]
    dup = case == 5  # two symbols share one name: they form one group
    if dup:
        names[1] = names[0]
    order = r.permutation(nsym)
    d = list(dict.fromkeys(r.permutation(names).tolist())) if not dup else list(r.permutation(names))
    idx = {s: i for i, s in enumerate(d)}
    name_codes = np.array([idx[names[i]] for i in order], dtype=np.int32)
    name_valid = None if case % 2 == 0 else (r.random(nsym) >= 0.1)
    sym_tab = [col(order, I32), (name_codes, name_valid, STR, d)]
    n = int(r.choice([10_000, 1_000_000]))
    trades = [col(r.integers(0, nsym + nsym // 10, n), I32), col(np.round(r.uniform(0, 200, n), 2), F64)]
    aggs = [(CS, 0), (AVG, 1), (MIN, 3), (MAX, 3), (CNT, 3)]
    u = Uploaded(ctx, trades + sym_tab)
    try:
        t = ctx.join_aggregate(u.dev[:2], u.dev[2:], [0], [0], 0, [3], aggs)
        try:
            got = table_rows(t)
        finally:
            t.free()
    finally:
        u.free()
    assert_same(got, O.join_aggregate(trades, sym_tab, [0], [0], 0, [3], aggs))


def test_join_aggregate_on_str_keys_cross_dictionary(ctx):
    # join on a STR key (independent dictionaries) and group by the probe's STR column
    r = np.random.default_rng(2400)
    n, m = 50_000, 400
    probe = [rand_str_col(r, n, 5), rand_str_col(r, n, 5), col(r.integers(0, 100, n), I64)]
    build = [rand_str_col(r, m, 5), col(np.arange(m), I64)]
    aggs = [(CS, 0), (SUM, 2), (MIN, 1), (MAX, 0)]
    for jt in (0, 1):
        u = Uploaded(ctx, probe + build)
        try:
            t = ctx.join_aggregate(u.dev[:3], u.dev[3:], [0], [0], jt, [1], aggs)
            try:
                got = table_rows(t)
            finally:
                t.free()
        finally:
            u.free()
        assert_same(got, O.join_aggregate(probe, build, [0], [0], jt, [1], aggs), rel=0, abs_=0)


def test_outputs_keep_their_dictionary(ctx):
    c = str_col(["b", "a", None, "c", "a"] * 100)
    u = Uploaded(ctx, [c, col(np.arange(500), I64)])
    try:
        t = ctx.filter(u.dev, [(1, O.LT, I64, 7)])
        assert t.col(0).strings() == ["b", "a", None, "c", "a", "b", "a"]
        t.free()
        pairs = ctx.hash_join([u.dev[0]], [u.dev[0]])
        g = pairs.gather([u.dev[0]], [u.dev[1], u.dev[0]])
        p, b = pairs.numpy()
        strs = [None if (c[1] is not None and not c[1][i]) else c[3][c[0][i]] for i in range(500)]
        assert g.col(0).strings() == [strs[i] for i in p.tolist()]
        assert g.col(2).strings() == [strs[i] for i in b.tolist()]
        g.free()
        pairs.free()
        sl = ctx.slice(u.dev[0], 64, 10)
        assert sl.strings() == strs[64:74]
        sl.free()
    finally:
        u.free()


def test_sort_by_str_and_f64_keys(ctx):
    r = np.random.default_rng(2500)
    n = 5000
    s = rand_str_col(r, n, 10)
    f = r.normal(size=n)
    f[r.random(n) < 0.05] = np.nan
    f[r.random(n) < 0.05] = 0.0
    f[r.random(n) < 0.05] = -0.0
    fvalid = r.random(n) >= 0.05
    u = Uploaded(ctx, [s, (f, fvalid, F64)])
    try:
        strs = [None if (s[1] is not None and not s[1][i]) else s[3][s[0][i]] for i in range(n)]
        for desc in (0, 1):
            for nf in (0, 1):
                perm = ctx.sort_perm([u.dev[0]], [desc], [nf]).tolist()
                nulls = [i for i in range(n) if strs[i] is None]
                # stable either way: equal values keep their input order
                vals = sorted((i for i in range(n) if strs[i] is not None), key=lambda i: strs[i],
                              reverse=bool(desc))
                want = nulls + vals if nf else vals + nulls
                assert perm == want, (desc, nf)
        # F64: NaN greatest (all NaNs tie), -0 == +0, NULLs placed absolutely
        fv = [None if not fvalid[i] else float(f[i]) for i in range(n)]

        def fkey(i):
            x = fv[i]
            return (1, 0.0) if np.isnan(x) else (0, x + 0.0)
        for desc in (0, 1):
            perm = ctx.sort_perm([u.dev[1]], [desc], [0]).tolist()
            vals = sorted((i for i in range(n) if fv[i] is not None), key=fkey, reverse=bool(desc))
            assert perm == vals + [i for i in range(n) if fv[i] is None], desc
    finally:
        u.free()


def test_asof_on_str_keys_matches_id_keys(ctx):
    # The same as-of over STR keys (build and probe on different dictionaries)
    # and over I32 value ids must agree.
    r = np.random.default_rng(2600)
    nb, np_ = 20000, 5000
    vocab = [f"K{i}" for i in range(40)]
    bvals = [vocab[i] for i in r.integers(0, 40, nb)]
    pvals = [vocab[i] if i < 40 else "missing" for i in r.integers(0, 44, np_)]
    bts = r.integers(0, 10**6, nb)
    pts = r.integers(0, 10**6, np_)
    bd, pd = ctx.dict(list(r.permutation(vocab))), ctx.dict(list(r.permutation(vocab + ["missing"])))
    bs, ps = ctx.upload_strings(bvals, bd), ctx.upload_strings(pvals, pd)
    vid = {v: i for i, v in enumerate(vocab + ["missing"])}
    bi = ctx.upload(np.array([vid[v] for v in bvals], np.int32), I32)
    pi = ctx.upload(np.array([vid[v] for v in pvals], np.int32), I32)
    bt, pt = ctx.upload(bts, I64), ctx.upload(pts, I64)
    try:
        a = ctx.asof_build([bs], bt)
        b = ctx.asof_build([bi], bt)
        try:
            assert (a.probe([ps], pt) == b.probe([pi], pt)).all()
        finally:
            a.free()
            b.free()
    finally:
        for x in (bs, ps, bi, pi, bt, pt):
            x.free()
        bd.free()
        pd.free()


def test_str_errors(ctx, qe):
    c = ctx.upload(np.array([0, 1, 5], np.int32), STR)  # no dictionary
    d = ctx.dict(["a", "b"])
    try:
        with pytest.raises(qe.InvalidArgument):
            ctx.aggregate([c], [0], [(CS, 0)])
        c.set_dict(d)
        with pytest.raises(qe.InvalidArgument):  # code 5 is outside the dictionary
            ctx.aggregate([c], [], [(MIN, 0)])
        with pytest.raises(qe.InvalidArgument):  # SUM over STR (ops/aggregate.cpp:84-90)
            ctx.aggregate([c], [], [(SUM, 0)])
        with pytest.raises(qe.InvalidArgument):
            ctx.remap(c, [0, 1], I32)
        i = ctx.upload(np.array([1, 2], np.int32), I32)
        with pytest.raises(qe.InvalidArgument):  # only STR columns carry a dictionary
            i.set_dict(d)
        i.free()
    finally:
        c.free()
        d.free()
