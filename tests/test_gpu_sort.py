"""Device stable sort permutation (qe_gpu_sort_perm) against the oracle's
restatement of Sort's radix path (ops/sort.cpp:125-188), which is itself
pinned to numpy's stable lexsort: multi-key, ASC/DESC, NULLS FIRST/LAST,
duplicate keys (stability), negative values, 1..several digit passes."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu
I32, I64, TS = O.I32, O.I64, O.TS


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if len(nulls):
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


@pytest.mark.parametrize("case", range(12))
def test_sort_perm_matches_oracle(ctx, case):
    r = np.random.default_rng(6000 + case)
    n = int(r.choice([1, 63, 4096, 4097, 50000, 300000]))
    nk = 1 + case % 3
    cols = []
    for j in range(nk):
        t = [I32, I64, TS][(case + j) % 3]
        span = int(r.choice([3, 1000, 2**40]))
        lo = -span if t != I32 else -min(span, 2**30)
        hi = span if t != I32 else min(span, 2**30)
        nulls = tuple(np.nonzero(r.random(n) < 0.05)[0]) if case % 2 else ()
        cols.append(col(r.integers(lo, hi, n), t, nulls))
    desc = [int(r.integers(0, 2)) for _ in range(nk)]
    nf = [int(r.integers(0, 2)) for _ in range(nk)]
    d = [ctx.upload(v, t, valid) for v, valid, t in cols]
    try:
        got = ctx.sort_perm(d, desc, nf)
    finally:
        for x in d:
            x.free()
    want = O.sort_perm(cols, list(range(nk)), desc, nf)
    assert (got == want).all()


F64, BOOL, STR = O.F64, O.BOOL, O.STR


@pytest.mark.parametrize("case", range(10))
def test_sort_perm_f64_str_bool_mixed(ctx, case):
    # The comparison path's key types (ops/sort.cpp:60-100): F64 in the
    # NaN-greatest total order with -0 == +0, STR by value, BOOL as 0/1, mixed
    # with integer keys; the oracle's comparison path is pinned to the
    # compiled reference Sort.
    from test_oracle_golden import _rand_sort_cols
    r = np.random.default_rng(6100 + case)
    n = int(r.choice([1, 100, 4097, 60000, 250000]))
    cols = _rand_sort_cols(r, n)
    nk = 1 + case % 3
    keys = [int(k) for k in r.permutation(4)[:nk]]
    desc = [int(r.integers(0, 2)) for _ in range(nk)]
    nf = [int(r.integers(0, 2)) for _ in range(nk)]
    dicts = []
    dev = []
    for c in cols:
        d = None
        if c[2] == STR:
            d = ctx.dict(c[3])
            dicts.append(d)
        dev.append(ctx.upload(c[0], c[2], c[1], dict=d))
    try:
        got = ctx.sort_perm([dev[k] for k in keys], desc, nf)
    finally:
        for x in dev:
            x.free()
        for d in dicts:
            d.free()
    want = O.sort_perm(cols, keys, desc, nf)
    assert (got == want).all()


def test_take_gathers_rows_and_rejects_bad_indices(ctx, qe):
    r = np.random.default_rng(6200)
    n = 10000
    v = r.integers(-100, 100, n)
    valid = r.random(n) >= 0.1
    a = ctx.upload(v, I64, valid)
    idx = np.concatenate([r.permutation(n)[:5000], r.integers(0, n, 3000)]).astype(np.uint32)
    buf = qe.DevBuffer(ctx, 4 * len(idx))
    try:
        buf.upload(idx)
        t = ctx.take([a], buf, len(idx))
        gv, gvalid = t.col(0).numpy()
        t.free()
        assert (gvalid == valid[idx]).all() and (gv[gvalid] == v[idx][valid[idx]]).all()
        bad = idx.copy()
        bad[17] = n
        buf.upload(bad)
        with pytest.raises(qe.InvalidArgument):
            ctx.take([a], buf, len(idx))
    finally:
        buf.free()
        a.free()


@pytest.mark.parametrize("m", [0, 1, 63, 64, 65, 4097])
def test_take_every_width_one_null_count_read_back(ctx, qe, m):
    # BOOL, I32, F64 and TS columns gathered together (one NULL-count read-back
    # for all of them): values and validity per row, all-valid sources give
    # all-valid outputs, a nullable source whose gathered rows are all valid
    # drops its bitmap too.
    r = np.random.default_rng(6300 + m)
    n = 5000
    cols = [(r.integers(0, 2, n).astype(np.uint8), O.BOOL, r.random(n) >= 0.2),
            (r.integers(-2**31, 2**31, n).astype(np.int32), I32, None),
            (r.normal(size=n), O.F64, r.random(n) >= 0.3),
            (r.integers(-2**60, 2**60, n), TS, np.ones(n, dtype=bool))]
    dev = [ctx.upload(v, t, valid) for v, t, valid in cols]
    idx = r.integers(0, n, m).astype(np.uint32)
    buf = qe.DevBuffer(ctx, 4 * max(m, 1))
    try:
        if m:
            buf.upload(idx)
        t = ctx.take(dev, buf, m)
        assert t.rows == m
        for c, (v, typ, valid) in enumerate(cols):
            gv, gvalid = t.col(c).numpy()
            gvalid = np.ones(m, dtype=bool) if gvalid is None else gvalid
            want_valid = np.ones(m, dtype=bool) if valid is None else valid[idx]
            assert (gvalid == want_valid).all(), c
            assert (gv[want_valid] == v[idx][want_valid]).all(), c
            if want_valid.all():
                assert t.col(c).all_valid, c
        t.free()
    finally:
        buf.free()
        for d in dev:
            d.free()
