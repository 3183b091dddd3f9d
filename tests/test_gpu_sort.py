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
