"""qe_gpu_filter_select_noalloc: the selection entry point that allocates
nothing, checked against qe_gpu_filter_select (bit-exact indices and counts).

Two calls in a row share the GPU's static scratch array: the first at the
maximum size (2^27 rows), the second smaller, so stale bitmap words or tile
offsets from the first call would show in the second. A device memory limit
of one byte makes any allocation through the library fail, so the no-alloc
call passing under it shows that it allocates nothing, while the allocating
entry point fails under the same limit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _select_both(ctx, qe, cols, terms):
    want = ctx.filter_select(cols, terms)
    out = qe.DevBuffer(ctx, 4 * max(cols[0].len, 1))
    try:
        n = ctx.filter_select_noalloc_into(cols, terms, out)
        got = out.numpy(np.uint32, n)
    finally:
        out.free()
    return got, want


def test_noalloc_matches_filter_select_max_then_smaller(ctx, qe):
    # Call 1: 2^27 rows, I32 in [0, 1000) and F64 in [0, 200) with 2% NULLs,
    # two terms (about 30% pass).
    n1 = 1 << 27
    a = ctx.gen_uniform(41, 0, qe.I32, n1, 0, 1000)
    b = ctx.gen_uniform(41, 1, qe.F64, n1, 0.0, 200.0, null_frac=0.02)
    try:
        terms1 = [(0, qe.CMP_LT, qe.I32, 600), (1, qe.CMP_GE, qe.F64, 100.0)]
        got1, want1 = _select_both(ctx, qe, [a, b], terms1)
    finally:
        a.free()
        b.free()
    assert 0.25 * n1 < len(want1) < 0.35 * n1
    assert np.array_equal(got1, want1)

    # Call 2: a smaller, ragged input with another predicate; most rows fail,
    # so leftover passing bits from call 1 would add rows.
    n2 = 70_001
    c = ctx.gen_uniform(42, 0, qe.I64, n2, -500, 500)
    try:
        terms2 = [(0, qe.CMP_GT, qe.I64, 450)]
        got2, want2 = _select_both(ctx, qe, [c], terms2)
    finally:
        c.free()
    assert 0 < len(want2) < 0.1 * n2
    assert np.array_equal(got2, want2)


def test_noalloc_allocates_nothing(ctx, qe):
    n = 1 << 20
    a = ctx.gen_uniform(43, 0, qe.I32, n, 0, 100)
    out = qe.DevBuffer(ctx, 4 * n)
    terms = [(0, qe.CMP_LE, qe.I32, 49)]
    try:
        want = ctx.filter_select([a], terms)
        ctx.set_mem_limit(1)
        try:
            m = ctx.filter_select_noalloc_into([a], terms, out)
            with pytest.raises(qe.OutOfMemory):
                ctx.filter_select([a], terms)
        finally:
            ctx.set_mem_limit(0)
        assert np.array_equal(out.numpy(np.uint32, m), want)
    finally:
        out.free()
        a.free()


def test_noalloc_limits_and_edges(ctx, qe):
    # Past the maximum: a clear error, nothing launched.
    big = ctx.gen_uniform(44, 0, qe.I32, (1 << 27) + 64, 0, 10)
    out = qe.DevBuffer(ctx, 4 * 64)
    try:
        with pytest.raises(qe.Unsupported, match="2\\^27"):
            ctx.filter_select_noalloc_into([big], [(0, qe.CMP_LT, qe.I32, 5)], out)
    finally:
        big.free()
    # A NULL literal passes nothing; every row passing; an empty input.
    c = ctx.gen_uniform(45, 0, qe.I32, 4099, 0, 10)
    empty = ctx.upload(np.zeros(0, dtype=np.int32))
    big_out = qe.DevBuffer(ctx, 4 * 4099)
    try:
        assert ctx.filter_select_noalloc_into([c], [(0, qe.CMP_LT, qe.I32, None)], big_out) == 0
        m = ctx.filter_select_noalloc_into([c], [(0, qe.CMP_GE, qe.I32, 0)], big_out)
        assert m == 4099 and np.array_equal(big_out.numpy(np.uint32, m), np.arange(4099, dtype=np.uint32))
        assert ctx.filter_select_noalloc_into([empty], [(0, qe.CMP_GE, qe.I32, 0)], out) == 0
    finally:
        for x in (c, empty, big_out, out):
            x.free()
