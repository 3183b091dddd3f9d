"""Failure handling forced from the host: every kernel still exits normally.
Invalid arguments map to InvalidArgument (std::invalid_argument), a device
memory cap to OutOfMemory (std::bad_alloc); the context stays usable."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_out_of_memory_is_reported_and_recoverable(ctx, qe):
    k = ctx.upload(np.arange(100000, dtype=np.int64))
    try:
        ctx.set_mem_limit(1 << 20)  # far below what a 100k-row group-by needs
        with pytest.raises(qe.OutOfMemory):
            ctx.aggregate([k], [0], [(qe.COUNT_STAR, 0)]).free()
    finally:
        ctx.set_mem_limit(0)
    t = ctx.aggregate([k], [0], [(qe.COUNT_STAR, 0)])
    assert t.rows == 100000
    t.free()
    k.free()


def test_invalid_arguments(ctx, qe):
    a = ctx.upload(np.arange(4, dtype=np.int32))
    b = ctx.upload(np.arange(5, dtype=np.int32))
    f = ctx.upload(np.arange(4, dtype=np.float64))
    try:
        with pytest.raises(qe.InvalidArgument):
            ctx.aggregate([a, b], [0], [(qe.SUM, 1)])  # unequal column lengths
        with pytest.raises(qe.InvalidArgument):
            ctx.aggregate([a], [3], [(qe.SUM, 0)])  # key column out of range
        with pytest.raises(qe.InvalidArgument, match="time columns"):
            ctx.asof_build([a], f)  # F64 timestamp
        idx = ctx.asof_build([a], a)
        with pytest.raises(qe.InvalidArgument, match="train"):
            idx.probe([a], a, mode=qe.ASOF_LEARNED)  # learned probe before training
        with pytest.raises(qe.InvalidArgument, match="key column types"):
            idx.probe([f], a)
        idx.free()
    finally:
        for c in (a, b, f):
            c.free()


def test_upload_rejects_nulls_without_bitmap(qe, ctx):
    import ctypes as C
    v = np.arange(3, dtype=np.int64)
    d = qe.ColDesc(qe.I64, 0, 3, v.ctypes.data, None)  # claims NULLs, no bitmap
    out = C.c_void_p()
    st = qe.lib().qe_gpu_column_upload(ctx.h, C.byref(d), C.byref(out))
    assert st == qe.InvalidArgument.status


def test_inputs_of_2_32_rows_are_rejected(ctx, qe):
    """Row ids, join pairs, permutations and first-seen ids are u32 with
    0xFFFFFFFF as the no-row sentinel: every entry point rejects 2^32-1 rows or
    more with Unsupported before a kernel reads the (uninitialised) input."""
    import ctypes as C
    L = qe.lib()
    n = (1 << 32) - 1
    big = ctx.alloc(qe.I32, n)  # 16 GiB, never initialised or read
    small = ctx.upload(np.arange(4, dtype=np.int32))
    buf = qe.DevBuffer(ctx, 4096)
    try:
        with pytest.raises(qe.Unsupported, match="32-bit"):
            ctx.aggregate([big], [0], [(qe.COUNT_STAR, 0)])
        with pytest.raises(qe.Unsupported, match="32-bit"):
            ctx.hash_join([big], [small])
        with pytest.raises(qe.Unsupported, match="32-bit"):
            ctx.join_aggregate([big], [small, small], [0], [0], qe.INNER, [2], [(qe.COUNT_STAR, 0)])
        with pytest.raises(qe.Unsupported, match="32-bit"):
            ctx.sort_perm_into([big], buf)
        with pytest.raises(qe.Unsupported, match="32-bit"):
            ctx.asof_build([big], big)
        idx = ctx.asof_build([small], small)
        try:
            with pytest.raises(qe.Unsupported, match="32-bit"):
                idx.probe_into([big], big, buf)
        finally:
            idx.free()
        cnt = C.c_uint64()
        terms = ctx._terms([(0, qe.CMP_GT, qe.I32, 0)])
        st = L.qe_gpu_filter_select(ctx.h, qe._cols([big]), 1, C.cast(terms, C.c_void_p), 1, buf.p, C.byref(cnt))
        assert st == qe.Unsupported.status
        ht = ctx.hash_table([qe.I32])
        try:
            st = L.qe_gpu_ht_insert_or_find(ctx.h, ht.h, qe._cols([big]), None, n, buf.p)
            assert st == qe.Unsupported.status
            st = L.qe_gpu_ht_find(ctx.h, ht.h, qe._cols([big]), None, n, buf.p)
            assert st == qe.Unsupported.status
        finally:
            ht.free()
    finally:
        buf.free()
        small.free()
        big.free()
