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
