"""ctypes binding of libqe_gpu.so, the MI355X (gfx950) execution path for the
query engine's hash table, hash join, hash group-by and as-of join.

This is thin plumbing over the C ABI declared in include/qe_gpu.h: it moves
numpy columns to HBM and back and turns status codes into the exception types
the reference operators throw (std::invalid_argument -> InvalidArgument, a
ValueError; std::overflow_error -> Overflow; std::bad_alloc -> OutOfMemory).
Nothing here computes a result: every operator runs in the HIP library, and a
missing library raises instead of falling back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libqe_gpu.so"

# qe::Type order (core/types.h:18)
I32, I64, F64, BOOL, TS, STR = range(6)
TYPE_NAMES = {I32: "I32", I64: "I64", F64: "F64", BOOL: "BOOL", TS: "TS", STR: "STR"}
NP_DTYPE = {I32: np.int32, I64: np.int64, F64: np.float64, BOOL: np.uint8, TS: np.int64, STR: np.int32}
# qe::AggFunc order (ops/aggregate.h:61)
COUNT_STAR, COUNT, SUM, MIN, MAX, AVG = range(6)
INNER, LEFT = 0, 1
ASOF_BINARY, ASOF_LEARNED, ASOF_LEARNED_NO_MLP = 0, 1, 2
ASOF_OUT_SORTED = 0x100  # or into a mode: matches as sorted positions of the index
NO_ROW = 0xFFFFFFFF


class QeError(RuntimeError):
    status = -1


class InvalidArgument(QeError, ValueError):
    status = 1


class OutOfMemory(QeError, MemoryError):
    status = 2


class Overflow(QeError, OverflowError):
    status = 3


class HipError(QeError):
    status = 4


class Unsupported(QeError, NotImplementedError):
    status = 5


_ERRORS = {1: InvalidArgument, 2: OutOfMemory, 3: Overflow, 4: HipError, 5: Unsupported}


class ColDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("all_valid", C.c_int32), ("len", C.c_uint64),
                ("data", C.c_void_p), ("validity", C.c_void_p)]


class AggSpec(C.Structure):
    _fields_ = [("func", C.c_int32), ("input_col", C.c_uint32)]


class PredTerm(C.Structure):
    _fields_ = [("col", C.c_uint32), ("op", C.c_int32), ("lit_type", C.c_int32), ("lit_is_null", C.c_int32),
                ("lit_i", C.c_int64), ("lit_f", C.c_double)]


class AsofStats(C.Structure):
    _fields_ = [("build_rows", C.c_uint64), ("partitions", C.c_uint64), ("pieces", C.c_uint64),
                ("max_piece_error", C.c_uint64), ("router_buckets", C.c_uint64),
                ("mean_bucket_knots", C.c_double), ("mean_bucket_knots_linear", C.c_double),
                ("fallbacks", C.c_uint64), ("final_loss", C.c_double),
                ("mlp_candidates", C.c_uint64), ("mlp_partitions", C.c_uint64),
                ("mean_bucket_knots_mlp", C.c_double), ("sample_probe_ms_mlp", C.c_double),
                ("sample_probe_ms_linear", C.c_double), ("data_line_buckets", C.c_uint64)]


P = C.c_void_p
_SIGS = {
    "qe_gpu_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "qe_gpu_ctx_destroy": (None, [P]),
    "qe_gpu_last_error": (C.c_char_p, [P]),
    "qe_gpu_sync": (C.c_int, [P]),
    "qe_gpu_set_mem_limit": (C.c_int, [P, C.c_uint64]),
    "qe_gpu_mem_usage": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "qe_gpu_set_timing": (C.c_int, [P, C.c_int]),
    "qe_gpu_filter": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32, C.POINTER(P)]),
    "qe_gpu_filter_select": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32, P, C.POINTER(C.c_uint64)]),
    "qe_gpu_filter_select_noalloc": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32, P, C.POINTER(C.c_uint64)]),
    "qe_gpu_set_tuning": (C.c_int, [P, C.c_int32, C.c_int64]),
    "qe_gpu_xcd_of_blocks": (C.c_int, [P, C.c_uint32, P]),
    "qe_gpu_timing": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "qe_gpu_column_upload": (C.c_int, [P, C.POINTER(ColDesc), C.POINTER(P)]),
    "qe_gpu_column_alloc": (C.c_int, [P, C.c_int32, C.c_uint64, C.c_int32, C.POINTER(P)]),
    "qe_gpu_column_download": (C.c_int, [P, P, P, P, C.POINTER(C.c_int32)]),
    "qe_gpu_column_download_range": (C.c_int, [P, P, C.c_uint64, C.c_uint64, P, P]),
    "qe_gpu_column_free": (None, [P, P]),
    "qe_gpu_column_len": (C.c_uint64, [P]),
    "qe_gpu_column_type": (C.c_int32, [P]),
    "qe_gpu_column_all_valid": (C.c_int32, [P]),
    "qe_gpu_column_slice": (C.c_int, [P, P, C.c_uint64, C.c_uint64, C.POINTER(P)]),
    "qe_gpu_dict_create": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(P)]),
    "qe_gpu_dict_free": (None, [P, P]),
    "qe_gpu_dict_size": (C.c_uint32, [P]),
    "qe_gpu_column_set_dict": (C.c_int, [P, P, P]),
    "qe_gpu_column_dict": (P, [P]),
    "qe_gpu_column_remap": (C.c_int, [P, P, P, C.c_uint32, C.c_int32, C.POINTER(P)]),
    "qe_gpu_take": (C.c_int, [P, C.POINTER(P), C.c_uint32, P, C.c_uint64, C.POINTER(P)]),
    "qe_gpu_table_rows": (C.c_uint64, [P]),
    "qe_gpu_table_cols": (C.c_uint32, [P]),
    "qe_gpu_table_col": (P, [P, C.c_uint32]),
    "qe_gpu_table_free": (None, [P, P]),
    "qe_gpu_aggregate": (C.c_int, [P, C.POINTER(P), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                   C.POINTER(AggSpec), C.c_uint32, C.POINTER(P)]),
    "qe_gpu_hash_join": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.c_uint32, C.c_int32, C.POINTER(P)]),
    "qe_gpu_pairs_count": (C.c_uint64, [P]),
    "qe_gpu_pairs_download": (C.c_int, [P, P, C.c_uint64, C.c_uint64, P, P]),
    "qe_gpu_pairs_gather": (C.c_int, [P, P, C.c_uint64, C.c_uint64, C.POINTER(P), C.c_uint32,
                                      C.POINTER(P), C.c_uint32, C.POINTER(P)]),
    "qe_gpu_pairs_free": (None, [P, P]),
    "qe_gpu_join_aggregate": (C.c_int, [P, C.POINTER(P), C.c_uint32, C.POINTER(P), C.c_uint32,
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32, C.c_int32,
                                        C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(AggSpec), C.c_uint32,
                                        C.POINTER(P)]),
    "qe_gpu_ht_create": (C.c_int, [P, C.POINTER(C.c_int32), C.c_uint32, C.c_int32, C.c_uint64, C.c_double,
                                   C.c_uint64, C.POINTER(P)]),
    "qe_gpu_ht_insert_or_find": (C.c_int, [P, P, C.POINTER(P), P, C.c_uint64, P]),
    "qe_gpu_ht_find": (C.c_int, [P, P, C.POINTER(P), P, C.c_uint64, P]),
    "qe_gpu_ht_num_groups": (C.c_uint64, [P]),
    "qe_gpu_ht_capacity": (C.c_uint64, [P]),
    "qe_gpu_ht_key_words": (C.c_int, [P, P, C.c_uint32, P, P]),
    "qe_gpu_ht_free": (None, [P, P]),
    "qe_gpu_sort_perm": (C.c_int, [P, C.POINTER(P), C.c_uint32, P, P, P]),
    "qe_gpu_asof_build": (C.c_int, [P, C.POINTER(P), C.c_uint32, P, C.POINTER(P)]),
    "qe_gpu_asof_train": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_float]),
    "qe_gpu_asof_probe": (C.c_int, [P, P, C.POINTER(P), C.c_uint32, P, C.c_int32, C.c_int64, C.c_int32, P]),
    "qe_gpu_asof_pairs": (C.c_int, [P, P, C.c_uint64, C.c_int32, C.POINTER(P)]),
    "qe_gpu_asof_take_sorted": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(P)]),
    "qe_gpu_asof_cluster": (C.c_int, [P, P, P, C.c_uint32, P, P, C.c_uint32, C.POINTER(P)]),
    "qe_gpu_asof_stats": (C.c_int, [P, P, C.POINTER(AsofStats)]),
    "qe_gpu_asof_sorted": (C.c_int, [P, P, P, P, P]),
    "qe_gpu_asof_free": (None, [P, P]),
    "qe_gpu_mlp_forward": (C.c_int, [P, P, P, C.c_uint64, P]),
    "qe_gpu_mlp_grad": (C.c_int, [P, P, P, P, C.c_uint32, P, P]),
    "qe_gpu_batch_prepare": (C.c_int, [P, P, C.c_uint32, C.POINTER(P)]),
    "qe_gpu_batch_run": (C.c_int, [P, P, P, C.c_int32, P, P]),
    "qe_gpu_batch_autotune": (C.c_int, [P, P, P]),
    "qe_gpu_batch_strategies": (C.c_int, [P, P, P, P]),
    "qe_gpu_batch_free": (None, [P, P]),
    "qe_gpu_gen_uniform": (C.c_int, [P, C.c_uint64, C.c_uint64, P, C.c_double, C.c_double, C.c_double]),
    "qe_gpu_malloc": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
    "qe_gpu_free": (C.c_int, [P, P]),
    "qe_gpu_memcpy_d2h": (C.c_int, [P, P, P, C.c_uint64]),
    "qe_gpu_memcpy_h2d": (C.c_int, [P, P, P, C.c_uint64]),
    "qe_gpu_stream_copy": (C.c_int, [P, P, P, C.c_uint64, C.POINTER(C.c_double)]),
    "qe_gpu_stream_read": (C.c_int, [P, P, C.c_uint64, P, C.POINTER(C.c_double)]),
    "qe_gpu_map_gather": (C.c_int, [P, P, C.c_int64, C.c_uint64, P, P, C.POINTER(C.c_double)]),
    "qe_gpu_join_autotune": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32, P, P, C.c_uint32, C.c_int32, P, C.c_uint32,
                                       P, C.c_uint32, P]),
    "qe_gpu_device_info": (C.c_int, [P, P]),
    "qe_gpu_gen_groupby": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, P, P]),
    "qe_gpu_gen_join_build": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint32, P, P]),
    "qe_gpu_gen_join_probe": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, P, P]),
    "qe_gpu_gen_asof": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_uint32,
                                  P, P, P, P, P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libqe_gpu.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python query-engine_amd/build.py` "
                               "(the HIP path has no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:  # reported by the ABI test; calling it raises AttributeError
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pack_validity(valid: np.ndarray) -> np.ndarray:
    """bool[n] -> u64 words, LSB-first (core/validity.h:21-64); padding bits set."""
    n = len(valid)
    nw = (n + 63) // 64
    padded = np.ones(nw * 64, dtype=bool)
    padded[:n] = valid
    return np.packbits(padded, bitorder="little").view(np.uint64).copy()


def unpack_validity(words: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


class Dict:
    """A string dictionary for STR columns (the StringDict layout: strings
    concatenated plus offsets). `strings[code]` is a code's value."""

    def __init__(self, ctx: "Context", strings):
        self.ctx = ctx
        self.strings = list(strings)
        enc = [x.encode() for x in self.strings]
        off = np.zeros(len(enc) + 1, dtype=np.uint32)
        if enc:
            off[1:] = np.cumsum([len(e) for e in enc])
        blob = np.frombuffer(b"".join(enc) + b"\0", dtype=np.uint8).copy()
        h = P()
        ctx._check(lib().qe_gpu_dict_create(ctx.h, _ptr(blob), _ptr(off), len(enc), C.byref(h)))
        self.h = h
        ctx._dicts[h.value] = self

    def __len__(self):
        return len(self.strings)

    def free(self):
        if self.h:
            self.ctx._dicts.pop(self.h.value, None)
            lib().qe_gpu_dict_free(self.ctx.h, self.h)
        self.h = None


class DevCol:
    """A device column (HBM). Owned columns are freed with the context."""

    def __init__(self, ctx: "Context", handle, owned: bool = True):
        self.ctx, self.h, self.owned = ctx, handle, owned

    @property
    def dict(self) -> "Dict | None":
        """The dictionary of a STR column (a Dict of this context), or None."""
        d = lib().qe_gpu_column_dict(self.h)
        return self.ctx._dicts.get(d) if d else None

    def set_dict(self, d: "Dict | None"):
        self.ctx._check(lib().qe_gpu_column_set_dict(self.ctx.h, self.h, d.h if d is not None else None))

    def strings(self):
        """A STR column's values as Python strings (None for NULL)."""
        vals, valid = self.numpy()
        d = self.dict
        assert d is not None, "STR column without a dictionary"
        return [None if (valid is not None and not valid[i]) else d.strings[c] for i, c in enumerate(vals.tolist())]

    @property
    def len(self) -> int:
        return int(lib().qe_gpu_column_len(self.h))

    @property
    def type(self) -> int:
        return int(lib().qe_gpu_column_type(self.h))

    @property
    def all_valid(self) -> bool:
        return bool(lib().qe_gpu_column_all_valid(self.h))

    def numpy(self):
        """(values, valid) with valid None when the column has no NULLs."""
        n, t = self.len, self.type
        data = np.empty(n, dtype=NP_DTYPE[t])
        words = np.empty(max((n + 63) // 64, 1), dtype=np.uint64)
        av = C.c_int32(1)
        self.ctx._check(lib().qe_gpu_column_download(self.ctx.h, self.h, _ptr(data), _ptr(words), C.byref(av)))
        return data, (None if av.value else unpack_validity(words, n))

    def free(self):
        if self.h and self.owned:
            lib().qe_gpu_column_free(self.ctx.h, self.h)
        self.h = None


class Table:
    def __init__(self, ctx: "Context", handle):
        self.ctx, self.h = ctx, handle

    @property
    def rows(self) -> int:
        return int(lib().qe_gpu_table_rows(self.h))

    @property
    def ncols(self) -> int:
        return int(lib().qe_gpu_table_cols(self.h))

    def col(self, i: int) -> DevCol:
        return DevCol(self.ctx, lib().qe_gpu_table_col(self.h, i), owned=False)

    def numpy(self):
        return [self.col(i).numpy() for i in range(self.ncols)]

    def free(self):
        if self.h:
            lib().qe_gpu_table_free(self.ctx.h, self.h)
        self.h = None


class Pairs:
    def __init__(self, ctx: "Context", handle):
        self.ctx, self.h = ctx, handle

    @property
    def count(self) -> int:
        return int(lib().qe_gpu_pairs_count(self.h))

    def numpy(self, offset: int = 0, count: int | None = None):
        n = self.count - offset if count is None else count
        p = np.empty(n, dtype=np.uint32)
        b = np.empty(n, dtype=np.uint32)
        self.ctx._check(lib().qe_gpu_pairs_download(self.ctx.h, self.h, offset, n, _ptr(p), _ptr(b)))
        return p, b

    def gather(self, probe_cols, build_cols, offset=0, count=None) -> Table:
        n = self.count - offset if count is None else count
        out = P()
        pa = (P * max(len(probe_cols), 1))(*[c.h for c in probe_cols])
        ba = (P * max(len(build_cols), 1))(*[c.h for c in build_cols])
        self.ctx._check(lib().qe_gpu_pairs_gather(self.ctx.h, self.h, offset, n, pa, len(probe_cols), ba,
                                                  len(build_cols), C.byref(out)))
        return Table(self.ctx, out)

    def free(self):
        if self.h:
            lib().qe_gpu_pairs_free(self.ctx.h, self.h)
        self.h = None


class DevBuffer:
    """Raw device allocation (e.g. a u32 match array)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx, self.nbytes = ctx, nbytes
        p = P()
        ctx._check(lib().qe_gpu_malloc(ctx.h, max(nbytes, 1), C.byref(p)))
        self.p = p

    def numpy(self, dtype, n):
        a = np.empty(n, dtype=dtype)
        self.ctx._check(lib().qe_gpu_memcpy_d2h(self.ctx.h, _ptr(a), self.p, a.nbytes))
        return a

    def upload(self, a: np.ndarray):
        self.ctx._check(lib().qe_gpu_memcpy_h2d(self.ctx.h, self.p, _ptr(a), a.nbytes))

    def free(self):
        if self.p:
            lib().qe_gpu_free(self.ctx.h, self.p)
        self.p = None


def _cols(cols):
    return (P * max(len(cols), 1))(*[c.h for c in cols])


def _u32(xs):
    return (C.c_uint32 * max(len(xs), 1))(*xs)


def _aggs(aggs):
    return (AggSpec * max(len(aggs), 1))(*[AggSpec(f, c) for f, c in aggs])


class Context:
    """One GPU context (one HIP stream, caching device allocator)."""

    def __init__(self, device: int = 0):
        h = P()
        st = lib().qe_gpu_ctx_create(device, C.byref(h))
        if st != 0:
            raise HipError(f"qe_gpu_ctx_create failed with status {st} (is a GPU visible?)")
        self.h = h
        self._dicts = {}  # dictionary handle -> Dict

    def close(self):
        if self.h:
            for d in list(self._dicts.values()):
                d.free()
            lib().qe_gpu_ctx_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int):
        if st != 0:
            msg = lib().qe_gpu_last_error(self.h).decode()
            raise _ERRORS.get(st, QeError)(msg)

    # ---- columns -----------------------------------------------------------
    def upload(self, values: np.ndarray, type: int | None = None, valid: np.ndarray | None = None,
               dict: "Dict | None" = None) -> DevCol:
        """A device column; a STR column's values are int32 codes into `dict`."""
        if type is None:
            type = {np.dtype(np.int32): I32, np.dtype(np.int64): I64, np.dtype(np.float64): F64,
                    np.dtype(np.uint8): BOOL, np.dtype(np.bool_): BOOL}[np.asarray(values).dtype]
        v = np.ascontiguousarray(values, dtype=NP_DTYPE[type])
        words = None if valid is None else pack_validity(np.asarray(valid, dtype=bool))
        d = ColDesc(type, 1 if valid is None else 0, len(v), _ptr(v), _ptr(words))
        out = P()
        self._check(lib().qe_gpu_column_upload(self.h, C.byref(d), C.byref(out)))
        col = DevCol(self, out)
        if dict is not None:
            col.set_dict(dict)
        return col

    def dict(self, strings) -> Dict:
        """A string dictionary (freed with the context or by Dict.free)."""
        return Dict(self, strings)

    def upload_strings(self, values, d: "Dict | None" = None) -> DevCol:
        """A STR column from Python strings (None = NULL), encoded in `d` or in a
        new dictionary of its distinct values in first-seen order."""
        if d is None:
            d = self.dict(list(dict.fromkeys(x for x in values if x is not None)))
        idx = {x: i for i, x in enumerate(d.strings)}
        codes = np.array([idx[x] if x is not None else 0 for x in values], dtype=np.int32)
        valid = None if all(x is not None for x in values) else np.array([x is not None for x in values])
        return self.upload(codes, STR, valid, dict=d)

    def stream_copy(self, dst: "DevBuffer", src: "DevBuffer", nbytes: int) -> float:
        """A plain streaming copy (the measured HBM ceiling); returns its kernel ms."""
        ms = C.c_double()
        self._check(lib().qe_gpu_stream_copy(self.h, dst.p, src.p, nbytes, C.byref(ms)))
        return ms.value

    def map_gather(self, keys: "DevCol", kmin: int, range_: int, map_: "DevBuffer", sink: "DevBuffer") -> float:
        """The join probe's map lookups alone over an I64 key column; returns kernel ms."""
        ms = C.c_double()
        self._check(lib().qe_gpu_map_gather(self.h, keys.h, kmin, range_, map_.p, sink.p, C.byref(ms)))
        return ms.value

    def device_info(self) -> dict:
        """The context's device (name, gfx arch, CUs, HBM bytes, clocks, HIP versions)."""
        d = DeviceInfo()
        self._check(lib().qe_gpu_device_info(self.h, C.byref(d)))
        out = {f: getattr(d, f) for f, _ in DeviceInfo._fields_}
        out["name"], out["arch"] = d.name.decode(), d.arch.decode()
        return out

    def stream_read(self, src: "DevBuffer", sink: "DevBuffer", nbytes: int) -> float:
        """A read-only stream over src (the measured read ceiling); returns its kernel ms."""
        ms = C.c_double()
        self._check(lib().qe_gpu_stream_read(self.h, src.p, nbytes, sink.p, C.byref(ms)))
        return ms.value

    def sort_perm_into(self, keys, out: "DevBuffer", desc=None, nulls_first=None):
        """Like sort_perm, with the permutation left in a device buffer."""
        d = None if desc is None else (C.c_int32 * len(keys))(*desc)
        nf = None if nulls_first is None else (C.c_int32 * len(keys))(*nulls_first)
        self._check(lib().qe_gpu_sort_perm(self.h, _cols(keys), len(keys), d, nf, out.p))

    def slice(self, col: DevCol, offset: int, n: int) -> DevCol:
        """Rows [offset, offset + n) of `col` without a copy (offset % 64 == 0)."""
        out = P()
        self._check(lib().qe_gpu_column_slice(self.h, col.h, offset, n, C.byref(out)))
        return DevCol(self, out)

    def asof_pairs(self, match: "DevBuffer", n: int, join_type=INNER) -> Pairs:
        """Join pairs from per-probe as-of matches (INNER drops unmatched rows)."""
        out = P()
        self._check(lib().qe_gpu_asof_pairs(self.h, match.p, n, join_type, C.byref(out)))
        return Pairs(self, out)

    def take(self, cols, idx: "DevBuffer", m: int) -> Table:
        """Rows idx[0..m) (a device u32 buffer) of `cols` as a new table."""
        out = P()
        self._check(lib().qe_gpu_take(self.h, _cols(cols), len(cols), idx.p, m, C.byref(out)))
        return Table(self, out)

    def remap(self, col: DevCol, table, out_type: int = I32) -> DevCol:
        t = np.ascontiguousarray(table, dtype=np.int32)
        out = P()
        self._check(lib().qe_gpu_column_remap(self.h, col.h, _ptr(t), len(t), out_type, C.byref(out)))
        return DevCol(self, out)

    def alloc(self, type: int, n: int, nullable: bool = False) -> DevCol:
        out = P()
        self._check(lib().qe_gpu_column_alloc(self.h, type, n, 1 if nullable else 0, C.byref(out)))
        return DevCol(self, out)

    def sync(self):
        self._check(lib().qe_gpu_sync(self.h))

    def set_timing(self, on: bool):
        self._check(lib().qe_gpu_set_timing(self.h, 1 if on else 0))

    def xcd_of_blocks(self, nblocks: int) -> np.ndarray:
        out = np.empty(nblocks, dtype=np.uint32)
        self._check(lib().qe_gpu_xcd_of_blocks(self.h, nblocks, _ptr(out)))
        return out

    def set_tuning(self, key: int, value: int):
        self._check(lib().qe_gpu_set_tuning(self.h, key, value))

    def timing(self, name: str):
        ms, n = C.c_double(), C.c_uint64()
        self._check(lib().qe_gpu_timing(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_mem_limit(self, nbytes: int):
        self._check(lib().qe_gpu_set_mem_limit(self.h, nbytes))

    def mem_usage(self) -> tuple[int, int]:
        """(bytes in live device buffers, bytes held from the device incl. cached blocks)."""
        live, held = C.c_uint64(), C.c_uint64()
        self._check(lib().qe_gpu_mem_usage(self.h, C.byref(live), C.byref(held)))
        return live.value, held.value

    # ---- operators ---------------------------------------------------------
    def aggregate(self, cols, keys, aggs) -> Table:
        out = P()
        self._check(lib().qe_gpu_aggregate(self.h, _cols(cols), len(cols), _u32(keys), len(keys), _aggs(aggs),
                                           len(aggs), C.byref(out)))
        return Table(self, out)

    @staticmethod
    def _terms(terms):
        """terms: (col, op, lit_type, value) with value None for a NULL literal."""
        arr = (PredTerm * max(len(terms), 1))()
        for i, (col, op, lt, v) in enumerate(terms):
            arr[i].col, arr[i].op, arr[i].lit_type = col, op, lt
            arr[i].lit_is_null = 1 if v is None else 0
            if v is not None:
                if lt == F64:
                    arr[i].lit_f = float(v)
                else:
                    arr[i].lit_i = int(v)
        return arr

    def filter(self, cols, terms) -> Table:
        """Rows where every `col op literal` term is TRUE, every column compacted."""
        out = P()
        self._check(lib().qe_gpu_filter(self.h, _cols(cols), len(cols), C.cast(self._terms(terms), P), len(terms),
                                        C.byref(out)))
        return Table(self, out)

    def filter_select(self, cols, terms) -> np.ndarray:
        """Passing row indices (ascending) of the same predicate."""
        n = cols[0].len
        buf = DevBuffer(self, 4 * max(n, 1))
        cnt = C.c_uint64()
        try:
            self._check(lib().qe_gpu_filter_select(self.h, _cols(cols), len(cols), C.cast(self._terms(terms), P),
                                                   len(terms), buf.p, C.byref(cnt)))
            return buf.numpy(np.uint32, cnt.value)
        finally:
            buf.free()

    def filter_select_noalloc_into(self, cols, terms, out: "DevBuffer") -> int:
        """The same selection through the entry point that allocates nothing:
        indices land in `out` (device u32[len]); returns their count."""
        cnt = C.c_uint64()
        self._check(lib().qe_gpu_filter_select_noalloc(self.h, _cols(cols), len(cols),
                                                       C.cast(self._terms(terms), P), len(terms), out.p,
                                                       C.byref(cnt)))
        return cnt.value

    def hash_join(self, probe_keys, build_keys, join_type=INNER) -> Pairs:
        out = P()
        self._check(lib().qe_gpu_hash_join(self.h, _cols(probe_keys), _cols(build_keys), len(probe_keys),
                                           join_type, C.byref(out)))
        return Pairs(self, out)

    def join_aggregate(self, probe_cols, build_cols, probe_keys, build_keys, join_type, group_keys,
                       aggs) -> Table:
        out = P()
        self._check(lib().qe_gpu_join_aggregate(
            self.h, _cols(probe_cols), len(probe_cols), _cols(build_cols), len(build_cols), _u32(probe_keys),
            _u32(build_keys), len(probe_keys), join_type, _u32(group_keys), len(group_keys), _aggs(aggs),
            len(aggs), C.byref(out)))
        return Table(self, out)

    def join_autotune(self, probe_cols, build_cols, probe_keys, build_keys, join_type, group_keys, aggs) -> dict:
        """Time the fused join -> group-by's launch shapes on this query; the
        fastest is left set on the context and returned."""
        t = JoinTuning()
        self._check(lib().qe_gpu_join_autotune(self.h, _cols(probe_cols), len(probe_cols), _cols(build_cols),
                                               len(build_cols), _u32(probe_keys), _u32(build_keys), len(probe_keys),
                                               join_type, _u32(group_keys), len(group_keys), _aggs(aggs), len(aggs),
                                               C.byref(t)))
        return {f: getattr(t, f) for f, _ in JoinTuning._fields_}

    def prepare_join_aggregate(self, probe_cols, build_cols, probe_keys, build_keys, join_type, group_keys,
                               aggs):
        """join_aggregate with its arguments marshalled once, for repeated calls
        on the same columns (which must stay alive): returns a callable."""
        fn = lib().qe_gpu_join_aggregate
        args = (self.h, _cols(probe_cols), len(probe_cols), _cols(build_cols), len(build_cols), _u32(probe_keys),
                _u32(build_keys), len(probe_keys), join_type, _u32(group_keys), len(group_keys), _aggs(aggs),
                len(aggs))

        def run() -> Table:
            out = P()
            self._check(fn(*args, C.byref(out)))
            return Table(self, out)

        return run

    def sort_perm(self, keys, desc=None, nulls_first=None) -> np.ndarray:
        n = keys[0].len
        out = DevBuffer(self, 4 * max(n, 1))
        d = None if desc is None else (C.c_int32 * len(keys))(*desc)
        nf = None if nulls_first is None else (C.c_int32 * len(keys))(*nulls_first)
        try:
            self._check(lib().qe_gpu_sort_perm(self.h, _cols(keys), len(keys), d, nf, out.p))
            return out.numpy(np.uint32, n)
        finally:
            out.free()

    # ---- hash table --------------------------------------------------------
    def hash_table(self, key_types, null_policy=0, initial_capacity=1024, max_load_factor=0.7,
                   seed=0x51E57A11AB1E5EED) -> "HashTable":
        return HashTable(self, key_types, null_policy, initial_capacity, max_load_factor, seed)

    # ---- as-of -------------------------------------------------------------
    def asof_build(self, build_keys, build_time) -> "AsofIndex":
        out = P()
        self._check(lib().qe_gpu_asof_build(self.h, _cols(build_keys), len(build_keys), build_time.h,
                                            C.byref(out)))
        return AsofIndex(self, out)

    # ---- learned-index router internals (verification) --------------------
    def mlp_forward(self, params: np.ndarray, x: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(params, dtype=np.float32)
        xv = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(len(xv), dtype=np.float32)
        self._check(lib().qe_gpu_mlp_forward(self.h, _ptr(p), _ptr(xv), len(xv), _ptr(y)))
        return y

    def mlp_grad(self, params: np.ndarray, x: np.ndarray, c: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        xv = np.ascontiguousarray(x, dtype=np.float32)
        cv = np.ascontiguousarray(c, dtype=np.float32)
        g = np.empty(321, dtype=np.float32)
        loss = np.empty(1, dtype=np.float32)
        self._check(lib().qe_gpu_mlp_grad(self.h, _ptr(p), _ptr(xv), _ptr(cv), len(xv), _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    # ---- generators --------------------------------------------------------
    def gen_groupby(self, seed, n, key_range, null_frac):
        key = self.alloc(I64, n)
        val = self.alloc(F64, n, nullable=null_frac > 0)
        self._check(lib().qe_gpu_gen_groupby(self.h, seed, n, key_range, null_frac, key.h, val.h))
        return key, val

    def gen_join(self, seed, build_n, probe_n, groups, zipf_s=1.0, miss_frac=0.1):
        bk, bg = self.alloc(I64, build_n), self.alloc(I32, build_n)
        pk, pp = self.alloc(I64, probe_n), self.alloc(F64, probe_n)
        self._check(lib().qe_gpu_gen_join_build(self.h, seed, build_n, groups, bk.h, bg.h))
        self._check(lib().qe_gpu_gen_join_probe(self.h, seed, probe_n, build_n, zipf_s, miss_frac, pk.h, pp.h))
        return (pk, pp), (bk, bg)

    def gen_asof(self, seed, right_n, left_n, nsym=5000, zipf_s=1.1, sessions=0):
        rs, rt, rp = self.alloc(I32, right_n), self.alloc(I64, right_n), self.alloc(F64, right_n)
        ls, lt = self.alloc(I32, left_n), self.alloc(I64, left_n)
        self._check(lib().qe_gpu_gen_asof(self.h, seed, right_n, left_n, nsym, zipf_s, sessions, rs.h, rt.h, rp.h,
                                          ls.h, lt.h))
        return (rs, rt, rp), (ls, lt)

    def gen_uniform(self, seed, stream, type, n, lo, hi, null_frac=0.0) -> DevCol:
        c = self.alloc(type, n, nullable=null_frac > 0)
        self._check(lib().qe_gpu_gen_uniform(self.h, seed, stream, c.h, lo, hi, null_frac))
        return c

    def batch(self, queries) -> "BatchPlan":
        return BatchPlan(self, queries)


class HashTable:
    """Device HashTable (ops/hashtable.h:108-190): first-seen dense ids."""

    def __init__(self, ctx, key_types, null_policy, initial_capacity, max_load_factor, seed):
        self.ctx = ctx
        kt = (C.c_int32 * len(key_types))(*key_types)
        out = P()
        ctx._check(lib().qe_gpu_ht_create(ctx.h, kt, len(key_types), null_policy, initial_capacity,
                                          max_load_factor, seed, C.byref(out)))
        self.h = out

    def _run(self, fn, keys, sel, n):
        out = DevBuffer(self.ctx, 4 * max(n, 1))
        selbuf = None
        if sel is not None:
            selbuf = DevBuffer(self.ctx, 4 * max(len(sel), 1))
            selbuf.upload(np.ascontiguousarray(sel, dtype=np.uint32))
        try:
            self.ctx._check(fn(self.ctx.h, self.h, _cols(keys), None if selbuf is None else selbuf.p, n, out.p))
            return out.numpy(np.uint32, n)
        finally:
            out.free()
            if selbuf is not None:
                selbuf.free()

    def insert_or_find(self, keys, n=None, sel=None):
        n = (len(sel) if sel is not None else keys[0].len) if n is None else n
        return self._run(lib().qe_gpu_ht_insert_or_find, keys, sel, n)

    def find(self, keys, n=None, sel=None):
        n = (len(sel) if sel is not None else keys[0].len) if n is None else n
        return self._run(lib().qe_gpu_ht_find, keys, sel, n)

    @property
    def num_groups(self) -> int:
        return int(lib().qe_gpu_ht_num_groups(self.h))

    @property
    def capacity(self) -> int:
        return int(lib().qe_gpu_ht_capacity(self.h))

    def key_words(self, col):
        g = self.num_groups
        w = np.empty(max(g, 1), dtype=np.uint64)
        nl = np.empty(max(g, 1), dtype=np.uint8)
        self.ctx._check(lib().qe_gpu_ht_key_words(self.ctx.h, self.h, col, _ptr(w), _ptr(nl)))
        return w[:g], nl[:g].astype(bool)

    def free(self):
        if self.h:
            lib().qe_gpu_ht_free(self.ctx.h, self.h)
        self.h = None


class JoinTuning(C.Structure):
    _fields_ = [("pairs", C.c_int32), ("blocks_per_cu", C.c_int32), ("probe_ms", C.c_double),
                ("default_ms", C.c_double), ("candidates", C.c_int32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("device", C.c_int32),
                ("compute_units", C.c_int32), ("hbm_bytes", C.c_uint64), ("lds_bytes_per_block", C.c_uint64),
                ("l2_bytes", C.c_uint64), ("clock_khz", C.c_int32), ("memory_clock_khz", C.c_int32),
                ("memory_bus_bits", C.c_int32), ("hip_runtime_version", C.c_int32),
                ("hip_driver_version", C.c_int32)]


class AsofIndex:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def train(self, piece_rows=64, steps=300, lr=0.003):
        self.ctx._check(lib().qe_gpu_asof_train(self.ctx.h, self.h, piece_rows, steps, lr))

    def probe_into(self, probe_keys, probe_time, out: DevBuffer, tolerance=None, mode=ASOF_BINARY):
        self.ctx._check(lib().qe_gpu_asof_probe(self.ctx.h, self.h, _cols(probe_keys), len(probe_keys),
                                                probe_time.h, 0 if tolerance is None else 1,
                                                0 if tolerance is None else int(tolerance), mode, out.p))

    def probe(self, probe_keys, probe_time, tolerance=None, mode=ASOF_BINARY) -> np.ndarray:
        n = probe_time.len
        out = DevBuffer(self.ctx, 4 * max(n, 1))
        try:
            self.probe_into(probe_keys, probe_time, out, tolerance, mode)
            return out.numpy(np.uint32, n)
        finally:
            out.free()

    def cluster(self, probe_keys, probe_time, cols) -> "Table":
        """`cols` (probe-side columns) reordered into groups by the index
        region their probes read (qe_gpu_asof_cluster)."""
        out = P()
        self.ctx._check(lib().qe_gpu_asof_cluster(self.ctx.h, self.h, _cols(probe_keys), len(probe_keys),
                                                  probe_time.h, _cols(cols), len(cols), C.byref(out)))
        return Table(self.ctx, out)

    def take_sorted(self, cols) -> "Table":
        """Build-side columns gathered into the index's sorted order."""
        out = P()
        self.ctx._check(lib().qe_gpu_asof_take_sorted(self.ctx.h, self.h, _cols(cols), len(cols), C.byref(out)))
        return Table(self.ctx, out)

    def stats(self) -> dict:
        s = AsofStats()
        self.ctx._check(lib().qe_gpu_asof_stats(self.ctx.h, self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in AsofStats._fields_}

    def sorted(self, n_rows: int, n_parts: int):
        ts = np.empty(max(n_rows, 1), dtype=np.int64)
        rows = np.empty(max(n_rows, 1), dtype=np.uint32)
        offs = np.empty(n_parts + 1, dtype=np.uint64)
        self.ctx._check(lib().qe_gpu_asof_sorted(self.ctx.h, self.h, _ptr(ts), _ptr(rows), _ptr(offs)))
        return ts[:n_rows], rows[:n_rows], offs

    def free(self):
        if self.h:
            lib().qe_gpu_asof_free(self.ctx.h, self.h)
        self.h = None


class BatchQuery(C.Structure):
    _fields_ = [("key", C.c_void_p), ("value", C.c_void_p), ("nmorsels", C.c_uint32),
                ("morsel_offsets", C.c_void_p), ("naggs", C.c_uint32), ("aggs", C.c_void_p),
                ("group_bound", C.c_uint64)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_rows", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("lds_slots", C.c_uint32)]


TUNE_JOIN_PAIRS, TUNE_ASOF_BUCKET_KNOTS, TUNE_JOIN_SPILL_CAP, TUNE_JOIN_BLOCKS_PER_CU = 0, 1, 2, 3
TUNE_ASOF_DATA_LINES = 4
CMP_LT, CMP_LE, CMP_GT, CMP_GE, CMP_EQ, CMP_NE = range(6)
BATCH_PERSISTENT, BATCH_PER_BATCH, BATCH_GRAPH = 0, 1, 2
BATCH_KINDS = {0: "small", 1: "part", 2: "hash", 3: "part_pass"}


class BatchPlan:
    """Prepared batch of independent GROUP BY queries (qe_gpu_batch_*)."""

    def __init__(self, ctx: "Context", queries):
        """queries: list of (key DevCol, value DevCol, morsel_offsets ndarray, aggs, group_bound)."""
        self.ctx = ctx
        self._keep = []
        arr = (BatchQuery * len(queries))()
        for i, (k, v, offs, aggs, bound) in enumerate(queries):
            o = np.ascontiguousarray(offs, dtype=np.uint64)
            a = _aggs(aggs)
            self._keep += [o, a]
            arr[i] = BatchQuery(k.h, v.h, len(o) - 1, o.ctypes.data, len(aggs), C.cast(a, C.c_void_p), bound)
        self.arr = arr
        self.nq = len(queries)
        h = P()
        ctx._check(lib().qe_gpu_batch_prepare(ctx.h, arr, self.nq, C.byref(h)))
        self.h = h

    def run(self, mode=BATCH_PERSISTENT, tuning=None, collect=True):
        outs = (P * self.nq)() if collect else None
        t = None if tuning is None else C.byref(BatchTuning(*tuning))
        self.ctx._check(lib().qe_gpu_batch_run(self.ctx.h, self.h, self.arr, mode, t, outs))
        return [Table(self.ctx, outs[i]) for i in range(self.nq)] if collect else None

    def autotune(self):
        t = BatchTuning()
        self.ctx._check(lib().qe_gpu_batch_autotune(self.ctx.h, self.h, C.byref(t)))
        return (t.tile_rows, t.blocks_per_cu, t.lds_slots)

    def strategies(self):
        """Per query: (strategy name, partitions), the strategy prepare chose."""
        kinds = (C.c_int32 * self.nq)()
        nparts = (C.c_uint32 * self.nq)()
        self.ctx._check(lib().qe_gpu_batch_strategies(self.ctx.h, self.h, kinds, nparts))
        return [(BATCH_KINDS[kinds[i]], nparts[i]) for i in range(self.nq)]

    def free(self):
        if self.h:
            lib().qe_gpu_batch_free(self.ctx.h, self.h)
        self.h = None


def exported_symbols() -> list[str]:
    """Every qe_gpu_* function name declared in include/qe_gpu.h."""
    import re
    hdr = (PKG_DIR.parent / "include" / "qe_gpu.h").read_text()
    return sorted(set(re.findall(r"\b(qe_gpu_[a-z0-9_]+)\s*\(", hdr)))
