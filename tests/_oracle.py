"""ctypes access to the CPU oracle (oracle/_build/libqe_oracle.so) and, where it
was built, the compiled reference harness (oracle/_ref/libqe_ref.so).

Test infrastructure: the checker the GPU path is compared against. Columns are
passed as (values ndarray, valid ndarray|None, type) triples; a STR column is a
4-tuple whose last item is its dictionary (a list of strings the int32 codes
index).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "_build" / "libqe_oracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libqe_ref.so"

I32, I64, F64, BOOL, TS, STR = range(6)
NP = {I32: np.int32, I64: np.int64, F64: np.float64, BOOL: np.uint8, TS: np.int64, STR: np.int32}
COUNT_STAR, COUNT, SUM, MIN, MAX, AVG = range(6)
NO_ROW = 0xFFFFFFFF


class OrcCol(C.Structure):
    _fields_ = [("type", C.c_int32), ("all_valid", C.c_int32), ("len", C.c_uint64),
                ("data", C.c_void_p), ("validity", C.c_void_p), ("dict_bytes", C.c_void_p),
                ("dict_off", C.c_void_p), ("dict_n", C.c_uint64)]


def dict_blob(strings):
    """(bytes ndarray, offsets ndarray) in StringDict's layout."""
    enc = [x.encode() for x in strings]
    off = np.zeros(len(enc) + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(e) for e in enc]) if enc else []
    blob = np.frombuffer(b"".join(enc) + b"\0", dtype=np.uint8).copy()
    return blob, off


P = C.c_void_p
_oracle = None
_ref = None


def _pack(valid):
    n = len(valid)
    nw = (n + 63) // 64
    padded = np.ones(max(nw, 1) * 64, dtype=bool)
    padded[:n] = valid
    return np.packbits(padded, bitorder="little").view(np.uint64).copy()


class Cols:
    """Keeps numpy buffers alive while the C structs point at them."""

    def __init__(self, cols):
        self.keep = []
        arr = (OrcCol * max(len(cols), 1))()
        for i, col in enumerate(cols):
            vals, valid, t = col[:3]
            v = np.ascontiguousarray(vals, dtype=NP[t])
            w = None if valid is None else _pack(np.asarray(valid, dtype=bool))
            self.keep += [v, w]
            db = do = None
            dn = 0
            if t == STR and len(col) > 3:
                db, do = dict_blob(col[3])
                dn = len(col[3])
                self.keep += [db, do]
            arr[i] = OrcCol(t, 1 if valid is None else 0, len(v), v.ctypes.data,
                            None if w is None else w.ctypes.data, None if db is None else db.ctypes.data,
                            None if do is None else do.ctypes.data, dn)
        self.arr = arr
        self.n = len(cols[0][0]) if cols else 0


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s"], check=True)
        L = C.CDLL(str(ORACLE_SO))
        sig = {
            "orc_ht_new": (P, [P, C.c_uint32, C.c_int32, C.c_uint64, C.c_double, C.c_uint64, C.c_int32]),
            "orc_ht_free": (None, [P]),
            "orc_ht_insert_or_find": (C.c_int, [P, P, P, C.c_uint64, P]),
            "orc_ht_find": (C.c_int, [P, P, P, C.c_uint64, P]),
            "orc_ht_num_groups": (C.c_uint64, [P]),
            "orc_ht_capacity": (C.c_uint64, [P]),
            "orc_ht_key_word": (C.c_uint64, [P, C.c_uint32, C.c_uint32]),
            "orc_ht_key_is_null": (C.c_int32, [P, C.c_uint32, C.c_uint32]),
            "orc_mix64": (C.c_uint64, [C.c_uint64]),
            "orc_agg_run": (P, [P, C.c_uint64, P, C.c_uint32, P, P, C.c_uint32, C.c_int32, C.c_int32]),
            "orc_agg_ngroups": (C.c_uint64, [P]),
            "orc_agg_keys": (None, [P, C.c_uint32, P, P]),
            "orc_agg_is_float": (C.c_int32, [P, C.c_uint32]),
            "orc_agg_values": (None, [P, C.c_uint32, P, P, P]),
            "orc_agg_free": (None, [P]),
            "orc_join_run": (P, [P, C.c_uint64, P, C.c_uint64, P, P, C.c_uint32, C.c_int32, C.c_int32]),
            "orc_pairs_count": (C.c_uint64, [P]),
            "orc_pairs_get": (None, [P, P, P]),
            "orc_pairs_free": (None, [P]),
            "orc_join_agg_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, C.c_uint64, P, P, C.c_uint32,
                                     C.c_int32, P, C.c_uint32, P, P, C.c_uint32, C.c_int32, C.c_int32]),
            "orc_asof_run": (P, [P, C.c_uint64, P, C.c_uint64, P, P, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_int32, C.c_int32, C.c_int64, C.c_int32]),
            "orc_sort_perm": (C.c_int, [P, P, P, P, C.c_uint32, C.c_uint64, P]),
            "orc_bl_groupby_i64_f64": (C.c_int64, [P, P, P, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                                   P, P, P, P]),
            "orc_bl_join_groupby": (C.c_int64, [P, P, C.c_uint64, P, P, C.c_uint64, C.c_uint32, C.c_uint64,
                                                C.c_uint64, P, P, P]),
            "orc_bl_asof_upper_bound": (None, [P, P, C.c_uint32, P, P, C.c_uint64, C.c_uint32, P]),
            "orc_bl_hw_threads": (C.c_uint32, []),
            "orc_filter_select": (C.c_uint64, [P, C.c_uint64, P, C.c_uint32, P, C.c_int32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _oracle = L
    return _oracle


def ref_available() -> bool:
    return REF_SO.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_SO))
        sig = {
            "ref_ht_new": (P, [P, C.c_uint32, C.c_int32, C.c_uint64, C.c_double, C.c_uint64]),
            "ref_ht_free": (None, [P]),
            "ref_ht_insert_or_find": (C.c_int, [P, P, P, C.c_uint64, P]),
            "ref_ht_find": (C.c_int, [P, P, P, C.c_uint64, P]),
            "ref_ht_num_groups": (C.c_uint64, [P]),
            "ref_ht_key_word": (C.c_uint64, [P, C.c_uint32, C.c_uint32]),
            "ref_ht_key_is_null": (C.c_int32, [P, C.c_uint32, C.c_uint32]),
            "ref_rs_nrows": (C.c_uint64, [P]),
            "ref_rs_ncols": (C.c_uint32, [P]),
            "ref_rs_type": (C.c_int32, [P, C.c_uint32]),
            "ref_rs_col": (None, [P, C.c_uint32, P, P, P]),
            "ref_rs_free": (None, [P]),
            "ref_rs_str_bytes": (C.c_uint64, [P, C.c_uint32]),
            "ref_rs_str_get": (None, [P, C.c_uint32, P, P]),
            "ref_agg_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, P, P, C.c_uint32, C.c_uint64]),
            "ref_join_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, C.c_uint64, P, P, C.c_uint32,
                                 C.c_int32, C.c_uint64]),
            "ref_join_agg_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, C.c_uint64, P, P, C.c_uint32,
                                     C.c_int32, P, C.c_uint32, P, P, C.c_uint32, C.c_uint64]),
            "ref_asof_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, C.c_uint64, P, P, C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.c_int64, C.c_uint64]),
            "ref_filter_run": (P, [P, C.c_uint32, C.c_uint64, P, C.c_uint32, C.c_uint64]),
            "ref_sort_run": (P, [P, C.c_uint32, C.c_uint64, P, P, P, C.c_uint32, C.c_uint64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _ref = L
    return _ref


def _u32(xs):
    return (C.c_uint32 * max(len(xs), 1))(*xs)


def _i32(xs):
    return (C.c_int32 * max(len(xs), 1))(*xs)


# ---- hash table ----------------------------------------------------------------
class OracleHT:
    def __init__(self, types, policy=0, init_cap=1024, lf=0.7, seed=0x51E57A11AB1E5EED, mutant=0, use_ref=False):
        self.use_ref = use_ref
        self.L = ref() if use_ref else oracle()
        self.ncols = len(types)
        if use_ref:
            self.h = self.L.ref_ht_new(_i32(types), len(types), policy, init_cap, lf, seed)
        else:
            self.h = self.L.orc_ht_new(_i32(types), len(types), policy, init_cap, lf, seed, mutant)

    def _call(self, fn, cols, sel):
        cs = Cols(cols)
        n = len(sel) if sel is not None else cs.n
        out = np.empty(max(n, 1), dtype=np.uint32)
        s = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint32)
        rc = fn(self.h, cs.arr, None if s is None else s.ctypes.data, n, out.ctypes.data)
        if rc != 0:
            raise OverflowError("group id space exhausted")
        return out[:n]

    def insert_or_find(self, cols, sel=None):
        return self._call(self.L.ref_ht_insert_or_find if self.use_ref else self.L.orc_ht_insert_or_find, cols, sel)

    def find(self, cols, sel=None):
        return self._call(self.L.ref_ht_find if self.use_ref else self.L.orc_ht_find, cols, sel)

    @property
    def num_groups(self):
        return int(self.L.ref_ht_num_groups(self.h) if self.use_ref else self.L.orc_ht_num_groups(self.h))

    def key_word(self, g, col):
        return int(self.L.ref_ht_key_word(self.h, g, col) if self.use_ref else self.L.orc_ht_key_word(self.h, g, col))

    def key_is_null(self, g, col):
        return bool(self.L.ref_ht_key_is_null(self.h, g, col) if self.use_ref
                    else self.L.orc_ht_key_is_null(self.h, g, col))

    def __del__(self):
        try:
            (self.L.ref_ht_free if self.use_ref else self.L.orc_ht_free)(self.h)
        except Exception:
            pass


# ---- result rows -----------------------------------------------------------------
def _decode_key(word, t):
    if t in (I32, STR):
        return int(np.array([word & 0xFFFFFFFF], dtype=np.uint32).view(np.int32)[0])
    if t == BOOL:
        return int(word & 1)
    if t == F64:
        return float(np.array([word], dtype=np.uint64).view(np.float64)[0])
    return int(np.array([word], dtype=np.uint64).view(np.int64)[0])


def aggregate(cols, keys, aggs, sum_mode=1, mutant=0):
    """Group-by rows [(key..., agg...)] in first-seen order; None for NULL; STR
    keys and STR MIN/MAX as strings."""
    L = oracle()
    cs = Cols(cols)
    funcs = _i32([f for f, _ in aggs])
    ins = _u32([c for _, c in aggs])
    h = L.orc_agg_run(cs.arr, cs.n, _u32(keys), len(keys), funcs, ins, len(aggs), sum_mode, mutant)
    try:
        rows = _agg_rows(L, h, [cols[k][2] for k in keys], len(aggs))
    finally:
        L.orc_agg_free(h)
    return _decode_str(rows, [cols[k] for k in keys], [(f, cols[c] if f not in (COUNT_STAR, COUNT) else None)
                                                       for f, c in aggs])


def _decode_str(rows, key_cols, agg_inputs):
    """Codes -> strings for STR key columns and STR MIN/MAX outputs."""
    dicts = [c[3] if (c[2] == STR and len(c) > 3) else None for c in key_cols]
    dicts += [c[3] if (c is not None and f in (MIN, MAX) and c[2] == STR and len(c) > 3) else None
              for f, c in agg_inputs]
    if not any(d is not None for d in dicts):
        return rows
    return [tuple(v if (d is None or v is None) else d[v] for v, d in zip(r, dicts)) for r in rows]


def _agg_rows(L, h, key_types, naggs):
    g = int(L.orc_agg_ngroups(h))
    colvals = []
    for j, t in enumerate(key_types):
        w = np.empty(max(g, 1), dtype=np.uint64)
        nl = np.empty(max(g, 1), dtype=np.uint8)
        L.orc_agg_keys(h, j, w.ctypes.data, nl.ctypes.data)
        colvals.append([None if nl[i] else _decode_key(int(w[i]), t) for i in range(g)])
    for a in range(naggs):
        iv = np.empty(max(g, 1), dtype=np.int64)
        dv = np.empty(max(g, 1), dtype=np.float64)
        nl = np.empty(max(g, 1), dtype=np.uint8)
        L.orc_agg_values(h, a, iv.ctypes.data, dv.ctypes.data, nl.ctypes.data)
        fl = L.orc_agg_is_float(h, a)
        colvals.append([None if nl[i] else (float(dv[i]) if fl else int(iv[i])) for i in range(g)])
    return [tuple(c[i] for c in colvals) for i in range(g)]


def join_pairs(probe_keys, build_keys, join_type=0, mutant=0):
    L = oracle()
    pc, bc = Cols(probe_keys), Cols(build_keys)
    nk = len(probe_keys)
    h = L.orc_join_run(pc.arr, pc.n, bc.arr, bc.n, _u32(range(nk)), _u32(range(nk)), nk, join_type, mutant)
    return _pairs(L, h)


def _pairs(L, h):
    try:
        n = int(L.orc_pairs_count(h))
        p = np.empty(max(n, 1), dtype=np.uint32)
        b = np.empty(max(n, 1), dtype=np.uint32)
        L.orc_pairs_get(h, p.ctypes.data, b.ctypes.data)
        return p[:n], b[:n]
    finally:
        L.orc_pairs_free(h)


def join_aggregate(probe_cols, build_cols, pkeys, bkeys, join_type, gkeys, aggs, sum_mode=1, mutant=0):
    L = oracle()
    pc, bc = Cols(probe_cols), Cols(build_cols)
    allcols = list(probe_cols) + list(build_cols)
    h = L.orc_join_agg_run(pc.arr, len(probe_cols), pc.n, bc.arr, len(build_cols), bc.n, _u32(pkeys),
                           _u32(bkeys), len(pkeys), join_type, _u32(gkeys), len(gkeys),
                           _i32([f for f, _ in aggs]), _u32([c for _, c in aggs]), len(aggs), sum_mode, mutant)
    try:
        rows = _agg_rows(L, h, [allcols[k][2] for k in gkeys], len(aggs))
    finally:
        L.orc_agg_free(h)
    return _decode_str(rows, [allcols[k] for k in gkeys],
                       [(f, allcols[c] if f not in (COUNT_STAR, COUNT) else None) for f, c in aggs])


def asof_pairs(probe_cols, build_cols, pkeys, bkeys, ptime, btime, join_type=0, tolerance=None, mutant=0):
    L = oracle()
    pc, bc = Cols(probe_cols), Cols(build_cols)
    h = L.orc_asof_run(pc.arr, pc.n, bc.arr, bc.n, _u32(pkeys), _u32(bkeys), len(pkeys), ptime, btime,
                       join_type, 0 if tolerance is None else 1, 0 if tolerance is None else tolerance, mutant)
    return _pairs(L, h)


# ---- filter ----------------------------------------------------------------------
LT, LE, GT, GE, EQ, NE = range(6)


class OrcPred(C.Structure):
    _fields_ = [("col", C.c_uint32), ("op", C.c_int32), ("lit_type", C.c_int32), ("lit_is_null", C.c_int32),
                ("lit_i", C.c_int64), ("lit_f", C.c_double)]


def _preds(terms):
    """terms: [(col, op, lit_type, value)] with value None for a NULL literal."""
    arr = (OrcPred * max(len(terms), 1))()
    for j, (c, op, lt, v) in enumerate(terms):
        isnull = v is None
        arr[j] = OrcPred(c, op, lt, 1 if isnull else 0, 0 if (isnull or lt == F64) else int(v),
                         float(v) if (lt == F64 and not isnull) else 0.0)
    return arr


def filter_select(cols, terms, mutant=0):
    """Passing row indices of a conjunction of column-vs-literal comparisons."""
    L = oracle()
    cs = Cols(cols)
    sel = np.empty(max(cs.n, 1), dtype=np.uint32)
    m = L.orc_filter_select(cs.arr, cs.n, _preds(terms), len(terms), sel.ctypes.data, mutant)
    return sel[:m].copy()


def sort_perm(cols, keys, desc=None, nulls_first=None):
    L = oracle()
    cs = Cols(cols)
    k = len(keys)
    perm = np.empty(max(cs.n, 1), dtype=np.uint32)
    rc = L.orc_sort_perm(cs.arr, _u32(keys), _i32(desc or [0] * k), _i32(nulls_first or [0] * k), k, cs.n,
                         perm.ctypes.data)
    assert rc == 0
    return perm[:cs.n]


# ---- compiled reference (this container only) ------------------------------------
def _ref_rows(h):
    L = ref()
    try:
        n, nc = int(L.ref_rs_nrows(h)), int(L.ref_rs_ncols(h))
        cols = []
        for c in range(nc):
            iv = np.empty(max(n, 1), dtype=np.int64)
            dv = np.empty(max(n, 1), dtype=np.float64)
            nl = np.empty(max(n, 1), dtype=np.uint8)
            L.ref_rs_col(h, c, iv.ctypes.data, dv.ctypes.data, nl.ctypes.data)
            ty = L.ref_rs_type(h, c)
            if ty == STR:
                nb = int(L.ref_rs_str_bytes(h, c))
                blob = np.empty(max(nb, 1), dtype=np.uint8)
                offs = np.empty(n + 1, dtype=np.uint64)
                L.ref_rs_str_get(h, c, blob.ctypes.data, offs.ctypes.data)
                raw = blob.tobytes()
                cols.append([None if nl[i] else raw[offs[i]:offs[i + 1]].decode() for i in range(n)])
                continue
            fl = ty == F64
            cols.append([None if nl[i] else (float(dv[i]) if fl else int(iv[i])) for i in range(n)])
        return [tuple(col[i] for col in cols) for i in range(n)]
    finally:
        L.ref_rs_free(h)


def ref_aggregate(cols, keys, aggs, batch_size=2048):
    cs = Cols(cols)
    h = ref().ref_agg_run(cs.arr, len(cols), cs.n, _u32(keys), len(keys), _i32([f for f, _ in aggs]),
                          _u32([c for _, c in aggs]), len(aggs), batch_size)
    return _ref_rows(h)


def ref_join_pairs(probe_cols, build_cols, pkeys, bkeys, join_type=0, batch_size=2048):
    pc, bc = Cols(probe_cols), Cols(build_cols)
    h = ref().ref_join_run(pc.arr, len(probe_cols), pc.n, bc.arr, len(build_cols), bc.n, _u32(pkeys),
                           _u32(bkeys), len(pkeys), join_type, batch_size)
    rows = _ref_rows(h)
    npc, nbc = len(probe_cols), len(build_cols)
    p = np.array([r[npc] for r in rows], dtype=np.int64)
    b = np.array([NO_ROW if r[npc + nbc + 1] is None else r[npc + nbc + 1] for r in rows], dtype=np.int64)
    return p.astype(np.uint32), b.astype(np.uint32)


def ref_join_aggregate(probe_cols, build_cols, pkeys, bkeys, join_type, gkeys, aggs, batch_size=2048):
    pc, bc = Cols(probe_cols), Cols(build_cols)
    h = ref().ref_join_agg_run(pc.arr, len(probe_cols), pc.n, bc.arr, len(build_cols), bc.n, _u32(pkeys),
                               _u32(bkeys), len(pkeys), join_type, _u32(gkeys), len(gkeys),
                               _i32([f for f, _ in aggs]), _u32([c for _, c in aggs]), len(aggs), batch_size)
    return _ref_rows(h)


def ref_asof_pairs(probe_cols, build_cols, pkeys, bkeys, ptime, btime, join_type=0, tolerance=None,
                   batch_size=2048):
    pc, bc = Cols(probe_cols), Cols(build_cols)
    h = ref().ref_asof_run(pc.arr, len(probe_cols), pc.n, bc.arr, len(build_cols), bc.n, _u32(pkeys),
                           _u32(bkeys), len(pkeys), ptime, btime, join_type, 0 if tolerance is None else 1,
                           0 if tolerance is None else tolerance, batch_size)
    rows = _ref_rows(h)
    npc, nbc = len(probe_cols), len(build_cols)
    p = np.array([r[npc] for r in rows], dtype=np.int64)
    b = np.array([NO_ROW if r[npc + nbc + 1] is None else r[npc + nbc + 1] for r in rows], dtype=np.int64)
    return p.astype(np.uint32), b.astype(np.uint32)


def ref_filter_select(cols, terms, batch_size=2048):
    cs = Cols(cols)
    h = ref().ref_filter_run(cs.arr, len(cols), cs.n, _preds(terms), len(terms), batch_size)
    rows = _ref_rows(h)
    return np.array([r[len(cols)] for r in rows], dtype=np.int64).astype(np.uint32)


def ref_sort_perm(cols, keys, desc=None, nulls_first=None, batch_size=2048):
    cs = Cols(cols)
    k = len(keys)
    h = ref().ref_sort_run(cs.arr, len(cols), cs.n, _u32(keys), _i32(desc or [0] * k), _i32(nulls_first or [0] * k),
                           k, batch_size)
    rows = _ref_rows(h)
    return np.array([r[len(cols)] for r in rows], dtype=np.int64).astype(np.uint32)
