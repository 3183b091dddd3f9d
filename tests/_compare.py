"""Result-set comparison with the reference comparator's semantics
(oracle/result_set.cpp:129-187 in the reference checkout): unordered results
are canonicalised by sorting rows on every column with NULLs first and F64 NaN
greatest; integers and nullness compare exactly; F64 compares within an
epsilon that is a parameter (the reference uses 1e-9 abs + 1e-9 rel; the GPU
bar here is 1e-12 relative, summation order being the only permitted
difference, against an accurately summed oracle).
"""
from __future__ import annotations

import math

REL = 1e-12
ABS = 1e-12


def _key(v):
    if v is None:
        return (0, 0.0)
    if isinstance(v, float):
        if math.isnan(v):
            return (2, 0.0)
        return (1, v)
    return (1, v)


def canon(rows):
    return sorted(rows, key=lambda r: tuple(_key(v) for v in r))


def f_eq(a, b, rel=REL, abs_=ABS):
    if isinstance(a, float) and isinstance(b, float):
        if math.isnan(a) or math.isnan(b):
            return math.isnan(a) and math.isnan(b)
        if math.isinf(a) or math.isinf(b):
            return a == b
        return abs(a - b) <= abs_ + rel * max(abs(a), abs(b))
    return a == b


def diff(got, want, rel=REL, abs_=ABS, ordered=False):
    """None if equal, else a message naming the first difference."""
    g = list(got) if ordered else canon(got)
    w = list(want) if ordered else canon(want)
    if len(g) != len(w):
        return f"row count {len(g)} != {len(w)}"
    for i, (rg, rw) in enumerate(zip(g, w)):
        if len(rg) != len(rw):
            return f"row {i}: width {len(rg)} != {len(rw)}"
        for c, (a, b) in enumerate(zip(rg, rw)):
            if (a is None) != (b is None):
                return f"row {i} col {c}: nullness {a!r} vs {b!r}"
            if a is None:
                continue
            if isinstance(a, float) or isinstance(b, float):
                if not f_eq(float(a), float(b), rel, abs_):
                    return f"row {i} col {c}: {a!r} vs {b!r}"
            elif a != b:
                return f"row {i} col {c}: {a!r} vs {b!r}"
    return None


def assert_same(got, want, rel=REL, abs_=ABS, ordered=False):
    d = diff(got, want, rel, abs_, ordered)
    assert d is None, d


def table_rows(cols):
    """[(values, valid)] per column (GPU table numpy) -> list of row tuples."""
    n = len(cols[0][0]) if cols else 0
    out = []
    pyc = []
    for vals, valid in cols:
        lst = vals.tolist()
        if valid is not None:
            lst = [x if ok else None for x, ok in zip(lst, valid.tolist())]
        pyc.append(lst)
    for i in range(n):
        out.append(tuple(c[i] for c in pyc))
    return out
