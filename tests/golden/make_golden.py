"""Writes the golden fixtures in this directory from the COMPILED REFERENCE
operators (oracle/_ref/libqe_ref.so, built by `make -C oracle ref` against the
reference checkout). Run here, where the reference exists; the JSON files are
committed so every other machine checks against the reference's own outputs.

Each fixture holds small seeded inputs and the reference's outputs:
  * hashtable   : first-seen ids of HashTable::insert_or_find (cap 8, LF 0.9)
  * groupby     : Aggregate(Scan) rows, F64 sums folded in row order
  * join_groupby: Aggregate(HashJoin(Scan, Scan)) rows
  * asof        : AsofJoin(Scan, Scan) (probe_row, build_row) pairs
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import _oracle as O  # noqa: E402

I32, I64, F64, TS = O.I32, O.I64, O.F64, O.TS


def spec(vals, t, nulls=()):
    v = [float(x) if t == F64 else int(x) for x in vals]
    d = {"type": t, "values": v}
    if len(nulls):
        d["nulls"] = [int(i) for i in nulls]
    return d


def as_col(s):
    valid = None
    if s.get("nulls"):
        valid = np.ones(len(s["values"]), dtype=bool)
        valid[s["nulls"]] = False
    return (np.array(s["values"], dtype=O.NP[s["type"]]), valid, s["type"])


def main():
    assert O.ref_available(), "build the reference harness first: make -C oracle ref"
    r = np.random.default_rng(20260614)
    out = {}

    # hashtable: collision storm + NULL keys under both policies.
    for pol in (0, 1):
        n = 3000
        keys = r.integers(0, 700, n)
        nulls = np.nonzero(r.random(n) < 0.02)[0]
        cols = [spec(keys, I64, nulls)]
        t = O.OracleHT([I64], policy=pol, init_cap=8, lf=0.9, use_ref=True)
        ids = t.insert_or_find([as_col(c) for c in cols]).tolist()
        out[f"hashtable_policy{pol}"] = {"kind": "hashtable", "types": [I64], "policy": pol, "init_cap": 8,
                                         "lf": 0.9, "cols": cols, "expected_ids": ids}

    # group-by: 4096 rows, I64 key in [0,64), F64 with 5% NULLs; COUNT(*)/SUM/AVG + MIN/MAX.
    n = 4096
    key = r.integers(0, 64, n)
    val = np.round(r.uniform(0, 1000, n), 3)
    vn = np.nonzero(r.random(n) < 0.05)[0]
    cols = [spec(key, I64), spec(val, F64, vn)]
    aggs = [(O.COUNT_STAR, 0), (O.SUM, 1), (O.AVG, 1), (O.MIN, 1), (O.MAX, 1), (O.COUNT, 1)]
    rows = O.ref_aggregate([as_col(c) for c in cols], [0], aggs)
    out["groupby_4096"] = {"kind": "groupby", "cols": cols, "keys": [0], "aggs": aggs, "expected": rows}

    # join + group-by: build 512 unique keys (permutation), probe 8192 skewed keys.
    nb, np_ = 512, 8192
    bkey = r.permutation(1024)[:nb]
    bgrp = r.integers(0, 37, nb)
    pkey = np.minimum(r.zipf(1.3, np_) - 1, 1023)
    pay = np.round(r.uniform(0, 1000, np_), 3)
    probe = [spec(pkey, I64), spec(pay, F64)]
    build = [spec(bkey, I64), spec(bgrp, I32)]
    jaggs = [(O.SUM, 1), (O.COUNT_STAR, 0)]
    jrows = O.ref_join_aggregate([as_col(c) for c in probe], [as_col(c) for c in build], [0], [0], 0, [3], jaggs)
    out["join_groupby_8192x512"] = {"kind": "join_groupby", "probe": probe, "build": build, "pkeys": [0],
                                    "bkeys": [0], "gkeys": [3], "aggs": jaggs, "expected": jrows}

    # as-of: 40 symbols, unique per-symbol times, LEFT join.
    nb, np_ = 3000, 3000
    bsym = r.integers(0, 40, nb)
    bts = r.permutation(nb * 4)[:nb]
    psym = r.integers(0, 42, np_)
    pts = r.integers(-10, nb * 4 + 10, np_)
    build = [spec(bsym, I32), spec(bts, TS)]
    probe = [spec(psym, I32), spec(pts, TS)]
    p, b = O.ref_asof_pairs([as_col(c) for c in probe], [as_col(c) for c in build], [0], [0], 1, 1, 1)
    out["asof_left_3000"] = {"kind": "asof", "probe": probe, "build": build, "pkeys": [0], "bkeys": [0],
                             "ptime": 1, "btime": 1, "join_type": 1,
                             "expected_pairs": [[int(x), int(y)] for x, y in zip(p, b)]}

    for name, fx in out.items():
        (HERE / f"{name}.json").write_text(json.dumps(fx, separators=(",", ":")))
        print(f"wrote {name}.json")


if __name__ == "__main__":
    main()
