"""The committed golden fixtures (tests/golden/, written by the compiled
reference through tests/golden/make_golden.py) run through the HIP path: the
reference's own outputs, not the restatement, are the expected values. Group
keys, COUNTs, hash ids and as-of pairs are exact; F64 SUM/AVG within 1e-12
relative of the reference's row-order sums (summation order differs)."""
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _fixture(name):
    return json.loads((GOLDEN / name).read_text())


def _upload(ctx, spec):
    out = []
    for c in spec:
        vals = np.array(c["values"], dtype=O.NP[c["type"]])
        valid = None
        if c.get("nulls"):
            valid = np.ones(len(vals), dtype=bool)
            valid[list(c["nulls"])] = False
        out.append(ctx.upload(vals, c["type"], valid))
    return out


def _free(cols):
    for c in cols:
        c.free()


def test_groupby_fixture(ctx):
    fx = _fixture("groupby_4096.json")
    cols = _upload(ctx, fx["cols"])
    try:
        t = ctx.aggregate(cols, fx["keys"], [tuple(a) for a in fx["aggs"]])
        got = table_rows(t.numpy())
        t.free()
    finally:
        _free(cols)
    assert_same(got, [tuple(r) for r in fx["expected"]])


def test_join_groupby_fixture(ctx, qe):
    fx = _fixture("join_groupby_8192x512.json")
    p, b = _upload(ctx, fx["probe"]), _upload(ctx, fx["build"])
    try:
        t = ctx.join_aggregate(p, b, fx["pkeys"], fx["bkeys"], qe.INNER, fx["gkeys"],
                               [tuple(a) for a in fx["aggs"]])
        got = table_rows(t.numpy())
        t.free()
    finally:
        _free(p + b)
    assert_same(got, [tuple(r) for r in fx["expected"]])


@pytest.mark.parametrize("mode", ["binary", "learned"])
def test_asof_left_fixture(ctx, qe, mode):
    fx = _fixture("asof_left_3000.json")
    assert fx["join_type"] == 1
    p, b = _upload(ctx, fx["probe"]), _upload(ctx, fx["build"])
    try:
        idx = ctx.asof_build([b[k] for k in fx["bkeys"]], b[fx["btime"]])
        if mode == "learned":
            idx.train()
        m = idx.probe([p[k] for k in fx["pkeys"]], p[fx["ptime"]],
                      mode=qe.ASOF_LEARNED if mode == "learned" else qe.ASOF_BINARY)
        idx.free()
    finally:
        _free(p + b)
    got = sorted(zip(range(len(m)), m.tolist()))
    assert got == sorted(tuple(x) for x in fx["expected_pairs"])


@pytest.mark.parametrize("name", ["hashtable_policy0.json", "hashtable_policy1.json"])
def test_hashtable_fixture(ctx, name):
    fx = _fixture(name)
    cols = _upload(ctx, fx["cols"])
    ht = ctx.hash_table(fx["types"], null_policy=fx["policy"], initial_capacity=fx["init_cap"],
                        max_load_factor=fx["lf"])
    try:
        ids = ht.insert_or_find(cols)
    finally:
        ht.free()
        _free(cols)
    assert ids.tolist() == fx["expected_ids"]
