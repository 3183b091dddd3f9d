"""The tick pipeline end to end through the C ABI, against the CPU oracle.

trades WHERE px > 100 -> backward as-of join to quotes on the symbol (binary
and learned index) -> join to a symbol table (I32 id, dictionary-encoded name)
-> GROUP BY name COUNT(*), AVG(px) -> ORDER BY the count descending. The
oracle runs the same steps (filter_select, asof_pairs, join_aggregate); group
rows are exact except AVG (1e-12 relative), and the device sort must order the
counts descending and keep every row.
"""
import numpy as np
import pytest

import _oracle as O
from _compare import assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nsym,nq,nt", [(300, 1 << 16, 1 << 14), (5000, 1 << 20, 1 << 18)])
def test_tick_pipeline_matches_oracle(ctx, qe, nsym, nq, nt):
    r = np.random.default_rng(nsym)
    (qs, qt, qp), (ts_, tt) = ctx.gen_asof(77, nq, nt, nsym, 1.1)
    tp = ctx.gen_uniform(77, 5, qe.F64, nt, 0.0, 200.0, null_frac=0.01)
    names = sorted({"".join(r.choice(list("ABCDEFGHIJ"), int(r.integers(3, 13)))) for _ in range(3 * nsym)})[:nsym]
    d = ctx.dict(list(r.permutation(names)))
    code = {x: i for i, x in enumerate(d.strings)}
    order = r.permutation(nsym)
    sym_id = ctx.upload(np.arange(nsym, dtype=np.int32), qe.I32)
    sym_name = ctx.upload(np.array([code[names[i]] for i in order], dtype=np.int32), qe.STR, dict=d)
    try:
        host = [c.numpy() for c in (qs, qt, ts_, tt, tp)]
        results = []
        for mode in (qe.ASOF_BINARY, qe.ASOF_LEARNED):
            f = ctx.filter([ts_, tt, tp], [(2, qe.CMP_GT, qe.F64, 100.0)])
            idx = ctx.asof_build([qs], qt)
            if mode == qe.ASOF_LEARNED:
                idx.train(64, 100, 0.003)
            match = qe.DevBuffer(ctx, 4 * max(f.rows, 1))
            idx.probe_into([f.col(0)], f.col(1), match, mode=mode)
            pairs = ctx.asof_pairs(match, f.rows, qe.INNER)
            g = pairs.gather([f.col(0), f.col(2)], [])
            agg = ctx.join_aggregate([g.col(0), g.col(1)], [sym_id, sym_name], [0], [0], qe.INNER, [3],
                                     [(qe.COUNT_STAR, 0), (qe.AVG, 1)])
            perm = qe.DevBuffer(ctx, 4 * max(agg.rows, 1))
            ctx.sort_perm_into([agg.col(1)], perm, desc=[1])
            out = ctx.take([agg.col(0), agg.col(1), agg.col(2)], perm, agg.rows)
            a, av = out.col(2).numpy()
            rows = list(zip(out.col(0).strings(), out.col(1).numpy()[0].tolist(),
                            [None if (av is not None and not av[i]) else x for i, x in enumerate(a.tolist())]))
            results.append(rows)
            for x in (out, perm, agg, g, pairs, match, idx, f):
                x.free()
        # the learned index returns the binary-search matches: same groups and
        # counts; AVG differs only by summation order (LDS atomics)
        assert_same(results[1], results[0])
        got = results[0]
        counts = [c for _, c, _ in got]
        assert all(x >= y for x, y in zip(counts, counts[1:]))
        (qsym, _), (qtsv, _), (tsym, _), (tts, _), (tpx, tpv) = host
        keep = O.filter_select([(tpx, tpv, O.F64)], [(0, O.GT, O.F64, 100.0)])
        fs, ft, fp = tsym[keep], tts[keep], tpx[keep]
        fpv = None if tpv is None else tpv[keep]
        p_idx, _ = O.asof_pairs([(fs, None, O.I32), (ft, None, O.I64)], [(qsym, None, O.I32), (qtsv, None, O.I64)],
                                [0], [0], 1, 1, 0)
        sym_tab = [(np.arange(nsym, dtype=np.int32), None, O.I32),
                   (np.array([code[names[i]] for i in order], dtype=np.int32), None, O.STR, d.strings)]
        want = O.join_aggregate([(fs[p_idx], None, O.I32), (fp[p_idx], None if fpv is None else fpv[p_idx], O.F64)],
                                sym_tab, [0], [0], 0, [3], [(O.COUNT_STAR, 0), (O.AVG, 1)])
        assert_same(got, want)
    finally:
        for c in (qs, qt, qp, ts_, tt, tp, sym_id, sym_name):
            c.free()
        d.free()


def test_tick_shape_groupby_by_name_takes_the_hot_set_path(ctx, qe):
    # The pipeline's last join -> group-by at the tick shape: 2^22 + 3 trades
    # whose symbols follow Zipf(1.1) over 50,000, joined to a 50,000-row symbol
    # table with dictionary-encoded names (3-12 characters), GROUP BY name
    # COUNT(*), AVG(px). 50,000 groups do not fit LDS, but the sampled hot set
    # (about 92% of the rows on the top LDS range of names) does: one probe pass
    # keeps it in LDS and the other rows are folded from the spill lists. Keys
    # and counts exact, AVG within 1e-12 of the oracle's Neumaier sums.
    nsym, nt = 50000, (1 << 22) + 3
    r = np.random.default_rng(4242)
    (qs, qt, qp), (ts_, tt) = ctx.gen_asof(4242, 1 << 16, nt, nsym, 1.1)
    tp = ctx.gen_uniform(4242, 5, qe.F64, nt, 0.0, 200.0)
    names = sorted({"".join(r.choice(list("ABCDEFGHIJ"), int(r.integers(3, 13)))) for _ in range(3 * nsym)})[:nsym]
    d = ctx.dict(list(r.permutation(names)))
    strings = list(d.strings)
    code = {x: i for i, x in enumerate(strings)}
    order = r.permutation(nsym)
    name_codes = np.array([code[names[i]] for i in order], dtype=np.int32)
    sym_id = ctx.upload(np.arange(nsym, dtype=np.int32), qe.I32)
    sym_name = ctx.upload(name_codes, qe.STR, dict=d)
    ctx.set_timing(True)
    try:
        for name in ("join_agg_hot_set", "join_agg_spill"):
            ctx.timing(name)
        agg = ctx.join_aggregate([ts_, tp], [sym_id, sym_name], [0], [0], qe.INNER, [3],
                                 [(qe.COUNT_STAR, 0), (qe.AVG, 1)])
        a, av = agg.col(2).numpy()
        got = list(zip(agg.col(0).strings(), agg.col(1).numpy()[0].tolist(),
                       [None if (av is not None and not av[i]) else x for i, x in enumerate(a.tolist())]))
        agg.free()
        _, sampled = ctx.timing("join_agg_hot_set")
        _, spill_passes = ctx.timing("join_agg_spill")
        tsym, _ = ts_.numpy()
        tpx, _ = tp.numpy()
    finally:
        ctx.set_timing(False)
        for c in (qs, qt, qp, ts_, tt, tp, sym_id, sym_name):
            c.free()
        d.free()
    assert sampled == 1 and spill_passes >= 1
    assert len(got) > 40000 and sum(c for _, c, _ in got) == nt
    sym_tab = [(np.arange(nsym, dtype=np.int32), None, O.I32), (name_codes, None, O.STR, strings)]
    want = O.join_aggregate([(tsym, None, O.I32), (tpx, None, O.F64)], sym_tab, [0], [0], 0, [3],
                            [(O.COUNT_STAR, 0), (O.AVG, 1)])
    assert_same(got, want)
