"""GPU hash join and fused join -> group-by against the CPU oracle.

Literal edge shapes follow tests/join_differential_test.cpp:103-296 (one key
with a NULL, composite (i32, f64), empty build / probe, all-match, 3000-row
fan-out past the 2048-row output batch, NULL keys on both sides); pair sets
are bit-exact; the fused aggregate is bit-exact for keys and COUNTs and within
1e-12 relative for F64 SUM.
"""
import numpy as np
import pytest

import _device_cases as DC
import _oracle as O
from _compare import assert_same, table_rows

pytestmark = pytest.mark.gpu
I32, I64, F64 = O.I32, O.I64, O.F64
CS, CNT, SUM, MIN, MAX, AVG = O.COUNT_STAR, O.COUNT, O.SUM, O.MIN, O.MAX, O.AVG
INNER, LEFT = 0, 1


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def up(ctx, cols):
    return [ctx.upload(v, t, valid) for v, valid, t in cols]


def gpu_pairs(ctx, pk, bk, jt):
    P, B = up(ctx, pk), up(ctx, bk)
    try:
        pr = ctx.hash_join(P, B, jt)
        p, b = pr.numpy()
        pr.free()
        return p, b
    finally:
        for d in P + B:
            d.free()


def pairset(p, b):
    return sorted(zip(p.tolist(), b.tolist()))


def check_pairs(ctx, pk, bk, jt):
    p, b = gpu_pairs(ctx, pk, bk, jt)
    wp, wb = O.join_pairs(pk, bk, jt)
    assert pairset(p, b) == pairset(wp, wb)
    return pairset(p, b)


@pytest.mark.parametrize("jt", [INNER, LEFT])
def test_edge_single_key(ctx, jt):
    bk = [col([1, 1, 2, 5, 5, 5], I32)]
    pk = [col([1, 2, 3, 5, 0], I32, nulls=(4,))]
    got = check_pairs(ctx, pk, bk, jt)
    want = [(0, 0), (0, 1), (1, 2), (3, 3), (3, 4), (3, 5)]
    if jt == LEFT:
        want += [(2, O.NO_ROW), (4, O.NO_ROW)]
    assert got == sorted(want)


@pytest.mark.parametrize("jt", [INNER, LEFT])
def test_edge_composite_i32_f64(ctx, jt):
    bk = [col([1, 1, 2, 2], I32), col([1.0, 2.0, 1.0, 2.0], F64)]
    pk = [col([1, 1, 2, 2], I32), col([2.0, 9.0, 1.0, 2.0], F64)]
    got = check_pairs(ctx, pk, bk, jt)
    assert (0, 1) in got and (2, 2) in got and (3, 3) in got


@pytest.mark.parametrize("jt", [INNER, LEFT])
def test_edge_empty_sides(ctx, jt):
    got = check_pairs(ctx, [col([1, 2, 3], I32)], [col([], I32)], jt)
    assert got == ([] if jt == INNER else [(i, O.NO_ROW) for i in range(3)])
    assert check_pairs(ctx, [col([], I32)], [col([1, 2, 3], I32)], jt) == []


@pytest.mark.parametrize("jt", [INNER, LEFT])
def test_edge_all_match_and_fanout(ctx, jt):
    got = check_pairs(ctx, [col([7, 7, 7], I32)], [col([7, 7, 7, 7], I32)], jt)
    assert len(got) == 12
    got = check_pairs(ctx, [col([7, 9, 7], I32)], [col([7] * 3000, I32)], jt)
    assert len(got) == 6000 + (1 if jt == LEFT else 0)


@pytest.mark.parametrize("jt", [INNER, LEFT])
def test_edge_null_keys_never_match(ctx, jt):
    got = check_pairs(ctx, [col([1, 0, 0], I32, nulls=(1, 2))], [col([1, 0, 2], I32, nulls=(1,))], jt)
    assert (0, 0) in got and all(b != 1 for _, b in got)


def test_f64_keys_canonical(ctx):
    nan = float("nan")
    got = check_pairs(ctx, [col([0.0, nan, 1.0], F64)], [col([-0.0, nan, 2.0], F64)], INNER)
    assert got == [(0, 0), (1, 1)]


@pytest.mark.parametrize("case", range(24))
def test_random_joins(ctx, case):
    r = np.random.default_rng(500 + case)
    nb, np_ = int(r.choice([0, 1, 100, 2049, 20000])), int(r.choice([0, 1, 100, 4097, 30000]))
    dom = int(r.choice([1, 10, 1000, 10**6, 2**40]))
    nk = int(r.integers(1, 3))
    t = int(r.choice([I32, I64]))
    bp = float(r.choice([0.0, 0.1]))

    def mk(n):
        cols = []
        for _ in range(nk):
            nulls = tuple(np.nonzero(r.random(n) < bp)[0]) if bp else ()
            cols.append(col(r.integers(0, min(dom, 2**31 - 1) if t == I32 else dom, n), t, nulls))
        return cols
    check_pairs(ctx, mk(np_), mk(nb), int(r.integers(0, 2)))


def test_gather_emits_probe_then_build_columns(ctx):
    pk = [col([1, 2, 3], I32)]
    pcols = pk + [col([10.5, 20.5, 30.5], F64, nulls=(1,))]
    bcols = [col([3, 1], I32), col([300, 100], I64)]
    P, B = up(ctx, pcols), up(ctx, bcols)
    try:
        pr = ctx.hash_join(P[:1], B[:1], LEFT)
        t = pr.gather(P, B)
        rows = sorted(table_rows(t.numpy()), key=lambda x: x[0])
        t.free()
        pr.free()
    finally:
        for d in P + B:
            d.free()
    assert rows == [(1, 10.5, 1, 100), (2, None, None, None), (3, 30.5, 3, 300)]


def join_agg_case(ctx, pcols, bcols, pk, bk, jt, gk, aggs):
    P, B = up(ctx, pcols), up(ctx, bcols)
    try:
        t = ctx.join_aggregate(P, B, pk, bk, jt, gk, aggs)
        got = table_rows(t.numpy())
        t.free()
    finally:
        for d in P + B:
            d.free()
    want = O.join_aggregate(pcols, bcols, pk, bk, jt, gk, aggs)
    assert_same(got, want)
    return got


def test_fused_join_groupby_small_config2_shape(ctx):
    # build 2^16 unique keys + I32 group col; probe 2^20 Zipf keys, 10% misses.
    (pk, pp), (bk, bg) = ctx.gen_join(7, 1 << 16, 1 << 20, 10000)
    try:
        cols_p = [pk.numpy()[0], pp.numpy()[0]]
        cols_b = [bk.numpy()[0], bg.numpy()[0]]
        t = ctx.join_aggregate([pk, pp], [bk, bg], [0], [0], INNER, [3], [(SUM, 1), (CS, 0)])
        got = table_rows(t.numpy())
        t.free()
    finally:
        for d in (pk, pp, bk, bg):
            d.free()
    assert len(np.unique(cols_b[0])) == len(cols_b[0])  # unique build keys
    pcols = [(cols_p[0], None, I64), (cols_p[1], None, F64)]
    bcols = [(cols_b[0], None, I64), (cols_b[1], None, I32)]
    want = O.join_aggregate(pcols, bcols, [0], [0], INNER, [3], [(SUM, 1), (CS, 0)])
    assert_same(got, want)
    matched = sum(r[2] for r in got)
    assert 0.85 * (1 << 20) < matched < 0.95 * (1 << 20)


def test_prepared_join_groupby_repeats(ctx):
    # The prepared call (arguments marshalled once) returns the oracle's groups
    # on every repetition, with results freed in between.
    (pk, pp), (bk, bg) = ctx.gen_join(8, 1 << 15, 1 << 18, 3000)
    try:
        pcols = [(pk.numpy()[0], None, I64), (pp.numpy()[0], None, F64)]
        bcols = [(bk.numpy()[0], None, I64), (bg.numpy()[0], None, I32)]
        want = O.join_aggregate(pcols, bcols, [0], [0], INNER, [3], [(SUM, 1), (CS, 0)])
        run = ctx.prepare_join_aggregate([pk, pp], [bk, bg], [0], [0], INNER, [3], [(SUM, 1), (CS, 0)])
        for _ in range(3):
            t = run()
            got = table_rows(t.numpy())
            t.free()
            assert_same(got, want)
    finally:
        for d in (pk, pp, bk, bg):
            d.free()


@pytest.mark.parametrize("pairs", [2, 4])
@pytest.mark.parametrize("case", range(8))
def test_fused_direct_stream_shapes(ctx, qe, pairs, case):
    # Direct unique build map with the streaming probe: odd and tiny probe
    # lengths (tail row), I32 / I64 keys, NULL probe keys, narrow and nullable
    # value columns, 0..3 distinct value columns (3 takes the per-row kernel).
    # Inputs shared with the device models' mutation catalog.
    pcols, bcols = DC.direct_stream_case(case)
    aggsets = [[(CS, 0)],
               [(SUM, 2), (CS, 0)],
               [(SUM, 1), (MIN, 1), (MAX, 1), (CNT, 1), (AVG, 1)],
               [(SUM, 2), (MAX, 2), (MIN, 3), (SUM, 3), (CS, 0)],
               [(SUM, 1), (SUM, 2), (SUM, 3)],
               [(SUM, 3), (CS, 0)]]
    ctx.set_tuning(qe.TUNE_JOIN_PAIRS, pairs)
    try:
        for aggs in aggsets:
            join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [5], aggs)
    finally:
        ctx.set_tuning(qe.TUNE_JOIN_PAIRS, 4)


@pytest.mark.parametrize("case", range(5))
def test_fused_stream_wide_maps(ctx, case):
    """The streaming probe at 2^20+ rows over maps of 2^16..2^17 keys: skewed
    and uniform keys with 10% misses, I32/I64 keys (negative offsets), one or
    two value columns (F64, I64, narrow I32: every width specialisation of the
    stream loads), an odd probe length, and (case 4) each probe tile's keys in
    one 4096-key stretch of the map. Against the oracle; inputs shared with the
    device models' mutation catalog."""
    pcols, bcols, aggs = DC.wide_map_case(case)
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [5], aggs)


def test_dense_map_duplicate_falls_back_with_clean_f64_state(ctx):
    # Dense build keys with one repeated key: the direct map is built and
    # probed, the duplicate is found with the emission, and the query runs
    # again through the row lists. F64 values include outliers on rows the
    # scale sample skips (every 16th row is sampled at 2^20 rows), a NaN and
    # an infinity: the second run must start from an empty outlier list and
    # cleared flags, so nothing is counted twice.
    r = np.random.default_rng(2024)
    n_b, n_p = 50000, 1 << 20
    keys = r.permutation(n_b)
    keys[7] = keys[8]  # one repeated key
    pk = r.integers(0, n_b + 100, n_p)
    pv = r.uniform(0, 1, n_p)
    out_rows = np.arange(1, n_p, 16)[::4000]  # never sampled
    pv[out_rows] = r.uniform(1e6, 1e7, len(out_rows)) * np.where(r.random(len(out_rows)) < 0.5, -1, 1)
    pv[5] = np.nan
    pv[17] = np.inf
    pcols = [col(pk, I64), col(pv, F64)]
    bcols = [col(keys, I64), col(r.integers(0, 300, n_b), I32)]
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [3], [(SUM, 1), (CS, 0), (AVG, 1)])


def test_fused_duplicate_build_keys_list_path(ctx, rng):
    n_b, n_p = 5000, 50000
    pcols = [col(rng.integers(0, 2000, n_p), I64), col(rng.uniform(0, 10, n_p), F64, nulls=(3, 5, 7))]
    bcols = [col(rng.integers(0, 1500, n_b), I64), col(rng.integers(0, 40, n_b), I32, nulls=(1,))]
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [3], [(SUM, 1), (CS, 0), (CNT, 1), (MIN, 1), (AVG, 1)])


def test_fused_hash_index_sparse_keys(ctx, rng):
    keys = rng.integers(-2**60, 2**60, 4000)
    pcols = [col(keys[rng.integers(0, 4000, 40000)], I64), col(rng.uniform(0, 10, 40000), F64)]
    bcols = [col(keys[:3000], I64), col(rng.integers(0, 30, 3000), I32)]
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [3], [(SUM, 1), (CS, 0)])


def test_general_path_left_join_probe_group_key(ctx, rng):
    pcols = [col(rng.integers(0, 50, 3000), I32), col(rng.integers(0, 7, 3000), I32)]
    bcols = [col(rng.integers(0, 40, 200), I32), col(rng.uniform(0, 1, 200), F64)]
    join_agg_case(ctx, pcols, bcols, [0], [0], LEFT, [1], [(CS, 0), (SUM, 3), (CNT, 3)])
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [], [(CS, 0), (SUM, 3)])


def test_key_type_mismatch_is_invalid_argument(ctx):
    P, B = up(ctx, [col([1], I32)]), up(ctx, [col([1], I64)])
    try:
        with pytest.raises(ValueError, match="types must match"):
            ctx.hash_join(P, B, INNER)
    finally:
        for d in P + B:
            d.free()


def test_xcd_of_blocks_reports_dies(ctx):
    # Block placement over the dies, printed for the record (consecutive
    # blocks on consecutive XCDs); the ids must be valid die numbers and all
    # eight dies must take blocks.
    ids = ctx.xcd_of_blocks(256)
    print("xcd of blocks 0..15:", ids[:16].tolist(),
          "round-robin:", bool((ids == np.arange(256) % 8).all()))
    assert ids.max() < 8
    assert len(set(ids.tolist())) == 8


@pytest.mark.parametrize("case", range(4))
def test_fused_many_groups_slot_range_passes(ctx, case):
    # More groups than one block's LDS holds (the tick pipeline's 50,000
    # symbol names): LDS partials per slot range, one probe pass per range,
    # each merged into its range of the result. Zipf-skewed probe keys put
    # heavy groups in every range; streamed (<= 2 value columns) and per-row
    # (3 value columns) probe kernels; I32 and I64 keys.
    r = np.random.default_rng(1000 + case)
    kt = I32 if case % 2 else I64
    n_b = 60000
    groups = [50000, 25000, 60000, 12000][case]
    n_p = [400000, 1000003, 200000, 300001][case]
    keys = r.permutation(n_b)
    zipf = np.minimum(r.zipf(1.3, n_p) - 1, n_b + 99)
    probe_keys = np.where(r.random(n_p) < 0.1, n_b + 5, keys[zipf % n_b])
    bcols = [col(keys, kt), col(r.integers(0, groups, n_b), I32, tuple(np.nonzero(r.random(n_b) < 0.01)[0]))]
    pcols = [col(probe_keys, kt), col(np.round(r.uniform(0, 200, n_p), 2), F64,
                                      tuple(np.nonzero(r.random(n_p) < 0.02)[0])),
             col(r.integers(-1000, 1000, n_p), I32), col(r.integers(-2**40, 2**40, n_p), I64)]
    aggs = [(CS, 0), (AVG, 1)] if case < 2 else [(CS, 0), (SUM, 1), (MIN, 2), (MAX, 3), (CNT, 1)]
    join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [5], aggs)


def hot_set_shape(case):
    """Build and probe columns of the hot-set shape (60,000 groups; 90% of 2^22
    probe rows on 4000 hot groups) for case 0-2, or spread over all groups for
    case 3, with the aggregates that case uses."""
    r = np.random.default_rng(1100 + case)
    kt = I32 if case % 2 else I64
    n_b, groups = 60000, 60000
    n_p = (1 << 22) + [3, 1, 0, 5][case]
    keys = r.permutation(n_b)
    grp = r.integers(0, groups, n_b)
    if case == 0:
        hot = (grp >= 20000) & (grp < 24000)
    else:
        hot = np.isin(grp, r.permutation(groups)[:4000])
    hot_rows = np.nonzero(hot)[0]
    u = r.random(n_p)
    if case < 3:
        probe_keys = np.where(u < 0.90, keys[hot_rows[r.integers(0, len(hot_rows), n_p)]],
                              np.where(u < 0.95, keys[r.integers(0, n_b, n_p)], n_b + 5))
    else:
        probe_keys = np.where(u < 0.95, keys[r.integers(0, n_b, n_p)], n_b + 5)
    bcols = [col(keys, kt), col(grp, I32, tuple(np.nonzero(r.random(n_b) < 0.01)[0]))]
    pcols = [col(probe_keys, kt), col(np.round(r.uniform(0, 200, n_p), 2), F64,
                                      tuple(np.nonzero(r.random(n_p) < 0.02)[0])),
             col(r.integers(-1000, 1000, n_p), I32), col(r.integers(-2**40, 2**40, n_p), I64)]
    aggs = [(CS, 0), (AVG, 1)] if case in (0, 3) else [(CS, 0), (SUM, 1), (MIN, 1), (MAX, 3), (SUM, 3), (CNT, 1)]
    return pcols, bcols, aggs


@pytest.mark.parametrize("case", range(4))
def test_fused_many_groups_hot_set_one_pass(ctx, qe, case):
    # More groups than LDS holds, but the probe puts nearly all rows on a few
    # thousand groups: a sample of 2^16 probe rows finds them, they are
    # relabelled to the first slots, and one probe pass keeps them in LDS while
    # the other rows fold with HBM atomics; the state returns to slot order.
    # Hot groups form one range of group values (case 0) or are scattered
    # (1, 2); case 3 spreads the rows over all groups and keeps one pass per
    # slot range. Case 2 caps the spill lists at 64 records per block, so most
    # spilled rows take the HBM-atomic fold. I32 and I64 keys; odd probe
    # lengths (tail row); NULL groups; the streamed probe (one or two value
    # columns).
    pcols, bcols, aggs = hot_set_shape(case)
    ctx.set_timing(True)
    if case == 2:
        ctx.set_tuning(qe.TUNE_JOIN_SPILL_CAP, 64)
    try:
        for name in ("join_agg_hot_set", "join_agg_probe", "join_agg_spill"):
            ctx.timing(name)
        join_agg_case(ctx, pcols, bcols, [0], [0], INNER, [5], aggs)
        _, sampled = ctx.timing("join_agg_hot_set")
        _, probes = ctx.timing("join_agg_probe")
        _, spill_passes = ctx.timing("join_agg_spill")
    finally:
        ctx.set_timing(False)
        ctx.set_tuning(qe.TUNE_JOIN_SPILL_CAP, 0)
    assert sampled == 1
    if case == 3:
        assert probes > 1 and spill_passes == 0
    else:
        assert probes == 1 and spill_passes >= 1


def test_hot_set_without_room_for_spill_lists_runs_the_passes(ctx, qe):
    # The hot-set shape with spill lists of 2^20 records per block (about 6 GiB)
    # under a device memory limit that leaves 1 GiB beyond the inputs: the
    # lists do not fit, so the query runs one probe pass per LDS slot range
    # instead of failing, and returns the oracle's groups.
    pcols, bcols, aggs = hot_set_shape(1)
    P, B = up(ctx, pcols), up(ctx, bcols)
    live, _ = ctx.mem_usage()
    ctx.set_timing(True)
    ctx.set_tuning(qe.TUNE_JOIN_SPILL_CAP, 1 << 20)
    ctx.set_mem_limit(live + (1 << 30))
    try:
        for name in ("join_agg_hot_set", "join_agg_probe", "join_agg_spill"):
            ctx.timing(name)
        t = ctx.join_aggregate(P, B, [0], [0], INNER, [5], aggs)
        got = table_rows(t.numpy())
        t.free()
        _, sampled = ctx.timing("join_agg_hot_set")
        _, probes = ctx.timing("join_agg_probe")
        _, spill_passes = ctx.timing("join_agg_spill")
    finally:
        ctx.set_mem_limit(0)
        ctx.set_timing(False)
        ctx.set_tuning(qe.TUNE_JOIN_SPILL_CAP, 0)
        for d in P + B:
            d.free()
    assert sampled == 1 and probes > 1 and spill_passes == 0
    assert_same(got, O.join_aggregate(pcols, bcols, [0], [0], INNER, [5], aggs))


def test_join_autotune_keeps_results(ctx, qe):
    # The autotuner times six launch shapes on this query and leaves the
    # fastest set; every shape returns the same groups (bit-identical sums).
    (pk, pp), (bk, bg) = ctx.gen_join(11, 1 << 16, 1 << 20, 3000)
    aggs = [(SUM, 1), (CS, 0)]
    try:
        before = ctx.join_aggregate([pk, pp], [bk, bg], [0], [0], INNER, [3], aggs)
        a = before.numpy()
        before.free()
        t = ctx.join_autotune([pk, pp], [bk, bg], [0], [0], INNER, [3], aggs)
        assert t["candidates"] == 6 and t["pairs"] in (2, 4) and t["blocks_per_cu"] in (1, 2, 4)
        assert 0 < t["probe_ms"] <= t["default_ms"]
        outs = []
        for pairs, bpc in ((t["pairs"], t["blocks_per_cu"]), (2, 4), (4, 2)):
            ctx.set_tuning(qe.TUNE_JOIN_PAIRS, pairs)
            ctx.set_tuning(qe.TUNE_JOIN_BLOCKS_PER_CU, bpc)
            r = ctx.join_aggregate([pk, pp], [bk, bg], [0], [0], INNER, [3], aggs)
            outs.append(r.numpy())
            r.free()
        with pytest.raises(qe.InvalidArgument):
            ctx.set_tuning(qe.TUNE_JOIN_BLOCKS_PER_CU, 9)
    finally:
        ctx.set_tuning(qe.TUNE_JOIN_PAIRS, 4)
        ctx.set_tuning(qe.TUNE_JOIN_BLOCKS_PER_CU, 1)
        for d in (pk, pp, bk, bg):
            d.free()
    for o in outs:
        assert len(o) == len(a)
        for (va, _), (vb, _) in zip(a, o):
            assert np.array_equal(np.sort(va.view(np.uint64)), np.sort(vb.view(np.uint64)))
