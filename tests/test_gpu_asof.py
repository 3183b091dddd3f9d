"""GPU as-of join: binary search and the learned index against the oracle.

Literal edge shapes follow tests/asof_differential_test.cpp:118-318 and
tests/catalog_checks_asof.cpp:38-92; match indices are bit-exact; the learned
probe must return exactly the binary-search matches (the exact answer) with and
without the MLP router; the device MLP forward pass and MFMA gradient kernel
are checked against the numpy oracle.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import _device_cases as DC
import _device_models as DM
import _oracle as O

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu
I32, I64, TS = O.I32, O.I64, O.TS
NO = O.NO_ROW
GOLDEN = Path(__file__).resolve().parent / "golden"


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if nulls:
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def oracle_matches(pcols, bcols, pk, bk, pt, bt, tol=None):
    p, b = O.asof_pairs(pcols, bcols, pk, bk, pt, bt, 1, tolerance=tol)
    out = np.full(len(pcols[pt][0]), NO, dtype=np.uint32)
    out[p] = b
    return out


def gpu_matches(ctx, qe, pcols, bcols, pk, bk, pt, bt, tol=None, modes=(0,), train=None):
    P = [ctx.upload(v, t, valid) for v, valid, t in pcols]
    B = [ctx.upload(v, t, valid) for v, valid, t in bcols]
    try:
        idx = ctx.asof_build([B[i] for i in bk], B[bt])
        if train is not None:
            idx.train(*train)
        res = [idx.probe([P[i] for i in pk], P[pt], tolerance=tol, mode=m) for m in modes]
        stats = idx.stats()
        idx.free()
        return res, stats
    finally:
        for d in P + B:
            d.free()


def check_all(ctx, qe, pcols, bcols, pk, bk, pt, bt, tol=None, piece_rows=8):
    # Every method, with bucket rows inline where they fit (the default) and
    # with knot records in every bucket line.
    want = oracle_matches(pcols, bcols, pk, bk, pt, bt, tol)
    stats = []
    for data_lines in (1, 0):
        ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, data_lines)
        try:
            res, st = gpu_matches(ctx, qe, pcols, bcols, pk, bk, pt, bt, tol, modes=(0, 1, 2),
                                  train=(piece_rows, 50, 0.01))
        finally:
            ctx.set_tuning(qe.TUNE_ASOF_DATA_LINES, 1)
        for m, got in zip((0, 1, 2), res):
            assert got.tolist() == want.tolist(), f"mode {m}, data lines {data_lines}"
        stats.append(st)
    assert stats[1]["data_line_buckets"] == 0
    return want, stats[0]


def test_literal_gap_tie_nearest(ctx, qe):
    b = [col([1, 1, 1, 2], I32), col([10, 20, 30, 15], TS), col([100, 200, 300, 150], I64)]
    p = [col([1, 1, 1, 2, 3, 1], I32), col([5, 10, 25, 15, 99, 50], TS)]
    want, _ = check_all(ctx, qe, p, b, [0], [0], 1, 1, piece_rows=2)
    assert want.tolist() == [NO, 0, 1, 3, NO, 2]
    w2, _ = check_all(ctx, qe, p, b, [0], [0], 1, 1, tol=10, piece_rows=2)
    assert w2.tolist() == [NO, 0, 1, 3, NO, NO]


def test_literal_zero_keys_nulls_composite(ctx, qe):
    w, _ = check_all(ctx, qe, [col([5, 25, 35], TS)], [col([10, 20, 30], TS)], [], [], 0, 0, piece_rows=2)
    assert w.tolist() == [NO, 1, 2]
    nb = [col([1, 1, 2, 7], I32, nulls=(3,)), col([10, 20, 0, 5], TS, nulls=(2,))]
    npb = [col([1, 2, 2, 9], I32, nulls=(3,)), col([25, 30, 0, 50], TS, nulls=(2,))]
    w, _ = check_all(ctx, qe, npb, nb, [0], [0], 1, 1, piece_rows=2)
    assert w.tolist() == [1, NO, NO, NO]
    cb = [col([1, 1, 1], I32), col([7, 7, 9], I32), col([10, 20, 10], TS)]
    cp = [col([1, 1], I32), col([7, 9], I32), col([25, 25], TS)]
    w, _ = check_all(ctx, qe, cp, cb, [0, 1], [0, 1], 2, 2, piece_rows=2)
    assert w.tolist() == [1, 2]


def test_catalog_boundary_ties(ctx, qe):
    kb = [col([1, 1, 2], I32), col([10, 20, 15], I64)]
    kp = [col([1, 1, 2], I32), col([20, 25, 15], I64)]
    w, _ = check_all(ctx, qe, kp, kb, [0], [0], 1, 1, piece_rows=2)
    assert w.tolist() == [1, 1, 2]


def test_empty_sides(ctx, qe):
    w, _ = check_all(ctx, qe, [col([1, 2], I32), col([5, 6], TS)], [col([], I32), col([], TS)], [0], [0], 1, 1)
    assert w.tolist() == [NO, NO]
    w, _ = check_all(ctx, qe, [col([], I32), col([], TS)], [col([1], I32), col([1], TS)], [0], [0], 1, 1)
    assert w.tolist() == []


@pytest.mark.parametrize("case", range(10))
def test_random_unsorted_inputs(ctx, qe, case):
    # Inputs shared with the device models' mutation catalog.
    pcols, bcols, tol, piece_rows = DC.asof_random_case(case)
    check_all(ctx, qe, pcols, bcols, [0], [0], 1, 1, tol=tol, piece_rows=piece_rows)


@pytest.mark.parametrize("piece_rows", [1, 2, 4, 16, 64])
def test_duplicate_time_runs_across_pieces(ctx, qe, piece_rows):
    # Runs of equal build timestamps longer than a piece (several knots share
    # one tau, and a piece's rows may all be equal), rows in shuffled order,
    # probes at, between, before and after the runs: the search inside a piece
    # starts from the spline's prediction and must end on the run's last row
    # in the stable (key, time) order.
    pcols, bcols = DC.asof_dup_runs_case(piece_rows)
    check_all(ctx, qe, pcols, bcols, [0], [0], 1, 1, piece_rows=piece_rows)


@pytest.mark.parametrize("piece_rows", [4, 16, 64])
def test_knots_at_router_bucket_edges(ctx, qe, piece_rows):
    # Knots a few hundred thousand microseconds past the linear router's
    # bucket edges over spans near 2^42 us: within the float32 resolution of
    # the normalised time, so a probe routed with other arithmetic than the
    # table's lands one bucket low and takes the wrong piece without checking
    # its successor. Probes at every knot, one microsecond either side and at
    # every row: the learned probe (MLP and linear routers) must equal the
    # binary search and the oracle on all of them.
    pcols, bcols = DC.asof_bucket_edge_case(piece_rows)
    check_all(ctx, qe, pcols, bcols, [0], [0], 1, 1, piece_rows=piece_rows)


@pytest.mark.parametrize("piece_rows", [4, 16, 64])
def test_router_buckets_filled_to_the_data_line_edge(ctx, qe, piece_rows):
    # Buckets of exactly 27-30 rows (a data line holds 28), empty and one-row
    # buckets, and buckets whose rows span more than 32-bit offsets: each
    # bucket is a data line or a knot line, and probes at every row, either
    # side of it, just inside every bucket edge (before the bucket's first row)
    # and past the ends must match the oracle, with and without a tolerance.
    pcols, bcols = DC.asof_bucket_fill_case(piece_rows)
    for tol in (None, 1 << 19, 1 << 33):
        _, st = check_all(ctx, qe, pcols, bcols, [0], [0], 1, 1, tol=tol, piece_rows=piece_rows)
        assert 0 < st["data_line_buckets"] < st["router_buckets"], st


def test_data_lines_every_output_mode(ctx, qe):
    # The data-line probe's three answers: the sorted position, the original
    # row, and the tolerance check on the matched time (taken from the line,
    # or from the sorted times for a match before the bucket's first row).
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(41, 1 << 21, 1 << 19, 300, 1.1)
    try:
        idx = ctx.asof_build([rs], rt)
        try:
            idx.train(16, 100, 0.003)
            st = idx.stats()
            rows = idx.sorted(st["build_rows"], st["partitions"])[1]
            for tol in (None, 0, 150, 700, 1 << 40):
                want = idx.probe([ls], lt, tolerance=tol, mode=qe.ASOF_BINARY)
                for mode in (qe.ASOF_LEARNED, qe.ASOF_LEARNED_NO_MLP):
                    assert np.array_equal(idx.probe([ls], lt, tolerance=tol, mode=mode), want), (mode, tol)
                    pos = idx.probe([ls], lt, tolerance=tol, mode=mode | qe.ASOF_OUT_SORTED)
                    hit = pos != NO
                    assert np.array_equal(hit, want != NO), (mode, tol)
                    assert np.array_equal(rows[pos[hit]], want[hit]), (mode, tol)
        finally:
            idx.free()
    finally:
        for c in (rs, rt, rp, ls, lt):
            c.free()
    # Most buckets of this shape hold their rows inline.
    assert st["data_line_buckets"] > 0.5 * st["router_buckets"], st


def _cluster_case(seed):
    """Build side of 40 symbols (ids 3..42, some absent) with up to 1500 rows
    each (under 64 pieces of 32 rows, so no router network is trained and the
    linear router is exact to model), and probes with NULL keys and times,
    absent symbols and times before and after each symbol's rows."""
    r = np.random.default_rng(seed)
    syms = np.array([s for s in range(3, 43) if s % 7 != 0])
    bsym = np.concatenate([np.full(int(r.integers(1, 1500)), s) for s in syms])
    bts = r.integers(0, 10**7, len(bsym)) * 3
    n = 30011
    psym = r.integers(0, 46, n)
    pts = r.integers(-10**6, 3 * 10**7 + 10**6, n)
    pkn = tuple(np.nonzero(r.random(n) < 0.02)[0])
    ptn = tuple(np.nonzero(r.random(n) < 0.02)[0])
    return ([col(bsym, I32), col(bts, I64)],
            [col(psym, I32, pkn), col(pts, I64, ptn), col(r.normal(size=n), O.F64, tuple(np.nonzero(r.random(n) < 0.1)[0])),
             col(r.integers(0, 2, n), O.BOOL), col(np.arange(n), I64)])


def _expected_groups(bcols, pcols, B, trained):
    """Each probe's group under the index layout: router-table entries laid out
    by partition (R = K entries + 1 per non-empty partition), the line of a
    probe its partition's first entry plus its linear-router bucket (the
    device arithmetic, modelled in tests/_device_models.py); untrained, the
    partition's first sorted row."""
    bsym, bts = bcols[0][0].astype(np.int64), bcols[1][0]
    kmin, kmax = int(bsym.min()), int(bsym.max())
    segs = [np.sort(bts[bsym == k]) for k in range(kmin, kmax + 1)]
    idx = DM.LearnedIndex(segs, B)
    tb, seg = [], []
    acc = acc_rows = 0
    for m in idx.parts:
        tb.append(acc)
        seg.append(acc_rows)
        acc += m["K"] + 1 if m["K"] else 0
        acc_rows += m["n"]
    ntab, nsorted = max(acc, 1), acc_rows
    psym, pts = pcols[0][0].astype(np.int64), pcols[1][0]
    pv = np.ones(len(psym), bool)
    for c in (pcols[0], pcols[1]):
        if c[1] is not None:
            pv &= c[1]
    g = np.zeros(len(psym), np.int64)
    for j in range(len(psym)):
        p = int(psym[j]) - kmin
        if not pv[j] or p < 0 or p > kmax - kmin:
            continue
        if trained:
            m = idx.parts[p]
            line = tb[p]
            if m["K"] and int(pts[j]) >= m["t0"]:
                line += idx._bucket(m, int(pts[j]))
            g[j] = min(line * 256 // ntab, 255)
        else:
            g[j] = min(seg[p] * 256 // nsorted, 255)
    return g


@pytest.mark.parametrize("trained", [True, False])
def test_cluster_groups_probes_by_index_region(ctx, qe, trained):
    # qe_gpu_asof_cluster: the probe rows reordered by the index region their
    # probe reads (stable within a group), every column carried with its
    # validity, and each row's match unchanged by the reordering.
    bcols, pcols = _cluster_case(11 if trained else 12)
    B = [ctx.upload(v, t, valid) for v, valid, t in bcols]
    P = [ctx.upload(v, t, valid) for v, valid, t in pcols]
    try:
        idx = ctx.asof_build([B[0]], B[1])
        try:
            if trained:
                idx.train(32, 50, 0.01)
                assert idx.stats()["mlp_candidates"] == 0
            modes = (qe.ASOF_BINARY, qe.ASOF_LEARNED) if trained else (qe.ASOF_BINARY,)
            orig = [idx.probe([P[0]], P[1], mode=m) for m in modes]
            cl = idx.cluster([P[0]], P[1], P)
            cl2 = idx.cluster([P[0]], P[1], P)
            try:
                got = [c.numpy() for c in (cl.col(i) for i in range(len(P)))]
                again = cl2.col(4).numpy()[0]
                moved = [idx.probe([cl.col(0)], cl.col(1), mode=m) for m in modes]
            finally:
                cl.free()
                cl2.free()
        finally:
            idx.free()
    finally:
        for d in P + B:
            d.free()
    rid = got[4][0]
    want = np.argsort(_expected_groups(bcols, pcols, 32, trained), kind="stable")
    assert np.array_equal(rid, want)
    assert np.array_equal(again, rid)  # deterministic
    for (vals, valid, _), (gv, gvalid) in zip(pcols, got):
        assert np.array_equal(gv, vals[rid])
        want_valid = np.ones(len(rid), bool) if valid is None else valid[rid]
        assert np.array_equal(np.ones(len(rid), bool) if gvalid is None else gvalid, want_valid)
    for o, m in zip(orig, moved):
        assert np.array_equal(m, o[rid])


@pytest.mark.parametrize("shape", ["zero_keys", "hashed_i64", "str_keys"])
def test_cluster_other_key_shapes(ctx, qe, shape):
    # The clustering's other partition functors: no keys (one partition),
    # hashed I64 keys (range too wide for direct partitions) and STR keys
    # (probe dictionary remapped onto the build's). The result is a
    # permutation of the probe rows, carrying every column, and each row's
    # match is unchanged.
    r = np.random.default_rng(77)
    nb, n = 50000, 60001
    if shape == "zero_keys":
        bk, pk = [], []
    elif shape == "hashed_i64":
        keys = r.integers(-2**62, 2**62, 40)
        bk = [col(keys[r.integers(0, 40, nb)], I64)]
        pk = [col(np.where(r.random(n) < 0.9, keys[r.integers(0, 40, n)], 12345), I64)]
    else:
        names = [f"sym{i}" for i in range(30)]
        bk = [("strs", [names[i] for i in r.integers(0, 30, nb)])]
        pk = [("strs", [names[i] if i < 30 else "none" for i in r.integers(0, 32, n)])]
    bts, pts = r.integers(0, 10**8, nb), r.integers(-10**6, 10**8 + 10**6, n)

    def up(c):
        if isinstance(c[0], str):  # a dictionary per side, in first-seen order
            sc = ctx.upload_strings(c[1])
            return sc, sc.dict
        return ctx.upload(c[0], c[2], c[1]), None

    B = [up(c) for c in bk]
    P = [up(c) for c in pk]
    bt, pt = ctx.upload(bts, I64), ctx.upload(pts, I64)
    rid = ctx.upload(np.arange(n), I64)
    try:
        idx = ctx.asof_build([b for b, _ in B], bt)
        try:
            idx.train(32, 50, 0.01)
            want = idx.probe([p for p, _ in P], pt, mode=qe.ASOF_LEARNED)
            cl = idx.cluster([p for p, _ in P], pt, [pt, rid] + [p for p, _ in P])
            try:
                perm = cl.col(1).numpy()[0]
                assert np.array_equal(np.sort(perm), np.arange(n))
                assert np.array_equal(cl.col(0).numpy()[0], pts[perm])
                keys = [cl.col(2 + i) for i in range(len(P))]
                got = idx.probe(keys, cl.col(0), mode=qe.ASOF_LEARNED)
                assert np.array_equal(got, want[perm])
            finally:
                cl.free()
        finally:
            idx.free()
    finally:
        for d, dd in B + P:
            d.free()
        for d in (bt, pt, rid):
            d.free()
        for _, dd in B + P:
            if dd is not None:
                dd.free()


def test_cluster_empty_and_tiny(ctx, qe):
    b = [ctx.upload(np.array([1, 1, 2], np.int32), I32), ctx.upload(np.array([5, 9, 7], np.int64), I64)]
    try:
        idx = ctx.asof_build([b[0]], b[1])
        try:
            for n in (0, 1, 5):
                pk = ctx.upload(np.array([1, 2, 9, 1, 2][:n], np.int32), I32)
                pt = ctx.upload(np.array([6, 8, 1, 10, 3][:n], np.int64), I64)
                try:
                    cl = idx.cluster([pk], pt, [pk, pt])
                    try:
                        assert cl.rows == n
                        ks, ts = cl.col(0).numpy()[0], cl.col(1).numpy()[0]
                        assert sorted(zip(ks.tolist(), ts.tolist())) == sorted(
                            zip([1, 2, 9, 1, 2][:n], [6, 8, 1, 10, 3][:n]))
                    finally:
                        cl.free()
                finally:
                    pk.free()
                    pt.free()
            with pytest.raises(qe.InvalidArgument):
                short = ctx.upload(np.array([1], np.int32), I32)
                try:
                    idx.cluster([b[0]], b[1], [short])
                finally:
                    short.free()
        finally:
            idx.free()
    finally:
        for d in b:
            d.free()


@pytest.mark.parametrize("n", [1, 63, 4095, 4096, 4097, 100_003])
def test_pairs_from_match_arrays(ctx, qe, n):
    # Per-probe matches to join pairs: INNER keeps the matched probes, LEFT
    # every probe with kNoRow for the unmatched; both in probe order. Match
    # arrays with none, all and a random half matched, across tile edges.
    r = np.random.default_rng(n)
    for frac in (0.0, 1.0, 0.5):
        match = np.where(r.random(n) < frac, r.integers(0, 1 << 31, n), NO).astype(np.uint32)
        buf = qe.DevBuffer(ctx, 4 * n)
        try:
            buf.upload(match)
            for jt in (qe.INNER, qe.LEFT):
                pr = ctx.asof_pairs(buf, n, jt)
                try:
                    p, b = pr.numpy()
                finally:
                    pr.free()
                keep = np.arange(n) if jt == qe.LEFT else np.nonzero(match != NO)[0]
                assert p.tolist() == keep.tolist()
                assert b.tolist() == match[keep].tolist()
        finally:
            buf.free()


def test_hash_partitioned_keys(ctx, qe):
    r = np.random.default_rng(77)
    nb, np_ = 20000, 20000
    keys = r.integers(-2**60, 2**60, 50)
    bcols = [col(keys[r.integers(0, 50, nb)], I64), col(r.permutation(nb) * 2, TS)]
    pcols = [col(keys[r.integers(0, 50, np_)], I64), col(r.integers(0, 2 * nb, np_), TS)]
    check_all(ctx, qe, pcols, bcols, [0], [0], 1, 1, piece_rows=32)


def test_golden_fixture(ctx, qe):
    fx = json.loads((GOLDEN / "asof_left_3000.json").read_text())
    mk = lambda spec: [col(c["values"], c["type"], tuple(c.get("nulls", []))) for c in spec]
    res, _ = gpu_matches(ctx, qe, mk(fx["probe"]), mk(fx["build"]), fx["pkeys"], fx["bkeys"], fx["ptime"],
                         fx["btime"], modes=(0, 1), train=(16, 50, 0.01))
    want = np.full(len(fx["probe"][0]["values"]), NO, dtype=np.uint32)
    for p, b in fx["expected_pairs"]:
        want[p] = b
    for got in res:
        assert got.tolist() == want.tolist()


def test_generated_tick_shape_learned_equals_binary(ctx, qe):
    # Scaled config shape: Zipf(1.1) symbols, bursty exponential gaps.
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(99, 1 << 22, 1 << 20, 500, 1.1)
    try:
        idx = ctx.asof_build([rs], rt)
        idx.train(64, 300, 0.003)
        b = idx.probe([ls], lt, mode=qe.ASOF_BINARY)
        l1 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED)
        st1 = idx.stats()
        l2 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED_NO_MLP)
        idx.free()
        rts, _ = rt.numpy()
        rsy, _ = rs.numpy()
        lsy, _ = ls.numpy()
        lts, _ = lt.numpy()
    finally:
        for c in (rs, rt, rp, ls, lt):
            c.free()
    assert (b == l1).all() and (b == l2).all()
    assert st1["fallbacks"] == 0
    # Spot check a sample against numpy searchsorted (per-symbol sorted runs).
    order = np.lexsort((rts, rsy))
    ssym, sts = rsy[order], rts[order]
    for j in np.random.default_rng(5).choice(len(lts), 2000, replace=False):
        lo = np.searchsorted(ssym, lsy[j], side="left")
        hi = np.searchsorted(ssym, lsy[j], side="right")
        k = np.searchsorted(sts[lo:hi], lts[j], side="right")
        assert b[j] == (NO if k == 0 else order[lo + k - 1])
    # Bursts here come in 512-row regimes, finer than any 16-unit router can
    # follow: the per-partition choice keeps the learned router no worse.
    assert st1["mean_bucket_knots"] <= st1["mean_bucket_knots_linear"]


def test_router_mlp_gain_on_session_shape(ctx, qe):
    # Five trading days per symbol with a U-shaped intraday intensity and
    # overnight gaps: coarse structure the trained router learns, so its radix
    # buckets hold fewer knots than the linear router's, with results unchanged.
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(123, 1 << 22, 1 << 20, 200, 1.1, sessions=5)
    try:
        idx = ctx.asof_build([rs], rt)
        idx.train(64, 300, 0.003)
        b = idx.probe([ls], lt, mode=qe.ASOF_BINARY)
        l1 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED)
        st = idx.stats()
        l2 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED_NO_MLP)
        idx.free()
        rts, _ = rt.numpy()
        rsy, _ = rs.numpy()
        lsy, _ = ls.numpy()
        lts, _ = lt.numpy()
    finally:
        for c in (rs, rt, rp, ls, lt):
            c.free()
    assert (b == l1).all() and (b == l2).all()
    assert st["fallbacks"] == 0
    order = np.lexsort((rts, rsy))
    ssym, sts = rsy[order], rts[order]
    for j in np.random.default_rng(6).choice(len(lts), 2000, replace=False):
        lo = np.searchsorted(ssym, lsy[j], side="left")
        hi = np.searchsorted(ssym, lsy[j], side="right")
        k = np.searchsorted(sts[lo:hi], lts[j], side="right")
        assert b[j] == (NO if k == 0 else order[lo + k - 1])
    # What the trained network gains: routed through it, a probe's bucket holds
    # far fewer knots than under the linear router.
    assert st["mean_bucket_knots_mlp"] < 0.8 * st["mean_bucket_knots_linear"], st
    # The probe keeps the MLP only where the timing says it is faster.
    assert st["sample_probe_ms_mlp"] > 0 and st["sample_probe_ms_linear"] > 0, st
    if st["mlp_partitions"] > 0:
        assert st["sample_probe_ms_mlp"] <= 0.98 * st["sample_probe_ms_linear"], st
        assert st["mean_bucket_knots"] < st["mean_bucket_knots_linear"], st
    else:
        assert st["sample_probe_ms_mlp"] > 0.98 * st["sample_probe_ms_linear"], st
        assert st["mean_bucket_knots"] == st["mean_bucket_knots_linear"], st


@pytest.mark.parametrize("kpb", [2, 8, 64])
def test_coarse_router_buckets_stay_exact(ctx, qe, kpb):
    # Several knots per router bucket (a smaller radix table): the learned
    # probe still returns exactly the binary-search matches, for both shapes.
    for sessions in (0, 3):
        (rs, rt, rp), (ls, lt) = ctx.gen_asof(77 + kpb, 1 << 20, 1 << 18, 64, 1.1, sessions=sessions)
        try:
            idx = ctx.asof_build([rs], rt)
            ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, kpb)
            try:
                idx.train(32, 200, 0.003)
            finally:
                ctx.set_tuning(qe.TUNE_ASOF_BUCKET_KNOTS, 1)
            b = idx.probe([ls], lt, mode=qe.ASOF_BINARY)
            l1 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED)
            st = idx.stats()
            l2 = idx.probe([ls], lt, mode=qe.ASOF_LEARNED_NO_MLP)
            idx.free()
        finally:
            for c in (rs, rt, rp, ls, lt):
                c.free()
        assert (b == l1).all() and (b == l2).all()
        assert st["fallbacks"] == 0
        assert st["router_buckets"] < st["pieces"]
        assert st["mean_bucket_knots"] <= st["mean_bucket_knots_linear"]


def test_sorted_position_output_and_take_sorted(ctx, qe):
    # QE_ASOF_OUT_SORTED: every method reports the sorted position whose
    # original row is the plain output's match, misses and tolerance drops in
    # the same places; take_sorted lays build columns out in that order.
    (rs, rt, rp), (ls, lt) = ctx.gen_asof(31, 1 << 21, 1 << 19, 300, 1.1)
    try:
        idx = ctx.asof_build([rs], rt)
        try:
            idx.train(64, 200, 0.003)
            st = idx.stats()
            sts, rows, offs = idx.sorted(st["build_rows"], st["partitions"])
            got = {}
            for mode in (qe.ASOF_BINARY, qe.ASOF_LEARNED, qe.ASOF_LEARNED_NO_MLP):
                for tol in (None, 700):
                    got[(mode, tol)] = (idx.probe([ls], lt, tolerance=tol, mode=mode),
                                        idx.probe([ls], lt, tolerance=tol, mode=mode | qe.ASOF_OUT_SORTED))
            t = idx.take_sorted([rp, rt])
            tv = t.numpy()
            t.free()
        finally:
            idx.free()
        rpn, _ = rp.numpy()
        rtn, _ = rt.numpy()
    finally:
        for c in (rs, rt, rp, ls, lt):
            c.free()
    for (mode, tol), (orig, pos) in got.items():
        hit = pos != NO
        assert np.array_equal(hit, orig != NO), (mode, tol)
        assert np.array_equal(rows[pos[hit]], orig[hit]), (mode, tol)
        assert (pos[hit] < len(rows)).all()
    # Tolerance drops some matches and keeps others.
    assert 0 < (got[(qe.ASOF_BINARY, 700)][0] != NO).sum() < (got[(qe.ASOF_BINARY, None)][0] != NO).sum()
    assert np.array_equal(tv[0][0], rpn[rows]) and np.array_equal(tv[1][0], sts)
    assert np.array_equal(tv[1][0], rtn[rows])


def test_take_sorted_skips_unmatchable_rows(ctx, qe):
    # Rows with a NULL key or time are not in the index: take_sorted returns
    # the matchable rows only, in (key, time) order.
    valid_k = np.array([1, 0, 1, 1, 1], dtype=bool)
    valid_t = np.array([1, 1, 1, 1, 0], dtype=bool)
    keys = ctx.upload(np.array([1, 5, 1, 2, 2], dtype=np.int32), I32, valid_k)
    ts = ctx.upload(np.array([10, 5, 20, 7, 3], dtype=np.int64), I64, valid_t)
    pay = ctx.upload(np.array([100, 101, 102, 103, 104], dtype=np.int64))
    pk = ctx.upload(np.array([1, 2], dtype=np.int32), I32)
    pt = ctx.upload(np.array([15, 9], dtype=np.int64), I64)
    try:
        idx = ctx.asof_build([keys], ts)
        try:
            t = idx.take_sorted([pay])
            vals = t.numpy()[0][0].tolist()
            t.free()
            pos = idx.probe([pk], pt, mode=qe.ASOF_BINARY | qe.ASOF_OUT_SORTED)
            orig = idx.probe([pk], pt, mode=qe.ASOF_BINARY)
        finally:
            idx.free()
    finally:
        for c in (keys, ts, pay, pk, pt):
            c.free()
    assert vals == [100, 102, 103]
    assert [vals[p] for p in pos] == [100, 103] and orig.tolist() == [0, 3]


def test_mlp_forward_matches_cpu(ctx):
    r = np.random.default_rng(8)
    p = M.identity_params() + np.abs(r.normal(scale=0.1, size=M.N_PARAMS))
    x = np.linspace(0, 1, 4097).astype(np.float32)
    y = ctx.mlp_forward(p.astype(np.float32), x)  # batched MFMA inference
    assert np.allclose(y, M.forward(p.astype(np.float32).astype(np.float64), x), rtol=1e-5, atol=1e-5)
    assert (np.diff(y) >= -1e-6).all()  # monotone up to float rounding


@pytest.mark.parametrize("n", [1, 15, 16, 17, 300, 2048])
def test_mlp_gradient_matches_cpu(ctx, n):
    r = np.random.default_rng(n)
    p = (M.identity_params() + r.normal(scale=0.05, size=M.N_PARAMS)).astype(np.float32)
    x = r.uniform(0, 1, n).astype(np.float32)
    c = np.sqrt(x).astype(np.float32)
    g, loss = ctx.mlp_grad(p, x, c)
    wl, wg = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), c.astype(np.float64))
    assert abs(loss - wl) <= 1e-5 * max(1.0, wl)
    scale = np.abs(wg).max() + 1e-12
    assert np.abs(g - wg).max() <= 2e-4 * scale, np.argmax(np.abs(g - wg))
