"""Seeded inputs shared by GPU tests and the CPU device models (the mutation
catalog of tests/test_device_mutants.py runs the models on exactly the inputs
the GPU tests give the kernels). Columns are (values, validity or None, type)."""
import numpy as np

import _oracle as O

I32, I64, F64, TS = O.I32, O.I64, O.F64, O.TS


def col(vals, t, nulls=()):
    vals = np.asarray(vals, dtype=O.NP[t])
    valid = None
    if len(nulls):
        valid = np.ones(len(vals), dtype=bool)
        valid[list(nulls)] = False
    return (vals, valid, t)


def direct_stream_case(case):
    """Direct unique build map with the streaming probe: odd and tiny probe
    lengths (tail row), I32 / I64 keys, NULL probe keys, narrow and nullable
    value columns. Build: key, group (37 groups); probe: key, I32, F64, I64."""
    r = np.random.default_rng(900 + case)
    kt = I32 if case % 2 else I64
    n_b = [1, 64, 3000, 70000][case % 4]
    n_p = [1, 3, 4097, 100001, 65536, 262143, 2, 5][case]
    keys = r.permutation(2 * n_b)[:n_b] - n_b  # unique, negative and positive, with gaps
    bcols = [col(keys, kt), col(r.integers(0, 37, n_b), I32)]
    knull = tuple(np.nonzero(r.random(n_p) < 0.03)[0]) if case % 3 == 0 else ()
    vnull = tuple(np.nonzero(r.random(n_p) < 0.05)[0]) if case % 2 else ()
    pcols = [col(r.integers(-n_b - 5, n_b + 5, n_p), kt, knull),
             col(r.integers(-1000, 1000, n_p), I32, vnull),
             col(r.normal(size=n_p) * 100, F64, vnull),
             col(r.integers(-2**40, 2**40, n_p), I64)]
    return pcols, bcols


def wide_map_case(case):
    """The streaming probe at 2^20+ rows over maps of 2^16..2^17 keys: skewed
    and uniform keys with 10% misses, I32/I64 keys (negative offsets), one or
    two value columns, an odd probe length, and (case 4) each probe tile's keys
    in one 4096-key stretch of the map. Returns (probe, build, aggregates)."""
    r = np.random.default_rng(1300 + case)
    n_b = [1 << 16, 1 << 17, 50000, 1 << 16, 1 << 16][case]
    n_p = [1 << 20, (1 << 20) + 3, 1 << 21, 1 << 20, 1 << 20][case]
    kt = I32 if case % 2 else I64
    kmin = [0, -5000, 123, 7, 0][case]
    keys = r.permutation(n_b) + kmin
    if case == 4:
        tile = 1024 * 4 * 2  # rows per probe tile (4 pairs per thread)
        t = np.arange(n_p) // tile
        pkv = kmin + (t % 16) * 4096 + r.integers(0, 4096, n_p)
    else:
        s = [1.1, 0.0, 1.3, 1.05][case]
        ranks = np.minimum(r.zipf(s, n_p) - 1, n_b - 1) if s > 0 else r.integers(0, n_b, n_p)
        pkv = np.where(r.random(n_p) < 0.1, kmin + n_b + r.integers(0, 1000, n_p), kmin + ranks)
    pcols = [col(pkv, kt), col(r.uniform(0, 1000, n_p), F64), col(r.integers(-2**40, 2**40, n_p), I64),
             col(r.integers(-1000, 1000, n_p), I32)]
    bcols = [col(keys, kt), col(r.integers(0, [10000, 3000, 500, 60000, 10000][case], n_b), I32)]
    aggs = [[(O.SUM, 1), (O.COUNT_STAR, 0)], [(O.SUM, 1), (O.SUM, 2), (O.COUNT_STAR, 0)],
            [(O.AVG, 3), (O.COUNT_STAR, 0)], [(O.SUM, 2), (O.AVG, 1), (O.COUNT_STAR, 0)],
            [(O.SUM, 1), (O.COUNT_STAR, 0)]][case]
    return pcols, bcols, aggs


def asof_random_case(case):
    """Unsorted build and probe sides with NULLs, 1-300 symbols, distinct build
    times, a tolerance on every third case. Returns (probe, build, tol, B)."""
    r = np.random.default_rng(4000 + case)
    nb, np_ = int(r.choice([1, 100, 5000, 60000])), int(r.choice([1, 100, 5000, 30000]))
    nsym = int(r.choice([1, 7, 300]))
    bsym = r.integers(0, nsym, nb)
    bts = r.permutation(nb * 5)[:nb] * 3 - 1000  # distinct, unsorted, some negative
    psym = r.integers(0, nsym + 2, np_)
    pts = r.integers(-2000, nb * 15 + 10, np_)
    bn = tuple(np.nonzero(r.random(nb) < 0.02)[0])
    pn = tuple(np.nonzero(r.random(np_) < 0.02)[0])
    tcol = int(r.choice([TS, I64]))
    bcols = [col(bsym, I32, bn), col(bts, tcol)]
    pcols = [col(psym, I32), col(pts, tcol, pn)]
    tol = None if case % 3 else int(r.integers(0, 40))
    return pcols, bcols, tol, int(r.choice([4, 16, 64]))


def asof_dup_runs_case(piece_rows):
    """Runs of equal build timestamps longer than a piece, rows shuffled,
    probes at, between, before and after the runs."""
    r = np.random.default_rng(808 + piece_rows)
    nb, np_ = 30000, 20000
    bsym = r.integers(0, 3, nb)
    bts = r.integers(0, 100, nb) * 10
    psym = r.integers(0, 4, np_)
    pts = r.integers(-20, 1020, np_)
    return [col(psym, I32), col(pts, TS)], [col(bsym, I32), col(bts, TS)]


def asof_bucket_edge_case(piece_rows):
    """Knots just past the linear router's bucket edges: per symbol, 4096
    build rows over a span near 2^42 us (the segment starts at a multiple of
    the piece size, so piece j starts at row j * B and there are K = 4096 / B
    pieces and router buckets); knot j (j = 1 .. K-2) sits a few hundred
    thousand microseconds past the first time of bucket j -- within the
    float32 resolution of the normalised time, where a router computed in
    other arithmetic than the table's can disagree -- and the rows of a piece
    lie between its knot and the next. Probes at every knot, one microsecond
    either side, and at every row. Returns (probe, build)."""
    r = np.random.default_rng(515 + piece_rows)
    nsym, per, B = 4, 4096, piece_rows
    K = per // B
    bsym, bts, psym, pts = [], [], [], []
    for s in range(nsym):
        t0 = int(r.integers(0, 1 << 30))
        span = (1 << 42) + int(r.integers(0, 1 << 40))
        knots = [t0]
        for j in range(1, K - 1):
            edge = t0 + (j * span + K - 1) // K  # first time with floor(x * K) = j (up to rounding)
            knots.append(edge + int(r.integers(0, 1 << 18)))
        knots.append(t0 + span)
        ts = []
        for j in range(K):
            nxt = knots[j + 1] if j + 1 < K else knots[j] + (span // K)
            inner = np.sort(r.integers(knots[j] + 1, nxt, B - 1)) if B > 1 else np.empty(0, np.int64)
            ts.extend([knots[j], *inner.tolist()])
        ts = np.asarray(ts, dtype=np.int64)
        bsym.append(np.full(per, s))
        bts.append(ts)
        kn = np.asarray(knots, dtype=np.int64)
        for d in (-1, 0, 1):
            psym.append(np.full(K, s))
            pts.append(kn + d)
        psym.append(np.full(per, s))
        pts.append(ts)
    order = r.permutation(nsym * per)
    bsym, bts = np.concatenate(bsym)[order], np.concatenate(bts)[order]
    psym, pts = np.concatenate(psym), np.concatenate(pts)
    return [col(psym, I32), col(pts, TS)], [col(bsym, I32), col(bts, TS)]


def asof_bucket_fill_case(piece_rows):
    """Router buckets with chosen row counts: per symbol K pieces of B rows
    (the segment starts at a multiple of B, so the linear router has R = K
    buckets of equal width w over [first time, first time of the last piece])
    and every bucket but the last filled with a set number of rows strictly
    inside it: counts at a data line's capacities (27-30 and 55-57 rows),
    empty and one-row buckets, fillers, each with its rows within 2^15 time
    units (16-bit offsets fit), within 2^20 (32-bit offsets), or spread over
    most of a 2^34-unit bucket (wider than 32-bit offsets); the last bucket
    holds the last piece. Probes at every row, one unit either side, just
    inside every bucket edge, before the first row and past the last, and for
    an absent symbol. Returns (probe, build)."""
    r = np.random.default_rng(2600 + piece_rows)
    B = piece_rows
    K = {4: 256, 16: 64, 64: 32}[B]
    w = 1 << 34
    edge = [27, 28, 29, 30, 55, 56, 57, 0, 1]
    widths = {0: 1 << 15, 1: 1 << 20, 2: w - (w >> 5)}
    bsym, bts, psym, pts = [], [], [], []
    for s in range(3):
        occ = [(c, wide) for c in edge for wide in (0, 1, 2)]
        rest = (K - 1) - len(occ)
        need = (K - 1) * B - sum(c for c, _ in occ)
        fill = r.multinomial(need, np.ones(rest) / rest) if rest > 0 else []
        occ += [(int(c), int(r.integers(0, 3))) for c in fill]
        occ = [occ[i] for i in r.permutation(len(occ))]
        t0 = int(r.integers(-(1 << 40), 1 << 40))
        tl = t0 + K * w  # the last piece's first row: x = 1, the last bucket
        ts = []
        for b, (c, wide) in enumerate(occ):
            lo = t0 + b * w + (w >> 6)
            span = widths[wide]
            ts.extend(sorted(int(v) for v in lo + r.integers(0, span, c)))
        if not ts or ts[0] != t0:
            # row 0 is the partition's first time (x = 0): put it at t0 in bucket 0
            first = next(i for i, (c, _) in enumerate(occ) if c)
            del ts[0]
            ts.insert(0, t0)
            if first:  # keep the counts: move one row of bucket `first` to bucket 0
                occ[0] = (occ[0][0] + 1, occ[0][1])
                occ[first] = (occ[first][0] - 1, occ[first][1])
        ts.extend([tl] + sorted(int(v) for v in tl + 1 + r.integers(0, 1 << 20, B - 1)))
        ts = np.asarray(ts, dtype=np.int64)
        assert len(ts) == K * B
        bsym.append(np.full(len(ts), s))
        bts.append(ts)
        edges = t0 + np.arange(K + 1, dtype=np.int64) * w
        for d in (-1, 0, 1):
            psym.append(np.full(len(ts), s))
            pts.append(ts + d)
        for d in (1, (w >> 7)):
            psym.append(np.full(K + 1, s))
            pts.append(edges + d)
        psym.append(np.array([s, s]))
        pts.append(np.array([t0 - 1, int(ts[-1]) + (1 << 30)]))
    psym.append(np.array([7, 7]))
    pts.append(np.array([0, 1 << 41]))
    order = r.permutation(sum(len(t) for t in bts))
    bsym, bts = np.concatenate(bsym)[order], np.concatenate(bts)[order]
    psym, pts = np.concatenate(psym), np.concatenate(pts)
    return [col(psym, I32), col(pts, TS)], [col(bsym, I32), col(bts, TS)]
