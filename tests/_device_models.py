"""CPU models of the HIP path's device algorithms, with the device-level
defects of the mutation catalog planted in them (test infrastructure: a
planted defect lives here, on the CPU, never in device code).

Each model follows its kernels' own decomposition (row-to-block assignment,
partitions and buckets, merge order, piece and router layout), so a planted
defect changes the result exactly where the same defect in the kernel would.
The clean model must equal the oracle on the GPU tests' inputs, and every
mutant must be flagged there by the comparison those tests make (the GPU test
files import the same input builders from tests/_device_cases.py).

fused_join_agg -- join.hip: the direct unique slot map built by key-range
  partitions (map_part_scatter16 into fixed buckets of 2^pb entries when the
  keys are dense, map_part_build writing each partition's slice), the
  streaming probe join_agg_stream (row pair p of grid tile t folded by block
  t mod blocks; the last row of an odd-length probe by block 0) and the
  block-order merge merge_partials_kernel (chunks of 16 blocks).
  Mutants (device hazards of the catalog):
    "bucket_edge" -- a partition's bucket keeps one entry fewer than it holds
                     (a partition skips a row on overflow);
    "slice_tail"  -- a partition's map slice is written one entry short;
    "pair_tail"   -- the wave-tail bound drops the probe's last row pair;
    "odd_tail"    -- the last row of an odd-length probe is not folded;
    "merge_drop"  -- the merge drops the last block partial of each chunk.

asof_learned -- asof.hip: matchable build rows sorted by (partition, time,
  row), pieces of B rows cut at global multiples of B, knots (first time,
  float32 slope, max error), the linear router (R = K buckets, floorf of the
  float32 normalised time), bucket lines of 8 records, the piece taken from
  the records and the outward (galloping) search from the spline's
  prediction, with the knot-range search and its verification past the 7th
  record.
  Buckets of at most 28 rows spanning less than 2^32 time units, or 56 rows
  spanning less than 2^16, hold the rows inline (data lines: the first time,
  the first sorted position r0 and each row's offset from the first time, in
  32 or 16 bits); the match is r0 + c - 1 for c rows at or before the probe's
  time.
  Mutants:
    "le_strict"  -- `<` for `<=` when the records or a data line's offsets
                    place the probe's time;
    "router_f32" -- probes normalise time in float32 arithmetic while the
                    router table was built with the double-precision form, so a
                    time near a bucket edge can route one bucket low and take
                    the last record there without checking its successor;
    "window_err" -- the search inside a piece is a window of the piece's
                    recorded error around the prediction, not bracketed;
    "dl_span"    -- data lines are built without the span check, so offsets
                    past their width (32 or 16 bits) wrap;
    "dl_cap"     -- a bucket of one row more than a data line holds (28 or
                    56) is still written as one, its last row dropped;
    "dl_below"   -- a probe before its data line's first time is not held at
                    zero rows: its offset wraps and counts every row.
"""
import numpy as np

NO_ROW = 0xFFFFFFFF

# ---- fused join -> group-by (direct unique map, streaming probe) --------------------------
MAP_TILE = 16384        # build rows per partitioning block (kMapTile)
MAX_MAP_PARTS = 4096    # kMaxMapParts
MERGE_CHUNK = 16        # blocks per first-stage merge chunk
JOIN_MUTANTS = ("bucket_edge", "slice_tail", "pair_tail", "odd_tail", "merge_drop")


def direct_map(bkeys, bslots, mutant=None):
    """gmap[key - kmin] = slot for unique integer build keys (all valid), built
    like build_direct_map for a 16-bit map; returns (gmap, kmin, dup)."""
    bkeys = np.asarray(bkeys, dtype=np.int64)
    kmin = int(bkeys.min())
    rng = int(bkeys.max()) - kmin + 1
    pb = 16 if rng > (MAX_MAP_PARTS << 15) else 15
    P = 1 << pb
    nparts = (rng + P - 1) >> pb
    off = bkeys - kmin
    part = off >> pb
    gmap = np.full(rng, 0xFFFF, dtype=np.int64)
    dup = False
    fixed = rng <= 2 * len(bkeys)
    cap = P - 1 if mutant == "bucket_edge" else P
    # Fixed buckets: tiles claim runs in tile order; a bucket keeps the first
    # `cap` entries it is given. The scanned layout keeps every entry.
    order = np.lexsort((np.arange(len(off)) // MAP_TILE, part))  # by partition, then tile
    kept = np.ones(len(off), dtype=bool)
    if fixed:
        starts = np.searchsorted(part[order], np.arange(nparts))
        within = np.arange(len(off)) - starts[part[order]]
        kept[order[within >= cap]] = False
        dup = bool((np.bincount(part, minlength=nparts) > P).any())
    seen = np.zeros(rng, dtype=bool)
    for i in np.nonzero(kept)[0]:
        if seen[off[i]]:
            dup = True
        seen[off[i]] = True
        gmap[off[i]] = int(bslots[i]) & 0xFFFF
    if mutant == "slice_tail":  # each slice written through min(P, range - lo) - 1
        for p in range(nparts):
            lo = p << pb
            cnt = min(P, rng - lo)
            gmap[lo + cnt - 1] = 0xFFFF
    return gmap, kmin, dup


def stream_blocks(n, blocks, pairs):
    """Block folding each probe row in join_agg_stream (1024-thread blocks,
    `pairs` row pairs per thread per tile); -1 for rows no block folds."""
    r = np.arange(n, dtype=np.int64)
    tile_pairs = 1024 * pairs
    blk = ((r >> 1) // tile_pairs) % blocks
    if n & 1:
        blk[n - 1] = 0  # the last row of an odd-length column: block 0, thread 0
    return blk


def fused_join_agg(pkeys, pvalid, values, bkeys, bgroups, blocks, pairs, mutant=None):
    """Groups of SELECT g, SUM(v) ..., COUNT(*) FROM probe JOIN build USING (k)
    GROUP BY g for unique integer build keys and a dense integer group column:
    rows (g, sum_0, ..., count) for every group with matches, in slot order.
    Integer value columns sum exactly; F64 columns in float64 per block
    partial (the device's fixed-point partials are exact; comparisons use the
    oracle's 1e-12 bound)."""
    pkeys = np.asarray(pkeys, dtype=np.int64)
    n = len(pkeys)
    bgroups = np.asarray(bgroups, dtype=np.int64)
    gmin = int(bgroups.min())
    S = int(bgroups.max()) - gmin + 1
    gmap, kmin, dup = direct_map(bkeys, bgroups - gmin, mutant)
    assert not dup, "the model covers unique build keys"
    blk = stream_blocks(n, blocks, pairs)
    live = np.ones(n, dtype=bool) if pvalid is None else np.asarray(pvalid, dtype=bool).copy()
    if mutant == "pair_tail" and n >= 2:
        npairs = n >> 1
        live[2 * (npairs - 1): 2 * npairs] = False
    if mutant == "odd_tail" and (n & 1):
        live[n - 1] = False
    off = pkeys - kmin
    inr = live & (off >= 0) & (off < len(gmap))
    slot = np.full(n, 0xFFFF, dtype=np.int64)
    slot[inr] = gmap[off[inr]]
    m = inr & (slot != 0xFFFF)
    # Per-block partials.
    cnt = np.zeros((blocks, S), dtype=np.int64)
    np.add.at(cnt, (blk[m], slot[m]), 1)
    sums = []
    for v in values:
        v = np.asarray(v)
        acc = np.zeros((blocks, S), dtype=np.int64 if v.dtype.kind in "iu" else np.float64)
        np.add.at(acc, (blk[m], slot[m]), v[m].astype(acc.dtype))
        sums.append(acc)
    # Block-order merge in chunks of MERGE_CHUNK.
    keep = np.ones(blocks, dtype=bool)
    if mutant == "merge_drop":
        for c0 in range(0, blocks, MERGE_CHUNK):
            keep[min(blocks, c0 + MERGE_CHUNK) - 1] = False
    tot_cnt = cnt[keep].sum(axis=0)
    tot = [a[keep].sum(axis=0) for a in sums]
    rows = []
    for s in np.nonzero(tot_cnt)[0]:
        rows.append((int(s) + gmin, *[(float(t[s]) if t.dtype.kind == "f" else int(t[s])) for t in tot],
                     int(tot_cnt[s])))
    return rows


def stream_grid_blocks(n, num_cus=256, blocks_per_cu=1):
    """Blocks the streaming probe launches for n probe rows."""
    return max(1, min(num_cus * blocks_per_cu, (n + 1023) // 1024))


# ---- learned as-of probe (linear router) -----------------------------------------------
BUCKET_RECS = 8
ASOF_MUTANTS = ("le_strict", "router_f32", "window_err", "dl_span", "dl_cap", "dl_below")
DATA_ROWS, DATA_ROWS16 = 28, 56
I64_MIN, I64_MAX = -(1 << 63), (1 << 63) - 1


def _f32(x):
    return np.float32(x)


def piece_pred(tau, slope, cnt, t):
    """The device's piece_pred: float32 fma of the time offset and slope."""
    dt = _f32(np.uint64((t - tau) & 0xFFFFFFFFFFFFFFFF))
    p = np.float32(np.float64(dt) * np.float64(slope) + 0.5)  # fmaf rounds once
    if not p >= 0:
        p = np.float32(0)
    hi = _f32(cnt - 1)
    if p > hi:
        p = hi
    return int(p)


class LearnedIndex:
    """The learned index of one set of partitions (sorted timestamps per
    partition, laid out back to back like the device's sorted array), linear
    router, B rows per piece, one knot per router bucket, data lines where a
    bucket's rows fit (data_lines)."""

    def __init__(self, seg_ts, B, data_lines=True, mutant=None):
        self.B = B
        self.data_lines = data_lines
        self.build_mutant = mutant
        self.parts = []
        seg = 0
        for ts in seg_ts:
            ts = np.asarray(ts, dtype=np.int64)
            n = len(ts)
            K = ((seg + n + B - 1) // B - seg // B) if n else 0
            starts = [max(seg, (seg // B + j) * B) - seg for j in range(K)]
            ends = [min(seg + n, (seg // B + j + 1) * B) - seg for j in range(K)]
            knots = []
            for j in range(K):
                a, e = starts[j], ends[j]
                cnt = e - a
                t0, tl = int(ts[a]), int(ts[e - 1])
                span = (tl - t0) & 0xFFFFFFFFFFFFFFFF
                s = _f32(np.float64(cnt - 1) / np.float64(span)) if cnt > 1 and span else _f32(0)
                err = max(abs(piece_pred(t0, s, cnt, int(ts[a + i])) - i) for i in range(cnt))
                knots.append((t0, s, err))
            meta = {"ts": ts, "seg": seg, "n": n, "K": K, "starts": starts, "ends": ends, "knots": knots, "R": K}
            if K:
                meta["t0"] = int(ts[0])
                span = (int(ts[starts[K - 1]]) - int(ts[0])) & 0xFFFFFFFFFFFFFFFF
                meta["inv_span"] = 1.0 / float(span) if span else 0.0
                tab = self._table(meta)
                meta["tab"] = tab
                meta["lines"] = [self._line(meta, b) for b in range(K)]
                meta["dlines"] = [self._dline(meta, b) if data_lines else None for b in range(K)]
            self.parts.append(meta)
            seg += n

    @staticmethod
    def _norm(meta, t, f32=False):
        if f32:  # float32 arithmetic throughout
            x = np.float32(np.float32((t - meta["t0"]) & 0xFFFFFFFFFFFFFFFF) * np.float32(meta["inv_span"]))
            return np.float32(min(x, np.float32(1.0)))
        x = float((t - meta["t0"]) & 0xFFFFFFFFFFFFFFFF) * meta["inv_span"]
        return np.float32(min(x, 1.0))

    def _bucket(self, meta, t, f32=False):
        b = np.floor(np.float32(self._norm(meta, t, f32) * np.float32(meta["R"])))
        if not b >= 0:
            b = 0.0
        return int(min(b, meta["R"] - 1))

    def _table(self, meta):
        K, R = meta["K"], meta["R"]
        tab = [0] * (R + 1)
        prev = -1
        for j, (tau, _, _) in enumerate(meta["knots"]):
            bj = self._bucket(meta, tau)
            for b in range(prev + 1, bj + 1):
                tab[b] = j
            prev = bj
        for b in range(prev + 1, R + 1):
            tab[b] = K
        return tab

    @staticmethod
    def _line(meta, b):
        lo, hi = meta["tab"][b], meta["tab"][b + 1]
        kn = meta["knots"]
        recs = [(kn[lo - 1][0], lo - 1) if lo > 0 else (I64_MIN, 0)]
        for k in range(1, BUCKET_RECS):
            j = lo + k - 1
            recs.append((kn[j][0], j) if j < hi else (I64_MAX, 0))
        more = hi - lo > BUCKET_RECS - 1
        return recs, more

    def _first_row(self, meta, k, b):
        """First row (partition-local) whose bucket is >= b, given k = tab[b]:
        a binary search between knot k - 1's row and knot k's."""
        if k == 0:
            return 0
        lo = meta["starts"][k - 1] + 1
        hi = meta["starts"][k] if k < meta["K"] else meta["n"]
        ts = meta["ts"]
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if self._bucket(meta, int(ts[mid])) >= b:
                hi = mid
            else:
                lo = mid + 1
        return lo

    def _dline(self, meta, b):
        """(base, r0, offsets, max offset) when bucket b's rows fit a data line
        (32-bit offsets first, else 16-bit ones), else None."""
        r0 = self._first_row(meta, meta["tab"][b], b)
        r1 = self._first_row(meta, meta["tab"][b + 1], b + 1)
        cnt = r1 - r0
        ts = meta["ts"]
        extra = 1 if self.build_mutant == "dl_cap" else 0
        base = int(ts[r0]) if cnt else 0
        span = ((int(ts[r1 - 1]) - base) & 0xFFFFFFFFFFFFFFFF) if cnt else 0
        for rows, top in ((DATA_ROWS, 0xFFFFFFFE), (DATA_ROWS16, 0xFFFE)):
            if cnt > rows + extra or (span > top and self.build_mutant != "dl_span"):
                continue
            n = min(cnt, rows)
            offs = [((int(ts[r0 + k]) - base) & 0xFFFFFFFFFFFFFFFF) & (top + 1) for k in range(n)]  # all-ones mask
            return base, r0, offs, top
        return None

    def probe(self, p, t, mutant=None):
        """Sorted position (within the partition) of the match, or -1."""
        meta = self.parts[p]
        if meta["K"] == 0 or t < meta["t0"]:
            return -1
        b = self._bucket(meta, t, f32=mutant == "router_f32")
        le = (lambda a, x: a < x) if mutant == "le_strict" else (lambda a, x: a <= x)
        dl = meta["dlines"][b]
        if dl is not None:
            base, r0, offs, top = dl
            c = 0
            if t >= base or mutant == "dl_below":
                dt = min((t - base) & 0xFFFFFFFFFFFFFFFF, top)
                c = sum(1 for o in offs if le(o, dt))
            return r0 + c - 1
        recs, more = meta["lines"][b]
        c0 = sum(1 for tau, _ in recs if le(tau, t))
        kn = meta["knots"]
        if c0 > 0 and not (more and c0 == BUCKET_RECS):
            piece = recs[c0 - 1][1]
        else:  # knot-range search, verified, else the whole knot list
            lo = max(meta["tab"][b] - 1, 0)
            hi = min(max(meta["tab"][b + 1] - 1, 0), meta["K"] - 1)
            lo = min(lo, hi)
            piece = self._knot_search(kn, lo, hi, t)
            if not (kn[piece][0] <= t and (piece + 1 >= meta["K"] or kn[piece + 1][0] > t)):
                piece = self._knot_search(kn, 0, meta["K"] - 1, t)
        tau, slope, err = kn[piece]
        a, e = meta["starts"][piece], meta["ends"][piece]
        cnt = e - a
        ts = meta["ts"]
        if cnt == 1:
            c = 1
        else:
            pr = piece_pred(tau, slope, cnt, t)
            if mutant == "window_err":
                wlo, whi = max(0, pr - err), min(cnt, pr + err)
                c = wlo + int(np.searchsorted(ts[a + wlo:a + whi], t, side="right"))
            else:  # the galloping search is exact wherever pr lands
                c = int(np.searchsorted(ts[a:e], t, side="right"))
        return a + c - 1 if c > 0 else -1

    @staticmethod
    def _knot_search(kn, lo, hi, t):
        while lo < hi:
            mid = lo + (hi - lo + 1) // 2
            if kn[mid][0] <= t:
                lo = mid
            else:
                hi = mid - 1
        return lo


def asof_learned(bsym, bts, psym, pts, B, mutant=None, tol=None, data_lines=True):
    """Per-probe matched build row (NO_ROW for none) through the learned index
    model: one partition per integer key value, NULL-free inputs."""
    bsym, bts = np.asarray(bsym, dtype=np.int64), np.asarray(bts, dtype=np.int64)
    kmin = int(bsym.min())
    nparts = int(bsym.max()) - kmin + 1
    order = np.lexsort((np.arange(len(bts)), bts, bsym))  # stable (partition, time)
    sym_s, ts_s = bsym[order], bts[order]
    bounds = np.searchsorted(sym_s, np.arange(kmin, kmin + nparts + 1))
    segs = [ts_s[bounds[i]:bounds[i + 1]] for i in range(nparts)]
    idx = LearnedIndex(segs, B, data_lines, mutant if mutant in ("dl_span", "dl_cap") else None)
    out = np.full(len(pts), NO_ROW, dtype=np.uint32)
    for j, (s, t) in enumerate(zip(np.asarray(psym, dtype=np.int64), np.asarray(pts, dtype=np.int64))):
        p = int(s) - kmin
        if p < 0 or p >= nparts:
            continue
        m = idx.probe(p, int(t), mutant)
        if m < 0:
            continue
        pos = bounds[p] + m
        if tol is not None and ((int(t) - int(ts_s[pos])) & 0xFFFFFFFFFFFFFFFF) > tol:
            continue
        out[j] = order[pos]
    return out
