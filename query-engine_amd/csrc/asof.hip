// Backward as-of join: qe_gpu_asof_build / qe_gpu_asof_probe replace
// tsx::AsofJoin (tsx/asof.h:101-105). Semantics (tsx/asof.h:10-39,
// tsx/asof.cpp:234-308): for a probe row with key tuple k and time t, the
// build row with the same key and the greatest tb <= t (the `<=` boundary);
// a NULL key or NULL time never matches; zero keys = one global partition;
// an optional tolerance drops the match when (u64)t - (u64)tb > tol.
//
// Build: each matchable build row (non-NULL keys and time) gets a partition id
// (direct key - min for one integer key of modest range, else a device hash
// table over the key tuple), rows are radix-sorted by (partition, time) on
// the device, and CSR offsets delimit each partition's sorted timestamps.
//
// Probe, two methods with identical results:
//  * binary  : upper_bound over the partition's sorted timestamps;
//  * learned : pieces of B sorted rows per partition with a linear position
//    model and a recorded max error; a radix router maps t to a small range of
//    pieces through a monotone MLP (1-16-16-1, clamp-ramp units, non-negative
//    weights after the first layer, trained on the device per partition); the
//    probe searches that piece range, then an error-bounded window inside the
//    piece, verifying both boundaries and falling back to binary search when a
//    bound does not hold.
// The router is exact by construction: the MLP is monotone in t and the same
// device function (explicit fmaf) builds the radix table and routes probes,
// so a key's bucket never precedes the bucket of an earlier key.
// Training runs the forward and backward passes as MFMA tiles
// (v_mfma_f32_16x16x4f32): the 16x16 hidden layer, its weight gradient
// (reduction over the batch) and the back-propagated activation gradient.
#include "agg_common.hpp"
#include "hash_common.hpp"
#include "tile_rank.hpp"

#include <algorithm>
#include <cmath>

using namespace qe_gpu;

namespace qe_gpu {

bool is_int_like(int t);

constexpr int kH = 16;
constexpr int kParams = kH + kH + kH * kH + kH + kH + 1;  // 321
constexpr int OFF_W1 = 0, OFF_B1 = 16, OFF_W2 = 32, OFF_B2 = 288, OFF_W3 = 304, OFF_B3 = 320;
constexpr uint32_t kMinKnotsForMlp = 64;

struct PartMeta {
    uint64_t seg;     // first sorted position
    uint32_t n;       // rows
    uint32_t K;       // knots (pieces)
    uint64_t kb;      // first knot index
    uint64_t tb;      // first router-table entry
    uint32_t R;       // router buckets
    int32_t mlp;      // router of the learned probe: 1 MLP warp grid, 0 linear
    int64_t t0;       // first knot time (the partition's first row)
    double inv_span;  // 1 / (last knot time - t0), 0 if single knot
    int32_t mlp_ok;   // enough knots (and a non-zero span) to train a router
    int32_t pad;
};

}  // namespace qe_gpu

struct qe_dev_asof {
    int nkeys = 0;
    int32_t key_types[kMaxKeys] = {};
    const qe_gpu_dict* key_dicts[kMaxKeys] = {};  // STR keys: the build side's dictionary
    int mode = 0;  // 0 zero keys, 1 direct, 2 hash
    int64_t kmin = 0;
    uint64_t range = 0;
    DeviceHashTable ht;       // hash mode: key tuple -> slot
    Scratch part_of_slot;     // hash mode: slot -> dense partition id
    uint64_t nparts = 0;
    uint64_t n = 0;           // matchable build rows
    uint64_t nrows_in = 0;    // build rows given (the original row ids index these)
    Scratch ts, row, offs;    // sorted timestamps, original rows, CSR (nparts+1)
    // learned index
    bool trained = false;
    uint32_t B = 0, kpb = 1;  // rows per piece, knots per router bucket
    uint64_t nknots = 0, ntab = 0;
    Scratch meta, knots, tab_mlp, tab_lin, params, grid;
    Scratch route;  // RouteRec per partition, from meta (rebuilt when the routers change)
    Scratch bkt_mlp, bkt_lin;  // per router-table entry: the bucket's knots inline (BucketLine)
    double final_loss = 0, mean_knots_mlp = 0, mean_knots_lin = 0;
    double mean_knots_cand = 0;             // the MLP router in every trained partition
    double sample_ms_mlp = 0, sample_ms_lin = 0;  // router timing sample (0: not timed)
    bool mlp_rejected = false;              // the timing kept the linear router
    uint64_t max_err = 0, mlp_candidates = 0, mlp_parts = 0;
    uint64_t data_lines = 0;  // buckets of the learned table (as probed) holding their rows inline
    Scratch fallback;  // u64 counter of the last probe
    uint64_t last_fallbacks = 0;
};

namespace qe_gpu {

// ---- partition ids ---------------------------------------------------------------
struct PartKeys {
    int mode;
    int nkeys;
    ColView key[kMaxKeys];
    int64_t kmin;
    uint64_t range;
    HashTableView t;
    const uint32_t* part_of_slot;
    // Partition of row r, or kNoRow (NULL key or absent / out of range).
    __device__ uint32_t part(uint64_t r) const {
        if (mode == 0) return 0;
        if (mode == 1) {
            if (!is_valid(key[0], r)) return kNoRow;
            const uint64_t off = static_cast<uint64_t>(int_value(key[0], r) - kmin);
            return off < range ? static_cast<uint32_t>(off) : kNoRow;
        }
        HashKeys hk;
        hk.nkeys = nkeys;
        for (int j = 0; j < nkeys; ++j) hk.key[j] = key[j];
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);
        if (nm != 0) return kNoRow;
        const uint32_t s = ht_find(t, nkeys, w, nm, h);
        return s == kNoSlot ? kNoRow : part_of_slot[s];
    }
};

__device__ inline int64_t time_value(const ColView& c, uint64_t r) { return int_value(c, r); }

__global__ void asof_slot_kernel(HashKeys hk, uint64_t n, HashTableView t) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);
        if (nm == 0) ht_insert(t, hk.nkeys, w, nm, h);
    }
}
__global__ void slot_flag_kernel(const uint32_t* state, uint64_t cap, uint32_t* flag) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < cap;
         s += uint64_t(gridDim.x) * blockDim.x)
        flag[s] = state[s] == 2 ? 1u : 0u;
}
__global__ void slot_part_kernel(const uint32_t* flag, const uint64_t* rank, uint64_t cap, uint32_t* part) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < cap;
         s += uint64_t(gridDim.x) * blockDim.x)
        part[s] = flag[s] ? static_cast<uint32_t>(rank[s]) : kNoRow;
}

// Matchable flag per build row (non-NULL key tuple and time).
__global__ void asof_flag_kernel(PartKeys pk, ColView time, uint64_t n, uint32_t* flag) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x)
        flag[r] = (is_valid(time, r) && pk.part(r) != kNoRow) ? 1u : 0u;
}
// Compact matchable rows into (key = part << tbits | (ts - tmin), value = row).
__global__ void asof_pack_kernel(PartKeys pk, ColView time, uint64_t n, const uint32_t* flag,
                                 const uint64_t* pos, int64_t tmin, unsigned tbits, uint64_t* keys, uint32_t* vals) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        if (!flag[r]) continue;
        const uint64_t p = pk.part(r);
        const uint64_t t = static_cast<uint64_t>(time_value(time, r) - tmin);
        keys[pos[r]] = (tbits >= 64) ? t : ((p << tbits) | t);
        vals[pos[r]] = static_cast<uint32_t>(r);
    }
}
// Unpack sorted keys into timestamps and the CSR offsets: offs[p] = first
// sorted position whose partition >= p.
__global__ void asof_unpack_kernel(const uint64_t* keys, uint64_t m, int64_t tmin, unsigned tbits, uint64_t nparts,
                                   int64_t* ts, uint64_t* offs) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < m;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t k = keys[i];
        const uint64_t mask = (tbits >= 64) ? ~0ull : ((uint64_t(1) << tbits) - 1);
        ts[i] = static_cast<int64_t>((k & mask) + static_cast<uint64_t>(tmin));
        const uint64_t p = (tbits >= 64) ? 0 : (k >> tbits);
        const uint64_t prev = (i == 0) ? 0 : (((tbits >= 64) ? 0 : (keys[i - 1] >> tbits)) + 1);
        for (uint64_t q = prev; q <= p; ++q) offs[q] = i;
        if (i == m - 1)
            for (uint64_t q = p + 1; q <= nparts; ++q) offs[q] = m;
    }
}
__global__ void fill_offs_kernel(uint64_t* offs, uint64_t nparts1, uint64_t v) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nparts1;
         i += uint64_t(gridDim.x) * blockDim.x)
        offs[i] = v;
}

static unsigned bits_for(uint64_t v) {  // bits needed to hold values 0..v
    unsigned b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

// ---- binary search probe ------------------------------------------------------------
// Largest i in [lo, hi) with ts[i] <= t, or hi-... returns `lo - 1` style via
// a flag: returns the count of entries <= t (upper_bound - lo).
__device__ inline uint64_t upper_bound_count(const int64_t* ts, uint64_t lo, uint64_t hi, int64_t t) {
    uint64_t first = lo, len = hi - lo;
    while (len > 0) {
        const uint64_t half = len >> 1;
        if (ts[first + half] <= t) {
            first += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return first - lo;
}

// Partition functors for the probe kernels: the general key tuple, and the two
// common shapes specialised so the kernels carry small arguments (fewer
// scalar registers, more waves in flight for these latency-bound searches).
struct ZeroKeyPart {
    __device__ uint32_t part(uint64_t) const { return 0; }
};
struct DirectKeyPart {  // one integer key, partition = key - kmin (probe keys: streamed reads)
    ColView key;
    int64_t kmin;
    uint64_t range;
    __device__ uint32_t part(uint64_t r) const {
        if (!is_valid(key, r)) return kNoRow;
        const uint64_t off = static_cast<uint64_t>(int_value_stream(key, r) - kmin);
        return off < range ? static_cast<uint32_t>(off) : kNoRow;
    }
};

struct GivenPart {  // partition ids given per probe (the router timing sample)
    const uint32_t* ids;
    __device__ uint32_t part(uint64_t r) const { return ids[r]; }
};

struct ProbeArgs {
    ColView time;
    uint64_t n;
    const int64_t* ts;
    const uint32_t* row;
    const uint64_t* offs;
    int has_tol;
    int sorted_out;  // report the sorted position m instead of the original row row[m]
    uint64_t tol;
    uint32_t* out;
};

__device__ inline void finish_match(const ProbeArgs& a, uint64_t j, int64_t t, uint64_t m, bool found) {
    uint32_t res = kNoRow;
    if (found) {
        const uint64_t span = static_cast<uint64_t>(t) - static_cast<uint64_t>(a.ts[m]);
        if (!a.has_tol || span <= a.tol) res = a.sorted_out ? static_cast<uint32_t>(m) : a.row[m];
    }
    __builtin_nontemporal_store(res, &a.out[j]);
}

template <class PK>
__global__ void asof_binary_kernel(PK pk, ProbeArgs a) {
    for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < a.n;
         j += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t p = is_valid(a.time, j) ? pk.part(j) : kNoRow;
        if (p == kNoRow) {
            __builtin_nontemporal_store(kNoRow, &a.out[j]);
            continue;
        }
        const int64_t t = time_value(a.time, j);
        const uint64_t lo = a.offs[p], hi = a.offs[p + 1];
        const uint64_t c = upper_bound_count(a.ts, lo, hi, t);
        finish_match(a, j, t, lo + c - 1, c > 0);
    }
}

// ---- learned index: pieces, router grid ------------------------------------------------
// Per partition: K pieces of B sorted rows (the spline knots, one 16-byte
// record each), a router of R = K buckets over normalised time, and a radix
// table tab[b] = first knot whose bucket >= b. The router is either linear
// (bucket = floor(x * R)) or a monotone warp y(x) given by kGrid + 1 values of
// the trained MLP (evaluated in batch with MFMA, made non-decreasing and
// normalised to [0, 1]) with linear interpolation inside a cell. The same
// device function routes knots (table build) and probes, and it is monotone
// in t, so the piece holding t always lies in the bucket's knot range.
constexpr int kGrid = 64;

__device__ inline float ramp(float z) { return fminf(fmaxf(z, 0.0f), 1.0f); }

struct Knot {
    int64_t tau;   // time of the piece's first row
    float slope;   // rows per time unit of the piece's linear position model
    uint32_t err;  // max |predicted - actual| position over the piece's rows
};

// A router bucket's knots inline in one 128-byte cache line, so the probe's
// routing reads one line instead of the table entry and then the knot records.
// Record 0 is the last knot of an earlier bucket (the piece of a time before
// the bucket's first knot; tau = INT64_MIN when there is none), records 1..7
// the bucket's first knots in order; unused records hold tau = INT64_MAX. The
// eight times come first (four 16-byte loads), then (slope, piece) pairs. A
// bucket of more than 7 knots sets kBucketMore in record 7's piece: a time at
// or past that knot is searched through the table and the knot records.
constexpr int kBucketRecs = 8;
constexpr uint32_t kBucketMore = 0x80000000u;
struct SlopePiece {
    float slope;
    uint32_t piece;  // knot index within the partition (| kBucketMore)
};
struct BucketLine {
    int64_t tau[kBucketRecs];
    SlopePiece sp[kBucketRecs];
};

// A bucket whose rows are few and close in time keeps the rows themselves in
// its line instead of knots: the first row's time, the bucket's first sorted
// position r0, and each row's time minus the first (ascending; unused entries
// all ones), as 28 32-bit offsets or, for up to 56 rows within 2^16 - 1 time
// units, 56 16-bit ones (kData16). The last word carries kDataLine [| kData16]
// | row count, in the place of a knot line's record-7 piece, which never has
// those bits while the index has fewer than 2^29 knots. The router is monotone
// in t and built the buckets, so every row before r0 lies in an earlier bucket
// and has time <= t: the match of a time routed here is sorted position
// r0 + c - 1, c the count of the bucket's rows with time <= t, with no further
// read.
constexpr int kDataRows = 28;
constexpr int kDataRows16 = 56;
constexpr uint32_t kDataLine = 0x40000000u;
constexpr uint32_t kData16 = 0x20000000u;
constexpr uint32_t kDataCount = 0xFFFFu;
constexpr uint32_t kDataPad = 0xFFFFFFFFu;
constexpr uint64_t kDataMaxSpan = 0xFFFFFFFEull;
constexpr uint64_t kDataMaxSpan16 = 0xFFFEull;
struct DataLine {
    int64_t base;
    uint32_t r0;
    uint32_t d[kDataRows];  // or uint16_t[kDataRows16] with kData16
    uint32_t flag;          // kDataLine [| kData16] | rows
};
static_assert(sizeof(BucketLine) == 128 && sizeof(DataLine) == 128, "one 128-byte line per bucket");

// Position inside a piece predicted for t >= tau, in [0, cnt). Non-decreasing
// in t: exact unsigned difference, monotone conversion and fmaf, clamps.
__device__ inline uint32_t piece_pred(int64_t tau, float slope, uint32_t cnt, int64_t t) {
    const float dt = static_cast<float>(static_cast<uint64_t>(t) - static_cast<uint64_t>(tau));
    float p = fmaf(dt, slope, 0.5f);
    if (!(p >= 0.0f)) p = 0.0f;
    const float hi = static_cast<float>(cnt - 1);
    if (p > hi) p = hi;
    return static_cast<uint32_t>(p);
}

// Normalised time in [0, 1] for t >= pm.t0. M: PartMeta or RouteRec (the
// fields t0, inv_span, R).
template <class M>
__device__ inline float norm_x(const M& pm, int64_t t) {
    double x = static_cast<double>(static_cast<uint64_t>(t) - static_cast<uint64_t>(pm.t0)) * pm.inv_span;
    x = x > 1.0 ? 1.0 : x;
    return static_cast<float>(x);
}

// Router output in [0, 1]: x itself (linear) or the interpolated warp grid W.
template <class M>
__device__ inline float route_y(const M& pm, const float* W, bool mlp, int64_t t) {
    const float x = norm_x(pm, t);
    if (!mlp) return x;
    const float s = x * static_cast<float>(kGrid);
    int g = static_cast<int>(s);
    if (g > kGrid - 1) g = kGrid - 1;
    const float fr = s - static_cast<float>(g);
    const float y = fmaf(W[g + 1] - W[g], fr, W[g]);
    return y < W[g + 1] ? y : W[g + 1];  // never past the next cell's start
}

template <class M>
__device__ inline uint32_t bucket_of(const M& pm, const float* W, bool mlp, int64_t t) {
    float b = floorf(route_y(pm, W, mlp, t) * static_cast<float>(pm.R));
    if (!(b >= 0.0f)) b = 0.0f;  // also maps NaN to 0
    if (b > static_cast<float>(pm.R - 1)) b = static_cast<float>(pm.R - 1);
    return static_cast<uint32_t>(b);
}

// Pieces cut a partition's sorted rows at the multiples of B of the whole
// sorted array, so a piece of at most 16 rows lies inside one 128-byte line of
// timestamps (the array starts 256-byte aligned) and its search reads that
// line alone. Piece j of a partition of n rows starting at seg covers
// [max(seg, (seg / B + j) B), min(seg + n, (seg / B + j + 1) B)); the first and
// last may be short.
__device__ inline uint32_t pieces_of(uint64_t seg, uint64_t n, uint32_t B) {
    return n ? static_cast<uint32_t>((seg + n + B - 1) / B - seg / B) : 0u;
}
__device__ inline uint64_t piece_start(uint64_t seg, uint32_t j, uint32_t B) {
    const uint64_t s = (seg / B + j) * B;
    return s > seg ? s : seg;
}
__device__ inline uint32_t piece_count(uint64_t seg, uint64_t n, uint32_t j, uint32_t B) {
    const uint64_t e = min(seg + n, (seg / B + j + 1) * B);
    return static_cast<uint32_t>(e - piece_start(seg, j, B));
}

// Knots K (pieces, as above) and buckets R = ceil(K / kpb) per partition.
__device__ inline uint32_t buckets_for(uint32_t k, uint32_t kpb) { return (k + kpb - 1) / kpb; }
__global__ void meta_counts_kernel(const uint64_t* offs, uint64_t nparts, uint32_t B, uint32_t kpb, uint32_t* K,
                                   uint32_t* R1) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < nparts;
         p += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t n = offs[p + 1] - offs[p];
        const uint32_t k = pieces_of(offs[p], n, B);
        K[p] = k;
        R1[p] = k ? buckets_for(k, kpb) + 1 : 0;  // R buckets -> R+1 table entries
    }
}
__global__ void meta_fill_kernel(const uint64_t* offs, uint64_t nparts, uint32_t B, uint32_t kpb, const uint32_t* K,
                                 const uint64_t* kb, const uint64_t* tb, const int64_t* ts, PartMeta* meta) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < nparts;
         p += uint64_t(gridDim.x) * blockDim.x) {
        PartMeta m{};
        m.seg = offs[p];
        m.n = static_cast<uint32_t>(offs[p + 1] - offs[p]);
        m.K = K[p];
        m.kb = kb[p];
        m.tb = tb[p];
        m.R = buckets_for(K[p], kpb);
        m.mlp_ok = K[p] >= kMinKnotsForMlp;
        m.mlp = 0;
        if (m.K) {
            m.t0 = ts[m.seg];
            const int64_t tl = ts[piece_start(m.seg, m.K - 1, B)];
            const uint64_t span = static_cast<uint64_t>(tl) - static_cast<uint64_t>(m.t0);
            m.inv_span = span ? 1.0 / static_cast<double>(span) : 0.0;
            if (!span) m.mlp_ok = 0;
        } else {
            m.t0 = 0;
            m.inv_span = 0.0;
        }
        meta[p] = m;
    }
}
__device__ inline uint32_t part_of_knot(const uint64_t* kb, uint64_t nparts, uint64_t g) {
    uint64_t lo = 0, hi = nparts;  // kb[lo] <= g < kb[hi]
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (kb[mid] <= g) lo = mid;
        else hi = mid;
    }
    return static_cast<uint32_t>(lo);
}
// Knot records: first time, slope and recorded max error of each piece.
__global__ void pieces_kernel(const PartMeta* meta, const uint64_t* kb, uint64_t nparts, uint64_t nknots, uint32_t B,
                              const int64_t* ts, Knot* knots, unsigned long long* max_err) {
    uint64_t local_max = 0;
    for (uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < nknots;
         g += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t p = part_of_knot(kb, nparts + 1, g);
        const PartMeta pm = meta[p];
        const uint32_t j = static_cast<uint32_t>(g - pm.kb);
        const uint64_t pos0 = piece_start(pm.seg, j, B);
        const uint32_t cnt = piece_count(pm.seg, pm.n, j, B);
        const int64_t t0 = ts[pos0], tl = ts[pos0 + cnt - 1];
        const uint64_t span = static_cast<uint64_t>(tl) - static_cast<uint64_t>(t0);
        const float s = (cnt > 1 && span) ? static_cast<float>(static_cast<double>(cnt - 1) / static_cast<double>(span))
                                          : 0.0f;
        uint32_t e = 0;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t pr = piece_pred(t0, s, cnt, ts[pos0 + i]);
            const uint32_t d = pr > i ? pr - i : i - pr;
            e = d > e ? d : e;
        }
        knots[g] = Knot{t0, s, e};
        local_max = e > local_max ? e : local_max;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint64_t o = __shfl_down(local_max, off);
        local_max = o > local_max ? o : local_max;
    }
    if (lane_id() == 0) atomicMax(max_err, static_cast<unsigned long long>(local_max));
}

// Router choice for a table: 0 linear everywhere, 1 the partition's chosen
// router, 2 the MLP wherever one was trained.
__device__ inline bool router_mlp(const PartMeta& pm, int which) {
    return which == 1 ? pm.mlp != 0 : (which == 2 ? pm.mlp_ok != 0 : false);
}

// Radix router table: tab[tb + b] = first knot whose bucket >= b (b in 0..R).
__global__ void router_table_kernel(const PartMeta* meta, const uint64_t* kb, uint64_t nparts, uint64_t nknots,
                                    const Knot* knots, const float* grid, int which, uint32_t* tab) {
    for (uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < nknots;
         g += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t p = part_of_knot(kb, nparts + 1, g);
        const PartMeta pm = meta[p];
        const uint32_t j = static_cast<uint32_t>(g - pm.kb);
        const bool mlp = router_mlp(pm, which);
        const float* W = grid + size_t(p) * (kGrid + 1);
        const int64_t bj = bucket_of(pm, W, mlp, knots[g].tau);
        const int64_t bprev = (j == 0) ? -1 : static_cast<int64_t>(bucket_of(pm, W, mlp, knots[g - 1].tau));
        for (int64_t b = bprev + 1; b <= bj; ++b) tab[pm.tb + b] = j;
        if (j == pm.K - 1)
            for (int64_t b = bj + 1; b <= static_cast<int64_t>(pm.R); ++b) tab[pm.tb + b] = pm.K;
    }
}

// First sorted row of the partition whose bucket is >= b, given k = tab[b]
// (the first knot whose bucket is >= b): knot k - 1 routes below b, so the row
// lies after that knot's row and at or before knot k's (the partition's end
// when k = K); a binary search of at most one piece.
__device__ inline uint64_t first_row_from_bucket(const PartMeta& pm, const float* W, bool mlp, const int64_t* ts,
                                                 uint32_t B, uint32_t k, uint32_t b) {
    if (k == 0) return pm.seg;
    uint64_t lo = piece_start(pm.seg, k - 1, B) + 1;
    uint64_t hi = k < pm.K ? piece_start(pm.seg, k, B) : pm.seg + pm.n;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (bucket_of(pm, W, mlp, ts[mid]) >= b) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Inline bucket lines from the router table (entries tb .. tb + R - 1 of each
// partition; entry tb + R is the table's end marker and gets no line): a data
// line where data_lines is set and the bucket's rows fit one, else its knots.
// which: the router the table was built with (router_mlp). ndata counts the
// data lines.
__global__ void bucket_lines_kernel(const PartMeta* meta, const uint64_t* tbs, uint64_t nparts, uint64_t nt,
                                    const Knot* knots, const uint32_t* tab, const int64_t* ts, const float* grid,
                                    int which, uint32_t B, int data_lines, BucketLine* out,
                                    unsigned long long* ndata) {
    uint32_t nd = 0;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nt;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t p = part_of_knot(tbs, nparts + 1, i);
        const PartMeta pm = meta[p];
        const uint32_t b = static_cast<uint32_t>(i - pm.tb);
        if (b >= pm.R) continue;
        const uint32_t lo = tab[i], hi = tab[i + 1];
        if (data_lines) {
            const float* W = grid + size_t(p) * (kGrid + 1);
            const bool mlp = router_mlp(pm, which);
            const uint64_t r0 = first_row_from_bucket(pm, W, mlp, ts, B, lo, b);
            const uint64_t r1 = first_row_from_bucket(pm, W, mlp, ts, B, hi, b + 1);
            const uint64_t cnt = r1 - r0;
            const uint64_t span = cnt ? static_cast<uint64_t>(ts[r1 - 1]) - static_cast<uint64_t>(ts[r0]) : 0;
            const bool w32 = cnt <= kDataRows && span <= kDataMaxSpan;
            if (w32 || (cnt <= kDataRows16 && span <= kDataMaxSpan16)) {
                DataLine D;
                D.base = cnt ? ts[r0] : 0;
                D.r0 = static_cast<uint32_t>(r0);
                auto off = [&](uint32_t k) {
                    return static_cast<uint32_t>(static_cast<uint64_t>(ts[r0 + k]) - static_cast<uint64_t>(D.base));
                };
                for (uint32_t k = 0; k < kDataRows; ++k) {
                    if (w32) {
                        D.d[k] = k < cnt ? off(k) : kDataPad;
                    } else {  // rows 2k and 2k + 1 in the low and high halves
                        const uint32_t lo = 2 * k < cnt ? off(2 * k) : 0xFFFFu;
                        const uint32_t hi = 2 * k + 1 < cnt ? off(2 * k + 1) : 0xFFFFu;
                        D.d[k] = lo | (hi << 16);
                    }
                }
                D.flag = kDataLine | (w32 ? 0u : kData16) | static_cast<uint32_t>(cnt);
                *reinterpret_cast<DataLine*>(out + i) = D;
                ++nd;
                continue;
            }
        }
        const Knot* kp = knots + pm.kb;
        BucketLine L;
        SlopePiece* sp = L.sp;
        L.tau[0] = lo > 0 ? kp[lo - 1].tau : INT64_MIN;
        sp[0] = SlopePiece{lo > 0 ? kp[lo - 1].slope : 0.0f, lo > 0 ? lo - 1 : 0u};
        for (uint32_t k = 1; k < kBucketRecs; ++k) {
            const uint32_t j = lo + k - 1;
            L.tau[k] = j < hi ? kp[j].tau : INT64_MAX;
            sp[k] = SlopePiece{j < hi ? kp[j].slope : 0.0f, j < hi ? j : 0u};
        }
        if (hi - lo > kBucketRecs - 1) sp[kBucketRecs - 1].piece |= kBucketMore;
        out[i] = L;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) nd += __shfl_down(nd, off);
    if (lane_id() == 0 && nd) atomicAdd(ndata, static_cast<unsigned long long>(nd));
}

// Knots a probe searches, summed per partition and overall (probes assumed to
// be distributed like the knots): for each knot, its bucket's occupancy + 1
// (the predecessor boundary). Partitions are contiguous knot ranges, so a wave
// whose lanes share one partition adds its sum with one atomic.
__global__ void router_stats_kernel(const PartMeta* meta, const uint64_t* kb, uint64_t nparts, uint64_t nknots,
                                    const Knot* knots, const float* grid, int which, const uint32_t* tab,
                                    double* per_part, double* total) {
    double local = 0;
    for (uint64_t g0 = blockIdx.x * uint64_t(blockDim.x); g0 < nknots; g0 += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t g = g0 + threadIdx.x;
        uint32_t p = 0xFFFFFFFFu;
        double v = 0;
        if (g < nknots) {
            p = part_of_knot(kb, nparts + 1, g);
            const PartMeta pm = meta[p];
            const uint32_t b = bucket_of(pm, grid + size_t(p) * (kGrid + 1), router_mlp(pm, which), knots[g].tau);
            v = static_cast<double>(tab[pm.tb + b + 1] - tab[pm.tb + b]) + 1.0;
        }
        local += v;
        if (per_part) {
            const uint32_t p0 = __shfl(p, 0);
            const bool same = __all(p == p0 || p == 0xFFFFFFFFu);
            if (same) {
                double s = v;
                for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_down(s, off);
                if (lane_id() == 0 && p0 != 0xFFFFFFFFu) atomicAdd(&per_part[p0], s);
            } else if (p != 0xFFFFFFFFu) {
                atomicAdd(&per_part[p], v);
            }
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) local += __shfl_down(local, off);
    if (lane_id() == 0) atomicAdd(total, local);
}

// Candidate learned router: the MLP where it searches fewer knots than the
// linear one. Whether the candidate is kept is decided by timing it against
// the linear router on a sample of the index's own rows (qe_gpu_asof_train).
__global__ void router_choose_kernel(PartMeta* meta, uint64_t nparts, const double* occ_lin, const double* occ_mlp,
                                     unsigned long long* chosen) {
    uint64_t c = 0;
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < nparts;
         p += uint64_t(gridDim.x) * blockDim.x) {
        const bool use = meta[p].mlp_ok && occ_mlp[p] < occ_lin[p];
        meta[p].mlp = use ? 1 : 0;
        c += use;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_down(c, off);
    if (lane_id() == 0 && c) atomicAdd(chosen, static_cast<unsigned long long>(c));
}

// The fields of PartMeta that route a time to its table entry, in 32 bytes
// (two 16-byte loads per probe row instead of four).
struct RouteRec {
    int64_t t0;
    double inv_span;
    uint64_t tb;
    uint32_t R;
    uint32_t flags;  // kRouteKnots: the partition has knots; kRouteMlp: the MLP router
};
constexpr uint32_t kRouteKnots = 1, kRouteMlp = 2;
static_assert(sizeof(RouteRec) == 32, "two 16-byte loads");

__global__ void route_recs_kernel(const PartMeta* meta, uint64_t nparts, RouteRec* out) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < nparts;
         p += uint64_t(gridDim.x) * blockDim.x) {
        const PartMeta pm = meta[p];
        out[p] = RouteRec{pm.t0, pm.inv_span, pm.tb, pm.R,
                          (pm.K != 0 ? kRouteKnots : 0u) | (pm.mlp != 0 ? kRouteMlp : 0u)};
    }
}

// ---- learned probe ---------------------------------------------------------------------
struct LearnedArgs {
    const RouteRec* route;  // per partition: what routes a time to its bucket line
    const PartMeta* meta;   // the rest, read on the knot-line path only
    const Knot* knots;
    const uint32_t* tab;
    const BucketLine* bkt;
    const float* grid;
    int use_mlp;
    uint32_t B;
    unsigned long long* fallbacks;
};

// Number of elements of the sorted run ts[base, base + cnt) that are <= t,
// searched outward from the predicted index p < cnt: steps of 1, 2, 4, ... away
// from p until the answer is bracketed, then a binary search of the bracket.
// Exact for any p; the elements read lie within twice the prediction's error of
// p, so a close prediction reads one or two cache lines.
__device__ inline uint32_t gallop_count(const int64_t* ts, uint64_t base, uint32_t cnt, uint32_t p, int64_t t) {
    uint32_t lo, hi;  // ts[base + lo - 1] <= t (or lo == 0) and ts[base + hi] > t (or hi == cnt)
    if (ts[base + p] <= t) {
        lo = p + 1;
        hi = cnt;
        for (uint32_t step = 1; p + step < cnt; step <<= 1) {
            if (ts[base + p + step] > t) {
                hi = p + step;
                break;
            }
            lo = p + step + 1;
        }
    } else {
        lo = 0;
        hi = p;
        for (uint32_t step = 1; step <= p; step <<= 1) {
            if (ts[base + p - step] <= t) {
                lo = p - step + 1;
                break;
            }
            hi = p - step;
        }
    }
    return lo + static_cast<uint32_t>(upper_bound_count(ts, base + lo, base + hi, t));
}

// Largest knot j in [lo, hi] with tau[kb+j] <= t (lo itself when tau[kb+lo] > t).
__device__ inline uint32_t knot_search(const Knot* knots, uint64_t kb, uint32_t lo, uint32_t hi, int64_t t) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (knots[kb + mid].tau <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class PK>
__global__ void asof_learned_kernel(PK pk, ProbeArgs a, LearnedArgs L) {
    uint32_t fb = 0;
    for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < a.n;
         j += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t p = is_valid(a.time, j) ? pk.part(j) : kNoRow;
        if (p == kNoRow) {
            __builtin_nontemporal_store(kNoRow, &a.out[j]);
            continue;
        }
        const int64_t t = int_value_stream(a.time, j);
        const RouteRec rr = L.route[p];
        if (!(rr.flags & kRouteKnots) || t < rr.t0) {  // before the partition's first row
            __builtin_nontemporal_store(kNoRow, &a.out[j]);
            continue;
        }
        // 1. route to a bucket and locate the piece (tau[piece] <= t < tau[piece+1]).
        //    The router is monotone in t and built the buckets, so the piece is
        //    the last of the bucket's records (its knots and the predecessor)
        //    with tau <= t.
        const bool mlp = L.use_mlp && (rr.flags & kRouteMlp);
        const uint32_t b = bucket_of(rr, L.grid + size_t(p) * (kGrid + 1), mlp, t);
        // The whole 128-byte line in eight 16-byte loads.
        const uint4* lp = reinterpret_cast<const uint4*>(L.bkt + (rr.tb + b));
        uint32_t w[32];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint4 v = lp[k];
            w[4 * k] = v.x;
            w[4 * k + 1] = v.y;
            w[4 * k + 2] = v.z;
            w[4 * k + 3] = v.w;
        }
        if (w[31] & kDataLine) {
            // Data line (DataLine): count the bucket's rows with time <= t.
            const int64_t base = static_cast<int64_t>((static_cast<uint64_t>(w[1]) << 32) | w[0]);
            uint32_t c = 0, dsel = 0;
            if (t >= base) {
                const uint64_t dt64 = static_cast<uint64_t>(t) - static_cast<uint64_t>(base);
                if (!(w[31] & kData16)) {
                    const uint32_t dt = dt64 < kDataMaxSpan ? static_cast<uint32_t>(dt64)
                                                            : static_cast<uint32_t>(kDataMaxSpan);
#pragma unroll
                    for (int k = 0; k < kDataRows; ++k) {
                        const bool le = w[3 + k] <= dt;  // ascending; padding never <= dt
                        c += le ? 1u : 0u;
                        dsel = le ? w[3 + k] : dsel;
                    }
                } else {
                    const uint32_t dt = dt64 < kDataMaxSpan16 ? static_cast<uint32_t>(dt64)
                                                              : static_cast<uint32_t>(kDataMaxSpan16);
#pragma unroll
                    for (int k = 0; k < kDataRows; ++k) {
                        const uint32_t lo = w[3 + k] & 0xFFFFu, hi = w[3 + k] >> 16;
                        const bool le0 = lo <= dt, le1 = hi <= dt;
                        c += (le0 ? 1u : 0u) + (le1 ? 1u : 0u);
                        dsel = le1 ? hi : (le0 ? lo : dsel);
                    }
                }
            }
            // t >= the partition's first time, so r0 + c >= 1 (an empty bucket
            // or a time before its first row matches the row before r0).
            const uint64_t m = static_cast<uint64_t>(w[2]) + c - 1;
            uint32_t res = a.sorted_out ? static_cast<uint32_t>(m) : kNoRow;
            if (a.has_tol) {
                const int64_t tm = c ? static_cast<int64_t>(static_cast<uint64_t>(base) + dsel) : a.ts[m];
                if (static_cast<uint64_t>(t) - static_cast<uint64_t>(tm) > a.tol) {
                    __builtin_nontemporal_store(kNoRow, &a.out[j]);
                    continue;
                }
            }
            if (!a.sorted_out) res = a.row[m];
            __builtin_nontemporal_store(res, &a.out[j]);
            continue;
        }
        const PartMeta pm = L.meta[p];
        int64_t rt[kBucketRecs];
#pragma unroll
        for (int k = 0; k < kBucketRecs; ++k)
            rt[k] = static_cast<int64_t>((static_cast<uint64_t>(w[2 * k + 1]) << 32) | w[2 * k]);
        uint32_t c0 = 0;
#pragma unroll
        for (int k = 0; k < kBucketRecs; ++k) c0 += rt[k] <= t ? 1u : 0u;
        SlopePiece sp{0.0f, kBucketMore};
#pragma unroll
        for (int k = 0; k < kBucketRecs; ++k)
            if (static_cast<uint32_t>(k) + 1 == c0) sp = SlopePiece{__uint_as_float(w[16 + 2 * k]), w[17 + 2 * k]};
        uint32_t piece;
        int64_t tau;
        float slope;
        if (!(sp.piece & kBucketMore)) {
            piece = sp.piece;
            tau = rt[0];
#pragma unroll
            for (int k = 1; k < kBucketRecs; ++k)
                if (static_cast<uint32_t>(k) == c0 - 1) tau = rt[k];
            slope = sp.slope;
        } else {
            // At or past the 8th record of a bucket of more than 7 knots: search its knot range, then verify
            // the piece by its own bound and its successor's (the records the
            // search ended on), else search the partition's whole knot list.
            const uint32_t t_lo = L.tab[pm.tb + b], t_hi = L.tab[pm.tb + b + 1];
            uint32_t lo = t_lo > 0 ? t_lo - 1 : 0;
            uint32_t hi = t_hi > 0 ? t_hi - 1 : 0;
            if (hi >= pm.K) hi = pm.K - 1;
            if (lo > hi) lo = hi;
            piece = knot_search(L.knots, pm.kb, lo, hi, t);
            Knot kn = L.knots[pm.kb + piece];
            if (!(kn.tau <= t && (piece + 1 >= pm.K || L.knots[pm.kb + piece + 1].tau > t))) {
                ++fb;
                piece = knot_search(L.knots, pm.kb, 0, pm.K - 1, t);
                kn = L.knots[pm.kb + piece];
            }
            tau = kn.tau;
            slope = kn.slope;
        }
        // 2. inside the piece, search outward from the spline's prediction; the
        //    result is exact wherever the prediction lands.
        const uint64_t base = piece_start(pm.seg, piece, L.B);
        const uint32_t cnt = piece_count(pm.seg, pm.n, piece, L.B);
        // A one-row piece is its knot: tau <= t already places the match there,
        // with no read of the timestamps.
        const uint32_t c = cnt == 1 ? 1u : gallop_count(a.ts, base, cnt, piece_pred(tau, slope, cnt, t), t);
        // tau[piece] <= t guarantees c >= 1.
        finish_match(a, j, t, base + c - 1, c > 0);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) fb += __shfl_down(fb, off);
    if (lane_id() == 0 && fb) atomicAdd(L.fallbacks, static_cast<unsigned long long>(fb));
}

// ---- probe-side clustering --------------------------------------------------------------
// The reference's AsofJoin sorts its probe child by (key, time) before merging
// (tsx/asof.cpp:163-166). Here the probe table is instead partitioned in one
// pass into kClGroups groups by the index region its probes will read: the
// learned router's line address (the table entry of the probe's bucket), or,
// untrained, the partition's first sorted position. Lines are laid out by
// partition and then bucket, so a group's probes read one contiguous 1/256 of
// the index, which keeps their random line reads within what the GPU's address
// translation covers. Rows that probe nothing (NULL key or time, no
// partition) go to group 0. Within a group rows keep their input order.
constexpr unsigned kClItems = 16;
constexpr unsigned kClTile = kRankBlock * kClItems;  // 4096 rows
constexpr unsigned kClGroups = kRankDigits;
constexpr int kClMaxCols = 8;

struct ClusterArgs {
    ColView time;
    uint64_t n;
    const RouteRec* route;  // nullptr: untrained index
    const float* grid;
    const uint64_t* offs;  // partition offsets into the sorted rows
    uint64_t nsorted;      // sorted rows
    uint64_t ntab;         // router-table entries (line addresses < ntab)
    uint8_t* grp;
    uint32_t* hist;        // hist[g * ntiles + tile]
    uint64_t ntiles;
};

template <class PK>
__global__ void __launch_bounds__(kRankBlock) cluster_group_kernel(PK pk, ClusterArgs a) {
    __shared__ uint32_t h[kClGroups];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = uint64_t(blockIdx.x) * kClTile;
    // Every row's partition and time first, so their loads are in flight together.
    uint32_t part[kClItems];
    int64_t tv[kClItems];
#pragma unroll
    for (unsigned i = 0; i < kClItems; ++i) {
        const uint64_t j = base + uint64_t(i) * kRankBlock + threadIdx.x;
        const bool in = j < a.n && is_valid(a.time, j);
        part[i] = in ? pk.part(j) : kNoRow;
        tv[i] = in ? time_value(a.time, j) : 0;
    }
#pragma unroll
    for (unsigned i = 0; i < kClItems; ++i) {
        const uint64_t j = base + uint64_t(i) * kRankBlock + threadIdx.x;
        if (j >= a.n) continue;
        uint64_t g = 0;
        const uint32_t p = part[i];
        if (p != kNoRow) {
            if (a.route) {
                const RouteRec rr = a.route[p];
                uint64_t line = rr.tb;
                if ((rr.flags & kRouteKnots) && tv[i] >= rr.t0)
                    line += bucket_of(rr, a.grid + size_t(p) * (kGrid + 1), (rr.flags & kRouteMlp) != 0, tv[i]);
                g = line * kClGroups / a.ntab;
            } else {
                g = a.nsorted ? a.offs[p] * kClGroups / a.nsorted : 0;
            }
            g = g < kClGroups ? g : kClGroups - 1;
        }
        a.grp[j] = static_cast<uint8_t>(g);
        atomicAdd(&h[g], 1u);
    }
    __syncthreads();
    a.hist[uint64_t(threadIdx.x) * a.ntiles + blockIdx.x] = h[threadIdx.x];
}

struct ClusterCol {
    const void* src;
    const uint64_t* svalid;  // nullptr: all valid
    void* dst;
    uint64_t* dvalid;        // preset to all ones; NULL rows cleared here
    int32_t width;           // 1, 4 or 8 bytes
};
struct ClusterCols {
    int32_t n;
    ClusterCol c[kClMaxCols];
};

__device__ inline uint64_t load_width(const void* p, int width, uint64_t r) {
    if (width == 8) return static_cast<const uint64_t*>(p)[r];
    if (width == 4) return static_cast<const uint32_t*>(p)[r];
    return static_cast<const uint8_t*>(p)[r];
}
__device__ inline void store_width(void* p, int width, uint64_t r, uint64_t v) {
    if (width == 8) static_cast<uint64_t*>(p)[r] = v;
    else if (width == 4) static_cast<uint32_t*>(p)[r] = static_cast<uint32_t>(v);
    else static_cast<uint8_t*>(p)[r] = static_cast<uint8_t>(v);
}

// Each tile ranks its rows stably by group, stages every column in LDS in
// group order and writes each group's run contiguously at its offset.
__global__ void __launch_bounds__(kRankBlock) cluster_scatter_kernel(const uint8_t* grp, uint64_t n,
                                                                     const uint64_t* offs, uint64_t ntiles,
                                                                     ClusterCols C) {
    __shared__ TileRankLds L;
    __shared__ uint64_t sval[kClTile];
    __shared__ uint16_t srank[kClTile];
    __shared__ uint8_t sgrp[kClTile];
    __shared__ uint32_t snull[kClTile / 32];  // NULL flag of each staged row (bitmask)
    __shared__ uint64_t goff[kClGroups];      // this tile's output offset of each group
    const uint64_t tile = die_tile(ntiles);
    if (tile >= ntiles) return;
    const uint64_t base = tile * kClTile;
    goff[threadIdx.x] = offs[uint64_t(threadIdx.x) * ntiles + tile];  // read before the ranking's barriers
    uint32_t d[kClItems];
    uint32_t at[kClItems];
    uint64_t v[kClItems];  // the next column's values, loaded ahead of its staging
    const uint32_t live_n = static_cast<uint32_t>(n - base < kClTile ? n - base : kClTile);
    auto live = [&](unsigned i) { return tile_row<kClItems>(i) < live_n; };
    auto load_col = [&](int c) {
        for (unsigned i = 0; i < kClItems; ++i)
            v[i] = live(i) ? load_width(C.c[c].src, C.c[c].width, base + tile_row<kClItems>(i)) : 0ull;
    };
    for (unsigned i = 0; i < kClItems; ++i) d[i] = live(i) ? grp[base + tile_row<kClItems>(i)] : 0u;
    load_col(0);  // in flight during the ranking
    auto digit = [&](unsigned i) { return d[i]; };
    tile_digit_ranks<kClItems>(L, digit, live, at);
    for (unsigned i = 0; i < kClItems; ++i) {
        if (!live(i)) continue;
        sgrp[at[i]] = static_cast<uint8_t>(d[i]);
        srank[at[i]] = static_cast<uint16_t>(at[i] - L.start[d[i]]);
    }
    for (int c = 0; c < C.n; ++c) {
        const ClusterCol& col = C.c[c];
        __syncthreads();  // the previous column's staging has been written out
        if (col.dvalid) {
            if (threadIdx.x < kClTile / 32) snull[threadIdx.x] = 0;
            __syncthreads();
        }
        for (unsigned i = 0; i < kClItems; ++i) {
            if (!live(i)) continue;
            const uint64_t j = base + tile_row<kClItems>(i);
            sval[at[i]] = v[i];
            if (col.dvalid && col.svalid && !((col.svalid[j >> 6] >> (j & 63)) & 1ull))
                atomicOr(&snull[at[i] >> 5], 1u << (at[i] & 31));
        }
        if (c + 1 < C.n) load_col(c + 1);  // in flight during this column's write-out
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < live_n; s += kRankBlock) {
            const uint64_t dst = goff[sgrp[s]] + srank[s];
            store_width(col.dst, col.width, dst, sval[s]);
            if (col.dvalid && ((snull[s >> 5] >> (s & 31)) & 1u))
                atomicAnd(reinterpret_cast<unsigned long long*>(&col.dvalid[dst >> 6]), ~(1ull << (dst & 63)));
        }
    }
}

// Probes for timing the routers: n_s rows of the index drawn by a hashed
// stride (unsorted, like real probes), each probing its own time, so every
// probe matches and the sample follows the rows' distribution over partitions
// and time.
__global__ void router_sample_kernel(const uint64_t* offs, uint64_t nparts, const int64_t* ts, uint64_t n,
                                     uint64_t n_s, uint32_t* ids, int64_t* t) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n_s;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t q = mix64(i * 0x9E3779B97F4A7C15ull + 0x51E57A11ull) % n;
        uint64_t lo = 0, hi = nparts;  // offs[lo] <= q < offs[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (offs[mid] <= q) lo = mid;
            else hi = mid;
        }
        ids[i] = static_cast<uint32_t>(lo);
        t[i] = ts[q];
    }
}

__global__ void clear_mlp_kernel(PartMeta* meta, uint64_t nparts) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < nparts;
         p += uint64_t(gridDim.x) * blockDim.x)
        meta[p].mlp = 0;
}

// ---- MLP training (MFMA) ---------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ inline floatx4 mfma16(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Gradients of the batch MSE (1/N) sum (y - c)^2 into g[kParams] (LDS, zeroed by
// the caller) and the summed squared error into *loss. P = parameters (LDS).
// Each wave takes 16-sample tiles. Lane l = 16q + r:
//   forward  Z2[s][m] = sum_k H1[s][k] W2[m][k] + b2[m]  (4 MFMAs, D: s = 4q+v, m = r)
//   backward dW2[m][k] += sum_s dZ2[s][m] H1[s][k]       (4 MFMAs, D: m = 4q+v, k = r)
//            dH1[s][k] = sum_m dZ2[s][m] W2[m][k]        (4 MFMAs, D: s = 4q+v, k = r)
__device__ void mlp_batch_grads(const float* P, const float* xs, const float* cs, uint32_t N, float* g, float* loss) {
    const unsigned lane = lane_id();
    const unsigned q = lane / 16, r = lane % 16;
    const unsigned wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
    const float invN = 1.0f / static_cast<float>(N);
    floatx4 dW2acc = {0, 0, 0, 0};
    float gw1 = 0, gb1 = 0, gb2 = 0, gw3 = 0, gb3 = 0, lsum = 0;
    const uint32_t ntiles = (N + 15) / 16;
    for (uint32_t tile = wave; tile < ntiles; tile += nwaves) {
        const uint32_t s0 = tile * 16;
        // ---- forward ----
        floatx4 z2;
        for (int v = 0; v < 4; ++v) z2[v] = P[OFF_B2 + r];
        for (int c = 0; c < 4; ++c) {
            const uint32_t s = s0 + r;
            const float x = s < N ? xs[s] : 0.0f;
            const int k = 4 * c + q;
            const float a = ramp(fmaf(P[OFF_W1 + k], x, P[OFF_B1 + k]));  // H1[s=r][k]
            const float bb = P[OFF_W2 + r * kH + k];                       // W2[m=r][k]
            z2 = mfma16(a, bb, z2);
        }
        float h2[4], y[4];
        for (int v = 0; v < 4; ++v) {
            h2[v] = ramp(z2[v]);
            float pv = P[OFF_W3 + r] * h2[v];
            for (int off = 8; off > 0; off >>= 1) pv += __shfl_xor(pv, off, 16);
            y[v] = pv + P[OFF_B3];
        }
        float dz2[4];
        for (int v = 0; v < 4; ++v) {
            const uint32_t s = s0 + 4 * q + v;
            const bool valid = s < N;
            const float diff = valid ? y[v] - cs[s] : 0.0f;
            const float dy = 2.0f * diff * invN;
            if (r == 0) {
                lsum += diff * diff;
                gb3 += dy;
            }
            gw3 += dy * h2[v];
            const float z = z2[v];
            dz2[v] = (z > 0.0f && z < 1.0f) ? dy * P[OFF_W3 + r] : 0.0f;
            gb2 += dz2[v];
        }
        // ---- dW2 += dZ2^T H1 ----
        for (int c = 0; c < 4; ++c) {
            // A[m=r][s'=q] = dZ2[4c+q][r], held by lane 16c + r in register q.
            const int src = 16 * c + r;
            const float t0 = __shfl(dz2[0], src), t1 = __shfl(dz2[1], src), t2 = __shfl(dz2[2], src),
                        t3 = __shfl(dz2[3], src);
            const float a = q == 0 ? t0 : (q == 1 ? t1 : (q == 2 ? t2 : t3));
            const uint32_t s = s0 + 4 * c + q;
            const float x = s < N ? xs[s] : 0.0f;
            const float bb = ramp(fmaf(P[OFF_W1 + r], x, P[OFF_B1 + r]));  // H1[4c+q][k=r]
            dW2acc = mfma16(a, bb, dW2acc);
        }
        // ---- dH1 = dZ2 W2 ----
        floatx4 dh1 = {0, 0, 0, 0};
        for (int c = 0; c < 4; ++c) {
            // A[s=r][m'=q] = dZ2[r][4c+q], held by lane 16*(r/4) + 4c+q in register r%4.
            const int src = 16 * (r / 4) + 4 * c + q;
            const float t0 = __shfl(dz2[0], src), t1 = __shfl(dz2[1], src), t2 = __shfl(dz2[2], src),
                        t3 = __shfl(dz2[3], src);
            const unsigned sel = r % 4;
            const float a = sel == 0 ? t0 : (sel == 1 ? t1 : (sel == 2 ? t2 : t3));
            const float bb = P[OFF_W2 + (4 * c + q) * kH + r];  // W2[m=4c+q][k=r]
            dh1 = mfma16(a, bb, dh1);
        }
        for (int v = 0; v < 4; ++v) {
            const uint32_t s = s0 + 4 * q + v;
            const float x = s < N ? xs[s] : 0.0f;
            const float z1 = fmaf(P[OFF_W1 + r], x, P[OFF_B1 + r]);
            const float dz1 = (z1 > 0.0f && z1 < 1.0f) ? dh1[v] : 0.0f;
            gw1 += dz1 * x;
            gb1 += dz1;
        }
    }
    // Reduce lane partials that share r (the 4 groups q) and fold into g.
    for (int off = 16; off < 64; off <<= 1) {
        gw1 += __shfl_xor(gw1, off);
        gb1 += __shfl_xor(gb1, off);
        gb2 += __shfl_xor(gb2, off);
        gw3 += __shfl_xor(gw3, off);
    }
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off);
        lsum += __shfl_xor(lsum, off);
    }
    if (q == 0) {
        atomicAdd(&g[OFF_W1 + r], gw1);
        atomicAdd(&g[OFF_B1 + r], gb1);
        atomicAdd(&g[OFF_B2 + r], gb2);
        atomicAdd(&g[OFF_W3 + r], gw3);
    }
    if (lane == 0) {
        atomicAdd(&g[OFF_B3], gb3);
        atomicAdd(loss, lsum);
    }
    for (int v = 0; v < 4; ++v) atomicAdd(&g[OFF_W2 + (4 * q + v) * kH + r], dW2acc[v]);  // dW2[m=4q+v][k=r]
}

__device__ inline void init_params(float* P) {
    // Identity router: unit k ramps over [k/16, (k+1)/16]; layer 2 passes it
    // through; the output averages the 16 ramps, so y(x) = x.
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
        float v = 0;
        if (i < OFF_B1) v = static_cast<float>(kH);
        else if (i < OFF_W2) v = -static_cast<float>(i - OFF_B1);
        else if (i < OFF_B2) v = ((i - OFF_W2) / kH == (i - OFF_W2) % kH) ? 1.0f : 0.0f;
        else if (i < OFF_W3) v = 0.0f;
        else if (i < OFF_B3) v = 1.0f / kH;
        else v = 0.0f;
        P[i] = v;
    }
}

constexpr uint32_t kTrainBatch = 2048;

// One block per partition with enough knots: Adam (cosine-decayed step size)
// on the batch MSE of predicting the knot CDF (j + 0.5) / K from the
// normalised knot time; projection keeps the router monotone (w1, W2, w3 >= 0).
__global__ void __launch_bounds__(256) mlp_train_kernel(const PartMeta* meta, const Knot* knots, uint64_t nparts,
                                                        uint32_t steps, float lr, float* params, float* losses) {
    __shared__ float P[kParams], G[kParams], M1[kParams], M2[kParams];
    __shared__ float xs[kTrainBatch], cs[kTrainBatch];
    __shared__ float loss;
    const uint64_t p = blockIdx.x;
    if (p >= nparts) return;
    const PartMeta pm = meta[p];
    if (!pm.mlp_ok) return;  // block-uniform exit
    init_params(P);
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) M1[i] = M2[i] = 0.0f;
    const uint32_t K = pm.K;
    const uint32_t N = K < kTrainBatch ? K : kTrainBatch;
    float last = 0;
    for (uint32_t step = 0; step < steps; ++step) {
        // Batch: N knots spread over the partition, phase shifted per step.
        const double phase = fmod(static_cast<double>(step) * 0.6180339887498949, 1.0);
        for (uint32_t s = threadIdx.x; s < N; s += blockDim.x) {
            uint32_t j = static_cast<uint32_t>((static_cast<double>(s) + phase) * K / N);
            if (j >= K) j = K - 1;
            xs[s] = norm_x(pm, knots[pm.kb + j].tau);
            cs[s] = (static_cast<float>(j) + 0.5f) / static_cast<float>(K);
        }
        for (int i = threadIdx.x; i < kParams; i += blockDim.x) G[i] = 0.0f;
        if (threadIdx.x == 0) loss = 0;
        __syncthreads();
        mlp_batch_grads(P, xs, cs, N, G, &loss);
        __syncthreads();
        const float a = 0.5f * lr * (1.0f + cosf(3.14159265f * static_cast<float>(step) / static_cast<float>(steps)));
        const float b1c = 1.0f - powf(0.9f, static_cast<float>(step + 1));
        const float b2c = 1.0f - powf(0.999f, static_cast<float>(step + 1));
        for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
            const float gi = G[i];
            M1[i] = 0.9f * M1[i] + 0.1f * gi;
            M2[i] = 0.999f * M2[i] + 0.001f * gi * gi;
            float v = P[i] - a * (M1[i] / b1c) / (sqrtf(M2[i] / b2c) + 1e-8f);
            const bool nonneg = (i < OFF_B1) || (i >= OFF_W2 && i < OFF_B2) || (i >= OFF_W3 && i < OFF_B3);
            if (nonneg && v < 0.0f) v = 0.0f;
            P[i] = v;
        }
        last = loss / static_cast<float>(N);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) params[p * kParams + i] = P[i];
    if (threadIdx.x == 0) losses[p] = last;
}

// Gradient check entry: one block computes the batch gradient for given params.
__global__ void __launch_bounds__(256) mlp_grad_kernel(const float* params, const float* x, const float* c,
                                                       uint32_t N, float* grads, float* loss_out) {
    __shared__ float P[kParams], G[kParams];
    __shared__ float loss;
    extern __shared__ float dyn[];  // xs[N], cs[N]
    float* xs = dyn;
    float* cs = dyn + N;
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
        P[i] = params[i];
        G[i] = 0.0f;
    }
    for (uint32_t s = threadIdx.x; s < N; s += blockDim.x) {
        xs[s] = x[s];
        cs[s] = c[s];
    }
    if (threadIdx.x == 0) loss = 0;
    __syncthreads();
    mlp_batch_grads(P, xs, cs, N, G, &loss);
    __syncthreads();
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) grads[i] = G[i];
    if (threadIdx.x == 0) *loss_out = loss / static_cast<float>(N);
}

// Batched inference of 16 samples on the matrix cores: lane 16q + r supplies
// sample r's input x_r; on return every lane of quarter q holds y of samples
// 4q .. 4q+3 (H1 [16x16] times W2^T as four 16x16x4 MFMAs, as in training).
__device__ inline void mlp_infer_tile(const float* P, float x_r, float y[4]) {
    const unsigned lane = lane_id();
    const unsigned q = lane / 16, r = lane % 16;
    floatx4 z2;
    for (int v = 0; v < 4; ++v) z2[v] = P[OFF_B2 + r];
    for (int c = 0; c < 4; ++c) {
        const int k = 4 * c + q;
        const float a = ramp(fmaf(P[OFF_W1 + k], x_r, P[OFF_B1 + k]));  // H1[s=r][k]
        z2 = mfma16(a, P[OFF_W2 + r * kH + k], z2);                    // W2[m=r][k]
    }
    for (int v = 0; v < 4; ++v) {
        float pv = P[OFF_W3 + r] * ramp(z2[v]);
        for (int off = 8; off > 0; off >>= 1) pv += __shfl_xor(pv, off, 16);
        y[v] = pv + P[OFF_B3];
    }
}

// Router warp grid of each trained partition: the MLP at x = i / kGrid
// (i = 0..kGrid) in batch, made non-decreasing and scaled to [0, 1]. A
// partition whose warp is flat or not finite keeps the linear router.
__global__ void __launch_bounds__(64) mlp_grid_kernel(PartMeta* meta, uint64_t nparts, const float* params,
                                                      float* grid) {
    __shared__ float P[kParams];
    __shared__ float W[80];
    const uint64_t p = blockIdx.x;
    if (p >= nparts || !meta[p].mlp_ok) return;  // block-uniform exit
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) P[i] = params[p * kParams + i];
    __syncthreads();
    const unsigned lane = lane_id(), q = lane / 16, r = lane % 16;
    for (int tile = 0; tile < 5; ++tile) {  // 80 >= kGrid + 1 samples
        const int s = 16 * tile + static_cast<int>(r);
        const float x = static_cast<float>(s < kGrid ? s : kGrid) / static_cast<float>(kGrid);
        float y[4];
        mlp_infer_tile(P, x, y);
        if (r == 0)
            for (int v = 0; v < 4; ++v) W[16 * tile + 4 * q + v] = y[v];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool ok = true;
        float run = W[0];
        for (int i = 0; i <= kGrid; ++i) {
            if (!(W[i] == W[i]) || W[i] > 3.0e38f || W[i] < -3.0e38f) ok = false;
            run = W[i] > run ? W[i] : run;
            W[i] = run;
        }
        const float lo = W[0], span = W[kGrid] - W[0];
        ok = ok && span > 0.0f;
        if (ok) {
            for (int i = 0; i <= kGrid; ++i) grid[p * (kGrid + 1) + i] = (W[i] - lo) / span;
            grid[p * (kGrid + 1) + kGrid] = 1.0f;
        } else {
            meta[p].mlp_ok = 0;
        }
    }
}

// Test entry: batched inference over n inputs (one wave per 16 samples).
__global__ void __launch_bounds__(256) mlp_forward_kernel(const float* params, const float* x, uint64_t n, float* y) {
    __shared__ float P[kParams];
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const unsigned lane = lane_id(), q = lane / 16, r = lane % 16;
    const uint64_t wave = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / kWave;
    const uint64_t nwaves = uint64_t(gridDim.x) * blockDim.x / kWave;
    for (uint64_t s0 = wave * 16; s0 < n; s0 += nwaves * 16) {  // wave-uniform loop
        const uint64_t s = s0 + r;
        float out[4];
        mlp_infer_tile(P, s < n ? x[s] : 0.0f, out);
        if (r == 0)
            for (int v = 0; v < 4; ++v)
                if (s0 + 4 * q + v < n) y[s0 + 4 * q + v] = out[v];
    }
}

static PartKeys part_keys(const qe_dev_asof* idx, const qe_dev_col* const* keys) {
    PartKeys pk{};
    pk.mode = idx->mode;
    pk.nkeys = idx->nkeys;
    for (int j = 0; j < idx->nkeys; ++j) pk.key[j] = view(keys[j]);
    pk.kmin = idx->kmin;
    pk.range = idx->range;
    pk.t = idx->ht.v;
    pk.part_of_slot = idx->part_of_slot.as<uint32_t>();
    return pk;
}

static bool time_type_ok(int t) { return t == QE_I32 || t == QE_I64 || t == QE_TS; }

}  // namespace qe_gpu

extern "C" {

}  // extern "C"

namespace qe_gpu {

static qe_gpu_status asof_build_codes(qe_gpu_ctx* ctx, const qe_dev_col* const* build_keys, uint32_t nkeys,
                                      const qe_dev_col* build_time, qe_dev_asof** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && build_time != nullptr, "null argument");
        // tsx/asof.cpp:144-152: time columns must be I32/I64/TS.
        invalid(time_type_ok(build_time->type), "AsofJoin: time columns must be I32/I64/TS");
        check(nkeys <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many as-of keys for the device path");
        const uint64_t n = build_time->len;
        for (uint32_t j = 0; j < nkeys; ++j) {
            invalid(build_keys[j] != nullptr, "null key column");
            invalid(build_keys[j]->len == n, "build key and time columns differ in length");
        }
        check_rows32(n);
        auto* idx = new qe_dev_asof();
        try {
            idx->nrows_in = n;
            idx->nkeys = static_cast<int>(nkeys);
            for (uint32_t j = 0; j < nkeys; ++j) idx->key_types[j] = build_keys[j]->type;
            if (nkeys == 0) {
                idx->mode = 0;
                idx->nparts = 1;
            } else if (nkeys == 1 && is_int_like(build_keys[0]->type)) {
                const MinMax mm = column_minmax(ctx, build_keys[0], n);
                const uint64_t R = mm.valid ? static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1 : 1;
                if (R <= std::max<uint64_t>(2 * n, 1 << 16) && R < (uint64_t(1) << 30)) {
                    idx->mode = 1;
                    idx->kmin = mm.valid ? mm.mn : 0;
                    idx->range = mm.valid ? R : 0;
                    idx->nparts = mm.valid ? R : 1;
                }
            }
            if (nkeys > 0 && idx->mode == 0) {
                idx->mode = 2;
                HashKeys hk{};
                hk.nkeys = static_cast<int>(nkeys);
                for (uint32_t j = 0; j < nkeys; ++j) hk.key[j] = view(build_keys[j]);
                idx->ht.create(ctx, static_cast<int>(nkeys), std::max<uint64_t>(n, 1));
                if (n)
                    asof_slot_kernel<<<grid_for(n, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(hk, n, idx->ht.v);
                QE_HIP(hipGetLastError());
                idx->ht.check_overflow(ctx);
                const uint64_t cap = idx->ht.v.cap;
                Scratch flag(ctx, cap * 4), rank(ctx, cap * 8);
                slot_flag_kernel<<<grid_for(cap, 256, 8192), 256, 0, ctx->stream>>>(idx->ht.v.state, cap,
                                                                                     flag.as<uint32_t>());
                const uint64_t parts = exclusive_scan_u32(ctx, flag.as<uint32_t>(), rank.as<uint64_t>(), cap);
                idx->part_of_slot = Scratch(ctx, cap * 4);
                slot_part_kernel<<<grid_for(cap, 256, 8192), 256, 0, ctx->stream>>>(
                    flag.as<uint32_t>(), rank.as<uint64_t>(), cap, idx->part_of_slot.as<uint32_t>());
                QE_HIP(hipGetLastError());
                idx->nparts = std::max<uint64_t>(parts, 1);
            }
            const PartKeys pk = part_keys(idx, build_keys);
            const ColView tv = view(build_time);
            Scratch flag(ctx, std::max<uint64_t>(n, 1) * 4), pos(ctx, std::max<uint64_t>(n, 1) * 8);
            if (n)
                asof_flag_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(pk, tv, n,
                                                                                                flag.as<uint32_t>());
            QE_HIP(hipGetLastError());
            const uint64_t m = exclusive_scan_u32(ctx, flag.as<uint32_t>(), pos.as<uint64_t>(), n);
            idx->n = m;
            idx->offs = Scratch(ctx, (idx->nparts + 1) * 8);
            idx->ts = Scratch(ctx, std::max<uint64_t>(m, 1) * 8);
            idx->row = Scratch(ctx, std::max<uint64_t>(m, 1) * 4);
            if (m == 0) {
                fill_offs_kernel<<<grid_for(idx->nparts + 1, 256, 4096), 256, 0, ctx->stream>>>(
                    idx->offs.as<uint64_t>(), idx->nparts + 1, 0);
                QE_HIP(hipGetLastError());
            } else {
                const MinMax tm = column_minmax(ctx, build_time, n);
                const uint64_t tspan = static_cast<uint64_t>(tm.mx) - static_cast<uint64_t>(tm.mn);
                const unsigned tbits = std::max(1u, bits_for(tspan));
                const unsigned pbits = bits_for(idx->nparts - 1);
                Scratch keys(ctx, m * 8);
                TimedRegion tr(ctx, "asof_build_sort");
                if (tbits + pbits <= 64) {
                    asof_pack_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                        pk, tv, n, flag.as<uint32_t>(), pos.as<uint64_t>(), tm.mn, tbits, keys.as<uint64_t>(),
                        idx->row.as<uint32_t>());
                    QE_HIP(hipGetLastError());
                    radix_sort_pairs(ctx, keys.as<uint64_t>(), idx->row.as<uint32_t>(), m, 0, tbits + pbits);
                    asof_unpack_kernel<<<grid_for(m, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                        keys.as<uint64_t>(), m, tm.mn, tbits,
                        idx->nparts, idx->ts.as<int64_t>(), idx->offs.as<uint64_t>());
                    QE_HIP(hipGetLastError());
                } else {
                    throw Error{QE_GPU_UNSUPPORTED, "as-of key: partition bits + time span exceed 64 bits"};
                }
            }
            QE_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete idx;
            throw;
        }
        *out = idx;
    });
}

}  // namespace qe_gpu

extern "C" {

// STR keys partition by value: a dictionary with repeated values is read as
// canonical codes (strings.hip); the index keeps the build dictionary so the
// probe side maps its codes onto the same values.
qe_gpu_status qe_gpu_asof_build(qe_gpu_ctx* ctx, const qe_dev_col* const* build_keys, uint32_t nkeys,
                                const qe_dev_col* build_time, qe_dev_asof** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && (build_keys != nullptr || nkeys == 0), "null argument");
        check(nkeys <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many as-of keys for the device path");
        OwnedCols own(ctx);
        std::vector<const qe_dev_col*> bk(build_keys, build_keys + nkeys);
        for (uint32_t j = 0; j < nkeys; ++j) {
            if (bk[j] == nullptr || bk[j]->type != QE_STR) continue;
            const qe_gpu_dict* d = need_dict(bk[j]);
            if (!d->unique) bk[j] = own.keep(remap_col(ctx, bk[j], d->canon, QE_STR, d));
        }
        const qe_gpu_status st = asof_build_codes(ctx, bk.data(), nkeys, build_time, out);
        if (st != QE_GPU_OK) throw Error{st, ctx->last_error};
        for (uint32_t j = 0; j < nkeys; ++j) (*out)->key_dicts[j] = bk[j]->type == QE_STR ? bk[j]->dict : nullptr;
    });
}

}  // extern "C"

namespace qe_gpu {

// The per-partition route records the learned probe and the clustering read.
static void build_routes(qe_gpu_ctx* ctx, qe_dev_asof* idx, uint64_t P) {
    idx->route = Scratch(ctx, std::max<uint64_t>(P, 1) * sizeof(RouteRec));
    if (P == 0) return;
    route_recs_kernel<<<grid_for(P, 256, 4096), 256, 0, ctx->stream>>>(idx->meta.as<PartMeta>(), P,
                                                                      idx->route.as<RouteRec>());
    QE_HIP(hipGetLastError());
}

// Times the candidate learned router (the MLP where it searches fewer knots)
// against the linear router on a sample of the index's rows, best of three
// launches each, and keeps the MLP only if it is at least 2% faster; else the
// learned probe uses the linear router in every partition. Both timings stay
// in the index's statistics.
static void time_routers(qe_gpu_ctx* ctx, qe_dev_asof* idx, uint64_t P, uint64_t nk, uint64_t nt) {
    const uint64_t ns = std::min<uint64_t>(idx->n, uint64_t(1) << 21);
    if (ns == 0) return;
    Scratch ids(ctx, ns * 4), tv(ctx, ns * 8), out(ctx, ns * 4), fb(ctx, 8);
    router_sample_kernel<<<grid_for(ns, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
        idx->offs.as<uint64_t>(), P, idx->ts.as<int64_t>(), idx->n, ns, ids.as<uint32_t>(), tv.as<int64_t>());
    QE_HIP(hipGetLastError());
    memset_async(ctx, fb.p, 0, 8);
    ProbeArgs a{};
    a.time = ColView{tv.p, nullptr, QE_I64, 0};
    a.n = ns;
    a.ts = idx->ts.as<int64_t>();
    a.row = idx->row.as<uint32_t>();
    a.offs = idx->offs.as<uint64_t>();
    a.out = out.as<uint32_t>();
    const GivenPart gp{ids.as<uint32_t>()};
    const LearnedArgs Lm{idx->route.as<RouteRec>(), idx->meta.as<PartMeta>(), idx->knots.as<Knot>(),
                         idx->tab_mlp.as<uint32_t>(),
                         idx->bkt_mlp.as<BucketLine>(), idx->grid.as<float>(), 1, idx->B,
                         fb.as<unsigned long long>()};
    LearnedArgs Ll = Lm;
    Ll.tab = idx->tab_lin.as<uint32_t>();
    Ll.bkt = idx->bkt_lin.as<BucketLine>();
    Ll.use_mlp = 0;
    const unsigned g = grid_for(ns, 256, ctx->num_cus * 32);
    hipEvent_t e0, e1;
    QE_HIP(hipEventCreate(&e0));
    QE_HIP(hipEventCreate(&e1));
    float best[2] = {1e30f, 1e30f};
    for (int rep = 0; rep < 4; ++rep) {
        for (int w = 0; w < 2; ++w) {
            QE_HIP(hipEventRecord(e0, ctx->stream));
            asof_learned_kernel<<<g, 256, 0, ctx->stream>>>(gp, a, w == 0 ? Lm : Ll);
            QE_HIP(hipGetLastError());
            QE_HIP(hipEventRecord(e1, ctx->stream));
            QE_HIP(hipEventSynchronize(e1));
            float ms = 0;
            QE_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) best[w] = std::min(best[w], ms);  // the first round warms the caches
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    idx->sample_ms_mlp = best[0];
    idx->sample_ms_lin = best[1];
    if (best[0] > 0.98f * best[1]) {
        clear_mlp_kernel<<<grid_for(P, 256, 4096), 256, 0, ctx->stream>>>(idx->meta.as<PartMeta>(), P);
        QE_HIP(hipGetLastError());
        QE_HIP(hipMemcpyAsync(idx->tab_mlp.p, idx->tab_lin.p, std::max<uint64_t>(nt, 1) * 4, hipMemcpyDeviceToDevice,
                              ctx->stream));
        QE_HIP(hipMemcpyAsync(idx->bkt_mlp.p, idx->bkt_lin.p, std::max<uint64_t>(nt, 1) * sizeof(BucketLine),
                              hipMemcpyDeviceToDevice, ctx->stream));
        idx->mlp_rejected = true;
    }
    (void)nk;
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_asof_train(qe_gpu_ctx* ctx, qe_dev_asof* idx, uint32_t piece_rows, uint32_t steps, float lr) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr, "null index");
        invalid(piece_rows >= 1 && piece_rows <= 4096, "piece_rows must be in [1, 4096]");
        const uint64_t P = idx->nparts;
        idx->B = piece_rows;
        idx->mlp_rejected = false;
        idx->sample_ms_mlp = idx->sample_ms_lin = 0;
        Scratch K(ctx, P * 4), R1(ctx, P * 4), kb(ctx, (P + 1) * 8), tb(ctx, (P + 1) * 8);
        const uint32_t kpb = static_cast<uint32_t>(ctx->tune_asof_bucket_knots);
        idx->kpb = kpb;
        meta_counts_kernel<<<grid_for(P, 256, 4096), 256, 0, ctx->stream>>>(idx->offs.as<uint64_t>(), P, piece_rows,
                                                                            kpb, K.as<uint32_t>(), R1.as<uint32_t>());
        QE_HIP(hipGetLastError());
        const uint64_t nk = exclusive_scan_u32(ctx, K.as<uint32_t>(), kb.as<uint64_t>(), P);
        const uint64_t nt = exclusive_scan_u32(ctx, R1.as<uint32_t>(), tb.as<uint64_t>(), P);
        QE_HIP(hipMemcpyAsync(kb.as<uint64_t>() + P, &nk, 8, hipMemcpyHostToDevice, ctx->stream));
        QE_HIP(hipMemcpyAsync(tb.as<uint64_t>() + P, &nt, 8, hipMemcpyHostToDevice, ctx->stream));
        idx->nknots = nk;
        idx->ntab = nt;
        idx->meta = Scratch(ctx, P * sizeof(PartMeta));
        meta_fill_kernel<<<grid_for(P, 256, 4096), 256, 0, ctx->stream>>>(
            idx->offs.as<uint64_t>(), P, piece_rows, kpb, K.as<uint32_t>(), kb.as<uint64_t>(), tb.as<uint64_t>(),
            idx->ts.as<int64_t>(), idx->meta.as<PartMeta>());
        QE_HIP(hipGetLastError());
        const uint64_t nk1 = std::max<uint64_t>(nk, 1);
        idx->knots = Scratch(ctx, nk1 * sizeof(Knot));
        idx->tab_mlp = Scratch(ctx, std::max<uint64_t>(nt, 1) * 4);
        idx->tab_lin = Scratch(ctx, std::max<uint64_t>(nt, 1) * 4);
        idx->bkt_mlp = Scratch(ctx, std::max<uint64_t>(nt, 1) * sizeof(BucketLine));
        idx->bkt_lin = Scratch(ctx, std::max<uint64_t>(nt, 1) * sizeof(BucketLine));
        idx->params = Scratch(ctx, std::max<uint64_t>(P, 1) * kParams * 4);
        idx->grid = Scratch(ctx, std::max<uint64_t>(P, 1) * (kGrid + 1) * 4);
        Scratch losses(ctx, std::max<uint64_t>(P, 1) * 4), maxe(ctx, 8), sums(ctx, 24),
            occ(ctx, std::max<uint64_t>(P, 1) * 16), chosen(ctx, 8), ndl(ctx, 16);
        memset_async(ctx, ndl.p, 0, 16);
        memset_async(ctx, losses.p, 0, std::max<uint64_t>(P, 1) * 4);
        memset_async(ctx, maxe.p, 0, 8);
        memset_async(ctx, sums.p, 0, 24);
        memset_async(ctx, occ.p, 0, std::max<uint64_t>(P, 1) * 16);
        memset_async(ctx, chosen.p, 0, 8);
        if (nk) {
            const unsigned gk = grid_for(nk, 256, ctx->num_cus * 16);
            PartMeta* meta = idx->meta.as<PartMeta>();
            const Knot* knots = idx->knots.as<Knot>();
            double* occ_lin = occ.as<double>();
            double* occ_mlp = occ_lin + P;
            pieces_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, piece_rows,
                                                       idx->ts.as<int64_t>(), idx->knots.as<Knot>(),
                                                       maxe.as<unsigned long long>());
            QE_HIP(hipGetLastError());
            {
                TimedRegion tr(ctx, "asof_mlp_train");
                mlp_train_kernel<<<P, 256, 0, ctx->stream>>>(meta, knots, P, steps, lr, idx->params.as<float>(),
                                                             losses.as<float>());
                QE_HIP(hipGetLastError());
            }
            mlp_grid_kernel<<<P, 64, 0, ctx->stream>>>(meta, P, idx->params.as<float>(), idx->grid.as<float>());
            QE_HIP(hipGetLastError());
            // Candidate MLP router and the linear one, with knots searched per partition.
            const float* grid = idx->grid.as<float>();
            router_table_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 2,
                                                             idx->tab_mlp.as<uint32_t>());
            router_stats_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 2,
                                                             idx->tab_mlp.as<uint32_t>(), occ_mlp,
                                                             sums.as<double>() + 2);
            router_table_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 0,
                                                             idx->tab_lin.as<uint32_t>());
            router_stats_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 0,
                                                             idx->tab_lin.as<uint32_t>(), occ_lin, sums.as<double>());
            router_choose_kernel<<<grid_for(P, 256, 4096), 256, 0, ctx->stream>>>(meta, P, occ_lin, occ_mlp,
                                                                                   chosen.as<unsigned long long>());
            // Candidate learned router: the MLP only where it searches fewer knots.
            router_table_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 1,
                                                             idx->tab_mlp.as<uint32_t>());
            router_stats_kernel<<<gk, 256, 0, ctx->stream>>>(meta, kb.as<uint64_t>(), P, nk, knots, grid, 1,
                                                             idx->tab_mlp.as<uint32_t>(), nullptr,
                                                             sums.as<double>() + 1);
            QE_HIP(hipGetLastError());
            const unsigned gt = grid_for(nt, 256, ctx->num_cus * 16);
            // Data lines need the kDataLine and kData16 bits clear in every knot line's last word.
            const int dl = ctx->tune_asof_data_lines && nk < (uint64_t(1) << 29) ? 1 : 0;
            bucket_lines_kernel<<<gt, 256, 0, ctx->stream>>>(
                meta, tb.as<uint64_t>(), P, nt, knots, idx->tab_mlp.as<uint32_t>(), idx->ts.as<int64_t>(), grid, 1,
                piece_rows, dl, idx->bkt_mlp.as<BucketLine>(), ndl.as<unsigned long long>());
            bucket_lines_kernel<<<gt, 256, 0, ctx->stream>>>(
                meta, tb.as<uint64_t>(), P, nt, knots, idx->tab_lin.as<uint32_t>(), idx->ts.as<int64_t>(), grid, 0,
                piece_rows, dl, idx->bkt_lin.as<BucketLine>(), ndl.as<unsigned long long>() + 1);
            QE_HIP(hipGetLastError());
            build_routes(ctx, idx, P);
            if (read_scalar(ctx, chosen.as<uint64_t>()) > 0) {
                time_routers(ctx, idx, P, nk, nt);
                build_routes(ctx, idx, P);  // the timing may have dropped the MLP routers
            }
        } else {
            build_routes(ctx, idx, P);
        }
        std::vector<float> hl(std::max<uint64_t>(P, 1));
        std::vector<PartMeta> hm(std::max<uint64_t>(P, 1));
        double hs[3] = {0, 0, 0};
        if (P) {
            QE_HIP(hipMemcpyAsync(hl.data(), losses.p, P * 4, hipMemcpyDeviceToHost, ctx->stream));
            QE_HIP(hipMemcpyAsync(hm.data(), idx->meta.p, P * sizeof(PartMeta), hipMemcpyDeviceToHost, ctx->stream));
        }
        QE_HIP(hipMemcpyAsync(hs, sums.p, 24, hipMemcpyDeviceToHost, ctx->stream));
        uint64_t hd[2] = {0, 0};
        QE_HIP(hipMemcpyAsync(hd, ndl.p, 16, hipMemcpyDeviceToHost, ctx->stream));
        idx->max_err = read_scalar(ctx, maxe.as<uint64_t>());
        idx->data_lines = idx->mlp_rejected ? hd[1] : hd[0];
        idx->mlp_parts = read_scalar(ctx, chosen.as<uint64_t>());
        double ls = 0;
        uint64_t cand = 0;
        for (uint64_t p = 0; p < P; ++p)
            if (hm[p].mlp_ok) {
                ls += hl[p];
                ++cand;
            }
        idx->mlp_candidates = cand;
        idx->final_loss = cand ? ls / cand : 0.0;
        idx->mean_knots_lin = nk ? hs[0] / nk : 0;
        idx->mean_knots_mlp = nk ? hs[1] / nk : 0;
        idx->mean_knots_cand = nk ? hs[2] / nk : 0;
        if (idx->mlp_rejected) {  // the learned probe routes linearly everywhere
            idx->mlp_parts = 0;
            idx->mean_knots_mlp = idx->mean_knots_lin;
        }
        idx->fallback = Scratch(ctx, 8);
        idx->trained = true;
    });
}

}  // extern "C"

namespace qe_gpu {

static qe_gpu_status asof_probe_codes(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* probe_keys,
                                      uint32_t nkeys, const qe_dev_col* probe_time, int32_t has_tolerance,
                                      int64_t tolerance, int32_t mode, uint32_t* match_out) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr && probe_time != nullptr, "null argument");
        invalid(time_type_ok(probe_time->type), "AsofJoin: time columns must be I32/I64/TS");
        invalid(static_cast<int>(nkeys) == idx->nkeys, "as-of needs equal-length key lists");
        for (uint32_t j = 0; j < nkeys; ++j) {
            invalid(probe_keys[j] != nullptr && probe_keys[j]->len == probe_time->len, "bad probe key column");
            invalid(probe_keys[j]->type == idx->key_types[j], "AsofJoin: key column types must match positionally");
        }
        const int sorted_out = (mode & QE_ASOF_OUT_SORTED) ? 1 : 0;
        mode &= ~QE_ASOF_OUT_SORTED;
        invalid(mode >= QE_ASOF_BINARY && mode <= QE_ASOF_LEARNED_NO_MLP, "unknown as-of mode");
        check(mode == QE_ASOF_BINARY || idx->trained, QE_GPU_INVALID_ARGUMENT,
              "learned as-of probe needs qe_gpu_asof_train first");
        const uint64_t n = probe_time->len;
        if (n == 0) return QE_GPU_OK;
        invalid(match_out != nullptr, "null output");
        ProbeArgs a{};
        const PartKeys pkeys = part_keys(idx, probe_keys);
        a.time = view(probe_time);
        a.n = n;
        a.ts = idx->ts.as<int64_t>();
        a.row = idx->row.as<uint32_t>();
        a.offs = idx->offs.as<uint64_t>();
        a.has_tol = has_tolerance ? 1 : 0;
        a.sorted_out = sorted_out;
        a.tol = static_cast<uint64_t>(tolerance);
        a.out = match_out;
        const unsigned g = grid_for(n, 256, ctx->num_cus * 32);
        // The partition functor for this index's key shape.
        auto with_part = [&](auto&& launch) {
            if (idx->mode == 0) {
                launch(ZeroKeyPart{});
            } else if (idx->mode == 1) {
                launch(DirectKeyPart{view(probe_keys[0]), idx->kmin, idx->range});
            } else {
                launch(pkeys);
            }
        };
        if (mode == QE_ASOF_BINARY) {
            TimedRegion tr(ctx, "asof_probe_binary");
            with_part([&](auto part) { asof_binary_kernel<<<g, 256, 0, ctx->stream>>>(part, a); });
            QE_HIP(hipGetLastError());
        } else {
            auto* mi = const_cast<qe_dev_asof*>(idx);
            memset_async(ctx, mi->fallback.p, 0, 8);
            const bool lm = mode == QE_ASOF_LEARNED;
            LearnedArgs L{idx->route.as<RouteRec>(), idx->meta.as<PartMeta>(), idx->knots.as<Knot>(),
                          lm ? idx->tab_mlp.as<uint32_t>() : idx->tab_lin.as<uint32_t>(),
                          lm ? idx->bkt_mlp.as<BucketLine>() : idx->bkt_lin.as<BucketLine>(), idx->grid.as<float>(),
                          lm ? 1 : 0, idx->B, mi->fallback.as<unsigned long long>()};
            {
                TimedRegion tr(ctx, mode == QE_ASOF_LEARNED ? "asof_probe_learned" : "asof_probe_learned_no_mlp");
                with_part([&](auto part) { asof_learned_kernel<<<g, 256, 0, ctx->stream>>>(part, a, L); });
                QE_HIP(hipGetLastError());
            }
            mi->last_fallbacks = read_scalar(ctx, mi->fallback.as<uint64_t>());
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // namespace qe_gpu

namespace qe_gpu {

// Probe key columns in the index's key space: STR probe keys map onto the
// build dictionary's values (strings.hip), a value the build side lacks
// getting an id no partition has.
static std::vector<const qe_dev_col*> probe_key_cols(qe_gpu_ctx* ctx, const qe_dev_asof* idx,
                                                     const qe_dev_col* const* probe_keys, uint32_t nkeys,
                                                     OwnedCols& own) {
    std::vector<const qe_dev_col*> pk(probe_keys, probe_keys + nkeys);
    for (uint32_t j = 0; j < nkeys; ++j) {
        if (pk[j] == nullptr || pk[j]->type != QE_STR || idx->key_types[j] != QE_STR) continue;
        const qe_gpu_dict* pd = need_dict(pk[j]);
        const qe_gpu_dict* bd = idx->key_dicts[j];
        invalid(bd != nullptr, "as-of index has no dictionary for a STR key");
        if (pd == bd && bd->unique) continue;
        pk[j] = own.keep(remap_col(ctx, pk[j], pd == bd ? bd->canon : probe_ids(bd, pd), QE_STR, nullptr));
    }
    return pk;
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_asof_probe(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* probe_keys,
                                uint32_t nkeys, const qe_dev_col* probe_time, int32_t has_tolerance,
                                int64_t tolerance, int32_t mode, uint32_t* match_out) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr && (probe_keys != nullptr || nkeys == 0), "null argument");
        invalid(static_cast<int>(nkeys) == idx->nkeys, "as-of needs equal-length key lists");
        invalid(probe_time != nullptr, "null argument");
        check_rows32(probe_time->len);
        OwnedCols own(ctx);
        const std::vector<const qe_dev_col*> pk = probe_key_cols(ctx, idx, probe_keys, nkeys, own);
        const qe_gpu_status st = asof_probe_codes(ctx, idx, pk.data(), nkeys, probe_time, has_tolerance, tolerance,
                                                  mode, match_out);
        if (st != QE_GPU_OK) throw Error{st, ctx->last_error};
    });
}

qe_gpu_status qe_gpu_asof_pairs(qe_gpu_ctx* ctx, const uint32_t* match, uint64_t n, int32_t join_type,
                                qe_dev_pairs** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        invalid(join_type == QE_JOIN_INNER || join_type == QE_JOIN_LEFT, "unknown join type");
        auto* pr = new qe_dev_pairs();
        try {
            if (n) {
                invalid(match != nullptr, "null match array");
                pr->count = select_pairs(ctx, match, n, join_type == QE_JOIN_LEFT, &pr->probe, &pr->build);
                QE_HIP(hipStreamSynchronize(ctx->stream));
            }
        } catch (...) {
            qe_gpu_pairs_free(ctx, pr);
            throw;
        }
        *out = pr;
    });
}

qe_gpu_status qe_gpu_asof_cluster(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* probe_keys,
                                  uint32_t nkeys, const qe_dev_col* probe_time, const qe_dev_col* const* cols,
                                  uint32_t ncols, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr && probe_time != nullptr && out != nullptr && cols != nullptr && ncols >= 1 &&
                    (probe_keys != nullptr || nkeys == 0),
                "null argument");
        invalid(static_cast<int>(nkeys) == idx->nkeys, "as-of needs equal-length key lists");
        invalid(time_type_ok(probe_time->type), "AsofJoin: time columns must be I32/I64/TS");
        const uint64_t n = probe_time->len;
        for (uint32_t j = 0; j < nkeys; ++j) {
            invalid(probe_keys[j] != nullptr && probe_keys[j]->len == n, "bad probe key column");
            invalid(probe_keys[j]->type == idx->key_types[j], "AsofJoin: key column types must match positionally");
        }
        for (uint32_t c = 0; c < ncols; ++c)
            invalid(cols[c] != nullptr && cols[c]->len == n, "clustered columns need the probe time's length");
        check_rows32(n);
        OwnedCols own(ctx);
        const std::vector<const qe_dev_col*> pk = probe_key_cols(ctx, idx, probe_keys, nkeys, own);
        qe_dev_table* t = new_table(n);
        try {
            for (uint32_t c = 0; c < ncols; ++c) {
                const bool nullable = !cols[c]->all_valid && cols[c]->validity != nullptr && n > 0;
                t->cols.push_back(new_col(ctx, cols[c]->type, n, nullable));
                t->cols.back()->dict = cols[c]->dict;
                // All ones, the all-valid padding convention past n included; the
                // scatter clears the NULL rows' bits.
                if (nullable) memset_async(ctx, t->cols.back()->validity, 0xFF, words_for(n) * 8);
            }
            if (n) {
                const uint64_t ntiles = (n + kClTile - 1) / kClTile;
                Scratch grp(ctx, n), hist(ctx, ntiles * kClGroups * 4), offs(ctx, ntiles * kClGroups * 8);
                ClusterArgs a{};
                a.time = view(probe_time);
                a.n = n;
                if (idx->trained && idx->nknots) a.route = idx->route.as<RouteRec>();
                a.grid = idx->grid.as<float>();
                a.offs = idx->offs.as<uint64_t>();
                a.nsorted = idx->n;
                a.ntab = std::max<uint64_t>(idx->ntab, 1);
                a.grp = grp.as<uint8_t>();
                a.hist = hist.as<uint32_t>();
                a.ntiles = ntiles;
                if (idx->mode == 0) {
                    cluster_group_kernel<<<ntiles, kRankBlock, 0, ctx->stream>>>(ZeroKeyPart{}, a);
                } else if (idx->mode == 1) {
                    cluster_group_kernel<<<ntiles, kRankBlock, 0, ctx->stream>>>(
                        DirectKeyPart{view(pk[0]), idx->kmin, idx->range}, a);
                } else {
                    cluster_group_kernel<<<ntiles, kRankBlock, 0, ctx->stream>>>(part_keys(idx, pk.data()), a);
                }
                QE_HIP(hipGetLastError());
                exclusive_scan_u32(ctx, hist.as<uint32_t>(), offs.as<uint64_t>(), ntiles * kClGroups);
                for (uint32_t c0 = 0; c0 < ncols; c0 += kClMaxCols) {
                    ClusterCols C{};
                    C.n = static_cast<int32_t>(std::min<uint32_t>(kClMaxCols, ncols - c0));
                    for (int i = 0; i < C.n; ++i) {
                        const qe_dev_col* src = cols[c0 + i];
                        qe_dev_col* dst = t->cols[c0 + i];
                        C.c[i] = ClusterCol{src->data, dst->validity ? src->validity : nullptr, dst->data,
                                            dst->validity, static_cast<int32_t>(type_width(src->type))};
                    }
                    TimedRegion tr(ctx, "asof_cluster_scatter");
                    cluster_scatter_kernel<<<tiles_per_die(ntiles) * kTileDies, kRankBlock, 0, ctx->stream>>>(
                        grp.as<uint8_t>(), n, offs.as<uint64_t>(), ntiles, C);
                    QE_HIP(hipGetLastError());
                }
            }
        } catch (...) {
            qe_gpu_table_free(ctx, t);
            throw;
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *out = t;
    });
}

qe_gpu_status qe_gpu_asof_take_sorted(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* cols,
                                      uint32_t ncols, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr && out != nullptr && cols != nullptr && ncols >= 1, "null argument");
        const uint64_t n = common_len(cols, ncols);
        invalid(n == idx->nrows_in, "take_sorted needs columns of the build side's length");
        const uint64_t m = idx->n;
        qe_dev_table* t = new_table(m);
        try {
            std::vector<const uint32_t*> idxs(ncols, idx->row.as<uint32_t>());
            if (m) t->cols = gather_cols(ctx, cols, idxs.data(), ncols, m);
            else
                for (uint32_t c = 0; c < ncols; ++c) t->cols.push_back(new_col(ctx, cols[c]->type, 0, false));
        } catch (...) {
            qe_gpu_table_free(ctx, t);
            throw;
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *out = t;
    });
}

qe_gpu_status qe_gpu_asof_stats(qe_gpu_ctx* ctx, const qe_dev_asof* idx, qe_asof_stats* out) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr && out != nullptr, "null argument");
        out->build_rows = idx->n;
        out->partitions = idx->nparts;
        out->pieces = idx->nknots;
        out->max_piece_error = idx->max_err;
        out->router_buckets = idx->ntab >= idx->nparts ? idx->ntab - idx->nparts : 0;
        out->mean_bucket_knots = idx->mean_knots_mlp;
        out->mean_bucket_knots_linear = idx->mean_knots_lin;
        out->fallbacks = idx->last_fallbacks;
        out->final_loss = idx->final_loss;
        out->mlp_candidates = idx->mlp_candidates;
        out->mlp_partitions = idx->mlp_parts;
        out->mean_bucket_knots_mlp = idx->mean_knots_cand;
        out->sample_probe_ms_mlp = idx->sample_ms_mlp;
        out->sample_probe_ms_linear = idx->sample_ms_lin;
        out->data_line_buckets = idx->data_lines;
    });
}

qe_gpu_status qe_gpu_asof_sorted(qe_gpu_ctx* ctx, const qe_dev_asof* idx, int64_t* ts, uint32_t* rows,
                                 uint64_t* offsets) {
    QE_ENTRY(ctx, {
        invalid(idx != nullptr, "null index");
        if (ts && idx->n) QE_HIP(hipMemcpyAsync(ts, idx->ts.p, idx->n * 8, hipMemcpyDeviceToHost, ctx->stream));
        if (rows && idx->n)
            QE_HIP(hipMemcpyAsync(rows, idx->row.p, idx->n * 4, hipMemcpyDeviceToHost, ctx->stream));
        if (offsets)
            QE_HIP(hipMemcpyAsync(offsets, idx->offs.p, (idx->nparts + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void qe_gpu_asof_free(qe_gpu_ctx* ctx, qe_dev_asof* idx) {
    if (!ctx || !idx) return;
    delete idx;  // Scratch members return their blocks to the context cache
}

qe_gpu_status qe_gpu_mlp_forward(qe_gpu_ctx* ctx, const float* params, const float* x, uint64_t n, float* y) {
    QE_ENTRY(ctx, {
        invalid(params != nullptr && (n == 0 || (x != nullptr && y != nullptr)), "null argument");
        Scratch P(ctx, kParams * 4), X(ctx, std::max<uint64_t>(n, 1) * 4), Y(ctx, std::max<uint64_t>(n, 1) * 4);
        QE_HIP(hipMemcpyAsync(P.p, params, kParams * 4, hipMemcpyHostToDevice, ctx->stream));
        if (n) QE_HIP(hipMemcpyAsync(X.p, x, n * 4, hipMemcpyHostToDevice, ctx->stream));
        if (n)
            mlp_forward_kernel<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(P.as<float>(), X.as<float>(), n,
                                                                                 Y.as<float>());
        QE_HIP(hipGetLastError());
        if (n) QE_HIP(hipMemcpyAsync(y, Y.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_mlp_grad(qe_gpu_ctx* ctx, const float* params, const float* x, const float* c, uint32_t n,
                              float* grads, float* loss) {
    QE_ENTRY(ctx, {
        invalid(params && x && c && grads && loss, "null argument");
        invalid(n >= 1 && n <= 8192, "batch size must be in [1, 8192]");
        Scratch P(ctx, kParams * 4), X(ctx, n * 4), Cc(ctx, n * 4), G(ctx, kParams * 4), L(ctx, 4);
        QE_HIP(hipMemcpyAsync(P.p, params, kParams * 4, hipMemcpyHostToDevice, ctx->stream));
        QE_HIP(hipMemcpyAsync(X.p, x, n * 4, hipMemcpyHostToDevice, ctx->stream));
        QE_HIP(hipMemcpyAsync(Cc.p, c, n * 4, hipMemcpyHostToDevice, ctx->stream));
        mlp_grad_kernel<<<1, 256, 2 * n * 4, ctx->stream>>>(P.as<float>(), X.as<float>(), Cc.as<float>(), n,
                                                            G.as<float>(), L.as<float>());
        QE_HIP(hipGetLastError());
        QE_HIP(hipMemcpyAsync(grads, G.p, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipMemcpyAsync(loss, L.p, 4, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
