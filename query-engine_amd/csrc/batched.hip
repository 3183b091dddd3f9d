// Batched group-by: many independent GROUP BY queries, each over its own
// morsels of uneven size, in a handful of persistent launches or one hipGraph
// (the engine's morsel driver, exec/parallel.cpp:452-480, fans the morsels of
// one query over CPU threads; here every query's morsels share the GPU at once).
//
// Each query picks a strategy at prepare time from its key column:
//  * small : integer keys whose value range fits a dense LDS table of at most
//            kSmallLds bytes; every tile aggregates into that table (slot =
//            key - min, NULL key in the last slot) and folds it into the
//            query's state with one atomic per touched slot field;
//  * part  : integer keys with a larger dense range and aggregates over the
//            value column; rows are radix-partitioned by slot (2^pbits slots
//            per partition: a histogram pass, one scan over every such query,
//            a scatter pass), then each partition is aggregated in LDS by one
//            block (its slots' final state written with plain stores) or, when
//            the query's partitions are large, by `chunks` blocks that each take
//            an equal slice of it and fold their tables in with atomics; a
//            query of at most kPassMaxParts partitions skips the histogram
//            and scatter: its blocks each take a slice of the input rows and
//            one partition, and fold that partition's rows of the slice;
//  * hash  : anything else (sparse ranges, aggregates over the key); rows probe
//            the shared device hash table, with an optional LDS slot cache.
// Each phase pulls work from one device queue shared by all queries; every
// block leaves when the queue is empty. Modes: persistent (the phase
// launches), graph (the same pipeline captured once in a hipGraph and
// replayed), per-batch (the baseline: one launch per 2048-row batch; direct
// slots take global atomics, hash queries go through the table). The
// autotuner times tile size x blocks per CU x LDS cache on the actual batch.
#include "agg_common.hpp"
#include "hash_common.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

using namespace qe_gpu;

namespace qe_gpu {

bool is_int_like(int t);

constexpr int kBatchMaxAcc = 6;
constexpr uint32_t kEmptyTag = 0xFFFFFFFFu;
constexpr size_t kSmallLds = 32 * 1024;  // LDS of a small query's dense table
constexpr size_t kPartLds = 128 * 1024;  // LDS of one partition's table (1024-thread block)
constexpr uint32_t kMaxParts = 8192;     // partitions of one query
constexpr uint32_t kPartAggThreads = 1024;
// Expected entries per partition-aggregation block: 2^19, so a block's table
// set-up and flush stay small against its entries, lowered (to 2^16 at least)
// while the batch's items would not cover every CU twice. On 64 queries of
// 1.24e9 rows (656M entries) 2^19 and 2^20 ran the pipeline in 13.41 ms,
// 2^18 in 13.45, 2^17 in 13.57, 2^16 in 13.76 and 2^15 in 14.34 ms.
constexpr uint64_t kChunkEntriesMin = uint64_t(1) << 16;
constexpr uint64_t kChunkEntriesMax = uint64_t(1) << 19;
constexpr uint32_t kSmallTileFactor = 4;     // small-kind tiles are this many times longer
// Partitioned queries of at most kPassMaxParts partitions skip the histogram
// and scatter: each partition's blocks read the key and value columns (every
// row, once per partition) and fold the partition's rows. Per row that is
// kPassMaxParts x (key + value) bytes, against key (histogram) + key + value
// (scatter) + 10 written + 10 read back through the entries.
constexpr uint32_t kPassMaxParts = 2;
// A partitioned entry's key is 16 bits: the slot within its partition (LDS
// holds at least 12 bytes per slot, so a 128 KiB partition has at most 2^13
// slots) | value validity << 15.
constexpr uint32_t kEntryValid = 15;
constexpr uint32_t kMaxPartBits = 15;
constexpr uint64_t kPassChunkRows = uint64_t(1) << 19;  // input rows per aggregation block of a pass query

enum BatchKind : int32_t { BK_SMALL = 0, BK_PART = 1, BK_HASH = 2 };

struct QueryDev {
    HashKeys hk;      // hk.key[0] = key column (every kind)
    HashTableView t;  // hash kind
    ColView val;      // value column
    uint64_t* st;     // state, compact layout over S slots
    AggProg p;
    int32_t kind;
    uint32_t pbits;   // part: 2^pbits slots per partition
    int64_t kmin;     // small / part: slot = key - kmin, NULL key -> S - 1
    uint64_t S;       // slots (hash: table capacity)
    uint32_t nparts;  // part
    uint32_t ntiles;  // part: tiles of this query (histogram columns)
    uint64_t hbase;   // part: first histogram entry, laid out [part][tile]
    uint64_t ebase;   // part: first partitioned entry
    uint64_t rows;
    uint32_t chunks;  // part: blocks per partition (slices of equal length)
    int32_t on_value; // every aggregate reads the value column (or is COUNT(*))
    int32_t kw, vw;   // key / value element width in bytes (1, 4 or 8)
    int32_t pass;     // part: aggregated straight from the key and value columns, no scatter
};

struct Tile {
    uint32_t q;
    uint32_t qt;  // tile index within its query
    uint64_t begin, end;
};
struct PartItem {
    uint32_t q;
    uint32_t part;
    uint32_t chunk;
    uint32_t pad;
};

__device__ inline uint32_t direct_slot(const QueryDev& q, uint64_t r) {
    const ColView& k = q.hk.key[0];
    if (!is_valid(k, r)) return static_cast<uint32_t>(q.S - 1);
    return static_cast<uint32_t>(static_cast<uint64_t>(int_value(k, r) - q.kmin));
}

// ---- typed row access for the streaming phases ---------------------------------------------
// Each phase picks a specialisation per tile (a uniform branch) on the key and
// value widths (the typed column access of internal.hpp).
//
// Direct slots of N rows per lane, rows r0 + u * blockDim.x + lane (kNoSlot
// past the tile end), with the value column's bits and validity when VAL.
constexpr int kRowUnroll = 4;  // rows per thread per step, loads issued first

template <int KW, int VW, bool VAL, int N = kRowUnroll>
__device__ inline void load_rows(const QueryDev& q, uint64_t r0, uint64_t end, uint32_t* slot, uint64_t* bits,
                                 uint32_t* valid) {
    const void* key = q.hk.key[0].data;
    const BitsView kv = bits_view(q.hk.key[0].validity);
    const void* val = q.val.data;
    const BitsView vv = bits_view(q.val.validity);
    uint64_t k[N];
    uint32_t kok[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
        const uint64_t r = r0 + uint64_t(u) * blockDim.x + threadIdx.x;
        const uint64_t rc = r < end ? r : end - 1;
        k[u] = elem_bits<KW>(key, rc);
        kok[u] = bit_at(kv, rc);
        if constexpr (VAL) {
            bits[u] = elem_bits<VW>(val, rc);
            valid[u] = bit_at(vv, rc);
        }
    }
    const uint32_t null_slot = static_cast<uint32_t>(q.S - 1);
#pragma unroll
    for (int u = 0; u < N; ++u) {
        const uint64_t r = r0 + uint64_t(u) * blockDim.x + threadIdx.x;
        const uint32_t s = kok[u] ? static_cast<uint32_t>(k[u] - static_cast<uint64_t>(q.kmin)) : null_slot;
        slot[u] = r < end ? s : kNoSlot;
    }
}

// Calls f.template run<KW, VW>() for the query's widths.
template <typename F>
__device__ inline void with_widths(const QueryDev& q, const F& f) {
    if (q.kw == 4) {
        if (q.vw == 8) f.template run<4, 8>();
        else if (q.vw == 4) f.template run<4, 4>();
        else f.template run<4, 1>();
    } else {
        if (q.vw == 8) f.template run<8, 8>();
        else if (q.vw == 4) f.template run<8, 4>();
        else f.template run<8, 1>();
    }
}

// One value's fold into LDS slot partials (the lds_bytes layout acc
// u64[nacc][S] | cnt u32[ncnt][S] | rows u32[S] | fhi u32[nfs][S]) for a query
// whose aggregates all read the value column: the same updates as
// lds_update_vals (exact F64 sums included), with the shared accumulators and
// the one non-NULL counter addressed directly. gslot_lo: the state slot of
// LDS slot 0.
__device__ inline void fold_value(char* smem, uint32_t S, const AggProg& p, uint32_t slot, uint64_t v,
                                  uint32_t valid, uint64_t gslot_lo = 0) {
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * p.nacc);
    uint32_t* rows = cnt + size_t(S) * p.ncnt;
    atomicAdd(&rows[slot], 1u);
    if (!valid) return;
    if (p.ncnt) atomicAdd(&cnt[slot], 1u);
    for (int a = 0; a < p.nacc; ++a) lds_acc_value(smem, S, p, a, p.acc_kind[a], slot, gslot_lo + slot, v);
}

// LDS partial slot s (the lds_bytes layout of S slots) folded into state slot
// g of the query's HBM state (SS slots) with order-free atomics: integer sums,
// 128-bit F64 sums, MIN/MAX.
__device__ inline void combine_atomic(int kind, uint64_t* dst, uint64_t v);
__device__ inline void flush_slot_atomic(const char* smem, uint32_t S, const AggProg& p, uint32_t s, uint64_t* st,
                                         uint64_t SS, uint64_t g) {
    const uint64_t* acc = reinterpret_cast<const uint64_t*>(smem);
    const uint32_t* cnt = reinterpret_cast<const uint32_t*>(smem + size_t(8) * S * p.nacc);
    const uint32_t* rows = cnt + size_t(S) * p.ncnt;
    const uint32_t* fhi = rows + S;
    if (rows[s] == 0) return;
    atomicAdd(reinterpret_cast<unsigned long long*>(&st[g]), static_cast<unsigned long long>(rows[s]));
    for (int a = 0; a < p.nacc; ++a) {
        const uint64_t v = acc[size_t(a) * S + s];
        if (p.acc_kind[a] == K_SUM_F) {
            const int j = p.fhi_of[a];
            global_add_fixed(&st[(1 + size_t(a)) * SS + g], &st[state_fhi(SS, p, j) + g], v,
                             static_cast<int64_t>(static_cast<int32_t>(fhi[size_t(j) * S + s])));
        } else {
            combine_atomic(p.acc_kind[a], &st[(1 + size_t(a)) * SS + g], v);
        }
    }
    for (int c = 0; c < p.ncnt; ++c)
        if (cnt[size_t(c) * S + s])
            atomicAdd(reinterpret_cast<unsigned long long*>(&st[(1 + size_t(p.nacc) + c) * SS + g]),
                      static_cast<unsigned long long>(cnt[size_t(c) * S + s]));
}

// The same when the block is the slot's only writer: plain stores.
__device__ inline void flush_slot_store(const char* smem, uint32_t S, const AggProg& p, uint32_t s, uint64_t* st,
                                        uint64_t SS, uint64_t g) {
    const uint64_t* acc = reinterpret_cast<const uint64_t*>(smem);
    const uint32_t* cnt = reinterpret_cast<const uint32_t*>(smem + size_t(8) * S * p.nacc);
    const uint32_t* rows = cnt + size_t(S) * p.ncnt;
    const uint32_t* fhi = rows + S;
    st[g] = rows[s];
    for (int a = 0; a < p.nacc; ++a) st[(1 + size_t(a)) * SS + g] = acc[size_t(a) * S + s];
    for (int c = 0; c < p.ncnt; ++c) st[(1 + size_t(p.nacc) + c) * SS + g] = cnt[size_t(c) * S + s];
    for (int j = 0; j < p.nfs; ++j)
        st[state_fhi(SS, p, j) + g] = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(fhi[size_t(j) * S + s])));
}

// ---- LDS slot cache of the hash kind ------------------------------------------------------
// Layout for C entries: tag u32[C] | rows u32[C] | acc u64[kBatchMaxAcc][C] | cnt u32[kBatchMaxAcc][C]
// | fhi u32[kBatchMaxAcc][C] (high words of exact F64 sums).
__host__ __device__ inline size_t cache_bytes(uint32_t C) {
    return size_t(C) * (4 + 4 + 8 * kBatchMaxAcc + 4 * kBatchMaxAcc + 4 * kBatchMaxAcc);
}
__device__ inline uint32_t* cache_fhi(char* smem, uint32_t C) {
    return reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(12) * kBatchMaxAcc * C);
}

__device__ inline void cache_reset(char* smem, uint32_t C, const AggProg& p) {
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
        tag[i] = kEmptyTag;
        rows[i] = 0;
        for (int a = 0; a < p.nacc; ++a) acc[size_t(a) * C + i] = acc_identity(p.acc_kind[a]);
        for (int c = 0; c < p.ncnt; ++c) cnt[size_t(c) * C + i] = 0;
        for (int j = 0; j < p.nfs; ++j) cache_fhi(smem, C)[size_t(j) * C + i] = 0;
    }
}

// Fold one row into the LDS cache entry for global slot s if the entry is (or
// becomes) s's; returns false when the entry holds another slot.
__device__ inline bool cache_update(char* smem, uint32_t C, const AggProg& p, uint32_t s, uint64_t r) {
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    const uint32_t e = s & (C - 1);
    const uint32_t old = atomicCAS(&tag[e], kEmptyTag, s);
    if (old != kEmptyTag && old != s) return false;
    atomicAdd(&rows[e], 1u);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR || !is_valid(p.in[a], r)) continue;
        if (p.cnt_of[a] >= 0 && p.own_cnt[a]) atomicAdd(&cnt[size_t(p.cnt_of[a]) * C + e], 1u);
        if (p.acc_of[a] < 0 || !p.own_acc[a]) continue;
        uint64_t* ap = &acc[size_t(p.acc_of[a]) * C + e];
        switch (k) {
            case K_SUM_I:
                atomicAdd(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(int_value(p.in[a], r)));
                break;
            case K_SUM_F: {
                const int i = p.acc_of[a];
                uint64_t l;
                int64_t h;
                const uint64_t vbits = value_bits(p.in[a], r);
                const uint32_t code = f64_fixed(vbits, p.fsc[i], p.fsc[kMaxAggs + i], l, h);
                if (code == 0)
                    lds_add_fixed(ap, cache_fhi(smem, C) + size_t(p.fhi_of[i]) * C + e, l, h);
                else
                    f64_special(p, p.fhi_of[i], s, code, vbits);
                break;
            }
            case K_MIN_I:
                atomicMin(reinterpret_cast<long long*>(ap), static_cast<long long>(int_value(p.in[a], r)));
                break;
            case K_MAX_I:
                atomicMax(reinterpret_cast<long long*>(ap), static_cast<long long>(int_value(p.in[a], r)));
                break;
            case K_MIN_F:
                atomicMin(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(f64_order_key(f64_value(p.in[a], r))));
                break;
            case K_MAX_F:
                atomicMax(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(f64_order_key(f64_value(p.in[a], r))));
                break;
            default:
                break;
        }
    }
    return true;
}

__device__ inline void combine_atomic(int kind, uint64_t* dst, uint64_t v) {
    switch (kind) {
        case K_SUM_I:
            atomicAdd(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        case K_MIN_I:
            atomicMin(reinterpret_cast<long long*>(dst), static_cast<long long>(v));
            break;
        case K_MAX_I:
            atomicMax(reinterpret_cast<long long*>(dst), static_cast<long long>(v));
            break;
        case K_MIN_F:
            atomicMin(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        case K_MAX_F:
            atomicMax(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        default:
            break;
    }
}

__device__ inline void cache_flush(char* smem, uint32_t C, const QueryDev& q) {
    const AggProg& p = q.p;
    const uint64_t S = q.S;
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
        const uint32_t s = tag[i];
        if (s == kEmptyTag) continue;
        atomicAdd(reinterpret_cast<unsigned long long*>(&q.st[s]), static_cast<unsigned long long>(rows[i]));
        for (int a = 0; a < p.nacc; ++a) {
            if (p.acc_kind[a] == K_SUM_F) {
                const int j = p.fhi_of[a];
                global_add_fixed(&q.st[(1 + size_t(a)) * S + s], &q.st[state_fhi(S, p, j) + s], acc[size_t(a) * C + i],
                                 static_cast<int64_t>(static_cast<int32_t>(cache_fhi(smem, C)[size_t(j) * C + i])));
            } else {
                combine_atomic(p.acc_kind[a], &q.st[(1 + size_t(a)) * S + s], acc[size_t(a) * C + i]);
            }
        }
        for (int c = 0; c < p.ncnt; ++c)
            atomicAdd(reinterpret_cast<unsigned long long*>(&q.st[(1 + size_t(p.nacc) + c) * S + s]),
                      static_cast<unsigned long long>(cnt[size_t(c) * C + i]));
    }
}

__device__ inline void hash_tile(const QueryDev& q, const Tile& tile, char* smem, uint32_t C) {
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = tile.begin + threadIdx.x; r < tile.end; r += blockDim.x) {
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(q.hk, r, w, nm, h);
        const uint32_t s = ht_insert(q.t, q.hk.nkeys, w, nm, h);
        if (s == kNoSlot) continue;  // overflow reported through the table flag
        if (C && cache_update(smem, C, q.p, s, r)) continue;
        for (int a = 0; a < q.p.n; ++a) vrow[a] = r;
        global_update(q.st, q.S, q.p, s, vrow);
    }
}

// Row loops of the small kind (fold into the LDS table) and of the part
// kind's scatter (each row's entry, slot within its partition | value validity
// << 31 and the value bits, to its partition's range of this tile).
struct SmallRows {
    const QueryDev& q;
    const Tile& tile;
    char* smem;
    template <int KW, int VW>
    __device__ void run() const {
        const uint32_t S = static_cast<uint32_t>(q.S);
        for (uint64_t r0 = tile.begin; r0 < tile.end; r0 += kRowUnroll * uint64_t(blockDim.x)) {
            uint32_t slot[kRowUnroll], valid[kRowUnroll];
            uint64_t bits[kRowUnroll];
            load_rows<KW, VW, true>(q, r0, tile.end, slot, bits, valid);
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (slot[u] != kNoSlot) fold_value(smem, S, q.p, slot[u], bits[u], valid[u]);
        }
    }
};

struct ScatterRows {
    const QueryDev& q;
    const Tile& tile;
    const uint64_t* base;
    uint32_t* cur;
    uint16_t* ekey;
    uint64_t* eval;
    template <int KW, int VW>
    __device__ void run() const {
        const uint32_t pbits = q.pbits;
        const uint32_t mask = (1u << pbits) - 1;
        for (uint64_t r0 = tile.begin; r0 < tile.end; r0 += kRowUnroll * uint64_t(blockDim.x)) {
            uint32_t slot[kRowUnroll], valid[kRowUnroll];
            uint64_t bits[kRowUnroll];
            load_rows<KW, VW, true>(q, r0, tile.end, slot, bits, valid);
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u) {
                if (slot[u] == kNoSlot) continue;
                const uint32_t part = slot[u] >> pbits;
                const uint64_t pos = base[part] + atomicAdd(&cur[part], 1u);
                ekey[pos] = static_cast<uint16_t>((slot[u] & mask) | (valid[u] << kEntryValid));
                eval[pos] = bits[u];
            }
        }
    }
};

// Small kind: dense LDS table over the query's S slots, folded into HBM with
// one atomic per touched slot field.
__device__ inline void small_tile(const QueryDev& q, const Tile& tile, char* smem) {
    const uint32_t S = static_cast<uint32_t>(q.S);
    lds_init(smem, S, q.p);
    __syncthreads();
    if (q.on_value) {
        with_widths(q, SmallRows{q, tile, smem});
    } else {
        uint64_t vrow[kMaxAggs];
        for (uint64_t r = tile.begin + threadIdx.x; r < tile.end; r += blockDim.x) {
            for (int a = 0; a < q.p.n; ++a) vrow[a] = r;
            lds_update(smem, S, q.p, direct_slot(q, r), vrow);
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) flush_slot_atomic(smem, S, q.p, s, q.st, q.S, s);
}

// Staged scatter for queries with few partitions: each step of kStageRows rows
// is counting-sorted by partition in LDS (rank within the partition from an
// LDS atomic, partition offsets from a block scan), then written out in LDS
// order, so consecutive lanes store consecutive entries of one partition's
// range instead of one entry in each of up to 64 ranges.
// LDS: gbase u64[np] | sval u64[S] | skey u16[S] | cnt u32[np] | ws u32[kMaxWaves + 1] | spart u16[S].
constexpr int kStageUnroll = 4 * kRowUnroll;               // rows per thread per step
constexpr uint32_t kStageRows = kStageUnroll * 256;        // rows per step (256-thread blocks)
constexpr uint32_t kStageMaxParts = 1024;                  // partitions of a staged query (block scan: 4 per thread)

__host__ __device__ inline size_t stage_lds_bytes(uint32_t np) {
    return size_t(np) * 12 + size_t(kStageRows) * 12 + 4 * (kMaxWaves + 1);
}

struct StagedScatterRows {
    const QueryDev& q;
    const Tile& tile;
    char* smem;
    uint16_t* ekey;
    uint64_t* eval;
    template <int KW, int VW>
    __device__ void run() const {
        const uint32_t np = q.nparts, pbits = q.pbits, mask = (1u << pbits) - 1;
        uint64_t* gbase = reinterpret_cast<uint64_t*>(smem);
        uint64_t* sval = gbase + np;
        uint16_t* skey = reinterpret_cast<uint16_t*>(sval + kStageRows);
        uint32_t* cnt = reinterpret_cast<uint32_t*>(skey + kStageRows);
        uint32_t* ws = cnt + np;
        uint16_t* spart = reinterpret_cast<uint16_t*>(ws + kMaxWaves + 1);
        for (uint64_t r0 = tile.begin; r0 < tile.end; r0 += kStageRows) {
            uint32_t slot[kStageUnroll], valid[kStageUnroll], rank[kStageUnroll];
            uint64_t bits[kStageUnroll];
            load_rows<KW, VW, true, kStageUnroll>(q, r0, tile.end, slot, bits, valid);
#pragma unroll
            for (int u = 0; u < int(kStageUnroll); ++u)
                if (slot[u] != kNoSlot) rank[u] = atomicAdd(&cnt[slot[u] >> pbits], 1u);
            __syncthreads();
            const uint32_t total = block_exclusive_scan<kStageMaxParts / 256>(cnt, np, ws);
#pragma unroll
            for (int u = 0; u < int(kStageUnroll); ++u) {
                if (slot[u] == kNoSlot) continue;
                const uint32_t part = slot[u] >> pbits;
                const uint32_t at = cnt[part] + rank[u];
                skey[at] = static_cast<uint16_t>((slot[u] & mask) | (valid[u] << kEntryValid));
                sval[at] = bits[u];
                spart[at] = static_cast<uint16_t>(part);
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < total; i += 256) {
                const uint32_t part = spart[i];
                const uint64_t g = gbase[part] + (i - cnt[part]);
                ekey[g] = skey[i];
                eval[g] = sval[i];
            }
            uint32_t mine[kStageMaxParts / 256];
#pragma unroll
            for (int j = 0; j < int(kStageMaxParts / 256); ++j) {
                const uint32_t t = threadIdx.x + 256u * j;
                mine[j] = t < np ? (t + 1 < np ? cnt[t + 1] : total) - cnt[t] : 0u;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < int(kStageMaxParts / 256); ++j) {
                const uint32_t t = threadIdx.x + 256u * j;
                if (t < np) {
                    gbase[t] += mine[j];
                    cnt[t] = 0;
                }
            }
            __syncthreads();
        }
    }
};

// Part kind, scatter (staged or direct) after this tile's partition bases are read.
__device__ inline void scatter_tile(const QueryDev& q, const Tile& tile, char* smem, const uint64_t* offs,
                                    uint16_t* ekey, uint64_t* eval) {
    const uint32_t np = q.nparts;
    const bool staged = np <= kStageMaxParts;
    uint64_t* base = reinterpret_cast<uint64_t*>(smem);
    uint32_t* cur = staged ? reinterpret_cast<uint32_t*>(smem + size_t(np) * 8 + size_t(kStageRows) * 10)
                           : reinterpret_cast<uint32_t*>(base + np);
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        base[i] = offs[q.hbase + uint64_t(i) * q.ntiles + tile.qt];
        cur[i] = 0;
    }
    __syncthreads();
    if (staged) with_widths(q, StagedScatterRows{q, tile, smem, ekey, eval});
    else with_widths(q, ScatterRows{q, tile, base, cur, ekey, eval});
}

// ---- kernels --------------------------------------------------------------------------------
// State of every query the phases do not write in full: small and hash
// queries (and, for the per-batch mode, part queries too).
__global__ void batch_init_kernel(const QueryDev* qs, uint32_t nq, int all) {
    for (uint32_t qi = 0; qi < nq; ++qi) {
        const QueryDev& q = qs[qi];
        const uint64_t S = q.S;
        if (q.kind == BK_PART && q.chunks == 1 && !all) {  // fully written; only the flags and outlier sums start at 0
            for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
                 s += uint64_t(gridDim.x) * blockDim.x)
                for (int j = 0; j < 2 * q.p.nfs; ++j) q.st[state_flags(S, q.p, 0) + size_t(j) * S + s] = 0;
            continue;
        }
        for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
             s += uint64_t(gridDim.x) * blockDim.x) {
            for (int j = 0; j < 3 * q.p.nfs; ++j) q.st[state_fhi(S, q.p, 0) + size_t(j) * S + s] = 0;
            if (q.kind == BK_HASH) q.t.state[s] = 0;
            q.st[s] = 0;
            for (int a = 0; a < q.p.nacc; ++a) q.st[(1 + size_t(a)) * S + s] = acc_identity(q.p.acc_kind[a]);
            for (int c = 0; c < q.p.ncnt; ++c) q.st[(1 + size_t(q.p.nacc) + c) * S + s] = 0;
        }
    }
}

__global__ void batch_reset_kernel(unsigned long long* next, unsigned int* ovf) {
    for (int k = 0; k < 5; ++k) next[k] = 0;
    *ovf = 0;
}

// The tile index every lane of the block read from `cur`, as a scalar: the
// tile and its query's fields are then uniform (scalar loads, uniform branches).
__device__ inline uint64_t uniform_index(uint64_t i) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(i));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(i >> 32));
    return (uint64_t(hi) << 32) | lo;
}

template <int KW>
__device__ inline void hist_rows(const QueryDev& q, const Tile& t, uint32_t* hcnt) {
    const uint32_t pbits = q.pbits;
    for (uint64_t r0 = t.begin; r0 < t.end; r0 += kRowUnroll * uint64_t(blockDim.x)) {
        uint32_t slot[kRowUnroll], valid[kRowUnroll];
        uint64_t bits[kRowUnroll];
        load_rows<KW, 8, false>(q, r0, t.end, slot, bits, valid);
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (slot[u] != kNoSlot) atomicAdd(&hcnt[slot[u] >> pbits], 1u);
    }
}

// Part queries, pass 1: per-(partition, tile) row counts.
__global__ void __launch_bounds__(256) batch_hist_kernel(const QueryDev* __restrict__ qs,
                                                         const Tile* __restrict__ tiles, uint64_t ntiles,
                                                         unsigned long long* next, uint32_t* hist) {
    extern __shared__ uint32_t hcnt[];
    __shared__ uint64_t cur;
    while (true) {
        if (threadIdx.x == 0) cur = atomicAdd(next, 1ull);
        __syncthreads();
        const uint64_t i = uniform_index(cur);
        if (i >= ntiles) break;
        const Tile t = tiles[i];
        const QueryDev& q = qs[t.q];
        for (uint32_t p = threadIdx.x; p < q.nparts; p += blockDim.x) hcnt[p] = 0;
        __syncthreads();
        if (q.kw == 4) hist_rows<4>(q, t, hcnt);
        else hist_rows<8>(q, t, hcnt);
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < q.nparts; p += blockDim.x) hist[q.hbase + uint64_t(p) * q.ntiles + t.qt] = hcnt[p];
        __syncthreads();
    }
}

// One persistent kernel per strategy over that strategy's tile list, so each
// gets its own register budget: small tables, part scatter, hash probes.
template <int KIND>
__global__ void __launch_bounds__(256) batch_tiles_kernel(const QueryDev* __restrict__ qs,
                                                          const Tile* __restrict__ tiles, uint64_t ntiles,
                                                          unsigned long long* next, uint32_t C, const uint64_t* offs,
                                                          uint16_t* ekey, uint64_t* eval) {
    extern __shared__ __align__(16) char smem[];
    __shared__ uint64_t cur;
    while (true) {
        if (threadIdx.x == 0) cur = atomicAdd(next, 1ull);
        __syncthreads();
        const uint64_t i = uniform_index(cur);
        if (i >= ntiles) break;
        const Tile tile = tiles[i];
        const QueryDev& q = qs[tile.q];
        if (KIND == BK_SMALL) {
            small_tile(q, tile, smem);
        } else if (KIND == BK_PART) {
            scatter_tile(q, tile, smem, offs, ekey, eval);
        } else {
            if (C) {
                cache_reset(smem, C, q.p);
                __syncthreads();
            }
            hash_tile(q, tile, smem, C);
            if (C) {
                __syncthreads();
                cache_flush(smem, C, q);
            }
        }
        __syncthreads();
    }
}

// Pass queries: one block folds the rows [b, e) of the input whose slot falls
// in partition `part` (LDS slots [0, cnt) hold state slots lo + s).
struct PassRows {
    const QueryDev& q;
    uint32_t part;
    uint64_t lo;
    uint32_t cnt;
    uint64_t b, e;
    char* smem;
    template <int KW, int VW>
    __device__ void run() const {
        const uint32_t pbits = q.pbits, mask = (1u << pbits) - 1;
        for (uint64_t r0 = b; r0 < e; r0 += kRowUnroll * uint64_t(blockDim.x)) {
            uint32_t slot[kRowUnroll], valid[kRowUnroll];
            uint64_t bits[kRowUnroll];
            load_rows<KW, VW, true>(q, r0, e, slot, bits, valid);
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (slot[u] != kNoSlot && (slot[u] >> pbits) == part)
                    fold_value(smem, cnt, q.p, slot[u] & mask, bits[u], valid[u], lo);
        }
    }
};

// Part queries, pass 2: one block per (query, partition) aggregates the
// partition's entries in LDS and writes its slots' final state.
__global__ void __launch_bounds__(kPartAggThreads) batch_part_agg_kernel(const QueryDev* __restrict__ qs,
                                                                         const PartItem* __restrict__ items,
                                                                         uint64_t nitems, unsigned long long* next,
                                                                         const uint64_t* offs, const uint16_t* ekey,
                                                                         const uint64_t* eval) {
    extern __shared__ __align__(16) char smem[];
    __shared__ uint64_t cur;
    while (true) {
        if (threadIdx.x == 0) cur = atomicAdd(next, 1ull);
        __syncthreads();
        const uint64_t i = uniform_index(cur);
        if (i >= nitems) break;
        const PartItem it = items[i];
        const QueryDev& q = qs[it.q];
        const uint64_t lo = uint64_t(it.part) << q.pbits;
        const uint32_t cnt = static_cast<uint32_t>(min<uint64_t>(uint64_t(1) << q.pbits, q.S - lo));
        lds_init(smem, cnt, q.p);
        __syncthreads();
        if (q.pass) {
            const uint64_t b = q.rows * it.chunk / q.chunks, e = q.rows * (it.chunk + 1) / q.chunks;
            with_widths(q, PassRows{q, it.part, lo, cnt, b, e, smem});
            __syncthreads();
            if (q.chunks == 1) {
                for (uint32_t s = threadIdx.x; s < cnt; s += blockDim.x) flush_slot_store(smem, cnt, q.p, s, q.st, q.S, lo + s);
            } else {
                for (uint32_t s = threadIdx.x; s < cnt; s += blockDim.x) flush_slot_atomic(smem, cnt, q.p, s, q.st, q.S, lo + s);
            }
            __syncthreads();
            continue;
        }
        const uint64_t pb = offs[q.hbase + uint64_t(it.part) * q.ntiles];
        const uint64_t pe = (it.part + 1 < q.nparts) ? offs[q.hbase + uint64_t(it.part + 1) * q.ntiles] : q.ebase + q.rows;
        const uint64_t b = pb + (pe - pb) * it.chunk / q.chunks, e = pb + (pe - pb) * (it.chunk + 1) / q.chunks;
        for (uint64_t k0 = b; k0 < e; k0 += kRowUnroll * uint64_t(blockDim.x)) {
            uint16_t w[kRowUnroll];
            uint64_t bits[kRowUnroll];
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u) {
                const uint64_t k = k0 + uint64_t(u) * blockDim.x + threadIdx.x;
                const uint64_t kc = k < e ? k : e - 1;
                w[u] = __builtin_nontemporal_load(ekey + kc);
                bits[u] = __builtin_nontemporal_load(eval + kc);
            }
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u) {
                const uint64_t k = k0 + uint64_t(u) * blockDim.x + threadIdx.x;
                if (k < e) fold_value(smem, cnt, q.p, w[u] & ((1u << kEntryValid) - 1), bits[u], w[u] >> kEntryValid, lo);
            }
        }
        __syncthreads();
        if (q.chunks == 1) {  // sole owner of these slots
            for (uint32_t s = threadIdx.x; s < cnt; s += blockDim.x) flush_slot_store(smem, cnt, q.p, s, q.st, q.S, lo + s);
        } else {
            for (uint32_t s = threadIdx.x; s < cnt; s += blockDim.x) flush_slot_atomic(smem, cnt, q.p, s, q.st, q.S, lo + s);
        }
        __syncthreads();
    }
}

// Per-batch baseline: one launch per <= 2048-row batch of one query.
__global__ void __launch_bounds__(256) batch_one_kernel(const QueryDev* qs, uint32_t qi, uint64_t begin, uint64_t end,
                                                        uint32_t C) {
    extern __shared__ __align__(16) char smem[];
    const QueryDev& q = qs[qi];
    const uint64_t per = (end - begin + gridDim.x - 1) / gridDim.x;
    Tile t{qi, 0, begin + per * blockIdx.x, min(end, begin + per * (blockIdx.x + 1))};
    if (q.kind != BK_HASH) {
        uint64_t vrow[kMaxAggs];
        for (uint64_t r = t.begin + threadIdx.x; r < t.end; r += blockDim.x) {
            for (int a = 0; a < q.p.n; ++a) vrow[a] = r;
            global_update(q.st, q.S, q.p, direct_slot(q, r), vrow);
        }
        return;
    }
    if (C) {
        cache_reset(smem, C, q.p);
        __syncthreads();
    }
    if (t.begin < t.end) hash_tile(q, t, smem, C);
    if (C) {
        __syncthreads();
        cache_flush(smem, C, q);
    }
}

}  // namespace qe_gpu

// ---- host side -----------------------------------------------------------------------------
struct BatchPlan {
    uint32_t nq = 0;
    std::vector<QueryDev> hq;
    std::vector<Scratch> bufs;  // per-query state and hash tables
    std::vector<std::vector<uint64_t>> morsels;
    Scratch dq, ovf, next;
    bool any_hash = false;
    size_t small_lds = 0;  // largest small-kind table
    uint32_t max_parts = 0, max_pbits = 0;
    size_t part_lds = 16;  // dynamic LDS of the part scatter
    // Layout for the current tile size.
    uint32_t tile_rows = 0;
    Scratch ktiles[3], htiles, items, hist, offs, sums, ekey, eval;  // ktiles: per strategy
    uint64_t n_ktiles[3] = {0, 0, 0}, n_htiles = 0, n_items = 0, nhist = 0, nent = 0;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_C = 0, graph_blocks = 0, graph_tile = 0;
    ~BatchPlan() {
        if (graph) hipGraphExecDestroy(graph);
    }
};

static void build_batch_plan(qe_gpu_ctx* ctx, BatchPlan& bp, const qe_batch_query* qs, uint32_t nq) {
    bp.nq = nq;
    bp.hq.resize(nq);
    bp.morsels.resize(nq);
    bp.ovf = Scratch(ctx, 8);
    bp.next = Scratch(ctx, 5 * 8);
    for (uint32_t i = 0; i < nq; ++i) {
        const qe_batch_query& q = qs[i];
        invalid(q.key != nullptr && q.value != nullptr, "batched query needs key and value columns");
        invalid(q.key->len == q.value->len, "key and value columns differ in length");
        check_rows32(q.key->len);
        invalid(q.morsel_offsets != nullptr && q.nmorsels >= 1, "batched query needs >=1 morsel");
        std::vector<uint64_t> offs(q.morsel_offsets, q.morsel_offsets + q.nmorsels + 1);
        invalid(offs.front() == 0 && offs.back() == q.key->len, "morsel offsets must cover the columns");
        for (size_t m = 0; m + 1 < offs.size(); ++m) invalid(offs[m] <= offs[m + 1], "morsel offsets must ascend");
        bp.morsels[i] = offs;
        const uint64_t n = q.key->len;
        ColView ins[kMaxAggs];
        int types[kMaxAggs];
        invalid(q.naggs >= 1 && q.aggs != nullptr, "batched query needs >=1 aggregate");
        bool on_value = true;
        for (uint32_t a = 0; a < q.naggs && a < uint32_t(kMaxAggs); ++a) {
            invalid(q.aggs[a].func == QE_AGG_COUNT_STAR || q.aggs[a].input_col <= 1,
                    "batched aggregate input must be 0 (key) or 1 (value)");
            const qe_dev_col* c = q.aggs[a].input_col == 0 ? q.key : q.value;
            if (q.aggs[a].func != QE_AGG_COUNT_STAR && q.aggs[a].input_col == 0) on_value = false;
            ins[a] = view(c);
            types[a] = c->type;
        }
        QueryDev d{};
        d.p = make_prog(q.aggs, q.naggs, ins, types);
        check(d.p.nacc <= kBatchMaxAcc && d.p.ncnt <= kBatchMaxAcc, QE_GPU_UNSUPPORTED,
              "too many aggregates for the batched path");
        d.hk.nkeys = 1;
        d.hk.key[0] = view(q.key);
        d.val = view(q.value);
        d.rows = n;
        d.on_value = on_value ? 1 : 0;
        d.kw = static_cast<int32_t>(type_width(q.key->type));
        d.vw = static_cast<int32_t>(type_width(q.value->type));
        d.kind = BK_HASH;
        // Dense integer keys: direct slots (small or partitioned).
        if (is_int_like(q.key->type) && q.key->type != QE_BOOL) {
            const MinMax mm = column_minmax(ctx, q.key, n);
            const uint64_t R = mm.valid ? static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1 : 0;
            if (R <= std::max<uint64_t>(2 * n, 1024) && R < (uint64_t(1) << 30)) {
                const uint64_t S = R + 1;
                d.kmin = mm.valid ? mm.mn : 0;
                d.S = S;
                if (lds_bytes(static_cast<uint32_t>(S), d.p) <= kSmallLds) {
                    d.kind = BK_SMALL;
                    bp.small_lds = std::max(bp.small_lds, lds_bytes(static_cast<uint32_t>(S), d.p));
                } else if (on_value && n < (uint64_t(1) << 32)) {  // LDS row counters are 32-bit
                    uint32_t pb = kMaxPartBits;
                    while (pb > 8 && lds_bytes(1u << pb, d.p) > kPartLds) --pb;
                    const uint64_t np64 = (S + (uint64_t(1) << pb) - 1) >> pb;
                    if (lds_bytes(1u << pb, d.p) <= kPartLds && np64 <= kPassMaxParts) {
                        d.kind = BK_PART;
                        d.pbits = pb;
                        d.nparts = static_cast<uint32_t>(np64);
                        d.pass = 1;
                    } else if (lds_bytes(1u << pb, d.p) <= kPartLds && np64 <= kMaxParts) {
                        d.kind = BK_PART;
                        d.pbits = pb;
                        d.nparts = static_cast<uint32_t>(np64);
                        bp.max_parts = std::max(bp.max_parts, d.nparts);
                        bp.part_lds = std::max(bp.part_lds, d.nparts <= kStageMaxParts ? stage_lds_bytes(d.nparts)
                                                                                       : size_t(d.nparts) * 12);
                        bp.max_pbits = std::max(bp.max_pbits, pb);
                    }
                }
            }
        }
        if (d.kind == BK_HASH) {
            // Capacity from the distinct-group bound the caller supplies (or the
            // row count), load <= 0.5.
            const uint64_t bound = q.group_bound ? std::min<uint64_t>(q.group_bound, n) : n;
            uint64_t cap = 1024;
            while (cap < 2 * bound + 2) cap <<= 1;
            check(cap <= (uint64_t(1) << 31), QE_GPU_UNSUPPORTED, "batched query too large");
            d.S = cap;
            Scratch state(ctx, cap * 4), hash(ctx, cap * 8), words(ctx, cap * 8), nmask(ctx, cap * 8);
            d.t = HashTableView{state.as<uint32_t>(), hash.as<uint64_t>(), words.as<uint64_t>(),
                                nmask.as<uint64_t>(), cap, bp.ovf.as<unsigned int>()};
            bp.bufs.push_back(std::move(state));
            bp.bufs.push_back(std::move(hash));
            bp.bufs.push_back(std::move(words));
            bp.bufs.push_back(std::move(nmask));
            bp.any_hash = true;
        }
        Scratch st(ctx, state_words(d.S, d.p) * 8);
        d.st = st.as<uint64_t>();
        bp.bufs.push_back(std::move(st));
        // Exact F64 sums: the plan runs many times over the same columns, so
        // the scales come from every row, once, here.
        Scratch fb;
        prepare_f64(ctx, d.p, d.st, d.S, n, true, fb);
        bp.bufs.push_back(std::move(fb));
        bp.hq[i] = d;
    }
    bp.dq = Scratch(ctx, nq * sizeof(QueryDev));
}

// Work lists for a tile size: all tiles (main phase), part-query tiles
// (histogram), (query, partition) items, and the histogram / entry layout.
static void make_layout(qe_gpu_ctx* ctx, BatchPlan& bp, uint32_t tile_rows) {
    if (bp.tile_rows == tile_rows) return;
    std::vector<std::vector<Tile>> per(bp.nq);
    for (uint32_t q = 0; q < bp.nq; ++q) {
        // Small tables fold into shared state once per tile: longer tiles, fewer flushes.
        const uint64_t tr = bp.hq[q].kind == BK_SMALL ? uint64_t(tile_rows) * kSmallTileFactor : tile_rows;
        uint32_t qt = 0;
        for (size_t m = 0; m + 1 < bp.morsels[q].size(); ++m)
            for (uint64_t b = bp.morsels[q][m]; b < bp.morsels[q][m + 1]; b += tr)
                per[q].push_back(Tile{q, qt++, b, std::min<uint64_t>(bp.morsels[q][m + 1], b + tr)});
    }
    // Interleave queries so concurrent blocks work on different tables.
    std::vector<Tile> kt[3], ht;
    for (size_t k = 0;; ++k) {
        bool any = false;
        for (uint32_t q = 0; q < bp.nq; ++q)
            if (k < per[q].size() && !bp.hq[q].pass) {
                kt[bp.hq[q].kind].push_back(per[q][k]);
                if (bp.hq[q].kind == BK_PART) ht.push_back(per[q][k]);
                any = true;
            }
        if (!any) break;
    }
    uint64_t hbase = 0, ebase = 0, part_entries = 0;
    for (uint32_t q = 0; q < bp.nq; ++q)
        if (bp.hq[q].kind == BK_PART && !bp.hq[q].pass) part_entries += bp.hq[q].rows;
    const uint64_t chunk_entries =
        std::clamp<uint64_t>(part_entries / (2 * uint64_t(ctx->num_cus)), kChunkEntriesMin, kChunkEntriesMax);
    std::vector<std::vector<PartItem>> qitems(bp.nq), pitems(bp.nq);
    for (uint32_t q = 0; q < bp.nq; ++q) {
        QueryDev& d = bp.hq[q];
        if (d.kind != BK_PART) continue;
        if (d.pass) {  // row slices of the input; a slice's partitions adjacent, so later passes find it in cache
            d.ntiles = 0;
            d.chunks = static_cast<uint32_t>(std::max<uint64_t>(1, (d.rows + kPassChunkRows - 1) / kPassChunkRows));
            for (uint32_t c = 0; c < d.chunks; ++c)
                for (uint32_t p = 0; p < d.nparts; ++p) pitems[q].push_back(PartItem{q, p, c, 0});
            continue;
        }
        d.ntiles = static_cast<uint32_t>(per[q].size());
        d.hbase = hbase;
        d.ebase = ebase;
        hbase += uint64_t(d.nparts) * d.ntiles;
        ebase += d.rows;
        // Blocks per partition from the expected partition length.
        const uint64_t expect = (d.rows + d.nparts - 1) / d.nparts;
        d.chunks = static_cast<uint32_t>(std::max<uint64_t>(1, (expect + chunk_entries - 1) / chunk_entries));
        for (uint32_t p = 0; p < d.nparts; ++p)
            for (uint32_t c = 0; c < d.chunks; ++c) qitems[q].push_back(PartItem{q, p, c, 0});
    }
    // Pass items first: they are the longest, and run while the short items fill in behind.
    std::vector<PartItem> it;
    for (auto* lists : {&pitems, &qitems})
        for (size_t k = 0;; ++k) {
            bool any = false;
            for (uint32_t q = 0; q < bp.nq; ++q)
                if (k < (*lists)[q].size()) {
                    it.push_back((*lists)[q][k]);
                    any = true;
                }
            if (!any) break;
        }
    for (int k = 0; k < 3; ++k) bp.n_ktiles[k] = kt[k].size();
    bp.n_htiles = ht.size();
    bp.n_items = it.size();
    bp.nhist = hbase;
    bp.nent = ebase;
    for (int k = 0; k < 3; ++k) bp.ktiles[k] = Scratch(ctx, std::max<uint64_t>(kt[k].size(), 1) * sizeof(Tile));
    bp.htiles = Scratch(ctx, std::max<uint64_t>(ht.size(), 1) * sizeof(Tile));
    bp.items = Scratch(ctx, std::max<uint64_t>(it.size(), 1) * sizeof(PartItem));
    bp.hist = Scratch(ctx, std::max<uint64_t>(hbase, 1) * 4);
    bp.offs = Scratch(ctx, std::max<uint64_t>(hbase, 1) * 8);
    bp.sums = Scratch(ctx, scan_sums_words(std::max<uint64_t>(hbase, 1)) * 8);
    bp.ekey = Scratch(ctx, std::max<uint64_t>(ebase, 1) * 2);
    bp.eval = Scratch(ctx, std::max<uint64_t>(ebase, 1) * 8);
    for (int k = 0; k < 3; ++k)
        if (!kt[k].empty())
            QE_HIP(hipMemcpyAsync(bp.ktiles[k].p, kt[k].data(), kt[k].size() * sizeof(Tile), hipMemcpyHostToDevice,
                                  ctx->stream));
    if (!ht.empty())
        QE_HIP(hipMemcpyAsync(bp.htiles.p, ht.data(), ht.size() * sizeof(Tile), hipMemcpyHostToDevice, ctx->stream));
    if (!it.empty())
        QE_HIP(hipMemcpyAsync(bp.items.p, it.data(), it.size() * sizeof(PartItem), hipMemcpyHostToDevice, ctx->stream));
    QE_HIP(hipMemcpyAsync(bp.dq.p, bp.hq.data(), bp.nq * sizeof(QueryDev), hipMemcpyHostToDevice, ctx->stream));
    QE_HIP(hipStreamSynchronize(ctx->stream));
    bp.tile_rows = tile_rows;
    if (bp.graph) {
        QE_HIP(hipGraphExecDestroy(bp.graph));
        bp.graph = nullptr;
    }
}


static void enqueue_persistent(qe_gpu_ctx* ctx, BatchPlan& bp, uint32_t blocks, uint32_t C) {
    unsigned long long* next = bp.next.as<unsigned long long>();
    batch_reset_kernel<<<1, 1, 0, ctx->stream>>>(next, bp.ovf.as<unsigned int>());
    batch_init_kernel<<<ctx->num_cus * 4, 256, 0, ctx->stream>>>(bp.dq.as<QueryDev>(), bp.nq, 0);
    if (bp.n_htiles) {
        batch_hist_kernel<<<blocks, 256, size_t(bp.max_parts) * 4, ctx->stream>>>(
            bp.dq.as<QueryDev>(), bp.htiles.as<Tile>(), bp.n_htiles, next, bp.hist.as<uint32_t>());
        exclusive_scan_u32_dev(ctx, bp.hist.as<uint32_t>(), bp.offs.as<uint64_t>(), bp.nhist, bp.sums.as<uint64_t>());
    }
    const uint64_t* offs = bp.offs.as<uint64_t>();
    if (bp.n_ktiles[BK_SMALL])
        batch_tiles_kernel<BK_SMALL><<<blocks, 256, std::max<size_t>(bp.small_lds, 16), ctx->stream>>>(
            bp.dq.as<QueryDev>(), bp.ktiles[BK_SMALL].as<Tile>(), bp.n_ktiles[BK_SMALL], next + 1, C, offs,
            bp.ekey.as<uint16_t>(), bp.eval.as<uint64_t>());
    if (bp.n_ktiles[BK_PART])
        batch_tiles_kernel<BK_PART><<<blocks, 256, bp.part_lds, ctx->stream>>>(
            bp.dq.as<QueryDev>(), bp.ktiles[BK_PART].as<Tile>(), bp.n_ktiles[BK_PART], next + 2, C, offs,
            bp.ekey.as<uint16_t>(), bp.eval.as<uint64_t>());
    if (bp.n_ktiles[BK_HASH])
        batch_tiles_kernel<BK_HASH><<<blocks, 256, std::max<size_t>(cache_bytes(C), 16), ctx->stream>>>(
            bp.dq.as<QueryDev>(), bp.ktiles[BK_HASH].as<Tile>(), bp.n_ktiles[BK_HASH], next + 3, C, offs,
            bp.ekey.as<uint16_t>(), bp.eval.as<uint64_t>());
    if (bp.n_items)
        batch_part_agg_kernel<<<ctx->num_cus, kPartAggThreads, kPartLds, ctx->stream>>>(
            bp.dq.as<QueryDev>(), bp.items.as<PartItem>(), bp.n_items, next + 4, offs, bp.ekey.as<uint16_t>(),
            bp.eval.as<uint64_t>());
}

static void run_batch(qe_gpu_ctx* ctx, BatchPlan& bp, int mode, const qe_batch_tuning& tn) {
    const uint32_t C = tn.lds_slots;
    check(C == 0 || (C & (C - 1)) == 0, QE_GPU_INVALID_ARGUMENT, "lds_slots must be 0 or a power of two");
    check(cache_bytes(C) <= 64 * 1024, QE_GPU_INVALID_ARGUMENT, "lds_slots too large");
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(tn.blocks_per_cu, 16)) * ctx->num_cus;
    // Tiles of at most 2^22 rows (2^24 for small tables) keep the per-tile LDS counters in 32 bits.
    make_layout(ctx, bp, std::min<uint32_t>(std::max<uint32_t>(256, tn.tile_rows), 1u << 22));
    if (mode == QE_BATCH_PER_BATCH) {
        batch_reset_kernel<<<1, 1, 0, ctx->stream>>>(bp.next.as<unsigned long long>(), bp.ovf.as<unsigned int>());
        batch_init_kernel<<<ctx->num_cus * 4, 256, 0, ctx->stream>>>(bp.dq.as<QueryDev>(), bp.nq, 1);
        TimedRegion tr(ctx, "batch_per_batch");
        for (uint32_t q = 0; q < bp.nq; ++q)
            for (size_t m = 0; m + 1 < bp.morsels[q].size(); ++m)
                for (uint64_t b = bp.morsels[q][m]; b < bp.morsels[q][m + 1]; b += 2048) {
                    const uint64_t e = std::min<uint64_t>(bp.morsels[q][m + 1], b + 2048);
                    batch_one_kernel<<<8, 256, cache_bytes(C), ctx->stream>>>(bp.dq.as<QueryDev>(), q, b, e, C);
                }
        QE_HIP(hipGetLastError());
        return;
    }
    if (mode == QE_BATCH_GRAPH) {
        if (!bp.graph || bp.graph_C != C || bp.graph_blocks != blocks || bp.graph_tile != bp.tile_rows) {
            if (bp.graph) QE_HIP(hipGraphExecDestroy(bp.graph));
            bp.graph = nullptr;
            hipStream_t cs;
            QE_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t g;
            hipStream_t saved = ctx->stream;
            ctx->stream = cs;
            QE_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            enqueue_persistent(ctx, bp, blocks, C);
            QE_HIP(hipStreamEndCapture(cs, &g));
            ctx->stream = saved;
            QE_HIP(hipGraphInstantiate(&bp.graph, g, nullptr, nullptr, 0));
            QE_HIP(hipGraphDestroy(g));
            QE_HIP(hipStreamDestroy(cs));
            bp.graph_C = C;
            bp.graph_blocks = blocks;
            bp.graph_tile = bp.tile_rows;
        }
        TimedRegion tr(ctx, "batch_graph");
        QE_HIP(hipGraphLaunch(bp.graph, ctx->stream));
        return;
    }
    TimedRegion tr(ctx, "batch_persistent");
    enqueue_persistent(ctx, bp, blocks, C);
    QE_HIP(hipGetLastError());
}

static void emit_batch(qe_gpu_ctx* ctx, BatchPlan& bp, const qe_batch_query* qs, qe_dev_table** outs) {
    check(read_scalar(ctx, bp.ovf.as<unsigned int>()) == 0, QE_GPU_OVERFLOW,
          "batched group table overflow (group_bound too small?)");
    for (uint32_t i = 0; i < bp.nq; ++i) {
        const QueryDev& d = bp.hq[i];
        KeyEmit ke{};
        ke.nkeys = 1;
        ke.types[0] = qs[i].key->type;
        if (d.kind == BK_HASH) {
            ke.mode = 1;
            ke.words = d.t.words;
            ke.nullmask = d.t.nullmask;
        } else {
            ke.mode = 0;
            ke.kmin = d.kmin;
            ke.null_slot = d.S - 1;
        }
        outs[i] = emit_groups(ctx, d.st, d.S, d.p, ke, false);
    }
}

struct qe_batch_plan {
    BatchPlan bp;
};

extern "C" {

qe_gpu_status qe_gpu_batch_prepare(qe_gpu_ctx* ctx, const qe_batch_query* queries, uint32_t nq, qe_batch_plan** out) {
    QE_ENTRY(ctx, {
        invalid(queries != nullptr && out != nullptr && nq >= 1, "batched group-by needs >=1 query");
        auto* p = new qe_batch_plan();
        try {
            build_batch_plan(ctx, p->bp, queries, nq);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

qe_gpu_status qe_gpu_batch_run(qe_gpu_ctx* ctx, qe_batch_plan* plan, const qe_batch_query* queries, int32_t mode,
                               const qe_batch_tuning* tuning, qe_dev_table** outs) {
    QE_ENTRY(ctx, {
        invalid(plan != nullptr && queries != nullptr, "null argument");
        invalid(mode >= QE_BATCH_PERSISTENT && mode <= QE_BATCH_GRAPH, "unknown batch mode");
        qe_batch_tuning tn = tuning ? *tuning : qe_batch_tuning{16384, 4, 0};
        run_batch(ctx, plan->bp, mode, tn);
        if (outs) emit_batch(ctx, plan->bp, queries, outs);
        else QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_batch_autotune(qe_gpu_ctx* ctx, qe_batch_plan* plan, qe_batch_tuning* out) {
    QE_ENTRY(ctx, {
        invalid(plan != nullptr && out != nullptr, "null argument");
        double best = 1e30;
        qe_batch_tuning bt{16384, 4, 0};
        std::vector<uint32_t> caches{0u};
        if (plan->bp.any_hash) caches.push_back(256u);
        for (uint32_t tile : {16384u, 65536u})
            for (uint32_t bpc : {2u, 4u})
                for (uint32_t C : caches) {
                    qe_batch_tuning tn{tile, bpc, C};
                    run_batch(ctx, plan->bp, QE_BATCH_PERSISTENT, tn);  // warm (and lay out)
                    QE_HIP(hipStreamSynchronize(ctx->stream));
                    const auto t0 = std::chrono::steady_clock::now();
                    run_batch(ctx, plan->bp, QE_BATCH_PERSISTENT, tn);
                    QE_HIP(hipStreamSynchronize(ctx->stream));
                    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (s < best) {
                        best = s;
                        bt = tn;
                    }
                }
        *out = bt;
    });
}

qe_gpu_status qe_gpu_batch_strategies(qe_gpu_ctx* ctx, const qe_batch_plan* plan, int32_t* kinds, uint32_t* nparts) {
    QE_ENTRY(ctx, {
        invalid(plan != nullptr && kinds != nullptr, "null argument");
        for (uint32_t q = 0; q < plan->bp.nq; ++q) {
            const QueryDev& d = plan->bp.hq[q];
            kinds[q] = d.kind == BK_PART && d.pass ? QE_BATCH_PART_PASS
                       : d.kind == BK_PART         ? QE_BATCH_PART
                       : d.kind == BK_SMALL        ? QE_BATCH_SMALL
                                                   : QE_BATCH_HASH;
            if (nparts) nparts[q] = d.kind == BK_PART ? d.nparts : 0;
        }
    });
}

void qe_gpu_batch_free(qe_gpu_ctx* ctx, qe_batch_plan* plan) {
    if (!ctx || !plan) return;
    hipStreamSynchronize(ctx->stream);
    delete plan;
}

__global__ void gen_uniform_kernel(uint64_t seed, uint64_t stream, uint64_t n, int type, double lo, double hi,
                                   void* out) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t h = mix64(mix64(seed ^ (stream * 0x9E3779B97F4A7C15ull)) ^ (r * 0xD1B54A32D192ED03ull));
        const double u = static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
        const double v = lo + u * (hi - lo);
        switch (type) {
            case QE_I32:
                static_cast<int32_t*>(out)[r] = static_cast<int32_t>(floor(v));
                break;
            case QE_F64:
                static_cast<double*>(out)[r] = v;
                break;
            default:
                static_cast<int64_t*>(out)[r] = static_cast<int64_t>(floor(v));
        }
    }
}
__global__ void gen_nulls_kernel(uint64_t seed, uint64_t stream, uint64_t n, double p, uint64_t* words) {
    const uint64_t nw = words_for(n);
    for (uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; w < nw;
         w += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t bits = 0;
        for (unsigned b = 0; b < 64; ++b) {
            const uint64_t r = w * 64 + b;
            const uint64_t h = mix64(mix64(seed ^ ((stream + 1) * 0x9E3779B97F4A7C15ull)) ^ (r * 0xD1B54A32D192ED03ull));
            const double u = static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
            bits |= uint64_t(r >= n || u >= p) << b;
        }
        words[w] = bits;
    }
}

qe_gpu_status qe_gpu_gen_uniform(qe_gpu_ctx* ctx, uint64_t seed, uint64_t stream, qe_dev_col* col, double lo,
                                 double hi, double null_frac) {
    QE_ENTRY(ctx, {
        invalid(col != nullptr, "null column");
        invalid(col->type == QE_I32 || col->type == QE_I64 || col->type == QE_TS || col->type == QE_F64,
                "uniform generator supports I32/I64/TS/F64");
        const uint64_t n = col->len;
        if (n) {
            gen_uniform_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(seed, stream, n, col->type,
                                                                                             lo, hi, col->data);
            if (null_frac > 0) {
                invalid(col->validity != nullptr, "column must be nullable for NULLs");
                gen_nulls_kernel<<<grid_for(words_for(n), 256, 4096), 256, 0, ctx->stream>>>(seed, stream, n, null_frac,
                                                                                            col->validity);
                col->all_valid = false;
            }
            QE_HIP(hipGetLastError());
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
