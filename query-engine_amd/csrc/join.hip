// Inner/left hash equi-join and the fused join -> group-by.
//
// qe_gpu_hash_join replaces qe::HashJoin (ops/join.h:79-81): the build side is
// indexed once (build_side, ops/join.cpp:133-197) and every probe row expands
// to one pair per matching build row (build_pairs_for_probe, :199-247); LEFT
// keeps unmatched probe rows with a NULL build row. NULL keys never match
// (NullPolicy::kNeverMatch); F64 keys compare canonically.
//
// Build index, chosen from the data at build time:
//  * direct  : one integer-like key whose value range R is at most a few times
//              the build row count -> an R-entry array indexed by key - min;
//  * hash    : otherwise, the shared device hash table over canonical tuples.
// Unique build keys (detected, not assumed) map a key straight to its build
// row; duplicate keys use CSR row lists.
//
// qe_gpu_join_aggregate fuses Aggregate(HashJoin(...)) (plan/plan.cpp:581-590).
// When every group key is a build column and every aggregate reads a probe
// column (or is COUNT(*)), the build side is pre-grouped: each build row gets
// the slot of its group key, the direct index stores that slot (u16 when the
// groups fit), and the probe streams once, folding each matching row into
// per-block LDS partials that a block-order merge combines. Other shapes run
// join pairs -> gather -> aggregate.
#include "agg_common.hpp"
#include "hash_common.hpp"

#include <algorithm>

using namespace qe_gpu;

namespace qe_gpu {

bool is_int_like(int t);

struct JoinProbeView {
    int mode;    // 0 direct, 1 hash
    int unique;  // 1: head[slot] is the build row
    int nkeys;
    ColView pkey[kMaxKeys];
    int64_t kmin;
    uint64_t range;
    const uint32_t* head;  // unique: slot -> build row (kNoRow = none)
    HashTableView t;
    const uint64_t* offs;  // lists: slot -> [offs[slot], offs[slot+1])
    const uint32_t* rows;

    // Slot of probe row p's key, or kNoSlot (NULL key, out of range, absent).
    __device__ uint64_t slot(uint64_t p) const {
        if (mode == 0) {
            if (!is_valid(pkey[0], p)) return kNoSlot;
            const uint64_t off = static_cast<uint64_t>(int_value(pkey[0], p) - kmin);
            return off < range ? off : kNoSlot;
        }
        HashKeys hk;
        hk.nkeys = nkeys;
        for (int j = 0; j < nkeys; ++j) hk.key[j] = pkey[j];
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, p, w, nm, h);
        if (nm != 0) return kNoSlot;  // kNeverMatch: a NULL in any key never joins
        const uint32_t s = ht_find(t, nkeys, w, nm, h);
        return s == kNoSlot ? kNoSlot : s;
    }
    // Number of matching build rows; `begin` indexes row_at().
    __device__ uint32_t probe(uint64_t p, uint64_t& begin) const {
        const uint64_t s = slot(p);
        if (s == kNoSlot) return 0;
        if (unique) {
            if (head[s] == kNoRow) return 0;
            begin = s;
            return 1;
        }
        begin = offs[s];
        return static_cast<uint32_t>(offs[s + 1] - offs[s]);
    }
    __device__ uint32_t row_at(uint64_t i) const { return unique ? head[i] : rows[i]; }
};

struct JoinIndex {
    int mode = 0;
    bool unique = true;
    bool empty = false;  // no matchable build row
    int64_t kmin = 0;
    uint64_t range = 0;
    uint64_t nslots = 0;
    DeviceHashTable ht;
    Scratch head, offs, rows;
    JoinProbeView view(const qe_dev_col* const* pkeys, int nkeys) const {
        JoinProbeView v{};
        v.mode = mode;
        v.unique = unique ? 1 : 0;
        v.nkeys = nkeys;
        for (int j = 0; j < nkeys; ++j) v.pkey[j] = qe_gpu::view(pkeys[j]);
        v.kmin = kmin;
        v.range = empty ? 0 : range;
        v.head = head.as<uint32_t>();
        v.t = ht.v;
        v.offs = offs.as<uint64_t>();
        v.rows = rows.as<uint32_t>();
        return v;
    }
};

__global__ void direct_fill_kernel(ColView key, uint64_t n, int64_t kmin, uint32_t* head,
                                   unsigned int* dup) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        if (!is_valid(key, r)) continue;
        const uint64_t off = static_cast<uint64_t>(int_value(key, r) - kmin);
        const uint32_t old = atomicCAS(&head[off], kNoRow, static_cast<uint32_t>(r));
        if (old != kNoRow) atomicOr(dup, 1u);
    }
}

__global__ void hash_fill_kernel(HashKeys hk, uint64_t n, HashTableView t, uint32_t* head,
                                 uint32_t* slot_of, unsigned int* dup) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);
        if (nm != 0) {  // NULL key: never matches, not indexed
            slot_of[r] = kNoRow;
            continue;
        }
        const uint32_t s = ht_insert(t, hk.nkeys, w, nm, h);
        slot_of[r] = s;
        if (s == kNoSlot) continue;
        const uint32_t old = atomicCAS(&head[s], kNoRow, static_cast<uint32_t>(r));
        if (old != kNoRow) atomicOr(dup, 1u);
    }
}

// Lists for duplicate keys: count per slot, scan, scatter rows.
__global__ void list_count_kernel(const JoinProbeView v, const uint32_t* slot_of, ColView key0, uint64_t n,
                                  uint32_t* counts) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t s;
        if (v.mode == 0) {
            if (!is_valid(key0, r)) continue;
            s = static_cast<uint64_t>(int_value(key0, r) - v.kmin);
        } else {
            s = slot_of[r];
            if (s == kNoRow) continue;
        }
        atomicAdd(&counts[s], 1u);
    }
}
__global__ void list_fill_kernel(const JoinProbeView v, const uint32_t* slot_of, ColView key0, uint64_t n,
                                 const uint64_t* offs, uint32_t* cursor, uint32_t* rows) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t s;
        if (v.mode == 0) {
            if (!is_valid(key0, r)) continue;
            s = static_cast<uint64_t>(int_value(key0, r) - v.kmin);
        } else {
            s = slot_of[r];
            if (s == kNoRow) continue;
        }
        const uint32_t k = atomicAdd(&cursor[s], 1u);
        rows[offs[s] + k] = static_cast<uint32_t>(r);
    }
}

// Within each list, order rows ascending (the reference's group_rows order);
// lists are short, so one thread per slot with insertion sort.
__global__ void list_sort_kernel(const uint64_t* offs, uint64_t nslots, uint32_t* rows) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < nslots;
         s += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t b = offs[s], e = offs[s + 1];
        if (e - b < 2 || e - b > 64) continue;
        for (uint64_t i = b + 1; i < e; ++i) {
            const uint32_t x = rows[i];
            uint64_t j = i;
            while (j > b && rows[j - 1] > x) {
                rows[j] = rows[j - 1];
                --j;
            }
            rows[j] = x;
        }
    }
}

void build_index(qe_gpu_ctx* ctx, JoinIndex& ix, const qe_dev_col* const* bkeys, uint32_t nkeys, uint64_t nb) {
    Scratch dup(ctx, 8);
    memset_async(ctx, dup.p, 0, 8);
    Scratch slot_of;
    const qe_dev_col* k0 = bkeys[0];
    bool direct = false;
    if (nkeys == 1 && is_int_like(k0->type)) {
        const MinMax mm = column_minmax(ctx, k0, nb);
        if (mm.valid == 0) {
            ix.empty = true;
            ix.mode = 0;
            ix.range = 0;
            return;
        }
        const uint64_t R = static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1;
        if (R <= std::max<uint64_t>(4 * nb, uint64_t(1) << 16) && R < (uint64_t(1) << 31)) {
            direct = true;
            ix.mode = 0;
            ix.kmin = mm.mn;
            ix.range = R;
            ix.nslots = R;
            ix.head = Scratch(ctx, R * 4);
            memset_async(ctx, ix.head.p, 0xFF, R * 4);
            direct_fill_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
                view(k0), nb, ix.kmin, ix.head.as<uint32_t>(), dup.as<unsigned int>());
            QE_HIP(hipGetLastError());
        }
    }
    if (!direct) {
        check(nkeys <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many join keys for the device path");
        ix.mode = 1;
        ix.ht.create(ctx, static_cast<int>(nkeys), nb);
        ix.nslots = ix.ht.v.cap;
        ix.range = ix.ht.v.cap;
        ix.head = Scratch(ctx, ix.nslots * 4);
        memset_async(ctx, ix.head.p, 0xFF, ix.nslots * 4);
        slot_of = Scratch(ctx, std::max<uint64_t>(nb, 1) * 4);
        HashKeys hk{};
        hk.nkeys = static_cast<int>(nkeys);
        for (uint32_t j = 0; j < nkeys; ++j) hk.key[j] = view(bkeys[j]);
        hash_fill_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
            hk, nb, ix.ht.v, ix.head.as<uint32_t>(), slot_of.as<uint32_t>(), dup.as<unsigned int>());
        QE_HIP(hipGetLastError());
        ix.ht.check_overflow(ctx);
    }
    ix.unique = read_scalar(ctx, dup.as<unsigned int>()) == 0;
    if (ix.unique) return;
    // Duplicate keys: CSR lists per slot.
    const uint64_t S = ix.nslots;
    Scratch counts(ctx, S * 4), cursor(ctx, S * 4);
    memset_async(ctx, counts.p, 0, S * 4);
    memset_async(ctx, cursor.p, 0, S * 4);
    ix.offs = Scratch(ctx, (S + 1) * 8);
    JoinProbeView v = ix.view(bkeys, static_cast<int>(nkeys));
    list_count_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
        v, slot_of.as<uint32_t>(), view(k0), nb, counts.as<uint32_t>());
    QE_HIP(hipGetLastError());
    const uint64_t total = exclusive_scan_u32(ctx, counts.as<uint32_t>(), ix.offs.as<uint64_t>(), S);
    QE_HIP(hipMemcpyAsync(ix.offs.as<uint64_t>() + S, &total, 8, hipMemcpyHostToDevice, ctx->stream));
    ix.rows = Scratch(ctx, std::max<uint64_t>(total, 1) * 4);
    list_fill_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
        v, slot_of.as<uint32_t>(), view(k0), nb, ix.offs.as<uint64_t>(), cursor.as<uint32_t>(),
        ix.rows.as<uint32_t>());
    list_sort_kernel<<<grid_for(S, 256, 8192), 256, 0, ctx->stream>>>(ix.offs.as<uint64_t>(), S,
                                                                       ix.rows.as<uint32_t>());
    QE_HIP(hipGetLastError());
    QE_HIP(hipStreamSynchronize(ctx->stream));  // `total` staging lives on the host stack
}

void check_key_types(const qe_dev_col* const* pk, const qe_dev_col* const* bk, uint32_t nkeys) {
    for (uint32_t j = 0; j < nkeys; ++j) {
        invalid(pk[j] != nullptr && bk[j] != nullptr, "null key column");
        // ops/join.cpp:118-122 (a real check in every build).
        invalid(pk[j]->type == bk[j]->type, "HashJoin: join key column types must match positionally");
    }
}

// ---- pairs -------------------------------------------------------------------------
__global__ void pair_count_kernel(JoinProbeView v, uint64_t n, int left, uint32_t* counts) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < n;
         p += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t b;
        const uint32_t c = v.probe(p, b);
        counts[p] = (c == 0 && left) ? 1u : c;
    }
}
__global__ void pair_write_kernel(JoinProbeView v, uint64_t n, int left, const uint64_t* offs,
                                  uint32_t* pp, uint32_t* bp) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < n;
         p += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t b = 0;
        const uint32_t c = v.probe(p, b);
        uint64_t o = offs[p];
        if (c == 0) {
            if (left) {
                pp[o] = static_cast<uint32_t>(p);
                bp[o] = kNoRow;
            }
            continue;
        }
        for (uint32_t i = 0; i < c; ++i, ++o) {
            pp[o] = static_cast<uint32_t>(p);
            bp[o] = v.row_at(b + i);
        }
    }
}

qe_dev_pairs* join_pairs(qe_gpu_ctx* ctx, const qe_dev_col* const* pk, const qe_dev_col* const* bk,
                         uint32_t nkeys, int join_type) {
    invalid(nkeys >= 1, "equi-join needs >=1 key");
    invalid(join_type == QE_JOIN_INNER || join_type == QE_JOIN_LEFT, "unknown join type");
    check_key_types(pk, bk, nkeys);
    const uint64_t np = common_len(pk, nkeys), nb = common_len(bk, nkeys);
    check(np < kNoRow && nb < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
    JoinIndex ix;
    if (nb > 0) build_index(ctx, ix, bk, nkeys, nb);
    else ix.empty = true;
    JoinProbeView v = ix.view(pk, static_cast<int>(nkeys));
    if (ix.empty) {
        v.mode = 0;
        v.range = 0;
        v.unique = 1;
    }
    const int left = join_type == QE_JOIN_LEFT;
    auto* out = new qe_dev_pairs();
    if (np == 0) return out;
    Scratch counts(ctx, np * 4), offs(ctx, np * 8);
    {
        TimedRegion tr(ctx, "join_count");
        pair_count_kernel<<<grid_for(np, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(v, np, left,
                                                                                          counts.as<uint32_t>());
        QE_HIP(hipGetLastError());
    }
    const uint64_t total = exclusive_scan_u32(ctx, counts.as<uint32_t>(), offs.as<uint64_t>(), np);
    out->count = total;
    try {
        out->probe = static_cast<uint32_t*>(dev_alloc(ctx, std::max<uint64_t>(total, 1) * 4));
        out->build = static_cast<uint32_t*>(dev_alloc(ctx, std::max<uint64_t>(total, 1) * 4));
        TimedRegion tr(ctx, "join_write");
        pair_write_kernel<<<grid_for(np, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
            v, np, left, offs.as<uint64_t>(), out->probe, out->build);
        QE_HIP(hipGetLastError());
        QE_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        qe_gpu_pairs_free(ctx, out);
        throw;
    }
    return out;
}

// ---- gather (emit) --------------------------------------------------------------------
// out[r] = src[idx[r]], NULL where idx[r] is kNoRow or the source row is NULL.
// Each wave takes 64 consecutive output rows, so the row validity is one ballot
// and the wave's lane 0 writes the output bitmap word (rows past the end read
// as valid) and counts the NULLs.
template <typename T>
__global__ void gather_kernel(const T* __restrict__ src, const uint64_t* __restrict__ src_valid,
                              const uint32_t* __restrict__ idx, uint64_t m, T* __restrict__ dst,
                              uint64_t* __restrict__ words, unsigned long long* nulls) {
    const uint64_t nw = words_for(m);
    const unsigned lane = lane_id();
    const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x / kWave);
    uint64_t local = 0;
    for (uint64_t w = blockIdx.x * uint64_t(blockDim.x / kWave) + threadIdx.x / kWave; w < nw; w += waves) {
        const uint64_t r = w * 64 + lane;
        bool ok = false;
        if (r < m) {
            const uint32_t s = idx[r];
            ok = s != kNoRow && (src_valid == nullptr || ((src_valid[s >> 6] >> (s & 63)) & 1ull));
            dst[r] = ok ? src[s] : T(0);
        }
        const uint64_t mask = __ballot(ok);
        if (lane == 0) {
            const uint64_t left = m - w * 64;
            const uint64_t in = left >= 64 ? ~0ull : ((1ull << left) - 1);
            words[w] = mask | ~in;
            local += __popcll(in & ~mask);
        }
    }
    if (lane == 0 && local) atomicAdd(nulls, static_cast<unsigned long long>(local));
}

std::vector<qe_dev_col*> gather_cols(qe_gpu_ctx* ctx, const qe_dev_col* const* srcs, const uint32_t* const* idxs,
                                     size_t n, uint64_t m) {
    std::vector<qe_dev_col*> out;
    try {
        Scratch cnt(ctx, std::max<size_t>(n, 1) * 8);
        if (m) memset_async(ctx, cnt.p, 0, n * 8);
        const unsigned g = grid_for(words_for(m) * kWave, 256, ctx->num_cus * 16);
        for (size_t c = 0; c < n; ++c) {
            const qe_dev_col* src = srcs[c];
            qe_dev_col* o = new_col(ctx, src->type, m, true);
            out.push_back(o);
            o->dict = src->dict;
            if (!m) continue;
            const uint64_t* sv = src->all_valid ? nullptr : src->validity;
            unsigned long long* nc = cnt.as<unsigned long long>() + c;
            switch (type_width(src->type)) {
                case 1:
                    gather_kernel<uint8_t><<<g, 256, 0, ctx->stream>>>(static_cast<const uint8_t*>(src->data), sv,
                                                                       idxs[c], m, static_cast<uint8_t*>(o->data),
                                                                       o->validity, nc);
                    break;
                case 4:
                    gather_kernel<uint32_t><<<g, 256, 0, ctx->stream>>>(static_cast<const uint32_t*>(src->data), sv,
                                                                        idxs[c], m, static_cast<uint32_t*>(o->data),
                                                                        o->validity, nc);
                    break;
                default:
                    gather_kernel<uint64_t><<<g, 256, 0, ctx->stream>>>(static_cast<const uint64_t*>(src->data), sv,
                                                                        idxs[c], m, static_cast<uint64_t*>(o->data),
                                                                        o->validity, nc);
            }
            QE_HIP(hipGetLastError());
        }
        // One read-back of every column's NULL count; a column without NULLs
        // drops its bitmap (all_valid).
        std::vector<uint64_t> nulls(n, 0);
        if (m && n) {
            uint64_t* h = static_cast<uint64_t*>(pinned(ctx, n * 8));
            QE_HIP(hipMemcpyAsync(h, cnt.p, n * 8, hipMemcpyDeviceToHost, ctx->stream));
            QE_HIP(hipStreamSynchronize(ctx->stream));
            std::copy(h, h + n, nulls.begin());
        }
        for (size_t c = 0; c < n; ++c) {
            if (nulls[c] == 0) {
                dev_free(ctx, out[c]->validity);
                out[c]->validity = nullptr;
                out[c]->all_valid = true;
            } else {
                out[c]->all_valid = false;
            }
        }
    } catch (...) {
        for (auto* c : out) free_col(ctx, c);
        throw;
    }
    return out;
}

qe_dev_col* gather_col(qe_gpu_ctx* ctx, const qe_dev_col* src, const uint32_t* idx, uint64_t m) {
    return gather_cols(ctx, &src, &idx, 1, m)[0];
}

// ---- fused join -> group-by -----------------------------------------------------------
struct ProbeAggLists {
    JoinProbeView v;
    const uint32_t* gslot;  // build row -> group slot
};

template <typename KeyT, typename MapT>
__global__ void __launch_bounds__(1024) join_agg_direct_lds(const KeyT* __restrict__ keys,
                                                            const uint64_t* key_valid, uint64_t n,
                                                            const MapT* __restrict__ gmap, int64_t kmin,
                                                            uint64_t range, uint32_t S, uint32_t slot_lo, AggProg p,
                                                            uint64_t* part) {
    extern __shared__ __align__(16) char smem[];
    lds_init(smem, S, p);
    __syncthreads();
    constexpr MapT kEmpty = static_cast<MapT>(~MapT(0));
    uint64_t vrow[kMaxAggs];
    const uint64_t nq = (n + 3) / 4;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; q < nq; q += stride) {
        const uint64_t r0 = q * 4;
        KeyT k[4];
        if (r0 + 4 <= n) {
#pragma unroll
            for (int j = 0; j < 4; ++j) k[j] = __builtin_nontemporal_load(keys + r0 + j);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) k[j] = (r0 + j < n) ? keys[r0 + j] : KeyT(0);
        }
        MapT g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t off = static_cast<uint64_t>(static_cast<int64_t>(k[j]) - kmin);
            bool ok = (r0 + j < n) && off < range;
            if (key_valid) ok = ok && ((key_valid[(r0 + j) >> 6] >> ((r0 + j) & 63)) & 1ull);
            g[j] = ok ? gmap[off] : kEmpty;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t sl = static_cast<uint32_t>(g[j]) - slot_lo;
            if (g[j] == kEmpty || sl >= S) continue;
            for (int a = 0; a < p.n; ++a) vrow[a] = r0 + j;
            lds_update(smem, S, p, sl, vrow, slot_lo);
        }
    }
    __syncthreads();
    lds_flush(smem, S, p, part);
}

template <typename KeyT, typename MapT>
__global__ void join_agg_direct_global(const KeyT* __restrict__ keys, const uint64_t* key_valid, uint64_t n,
                                       const MapT* __restrict__ gmap, int64_t kmin, uint64_t range, uint64_t S,
                                       AggProg p, uint64_t* st) {
    constexpr MapT kEmpty = static_cast<MapT>(~MapT(0));
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        if (key_valid && !((key_valid[r >> 6] >> (r & 63)) & 1ull)) continue;
        const uint64_t off = static_cast<uint64_t>(static_cast<int64_t>(keys[r]) - kmin);
        if (off >= range) continue;
        const MapT g = gmap[off];
        if (g == kEmpty) continue;
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        global_update(st, S, p, g, vrow);
    }
}

// ---- streaming probe: coalesced vector loads of key and value columns ----------------
// Value columns an aggregate program reads, deduplicated (at most kStreamCols).
constexpr int kStreamCols = 2;
struct StreamCols {
    int32_t wide[kStreamCols];  // 1: 8-byte values, 0: I32 (sign-extended)
    const void* data[kStreamCols];
    const uint64_t* validity[kStreamCols];  // nullptr when all-valid
    int32_t vcol[kMaxAggs];                 // agg -> value column, -1 for COUNT(*)
};

typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

// Rows 2*pr and 2*pr+1 of an 8-byte or 4-byte column in one coalesced load.
__device__ inline void load_pair(const void* base, bool wide, uint64_t pr, uint64_t& a, uint64_t& b) {
    if (wide) {
        const u64x2_t v = __builtin_nontemporal_load(static_cast<const u64x2_t*>(base) + pr);
        a = v.x;
        b = v.y;
    } else {
        const u32x2_t v = __builtin_nontemporal_load(static_cast<const u32x2_t*>(base) + pr);
        a = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(v.x)));
        b = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(v.y)));
    }
}

// Fused probe over a direct unique map with LDS partials. Each thread owns
// PAIRS row pairs per tile; pair i of thread t is tile_base + i*blockDim + t, so
// every load instruction of a wave reads one contiguous 1 KiB (8-byte column)
// or 512 B (4-byte column) span. NV = value columns read.
//
// Software-pipelined: vector memory loads complete in issue order (vmcnt), so
// a tile's map lookups are issued first, the NEXT tile's key and value loads
// second, and only then does the wave wait -- for the lookups alone -- and
// fold the current tile into LDS while the next tile's stream is in flight.
// Every load is issued unconditionally (rows past the end re-read the last
// pair, rows that cannot match look up entry 0; both are masked), so the
// number of loads after the lookups is static and the compiler's wait before
// the fold leaves the next tile's loads outstanding.
//
// Rows whose slot lies outside the LDS range are skipped (another launch covers
// their range) or, with SPILL, written to the block's spill list: record i of
// block b is words [(b * cap + i) * (1 + NV), ...) of `rec`, a header (slot |
// validity bits << 32) then the NV value words; `count[b]` receives the
// block's record count. Rows beyond `cap` records fold with HBM atomics into
// the state `st` of `S` slots. One probe pass then covers a slot space larger
// than LDS, and the lists (a few percent of the rows when the LDS range holds
// nearly all of them) are aggregated range by range afterwards.
struct Spill {
    uint64_t* st = nullptr;
    uint64_t S = 0;
    uint64_t* rec = nullptr;
    uint32_t cap = 0;
    uint32_t* count = nullptr;
};

template <int NV>
__device__ inline void spill_row(const Spill& spill, uint32_t* spill_n, const AggProg& p, uint32_t slot,
                                 const int32_t* vcol, const uint64_t* bits, uint32_t valid) {
    const uint32_t pos = atomicAdd(spill_n, 1u);
    if (pos < spill.cap) {
        uint64_t* r = spill.rec + (size_t(blockIdx.x) * spill.cap + pos) * (1 + NV);
        r[0] = slot | (static_cast<uint64_t>(valid) << 32);
#pragma unroll
        for (int c = 0; c < NV; ++c) r[1 + c] = bits[c];
    } else {
        global_update_vals(spill.st, spill.S, p, slot, vcol, bits, valid);
    }
}

// The key and value words of one tile (pairs t0 + i*blockDim + threadIdx.x),
// every load unconditional (pairs past the end re-read the last pair). Bit c
// of WMASK: value column c is 8 bytes wide (else I32, sign-extended).
template <typename KeyT, int PAIRS, int NV, int WMASK>
__device__ inline void stream_tile_load(const KeyT* __restrict__ keys, const StreamCols& sc, uint64_t t0,
                                        uint64_t npairs, uint64_t (&kw)[2 * PAIRS],
                                        uint64_t (&vb)[NV > 0 ? NV : 1][2 * PAIRS]) {
    constexpr bool kWideKey = sizeof(KeyT) == 8;
#pragma unroll
    for (int i = 0; i < PAIRS; ++i) {
        uint64_t pr = t0 + uint64_t(i) * blockDim.x + threadIdx.x;
        pr = pr < npairs ? pr : npairs - 1;
        load_pair(keys, kWideKey, pr, kw[2 * i], kw[2 * i + 1]);
#pragma unroll
        for (int c = 0; c < NV; ++c) load_pair(sc.data[c], ((WMASK >> c) & 1) != 0, pr, vb[c][2 * i], vb[c][2 * i + 1]);
    }
}

// FAST: the program has the fast fold's shape (lds_fold_fast), taken with its
// few fields in scalar registers instead of the general program loop.
template <typename KeyT, typename MapT, int PAIRS, int NV, int WMASK, bool SPILL, bool FAST>
__global__ void __launch_bounds__(1024) join_agg_stream(const KeyT* __restrict__ keys, const uint64_t* key_valid,
                                                        uint64_t n, const MapT* __restrict__ gmap, int64_t kmin,
                                                        uint64_t range, uint32_t S, uint32_t slot_lo, AggProg p,
                                                        StreamCols sc, uint64_t* part, Spill spill, FastFold ff) {
    // S slots of LDS partials for slots [slot_lo, slot_lo + S) of the map's
    // slot space.
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t spill_n;
    lds_init(smem, S, p);
    if (SPILL && threadIdx.x == 0) spill_n = 0;
    if constexpr (FAST) fast_fold_scales(ff, p);
    __syncthreads();
    constexpr MapT kEmpty = static_cast<MapT>(~MapT(0));
    constexpr int R = 2 * PAIRS;  // rows per thread per tile
    const uint64_t npairs = n >> 1;
    const uint64_t tile = uint64_t(blockDim.x) * PAIRS;
    const uint64_t step = uint64_t(gridDim.x) * tile;
    uint64_t t0 = blockIdx.x * tile;
    if (t0 < npairs) {
        uint64_t kw[R];
        uint64_t vb[NV > 0 ? NV : 1][R];
        stream_tile_load<KeyT, PAIRS, NV, WMASK>(keys, sc, t0, npairs, kw, vb);
        for (; t0 < npairs; t0 += step) {
            // 1. This tile's validity words, then its map lookups, issued first.
            uint32_t valid[R];
            bool ok[R];
#pragma unroll
            for (int i = 0; i < PAIRS; ++i) {
                const uint64_t pr = t0 + uint64_t(i) * blockDim.x + threadIdx.x;
                const uint64_t w = (2 * pr) >> 6;  // both rows of a pair share a validity word
                const uint32_t sh = static_cast<uint32_t>((2 * pr) & 63);
                uint64_t kvw = ~0ull;
                if (key_valid && pr < npairs) kvw = key_valid[w];
                uint32_t v0 = 0, v1 = 0;
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    uint64_t vw = ~0ull;
                    if (sc.validity[c] && pr < npairs) vw = sc.validity[c][w];
                    v0 |= static_cast<uint32_t>((vw >> sh) & 1ull) << c;
                    v1 |= static_cast<uint32_t>((vw >> (sh + 1)) & 1ull) << c;
                }
                valid[2 * i] = v0;
                valid[2 * i + 1] = v1;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const uint64_t off = kw[2 * i + e] - static_cast<uint64_t>(kmin);
                    ok[2 * i + e] = pr < npairs && off < range && ((kvw >> (sh + e)) & 1ull);
                }
            }
            MapT g[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t off = kw[j] - static_cast<uint64_t>(kmin);
                g[j] = gmap[ok[j] ? off : 0];
            }
            // 2. The next tile's stream, in flight while this tile folds.
            uint64_t kn[R];
            uint64_t vn[NV > 0 ? NV : 1][R];
            stream_tile_load<KeyT, PAIRS, NV, WMASK>(keys, sc, t0 + step, npairs, kn, vn);
            // Keep the fold below the prefetch: the scheduler would otherwise
            // hoist the first row's fold and wait for its lookup first.
            __builtin_amdgcn_sched_barrier(0);
            // 3. Fold this tile into the LDS partials.
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const MapT gj = ok[j] ? g[j] : kEmpty;
                const uint32_t sl = static_cast<uint32_t>(gj) - slot_lo;
                const bool in = gj != kEmpty && sl < S;
                if (!in && !(SPILL && gj != kEmpty)) continue;
                uint64_t bits[NV > 0 ? NV : 1];
#pragma unroll
                for (int c = 0; c < NV; ++c) bits[c] = vb[c][j];
                if (!SPILL || in) {
                    if constexpr (FAST)
                        lds_fold_fast<NV>(smem, S, ff, p, sl, bits, slot_lo);
                    else
                        lds_update_vals(smem, S, p, sl, sc.vcol, bits, valid[j], slot_lo);
                } else {
                    spill_row<NV>(spill, &spill_n, p, static_cast<uint32_t>(gj), sc.vcol, bits, valid[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < R; ++j) {
                kw[j] = kn[j];
#pragma unroll
                for (int c = 0; c < NV; ++c) vb[c][j] = vn[c][j];
            }
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {  // last row of an odd-length column
        const uint64_t r = n - 1;
        const uint64_t off = static_cast<uint64_t>(static_cast<int64_t>(keys[r])) - static_cast<uint64_t>(kmin);
        bool ok = off < range;
        if (key_valid) ok = ok && ((key_valid[r >> 6] >> (r & 63)) & 1ull);
        const MapT gg = ok ? gmap[off] : kEmpty;
        const bool in = gg != kEmpty && static_cast<uint32_t>(gg) - slot_lo < S;
        if (in || (SPILL && gg != kEmpty)) {
            uint64_t bits[NV > 0 ? NV : 1];
            uint32_t valid = 0;
#pragma unroll
            for (int c = 0; c < NV; ++c) {
                bits[c] = sc.wide[c] ? static_cast<const uint64_t*>(sc.data[c])[r]
                                     : static_cast<uint64_t>(static_cast<int64_t>(static_cast<const int32_t*>(sc.data[c])[r]));
                const bool v = sc.validity[c] == nullptr || ((sc.validity[c][r >> 6] >> (r & 63)) & 1ull);
                valid |= static_cast<uint32_t>(v) << c;
            }
            if (in)
                lds_update_vals(smem, S, p, static_cast<uint32_t>(gg) - slot_lo, sc.vcol, bits, valid, slot_lo);
            else
                spill_row<NV>(spill, &spill_n, p, static_cast<uint32_t>(gg), sc.vcol, bits, valid);
        }
    }
    __syncthreads();
    if (SPILL && threadIdx.x == 0) spill.count[blockIdx.x] = min(spill_n, spill.cap);
    lds_flush(smem, S, p, part);
}

// One slot range [slot_lo, slot_lo + S) of the spill lists in LDS partials;
// block b folds list b (one list per block of the probe pass).
template <int NV>
__global__ void __launch_bounds__(1024) spill_agg_kernel(const uint64_t* __restrict__ rec,
                                                         const uint32_t* __restrict__ count, uint32_t cap,
                                                         uint32_t slot_lo, uint32_t S, AggProg p, StreamCols sc,
                                                         uint64_t* part) {
    extern __shared__ __align__(16) char smem[];
    lds_init(smem, S, p);
    __syncthreads();
    const uint32_t n = count[blockIdx.x];
    const uint64_t* rb = rec + size_t(blockIdx.x) * cap * (1 + NV);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t h = rb[size_t(i) * (1 + NV)];
        const uint32_t sl = static_cast<uint32_t>(h) - slot_lo;
        if (sl >= S) continue;
        uint64_t bits[NV > 0 ? NV : 1];
#pragma unroll
        for (int c = 0; c < NV; ++c) bits[c] = rb[size_t(i) * (1 + NV) + 1 + c];
        lds_update_vals(smem, S, p, sl, sc.vcol, bits, static_cast<uint32_t>(h >> 32), slot_lo);
    }
    __syncthreads();
    lds_flush(smem, S, p, part);
}

// Group slots of every `stride`-th probe row, counted in cnt[S] (HBM atomics;
// about 2^16 samples, so the heaviest slot's atomics stay short).
template <typename KeyT, typename MapT>
__global__ void slot_sample_kernel(const KeyT* __restrict__ keys, const uint64_t* key_valid, uint64_t n,
                                   uint64_t stride, const MapT* __restrict__ gmap, int64_t kmin, uint64_t range,
                                   uint32_t* cnt) {
    constexpr MapT kEmpty = static_cast<MapT>(~MapT(0));
    const uint64_t ns = (n + stride - 1) / stride;
    for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < ns;
         j += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t r = j * stride;
        const uint64_t off = static_cast<uint64_t>(static_cast<int64_t>(keys[r])) - static_cast<uint64_t>(kmin);
        bool ok = off < range;
        if (key_valid) ok = ok && ((key_valid[r >> 6] >> (r & 63)) & 1ull);
        const MapT g = ok ? gmap[off] : kEmpty;
        if (g != kEmpty) atomicAdd(&cnt[g], 1u);
    }
}

// Hot set of at most W slots by sampled count: the slots with cnt >= c for the
// smallest c >= 1 that leaves at most W of them. perm[s] relabels the slots, the
// hot ones first (ids 0..h-1), each part in slot order. out = {h, samples in
// the hot set, all samples, largest count outside it}. One block of 1024
// threads; with `staged`, the threshold search reads the counts (capped at
// 65535, above any threshold it can pick) from S u16 of dynamic LDS.
constexpr uint32_t kHotStagedSlots = 75000;

__global__ void __launch_bounds__(1024) hot_set_kernel(const uint32_t* cnt, uint32_t S, uint32_t W, int staged,
                                                       uint32_t* perm, uint32_t* out) {
    extern __shared__ uint16_t lc[];
    __shared__ uint32_t csum[1024];
    __shared__ uint32_t ws[kMaxWaves + 1];
    __shared__ uint32_t n_at, top, total, cold_sum, cold_max;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        top = 0;
        total = 0;
        cold_sum = 0;
        cold_max = 0;
    }
    __syncthreads();
    uint32_t m = 0, sum = 0;
    for (uint32_t i = t; i < S; i += blockDim.x) {  // coalesced
        const uint32_t v = cnt[i];
        m = max(m, v);
        sum += v;
        if (staged) lc[i] = static_cast<uint16_t>(min(v, 65535u));
    }
    atomicMax(&top, m);
    atomicAdd(&total, sum);
    __syncthreads();
    // Counts capped at 65535 keep every comparison against c <= 65535 exact.
    auto val = [&](uint32_t i) -> uint32_t { return staged ? lc[i] : min(cnt[i], 65535u); };
    const uint32_t C = (S + blockDim.x - 1) / blockDim.x;
    const uint32_t b = min(S, t * C), e = min(S, b + C);
    // Smallest c in [1, top + 1] with #{cnt >= c} <= W (c = top + 1 leaves none).
    // #{cnt >= c} <= total / c, so c <= 65535 whenever W >= 2^17 / 65535.
    uint32_t lo = 1, hi = min(top, 65535u) + 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (t == 0) n_at = 0;
        __syncthreads();
        uint32_t k = 0;
        for (uint32_t i = b; i < e; ++i) k += val(i) >= mid ? 1u : 0u;
        atomicAdd(&n_at, k);
        __syncthreads();
        if (n_at <= W)
            hi = mid;
        else
            lo = mid + 1;
        __syncthreads();
    }
    const uint32_t c = lo;
    uint32_t k = 0, cs = 0, cm = 0;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = val(i);
        if (v >= c) {
            ++k;
        } else {  // below c, so exact
            cs += v;
            cm = max(cm, v);
        }
    }
    csum[t] = k;
    atomicAdd(&cold_sum, cs);
    atomicMax(&cold_max, cm);
    __syncthreads();  // the scan reads other threads' csum entries
    const uint32_t h = block_exclusive_scan(csum, blockDim.x, ws);
    uint32_t hp = csum[t];  // hot slots before slot b
    for (uint32_t i = b; i < e; ++i) {
        if (val(i) >= c) {
            perm[i] = hp++;
        } else {
            perm[i] = h + (i - hp);
        }
    }
    __syncthreads();
    if (t == 0) {
        out[0] = h;
        out[1] = total - cold_sum;
        out[2] = total;
        out[3] = cold_max;
    }
}

// gout[i] = perm[gin[i]]; the empty entry stays empty.
template <typename MapT>
__global__ void remap_slots_kernel(const MapT* __restrict__ gin, uint64_t n, const uint32_t* __restrict__ perm,
                                   MapT* __restrict__ gout) {
    constexpr MapT kEmpty = static_cast<MapT>(~MapT(0));
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const MapT g = gin[i];
        gout[i] = g == kEmpty ? kEmpty : static_cast<MapT>(perm[g]);
    }
}

// st[f * S + s] = src[f * S + perm[s]] for each of the state's `fields` fields.
__global__ void unpermute_state_kernel(const uint64_t* __restrict__ src, uint64_t S, uint32_t fields,
                                       const uint32_t* __restrict__ perm, uint64_t* __restrict__ st) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < S * fields;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t f = i / S, sl = i - f * S;
        st[i] = src[f * S + perm[sl]];
    }
}

__global__ void __launch_bounds__(1024) join_agg_lists_lds(ProbeAggLists L, uint64_t n, uint32_t S, AggProg p,
                                                           uint64_t* part) {
    extern __shared__ __align__(16) char smem[];
    lds_init(smem, S, p);
    __syncthreads();
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t b = 0;
        const uint32_t c = L.v.probe(r, b);
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        for (uint32_t i = 0; i < c; ++i) lds_update(smem, S, p, L.gslot[L.v.row_at(b + i)], vrow);
    }
    __syncthreads();
    lds_flush(smem, S, p, part);
}

__global__ void join_agg_lists_global(ProbeAggLists L, uint64_t n, uint64_t S, AggProg p, uint64_t* st) {
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t b = 0;
        const uint32_t c = L.v.probe(r, b);
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        for (uint32_t i = 0; i < c; ++i) global_update(st, S, p, L.gslot[L.v.row_at(b + i)], vrow);
    }
}

// Group slot of each build row over its group-key columns.
__global__ void gslot_direct_kernel(ColView g, uint64_t n, int64_t gmin, uint32_t null_slot, uint32_t* gslot) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x)
        gslot[r] = is_valid(g, r) ? static_cast<uint32_t>(static_cast<uint64_t>(int_value(g, r) - gmin)) : null_slot;
}
__global__ void gslot_hash_kernel(HashKeys hk, uint64_t n, HashTableView t, uint32_t* gslot) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);  // kEqual: NULL tuples group together
        gslot[r] = ht_insert(t, hk.nkeys, w, nm, h);
    }
}
template <typename MapT>
__global__ void gmap_fill_kernel(ColView key, uint64_t n, int64_t kmin, const uint32_t* gslot, MapT* gmap) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        if (!is_valid(key, r)) continue;
        gmap[static_cast<uint64_t>(int_value(key, r) - kmin)] = static_cast<MapT>(gslot[r]);
    }
}

// ---- direct unique map, built by key-range partitions ----------------------------
// gmap[key - kmin] = group slot of the build row holding that key. Build rows
// are first partitioned by the high bits of their key offset (2^pb keys per
// partition), then one block per partition assembles its slice of the map in
// LDS, where a bitmap flags a repeated key, and writes the slice out with
// contiguous stores: no scattered 2-byte writes to HBM, no separate clear.
constexpr uint32_t kMapTile = 16384;     // build rows per partitioning block (1024 threads)
constexpr uint32_t kMaxMapParts = 4096;  // partitions (LDS cursors per block)

constexpr int kMapRpt = kMapTile / 1024;  // build rows per thread

// Group slot of build row r for the map: from the computed array, or straight
// from a dense integer group column (value - gmin, NULL -> null_slot), both as
// unconditional loads.
struct SlotArray {
    const uint32_t* s;
    __device__ uint32_t operator()(uint64_t r) const { return gcast<uint32_t>(s)[r]; }
};
template <int GW>
struct SlotDirect {
    const void* g;
    const uint64_t* gvalid;
    int64_t gmin;
    uint32_t null_slot;
    __device__ uint32_t operator()(uint64_t r) const {
        const uint64_t v = elem_bits<GW>(g, r);
        return bit_at(bits_view(gvalid), r) ? static_cast<uint32_t>(v - static_cast<uint64_t>(gmin)) : null_slot;
    }
};

// Key offsets (key - kmin) of the tile's rows k0 .. k0 + N - 1 of this thread
// (row k * 1024 + thread), ok = 0 for NULL keys and rows past the end; all
// loads issued before any is used.
template <int KW, int N>
__device__ inline void map_tile_keys(const ColView& key, uint64_t t0, uint32_t m, int64_t kmin, int k0,
                                     uint64_t (&off)[N], uint32_t (&ok)[N]) {
    const BitsView kv = bits_view(key.validity);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const uint32_t i = (k0 + k) * 1024 + threadIdx.x;
        const uint64_t r = t0 + (i < m ? i : m - 1);
        off[k] = elem_bits<KW>(key.data, r) - static_cast<uint64_t>(kmin);
        ok[k] = bit_at(kv, r) & (i < m ? 1u : 0u);
    }
}

template <int KW>
__global__ void __launch_bounds__(1024) map_part_hist(ColView key, uint64_t n, int64_t kmin, uint32_t pb,
                                                     uint32_t nparts, uint32_t ntiles, uint32_t* hist) {
    extern __shared__ uint32_t hcnt[];
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) hcnt[i] = 0;
    __syncthreads();
    const uint64_t t0 = uint64_t(blockIdx.x) * kMapTile;
    const uint32_t m = static_cast<uint32_t>(min<uint64_t>(kMapTile, n - t0));
    uint64_t off[kMapRpt];
    uint32_t ok[kMapRpt];
    map_tile_keys<KW, kMapRpt>(key, t0, m, kmin, 0, off, ok);
#pragma unroll
    for (int k = 0; k < kMapRpt; ++k)
        if (ok[k]) atomicAdd(&hcnt[off[k] >> pb], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) hist[size_t(i) * ntiles + blockIdx.x] = hcnt[i];
}

__global__ void __launch_bounds__(1024) map_part_scatter(ColView key, uint64_t n, int64_t kmin, uint32_t pb,
                                                        uint32_t nparts, uint32_t ntiles, const uint64_t* offs,
                                                        const uint32_t* gslot, uint64_t* ent) {
    extern __shared__ __align__(8) char msm[];
    uint64_t* base = reinterpret_cast<uint64_t*>(msm);
    uint32_t* cur = reinterpret_cast<uint32_t*>(base + nparts);
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        base[i] = offs[size_t(i) * ntiles + blockIdx.x];
        cur[i] = 0;
    }
    __syncthreads();
    const uint64_t mask = (uint64_t(1) << pb) - 1;
    const uint64_t t0 = uint64_t(blockIdx.x) * kMapTile, t1 = min(n, t0 + kMapTile);
    for (uint64_t r = t0 + threadIdx.x; r < t1; r += blockDim.x) {
        if (!is_valid(key, r)) continue;
        const uint64_t off = static_cast<uint64_t>(int_value(key, r) - kmin);
        const uint32_t part = static_cast<uint32_t>(off >> pb);
        const uint32_t l = atomicAdd(&cur[part], 1u);
        ent[base[part] + l] = ((off & mask) << 32) | gslot[r];
    }
}

// Scatter for 16-bit maps (2^pb <= 2^16 keys per partition): the tile is sorted
// by partition in LDS first, then each partition's run is written contiguously.
// Entry = (key offset within the partition) << 16 | group slot.
// Two layouts of the entries:
//  * offs != nullptr: partition-major runs at the offsets of a scan of the
//    per-(partition, tile) counts (any key distribution);
//  * gcur != nullptr: a fixed bucket of 2^pb entries per partition (as many
//    as the partition has distinct keys), each block claiming its runs with
//    one atomic per partition on gcur; no count pass and no scan. A bucket
//    that would overflow can only hold repeated keys: the excess is dropped
//    and the build kernel reports duplicates from gcur.
template <int KW, class Slots>
__global__ void __launch_bounds__(1024) map_part_scatter16(ColView key, uint64_t n, int64_t kmin, uint32_t pb,
                                                          uint32_t nparts, uint32_t ntiles, const uint64_t* offs,
                                                          uint32_t* gcur, Slots slots, uint32_t* ent) {
    extern __shared__ __align__(16) char msm[];
    uint64_t* gbase = reinterpret_cast<uint64_t*>(msm);                 // [nparts]
    uint32_t* stage = reinterpret_cast<uint32_t*>(gbase + nparts);      // [kMapTile]
    uint16_t* spart = reinterpret_cast<uint16_t*>(stage + kMapTile);    // [kMapTile]
    uint32_t* cnt = reinterpret_cast<uint32_t*>(spart + kMapTile);      // [nparts] counts, then local bases
    uint32_t* ws = cnt + nparts;                                        // [kMaxWaves + 1]
    const bool fixed = gcur != nullptr;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        gbase[i] = fixed ? 0 : offs[size_t(i) * ntiles + blockIdx.x];
        cnt[i] = 0;
    }
    __syncthreads();
    constexpr int kRpt = kMapRpt;
    const uint64_t t0 = uint64_t(blockIdx.x) * kMapTile;
    const uint32_t m = static_cast<uint32_t>(min<uint64_t>(kMapTile, n - t0));
    uint32_t part[kRpt], pos[kRpt], ev[kRpt];
    // Two halves of eight rows: each half's loads in flight together, within
    // the 128 registers of a 1024-thread block.
    constexpr int kHalf = kRpt / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        uint64_t off[kHalf];
        uint32_t ok[kHalf], sl[kHalf];
        map_tile_keys<KW, kHalf>(key, t0, m, kmin, h * kHalf, off, ok);
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
            const uint32_t i = (h * kHalf + k) * 1024 + threadIdx.x;
            sl[k] = slots(t0 + (i < m ? i : m - 1));
        }
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
            const int kk = h * kHalf + k;
            part[kk] = 0xFFFFFFFFu;
            if (ok[k]) {
                part[kk] = static_cast<uint32_t>(off[k] >> pb);
                pos[kk] = atomicAdd(&cnt[part[kk]], 1u);
                ev[kk] = static_cast<uint32_t>((off[k] & ((uint64_t(1) << pb) - 1)) << 16) | (sl[k] & 0xFFFFu);
            }
        }
    }
    __syncthreads();
    if (fixed) {  // claim this tile's run in each partition's bucket
        for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
            const uint32_t c = cnt[i];
            if (c) gbase[i] = atomicAdd(gcur + i, c);
        }
    }
    const uint32_t total = block_exclusive_scan(cnt, nparts, ws);
#pragma unroll
    for (int k = 0; k < kRpt; ++k) {
        if (part[k] == 0xFFFFFFFFu) continue;
        const uint32_t at = cnt[part[k]] + pos[k];
        stage[at] = ev[k];
        spart[at] = static_cast<uint16_t>(part[k]);
    }
    __syncthreads();
    if (fixed) {
        const uint64_t cap = uint64_t(1) << pb;
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const uint32_t pp = spart[i];
            const uint64_t within = gbase[pp] + (i - cnt[pp]);
            if (within < cap) ent[(uint64_t(pp) << pb) + within] = stage[i];
        }
    } else {
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const uint32_t pp = spart[i];
            ent[gbase[pp] + (i - cnt[pp])] = stage[i];
        }
    }
}

// One block per partition: its entries (a run of the scan layout, or its
// fixed bucket when gcur != nullptr) land in an LDS copy of the partition's
// map range, which is then written out whole. A key seen twice (or a bucket
// that overflowed) sets *dup.
template <typename MapT, typename EntT>
__global__ void __launch_bounds__(1024) map_part_build(const EntT* ent, const uint64_t* offs, const uint32_t* gcur,
                                                       uint32_t ntiles, uint32_t nparts, uint64_t total, uint32_t pb,
                                                       uint64_t range, MapT* gmap, unsigned int* dup) {
    extern __shared__ __align__(16) char msm[];
    const uint32_t P = 1u << pb;
    MapT* m = reinterpret_cast<MapT*>(msm);
    uint32_t* bm = reinterpret_cast<uint32_t*>(msm + size_t(P) * sizeof(MapT));
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) m[i] = static_cast<MapT>(~MapT(0));
    for (uint32_t i = threadIdx.x; i < P / 32; i += blockDim.x) bm[i] = 0;
    __syncthreads();
    const uint32_t part = blockIdx.x;
    uint64_t b, e;
    bool d = false;
    if (gcur) {
        const uint64_t c = gcur[part];
        b = uint64_t(part) << pb;
        e = b + min<uint64_t>(c, P);
        d = c > P;
    } else {
        b = offs[size_t(part) * ntiles];
        e = (part + 1 < nparts) ? offs[size_t(part + 1) * ntiles] : total;
    }
    // Eight entries per thread in flight (unconditional loads at a clamped
    // index through a global pointer), then their LDS updates.
    constexpr int kShift = sizeof(EntT) == 4 ? 16 : 32;  // offset | slot packing of the scatter
    constexpr int U = 8;
    const gptr<EntT> eg = gcast<EntT>(ent);
    for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += uint64_t(U) * blockDim.x) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + uint64_t(u) * blockDim.x;
            v[u] = eg[i < e ? i : e - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + uint64_t(u) * blockDim.x >= e) break;
            const uint32_t o = static_cast<uint32_t>(v[u] >> kShift);
            const uint32_t bit = 1u << (o & 31);
            if (atomicOr(&bm[o >> 5], bit) & bit) d = true;
            m[o] = static_cast<MapT>(static_cast<uint32_t>(v[u] & ((uint64_t(1) << kShift) - 1)));
        }
    }
    if (d) atomicOr(dup, 1u);
    __syncthreads();
    const uint64_t lo = uint64_t(part) << pb;
    const uint64_t cnt = min<uint64_t>(P, range - lo);
    for (uint64_t i = threadIdx.x; i < cnt; i += blockDim.x) gmap[lo + i] = m[i];
}

enum class MapBuild { kOk, kDuplicate, kTooWide };

// Where the map's group slots come from: the per-row array, or (16-bit maps)
// a dense integer group column read by the scatter itself.
struct MapSlots {
    const uint32_t* gslot = nullptr;
    const qe_dev_col* g = nullptr;
    int64_t gmin = 0;
    uint32_t null_slot = 0;
};

template <int KW>
static void launch_scatter16(qe_gpu_ctx* ctx, uint32_t ntiles, size_t lds, const ColView& kv, uint64_t nb,
                             int64_t kmin, uint32_t pb, uint32_t nparts, const uint64_t* offs, uint32_t* gcur,
                             const MapSlots& ms, uint32_t* ent) {
    if (ms.gslot) {
        map_part_scatter16<KW, SlotArray><<<ntiles, 1024, lds, ctx->stream>>>(kv, nb, kmin, pb, nparts, ntiles, offs,
                                                                             gcur, SlotArray{ms.gslot}, ent);
        return;
    }
    const uint64_t* gv = ms.g->all_valid ? nullptr : ms.g->validity;
    switch (type_width(ms.g->type)) {
        case 1:
            map_part_scatter16<KW, SlotDirect<1>><<<ntiles, 1024, lds, ctx->stream>>>(
                kv, nb, kmin, pb, nparts, ntiles, offs, gcur, SlotDirect<1>{ms.g->data, gv, ms.gmin, ms.null_slot},
                ent);
            break;
        case 4:
            map_part_scatter16<KW, SlotDirect<4>><<<ntiles, 1024, lds, ctx->stream>>>(
                kv, nb, kmin, pb, nparts, ntiles, offs, gcur, SlotDirect<4>{ms.g->data, gv, ms.gmin, ms.null_slot},
                ent);
            break;
        default:
            map_part_scatter16<KW, SlotDirect<8>><<<ntiles, 1024, lds, ctx->stream>>>(
                kv, nb, kmin, pb, nparts, ntiles, offs, gcur, SlotDirect<8>{ms.g->data, gv, ms.gmin, ms.null_slot},
                ent);
    }
}

// `total` is the number of valid build keys (all within [kmin, kmin + range)).
// The duplicate-key flag is left in `dup` (device) for the caller to read once
// the probe is queued, so no host round trip sits between build and probe.
// 32-bit maps take the slots from the array (ms.gslot).
template <typename MapT>
MapBuild build_direct_map(qe_gpu_ctx* ctx, const qe_dev_col* key, uint64_t nb, uint64_t total, int64_t kmin,
                          uint64_t range, const MapSlots& ms, Scratch& gmap, Scratch& dup) {
    // 2^15 keys per partition (64 KiB of 16-bit map: two building blocks per
    // CU; 128 KiB of 32-bit map); 16-bit maps of very wide ranges take 2^16.
    const uint32_t pb = (sizeof(MapT) == 2 && range > (uint64_t(kMaxMapParts) << 15)) ? 16 : 15;
    const uint64_t np64 = (range + (uint64_t(1) << pb) - 1) >> pb;
    if (np64 == 0 || np64 > kMaxMapParts) return MapBuild::kTooWide;
    const uint32_t nparts = static_cast<uint32_t>(np64);
    const uint32_t ntiles = static_cast<uint32_t>((nb + kMapTile - 1) / kMapTile);
    const uint64_t nh = uint64_t(nparts) * ntiles;
    const ColView kv = view(key);
    const int kw = static_cast<int>(type_width(key->type));
    constexpr bool k16 = sizeof(MapT) == 2;  // 4-byte entries, LDS-sorted scatter
    const size_t build_lds = (size_t(1) << pb) * sizeof(MapT) + ((size_t(1) << pb) / 32) * 4;
    if (k16 && range <= 2 * total) {
        // Dense keys: fixed buckets of 2^pb entries per partition (at most 2x
        // the entries of a scanned layout), so no count pass and no scan. The
        // bucket cursors follow the duplicate flag in its allocation: one
        // clear for both.
        dup = Scratch(ctx, (size_t(nparts) + 1) * 4);
        uint32_t* gcur = dup.as<uint32_t>() + 1;
        Scratch ent(ctx, (size_t(nparts) << pb) * 4);
        memset_async(ctx, dup.p, 0, (size_t(nparts) + 1) * 4);
        const size_t lds = size_t(nparts) * 12 + size_t(kMapTile) * 6 + (kMaxWaves + 1) * 4;
        if (kw == 8)
            launch_scatter16<8>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, nullptr, gcur, ms, ent.as<uint32_t>());
        else if (kw == 4)
            launch_scatter16<4>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, nullptr, gcur, ms, ent.as<uint32_t>());
        else
            launch_scatter16<1>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, nullptr, gcur, ms, ent.as<uint32_t>());
        QE_HIP(hipGetLastError());
        gmap = Scratch(ctx, range * sizeof(MapT));
        map_part_build<MapT, uint32_t><<<nparts, 1024, build_lds, ctx->stream>>>(
            ent.as<uint32_t>(), nullptr, gcur, ntiles, nparts, total, pb, range, gmap.as<MapT>(),
            dup.as<unsigned int>());
        QE_HIP(hipGetLastError());
        return MapBuild::kOk;
    }
    dup = Scratch(ctx, 4);
    Scratch hist(ctx, nh * 4), offs(ctx, nh * 8), sums(ctx, scan_sums_words(nh) * 8);
    if (kw == 8)
        map_part_hist<8><<<ntiles, 1024, nparts * 4, ctx->stream>>>(kv, nb, kmin, pb, nparts, ntiles,
                                                                    hist.as<uint32_t>());
    else if (kw == 4)
        map_part_hist<4><<<ntiles, 1024, nparts * 4, ctx->stream>>>(kv, nb, kmin, pb, nparts, ntiles,
                                                                    hist.as<uint32_t>());
    else
        map_part_hist<1><<<ntiles, 1024, nparts * 4, ctx->stream>>>(kv, nb, kmin, pb, nparts, ntiles,
                                                                    hist.as<uint32_t>());
    QE_HIP(hipGetLastError());
    exclusive_scan_u32_dev(ctx, hist.as<uint32_t>(), offs.as<uint64_t>(), nh, sums.as<uint64_t>());
    Scratch ent(ctx, std::max<uint64_t>(total, 1) * (k16 ? 4 : 8));
    if (k16) {
        const size_t lds = size_t(nparts) * 12 + size_t(kMapTile) * 6 + (kMaxWaves + 1) * 4;
        if (kw == 8)
            launch_scatter16<8>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, offs.as<uint64_t>(), nullptr, ms,
                                 ent.as<uint32_t>());
        else if (kw == 4)
            launch_scatter16<4>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, offs.as<uint64_t>(), nullptr, ms,
                                 ent.as<uint32_t>());
        else
            launch_scatter16<1>(ctx, ntiles, lds, kv, nb, kmin, pb, nparts, offs.as<uint64_t>(), nullptr, ms,
                                 ent.as<uint32_t>());
    } else {
        invalid(ms.gslot != nullptr, "32-bit map build needs the slot array");
        map_part_scatter<<<ntiles, 1024, nparts * 12, ctx->stream>>>(kv, nb, kmin, pb, nparts, ntiles,
                                                                     offs.as<uint64_t>(), ms.gslot,
                                                                     ent.as<uint64_t>());
    }
    QE_HIP(hipGetLastError());
    gmap = Scratch(ctx, range * sizeof(MapT));
    memset_async(ctx, dup.p, 0, 4);
    if (k16)
        map_part_build<MapT, uint32_t><<<nparts, 1024, build_lds, ctx->stream>>>(
            ent.as<uint32_t>(), offs.as<uint64_t>(), nullptr, ntiles, nparts, total, pb, range, gmap.as<MapT>(),
            dup.as<unsigned int>());
    else
        map_part_build<MapT, uint64_t><<<nparts, 1024, build_lds, ctx->stream>>>(
            ent.as<uint64_t>(), offs.as<uint64_t>(), nullptr, ntiles, nparts, total, pb, range, gmap.as<MapT>(),
            dup.as<unsigned int>());
    QE_HIP(hipGetLastError());
    return MapBuild::kOk;
}

// Value columns of a program for the streaming probe: I32/I64/TS/F64 inputs,
// at most kStreamCols distinct columns, every column (and the key column)
// aligned for its pair loads.
bool stream_cols(const AggProg& p, const void* keys, size_t key_width, StreamCols& sc, int& nv) {
    auto aligned = [](const void* d, size_t w) { return (reinterpret_cast<uintptr_t>(d) % (2 * w)) == 0; };
    if (!aligned(keys, key_width)) return false;
    nv = 0;
    for (int a = 0; a < p.n; ++a) {
        sc.vcol[a] = -1;
        if (p.kind[a] == K_COUNT_STAR) continue;
        const ColView& v = p.in[a];
        const bool wide = v.type == QE_I64 || v.type == QE_TS || v.type == QE_F64;
        if (!wide && v.type != QE_I32) return false;
        int c = 0;
        while (c < nv && sc.data[c] != v.data) ++c;
        if (c == nv) {
            if (nv == kStreamCols || !aligned(v.data, wide ? 8 : 4)) return false;
            sc.data[c] = v.data;
            sc.validity[c] = v.validity;
            sc.wide[c] = wide ? 1 : 0;
            ++nv;
        }
        sc.vcol[a] = c;
    }
    return true;
}

// Sampled hot set for a slot space larger than LDS: at most W slots relabelled
// first by perm (caller's buffer of S entries), the sampled rows in the set and
// in all, and the heaviest slot outside it.
struct HotSet {
    uint32_t h = 0, inside = 0, total = 0, cold_max = 0;
    uint64_t stride = 1;  // probe rows per sample
};

template <typename KeyT, typename MapT>
HotSet hot_set(qe_gpu_ctx* ctx, const KeyT* keys, const uint64_t* kv, uint64_t n, const MapT* gm, int64_t kmin,
               uint64_t range, uint32_t S, uint32_t W, uint32_t* perm) {
    TimedRegion tr(ctx, "join_agg_hot_set");
    Scratch buf(ctx, (size_t(S) + 4) * 4);  // cnt[S] | out[4]
    uint32_t* cnt = buf.as<uint32_t>();
    memset_async(ctx, cnt, 0, size_t(S) * 4);
    const uint64_t stride = std::max<uint64_t>(1, n >> 16);
    const uint64_t ns = (n + stride - 1) / stride;
    slot_sample_kernel<KeyT, MapT><<<grid_for(ns, 256, 1024), 256, 0, ctx->stream>>>(keys, kv, n, stride, gm, kmin,
                                                                                     range, cnt);
    QE_HIP(hipGetLastError());
    const int staged = S <= kHotStagedSlots ? 1 : 0;
    hot_set_kernel<<<1, 1024, staged ? size_t(S) * 2 : 0, ctx->stream>>>(cnt, S, W, staged, perm, cnt + S);
    QE_HIP(hipGetLastError());
    uint32_t* h = static_cast<uint32_t*>(pinned(ctx, 16));
    QE_HIP(hipMemcpyAsync(h, cnt + S, 16, hipMemcpyDeviceToHost, ctx->stream));
    QE_HIP(hipStreamSynchronize(ctx->stream));
    HotSet hs;
    hs.h = h[0];
    hs.inside = h[1];
    hs.total = h[2];
    hs.cold_max = h[3];
    hs.stride = stride;
    return hs;
}

void launch_spill_agg(qe_gpu_ctx* ctx, int nv, unsigned blocks, size_t lb, const uint64_t* rec,
                      const uint32_t* count, uint32_t cap, uint32_t slot_lo, uint32_t S, const AggProg& p,
                      const StreamCols& sc, uint64_t* part) {
    switch (nv) {
        case 0:
            spill_agg_kernel<0><<<blocks, 1024, lb, ctx->stream>>>(rec, count, cap, slot_lo, S, p, sc, part);
            break;
        case 1:
            spill_agg_kernel<1><<<blocks, 1024, lb, ctx->stream>>>(rec, count, cap, slot_lo, S, p, sc, part);
            break;
        default:
            spill_agg_kernel<2><<<blocks, 1024, lb, ctx->stream>>>(rec, count, cap, slot_lo, S, p, sc, part);
            break;
    }
    QE_HIP(hipGetLastError());
}

template <typename KeyT, typename MapT, int PAIRS, int NV, int WMASK>
void launch_stream_w(qe_gpu_ctx* ctx, unsigned blocks, size_t lb, const KeyT* keys, const uint64_t* kv, uint64_t n,
                     const MapT* gm, int64_t kmin, uint64_t range, uint32_t S, uint32_t slot_lo, const AggProg& p,
                     const StreamCols& sc, uint64_t* part, const Spill& spill) {
    FastFold ff;
    const bool fast = make_fast_fold(p, sc.vcol, ff);
    if (spill.st) {
        if (fast)
            join_agg_stream<KeyT, MapT, PAIRS, NV, WMASK, true, true><<<blocks, 1024, lb, ctx->stream>>>(
                keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part, spill, ff);
        else
            join_agg_stream<KeyT, MapT, PAIRS, NV, WMASK, true, false><<<blocks, 1024, lb, ctx->stream>>>(
                keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part, spill, ff);
    } else {
        if (fast)
            join_agg_stream<KeyT, MapT, PAIRS, NV, WMASK, false, true><<<blocks, 1024, lb, ctx->stream>>>(
                keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part, spill, ff);
        else
            join_agg_stream<KeyT, MapT, PAIRS, NV, WMASK, false, false><<<blocks, 1024, lb, ctx->stream>>>(
                keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part, spill, ff);
    }
    QE_HIP(hipGetLastError());
}

// Value widths as a template mask (bit c: column c is 8 bytes): no width
// branch around the stream loads.
template <typename KeyT, typename MapT, int PAIRS, int NV>
void launch_stream_nv(qe_gpu_ctx* ctx, unsigned blocks, size_t lb, const KeyT* keys, const uint64_t* kv, uint64_t n,
                      const MapT* gm, int64_t kmin, uint64_t range, uint32_t S, uint32_t slot_lo, const AggProg& p,
                      const StreamCols& sc, uint64_t* part, const Spill& spill) {
    int wmask = 0;
    for (int c = 0; c < NV; ++c) wmask |= (sc.wide[c] ? 1 : 0) << c;
    if constexpr (NV == 0) {
        launch_stream_w<KeyT, MapT, PAIRS, NV, 0>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc,
                                                  part, spill);
    } else if constexpr (NV == 1) {
        if (wmask)
            launch_stream_w<KeyT, MapT, PAIRS, NV, 1>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p,
                                                      sc, part, spill);
        else
            launch_stream_w<KeyT, MapT, PAIRS, NV, 0>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p,
                                                      sc, part, spill);
    } else {
        switch (wmask) {
            case 0:
                launch_stream_w<KeyT, MapT, PAIRS, NV, 0>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo,
                                                          p, sc, part, spill);
                break;
            case 1:
                launch_stream_w<KeyT, MapT, PAIRS, NV, 1>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo,
                                                          p, sc, part, spill);
                break;
            case 2:
                launch_stream_w<KeyT, MapT, PAIRS, NV, 2>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo,
                                                          p, sc, part, spill);
                break;
            default:
                launch_stream_w<KeyT, MapT, PAIRS, NV, 3>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo,
                                                          p, sc, part, spill);
                break;
        }
    }
}

template <typename KeyT, typename MapT, int PAIRS>
void launch_stream_p(qe_gpu_ctx* ctx, int nv, unsigned blocks, size_t lb, const KeyT* keys, const uint64_t* kv,
                     uint64_t n, const MapT* gm, int64_t kmin, uint64_t range, uint32_t S, uint32_t slot_lo,
                     const AggProg& p, const StreamCols& sc, uint64_t* part, const Spill& spill) {
    switch (nv) {
        case 0:
            launch_stream_nv<KeyT, MapT, PAIRS, 0>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc,
                                                   part, spill);
            break;
        case 1:
            launch_stream_nv<KeyT, MapT, PAIRS, 1>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc,
                                                   part, spill);
            break;
        default:
            launch_stream_nv<KeyT, MapT, PAIRS, 2>(ctx, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc,
                                                   part, spill);
            break;
    }
}

// Row pairs per thread per tile: 2 or 4 (tuning knob QE_TUNE_JOIN_PAIRS).
template <typename KeyT, typename MapT>
void launch_stream(qe_gpu_ctx* ctx, int pairs, int nv, unsigned blocks, size_t lb, const KeyT* keys,
                   const uint64_t* kv, uint64_t n, const MapT* gm, int64_t kmin, uint64_t range, uint32_t S,
                   uint32_t slot_lo, const AggProg& p, const StreamCols& sc, uint64_t* part,
                   const Spill& spill = Spill{}) {
    if (pairs <= 2)
        launch_stream_p<KeyT, MapT, 2>(ctx, nv, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part,
                                       spill);
    else
        launch_stream_p<KeyT, MapT, 4>(ctx, nv, blocks, lb, keys, kv, n, gm, kmin, range, S, slot_lo, p, sc, part,
                                       spill);
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_hash_join(qe_gpu_ctx* ctx, const qe_dev_col* const* probe_keys,
                               const qe_dev_col* const* build_keys, uint32_t nkeys, int32_t join_type,
                               qe_dev_pairs** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && probe_keys != nullptr && build_keys != nullptr, "null argument");
        check_key_types(probe_keys, build_keys, nkeys);
        OwnedCols own(ctx);  // STR keys compare by value (strings.hip)
        std::vector<const qe_dev_col*> pk, bk;
        str_join_keys(ctx, probe_keys, build_keys, nkeys, pk, bk, own);
        *out = join_pairs(ctx, pk.data(), bk.data(), nkeys, join_type);
    });
}

qe_gpu_status qe_gpu_pairs_gather(qe_gpu_ctx* ctx, const qe_dev_pairs* p, uint64_t offset, uint64_t count,
                                  const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                  const qe_dev_col* const* build_cols, uint32_t nbuild, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(p != nullptr && out != nullptr, "null argument");
        invalid(offset + count <= p->count, "pair range out of bounds");
        common_len(probe_cols, nprobe);
        common_len(build_cols, nbuild);
        qe_dev_table* t = new_table(count);
        try {
            std::vector<const qe_dev_col*> src;
            std::vector<const uint32_t*> idx;
            for (uint32_t c = 0; c < nprobe; ++c) {
                src.push_back(probe_cols[c]);
                idx.push_back(p->probe + offset);
            }
            for (uint32_t c = 0; c < nbuild; ++c) {
                src.push_back(build_cols[c]);
                idx.push_back(p->build + offset);
            }
            t->cols = gather_cols(ctx, src.data(), idx.data(), src.size(), count);
        } catch (...) {
            qe_gpu_table_free(ctx, t);
            throw;
        }
        *out = t;
    });
}

}  // extern "C"

namespace qe_gpu {

// Most LDS slot-range passes over the probe before HBM atomics are used.
constexpr uint64_t kMaxSlotPasses = 32;
// Fewest probe rows for which a sampled hot set may replace the passes, and the
// largest spill-list allocation it may take.
constexpr uint64_t kSpillMinRows = uint64_t(1) << 22;
constexpr size_t kSpillMaxBytes = size_t(8) << 30;

// The fused join -> group-by over prepared columns (no STR rewriting).
qe_gpu_status join_aggregate_codes(qe_gpu_ctx* ctx, const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                   const qe_dev_col* const* build_cols, uint32_t nbuild, const uint32_t* probe_keys,
                                   const uint32_t* build_keys, uint32_t nkeys, int32_t join_type,
                                   const uint32_t* group_keys, uint32_t ngk, const qe_agg_spec* aggs, uint32_t naggs,
                                   qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        invalid(nprobe >= 1 && nbuild >= 1, "join inputs need >=1 column each");
        invalid(nkeys >= 1, "equi-join needs >=1 key");
        invalid(join_type == QE_JOIN_INNER || join_type == QE_JOIN_LEFT, "unknown join type");
        check(nkeys <= kMaxKeys && ngk <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many keys for the device path");
        const uint64_t np = common_len(probe_cols, nprobe), nb = common_len(build_cols, nbuild);
        check(np < kNoRow && nb < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
        const uint32_t ncols = nprobe + nbuild;
        std::vector<const qe_dev_col*> pk(nkeys), bk(nkeys);
        for (uint32_t j = 0; j < nkeys; ++j) {
            invalid(probe_keys[j] < nprobe && build_keys[j] < nbuild, "join key column out of range");
            pk[j] = probe_cols[probe_keys[j]];
            bk[j] = build_cols[build_keys[j]];
        }
        check_key_types(pk.data(), bk.data(), nkeys);
        auto col_at = [&](uint32_t c) { return c < nprobe ? probe_cols[c] : build_cols[c - nprobe]; };
        for (uint32_t k = 0; k < ngk; ++k) invalid(group_keys[k] < ncols, "group key column out of range");
        bool fast = join_type == QE_JOIN_INNER && ngk >= 1;
        for (uint32_t k = 0; k < ngk; ++k) fast = fast && group_keys[k] >= nprobe;
        ColView ins[kMaxAggs];
        int types[kMaxAggs];
        for (uint32_t a = 0; a < naggs && a < uint32_t(kMaxAggs); ++a) {
            if (aggs[a].func == QE_AGG_COUNT_STAR) {
                ins[a] = ColView{nullptr, nullptr, QE_I64, 0};
                types[a] = QE_I64;
                continue;
            }
            invalid(aggs[a].input_col < ncols, "aggregate input column out of range");
            fast = fast && aggs[a].input_col < nprobe;
            ins[a] = view(col_at(aggs[a].input_col));
            types[a] = col_at(aggs[a].input_col)->type;
        }
        const AggProg p0 = make_prog(aggs, naggs, ins, types);

        if (!fast) {
            // General composition: pairs -> gather referenced columns -> aggregate.
            qe_dev_pairs* pr = join_pairs(ctx, pk.data(), bk.data(), nkeys, join_type);
            std::vector<qe_dev_col*> gathered(ncols, nullptr);
            std::vector<qe_dev_col*> owned;
            qe_dev_table* res = nullptr;
            try {
                std::vector<char> need(ncols, 0);
                for (uint32_t k = 0; k < ngk; ++k) need[group_keys[k]] = 1;
                for (uint32_t a = 0; a < naggs; ++a)
                    if (aggs[a].func != QE_AGG_COUNT_STAR) need[aggs[a].input_col] = 1;
                if (ngk == 0 && naggs > 0) need[0] = 1;  // keep the row count
                std::vector<const qe_dev_col*> src;
                std::vector<const uint32_t*> idx;
                std::vector<uint32_t> at;
                for (uint32_t c = 0; c < ncols; ++c) {
                    if (!need[c]) continue;
                    src.push_back(col_at(c));
                    idx.push_back(c < nprobe ? pr->probe : pr->build);
                    at.push_back(c);
                }
                owned = gather_cols(ctx, src.data(), idx.data(), src.size(), pr->count);
                for (size_t k = 0; k < at.size(); ++k) gathered[at[k]] = owned[k];
                // Unreferenced positions get a referenced column of the right
                // length so indices stay those of the join output.
                qe_dev_col* any = owned.empty() ? nullptr : owned[0];
                for (uint32_t c = 0; c < ncols; ++c)
                    if (!gathered[c]) gathered[c] = any;
                const qe_gpu_status st = aggregate_codes(ctx, gathered.data(), ncols, group_keys, ngk, aggs, naggs, &res);
                if (st != QE_GPU_OK) throw Error{st, ctx->last_error};
            } catch (...) {
                for (auto* c : owned) free_col(ctx, c);
                qe_gpu_pairs_free(ctx, pr);
                throw;
            }
            for (auto* c : owned) free_col(ctx, c);
            qe_gpu_pairs_free(ctx, pr);
            *out = res;
            return QE_GPU_OK;
        }

        // ---- fast path: pre-group the build side ---------------------------------
        // F64 sums scaled from a sample; a value past the sampled bound makes
        // the emission throw ScaleRetry and the query runs again with bounds
        // taken from every row.
        auto run = [&](bool exact) -> qe_dev_table* {
        AggProg p = p0;
            KeyEmit ke{};
            ke.nkeys = static_cast<int>(ngk);
            for (uint32_t k = 0; k < ngk; ++k) ke.types[k] = col_at(group_keys[k])->type;
            Scratch gslot(ctx, std::max<uint64_t>(nb, 1) * 4);
            DeviceHashTable gtab;
            uint64_t S = 0;
            const qe_dev_col* g0 = col_at(group_keys[0]);
            bool gslot_ready = true;
            // The per-row group slots of a direct group key, when a path needs the array.
            auto ensure_gslot = [&]() {
                if (gslot_ready) return;
                gslot_direct_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
                    view(g0), nb, ke.kmin, static_cast<uint32_t>(ke.null_slot), gslot.as<uint32_t>());
                QE_HIP(hipGetLastError());
                gslot_ready = true;
            };
            bool gdirect = false;
            const bool int_key = nkeys == 1 && (pk[0]->type == QE_I64 || pk[0]->type == QE_TS || pk[0]->type == QE_I32);
            // Group-key and build-key ranges in one pass and one read-back.
            const bool want_g = ngk == 1 && is_int_like(g0->type) && nb > 0;
            const bool want_k = int_key && nb > 0 && np > 0;
            MinMax g_mm{}, k_mm{};
            Scratch fbuf;  // exact F64 sums: scales, flags, outliers
            {
                const qe_dev_col* mc[2];
                uint32_t nm = 0;
                if (want_g) mc[nm++] = g0;
                if (want_k) mc[nm++] = bk[0];
                MinMax r[2];
                MinMaxReadback rb;
                if (nm) rb = column_minmax_start(ctx, mc, nm, nb);
                // The F64 scales need only the probe's values: queued here,
                // they run while the host waits for the ranges.
                prepare_f64_scales(ctx, p, np, exact, fbuf);
                if (nm) column_minmax_finish(ctx, rb, r);
                if (want_g) g_mm = r[0];
                if (want_k) k_mm = r[want_g ? 1 : 0];
            }
            if (want_g) {
                const MinMax mm = g_mm;
                const uint64_t R = mm.valid ? static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1 : 0;
                if (R <= std::max<uint64_t>(2 * nb, 1024) && R < (uint64_t(1) << 30)) {
                    gdirect = true;
                    S = R + 1;
                    ke.mode = 0;
                    ke.kmin = mm.valid ? mm.mn : 0;
                    ke.null_slot = S - 1;
                    gslot_ready = false;  // computed only by the paths that read the array
                }
            }
            if (!gdirect) {
                HashKeys hk{};
                hk.nkeys = static_cast<int>(ngk);
                for (uint32_t k = 0; k < ngk; ++k) hk.key[k] = view(col_at(group_keys[k]));
                gtab.create(ctx, static_cast<int>(ngk), std::max<uint64_t>(nb, 1));
                S = gtab.v.cap;
                if (nb)
                    gslot_hash_kernel<<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(hk, nb, gtab.v,
                                                                                                     gslot.as<uint32_t>());
                QE_HIP(hipGetLastError());
                gtab.check_overflow(ctx);
                ke.mode = 1;
                ke.words = gtab.v.words;
                ke.nullmask = gtab.v.nullmask;
            }

            Scratch st(ctx, state_words(S, p) * 8);
            const size_t lb = lds_bytes(static_cast<uint32_t>(std::min<uint64_t>(S, UINT32_MAX)), p);
            const bool use_lds = S < (uint64_t(1) << 20) && lb <= kLdsBudget;
            // A slot space larger than LDS: LDS partials for one slot range per
            // pass over the probe (each pass streams the probe columns again),
            // instead of HBM atomics that serialise on the heavy groups.
            const uint64_t pass_slots = std::max<uint64_t>(1, kLdsBudget / lds_bytes(1, p));
            const uint64_t npasses = (S + pass_slots - 1) / pass_slots;
            const bool multipass = !use_lds && npasses <= kMaxSlotPasses;
            // Blocks per CU as tuned, as far as the LDS partials leave room.
            const uint64_t lds_per_block = lds_bytes(static_cast<uint32_t>(std::min<uint64_t>(S, pass_slots)), p);
            const uint64_t bpc = std::max<uint64_t>(
                1, std::min<uint64_t>(ctx->tune_join_blocks_per_cu,
                                      lds_per_block ? (kLdsBudget + 256) / lds_per_block : 8));
            const unsigned lds_blocks = static_cast<unsigned>(
                std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(ctx->num_cus) * bpc, (np + 1023) / 1024)));
            // The state's F64 NaN/inf flags and outlier sums cleared: before
            // the first kernel that updates the state or the emission, on every
            // path, and as late as each path allows (the dense path clears them
            // after the map build's launches, so the build starts sooner).
            bool f64_state_ready = false;
            auto ensure_f64_state = [&]() {
                if (f64_state_ready) return;
                prepare_f64_state(ctx, p, st.as<uint64_t>(), S);
                f64_state_ready = true;
            };
            const bool small = S < 0xFFFF;
            const uint64_t* kv = pk[0]->all_valid ? nullptr : pk[0]->validity;

            // Probe over a direct unique map (gmap[key - kmin] = group slot).
            auto probe_direct = [&](const Scratch& gmap, int64_t kmin, uint64_t range) {
                ensure_f64_state();
                auto launch = [&](auto key_tag, auto map_tag) {
                    using KeyT = decltype(key_tag);
                    using MapT = decltype(map_tag);
                    const KeyT* keys = static_cast<const KeyT*>(pk[0]->data);
                    const MapT* gm = static_cast<const MapT*>(gmap.p);
                    StreamCols sc{};
                    int nv = 0;
                    const bool streamable = (use_lds || multipass) && stream_cols(p, keys, sizeof(KeyT), sc, nv);
                    // More slots than LDS holds: when a sample of the probe puts at
                    // least 7/8 of the rows in at most an LDS range of slots, and no
                    // slot outside them carries more than 1/4096 of the rows or
                    // about 2^16 rows, those slots are relabelled first (the map is
                    // rewritten through the relabelling) and one pass keeps them in
                    // LDS while the other rows go to per-block spill lists; the lists
                    // are then folded one LDS range at a time, and the relabelled
                    // state returns to slot order.
                    bool spill = false;
                    if (multipass && streamable && np >= kSpillMinRows) {
                        const uint32_t S32 = static_cast<uint32_t>(S), W = static_cast<uint32_t>(pass_slots);
                        Scratch perm(ctx, size_t(S) * 4);
                        const HotSet hs = hot_set<KeyT, MapT>(ctx, keys, kv, np, gm, kmin, range, S32, W,
                                                              perm.as<uint32_t>());
                        // Spill lists sized for a quarter of each block's rows (the
                        // sample put at most an eighth outside the LDS range).
                        const uint64_t per_block = (np + lds_blocks - 1) / lds_blocks;
                        const uint32_t cap = ctx->tune_join_spill_cap
                                                 ? ctx->tune_join_spill_cap
                                                 : static_cast<uint32_t>(std::min<uint64_t>(per_block / 4 + 1024, 1u << 30));
                        const size_t rec_bytes = size_t(lds_blocks) * cap * (1 + nv) * 8;
                        // Its working set (relabelled map and state, partials, spill
                        // lists) must fit next to what is live, else the passes run.
                        const size_t spill_bytes = size_t(range) * sizeof(MapT) + state_words(S, p) * 8 +
                                                   size_t(lds_blocks) * part_words(W, p) * 8 + rec_bytes +
                                                   size_t(lds_blocks) * 4;
                        spill = hs.total >= 1024 && uint64_t(hs.total - hs.inside) * 8 <= hs.total &&
                                (uint64_t(hs.cold_max) * 4096 <= hs.total ||
                                 uint64_t(hs.cold_max) * hs.stride <= (uint64_t(1) << 16)) &&
                                rec_bytes <= kSpillMaxBytes && device_room(ctx, spill_bytes);
                        Scratch gm2, st2, part, rec, rcount;
                        if (spill) {
                            try {  // every allocation before any launch: a failure leaves no trace
                                gm2 = Scratch(ctx, size_t(range) * sizeof(MapT));
                                st2 = Scratch(ctx, state_words(S, p) * 8);
                                part = Scratch(ctx, size_t(lds_blocks) * part_words(W, p) * 8);
                                rec = Scratch(ctx, rec_bytes);
                                rcount = Scratch(ctx, size_t(lds_blocks) * 4);
                            } catch (const Error& e) {
                                if (e.status != QE_GPU_OUT_OF_MEMORY) throw;
                                spill = false;
                                gm2 = Scratch();
                                st2 = Scratch();
                                part = Scratch();
                                rec = Scratch();
                                rcount = Scratch();
                            }
                        }
                        if (spill) {
                            remap_slots_kernel<MapT><<<grid_for(range, 256, 4096), 256, 0, ctx->stream>>>(
                                gm, range, perm.as<uint32_t>(), gm2.as<MapT>());
                            QE_HIP(hipGetLastError());
                            init_state(ctx, st2.as<uint64_t>(), S, p);
                            // The relabelled state's own NaN/inf flags (moved with it).
                            AggProg p2 = p;
                            p2.fflag = st2.as<uint64_t>() + state_flags(S, p, 0);
                            Spill sp;
                            sp.st = st2.as<uint64_t>();
                            sp.S = S;
                            sp.rec = rec.as<uint64_t>();
                            sp.cap = cap;
                            sp.count = rcount.as<uint32_t>();
                            {
                                TimedRegion tr(ctx, "join_agg_probe");
                                launch_stream<KeyT, MapT>(ctx, ctx->tune_join_pairs, nv, lds_blocks, lds_bytes(W, p), keys,
                                                          kv, np, gm2.as<MapT>(), kmin, range, W, 0, p2, sc,
                                                          part.as<uint64_t>(), sp);
                            }
                            merge_partials(ctx, part.as<uint64_t>(), lds_blocks, st2.as<uint64_t>(), W, p, S, 0, true);
                            for (uint64_t lo = W; lo < S; lo += W) {
                                const uint32_t Sp = static_cast<uint32_t>(std::min<uint64_t>(W, S - lo));
                                {
                                    TimedRegion tr(ctx, "join_agg_spill");
                                    launch_spill_agg(ctx, nv, lds_blocks, lds_bytes(Sp, p), rec.as<uint64_t>(),
                                                     rcount.as<uint32_t>(), cap, static_cast<uint32_t>(lo), Sp, p2, sc,
                                                     part.as<uint64_t>());
                                }
                                merge_partials(ctx, part.as<uint64_t>(), lds_blocks, st2.as<uint64_t>(), Sp, p, S, lo,
                                               true);
                            }
                            fold_outliers(ctx, p2, st2.as<uint64_t>(), S);  // relabelled slots: before the move
                            const uint32_t fields = static_cast<uint32_t>(state_words(1, p));
                            unpermute_state_kernel<<<grid_for(S * fields, 256, 4096), 256, 0, ctx->stream>>>(
                                st2.as<uint64_t>(), S, fields, perm.as<uint32_t>(), st.as<uint64_t>());
                            QE_HIP(hipGetLastError());
                        }
                    }
                    if (spill) return;
                    if (use_lds || multipass) {
                        const uint64_t per = use_lds ? S : pass_slots;
                        Scratch part(ctx, size_t(lds_blocks) * part_words(static_cast<uint32_t>(per), p) * 8);
                        for (uint64_t lo = 0; lo < S; lo += per) {
                            const uint32_t Sp = static_cast<uint32_t>(std::min<uint64_t>(per, S - lo));
                            const size_t lbp = lds_bytes(Sp, p);
                            {
                                TimedRegion tr(ctx, "join_agg_probe");
                                if (streamable) {
                                    launch_stream<KeyT, MapT>(ctx, ctx->tune_join_pairs, nv, lds_blocks, lbp, keys, kv, np,
                                                              gm, kmin, range, Sp, static_cast<uint32_t>(lo), p, sc,
                                                              part.as<uint64_t>());
                                } else {
                                    join_agg_direct_lds<KeyT, MapT><<<lds_blocks, 1024, lbp, ctx->stream>>>(
                                        keys, kv, np, gm, kmin, range, Sp, static_cast<uint32_t>(lo), p,
                                        part.as<uint64_t>());
                                    QE_HIP(hipGetLastError());
                                }
                            }
                            merge_partials(ctx, part.as<uint64_t>(), lds_blocks, st.as<uint64_t>(), Sp, p, S, lo);
                        }
                    } else {
                        init_state(ctx, st.as<uint64_t>(), S, p);
                        TimedRegion tr(ctx, "join_agg_probe");
                        join_agg_direct_global<KeyT, MapT><<<grid_for(np, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                            keys, kv, np, gm, kmin, range, S, p, st.as<uint64_t>());
                        QE_HIP(hipGetLastError());
                    }
                };
                if (pk[0]->type == QE_I32) {
                    if (small) launch(int32_t(0), uint16_t(0));
                    else launch(int32_t(0), uint32_t(0));
                } else {
                    if (small) launch(int64_t(0), uint16_t(0));
                    else launch(int64_t(0), uint32_t(0));
                }
            };

            // Dense unique integer keys: partitioned map build, then the streaming probe.
            bool done = false;
            if (want_k) {
                const MinMax mm = k_mm;
                if (mm.valid == 0) {  // every build key NULL: nothing matches
                    ensure_f64_state();
                    init_state(ctx, st.as<uint64_t>(), S, p);
                    done = true;
                } else {
                    const uint64_t R = static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1;
                    if (R <= std::max<uint64_t>(4 * nb, uint64_t(1) << 16) && R < (uint64_t(1) << 31)) {
                        Scratch gmap, dup;
                        MapSlots ms;
                        if (small && !gslot_ready) {  // the scatter reads the group column itself
                            ms.g = g0;
                            ms.gmin = ke.kmin;
                            ms.null_slot = static_cast<uint32_t>(ke.null_slot);
                        } else {
                            ensure_gslot();
                            ms.gslot = gslot.as<uint32_t>();
                        }
                        const MapBuild mbld = small ? build_direct_map<uint16_t>(ctx, bk[0], nb, mm.valid, mm.mn, R, ms,
                                                                                 gmap, dup)
                                                    : build_direct_map<uint32_t>(ctx, bk[0], nb, mm.valid, mm.mn, R, ms,
                                                                                 gmap, dup);
                        if (mbld == MapBuild::kOk) {
                            probe_direct(gmap, mm.mn, R);
                            // Repeated build keys make the map wrong: then the
                            // groups are discarded and the row-list path runs. The
                            // flag comes back with the emission's counts.
                            unsigned int dupv = 0;
                            qe_dev_table* t = emit_groups(ctx, st.as<uint64_t>(), S, p, ke, false,
                                                          dup.as<unsigned int>(), &dupv);
                            if (dupv == 0) return t;
                            qe_gpu_table_free(ctx, t);
                            // The row-list path starts over: an empty outlier
                            // list and cleared flags.
                            prepare_f64_scales(ctx, p, np, exact, fbuf);
                            f64_state_ready = false;
                        }
                    }
                }
            }

            ensure_f64_state();
            JoinIndex ix;
            if (done) {
            } else if (nb > 0) {
                build_index(ctx, ix, bk.data(), nkeys, nb);
            } else {
                ix.empty = true;
            }
            if (!done) ensure_gslot();
            if (done) {
            } else if (ix.empty || np == 0) {
                init_state(ctx, st.as<uint64_t>(), S, p);
            } else if (ix.mode == 0 && ix.unique && int_key) {
                // Direct unique index too wide for the partitioned build.
                const size_t mb = ix.range * (small ? 2 : 4);
                Scratch gmap(ctx, mb);
                memset_async(ctx, gmap.p, 0xFF, mb);
                if (small)
                    gmap_fill_kernel<uint16_t><<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
                        view(bk[0]), nb, ix.kmin, gslot.as<uint32_t>(), gmap.as<uint16_t>());
                else
                    gmap_fill_kernel<uint32_t><<<grid_for(nb, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
                        view(bk[0]), nb, ix.kmin, gslot.as<uint32_t>(), gmap.as<uint32_t>());
                QE_HIP(hipGetLastError());
                probe_direct(gmap, ix.kmin, ix.range);
            } else {
                ProbeAggLists L{ix.view(pk.data(), static_cast<int>(nkeys)), gslot.as<uint32_t>()};
                if (use_lds) {
                    Scratch part(ctx, size_t(lds_blocks) * part_words(static_cast<uint32_t>(S), p) * 8);
                    {
                        TimedRegion tr(ctx, "join_agg_probe");
                        join_agg_lists_lds<<<lds_blocks, 1024, lb, ctx->stream>>>(L, np, static_cast<uint32_t>(S), p,
                                                                                  part.as<uint64_t>());
                        QE_HIP(hipGetLastError());
                    }
                    merge_partials(ctx, part.as<uint64_t>(), lds_blocks, st.as<uint64_t>(), S, p);
                } else {
                    init_state(ctx, st.as<uint64_t>(), S, p);
                    TimedRegion tr(ctx, "join_agg_probe");
                    join_agg_lists_global<<<grid_for(np, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                        L, np, S, p, st.as<uint64_t>());
                    QE_HIP(hipGetLastError());
                }
            }
            return emit_groups(ctx, st.as<uint64_t>(), S, p, ke, false);
        };
        try {
            *out = run(false);
        } catch (const ScaleRetry&) {
            *out = run(true);
        }
    });
}

}  // namespace qe_gpu

extern "C" qe_gpu_status qe_gpu_join_aggregate(qe_gpu_ctx* ctx, const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                               const qe_dev_col* const* build_cols, uint32_t nbuild,
                                               const uint32_t* probe_keys, const uint32_t* build_keys,
                                               uint32_t nkeys, int32_t join_type, const uint32_t* group_keys,
                                               uint32_t ngk, const qe_agg_spec* aggs, uint32_t naggs,
                                               qe_dev_table** out) {
    // STR join keys, group keys and MIN/MAX inputs compare by value
    // (strings.hip): the rewritten columns are appended to their own side, and
    // the indices of the join output (probe columns ++ build columns) follow.
    QE_ENTRY(ctx, {
        invalid(out != nullptr && probe_cols != nullptr && build_cols != nullptr, "null argument");
        invalid(nprobe >= 1 && nbuild >= 1, "join inputs need >=1 column each");
        invalid(nkeys >= 1 && probe_keys != nullptr && build_keys != nullptr, "equi-join needs >=1 key");
        invalid(group_keys != nullptr || ngk == 0, "null group key list");
        invalid(aggs != nullptr || naggs == 0, "null aggregate list");
        const uint32_t ncols = nprobe + nbuild;
        auto col_at = [&](uint32_t c) { return c < nprobe ? probe_cols[c] : build_cols[c - nprobe]; };
        std::vector<const qe_dev_col*> P(probe_cols, probe_cols + nprobe), B(build_cols, build_cols + nbuild);
        OwnedCols own(ctx);
        StrAggFix fix;
        fix.key_dict.assign(ngk, nullptr);
        fix.rank_dict.assign(naggs, nullptr);
        bool any = false;
        for (uint32_t k = 0; k < ngk; ++k) {
            invalid(group_keys[k] < ncols, "group key column out of range");
            const qe_dev_col* c = col_at(group_keys[k]);
            if (c->type != QE_STR) continue;
            const qe_gpu_dict* d = need_dict(c);
            fix.key_dict[k] = d;
            any = true;
            if (d->unique) continue;
            qe_dev_col* r = own.keep(remap_col(ctx, c, d->canon, QE_STR, d));
            if (group_keys[k] < nprobe) P[group_keys[k]] = r;
            else B[group_keys[k] - nprobe] = r;
        }
        // Appended rank columns: (side, index within that side).
        std::vector<std::pair<int, uint32_t>> rank_at(naggs, {-1, 0});
        for (uint32_t a = 0; a < naggs; ++a) {
            if (aggs[a].func != QE_AGG_MIN && aggs[a].func != QE_AGG_MAX) continue;
            invalid(aggs[a].input_col < ncols, "aggregate input column out of range");
            const qe_dev_col* c = col_at(aggs[a].input_col);
            if (c->type != QE_STR) continue;
            const qe_gpu_dict* d = need_dict(c);
            fix.rank_dict[a] = d;
            any = true;
            qe_dev_col* r = own.keep(remap_col(ctx, c, d->rank, QE_I32, nullptr));
            if (aggs[a].input_col < nprobe) {
                P.push_back(r);
                rank_at[a] = {0, static_cast<uint32_t>(P.size() - 1)};
            } else {
                B.push_back(r);
                rank_at[a] = {1, static_cast<uint32_t>(B.size() - 1)};
            }
        }
        std::vector<uint32_t> pk2(probe_keys, probe_keys + nkeys), bk2(build_keys, build_keys + nkeys);
        {
            std::vector<const qe_dev_col*> pkc(nkeys), bkc(nkeys), pko, bko;
            for (uint32_t j = 0; j < nkeys; ++j) {
                invalid(probe_keys[j] < nprobe && build_keys[j] < nbuild, "join key column out of range");
                pkc[j] = probe_cols[probe_keys[j]];
                bkc[j] = build_cols[build_keys[j]];
            }
            check_key_types(pkc.data(), bkc.data(), nkeys);
            str_join_keys(ctx, pkc.data(), bkc.data(), nkeys, pko, bko, own);
            for (uint32_t j = 0; j < nkeys; ++j) {
                if (pko[j] != pkc[j]) {
                    P.push_back(pko[j]);
                    pk2[j] = static_cast<uint32_t>(P.size() - 1);
                }
                if (bko[j] != bkc[j]) {
                    B.push_back(bko[j]);
                    bk2[j] = static_cast<uint32_t>(B.size() - 1);
                }
            }
        }
        const uint32_t np2 = static_cast<uint32_t>(P.size());
        auto remap_index = [&](uint32_t c) { return c < nprobe ? c : c - nprobe + np2; };
        std::vector<uint32_t> gk2(ngk);
        for (uint32_t k = 0; k < ngk; ++k) gk2[k] = remap_index(group_keys[k]);
        std::vector<qe_agg_spec> a2(aggs, aggs + naggs);
        for (uint32_t a = 0; a < naggs; ++a) {
            if (rank_at[a].first == 0) a2[a].input_col = rank_at[a].second;
            else if (rank_at[a].first == 1) a2[a].input_col = np2 + rank_at[a].second;
            else if (a2[a].func != QE_AGG_COUNT_STAR && a2[a].input_col < ncols) a2[a].input_col = remap_index(a2[a].input_col);
        }
        qe_dev_table* t = nullptr;
        const qe_gpu_status st = join_aggregate_codes(ctx, P.data(), np2, B.data(), static_cast<uint32_t>(B.size()),
                                                      pk2.data(), bk2.data(), nkeys, join_type, gk2.data(), ngk,
                                                      a2.data(), naggs, &t);
        if (st != QE_GPU_OK) throw Error{st, ctx->last_error};
        if (any) {
            try {
                str_agg_finish(ctx, t, fix);
            } catch (...) {
                qe_gpu_table_free(ctx, t);
                throw;
            }
        }
        *out = t;
    });
}

extern "C" qe_gpu_status qe_gpu_join_autotune(qe_gpu_ctx* ctx, const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                              const qe_dev_col* const* build_cols, uint32_t nbuild,
                                              const uint32_t* probe_keys, const uint32_t* build_keys, uint32_t nkeys,
                                              int32_t join_type, const uint32_t* group_keys, uint32_t ngk,
                                              const qe_agg_spec* aggs, uint32_t naggs, qe_join_tuning* out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null output");
        const bool timing = ctx->timing;
        const int pairs0 = ctx->tune_join_pairs, bpc0 = ctx->tune_join_blocks_per_cu;
        ctx->timing = true;
        qe_join_tuning best{pairs0, bpc0, 1e30, 0.0, 0};
        try {
            for (int pairs : {4, 2}) {
                for (int bpc : {1, 2, 4}) {
                    ctx->tune_join_pairs = pairs;
                    ctx->tune_join_blocks_per_cu = bpc;
                    double t = 1e30;
                    for (int rep = 0; rep < 2; ++rep) {
                        resolve_timings(ctx);
                        ctx->timings.erase("join_agg_probe");
                        qe_dev_table* tab = nullptr;
                        const qe_gpu_status st = qe_gpu_join_aggregate(ctx, probe_cols, nprobe, build_cols, nbuild,
                                                                       probe_keys, build_keys, nkeys, join_type,
                                                                       group_keys, ngk, aggs, naggs, &tab);
                        if (st != QE_GPU_OK) throw Error{st, ctx->last_error};
                        qe_gpu_table_free(ctx, tab);
                        resolve_timings(ctx);
                        const auto it = ctx->timings.find("join_agg_probe");
                        const double ms = it == ctx->timings.end() ? 0.0 : it->second.ms;
                        t = std::min(t, ms);
                    }
                    ++best.candidates;
                    if (pairs == 4 && bpc == 1) best.default_ms = t;
                    if (t < best.probe_ms) {
                        best.probe_ms = t;
                        best.pairs = pairs;
                        best.blocks_per_cu = bpc;
                    }
                }
            }
        } catch (...) {
            ctx->timing = timing;
            ctx->tune_join_pairs = pairs0;
            ctx->tune_join_blocks_per_cu = bpc0;
            throw;
        }
        ctx->timing = timing;
        resolve_timings(ctx);
        ctx->timings.erase("join_agg_probe");
        ctx->tune_join_pairs = best.pairs;
        ctx->tune_join_blocks_per_cu = best.blocks_per_cu;
        *out = best;
    });
}
