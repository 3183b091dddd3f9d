// Device HashTable: key tuple -> dense group id in first-seen row order,
// replacing qe::HashTable (ops/hashtable.h:108-190) with the same contract:
//  * insert_or_find (ops/hashtable.cpp:93-133): identical keys share an id; a
//    new key gets the next id in first-seen order; ids never change, across
//    calls and across growth; NullPolicy::kNeverMatch gives every key tuple with
//    a NULL a fresh, unfindable ("dead") id, interleaved in the same order;
//  * find (:135-163): the id of an existing key or kNoGroup; never mutates;
//  * key store read-back (group_key_word / group_is_null, :143-155).
//
// Parallel construction of first-seen ids: every row claims or finds its slot
// and records the smallest row index that inserted the key (atomicMin); a row
// is "first" if it is that row of a new slot, or a dead NULL-key row; an
// exclusive scan over the first flags assigns next_id + rank; every other row
// reads its slot's id. This reproduces the sequential assignment exactly.
#include "agg_common.hpp"
#include "hash_common.hpp"

#include <algorithm>

using namespace qe_gpu;

namespace qe_gpu {
constexpr uint32_t kPending = 0xFFFFFFFEu;
constexpr uint32_t kDead = 0xFFFFFFFEu;  // slot_of marker for a dead NULL-key row
}  // namespace qe_gpu

struct qe_dev_ht {
    int nkeys = 0;
    int32_t types[kMaxKeys] = {};
    int policy = 0;
    double lf = 0.7;
    uint64_t seed = kHashSeed;
    uint64_t cap = 0;
    uint64_t next_id = 0;
    uint64_t occupied = 0;
    // slot arrays
    uint32_t* state = nullptr;
    uint64_t* hash = nullptr;
    uint64_t* words = nullptr;  // [nkeys][cap]
    uint64_t* nullmask = nullptr;
    uint32_t* gid = nullptr;
    uint32_t* first = nullptr;
    unsigned int* overflow = nullptr;
    // key store by id
    uint64_t id_cap = 0;
    uint64_t* id_words = nullptr;  // [nkeys][id_cap]
    uint64_t* id_null = nullptr;   // [id_cap]
};

namespace qe_gpu {

struct HtDev {
    HashTableView t;
    uint32_t* gid;
    uint32_t* first;
};

__device__ inline uint32_t ht_insert_first(const HtDev& d, int nkeys, const uint64_t* w, uint64_t nm, uint64_t h,
                                           uint32_t row) {
    const HashTableView& t = d.t;
    const uint64_t mask = t.cap - 1;
    uint64_t s = h & mask;
    uint64_t probes = 0;
    uint32_t rereads = 0;
    while (true) {
        const uint32_t st = __hip_atomic_load(&t.state[s], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (st != 2) {
            if (++rereads > kMaxRereads) {
                atomicOr(t.overflow, 2u);
                return kNoSlot;
            }
            if (st == 0) {
                uint32_t expected = 0;
                if (__hip_atomic_compare_exchange_strong(&t.state[s], &expected, 1u, __ATOMIC_ACQ_REL,
                                                         __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) {
                    for (int j = 0; j < nkeys; ++j) t.words[size_t(j) * t.cap + s] = w[j];
                    t.nullmask[s] = nm;
                    t.hash[s] = h;
                    d.gid[s] = kPending;
                    d.first[s] = row;
                    __hip_atomic_store(&t.state[s], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    return static_cast<uint32_t>(s);
                }
            }
            continue;
        }
        if (slot_equal(t, nkeys, s, w, nm, h)) {
            if (d.gid[s] == kPending) atomicMin(&d.first[s], row);
            return static_cast<uint32_t>(s);
        }
        s = (s + 1) & mask;
        rereads = 0;
        if (++probes >= t.cap) {
            atomicOr(t.overflow, 1u);
            return kNoSlot;
        }
    }
}

__global__ void ht_insert_kernel(HashKeys hk, const uint32_t* sel, uint64_t n, int never, HtDev d,
                                 uint32_t* slot_of) {
    for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < n;
         k += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t r = sel ? sel[k] : k;
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);
        if (never && nm != 0) {
            slot_of[k] = kDead;
            continue;
        }
        slot_of[k] = ht_insert_first(d, hk.nkeys, w, nm, h, static_cast<uint32_t>(k));
    }
}

__global__ void ht_flag_kernel(const uint32_t* slot_of, uint64_t n, const uint32_t* gid, const uint32_t* first,
                               uint32_t* is_new) {
    for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < n;
         k += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = slot_of[k];
        is_new[k] = (s == kDead) || (s != kNoSlot && gid[s] == kPending && first[s] == k);
    }
}

__global__ void ht_assign_kernel(HashKeys hk, const uint32_t* sel, uint64_t n, const uint32_t* slot_of,
                                 const uint32_t* is_new, const uint64_t* rank, uint64_t next_id, HtDev d,
                                 uint64_t* id_words, uint64_t* id_null, uint64_t id_cap, uint32_t* out) {
    for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < n;
         k += uint64_t(gridDim.x) * blockDim.x) {
        if (!is_new[k]) continue;
        const uint64_t id = next_id + rank[k];
        const uint32_t s = slot_of[k];
        if (s == kDead) {
            uint64_t w[kMaxKeys], nm, h;
            normalize_row(hk, sel ? sel[k] : k, w, nm, h);
            for (int j = 0; j < hk.nkeys; ++j) id_words[size_t(j) * id_cap + id] = w[j];
            id_null[id] = nm;
            out[k] = static_cast<uint32_t>(id);
        } else {
            for (int j = 0; j < hk.nkeys; ++j) id_words[size_t(j) * id_cap + id] = d.t.words[size_t(j) * d.t.cap + s];
            id_null[id] = d.t.nullmask[s];
            d.gid[s] = static_cast<uint32_t>(id);
        }
    }
}

__global__ void ht_resolve_kernel(const uint32_t* slot_of, uint64_t n, const uint32_t* gid, uint32_t* out) {
    for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < n;
         k += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = slot_of[k];
        if (s != kDead && s != kNoSlot) out[k] = gid[s];
    }
}

__global__ void ht_find_kernel(HashKeys hk, const uint32_t* sel, uint64_t n, int never, HashTableView t,
                               const uint32_t* gid, uint32_t* out) {
    for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < n;
         k += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t r = sel ? sel[k] : k;
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(hk, r, w, nm, h);
        if (never && nm != 0) {
            out[k] = kNoRow;
            continue;
        }
        const uint32_t s = ht_find(t, hk.nkeys, w, nm, h);
        out[k] = (s == kNoSlot) ? kNoRow : gid[s];
    }
}

// Re-slot every live id into a fresh table (growth: ids unchanged).
__global__ void ht_rehash_kernel(int nkeys, uint64_t nids, const uint64_t* id_words, const uint64_t* id_null,
                                 uint64_t id_cap, int never, HtDev d) {
    for (uint64_t id = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; id < nids;
         id += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t nm = id_null[id];
        if (never && nm != 0) continue;  // dead ids are never slotted
        uint64_t h = kHashSeed;
        for (int j = 0; j < nkeys; ++j) h = hash_combine(h, id_words[size_t(j) * id_cap + id]);
        const uint64_t mask = d.t.cap - 1;
        uint64_t s = h & mask;
        for (uint64_t probes = 0; probes < d.t.cap; ++probes, s = (s + 1) & mask) {
            uint32_t expected = 0;
            if (__hip_atomic_compare_exchange_strong(&d.t.state[s], &expected, 1u, __ATOMIC_ACQ_REL,
                                                     __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) {
                for (int j = 0; j < nkeys; ++j) d.t.words[size_t(j) * d.t.cap + s] = id_words[size_t(j) * id_cap + id];
                d.t.nullmask[s] = nm;
                d.t.hash[s] = h;
                d.gid[s] = static_cast<uint32_t>(id);
                d.first[s] = 0;
                __hip_atomic_store(&d.t.state[s], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
}

static HtDev dev_view(qe_dev_ht* ht) {
    HtDev d;
    d.t = HashTableView{ht->state, ht->hash, ht->words, ht->nullmask, ht->cap, ht->overflow};
    d.gid = ht->gid;
    d.first = ht->first;
    return d;
}

static void free_slots(qe_gpu_ctx* ctx, qe_dev_ht* ht) {
    dev_free(ctx, ht->state);
    dev_free(ctx, ht->hash);
    dev_free(ctx, ht->words);
    dev_free(ctx, ht->nullmask);
    dev_free(ctx, ht->gid);
    dev_free(ctx, ht->first);
    ht->state = nullptr;
}

static void alloc_slots(qe_gpu_ctx* ctx, qe_dev_ht* ht, uint64_t cap) {
    ht->cap = cap;
    ht->state = static_cast<uint32_t*>(dev_alloc(ctx, cap * 4));
    ht->hash = static_cast<uint64_t*>(dev_alloc(ctx, cap * 8));
    ht->words = static_cast<uint64_t*>(dev_alloc(ctx, cap * 8 * ht->nkeys));
    ht->nullmask = static_cast<uint64_t*>(dev_alloc(ctx, cap * 8));
    ht->gid = static_cast<uint32_t*>(dev_alloc(ctx, cap * 4));
    ht->first = static_cast<uint32_t*>(dev_alloc(ctx, cap * 4));
    memset_async(ctx, ht->state, 0, cap * 4);
}

static void ensure_capacity(qe_gpu_ctx* ctx, qe_dev_ht* ht, uint64_t upcoming) {
    // Grow (doubling) so that even if every upcoming row were a new key the
    // load stays within the configured factor (ops/hashtable.cpp:75-91 grows
    // after the insert that crosses it; growing ahead keeps device inserts
    // lock-free).
    uint64_t cap = ht->cap;
    while (static_cast<double>(ht->occupied + upcoming) > ht->lf * static_cast<double>(cap)) cap *= 2;
    if (cap == ht->cap) return;
    check(cap <= (uint64_t(1) << 31), QE_GPU_UNSUPPORTED, "hash table capacity beyond 2^31 slots");
    free_slots(ctx, ht);
    alloc_slots(ctx, ht, cap);
    if (ht->next_id)
        ht_rehash_kernel<<<grid_for(ht->next_id, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
            ht->nkeys, ht->next_id, ht->id_words, ht->id_null, ht->id_cap,
            ht->policy == QE_NULL_NEVER_MATCH, dev_view(ht));
    QE_HIP(hipGetLastError());
}

static void ensure_id_capacity(qe_gpu_ctx* ctx, qe_dev_ht* ht, uint64_t need) {
    if (need <= ht->id_cap) return;
    uint64_t nc = std::max<uint64_t>(ht->id_cap ? ht->id_cap : 1024, 1024);
    while (nc < need) nc *= 2;
    auto* nw = static_cast<uint64_t*>(dev_alloc(ctx, nc * 8 * ht->nkeys));
    auto* nn = static_cast<uint64_t*>(dev_alloc(ctx, nc * 8));
    for (int j = 0; j < ht->nkeys && ht->next_id; ++j)
        QE_HIP(hipMemcpyAsync(nw + size_t(j) * nc, ht->id_words + size_t(j) * ht->id_cap, ht->next_id * 8,
                              hipMemcpyDeviceToDevice, ctx->stream));
    if (ht->next_id)
        QE_HIP(hipMemcpyAsync(nn, ht->id_null, ht->next_id * 8, hipMemcpyDeviceToDevice, ctx->stream));
    dev_free(ctx, ht->id_words);
    dev_free(ctx, ht->id_null);
    ht->id_words = nw;
    ht->id_null = nn;
    ht->id_cap = nc;
}

static HashKeys keys_of(const qe_dev_ht* ht, const qe_dev_col* const* keys) {
    invalid(keys != nullptr, "null key columns");
    HashKeys hk{};
    hk.nkeys = ht->nkeys;
    for (int j = 0; j < ht->nkeys; ++j) {
        invalid(keys[j] != nullptr, "null key column");
        invalid(keys[j]->type == ht->types[j], "key column type differs from the table's key type");
        hk.key[j] = view(keys[j]);
    }
    return hk;
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_ht_create(qe_gpu_ctx* ctx, const int32_t* key_types, uint32_t nkeys, int32_t null_policy,
                               uint64_t initial_capacity, double max_load_factor, uint64_t seed, qe_dev_ht** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && key_types != nullptr, "null argument");
        invalid(nkeys >= 1, "key column count must be >= 1");
        check(nkeys <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many key columns for the device table");
        invalid(null_policy == QE_NULL_EQUAL || null_policy == QE_NULL_NEVER_MATCH, "unknown null policy");
        auto* ht = new qe_dev_ht();
        ht->nkeys = static_cast<int>(nkeys);
        for (uint32_t j = 0; j < nkeys; ++j) {
            invalid(valid_type(key_types[j]), "unknown key type");
            ht->types[j] = key_types[j];
        }
        ht->policy = null_policy;
        // HashTableConfig normalisation (ops/hashtable.cpp:43-52).
        double lf = max_load_factor;
        if (!(lf > 0.0)) lf = 0.7;
        if (lf > 0.95) lf = 0.95;
        ht->lf = lf;
        ht->seed = seed;
        uint64_t cap = 8;
        while (cap < initial_capacity) cap <<= 1;
        ht->overflow = static_cast<unsigned int*>(dev_alloc(ctx, 8));
        memset_async(ctx, ht->overflow, 0, 8);
        alloc_slots(ctx, ht, cap);
        *out = ht;
    });
}

qe_gpu_status qe_gpu_ht_insert_or_find(qe_gpu_ctx* ctx, qe_dev_ht* ht, const qe_dev_col* const* keys,
                                       const uint32_t* sel, uint64_t n, uint32_t* out_groups) {
    QE_ENTRY(ctx, {
        invalid(ht != nullptr, "null table");
        if (n == 0) return QE_GPU_OK;
        invalid(out_groups != nullptr, "null output");
        check_rows32(n);
        const HashKeys hk = keys_of(ht, keys);
        ensure_capacity(ctx, ht, n);
        Scratch slot_of(ctx, n * 4), is_new(ctx, n * 4), rank(ctx, n * 8);
        const int never = ht->policy == QE_NULL_NEVER_MATCH;
        const unsigned g = grid_for(n, 256, ctx->num_cus * 16);
        HtDev d = dev_view(ht);
        ht_insert_kernel<<<g, 256, 0, ctx->stream>>>(hk, sel, n, never, d, slot_of.as<uint32_t>());
        QE_HIP(hipGetLastError());
        check(read_scalar(ctx, ht->overflow) == 0, QE_GPU_OVERFLOW, "device hash table overflow");
        ht_flag_kernel<<<g, 256, 0, ctx->stream>>>(slot_of.as<uint32_t>(), n, ht->gid, ht->first,
                                                   is_new.as<uint32_t>());
        QE_HIP(hipGetLastError());
        const uint64_t fresh = exclusive_scan_u32(ctx, is_new.as<uint32_t>(), rank.as<uint64_t>(), n);
        // ops/hashtable.cpp:60-65: ids 0..kNoGroup-1; exhausting them is an error.
        check(ht->next_id + fresh <= uint64_t(kNoRow), QE_GPU_OVERFLOW,
              "HashTable: group id space exhausted (2^32-1)");
        ensure_id_capacity(ctx, ht, ht->next_id + fresh);
        ht_assign_kernel<<<g, 256, 0, ctx->stream>>>(hk, sel, n, slot_of.as<uint32_t>(), is_new.as<uint32_t>(),
                                                     rank.as<uint64_t>(), ht->next_id, d, ht->id_words,
                                                     ht->id_null, ht->id_cap, out_groups);
        ht_resolve_kernel<<<g, 256, 0, ctx->stream>>>(slot_of.as<uint32_t>(), n, ht->gid, out_groups);
        QE_HIP(hipGetLastError());
        ht->next_id += fresh;
        ht->occupied += fresh;
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_ht_find(qe_gpu_ctx* ctx, const qe_dev_ht* ht, const qe_dev_col* const* keys,
                             const uint32_t* sel, uint64_t n, uint32_t* out_groups) {
    QE_ENTRY(ctx, {
        invalid(ht != nullptr, "null table");
        if (n == 0) return QE_GPU_OK;
        invalid(out_groups != nullptr, "null output");
        check_rows32(n);
        const HashKeys hk = keys_of(ht, keys);
        HashTableView t{ht->state, ht->hash, ht->words, ht->nullmask, ht->cap, ht->overflow};
        ht_find_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
            hk, sel, n, ht->policy == QE_NULL_NEVER_MATCH, t, ht->gid, out_groups);
        QE_HIP(hipGetLastError());
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

uint64_t qe_gpu_ht_num_groups(const qe_dev_ht* ht) { return ht ? ht->next_id : 0; }
uint64_t qe_gpu_ht_capacity(const qe_dev_ht* ht) { return ht ? ht->cap : 0; }

qe_gpu_status qe_gpu_ht_key_words(qe_gpu_ctx* ctx, const qe_dev_ht* ht, uint32_t col, uint64_t* words,
                                  uint8_t* nulls) {
    QE_ENTRY(ctx, {
        invalid(ht != nullptr, "null table");
        invalid(col < uint32_t(ht->nkeys), "key column out of range");
        const uint64_t g = ht->next_id;
        if (g == 0) return QE_GPU_OK;
        std::vector<uint64_t> nm(g);
        if (words)
            QE_HIP(hipMemcpyAsync(words, ht->id_words + size_t(col) * ht->id_cap, g * 8, hipMemcpyDeviceToHost,
                                  ctx->stream));
        QE_HIP(hipMemcpyAsync(nm.data(), ht->id_null, g * 8, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
        if (nulls)
            for (uint64_t i = 0; i < g; ++i) nulls[i] = (nm[i] >> col) & 1ull;
    });
}

void qe_gpu_ht_free(qe_gpu_ctx* ctx, qe_dev_ht* ht) {
    if (!ctx || !ht) return;
    free_slots(ctx, ht);
    dev_free(ctx, ht->overflow);
    dev_free(ctx, ht->id_words);
    dev_free(ctx, ht->id_null);
    delete ht;
}

}  // extern "C"
