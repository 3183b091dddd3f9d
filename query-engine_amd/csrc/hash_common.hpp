// Device open-addressing hash table over canonical key tuples, shared by the
// group-by, the join build side and the hash-table API. Keys are normalised
// as ops/hash_internal.cpp:62-97 does (canonical words, NULL mask, seeded
// splitmix64 fold); equality compares the NULL mask, then non-NULL words
// (ops/hash_internal.h:63-75).
#pragma once

#include "agg_common.hpp"

namespace qe_gpu {

// Open addressing over canonical key tuples. state: 0 empty, 1 being written,
// 2 published. A lane that claims a slot writes the key and publishes it in the
// same loop iteration; lanes that find a slot being written re-read it on a
// later iteration (no lane ever waits inside a branch another lane must leave).
struct HashKeys {
    int nkeys;
    ColView key[kMaxKeys];
};

struct HashTableView {
    uint32_t* state;
    uint64_t* hash;
    uint64_t* words;     // [nkeys][cap]
    uint64_t* nullmask;  // [cap]
    uint64_t cap;        // power of two
    unsigned int* overflow;
};

__device__ inline void normalize_row(const HashKeys& hk, uint64_t r, uint64_t* w, uint64_t& nm,
                                     uint64_t& h) {
    nm = 0;
    h = kHashSeed;
    for (int j = 0; j < hk.nkeys; ++j) {
        if (!is_valid(hk.key[j], r)) {
            w[j] = kNullWord;
            nm |= (1ull << j);
        } else {
            w[j] = key_word(hk.key[j], r);
        }
        h = hash_combine(h, w[j]);
    }
}

__device__ inline bool slot_equal(const HashTableView& t, int nkeys, uint64_t s, const uint64_t* w,
                                  uint64_t nm, uint64_t h) {
    if (t.hash[s] != h || t.nullmask[s] != nm) return false;
    for (int j = 0; j < nkeys; ++j) {
        if ((nm >> j) & 1ull) continue;
        if (t.words[size_t(j) * t.cap + s] != w[j]) return false;
    }
    return true;
}

// Insert-or-find; returns the slot, or kNoSlot after setting the overflow flag.
// Every loop is bounded: a full sweep of the table, or (for a slot seen in the
// "being written" state) a re-read budget far above any real publish delay;
// exhausting either reports overflow instead of waiting on.
constexpr uint32_t kMaxRereads = 1u << 22;
__device__ inline uint32_t ht_insert(const HashTableView& t, int nkeys, const uint64_t* w, uint64_t nm,
                                     uint64_t h) {
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
                    __hip_atomic_store(&t.state[s], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    return static_cast<uint32_t>(s);
                }
            }
            continue;  // lost the race or slot being published: re-read this slot
        }
        if (slot_equal(t, nkeys, s, w, nm, h)) return static_cast<uint32_t>(s);
        s = (s + 1) & mask;
        rereads = 0;
        if (++probes >= t.cap) {
            atomicOr(t.overflow, 1u);
            return kNoSlot;
        }
    }
}

// Lookup without insert: the slot holding the tuple, or kNoSlot.
__device__ inline uint32_t ht_find(const HashTableView& t, int nkeys, const uint64_t* w, uint64_t nm,
                                   uint64_t h) {
    const uint64_t mask = t.cap - 1;
    uint64_t s = h & mask;
    for (uint64_t probes = 0; probes < t.cap; ++probes) {
        const uint32_t st = __hip_atomic_load(&t.state[s], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (st == 0) return kNoSlot;
        if (st == 2 && slot_equal(t, nkeys, s, w, nm, h)) return static_cast<uint32_t>(s);
        s = (s + 1) & mask;
    }
    return kNoSlot;
}

// Allocation of a table with capacity >= 2*n (power of two, >= 1024).
struct DeviceHashTable {
    Scratch state, hash, words, nullmask, overflow;
    HashTableView v{};
    int nkeys = 0;
    void create(qe_gpu_ctx* ctx, int nk, uint64_t n) {
        uint64_t cap = 1024;
        while (cap < 2 * n) cap <<= 1;
        check(cap <= (uint64_t(1) << 31), QE_GPU_UNSUPPORTED, "input too large for one device hash table");
        nkeys = nk;
        state = Scratch(ctx, cap * 4);
        hash = Scratch(ctx, cap * 8);
        words = Scratch(ctx, cap * 8 * (nk ? nk : 1));
        nullmask = Scratch(ctx, cap * 8);
        overflow = Scratch(ctx, 8);
        memset_async(ctx, state.p, 0, cap * 4);
        memset_async(ctx, overflow.p, 0, 8);
        v = HashTableView{state.as<uint32_t>(), hash.as<uint64_t>(), words.as<uint64_t>(),
                          nullmask.as<uint64_t>(), cap, overflow.as<unsigned int>()};
    }
    void check_overflow(qe_gpu_ctx* ctx) {
        check(read_scalar(ctx, overflow.as<unsigned int>()) == 0, QE_GPU_OVERFLOW, "device hash table overflow");
    }
};

}  // namespace qe_gpu
