// qe_baseline — morsel-parallel CPU restatement timed beside the GPU.
//
// TEST INFRASTRUCTURE ONLY (the bench's cpu_baseline leg and tests load it).
// Restates the reference's morsel-parallel driver, exec/parallel.cpp:431-599:
// an atomic morsel cursor over the driving input (exec/parallel.cpp:452-480,
// default 1024-row morsels, exec/parallel.h:58), per-worker private partial
// aggregates folded with plain combine (:282-317), AVG split into SUM+COUNT
// (:189-246), and a cross-worker merge after join() (:485-543).
//
// Deviations, stated with every result that uses them:
//  * join build side is built ONCE and shared read-only; the reference rebuilds
//    it per morsel (exec/parallel.cpp:101-105), infeasible at 2^24 build rows;
//  * per-thread probe scratch (the reference find() is not re-entrant);
//  * matched pairs are folded straight into the worker's partial map instead of
//    being materialised into 2048-row join batches first (this only makes the
//    CPU baseline faster);
//  * as-of: the reference has no parallel as-of (exec/parallel.h:37-40); the
//    parallel leg here is an upper_bound probe over the sorted right side.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

extern "C" {

// Dense single-key I64 group-by of F64 values with validity:
// COUNT(*), SUM(v), AVG(v) (the 1M-row group-by's query shape).
// out arrays are sized by the caller to max_groups; returns #groups or -1.
int64_t orc_bl_groupby_i64_f64(const int64_t* key, const double* val,
                               const uint64_t* val_validity, uint64_t n,
                               uint32_t threads, uint64_t morsel_rows,
                               uint64_t max_groups, int64_t* out_key, int64_t* out_cnt,
                               double* out_sum, int64_t* out_nonnull) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    if (morsel_rows == 0) morsel_rows = 1024;
    struct Acc {
        int64_t cnt = 0, nn = 0;
        double sum = 0;
    };
    std::vector<std::unordered_map<int64_t, Acc>> part(threads);
    std::atomic<uint64_t> cursor{0};
    const uint64_t nm = (n + morsel_rows - 1) / morsel_rows;
    auto work = [&](unsigned w) {
        auto& mp = part[w];
        uint64_t m;
        while ((m = cursor.fetch_add(1)) < nm) {
            const uint64_t b = m * morsel_rows, e = std::min(n, b + morsel_rows);
            for (uint64_t r = b; r < e; ++r) {
                Acc& a = mp[key[r]];
                ++a.cnt;
                const bool v = !val_validity || ((val_validity[r >> 6] >> (r & 63)) & 1);
                if (v) {
                    a.sum += val[r];
                    ++a.nn;
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < threads; ++w) pool.emplace_back(work, w);
    work(0);
    for (auto& t : pool) t.join();
    std::unordered_map<int64_t, Acc> merged;
    for (auto& mp : part)
        for (auto& kv : mp) {
            Acc& d = merged[kv.first];
            d.cnt += kv.second.cnt;
            d.nn += kv.second.nn;
            d.sum += kv.second.sum;
        }
    if (merged.size() > max_groups) return -1;
    int64_t g = 0;
    for (auto& kv : merged) {
        out_key[g] = kv.first;
        out_cnt[g] = kv.second.cnt;
        out_sum[g] = kv.second.sum;
        out_nonnull[g] = kv.second.nn;
        ++g;
    }
    return g;
}

// Inner join on a unique-or-duplicate I64 build key, then GROUP BY a build-side
// I32 column: SUM(probe F64 payload), COUNT(*) (the join -> group-by workload's query shape).
// Build: open-addressing table key -> build row list head (shared read-only).
int64_t orc_bl_join_groupby(const int64_t* pkey, const double* ppay, uint64_t np,
                            const int64_t* bkey, const int32_t* bgrp, uint64_t nb,
                            uint32_t threads, uint64_t morsel_rows, uint64_t max_groups,
                            int32_t* out_g, double* out_sum, int64_t* out_cnt) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    if (morsel_rows == 0) morsel_rows = 1024;
    auto mix = [](uint64_t x) {
        x ^= x >> 30;
        x *= 0xbf58476d1ce4e5b9ull;
        x ^= x >> 27;
        x *= 0x94d049bb133111ebull;
        x ^= x >> 31;
        return x;
    };
    // Build (single-threaded, like the reference build_side): linear probing,
    // capacity the next power of two above nb / 0.7 (ops/hashtable.h:83-94).
    uint64_t cap = 8;
    while (static_cast<double>(cap) * 0.7 < static_cast<double>(nb) + 1) cap <<= 1;
    const uint64_t mask = cap - 1;
    std::vector<uint32_t> slot_head(cap, UINT32_MAX);
    std::vector<uint32_t> next_row(nb, UINT32_MAX);
    for (uint64_t r = 0; r < nb; ++r) {
        uint64_t s = mix(0x51E57A11AB1E5EEDull ^ mix(static_cast<uint64_t>(bkey[r]))) & mask;
        while (slot_head[s] != UINT32_MAX && bkey[slot_head[s]] != bkey[r]) s = (s + 1) & mask;
        if (slot_head[s] == UINT32_MAX) {
            slot_head[s] = static_cast<uint32_t>(r);
        } else {  // duplicate key: append to the group's row list (kept in order)
            uint32_t t = slot_head[s];
            while (next_row[t] != UINT32_MAX) t = next_row[t];
            next_row[t] = static_cast<uint32_t>(r);
        }
    }
    struct Acc {
        double sum = 0;
        int64_t cnt = 0;
    };
    std::vector<std::unordered_map<int32_t, Acc>> part(threads);
    std::atomic<uint64_t> cursor{0};
    const uint64_t nm = (np + morsel_rows - 1) / morsel_rows;
    auto work = [&](unsigned w) {
        auto& mp = part[w];
        uint64_t m;
        while ((m = cursor.fetch_add(1)) < nm) {
            const uint64_t b = m * morsel_rows, e = std::min(np, b + morsel_rows);
            for (uint64_t r = b; r < e; ++r) {
                const uint64_t k = static_cast<uint64_t>(pkey[r]);
                uint64_t s = mix(0x51E57A11AB1E5EEDull ^ mix(k)) & mask;
                while (slot_head[s] != UINT32_MAX &&
                       static_cast<uint64_t>(bkey[slot_head[s]]) != k)
                    s = (s + 1) & mask;
                for (uint32_t br = slot_head[s]; br != UINT32_MAX; br = next_row[br]) {
                    Acc& a = mp[bgrp[br]];
                    a.sum += ppay[r];
                    ++a.cnt;
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < threads; ++w) pool.emplace_back(work, w);
    work(0);
    for (auto& t : pool) t.join();
    std::unordered_map<int32_t, Acc> merged;
    for (auto& mp : part)
        for (auto& kv : mp) {
            Acc& d = merged[kv.first];
            d.sum += kv.second.sum;
            d.cnt += kv.second.cnt;
        }
    if (merged.size() > max_groups) return -1;
    int64_t g = 0;
    for (auto& kv : merged) {
        out_g[g] = kv.first;
        out_sum[g] = kv.second.sum;
        out_cnt[g] = kv.second.cnt;
        ++g;
    }
    return g;
}

// Parallel as-of probe over a right side already sorted by (sym, ts) with CSR
// offsets seg[sym]..seg[sym+1]: match = upper_bound(ts, t) - 1 within the
// segment (the greatest tb <= t), reported as the SORTED right index, or
// UINT32_MAX for no match. Syms outside [0, nsym) never match.
void orc_bl_asof_upper_bound(const int64_t* sorted_ts, const uint64_t* seg, uint32_t nsym,
                             const int32_t* psym, const int64_t* pts, uint64_t np,
                             uint32_t threads, uint32_t* out) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> cursor{0};
    const uint64_t chunk = 1 << 16;
    auto work = [&]() {
        uint64_t c;
        while ((c = cursor.fetch_add(chunk)) < np) {
            const uint64_t e = std::min(np, c + chunk);
            for (uint64_t i = c; i < e; ++i) {
                const int32_t s = psym[i];
                if (s < 0 || static_cast<uint32_t>(s) >= nsym) {
                    out[i] = UINT32_MAX;
                    continue;
                }
                const int64_t* b = sorted_ts + seg[s];
                const int64_t* en = sorted_ts + seg[s + 1];
                const int64_t* u = std::upper_bound(b, en, pts[i]);
                out[i] = (u == b) ? UINT32_MAX : static_cast<uint32_t>(u - sorted_ts - 1);
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < threads; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

uint32_t orc_bl_hw_threads() { return std::max(1u, std::thread::hardware_concurrency()); }

}  // extern "C"
