// qe_oracle — CPU restatement of the reference engine's hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker the GPU path is
// compared against (and the "port" CPU baseline timed beside it). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product
// library (libqe_gpu.so) never links or calls it.
//
// Every function restates a reference algorithm from its documented behaviour
// and cites the reference file:line it follows. Parity pins (what checks each
// function against something other than itself):
//   * hashtable (Oht)       : tests/test_oracle_golden.py literal cases from
//                             tests/hashtable_test.cpp:32-189 and
//                             tests/catalog_checks_hash.cpp:35-53, an
//                             independent Python dict model, and the compiled
//                             reference (oracle/_ref) where it builds.
//   * aggregate             : literal cases of tests/aggregate_differential_test.cpp
//                             :84-237, a brute-force Python group-by, _ref.
//   * join                  : literal cases of tests/join_differential_test.cpp
//                             :103-296, brute-force nested loop, _ref.
//   * asof                  : literal cases of tests/asof_differential_test.cpp
//                             :118-318 and tests/catalog_checks_asof.cpp:38-92,
//                             brute force O(n*m), numpy searchsorted, _ref.
//   * sort_perm             : numpy lexsort (stable) on random keys.
//   * filter_select         : literal cases of tests/ops_test.cpp:99-146 and
//                             tests/expr_null_truthtable_test.cpp, an
//                             independent numpy model, _ref (Filter over
//                             expr::cmp / logic_and).
//   * morsel driver         : equality with the single-thread path.
//
// Mutants (the mutation catalog) are planted HERE, on the CPU, selected by an
// integer flag; the GPU is never given a broken kernel. A mutant oracle must
// disagree with the GPU result where the clean oracle agrees.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <vector>

extern "C" {

// Mirrors the reference Column view (core/column.h:55-67): type, length, data,
// LSB-first u64 validity words, all_valid authoritative (core/column.h:42-54).
// STR columns carry their dictionary as the reference's StringDict lays it out
// (core/string_dict.h:21-23): all strings concatenated, dict_n + 1 offsets.
typedef struct {
    int32_t type;  // 0 I32, 1 I64, 2 F64, 3 BOOL, 4 TS, 5 STR (core/types.h:18)
    int32_t all_valid;
    uint64_t len;
    const void* data;
    const uint64_t* validity;
    const char* dict_bytes;
    const uint32_t* dict_off;
    uint64_t dict_n;
} orc_col;

}  // extern "C"

namespace {

enum Ty { I32 = 0, I64 = 1, F64 = 2, BOOL = 3, TS = 4, STR = 5 };
enum Policy { kEqual = 0, kNeverMatch = 1 };
enum Agg { CountStar = 0, Count = 1, Sum = 2, Min = 3, Max = 4, Avg = 5 };

constexpr uint32_t kNoGroup = UINT32_MAX;               // ops/hashtable.h:58
constexpr uint64_t kNullWord = 0x9E3779B97F4A7C15ull;   // ops/hash_internal.h:36
constexpr uint64_t kCanonNaN = 0x7ff8000000000000ull;   // ops/hash_internal.cpp:19

size_t width(int t) { return (t == I32 || t == STR) ? 4 : (t == BOOL ? 1 : 8); }

bool valid_at(const orc_col& c, uint64_t r) {
    if (c.all_valid || c.validity == nullptr) return true;
    return (c.validity[r >> 6] >> (r & 63)) & 1ull;
}

// splitmix64 finalizer, ops/hash_kernels.h:7-12 (documented constants).
inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Canonical key word, ops/hash_internal.cpp:29-60: I32/STR zero-extended,
// I64/TS bitwise, BOOL byte, F64 with NaN -> canonical NaN and +-0 -> 0.
uint64_t key_word(const orc_col& c, uint64_t r) {
    const char* p = static_cast<const char*>(c.data) + r * width(c.type);
    switch (c.type) {
        case I32:
        case STR: {
            uint32_t v;
            std::memcpy(&v, p, 4);
            return v;
        }
        case I64:
        case TS: {
            uint64_t v;
            std::memcpy(&v, p, 8);
            return v;
        }
        case BOOL:
            return static_cast<uint8_t>(*p);
        case F64: {
            double d;
            std::memcpy(&d, p, 8);
            if (std::isnan(d)) return kCanonNaN;
            if (d == 0.0) return 0;
            uint64_t v;
            std::memcpy(&v, &d, 8);
            return v;
        }
    }
    return 0;
}

int64_t int_at(const orc_col& c, uint64_t r) {
    const char* p = static_cast<const char*>(c.data);
    switch (c.type) {
        case I32:
        case STR: {
            int32_t v;
            std::memcpy(&v, p + r * 4, 4);
            return v;
        }
        case BOOL:
            return static_cast<uint8_t>(p[r]);
        default: {
            int64_t v;
            std::memcpy(&v, p + r * 8, 8);
            return v;
        }
    }
}

double f64_at(const orc_col& c, uint64_t r) {
    double d;
    std::memcpy(&d, static_cast<const char*>(c.data) + r * 8, 8);
    return d;
}

// The bytes of a STR code (StringDict::at, core/string_dict.h:54-61).
std::string_view str_of_code(const orc_col& c, int64_t code) {
    if (c.dict_off == nullptr || code < 0 || static_cast<uint64_t>(code) >= c.dict_n)
        throw std::out_of_range("STR code outside its dictionary");
    const uint32_t lo = c.dict_off[code], hi = c.dict_off[code + 1];
    return std::string_view(c.dict_bytes + lo, hi - lo);
}

// A column this library owns (canonicalised keys, gathered join output).
struct OwnedCol {
    int32_t type = I32;
    std::vector<char> data;
    std::vector<uint64_t> validity;
    bool all_valid = true;
    const char* dict_bytes = nullptr;
    const uint32_t* dict_off = nullptr;
    uint64_t dict_n = 0;
    orc_col view() const {
        orc_col c;
        c.type = type;
        c.all_valid = all_valid;
        c.len = data.size() / width(type);
        c.data = data.data();
        c.validity = all_valid ? nullptr : validity.data();
        c.dict_bytes = dict_bytes;
        c.dict_off = dict_off;
        c.dict_n = dict_n;
        return c;
    }
};

// STR key canonicalisation (ops/join.cpp:48-71, ops/aggregate.cpp:164-187): the
// codes of a STR key column become I32 value ids interned by string VALUE in
// row order into `ids` (shared by both join sides), so equal strings from
// different dictionaries get one id; NULL rows stay NULL.
OwnedCol canon_str(const orc_col& c, uint64_t n, std::unordered_map<std::string, int32_t>& ids) {
    OwnedCol o;
    o.type = I32;
    o.data.assign(n * 4, 0);
    o.validity.assign((n + 63) / 64, ~0ull);
    for (uint64_t r = 0; r < n; ++r) {
        if (!valid_at(c, r)) {
            o.all_valid = false;
            o.validity[r >> 6] &= ~(1ull << (r & 63));
            continue;
        }
        auto it = ids.emplace(std::string(str_of_code(c, int_at(c, r))), static_cast<int32_t>(ids.size())).first;
        std::memcpy(o.data.data() + r * 4, &it->second, 4);
    }
    return o;
}

// ---------------------------------------------------------------------------
// Hash table: ops/hashtable.cpp:35-163 (linear probing, power-of-two capacity,
// first-seen dense ids, kNeverMatch dead ids, grow-after-insert doubling).
// Mutants: 1 = probe stops one slot early on a collision
// (ops/hashtable_mutants.h:25-34 kFindStopEarly); 2 = growth drops one group
// when re-slotting (kSkipRehashOne).
struct Oht {
    std::vector<int32_t> types;
    int policy;
    uint64_t seed;
    double lf;
    size_t cap, mask, occupied = 0, threshold;
    uint32_t next_id = 0;
    std::vector<uint32_t> slots;
    std::vector<std::vector<uint64_t>> words;  // [col][group]
    std::vector<uint64_t> null_mask, hashes;   // [group]
    int mutant = 0;

    Oht(const int32_t* t, uint32_t ncols, int pol, uint64_t init_cap, double max_lf,
        uint64_t sd)
        : types(t, t + ncols), policy(pol), seed(sd) {
        // ops/hashtable.cpp:43-52 normalisation of the growth policy.
        double l = max_lf;
        if (!(l > 0.0)) l = 0.7;
        if (l > 0.95) l = 0.95;
        lf = l;
        cap = 8;
        while (cap < init_cap) cap <<= 1;
        mask = cap - 1;
        threshold = static_cast<size_t>(static_cast<double>(cap) * lf);
        slots.assign(cap, kNoGroup);
        words.resize(ncols);
    }

    // ops/hash_internal.cpp:62-97: canonical words, null mask, seeded hash.
    void normalize(const orc_col* cols, const uint32_t* sel, uint64_t k,
                   std::vector<uint64_t>& w, uint64_t& nm, uint64_t& h) const {
        const uint64_t phys = sel ? sel[k] : k;
        nm = 0;
        h = seed;
        for (size_t j = 0; j < types.size(); ++j) {
            if (!valid_at(cols[j], phys)) {
                w[j] = kNullWord;
                nm |= (1ull << j);
            } else {
                w[j] = key_word(cols[j], phys);
            }
            h = mix64(h ^ mix64(w[j]));  // ops/hash_scalar.cpp:24-29
        }
    }

    // ops/hash_internal.h:63-75
    bool equal(const std::vector<uint64_t>& w, uint64_t nm, uint32_t g) const {
        if (nm != null_mask[g]) return false;
        for (size_t j = 0; j < types.size(); ++j) {
            if ((nm >> j) & 1ull) continue;
            if (w[j] != words[j][g]) return false;
        }
        return true;
    }

    uint32_t new_group(const std::vector<uint64_t>& w, uint64_t nm, uint64_t h) {
        if (next_id == kNoGroup) throw std::overflow_error("group id space exhausted");
        const uint32_t g = next_id++;
        for (size_t j = 0; j < types.size(); ++j) words[j].push_back(w[j]);
        null_mask.push_back(nm);
        hashes.push_back(h);
        return g;
    }

    void grow() {  // ops/hashtable.cpp:75-91
        const size_t ncap = cap * 2, nmask = ncap - 1;
        std::vector<uint32_t> ns(ncap, kNoGroup);
        bool skipped = false;
        for (size_t s = 0; s < cap; ++s) {
            const uint32_t g = slots[s];
            if (g == kNoGroup) continue;
            if (mutant == 2 && !skipped) {  // planted: one group lost on rehash
                skipped = true;
                continue;
            }
            size_t slot = hashes[g] & nmask;
            while (ns[slot] != kNoGroup) slot = (slot + 1) & nmask;
            ns[slot] = g;
        }
        slots.swap(ns);
        cap = ncap;
        mask = nmask;
        threshold = static_cast<size_t>(static_cast<double>(cap) * lf);
    }

    uint32_t lookup(const std::vector<uint64_t>& w, uint64_t nm, uint64_t h,
                    size_t& slot_out) const {
        size_t slot = h & mask;
        while (slots[slot] != kNoGroup) {
            const uint32_t g = slots[slot];
            if (hashes[g] == h && equal(w, nm, g)) {
                slot_out = slot;
                return g;
            }
            if (mutant == 1 && slots[(slot + 1) & mask] != kNoGroup &&
                (slot & 1)) {  // planted: gives up early on odd slots in a chain
                slot_out = slot;
                return kNoGroup;
            }
            slot = (slot + 1) & mask;
        }
        slot_out = slot;
        return kNoGroup;
    }

    void insert_or_find(const orc_col* cols, const uint32_t* sel, uint64_t n,
                        uint32_t* out) {  // ops/hashtable.cpp:93-133
        std::vector<uint64_t> w(types.size());
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t nm, h;
            normalize(cols, sel, k, w, nm, h);
            if (policy == kNeverMatch && nm != 0) {
                out[k] = new_group(w, nm, h);  // dead, unslotted id
                continue;
            }
            size_t slot;
            uint32_t g = lookup(w, nm, h, slot);
            if (g == kNoGroup) {
                if (slots[slot] != kNoGroup) {  // mutant early exit: find a free slot
                    while (slots[slot] != kNoGroup) slot = (slot + 1) & mask;
                }
                g = new_group(w, nm, h);
                slots[slot] = g;
                ++occupied;
                if (occupied > threshold) grow();
            }
            out[k] = g;
        }
    }

    void find(const orc_col* cols, const uint32_t* sel, uint64_t n,
              uint32_t* out) const {  // ops/hashtable.cpp:135-163
        std::vector<uint64_t> w(types.size());
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t nm, h;
            normalize(cols, sel, k, w, nm, h);
            if (policy == kNeverMatch && nm != 0) {
                out[k] = kNoGroup;
                continue;
            }
            size_t slot;
            out[k] = lookup(w, nm, h, slot);
        }
    }
};

// ---------------------------------------------------------------------------
// Aggregate: ops/aggregate.cpp:118-155 (scatter_row), ops/agg_internal.h:34-81
// (NaN-greatest total order, identities), :169-229 (finalisation),
// ops/aggregate.cpp:78-97 (result types).
// Sum modes: 0 = naive sequential fold in row order (the reference's bits);
// 1 = Neumaier-compensated (the accurate reference the 1e-12 bar is measured
// against). Mutants: 1 = NULL folded into SUM as 0 and counted
// (aggregate_mutants kFoldNullInSum); 2 = empty group emits 0 instead of NULL
// (kEmptyGroupZero); 3 = MAX drops NaN (kMaxDropsNan); 4 = the last row of the
// input is not aggregated (a tail off-by-one); 5 = STR keys grouped by raw code
// instead of by value.

bool f64_less_total(double a, double b) {  // ops/agg_internal.h:34-38
    if (std::isnan(b)) return !std::isnan(a);
    if (std::isnan(a)) return false;
    return a < b;
}

struct Cell {
    int64_t i = 0;
    double d = 0.0, comp = 0.0;  // comp: Neumaier compensation term
    int64_t cnt = 0;
};

Cell init_cell(int func) {  // ops/agg_internal.h:65-81
    Cell c;
    if (func == Min) {
        c.i = INT64_MAX;
        c.d = std::numeric_limits<double>::quiet_NaN();
    } else if (func == Max) {
        c.i = INT64_MIN;
        c.d = -std::numeric_limits<double>::infinity();
    }
    return c;
}

void neumaier_add(Cell& c, double v) {
    const double t = c.d + v;
    if (std::fabs(c.d) >= std::fabs(v))
        c.comp += (c.d - t) + v;
    else
        c.comp += (v - t) + c.d;
    c.d = t;
}

void scatter(Cell& c, int func, bool is_float, bool valid, int64_t iv, double dv,
             int sum_mode, int mutant) {
    switch (func) {
        case CountStar:
            ++c.cnt;
            return;
        case Count:
            if (valid) ++c.cnt;
            return;
        case Sum:
        case Avg:
            if (!valid) {
                if (mutant == 1) ++c.cnt;  // planted: NULL counted as a 0 addend
                return;
            }
            if (is_float) {
                if (sum_mode == 1) neumaier_add(c, dv);
                else c.d += dv;
            } else {
                c.i = static_cast<int64_t>(static_cast<uint64_t>(c.i) +
                                           static_cast<uint64_t>(iv));
            }
            ++c.cnt;
            return;
        case Min:
            if (!valid) return;
            if (is_float) {
                if (f64_less_total(dv, c.d)) c.d = dv;
            } else {
                c.i = std::min(c.i, iv);
            }
            ++c.cnt;
            return;
        case Max:
            if (!valid) return;
            if (is_float) {
                if (mutant == 3) {
                    if (!std::isnan(dv) && (std::isnan(c.d) || dv > c.d || c.cnt == 0))
                        c.d = dv;
                } else if (f64_less_total(c.d, dv)) {
                    c.d = dv;
                }
            } else {
                c.i = std::max(c.i, iv);
            }
            ++c.cnt;
            return;
    }
}

struct AggResult {
    uint32_t nkeys = 0, naggs = 0;
    uint64_t ngroups = 0;
    std::vector<std::vector<uint64_t>> key_words;  // [key][group]
    std::vector<std::vector<uint8_t>> key_null;
    std::vector<int32_t> out_is_float;             // [agg]
    std::vector<std::vector<int64_t>> out_i;       // [agg][group]
    std::vector<std::vector<double>> out_d;
    std::vector<std::vector<uint8_t>> out_null;
};

// Core grouped/global aggregation over whole columns (row order == the
// reference's batch-by-batch order, since batches are contiguous windows).
AggResult* aggregate_impl(const orc_col* cols, uint64_t nrows, const uint32_t* key_cols,
                          uint32_t nkeys, const int32_t* funcs, const uint32_t* inputs,
                          uint32_t naggs, int sum_mode, int mutant) {
    auto* res = new AggResult();
    res->nkeys = nkeys;
    res->naggs = naggs;
    std::vector<int> isf(naggs);
    for (uint32_t a = 0; a < naggs; ++a)
        isf[a] = (funcs[a] != CountStar && funcs[a] != Count &&
                  cols[inputs[a]].type == F64);
    const uint64_t nrun = (mutant == 4 && nrows > 0) ? nrows - 1 : nrows;

    std::vector<std::vector<Cell>> cells(naggs);
    std::vector<uint32_t> gid(nrows, 0);
    Oht* ht = nullptr;
    std::vector<orc_col> kc;
    std::vector<OwnedCol> canon(nkeys);
    if (nkeys > 0) {
        std::vector<int32_t> kt;
        for (uint32_t j = 0; j < nkeys; ++j) {
            kc.push_back(cols[key_cols[j]]);
            if (kc.back().type == STR) {  // group STR keys by value (mutant 5: by raw code)
                if (mutant != 5) {
                    std::unordered_map<std::string, int32_t> ids;
                    canon[j] = canon_str(kc.back(), nrows, ids);
                    kc.back() = canon[j].view();
                } else {
                    kc.back().type = I32;
                }
            }
            kt.push_back(kc.back().type);
        }
        ht = new Oht(kt.data(), nkeys, kEqual, 1024, 0.7, 0x51E57A11AB1E5EEDull);
        if (nrows) ht->insert_or_find(kc.data(), nullptr, nrows, gid.data());
        res->ngroups = ht->next_id;
    } else {
        res->ngroups = 1;  // global: exactly one row, even over empty input
    }
    for (uint32_t a = 0; a < naggs; ++a)
        cells[a].assign(res->ngroups, init_cell(funcs[a]));
    for (uint32_t a = 0; a < naggs; ++a) {
        const int f = funcs[a];
        const orc_col* in = (f == CountStar) ? nullptr : &cols[inputs[a]];
        for (uint64_t r = 0; r < nrun; ++r) {
            bool v = true;
            int64_t iv = 0;
            double dv = 0;
            if (in) {
                v = valid_at(*in, r);
                if (v) {
                    if (isf[a]) dv = f64_at(*in, r);
                    else iv = int_at(*in, r);
                }
            }
            if (in && in->type == STR && (f == Min || f == Max)) {
                // ops/aggregate.cpp:189-205: the extremum by lexicographic byte
                // value; the cell keeps the winning row's code.
                Cell& c = cells[a][gid[r]];
                if (!v) continue;
                if (c.cnt == 0) {
                    c.i = iv;
                } else {
                    const std::string_view sv = str_of_code(*in, iv), cur = str_of_code(*in, c.i);
                    if (f == Min ? sv < cur : sv > cur) c.i = iv;
                }
                ++c.cnt;
                continue;
            }
            scatter(cells[a][gid[r]], f, isf[a], v, iv, dv, sum_mode, mutant);
        }
    }
    res->key_words.assign(nkeys, std::vector<uint64_t>(res->ngroups));
    res->key_null.assign(nkeys, std::vector<uint8_t>(res->ngroups));
    for (uint32_t j = 0; j < nkeys; ++j)
        for (uint64_t g = 0; g < res->ngroups; ++g) {
            res->key_words[j][g] = ht->words[j][g];
            res->key_null[j][g] = (ht->null_mask[g] >> j) & 1ull;
        }
    // A STR key reads back as the source code of its group's first row (a code of
    // the input's dictionary with the group's value).
    std::vector<uint64_t> first_row;
    for (uint32_t j = 0; j < nkeys; ++j) {
        if (cols[key_cols[j]].type != STR) continue;
        if (first_row.empty()) {
            first_row.assign(res->ngroups, UINT64_MAX);
            for (uint64_t r = 0; r < nrows; ++r)
                if (first_row[gid[r]] == UINT64_MAX) first_row[gid[r]] = r;
        }
        for (uint64_t g = 0; g < res->ngroups; ++g)
            if (!res->key_null[j][g] && first_row[g] != UINT64_MAX)
                res->key_words[j][g] = static_cast<uint32_t>(int_at(cols[key_cols[j]], first_row[g]));
    }
    res->out_is_float.resize(naggs);
    res->out_i.assign(naggs, std::vector<int64_t>(res->ngroups, 0));
    res->out_d.assign(naggs, std::vector<double>(res->ngroups, 0.0));
    res->out_null.assign(naggs, std::vector<uint8_t>(res->ngroups, 0));
    for (uint32_t a = 0; a < naggs; ++a) {
        const int f = funcs[a];
        // ops/aggregate.cpp:78-97: COUNT -> I64, SUM(F64)/AVG -> F64, MIN/MAX keep type.
        const bool float_out = (f == Avg) || ((f == Sum || f == Min || f == Max) && isf[a]);
        res->out_is_float[a] = float_out;
        for (uint64_t g = 0; g < res->ngroups; ++g) {
            const Cell& c = cells[a][g];
            if (f == CountStar || f == Count) {
                res->out_i[a][g] = c.cnt;
                continue;
            }
            if (c.cnt == 0) {  // ops/agg_internal.h:186-229: no non-null input -> NULL
                if (mutant == 2) {
                    res->out_i[a][g] = 0;
                    res->out_d[a][g] = 0.0;
                } else {
                    res->out_null[a][g] = 1;
                }
                continue;
            }
            // A non-finite running sum is already the row-order fold's answer
            // (+/-inf, or NaN from a NaN or from both infinities); the
            // compensation term is then NaN (inf - inf) and is dropped.
            const double dsum = (sum_mode == 1 && std::isfinite(c.d)) ? c.d + c.comp : c.d;
            if (f == Avg)
                res->out_d[a][g] = (isf[a] ? dsum : static_cast<double>(c.i)) /
                                   static_cast<double>(c.cnt);
            else if (float_out)
                res->out_d[a][g] = (f == Sum) ? dsum : c.d;
            else
                res->out_i[a][g] = c.i;
        }
    }
    delete ht;
    return res;
}

// ---------------------------------------------------------------------------
// Hash join: ops/join.cpp:133-247 (build into kNeverMatch table + per-group
// build-row lists in build row order; probe rows in order; each match fans out
// to every build row of its group; LEFT emits kNullBuildRow for a miss).
// Mutants: 1 = the last build row of every multi-row group is dropped
// (join_mutants kFanoutTailOffByOne); 2 = the first matched probe row is
// dropped (kDropProbeMatch); 3 = composite keys compare only the first
// column (kCompositeFirstKeyOnly); 4 = STR keys compared by raw dictionary code
// (string_key_mutants kHashRawCode).
struct JoinResult {
    std::vector<uint32_t> probe_rows, build_rows;
};

JoinResult* join_impl(const orc_col* pcols, uint64_t np, const orc_col* bcols,
                      uint64_t nb, const uint32_t* pkeys, const uint32_t* bkeys,
                      uint32_t nkeys, int type, int mutant) {
    auto* res = new JoinResult();
    const uint32_t nk = (mutant == 3) ? 1 : nkeys;
    std::vector<orc_col> bk, pk;
    std::vector<int32_t> kt;
    std::vector<OwnedCol> canon(2 * nk);
    for (uint32_t j = 0; j < nk; ++j) {
        bk.push_back(bcols[bkeys[j]]);
        pk.push_back(pcols[pkeys[j]]);
        if (bk.back().type == STR && pk.back().type == STR) {
            if (mutant != 4) {  // mutant 4: STR keys by raw code (string_key_mutants kHashRawCode)
                std::unordered_map<std::string, int32_t> ids;  // one id space per key position
                canon[2 * j] = canon_str(bk.back(), nb, ids);
                canon[2 * j + 1] = canon_str(pk.back(), np, ids);
                bk.back() = canon[2 * j].view();
                pk.back() = canon[2 * j + 1].view();
            } else {
                bk.back().type = I32;
                pk.back().type = I32;
            }
        }
        kt.push_back(bk.back().type);
    }
    Oht ht(kt.data(), nk, kNeverMatch, 1024, 0.7, 0x51E57A11AB1E5EEDull);
    std::vector<uint32_t> bg(nb), pg(np);
    if (nb) ht.insert_or_find(bk.data(), nullptr, nb, bg.data());
    std::vector<std::vector<uint32_t>> rows(ht.next_id);
    for (uint64_t r = 0; r < nb; ++r) rows[bg[r]].push_back(static_cast<uint32_t>(r));
    if (np) ht.find(pk.data(), nullptr, np, pg.data());
    uint64_t matched = 0;
    for (uint64_t r = 0; r < np; ++r) {
        const uint32_t g = pg[r];
        if (g != kNoGroup) {
            ++matched;
            if (mutant == 2 && matched == 1) continue;  // planted: first matched probe row dropped
            const auto& lst = rows[g];
            size_t m = lst.size();
            if (mutant == 1 && m > 1) --m;
            for (size_t i = 0; i < m; ++i) {
                res->probe_rows.push_back(static_cast<uint32_t>(r));
                res->build_rows.push_back(lst[i]);
            }
        } else if (type == 1) {
            res->probe_rows.push_back(static_cast<uint32_t>(r));
            res->build_rows.push_back(kNoGroup);  // kNullBuildRow, ops/join_internal.h:31
        }
    }
    return res;
}

// Gather the join-output column `c` (probe cols ++ build cols) for the pair list
// into an owned column (ops/join_internal.h:136-179 emit semantics); a STR
// column keeps its codes and dictionary.
OwnedCol gather_col(const orc_col& src, const std::vector<uint32_t>& idx) {
    OwnedCol o;
    o.type = src.type;
    o.dict_bytes = src.dict_bytes;
    o.dict_off = src.dict_off;
    o.dict_n = src.dict_n;
    const size_t w = width(src.type);
    o.data.assign(idx.size() * w, 0);
    o.validity.assign((idx.size() + 63) / 64, ~0ull);
    for (size_t r = 0; r < idx.size(); ++r) {
        const uint32_t s = idx[r];
        if (s == kNoGroup || !valid_at(src, s)) {
            o.all_valid = false;
            o.validity[r >> 6] &= ~(1ull << (r & 63));
            continue;
        }
        std::memcpy(o.data.data() + r * w, static_cast<const char*>(src.data) + s * w, w);
    }
    return o;
}

// ---------------------------------------------------------------------------
// Stable multi-key sort permutation: ops/sort.cpp:125-188 (LSD byte radix over
// [null byte][big-endian sign-flipped value] per key, key 0 most significant).
// Integer keys (I32/I64/TS) only, matching the reference radix path.
std::vector<uint32_t> sort_perm_impl(const orc_col* cols, const uint32_t* keys,
                                     const int32_t* desc, const int32_t* nulls_first,
                                     uint32_t nk, uint64_t n) {
    std::vector<size_t> vb(nk), off(nk);
    size_t W = 0;
    for (uint32_t j = 0; j < nk; ++j) {
        vb[j] = (cols[keys[j]].type == I32) ? 4 : 8;
        off[j] = W;
        W += 1 + vb[j];
    }
    std::vector<uint8_t> kb(n * W, 0);
    for (uint64_t r = 0; r < n; ++r) {
        uint8_t* row = kb.data() + r * W;
        for (uint32_t j = 0; j < nk; ++j) {
            const orc_col& c = cols[keys[j]];
            const bool isnull = !valid_at(c, r);
            row[off[j]] = nulls_first[j] ? (isnull ? 0 : 1) : (isnull ? 1 : 0);
            uint64_t u = 0;
            if (!isnull) {
                if (c.type == I32)
                    u = static_cast<uint32_t>(static_cast<int32_t>(int_at(c, r))) ^ 0x80000000u;
                else
                    u = static_cast<uint64_t>(int_at(c, r)) ^ 0x8000000000000000ull;
            }
            for (size_t b = 0; b < vb[j]; ++b) {
                uint8_t v = static_cast<uint8_t>(u >> (8 * (vb[j] - 1 - b)));
                if (desc[j]) v = static_cast<uint8_t>(~v);
                row[off[j] + 1 + b] = v;
            }
        }
    }
    std::vector<uint32_t> perm(n), tmp(n);
    std::iota(perm.begin(), perm.end(), 0u);
    for (size_t pos = W; pos-- > 0;) {
        size_t cnt[257] = {0};
        for (uint64_t i = 0; i < n; ++i) ++cnt[kb[perm[i] * W + pos] + 1];
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (uint64_t i = 0; i < n; ++i) tmp[cnt[kb[perm[i] * W + pos]]++] = perm[i];
        perm.swap(tmp);
    }
    return perm;
}

// The comparison path (ops/sort.cpp:60-100): a stable sort by a row "less"
// over the keys in order; NULL placement is absolute, F64 uses the
// NaN-greatest total order with all NaNs tied and -0 == +0
// (ops/sort_internal.h:100-104), STR compares the strings, BOOL compares 0/1.
int cmp_key(const orc_col& c, uint64_t a, uint64_t b) {
    if (c.type == F64) {
        const double x = f64_at(c, a), y = f64_at(c, b);
        const bool nx = std::isnan(x), ny = std::isnan(y);
        if (nx || ny) return nx == ny ? 0 : (nx ? 1 : -1);
        return x < y ? -1 : (x > y ? 1 : 0);
    }
    if (c.type == STR) {
        const int r = str_of_code(c, int_at(c, a)).compare(str_of_code(c, int_at(c, b)));
        return r < 0 ? -1 : (r > 0 ? 1 : 0);
    }
    int64_t x = int_at(c, a), y = int_at(c, b);
    if (c.type == BOOL) {
        x = x != 0;
        y = y != 0;
    }
    return x < y ? -1 : (x > y ? 1 : 0);
}

std::vector<uint32_t> sort_perm_comparison(const orc_col* cols, const uint32_t* keys, const int32_t* desc,
                                           const int32_t* nulls_first, uint32_t nk, uint64_t n) {
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
        for (uint32_t j = 0; j < nk; ++j) {
            const orc_col& c = cols[keys[j]];
            const bool av = valid_at(c, a), bv = valid_at(c, b);
            if (!av || !bv) {
                if (!av && !bv) continue;
                return nulls_first[j] ? !av : av;
            }
            const int r = cmp_key(c, a, b);
            if (r != 0) return desc[j] ? r > 0 : r < 0;
        }
        return false;
    });
    return perm;
}

// ---------------------------------------------------------------------------
// Backward as-of join: tsx/asof.cpp:122-308. Both sides stably sorted by
// (keys ASC NULLS LAST, ts ASC NULLS LAST) (:51-59, :163-166); build rows with
// non-NULL key AND ts go into per-group lists in sorted order (:217-229);
// probes walk in sorted order with a per-group advancing cursor (:271-280),
// match = last build ts <= t (the `<=` boundary), tolerance in u64 (:281-294),
// LEFT emits kNullBuildRow. Output pairs are ORIGINAL row indices, emitted in
// the sorted-probe order of the reference.
// Mutants: 1 = strict boundary `<` (asof_mutants kBoundaryStrict);
// 2 = nearest following instead of preceding (kNearestFollowing);
// 3 = the last key column is ignored (kIgnoreLastKey); 4 = a LEFT probe row
// without a match carries build row 0 instead of NULL (kLeftWrongNull,
// tsx/asof_mutants.cpp:233-239).
int64_t time_at(const orc_col& c, uint64_t r) { return int_at(c, r); }

JoinResult* asof_impl(const orc_col* pcols, uint64_t np, const orc_col* bcols,
                      uint64_t nb, const uint32_t* pkeys, const uint32_t* bkeys,
                      uint32_t nkeys, uint32_t ptime, uint32_t btime, int type,
                      int has_tol, int64_t tol, int mutant) {
    auto* res = new JoinResult();
    const uint32_t nk = (mutant == 3 && nkeys > 0) ? nkeys - 1 : nkeys;
    std::vector<int32_t> zero(nk + 1, 0);
    // Sort keys: keys..., time; all ASC NULLS LAST.
    std::vector<uint32_t> bsk(bkeys, bkeys + nk), psk(pkeys, pkeys + nk);
    bsk.push_back(btime);
    psk.push_back(ptime);
    const auto bperm = sort_perm_impl(bcols, bsk.data(), zero.data(), zero.data(), nk + 1, nb);
    const auto pperm = sort_perm_impl(pcols, psk.data(), zero.data(), zero.data(), nk + 1, np);

    std::vector<int32_t> kt;
    std::vector<orc_col> bk, pk;
    for (uint32_t j = 0; j < nk; ++j) {
        kt.push_back(bcols[bkeys[j]].type);
        bk.push_back(bcols[bkeys[j]]);
        pk.push_back(pcols[pkeys[j]]);
    }
    std::vector<std::vector<uint32_t>> lists;  // group -> original build rows, ts-ascending
    Oht* ht = nullptr;
    if (nk > 0) {
        ht = new Oht(kt.data(), nk, kNeverMatch, 1024, 0.7, 0x51E57A11AB1E5EEDull);
        std::vector<uint32_t> g(nb);
        if (nb) ht->insert_or_find(bk.data(), bperm.data(), nb, g.data());
        lists.resize(ht->next_id);
        for (uint64_t k = 0; k < nb; ++k) {
            const uint32_t r = bperm[k];
            bool keynull = false;
            for (uint32_t j = 0; j < nk; ++j) keynull |= !valid_at(bk[j], r);
            if (!keynull && valid_at(bcols[btime], r)) lists[g[k]].push_back(r);
        }
    } else {
        lists.resize(1);
        for (uint64_t k = 0; k < nb; ++k) {
            const uint32_t r = bperm[k];
            if (valid_at(bcols[btime], r)) lists[0].push_back(r);
        }
    }
    std::vector<uint32_t> pg(np, 0);
    if (nk > 0 && np) ht->find(pk.data(), pperm.data(), np, pg.data());
    uint32_t last = kNoGroup;
    size_t cursor = 0;
    for (uint64_t k = 0; k < np; ++k) {
        const uint32_t r = pperm[k];
        uint32_t matched = kNoGroup;
        const uint32_t g = pg[k];
        bool keynull = false;
        for (uint32_t j = 0; j < nk; ++j) keynull |= !valid_at(pk[j], r);
        if (g != kNoGroup && !keynull && valid_at(pcols[ptime], r)) {
            const int64_t t = time_at(pcols[ptime], r);
            const auto& bl = lists[g];
            if (g != last) {
                last = g;
                cursor = 0;
            }
            if (mutant == 1) {
                while (cursor < bl.size() && time_at(bcols[btime], bl[cursor]) < t) ++cursor;
            } else {
                while (cursor < bl.size() && time_at(bcols[btime], bl[cursor]) <= t) ++cursor;
            }
            uint32_t cand = kNoGroup;
            if (mutant == 2) {
                // nearest following: first build ts >= t
                size_t c2 = 0;
                while (c2 < bl.size() && time_at(bcols[btime], bl[c2]) < t) ++c2;
                if (c2 < bl.size()) cand = bl[c2];
            } else if (cursor > 0) {
                cand = bl[cursor - 1];
            }
            if (cand != kNoGroup) {
                const int64_t tb = time_at(bcols[btime], cand);
                if (!has_tol || static_cast<uint64_t>(t) - static_cast<uint64_t>(tb) <=
                                    static_cast<uint64_t>(tol))
                    matched = cand;
            }
        }
        if (matched != kNoGroup) {
            res->probe_rows.push_back(r);
            res->build_rows.push_back(matched);
        } else if (type == 1) {
            res->probe_rows.push_back(r);
            res->build_rows.push_back((mutant == 4 && nb > 0) ? 0u : kNoGroup);
        }
    }
    delete ht;
    return res;
}

}  // namespace

// ===========================================================================
// C ABI (ctypes). Every call returns 0 on success, nonzero on error.
extern "C" {

void* orc_ht_new(const int32_t* types, uint32_t ncols, int32_t policy,
                 uint64_t init_cap, double lf, uint64_t seed, int32_t mutant) {
    auto* h = new Oht(types, ncols, policy, init_cap, lf, seed);
    h->mutant = mutant;
    return h;
}
void orc_ht_free(void* h) { delete static_cast<Oht*>(h); }
int orc_ht_insert_or_find(void* h, const orc_col* cols, const uint32_t* sel,
                          uint64_t n, uint32_t* out) {
    try {
        static_cast<Oht*>(h)->insert_or_find(cols, sel, n, out);
    } catch (...) {
        return 1;
    }
    return 0;
}
int orc_ht_find(void* h, const orc_col* cols, const uint32_t* sel, uint64_t n,
                uint32_t* out) {
    static_cast<Oht*>(h)->find(cols, sel, n, out);
    return 0;
}
uint64_t orc_ht_num_groups(void* h) { return static_cast<Oht*>(h)->next_id; }
uint64_t orc_ht_capacity(void* h) { return static_cast<Oht*>(h)->cap; }
uint64_t orc_ht_key_word(void* h, uint32_t g, uint32_t col) {
    return static_cast<Oht*>(h)->words[col][g];
}
int32_t orc_ht_key_is_null(void* h, uint32_t g, uint32_t col) {
    return (static_cast<Oht*>(h)->null_mask[g] >> col) & 1ull;
}
uint64_t orc_mix64(uint64_t x) { return mix64(x); }
uint64_t orc_key_word(const orc_col* c, uint64_t r) { return key_word(*c, r); }

// --- aggregate -------------------------------------------------------------
void* orc_agg_run(const orc_col* cols, uint64_t nrows, const uint32_t* key_cols,
                  uint32_t nkeys, const int32_t* funcs, const uint32_t* inputs,
                  uint32_t naggs, int32_t sum_mode, int32_t mutant) {
    return aggregate_impl(cols, nrows, key_cols, nkeys, funcs, inputs, naggs, sum_mode,
                          mutant);
}
uint64_t orc_agg_ngroups(void* h) { return static_cast<AggResult*>(h)->ngroups; }
void orc_agg_keys(void* h, uint32_t j, uint64_t* words, uint8_t* nulls) {
    auto* r = static_cast<AggResult*>(h);
    std::memcpy(words, r->key_words[j].data(), r->ngroups * 8);
    std::memcpy(nulls, r->key_null[j].data(), r->ngroups);
}
int32_t orc_agg_is_float(void* h, uint32_t a) {
    return static_cast<AggResult*>(h)->out_is_float[a];
}
void orc_agg_values(void* h, uint32_t a, int64_t* iv, double* dv, uint8_t* nulls) {
    auto* r = static_cast<AggResult*>(h);
    std::memcpy(iv, r->out_i[a].data(), r->ngroups * 8);
    std::memcpy(dv, r->out_d[a].data(), r->ngroups * 8);
    std::memcpy(nulls, r->out_null[a].data(), r->ngroups);
}
void orc_agg_free(void* h) { delete static_cast<AggResult*>(h); }

// --- join ------------------------------------------------------------------
void* orc_join_run(const orc_col* pcols, uint64_t np, const orc_col* bcols, uint64_t nb,
                   const uint32_t* pkeys, const uint32_t* bkeys, uint32_t nkeys,
                   int32_t type, int32_t mutant) {
    return join_impl(pcols, np, bcols, nb, pkeys, bkeys, nkeys, type, mutant);
}
uint64_t orc_pairs_count(void* h) { return static_cast<JoinResult*>(h)->probe_rows.size(); }
void orc_pairs_get(void* h, uint32_t* probe_rows, uint32_t* build_rows) {
    auto* r = static_cast<JoinResult*>(h);
    std::memcpy(probe_rows, r->probe_rows.data(), r->probe_rows.size() * 4);
    std::memcpy(build_rows, r->build_rows.data(), r->build_rows.size() * 4);
}
void orc_pairs_free(void* h) { delete static_cast<JoinResult*>(h); }

// Aggregate(HashJoin(probe, build)) — plan/plan.cpp:581-590 composition. Group
// keys and agg inputs index the join output (probe cols ++ build cols).
void* orc_join_agg_run(const orc_col* pcols, uint32_t npc, uint64_t np,
                       const orc_col* bcols, uint32_t nbc, uint64_t nb,
                       const uint32_t* pkeys, const uint32_t* bkeys, uint32_t nkeys,
                       int32_t type, const uint32_t* gkeys, uint32_t ngk,
                       const int32_t* funcs, const uint32_t* inputs, uint32_t naggs,
                       int32_t sum_mode, int32_t mutant) {
    JoinResult* j = join_impl(pcols, np, bcols, nb, pkeys, bkeys, nkeys, type,
                              mutant >= 10 ? mutant - 10 : 0);
    // Materialise only the columns the aggregate reads.
    std::vector<OwnedCol> owned(npc + nbc);
    std::vector<orc_col> view(npc + nbc);
    std::vector<char> need(npc + nbc, 0);
    for (uint32_t k = 0; k < ngk; ++k) need[gkeys[k]] = 1;
    for (uint32_t a = 0; a < naggs; ++a)
        if (funcs[a] != CountStar) need[inputs[a]] = 1;
    for (uint32_t c = 0; c < npc + nbc; ++c) {
        if (!need[c]) {
            view[c] = (c < npc) ? pcols[c] : bcols[c - npc];
            view[c].len = j->probe_rows.size();
            continue;
        }
        owned[c] = (c < npc) ? gather_col(pcols[c], j->probe_rows)
                             : gather_col(bcols[c - npc], j->build_rows);
        view[c] = owned[c].view();
    }
    AggResult* r = aggregate_impl(view.data(), j->probe_rows.size(), gkeys, ngk, funcs,
                                  inputs, naggs, sum_mode, mutant < 10 ? mutant : 0);
    delete j;
    return r;
}

// --- as-of -----------------------------------------------------------------
void* orc_asof_run(const orc_col* pcols, uint64_t np, const orc_col* bcols, uint64_t nb,
                   const uint32_t* pkeys, const uint32_t* bkeys, uint32_t nkeys,
                   uint32_t ptime, uint32_t btime, int32_t type, int32_t has_tol,
                   int64_t tol, int32_t mutant) {
    return asof_impl(pcols, np, bcols, nb, pkeys, bkeys, nkeys, ptime, btime, type,
                     has_tol, tol, mutant);
}

// --- sort ------------------------------------------------------------------
int orc_sort_perm(const orc_col* cols, const uint32_t* keys, const int32_t* desc,
                  const int32_t* nulls_first, uint32_t nk, uint64_t n, uint32_t* perm) {
    bool radix = true;  // the reference's kAuto choice (ops/sort.cpp:192-215)
    for (uint32_t j = 0; j < nk; ++j) {
        const int t = cols[keys[j]].type;
        if (t != I32 && t != I64 && t != TS) radix = false;
    }
    const auto p = radix ? sort_perm_impl(cols, keys, desc, nulls_first, nk, n)
                         : sort_perm_comparison(cols, keys, desc, nulls_first, nk, n);
    std::memcpy(perm, p.data(), n * 4);
    return 0;
}

}  // extern "C"

// --- filter ----------------------------------------------------------------
// Filter with a conjunction of column-versus-literal comparisons, restating
// ops/filter.cpp:28-66 (a row passes iff the predicate is valid AND true;
// passing rows keep input order) over expr::cmp (expr/expr.cpp:133-151: two
// numerics are promoted to the higher of F64 > I64 > I32, otherwise BOOL/BOOL
// and TS/TS compare in place), null-if-any comparison and IEEE NaN
// (expr/expr.h:22-40, expr/compare_scalar.cpp:1-40), and three-valued AND
// (a conjunction is TRUE only if every operand is TRUE).
//
// Mutants: 1 = a NULL comparison passes; 2 = `<=`/`>=` evaluated strictly;
// 3 = an F64 comparison done on integers (literal truncated, no promotion);
// 4 = the decision inverted (oracle_mutation_test.cpp:48-79 BuggyFilter).

extern "C" {
typedef struct {
    uint32_t col;
    int32_t op;  // 0 Lt, 1 Le, 2 Gt, 3 Ge, 4 Eq, 5 Ne (expr/expr.h:79)
    int32_t lit_type;
    int32_t lit_is_null;
    int64_t lit_i;
    double lit_f;
} orc_pred;
}

namespace {

template <typename T>
bool cmp_apply(int op, T a, T b, int mutant) {
    if (mutant == 2 && op == 1) op = 0;
    if (mutant == 2 && op == 3) op = 2;
    switch (op) {
        case 0: return a < b;
        case 1: return a <= b;
        case 2: return a > b;
        case 3: return a >= b;
        case 4: return a == b;
        default: return a != b;
    }
}

// Three-valued result of one comparison: -1 NULL, 0 FALSE, 1 TRUE.
int cmp_term(const orc_col& c, const orc_pred& t, uint64_t r, int mutant) {
    if (t.lit_is_null || !valid_at(c, r)) return -1;
    const bool as_float = (c.type == F64 || t.lit_type == F64) && mutant != 3;
    if (as_float) {
        const double a = c.type == F64 ? f64_at(c, r) : static_cast<double>(int_at(c, r));
        const double b = t.lit_type == F64 ? t.lit_f : static_cast<double>(t.lit_i);
        return cmp_apply<double>(t.op, a, b, mutant) ? 1 : 0;
    }
    const int64_t a = c.type == F64 ? static_cast<int64_t>(f64_at(c, r)) : int_at(c, r);
    const int64_t b = t.lit_type == F64 ? static_cast<int64_t>(t.lit_f) : t.lit_i;
    return cmp_apply<int64_t>(t.op, a, b, mutant) ? 1 : 0;
}

}  // namespace

extern "C" {

// Writes the passing row indices (ascending) to `sel`; returns their count.
uint64_t orc_filter_select(const orc_col* cols, uint64_t n, const orc_pred* terms,
                           uint32_t nterms, uint32_t* sel, int32_t mutant) {
    uint64_t m = 0;
    for (uint64_t r = 0; r < n; ++r) {
        int v = 1;  // AND identity
        for (uint32_t j = 0; j < nterms; ++j) {
            const int x = cmp_term(cols[terms[j].col], terms[j], r, mutant);
            if (x == 0) { v = 0; break; }  // FALSE AND anything = FALSE
            if (x < 0) v = -1;             // NULL AND TRUE = NULL
        }
        bool pass = v == 1 || (mutant == 1 && v < 0);
        if (mutant == 4) pass = !(v == 1);
        if (pass) sel[m++] = static_cast<uint32_t>(r);
    }
    return m;
}

}  // extern "C"
