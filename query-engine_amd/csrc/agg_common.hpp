// Slot aggregation machinery shared by the group-by and the fused
// join->group-by: an aggregate "program", per-slot state updates in LDS (block
// partials) or HBM (atomics), the deterministic block-order merge and the
// output emission. Semantics follow ops/aggregate.cpp:118-155 (scatter_row),
// ops/agg_internal.h:34-81 (NaN-greatest order, identities) and :169-229
// (finalisation), with the result types of ops/aggregate.cpp:78-97.
#pragma once

#include "internal.hpp"

namespace qe_gpu {

constexpr int kMaxAggs = 16;
constexpr int kMaxKeys = 8;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

enum AggKind : int32_t {
    K_COUNT_STAR = 0,
    K_COUNT = 1,
    K_SUM_I = 2,
    K_SUM_F = 3,
    K_MIN_I = 4,
    K_MAX_I = 5,
    K_MIN_F = 6,
    K_MAX_F = 7,
};

struct AggProg {
    int32_t n;
    int32_t kind[kMaxAggs];
    int32_t func[kMaxAggs];   // qe_agg_func (AVG finalises differently from SUM)
    ColView in[kMaxAggs];
    // Compact partial layout: which aggs keep an accumulator / a non-NULL count.
    // A count is only tracked for nullable inputs; for all-valid inputs the
    // non-NULL count equals the group's row count.
    int32_t acc_of[kMaxAggs];  // index among accumulators, or -1
    int32_t cnt_of[kMaxAggs];  // index among counters, or -1
    // Aggregates over the same column share state: SUM(x) and AVG(x) one
    // accumulator, COUNT/SUM/AVG/MIN/MAX(x) one non-NULL counter. Only the
    // owner (first such aggregate) updates a shared field.
    int32_t own_acc[kMaxAggs];
    int32_t own_cnt[kMaxAggs];
    int32_t acc_kind[kMaxAggs];  // AggKind of accumulator i
    int32_t nacc, ncnt;
};

// Result type of an aggregate (ops/aggregate.cpp:78-97).
inline int agg_result_type(int func, int in_type) {
    switch (func) {
        case QE_AGG_COUNT_STAR:
        case QE_AGG_COUNT:
            return QE_I64;
        case QE_AGG_SUM:
            return in_type == QE_F64 ? QE_F64 : QE_I64;
        case QE_AGG_AVG:
            return QE_F64;
        default:
            return in_type;
    }
}

// Build the device program; validates like agg_result_type (SUM/AVG of STR are
// invalid_argument; MIN/MAX of STR compare by dictionary value, which the device
// path does not hold, so they are reported unsupported).
AggProg make_prog(const qe_agg_spec* aggs, uint32_t naggs, const ColView* inputs,
                  const int* input_types);

// Identity of each state field.
__host__ __device__ inline uint64_t acc_identity(int kind) {
    switch (kind) {
        case K_MIN_I:
            return static_cast<uint64_t>(INT64_MAX);
        case K_MAX_I:
            return static_cast<uint64_t>(INT64_MIN);
        case K_MIN_F:
            return 0xFFFFFFFFFFFFFFFFull;  // order key above +qNaN: "no value yet"
        case K_MAX_F:
            return 0;  // below every order key
        default:
            return 0;
    }
}

// ---- LDS partial layout: acc u64[nacc][S] | cnt u32[ncnt][S] | rows u32[S] ------
__host__ __device__ inline size_t lds_bytes(uint32_t S, const AggProg& p) {
    return size_t(S) * (8 * p.nacc + 4 * p.ncnt + 4);
}
// Compact partial (per block) layout in HBM: rows | acc[nacc] | cnt[ncnt], u64 each.
__host__ __device__ inline size_t part_words(uint32_t S, const AggProg& p) {
    return size_t(S) * (1 + p.nacc + p.ncnt);
}

__device__ inline void lds_init(char* smem, uint32_t S, const AggProg& p) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * p.nacc);
    uint32_t* rows = cnt + size_t(S) * p.ncnt;
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        rows[s] = 0;
        for (int a = 0; a < p.n; ++a)
            if (p.acc_of[a] >= 0) acc[size_t(p.acc_of[a]) * S + s] = acc_identity(p.kind[a]);
        for (int c = 0; c < p.ncnt; ++c) cnt[size_t(c) * S + s] = 0;
    }
}

// Fold one row into `slot` of the block's LDS partials; value inputs of agg a
// are read at row vrow[a] (the join path reads probe- or build-side rows).
__device__ inline void lds_update(char* smem, uint32_t S, const AggProg& p, uint32_t slot,
                                  const uint64_t* vrow) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * p.nacc);
    uint32_t* rows = cnt + size_t(S) * p.ncnt;
    atomicAdd(&rows[slot], 1u);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR) continue;
        const uint64_t r = vrow[a];
        if (!is_valid(p.in[a], r)) continue;
        if (p.cnt_of[a] >= 0 && p.own_cnt[a]) atomicAdd(&cnt[size_t(p.cnt_of[a]) * S + slot], 1u);
        if (p.acc_of[a] < 0 || !p.own_acc[a]) continue;
        uint64_t* ap = &acc[size_t(p.acc_of[a]) * S + slot];
        switch (k) {
            case K_SUM_I:
                atomicAdd(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(int_value(p.in[a], r)));
                break;
            case K_SUM_F:
                atomicAdd(reinterpret_cast<double*>(ap), f64_value(p.in[a], r));
                break;
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
}

// Same fold with the values already in registers: agg a reads value column
// vcol[a] (-1 for COUNT(*)) whose 64-bit pattern is bits[vcol[a]] (integers
// sign-extended, F64 raw) and whose validity is bit vcol[a] of `valid`.
__device__ inline void lds_update_vals(char* smem, uint32_t S, const AggProg& p, uint32_t slot,
                                       const int32_t* vcol, const uint64_t* bits, uint32_t valid) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * p.nacc);
    uint32_t* rows = cnt + size_t(S) * p.ncnt;
    atomicAdd(&rows[slot], 1u);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR) continue;
        const int c = vcol[a];
        if (!((valid >> c) & 1u)) continue;
        if (p.cnt_of[a] >= 0 && p.own_cnt[a]) atomicAdd(&cnt[size_t(p.cnt_of[a]) * S + slot], 1u);
        if (p.acc_of[a] < 0 || !p.own_acc[a]) continue;
        uint64_t* ap = &acc[size_t(p.acc_of[a]) * S + slot];
        const uint64_t v = bits[c];
        switch (k) {
            case K_SUM_I:
                atomicAdd(reinterpret_cast<unsigned long long*>(ap), static_cast<unsigned long long>(v));
                break;
            case K_SUM_F:
                atomicAdd(reinterpret_cast<double*>(ap), __longlong_as_double(static_cast<long long>(v)));
                break;
            case K_MIN_I:
                atomicMin(reinterpret_cast<long long*>(ap), static_cast<long long>(v));
                break;
            case K_MAX_I:
                atomicMax(reinterpret_cast<long long*>(ap), static_cast<long long>(v));
                break;
            case K_MIN_F:
                atomicMin(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(
                              f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
                break;
            case K_MAX_F:
                atomicMax(reinterpret_cast<unsigned long long*>(ap),
                          static_cast<unsigned long long>(
                              f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
                break;
            default:
                break;
        }
    }
}

// A fold specialised for the programs a streaming probe mostly runs: COUNT(*)
// (and COUNT of all-valid columns) plus at most two SUM/AVG accumulators over
// all-valid value columns (no non-NULL counters, ncnt == 0). Its few fields
// are kernel arguments held in scalar registers, so a row's fold is one LDS
// atomic per field with no reload of the aggregate program.
struct FastFold {
    int32_t nacc;    // accumulators (0..2), the program's own
    int32_t f64[2];  // accumulator i sums F64 (else a 64-bit integer)
    int32_t col[2];  // value column (0 or 1) accumulator i reads
};

// Host: the fast fold of `p` when it has that shape (vcol: agg -> value column).
inline bool make_fast_fold(const AggProg& p, const int32_t* vcol, FastFold& ff) {
    ff = FastFold{};
    if (p.ncnt != 0 || p.nacc > 2) return false;
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR || k == K_COUNT) continue;
        if (k != K_SUM_F && k != K_SUM_I) return false;
        if (!p.own_acc[a]) continue;
        const int i = p.acc_of[a];
        if (i < 0 || i > 1 || vcol[a] < 0 || vcol[a] > 1) return false;
        ff.f64[i] = k == K_SUM_F ? 1 : 0;
        ff.col[i] = vcol[a];
    }
    ff.nacc = p.nacc;
    return true;
}

template <int NV>
__device__ inline void lds_fold_fast(char* smem, uint32_t S, const FastFold& ff, uint32_t slot,
                                     const uint64_t* bits) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* rows = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * ff.nacc);
    atomicAdd(&rows[slot], 1u);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (i >= ff.nacc) break;
        const uint64_t v = (NV > 1 && ff.col[i]) ? bits[NV > 1 ? 1 : 0] : bits[0];
        uint64_t* ap = &acc[size_t(i) * S + slot];
        if (ff.f64[i])
            atomicAdd(reinterpret_cast<double*>(ap), __longlong_as_double(static_cast<long long>(v)));
        else
            atomicAdd(reinterpret_cast<unsigned long long*>(ap), static_cast<unsigned long long>(v));
    }
}

// Flush the block's LDS partials into partial[block] (compact layout).
__device__ inline void lds_flush(const char* smem, uint32_t S, const AggProg& p, uint64_t* part) {
    const uint64_t* acc = reinterpret_cast<const uint64_t*>(smem);
    const uint32_t* cnt = reinterpret_cast<const uint32_t*>(smem + size_t(8) * S * p.nacc);
    const uint32_t* rows = cnt + size_t(S) * p.ncnt;
    uint64_t* dst = part + size_t(blockIdx.x) * part_words(S, p);
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        dst[s] = rows[s];
        for (int i = 0; i < p.nacc; ++i) dst[(1 + size_t(i)) * S + s] = acc[size_t(i) * S + s];
        for (int c = 0; c < p.ncnt; ++c) dst[(1 + size_t(p.nacc) + c) * S + s] = cnt[size_t(c) * S + s];
    }
}

// ---- HBM state (final, after merge): rows | acc[nacc] | cnt[ncnt] (same compact
// layout, u64 words). Atomic updates straight into HBM for slot counts that do
// not fit the LDS budget.
__host__ __device__ inline size_t state_words(uint64_t S, const AggProg& p) {
    return size_t(S) * (1 + p.nacc + p.ncnt);
}

__device__ inline void global_update(uint64_t* st, uint64_t S, const AggProg& p, uint64_t slot,
                                     const uint64_t* vrow) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&st[slot]), 1ull);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR) continue;
        const uint64_t r = vrow[a];
        if (!is_valid(p.in[a], r)) continue;
        if (p.cnt_of[a] >= 0 && p.own_cnt[a])
            atomicAdd(reinterpret_cast<unsigned long long*>(&st[(1 + size_t(p.nacc) + p.cnt_of[a]) * S + slot]),
                      1ull);
        if (p.acc_of[a] < 0 || !p.own_acc[a]) continue;
        uint64_t* acc = st + (1 + size_t(p.acc_of[a])) * S + slot;
        switch (k) {
            case K_SUM_I:
                atomicAdd(reinterpret_cast<unsigned long long*>(acc),
                          static_cast<unsigned long long>(int_value(p.in[a], r)));
                break;
            case K_SUM_F:
                unsafeAtomicAdd(reinterpret_cast<double*>(acc), f64_value(p.in[a], r));
                break;
            case K_MIN_I:
                atomicMin(reinterpret_cast<long long*>(acc), static_cast<long long>(int_value(p.in[a], r)));
                break;
            case K_MAX_I:
                atomicMax(reinterpret_cast<long long*>(acc), static_cast<long long>(int_value(p.in[a], r)));
                break;
            case K_MIN_F:
                atomicMin(reinterpret_cast<unsigned long long*>(acc),
                          static_cast<unsigned long long>(f64_order_key(f64_value(p.in[a], r))));
                break;
            case K_MAX_F:
                atomicMax(reinterpret_cast<unsigned long long*>(acc),
                          static_cast<unsigned long long>(f64_order_key(f64_value(p.in[a], r))));
                break;
            default:
                break;
        }
    }
}

// Same fold as global_update with the values already in registers, as
// lds_update_vals takes them (64-bit counters in the HBM state).
__device__ inline void global_update_vals(uint64_t* st, uint64_t S, const AggProg& p, uint64_t slot,
                                          const int32_t* vcol, const uint64_t* bits, uint32_t valid) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&st[slot]), 1ull);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR) continue;
        const int c = vcol[a];
        if (!((valid >> c) & 1u)) continue;
        if (p.cnt_of[a] >= 0 && p.own_cnt[a])
            atomicAdd(reinterpret_cast<unsigned long long*>(&st[(1 + size_t(p.nacc) + p.cnt_of[a]) * S + slot]),
                      1ull);
        if (p.acc_of[a] < 0 || !p.own_acc[a]) continue;
        uint64_t* acc = st + (1 + size_t(p.acc_of[a])) * S + slot;
        const uint64_t v = bits[c];
        switch (k) {
            case K_SUM_I:
                atomicAdd(reinterpret_cast<unsigned long long*>(acc), static_cast<unsigned long long>(v));
                break;
            case K_SUM_F:
                unsafeAtomicAdd(reinterpret_cast<double*>(acc), __longlong_as_double(static_cast<long long>(v)));
                break;
            case K_MIN_I:
                atomicMin(reinterpret_cast<long long*>(acc), static_cast<long long>(v));
                break;
            case K_MAX_I:
                atomicMax(reinterpret_cast<long long*>(acc), static_cast<long long>(v));
                break;
            case K_MIN_F:
                atomicMin(reinterpret_cast<unsigned long long*>(acc),
                          static_cast<unsigned long long>(
                              f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
                break;
            case K_MAX_F:
                atomicMax(reinterpret_cast<unsigned long long*>(acc),
                          static_cast<unsigned long long>(
                              f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
                break;
            default:
                break;
        }
    }
}

// Host launchers (agg_common.hip).
void init_state(qe_gpu_ctx* ctx, uint64_t* st, uint64_t S, const AggProg& p);
// Merge nblocks partials of S slots into st in block order (deterministic). The
// partials may cover slots [out_lo, out_lo + S) of a state laid out for out_S
// slots (out_S = 0: st is laid out for S slots). With `accumulate` the merge
// folds into the state's values instead of replacing them.
void merge_partials(qe_gpu_ctx* ctx, const uint64_t* part, uint32_t nblocks, uint64_t* st,
                    uint64_t S, const AggProg& p, uint64_t out_S = 0, uint64_t out_lo = 0,
                    bool accumulate = false);

// Key source for emission.
struct KeyEmit {
    int mode;            // 0 = direct range (key = kmin + slot, slot S-1 is NULL), 1 = words
    int nkeys;
    int32_t types[kMaxKeys];
    int64_t kmin;
    uint64_t null_slot;  // direct mode
    const uint64_t* words;     // words mode: [nkeys][S]
    const uint64_t* nullmask;  // words mode: [S]
};

// Emit one output row per slot with rows > 0 (or every slot when `all_slots`):
// key columns then finalised aggregates. Returns the table. A device flag
// (`flag`, optional) is read back with the emission's own counts into
// *flag_out, so a caller's check of the aggregation costs no extra round trip.
qe_dev_table* emit_groups(qe_gpu_ctx* ctx, const uint64_t* st, uint64_t S, const AggProg& p,
                          const KeyEmit& ke, bool all_slots, const unsigned int* flag = nullptr,
                          unsigned int* flag_out = nullptr);

// LDS budget per block for partial aggregation (bytes). One block per CU can
// take up to 160 KiB on gfx950; smaller footprints allow several.
constexpr size_t kLdsBudget = 152 * 1024;

}  // namespace qe_gpu
