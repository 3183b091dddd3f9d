// Slot aggregation machinery shared by the group-by and the fused
// join->group-by: an aggregate "program", per-slot state updates in LDS (block
// partials) or HBM (atomics), the merge of block partials and the output
// emission.
//
// F64 SUM/AVG are exact and bit-reproducible run to run. Each finite value x
// of a summed F64 column becomes the integer X = round(x * 2^s) (rounded half
// away from zero; exact whenever ulp(x) >= 2^-s), and the sums are integer
// sums: 96-bit in LDS (a 64-bit low word whose atomic add returns the carry,
// and a 32-bit high word), 128-bit in HBM. Integer addition is associative,
// so the order in which rows, waves and blocks add never changes a bit. The
// scale s of each column is computed on the device before the aggregation:
// s = 63 - e, where |x| < 2^e for the column's values (e = the largest
// exponent of a 2^16-row sample + 2 bits of headroom), so |X| < 2^63 and a
// 96-bit partial holds 2^32 terms; values at or past 2^e (outliers) go to a
// per-query list summed in sorted order at emission, and a full list makes
// the query rerun with e from every row. NaN and +/-inf set per-slot flags
// (the sum becomes NaN / +inf / -inf as the row-order fold would: NaN if any
// NaN or both infinities). The finished sum is the exact integer total scaled
// by 2^-s, rounded once to a double (plus the slot's outlier sum, if any).
// Terms whose magnitude is below 2^(e - 11) are rounded to the quantum 2^-s,
// so a group made only of values much smaller than its column's largest keeps
// an absolute error of at most (its rows) * 2^(e - 64). Semantics follow ops/aggregate.cpp:118-155 (scatter_row),
// ops/agg_internal.h:34-81 (NaN-greatest order, identities) and :169-229
// (finalisation), with the result types of ops/aggregate.cpp:78-97.
#pragma once

#include "internal.hpp"

namespace qe_gpu {

constexpr int kMaxAggs = 16;
constexpr int kMaxKeys = 8;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
// Summed F64 values past a column's sampled bound (outliers) kept per query.
constexpr uint32_t kMaxOutliers = 4096;

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
    // Exact F64 sums: accumulator i of kind K_SUM_F keeps its low word in the
    // accumulator field and its high word at index fhi_of[i] (-1 otherwise).
    int32_t fhi_of[kMaxAggs];
    int32_t nfs;
    const int32_t* fsc = nullptr;  // device [2 * kMaxAggs]: scale s of accumulator i, then its exponent bound e
    uint64_t* fflag = nullptr;     // device [nfs][fflag_stride]: non-finite flags by state slot
    uint64_t fflag_stride = 0;
    unsigned int* fovf = nullptr;  // device: the outlier list overflowed (the query reruns with exact bounds)
    uint64_t* fout = nullptr;      // device [kMaxOutliers][2]: (F64 sum index << 48 | state slot, value bits)
    unsigned int* fout_n = nullptr;
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

// ---- LDS partial layout: acc u64[nacc][S] | cnt u32[ncnt][S] | rows u32[S] | fhi u32[nfs][S]
__host__ __device__ inline size_t lds_bytes(uint32_t S, const AggProg& p) {
    return size_t(S) * (8 * p.nacc + 4 * p.ncnt + 4 + 4 * p.nfs);
}
// Merged fields per slot (block partials and the state's summed part):
// rows | acc[nacc] | cnt[ncnt] | fhi[nfs], u64 words (fhi sign-extended).
__host__ __device__ inline size_t merge_words(uint64_t S, const AggProg& p) {
    return size_t(S) * (1 + p.nacc + p.ncnt + p.nfs);
}
__host__ __device__ inline size_t part_words(uint32_t S, const AggProg& p) { return merge_words(S, p); }

// ---- exact F64 sums -----------------------------------------------------------------------
// x as round(x * 2^s) in 128-bit two's complement (lo, hi; hi is 0 or -1).
// Returns 0, or a flag: 1 NaN, 2 +inf, 4 -inf, 8 |x| >= 2^e (the scale does
// not cover x).
__device__ inline uint32_t f64_fixed(uint64_t bits, int s, int e, uint64_t& lo, int64_t& hi) {
    lo = 0;
    hi = 0;
    const int bexp = static_cast<int>((bits >> 52) & 0x7FF);
    const bool neg = (bits >> 63) != 0;
    if (bexp == 0x7FF) return (bits & ((uint64_t(1) << 52) - 1)) ? 1u : (neg ? 4u : 2u);
    uint64_t m = bits & ((uint64_t(1) << 52) - 1);
    if (bexp) m |= uint64_t(1) << 52;
    if (m == 0) return 0;
    if (bexp - 1023 >= e) return 8u;
    const int k = (bexp ? bexp - 1075 : -1074) + s;  // |x| = m * 2^(k - s)
    // With s = 63 - e and |x| < 2^e, |x| 2^s < 2^63: k <= 10 (11 for a subnormal,
    // whose m < 2^52) and the value fits the low word. A scale that breaks
    // this treats x as outside it.
    if (k > (bexp ? 10 : 11)) return 8u;
    // Both shifts computed and one selected (no divergent branches): m 2^k for
    // k >= 0, else m 2^k with half a quantum rounding away from zero (0 below
    // half a quantum, k < -53).
    const int r = k < 0 ? (k < -63 ? 63 : -k) : 1;
    const uint64_t up = m << (k < 0 ? 0 : k);
    const uint64_t dn = (m + (uint64_t(1) << (r - 1))) >> r;
    uint64_t l = k >= 0 ? up : (k < -53 ? 0 : dn);
    if (neg) l = ~l + 1;
    lo = l;
    hi = (neg && l != 0) ? -1 : 0;
    return 0;
}

// The exact (hi:lo) * 2^-s rounded once to a double (sticky bit below the
// 64 leading bits keeps the u64 -> double rounding correct).
__host__ __device__ inline double fixed_to_f64(uint64_t lo, int64_t hi, int s) {
    const bool neg = hi < 0;
    uint64_t l = lo, h = static_cast<uint64_t>(hi);
    if (neg) {
        l = ~l + 1;
        h = ~h + (l == 0 ? 1 : 0);
    }
    if (h == 0 && l == 0) return 0.0;
    double d;
    int ex = -s;
    if (h == 0) {
        d = static_cast<double>(l);
    } else {
        int lz = 0;
        while (!((h >> (63 - lz)) & 1)) ++lz;
        const int sh = 64 - lz;  // bits of h: the value is h * 2^64 + l; keep its 64 leading bits
        uint64_t m = sh == 64 ? h : ((h << (64 - sh)) | (l >> sh));
        const uint64_t below = sh == 64 ? l : (l & ((uint64_t(1) << sh) - 1));
        m |= below != 0 ? 1 : 0;
        d = static_cast<double>(m);
        ex += sh;
    }
    d = ldexp(d, ex);
    return neg ? -d : d;
}

// The sum field as finished: the flags win (NaN, or an infinity), else the
// scaled integer total.
__host__ __device__ inline double fixed_sum_value(uint64_t lo, int64_t hi, int s, uint64_t flags) {
    if ((flags & 1) || ((flags & 6) == 6)) return __builtin_nan("");
    if (flags & 2) return __builtin_inf();
    if (flags & 4) return -__builtin_inf();
    return fixed_to_f64(lo, hi, s);
}

// 96-bit LDS accumulation: the low word's atomic returns the value it added
// to; a wrap past 2^64 carries into the high word. The total is exact mod
// 2^96 whatever the order of the adds.
__device__ inline void lds_add_fixed(uint64_t* lo, uint32_t* hi, uint64_t l, int64_t h) {
    const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long*>(lo), static_cast<unsigned long long>(l));
    const uint32_t ha = static_cast<uint32_t>(h) + ((old + l < old) ? 1u : 0u);
    if (ha) atomicAdd(hi, ha);
}
// 128-bit HBM accumulation, the same way.
__device__ inline void global_add_fixed(uint64_t* lo, uint64_t* hi, uint64_t l, int64_t h) {
    const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long*>(lo), static_cast<unsigned long long>(l));
    const uint64_t ha = static_cast<uint64_t>(h) + ((old + l < old) ? 1u : 0u);
    if (ha) atomicAdd(reinterpret_cast<unsigned long long*>(hi), static_cast<unsigned long long>(ha));
}
// A value the integer sum cannot take: NaN/inf flag its slot; a finite value
// at or beyond the column's sampled bound goes to the outlier list (summed
// in a fixed order at emission), and a full list flags the query.
__device__ inline void f64_special(const AggProg& p, int fi, uint64_t gslot, uint32_t code, uint64_t bits) {
    if (code & 8u) {
        const unsigned int at = atomicAdd(p.fout_n, 1u);
        if (at < kMaxOutliers) {
            p.fout[2 * at] = (static_cast<uint64_t>(fi) << 48) | gslot;
            p.fout[2 * at + 1] = bits;
        } else {
            atomicOr(p.fovf, 1u);
        }
    } else {
        atomicOr(reinterpret_cast<unsigned long long*>(&p.fflag[size_t(fi) * p.fflag_stride + gslot]),
                 static_cast<unsigned long long>(code));
    }
}

__device__ inline uint32_t* lds_fhi(char* smem, uint32_t S, const AggProg& p) {
    return reinterpret_cast<uint32_t*>(smem + size_t(S) * (8 * p.nacc + 4 * p.ncnt + 4));
}

__device__ inline void lds_init(char* smem, uint32_t S, const AggProg& p) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * p.nacc);
    uint32_t* rows = cnt + size_t(S) * p.ncnt;
    uint32_t* fhi = rows + S;
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        rows[s] = 0;
        for (int a = 0; a < p.n; ++a)
            if (p.acc_of[a] >= 0) acc[size_t(p.acc_of[a]) * S + s] = acc_identity(p.kind[a]);
        for (int c = 0; c < p.ncnt; ++c) cnt[size_t(c) * S + s] = 0;
        for (int j = 0; j < p.nfs; ++j) fhi[size_t(j) * S + s] = 0;
    }
}

// One value of accumulator i (kind k) into its LDS field. gslot = the state
// slot of LDS slot `slot` (for the non-finite flags).
__device__ inline void lds_acc_value(char* smem, uint32_t S, const AggProg& p, int i, int k, uint32_t slot,
                                     uint64_t gslot, uint64_t v) {
    uint64_t* ap = reinterpret_cast<uint64_t*>(smem) + size_t(i) * S + slot;
    switch (k) {
        case K_SUM_I:
            atomicAdd(reinterpret_cast<unsigned long long*>(ap), static_cast<unsigned long long>(v));
            break;
        case K_SUM_F: {
            uint64_t l;
            int64_t h;
            const uint32_t code = f64_fixed(v, p.fsc[i], p.fsc[kMaxAggs + i], l, h);
            if (code == 0)
                lds_add_fixed(ap, lds_fhi(smem, S, p) + size_t(p.fhi_of[i]) * S + slot, l, h);
            else
                f64_special(p, p.fhi_of[i], gslot, code, v);
            break;
        }
        case K_MIN_I:
            atomicMin(reinterpret_cast<long long*>(ap), static_cast<long long>(v));
            break;
        case K_MAX_I:
            atomicMax(reinterpret_cast<long long*>(ap), static_cast<long long>(v));
            break;
        case K_MIN_F:
            atomicMin(reinterpret_cast<unsigned long long*>(ap),
                      static_cast<unsigned long long>(f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
            break;
        case K_MAX_F:
            atomicMax(reinterpret_cast<unsigned long long*>(ap),
                      static_cast<unsigned long long>(f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
            break;
        default:
            break;
    }
}

// The 64-bit pattern of input value r of agg a (integers sign-extended, F64 raw).
__device__ inline uint64_t value_bits(const ColView& c, uint64_t r) {
    if (c.type == QE_F64) return static_cast<const uint64_t*>(c.data)[r];
    return static_cast<uint64_t>(int_value(c, r));
}

// Fold one row into `slot` of the block's LDS partials; value inputs of agg a
// are read at row vrow[a] (the join path reads probe- or build-side rows).
__device__ inline void lds_update(char* smem, uint32_t S, const AggProg& p, uint32_t slot, const uint64_t* vrow,
                                  uint64_t slot_lo = 0) {
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
        lds_acc_value(smem, S, p, p.acc_of[a], k, slot, slot_lo + slot, value_bits(p.in[a], r));
    }
}

// Same fold with the values already in registers: agg a reads value column
// vcol[a] (-1 for COUNT(*)) whose 64-bit pattern is bits[vcol[a]] (integers
// sign-extended, F64 raw) and whose validity is bit vcol[a] of `valid`.
__device__ inline void lds_update_vals(char* smem, uint32_t S, const AggProg& p, uint32_t slot,
                                       const int32_t* vcol, const uint64_t* bits, uint32_t valid,
                                       uint64_t slot_lo = 0) {
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
        lds_acc_value(smem, S, p, p.acc_of[a], k, slot, slot_lo + slot, bits[c]);
    }
}

// A fold specialised for the programs a streaming probe mostly runs: COUNT(*)
// (and COUNT of all-valid columns) plus at most two SUM/AVG accumulators over
// all-valid value columns (no non-NULL counters, ncnt == 0). Its fields are
// kernel arguments in scalar registers (the F64 scales are read from the
// device once per block by fast_fold_scales), so a row's fold is a fixed
// sequence of LDS atomics with no reload of the aggregate program.
struct FastFold {
    int32_t nacc;    // accumulators (0..2), the program's own
    int32_t f64[2];  // accumulator i sums F64 (else a 64-bit integer)
    int32_t col[2];  // value column (0 or 1) accumulator i reads
    int32_t fhi[2];  // F64: index of the high word
    int32_t scale[2], elim[2];  // F64: filled on the device
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
        ff.fhi[i] = p.fhi_of[i];
    }
    ff.nacc = p.nacc;
    return true;
}

// Device: the F64 scales of the fast fold, read once per block.
__device__ inline void fast_fold_scales(FastFold& ff, const AggProg& p) {
    for (int i = 0; i < 2; ++i)
        if (i < ff.nacc && ff.f64[i]) {
            ff.scale[i] = p.fsc[i];
            ff.elim[i] = p.fsc[kMaxAggs + i];
        }
}

template <int NV>
__device__ inline void lds_fold_fast(char* smem, uint32_t S, const FastFold& ff, const AggProg& p, uint32_t slot,
                                     const uint64_t* bits, uint64_t slot_lo = 0) {
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
    uint32_t* rows = reinterpret_cast<uint32_t*>(smem + size_t(8) * S * ff.nacc);
    uint32_t* fhi = rows + S;
    atomicAdd(&rows[slot], 1u);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (i >= ff.nacc) break;
        const uint64_t v = (NV > 1 && ff.col[i]) ? bits[NV > 1 ? 1 : 0] : bits[0];
        uint64_t* ap = &acc[size_t(i) * S + slot];
        if (ff.f64[i]) {
            uint64_t l;
            int64_t h;
            const uint32_t code = f64_fixed(v, ff.scale[i], ff.elim[i], l, h);
            if (code == 0)
                lds_add_fixed(ap, &fhi[size_t(ff.fhi[i]) * S + slot], l, h);
            else
                f64_special(p, ff.fhi[i], slot_lo + slot, code, v);
        } else {
            atomicAdd(reinterpret_cast<unsigned long long*>(ap), static_cast<unsigned long long>(v));
        }
    }
}

// Flush the block's LDS partials into partial[block] (merge layout).
__device__ inline void lds_flush(const char* smem, uint32_t S, const AggProg& p, uint64_t* part) {
    const uint64_t* acc = reinterpret_cast<const uint64_t*>(smem);
    const uint32_t* cnt = reinterpret_cast<const uint32_t*>(smem + size_t(8) * S * p.nacc);
    const uint32_t* rows = cnt + size_t(S) * p.ncnt;
    const uint32_t* fhi = rows + S;
    uint64_t* dst = part + size_t(blockIdx.x) * part_words(S, p);
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        dst[s] = rows[s];
        for (int i = 0; i < p.nacc; ++i) dst[(1 + size_t(i)) * S + s] = acc[size_t(i) * S + s];
        for (int c = 0; c < p.ncnt; ++c) dst[(1 + size_t(p.nacc) + c) * S + s] = cnt[size_t(c) * S + s];
        for (int j = 0; j < p.nfs; ++j)
            dst[(1 + size_t(p.nacc) + p.ncnt + j) * S + s] =
                static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(fhi[size_t(j) * S + s])));
    }
}

// ---- HBM state: rows | acc[nacc] | cnt[ncnt] | fhi[nfs] | flags[nfs] | osum[nfs],
// u64 words (fhi: the high words of the 128-bit F64 sums; flags: their NaN/inf
// bits; osum: the F64 sum of the slot's outliers). Atomic updates straight into
// HBM for slot counts that do not fit the LDS budget. The flags and outlier
// sums follow the merged fields and no merge writes them.
__host__ __device__ inline size_t state_words(uint64_t S, const AggProg& p) {
    return size_t(S) * (1 + p.nacc + p.ncnt + 3 * p.nfs);
}
__host__ __device__ inline size_t state_osum(uint64_t S, const AggProg& p, int j) {
    return size_t(S) * (1 + p.nacc + p.ncnt + 2 * p.nfs + j);
}
__host__ __device__ inline size_t state_fhi(uint64_t S, const AggProg& p, int j) {
    return size_t(S) * (1 + p.nacc + p.ncnt + j);
}
__host__ __device__ inline size_t state_flags(uint64_t S, const AggProg& p, int j) {
    return size_t(S) * (1 + p.nacc + p.ncnt + p.nfs + j);
}

__device__ inline void global_acc_value(uint64_t* st, uint64_t S, const AggProg& p, int i, int k, uint64_t slot,
                                        uint64_t v) {
    uint64_t* acc = st + (1 + size_t(i)) * S + slot;
    switch (k) {
        case K_SUM_I:
            atomicAdd(reinterpret_cast<unsigned long long*>(acc), static_cast<unsigned long long>(v));
            break;
        case K_SUM_F: {
            uint64_t l;
            int64_t h;
            const uint32_t code = f64_fixed(v, p.fsc[i], p.fsc[kMaxAggs + i], l, h);
            if (code == 0)
                global_add_fixed(acc, st + state_fhi(S, p, p.fhi_of[i]) + slot, l, h);
            else
                f64_special(p, p.fhi_of[i], slot, code, v);
            break;
        }
        case K_MIN_I:
            atomicMin(reinterpret_cast<long long*>(acc), static_cast<long long>(v));
            break;
        case K_MAX_I:
            atomicMax(reinterpret_cast<long long*>(acc), static_cast<long long>(v));
            break;
        case K_MIN_F:
            atomicMin(reinterpret_cast<unsigned long long*>(acc),
                      static_cast<unsigned long long>(f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
            break;
        case K_MAX_F:
            atomicMax(reinterpret_cast<unsigned long long*>(acc),
                      static_cast<unsigned long long>(f64_order_key(__longlong_as_double(static_cast<long long>(v)))));
            break;
        default:
            break;
    }
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
        global_acc_value(st, S, p, p.acc_of[a], k, slot, value_bits(p.in[a], r));
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
        global_acc_value(st, S, p, p.acc_of[a], k, slot, bits[c]);
    }
}

// Host: the exact F64 sums of a query over n_rows input rows. `prepare_f64`
// launches the scale computation for every F64-summed column (a sample of
// 2^16 valid rows, or every row when `exact`), clears the flag and outlier-sum
// words of `st` (state of S slots) and points the program at the device
// scales, those words, the outlier list and the overflow word (all in `buf`).
// No host round trip: the aggregation kernels read the scales on the device.
void prepare_f64(qe_gpu_ctx* ctx, AggProg& p, uint64_t* st, uint64_t S, uint64_t n_rows, bool exact, Scratch& buf);
// Its two halves: the column scales (sample, exponent bound; needs only the
// input rows, so it can be queued before the slot count is known) and the
// state's NaN/inf flags and outlier sums (needs the state).
void prepare_f64_scales(qe_gpu_ctx* ctx, AggProg& p, uint64_t n_rows, bool exact, Scratch& buf);
void prepare_f64_state(qe_gpu_ctx* ctx, AggProg& p, uint64_t* st, uint64_t S);
// Sum the outlier list into the state's osum words (emit_groups does this; a
// path whose state is relabelled afterwards calls it before the relabelling).
void fold_outliers(qe_gpu_ctx* ctx, const AggProg& p, uint64_t* st, uint64_t S);

// Host launchers (agg_common.hip).
void init_state(qe_gpu_ctx* ctx, uint64_t* st, uint64_t S, const AggProg& p);
// Merge nblocks partials of S slots into st (integer sums, order-free). The
// partials may cover slots [out_lo, out_lo + S) of a state laid out for out_S
// slots (out_S = 0: st is laid out for S slots). With `accumulate` the merge
// folds into the state's values instead of replacing them.
void merge_partials(qe_gpu_ctx* ctx, const uint64_t* part, uint32_t nblocks, uint64_t* st,
                    uint64_t S, const AggProg& p, uint64_t out_S = 0, uint64_t out_lo = 0,
                    bool accumulate = false);

// Thrown by emit_groups when more summed F64 values than the outlier list
// holds passed their sampled bound: the caller reruns the query with bounds
// from every row (prepare_f64 exact).
struct ScaleRetry : Error {
    ScaleRetry() : Error{QE_GPU_HIP_ERROR, "internal: F64 sum scale retry not handled"} {}
};

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
// *flag_out, so a caller's check of the aggregation costs no extra round trip;
// so is the program's F64 overflow word (ScaleRetry when set).
qe_dev_table* emit_groups(qe_gpu_ctx* ctx, const uint64_t* st, uint64_t S, const AggProg& p,
                          const KeyEmit& ke, bool all_slots, const unsigned int* flag = nullptr,
                          unsigned int* flag_out = nullptr);

// LDS budget per block for partial aggregation (bytes): a gfx950 workgroup
// can take the CU's whole 160 KiB (hipDeviceProp sharedMemPerBlock = 163840),
// less 256 B for the kernels' static words. Smaller footprints allow several
// blocks per CU. 10,001 slots of COUNT(*) + an exact F64 SUM (16 B each) fit.
constexpr size_t kLdsBudget = 160 * 1024 - 256;

}  // namespace qe_gpu
