// Filter with stable compaction. qe_gpu_filter replaces qe::Filter
// (ops/filter.h, ops/filter.cpp:28-66) for predicates that are a conjunction of
// column-versus-literal comparisons (expr/expr.h: cmp/lt/le/gt/ge/eq/ne joined
// by logic_and):
//  * a row passes iff every comparison is TRUE (SQL WHERE): a NULL operand or a
//    NULL literal makes the comparison NULL, which does not pass; NaN compares
//    FALSE except under != (IEEE, as the reference's compare kernels);
//  * numeric operands are promoted as expr::cmp promotes them: any F64 operand
//    gives a double comparison, otherwise a 64-bit integer comparison (I32
//    sign-extends); TS compares TS and BOOL compares BOOL;
//  * passing rows keep their order, and every column is compacted with its
//    validity (compact_column, core/owned_batch.cpp:101-166).
//
// Kernels: pred_count_kernel evaluates a 4096-row tile (16 rows per thread,
// rows k*256 + t so each wave's rows are consecutive) and counts the passing
// rows with wave64 ballots; one scan over the tiles; pred_emit_kernel repeats
// the ballots and writes each passing row's index at tile offset + rows of
// earlier rounds + rows of earlier waves + its rank among the wave's passing
// lanes; the columns are then gathered through those indices.
#include "internal.hpp"

using namespace qe_gpu;

namespace qe_gpu {

constexpr int kMaxPredTerms = 8;
constexpr uint32_t kPredBlock = 256;
constexpr uint32_t kPredRounds = 16;
constexpr uint32_t kPredTile = kPredBlock * kPredRounds;
constexpr uint32_t kPredWaves = kPredBlock / kWave;

struct PredTerm {
    ColView c;
    int32_t op;        // qe_cmp_op
    int32_t as_float;  // compare as double
    int64_t li;        // literal (integer comparison)
    double lf;         // literal (double comparison)
};
struct Pred {
    int32_t n;
    int32_t never;  // a NULL literal: nothing passes
    PredTerm t[kMaxPredTerms];
};

template <typename T>
__device__ inline bool cmp_vals(int op, T a, T b) {
    switch (op) {
        case QE_CMP_LT:
            return a < b;
        case QE_CMP_LE:
            return a <= b;
        case QE_CMP_GT:
            return a > b;
        case QE_CMP_GE:
            return a >= b;
        case QE_CMP_EQ:
            return a == b;
        default:
            return a != b;
    }
}

__device__ inline bool pred_row(const Pred& P, uint64_t r) {
    if (P.never) return false;
    for (int j = 0; j < P.n; ++j) {
        const PredTerm& t = P.t[j];
        if (!is_valid(t.c, r)) return false;  // NULL comparison: not TRUE
        bool res;
        if (t.as_float) {
            const double a = t.c.type == QE_F64 ? f64_value(t.c, r) : static_cast<double>(int_value(t.c, r));
            res = cmp_vals<double>(t.op, a, t.lf);
        } else {
            res = cmp_vals<int64_t>(t.op, int_value(t.c, r), t.li);
        }
        if (!res) return false;
    }
    return true;
}

__global__ void __launch_bounds__(kPredBlock) pred_count_kernel(Pred P, uint64_t n, uint32_t* tile_cnt) {
    __shared__ uint32_t wsum[kPredWaves];
    const uint64_t base = uint64_t(blockIdx.x) * kPredTile;
    uint32_t cnt = 0;
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t r = base + k * kPredBlock + threadIdx.x;
        cnt += (r < n && pred_row(P, r)) ? 1u : 0u;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kPredWaves; ++w) s += wsum[w];
        tile_cnt[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kPredBlock) pred_emit_kernel(Pred P, uint64_t n, const uint64_t* tile_off,
                                                               uint32_t* sel) {
    __shared__ uint32_t cnt[kPredRounds * kPredWaves];  // passing rows per (round, wave), then exclusive offsets
    const uint64_t base = uint64_t(blockIdx.x) * kPredTile;
    const unsigned lane = lane_id(), w = threadIdx.x / kWave;
    uint64_t mask[kPredRounds];
    bool pass[kPredRounds];
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t r = base + k * kPredBlock + threadIdx.x;
        pass[k] = r < n && pred_row(P, r);
        mask[k] = __ballot(pass[k]);
        if (lane == 0) cnt[k * kPredWaves + w] = static_cast<uint32_t>(__popcll(mask[k]));
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // rows are ordered by (round, wave, lane)
        uint32_t run = 0;
        for (uint32_t i = 0; i < kPredRounds * kPredWaves; ++i) {
            const uint32_t c = cnt[i];
            cnt[i] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint64_t toff = tile_off[blockIdx.x];
    const uint64_t below = (uint64_t(1) << lane) - 1;
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        if (!pass[k]) continue;
        const uint64_t pos = toff + cnt[k * kPredWaves + w] + __popcll(mask[k] & below);
        sel[pos] = static_cast<uint32_t>(base + k * kPredBlock + threadIdx.x);
    }
}

// Validates the terms against the columns (the type rules of expr::cmp) and
// builds the device predicate.
Pred make_pred(const qe_dev_col* const* cols, uint32_t ncols, const qe_pred_term* terms, uint32_t nterms) {
    invalid(nterms >= 1 && terms != nullptr, "filter needs >=1 comparison");
    check(nterms <= kMaxPredTerms, QE_GPU_UNSUPPORTED, "too many filter terms for the device path");
    Pred P{};
    P.n = static_cast<int32_t>(nterms);
    for (uint32_t j = 0; j < nterms; ++j) {
        const qe_pred_term& t = terms[j];
        invalid(t.col < ncols, "filter column out of range");
        invalid(t.op >= QE_CMP_LT && t.op <= QE_CMP_NE, "unknown comparison");
        const int ct = cols[t.col]->type, lt = t.lit_type;
        check(ct != QE_STR && lt != QE_STR, QE_GPU_UNSUPPORTED,
              "STR comparisons resolve dictionary values; not on the device path");
        const bool cnum = ct == QE_I32 || ct == QE_I64 || ct == QE_F64;
        const bool lnum = lt == QE_I32 || lt == QE_I64 || lt == QE_F64;
        invalid((cnum && lnum) || (ct == lt && (ct == QE_TS || ct == QE_BOOL)),
                "comparison operand types are incompatible");
        // Literal payloads as expr::Scalar carries them (expr/expr.h:84-90).
        invalid(t.lit_is_null || lt != QE_I32 || (t.lit_i >= INT32_MIN && t.lit_i <= INT32_MAX),
                "I32 literal out of int32 range");
        invalid(t.lit_is_null || lt != QE_BOOL || t.lit_i == 0 || t.lit_i == 1, "BOOL literal must be 0 or 1");
        PredTerm& d = P.t[j];
        d.c = view(cols[t.col]);
        d.op = t.op;
        d.as_float = (ct == QE_F64 || lt == QE_F64) ? 1 : 0;
        d.li = t.lit_i;
        d.lf = lt == QE_F64 ? t.lit_f : static_cast<double>(t.lit_i);
        if (t.lit_is_null) P.never = 1;
    }
    return P;
}

// Passing row indices (in order) into `sel`; returns their count.
uint64_t filter_select(qe_gpu_ctx* ctx, const Pred& P, uint64_t n, uint32_t* sel) {
    if (n == 0) return 0;
    const uint64_t tiles = (n + kPredTile - 1) / kPredTile;
    Scratch cnt(ctx, tiles * 4), off(ctx, tiles * 8);
    pred_count_kernel<<<tiles, kPredBlock, 0, ctx->stream>>>(P, n, cnt.as<uint32_t>());
    QE_HIP(hipGetLastError());
    const uint64_t total = exclusive_scan_u32(ctx, cnt.as<uint32_t>(), off.as<uint64_t>(), tiles);
    pred_emit_kernel<<<tiles, kPredBlock, 0, ctx->stream>>>(P, n, off.as<uint64_t>(), sel);
    QE_HIP(hipGetLastError());
    return total;
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_filter(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                            const qe_pred_term* terms, uint32_t nterms, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && cols != nullptr && ncols >= 1, "filter needs >=1 column");
        const uint64_t n = common_len(cols, ncols);
        check(n < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
        const Pred P = make_pred(cols, ncols, terms, nterms);
        Scratch sel(ctx, std::max<uint64_t>(n, 1) * 4);
        const uint64_t m = filter_select(ctx, P, n, sel.as<uint32_t>());
        qe_dev_table* t = new_table(m);
        try {
            for (uint32_t c = 0; c < ncols; ++c) t->cols.push_back(gather_col(ctx, cols[c], sel.as<uint32_t>(), m));
        } catch (...) {
            qe_gpu_table_free(ctx, t);
            throw;
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *out = t;
    });
}

qe_gpu_status qe_gpu_filter_select(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                   const qe_pred_term* terms, uint32_t nterms, uint32_t* sel_out,
                                   uint64_t* count_out) {
    QE_ENTRY(ctx, {
        invalid(cols != nullptr && ncols >= 1 && count_out != nullptr, "filter needs >=1 column");
        const uint64_t n = common_len(cols, ncols);
        check(n < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
        invalid(n == 0 || sel_out != nullptr, "null selection output");
        const Pred P = make_pred(cols, ncols, terms, nterms);
        *count_out = filter_select(ctx, P, n, sel_out);
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
