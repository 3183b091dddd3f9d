// Filter with stable compaction. qe_gpu_filter replaces qe::Filter
// (ops/filter.h, ops/filter.cpp:28-66) for predicates that are a conjunction of
// column-versus-literal comparisons (expr/expr.h: cmp/lt/le/gt/ge/eq/ne joined
// by logic_and):
//  * a row passes iff every comparison is TRUE (SQL WHERE): a NULL operand or a
//    NULL literal makes the comparison NULL, which does not pass; NaN compares
//    FALSE except under != (IEEE, expr/expr.h:36-40);
//  * numeric operands are promoted as expr::cmp promotes them: any F64 operand
//    gives a double comparison, otherwise a 64-bit integer comparison (I32
//    sign-extends); TS compares TS and BOOL compares BOOL;
//  * passing rows keep their order, and every column is compacted with its
//    validity; an output column is all-valid when no NULL survives
//    (compact_column, core/owned_batch.cpp:101-166).
//
// Two passes over 4096-row tiles (256 threads x 16 rounds; round k of wave w
// covers the 64-aligned rows base + 256k + 64w + lane, i.e. one bitmap word):
//  1. pred_mask_kernel evaluates the predicate term by term over all 16 rounds
//     (independent loads in flight), writes the passing-row bitmap (one ballot
//     per word) and the tile's count;
//  2. after an exclusive scan of the tile counts, compact_kernel reads the
//     bitmap and ranks each passing lane within the tile; per output column it
//     loads all 16 rounds (full-wave loads, no branches), stages the passing
//     values in LDS at their tile-local positions and writes the tile's
//     contiguous output range with coalesced stores (no selection-vector round
//     trip). Output validity is staged the same way, one byte per output row,
//     and each output word is then one wave ballot; only the tile's two edge
//     words, shared with the neighbouring tiles, use global atomics.
// qe_gpu_filter_select runs the same passes with the row indices as the only
// output.
#include "internal.hpp"

#include <mutex>

using namespace qe_gpu;

namespace qe_gpu {

constexpr int kMaxPredTerms = 8;
constexpr uint32_t kPredBlock = 256;
constexpr uint32_t kPredRounds = 16;
constexpr uint32_t kPredTile = kPredBlock * kPredRounds;
constexpr uint32_t kPredWaves = kPredBlock / kWave;
constexpr uint32_t kMaxCompactCols = 8;

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

struct CompactCol {
    const void* src;
    const uint64_t* sv;  // source validity, nullptr when all-valid
    void* dst;
    uint64_t* dv;        // output validity (zeroed), nullptr when all-valid
    int32_t width;
    int32_t pad;
};
struct CompactCols {
    int32_t n;
    CompactCol c[kMaxCompactCols];
};

// Lane k of a wave holds word k of the wave's 16 rounds; round k reads it back
// as a wave-uniform value.
__device__ inline uint64_t round_word(uint64_t lane_words, uint32_t k) {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(lane_words), k);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(lane_words >> 32), k);
    return (uint64_t(hi) << 32) | lo;
}

// pass[k] &= (v[k] op lit) for the 16 rounds; op is uniform, so each case is a
// straight unrolled loop.
template <typename T>
__device__ inline void cmp_rounds(int op, const T (&v)[kPredRounds], T lit, bool (&pass)[kPredRounds]) {
#define QE_CMP_LOOP(expr)                                              \
    _Pragma("unroll") for (uint32_t k = 0; k < kPredRounds; ++k) { \
        const T a = v[k];                                              \
        pass[k] = pass[k] && (expr);                                   \
    }                                                                  \
    break;
    switch (op) {
        case QE_CMP_LT:
            QE_CMP_LOOP(a < lit)
        case QE_CMP_LE:
            QE_CMP_LOOP(a <= lit)
        case QE_CMP_GT:
            QE_CMP_LOOP(a > lit)
        case QE_CMP_GE:
            QE_CMP_LOOP(a >= lit)
        case QE_CMP_EQ:
            QE_CMP_LOOP(a == lit)
        default:
            QE_CMP_LOOP(a != lit)
    }
#undef QE_CMP_LOOP
}

// One term over the 16 rounds of this thread: loads for every still-passing
// row first, then the comparison in the promoted type.
template <typename S>
__device__ inline void term_rounds(const PredTerm& t, uint64_t base, uint64_t n, bool (&pass)[kPredRounds]) {
    const S* src = static_cast<const S*>(t.c.data);
    const uint64_t last = n - 1;  // rows past the end load the last row; pass[k] is false there
    S raw[kPredRounds];
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t r = base + k * kPredBlock;
        raw[k] = src[r < last ? r : last];
    }
    if (t.as_float) {
        double v[kPredRounds];
#pragma unroll
        for (uint32_t k = 0; k < kPredRounds; ++k) v[k] = static_cast<double>(raw[k]);
        cmp_rounds<double>(t.op, v, t.lf, pass);
    } else {
        int64_t v[kPredRounds];
#pragma unroll
        for (uint32_t k = 0; k < kPredRounds; ++k) v[k] = static_cast<int64_t>(raw[k]);
        cmp_rounds<int64_t>(t.op, v, t.li, pass);
    }
}

// pass[k] for rows base + 256k (base = tile start + threadIdx.x); word0 is the
// bitmap word of round 0 of this wave (wave-uniform).
__device__ inline void eval_pred(const Pred& P, uint64_t base, uint64_t word0, uint64_t n, bool (&pass)[kPredRounds]) {
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) pass[k] = !P.never && base + k * kPredBlock < n;
    for (int j = 0; j < P.n; ++j) {
        const PredTerm& t = P.t[j];
        if (t.c.validity) {  // NULL operand: the comparison is NULL, not TRUE
            // Round k of this wave needs one validity word: lane k fetches it.
            const uint32_t lane = lane_id();
            const uint64_t my_word = word0 + lane * kPredWaves;
            const uint64_t my_valid = (lane < kPredRounds && my_word < words_for(n)) ? t.c.validity[my_word] : 0;
#pragma unroll
            for (uint32_t k = 0; k < kPredRounds; ++k)
                pass[k] = pass[k] && ((round_word(my_valid, k) >> lane) & 1ull);
        }
        switch (t.c.type) {
            case QE_F64:
                term_rounds<double>(t, base, n, pass);
                break;
            case QE_I32:
                term_rounds<int32_t>(t, base, n, pass);
                break;
            case QE_BOOL:
                term_rounds<uint8_t>(t, base, n, pass);
                break;
            default:  // I64, TS
                term_rounds<int64_t>(t, base, n, pass);
        }
    }
}

__global__ void __launch_bounds__(kPredBlock) pred_mask_kernel(Pred P, uint64_t n, uint64_t* mask, uint32_t* tile_cnt) {
    __shared__ uint32_t wsum[kPredWaves];
    const uint64_t tile = blockIdx.x;
    const uint64_t base = tile * kPredTile + threadIdx.x;
    const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t nwords = words_for(n);
    bool pass[kPredRounds];
    eval_pred(P, base, tile * (kPredTile / 64) + w, n, pass);
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t m = __ballot(pass[k]);
        const uint64_t word = tile * (kPredTile / 64) + k * kPredWaves + w;
        if (lane_id() == 0 && word < nwords) mask[word] = m;
        cnt += static_cast<uint32_t>(__popcll(m));
    }
    if (lane_id() == 0) wsum[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t i = 0; i < kPredWaves; ++i) s += wsum[i];
        tile_cnt[tile] = s;
    }
}

// Loads round k's row (clamped into range, so every load is a full-wave load
// with no branch around it) and stages the passing values at their tile-local
// output positions in LDS.
template <typename T>
__device__ inline void stage_rounds(const CompactCol& c, uint64_t base, uint64_t last, const uint32_t (&lp)[kPredRounds],
                                    uint32_t bits, void* stage) {
    const T* src = static_cast<const T*>(c.src);
    T v[kPredRounds];
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t r = base + k * kPredBlock;
        v[k] = __builtin_nontemporal_load(src + (r < last ? r : last));
    }
    T* st = static_cast<T*>(stage);
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k)
        if ((bits >> k) & 1u) st[lp[k]] = v[k];
}

template <typename T>
__device__ inline void copy_out(const void* stage, void* dst, uint64_t toff, uint32_t cnt) {
    const T* st = static_cast<const T*>(stage);
    T* d = static_cast<T*>(dst) + toff;
    for (uint32_t i = threadIdx.x; i < cnt; i += kPredBlock) d[i] = st[i];
}

// sel (optional) receives the passing row indices; cols are written at the
// same positions. nulls[c] counts the NULLs that survive in column c.
__global__ void __launch_bounds__(kPredBlock) compact_kernel(const uint64_t* __restrict__ mask, uint64_t n,
                                                             const uint64_t* __restrict__ tile_off, uint64_t total,
                                                             CompactCols C, uint32_t* __restrict__ sel,
                                                             unsigned long long* nulls) {
    // Output staging for one column of one tile; its validity as one byte per
    // output row, placed at (toff mod 64) + position so that byte i*64 + lane
    // is bit `lane` of output word (toff / 64) + i; the (round, wave) offsets.
    __shared__ __align__(16) uint64_t stage[kPredTile];
    __shared__ uint8_t vstage[kPredTile + 128];
    __shared__ uint32_t off[kPredRounds * kPredWaves];
    __shared__ uint32_t valid_sum;
    const uint64_t tile = blockIdx.x;
    const unsigned lane = lane_id();
    const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint64_t nwords = words_for(n);
    const uint64_t below = (uint64_t(1) << lane) - 1;
    // Lane k < 16 fetches the bitmap word of round k (one load per wave).
    const uint64_t my_word = tile * (kPredTile / 64) + lane * kPredWaves + w;
    const bool word_lane = lane < kPredRounds && my_word < nwords;
    const uint64_t my_mask = word_lane ? mask[my_word] : 0;
    if (lane < kPredRounds) off[lane * kPredWaves + w] = static_cast<uint32_t>(__popcll(my_mask));
    __syncthreads();
    if (threadIdx.x < kWave) {  // wave 0: exclusive scan of the 64 (round, wave) counts
        static_assert(kPredRounds * kPredWaves == kWave, "one count per lane");
        const uint32_t c = off[threadIdx.x];
        uint32_t s = c;
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t u = __shfl_up(s, d);
            if (static_cast<int>(threadIdx.x) >= d) s += u;
        }
        off[threadIdx.x] = s - c;
    }
    __syncthreads();
    const uint64_t toff = tile_off[tile];
    const uint32_t tcnt = static_cast<uint32_t>((tile + 1 < gridDim.x ? tile_off[tile + 1] : total) - toff);
    if (tcnt == 0) return;  // uniform across the block
    const uint64_t base = tile * kPredTile + threadIdx.x;
    const uint64_t last = n - 1;
    uint32_t lp[kPredRounds];  // tile-local output position of round k's row
    uint32_t bits = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPredRounds; ++k) {
        const uint64_t m = round_word(my_mask, k);
        bits |= static_cast<uint32_t>((m >> lane) & 1ull) << k;
        lp[k] = off[k * kPredWaves + w] + static_cast<uint32_t>(__popcll(m & below));
    }
    if (sel) {
        uint32_t* st = reinterpret_cast<uint32_t*>(stage);
#pragma unroll
        for (uint32_t k = 0; k < kPredRounds; ++k)
            if ((bits >> k) & 1u) st[lp[k]] = static_cast<uint32_t>(base + k * kPredBlock);
        __syncthreads();
        copy_out<uint32_t>(stage, sel, toff, tcnt);
        __syncthreads();
    }
    const unsigned sh0 = static_cast<unsigned>(toff & 63);
    const uint32_t vwords = (sh0 + tcnt + 63) / 64;
    for (int c = 0; c < C.n; ++c) {
        const CompactCol& col = C.c[c];
        const bool nullable = col.dv != nullptr;
        switch (col.width) {
            case 1:
                stage_rounds<uint8_t>(col, base, last, lp, bits, stage);
                break;
            case 4:
                stage_rounds<uint32_t>(col, base, last, lp, bits, stage);
                break;
            default:
                stage_rounds<uint64_t>(col, base, last, lp, bits, stage);
        }
        if (nullable) {
            // Round k of this wave reads source validity word my_word of lane k.
            const uint64_t my_valid = word_lane ? col.sv[my_word] : 0;
#pragma unroll
            for (uint32_t k = 0; k < kPredRounds; ++k) {
                const uint64_t vw = round_word(my_valid, k);
                if ((bits >> k) & 1u) vstage[sh0 + lp[k]] = static_cast<uint8_t>((vw >> lane) & 1ull);
            }
            if (threadIdx.x == 0) valid_sum = 0;
        }
        __syncthreads();
        switch (col.width) {
            case 1:
                copy_out<uint8_t>(stage, col.dst, toff, tcnt);
                break;
            case 4:
                copy_out<uint32_t>(stage, col.dst, toff, tcnt);
                break;
            default:
                copy_out<uint64_t>(stage, col.dst, toff, tcnt);
        }
        if (nullable) {
            // Output word (toff / 64) + i from one ballot; interior words belong
            // to this tile alone, the first and last are shared with the
            // neighbouring tiles.
            unsigned long long* dv = reinterpret_cast<unsigned long long*>(col.dv) + (toff >> 6);
            uint32_t valid = 0;
            for (uint32_t i = w; i < vwords; i += kPredWaves) {
                const uint32_t idx = i * 64 + lane;
                const bool b = idx >= sh0 && idx < sh0 + tcnt && vstage[idx] != 0;
                const uint64_t word = __ballot(b);
                valid += static_cast<uint32_t>(__popcll(word));
                if (lane == 0) {
                    if (i == 0 || i + 1 == vwords) {
                        if (word) atomicOr(dv + i, static_cast<unsigned long long>(word));
                    } else {
                        dv[i] = word;
                    }
                }
            }
            if (lane == 0 && valid) atomicAdd(&valid_sum, valid);
            __syncthreads();
            if (threadIdx.x == 0) {
                if (valid_sum != tcnt) atomicAdd(nulls + c, static_cast<unsigned long long>(tcnt - valid_sum));
                // Padding bits past the last row are set (the all-valid padding convention).
                if (tile + 1 == gridDim.x && (total & 63))
                    atomicOr(reinterpret_cast<unsigned long long*>(col.dv) + (total >> 6),
                             static_cast<unsigned long long>(~uint64_t(0) << (total & 63)));
            }
        }
        __syncthreads();  // stage / vstage are reused by the next column
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

// Bitmap + tile offsets of the passing rows; returns their count.
struct FilterMask {
    Scratch mask, off;
    uint64_t tiles = 0, total = 0;
};
FilterMask filter_mask(qe_gpu_ctx* ctx, const Pred& P, uint64_t n) {
    FilterMask f;
    if (n == 0) return f;
    f.tiles = (n + kPredTile - 1) / kPredTile;
    f.mask = Scratch(ctx, words_for(n) * 8);
    f.off = Scratch(ctx, f.tiles * 8);
    Scratch cnt(ctx, f.tiles * 4);
    {
        TimedRegion tr(ctx, "filter_pred_mask");
        pred_mask_kernel<<<f.tiles, kPredBlock, 0, ctx->stream>>>(P, n, f.mask.as<uint64_t>(), cnt.as<uint32_t>());
        QE_HIP(hipGetLastError());
    }
    f.total = exclusive_scan_u32(ctx, cnt.as<uint32_t>(), f.off.as<uint64_t>(), f.tiles);
    return f;
}

void launch_compact(qe_gpu_ctx* ctx, const FilterMask& f, uint64_t n, const CompactCols& C, uint32_t* sel,
                    unsigned long long* nulls) {
    TimedRegion tr(ctx, "filter_compact");
    compact_kernel<<<f.tiles, kPredBlock, 0, ctx->stream>>>(f.mask.as<uint64_t>(), n, f.off.as<uint64_t>(), f.total,
                                                            C, sel, nulls);
    QE_HIP(hipGetLastError());
}

// Join pairs from per-probe matches (kNoRow = no match): every row when
// keep_all, else the rows with a match; probe = row index, build = match, in
// row order. The same two passes as a filter on `match != kNoRow` with the row
// indices as selection output and `match` as the one compacted column.
uint64_t select_pairs(qe_gpu_ctx* ctx, const uint32_t* match, uint64_t n, bool keep_all, uint32_t** probe,
                      uint32_t** build) {
    check(n < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
    Pred P{};
    if (!keep_all) {
        P.n = 1;
        P.t[0].c = ColView{match, nullptr, QE_I32, 0};
        P.t[0].op = QE_CMP_NE;
        P.t[0].as_float = 0;
        P.t[0].li = static_cast<int32_t>(kNoRow);  // the I32 reading of kNoRow (-1)
        P.t[0].lf = 0.0;
    }
    FilterMask f = filter_mask(ctx, P, n);
    *probe = static_cast<uint32_t*>(dev_alloc(ctx, std::max<uint64_t>(f.total, 1) * 4));
    *build = static_cast<uint32_t*>(dev_alloc(ctx, std::max<uint64_t>(f.total, 1) * 4));
    if (f.total > 0) {
        CompactCols C{};
        C.n = 1;
        C.c[0] = CompactCol{match, nullptr, *build, nullptr, 4, 0};
        launch_compact(ctx, f, n, C, *probe, nullptr);
    }
    return f.total;
}

// ---- selection with no allocation ---------------------------------------------
// qe_gpu_filter_select_noalloc keeps its bitmap, tile counts, tile offsets and
// scan partials in one file-scope device array (below) instead of allocator
// blocks. Layout for at most kNoAllocMaxRows rows (256-byte aligned parts):
//   [0, 16 MiB)                 passing-row bitmap, one bit per row
//   [16 MiB, +128 KiB)          per-tile passing counts (u32)
//   [16 MiB + 128 KiB, +256 KiB) per-tile output offsets (u64)
//   [16 MiB + 384 KiB, +4 KiB)   scan partials (u64)
// Every word is written before it is read (the mask kernel writes every bitmap
// word and every tile count, the scan writes every partial and offset), so no
// part needs clearing between calls.
constexpr uint64_t kNoAllocMaxRows = uint64_t(1) << 27;
constexpr uint64_t kNoAllocTiles = kNoAllocMaxRows / kPredTile;
constexpr uint64_t kNoAllocMaskOff = 0;
constexpr uint64_t kNoAllocCntOff = kNoAllocMaskOff + kNoAllocMaxRows / 8;
constexpr uint64_t kNoAllocTileOff = kNoAllocCntOff + kNoAllocTiles * 4;
constexpr uint64_t kNoAllocSumsOff = kNoAllocTileOff + kNoAllocTiles * 8;
constexpr uint64_t kNoAllocSumsBytes = 4096;
constexpr uint64_t kNoAllocBytes = kNoAllocSumsOff + kNoAllocSumsBytes;
static_assert(kNoAllocMaxRows % kPredTile == 0 && kNoAllocCntOff % 256 == 0 && kNoAllocTileOff % 256 == 0 &&
                  kNoAllocSumsOff % 256 == 0,
              "scratch parts stay 256-byte aligned");

}  // namespace qe_gpu

// The no-allocation path's device memory: one copy per GPU, held by the loaded
// code object until the process ends. Host code never uses this name as an
// address (on the host it is a stand-in, not device memory): it uses the
// per-device address that qe_gpu::noalloc_setup resolves.
static __device__ __align__(256) uint8_t scratch[qe_gpu::kNoAllocBytes];

namespace qe_gpu {
namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices];
int g_noalloc_state[kNoAllocMaxDevices];  // 0 not yet resolved, 1 resolved, -1 failed
}  // namespace

void noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_state[device] != 0) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p != nullptr) {
        g_noalloc_ptr[device] = p;
        g_noalloc_state[device] = 1;
    } else {
        (void)hipGetLastError();
        g_noalloc_state[device] = -1;
    }
}

static uint8_t* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    check(device >= 0 && device < kNoAllocMaxDevices && g_noalloc_state[device] == 1, QE_GPU_HIP_ERROR,
          "the no-allocation scratch array has no device address on this GPU");
    return static_cast<uint8_t*>(g_noalloc_ptr[device]);
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
        FilterMask f = filter_mask(ctx, P, n);
        const uint64_t m = f.total;
        qe_dev_table* t = new_table(m);
        try {
            for (uint32_t c = 0; c < ncols; ++c) {
                const bool nullable = !cols[c]->all_valid && cols[c]->validity != nullptr && m > 0;
                t->cols.push_back(new_col(ctx, cols[c]->type, m, nullable));
                t->cols.back()->dict = cols[c]->dict;
                if (nullable) memset_async(ctx, t->cols.back()->validity, 0, words_for(m) * 8);
            }
            if (m > 0) {
                Scratch nulls(ctx, kMaxCompactCols * 8);
                for (uint32_t c0 = 0; c0 < ncols; c0 += kMaxCompactCols) {
                    CompactCols C{};
                    C.n = static_cast<int32_t>(std::min<uint32_t>(kMaxCompactCols, ncols - c0));
                    for (int i = 0; i < C.n; ++i) {
                        const qe_dev_col* s = cols[c0 + i];
                        qe_dev_col* d = t->cols[c0 + i];
                        C.c[i] = CompactCol{s->data, d->validity ? s->validity : nullptr, d->data, d->validity,
                                            static_cast<int32_t>(type_width(s->type)), 0};
                    }
                    memset_async(ctx, nulls.p, 0, kMaxCompactCols * 8);
                    launch_compact(ctx, f, n, C, nullptr, nulls.as<unsigned long long>());
                    uint64_t h[kMaxCompactCols];
                    QE_HIP(hipMemcpyAsync(h, nulls.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
                    QE_HIP(hipStreamSynchronize(ctx->stream));
                    for (int i = 0; i < C.n; ++i) {  // refresh_all_valid: no NULL survived
                        qe_dev_col* d = t->cols[c0 + i];
                        if (d->validity && h[i] == 0) {
                            dev_free(ctx, d->validity);
                            d->validity = nullptr;
                            d->all_valid = true;
                        }
                    }
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

qe_gpu_status qe_gpu_filter_select(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                   const qe_pred_term* terms, uint32_t nterms, uint32_t* sel_out,
                                   uint64_t* count_out) {
    QE_ENTRY(ctx, {
        invalid(cols != nullptr && ncols >= 1 && count_out != nullptr, "filter needs >=1 column");
        const uint64_t n = common_len(cols, ncols);
        check(n < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
        invalid(n == 0 || sel_out != nullptr, "null selection output");
        const Pred P = make_pred(cols, ncols, terms, nterms);
        FilterMask f = filter_mask(ctx, P, n);
        if (f.total > 0) launch_compact(ctx, f, n, CompactCols{}, sel_out, nullptr);
        *count_out = f.total;
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}


qe_gpu_status qe_gpu_filter_select_noalloc(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                           const qe_pred_term* terms, uint32_t nterms, uint32_t* sel_out,
                                           uint64_t* count_out) {
    QE_ENTRY(ctx, {
        invalid(cols != nullptr && ncols >= 1 && count_out != nullptr, "filter needs >=1 column");
        const uint64_t n = common_len(cols, ncols);
        check(n <= kNoAllocMaxRows, QE_GPU_UNSUPPORTED,
              "filter_select_noalloc takes at most 2^27 rows (the size of its static scratch); "
              "use qe_gpu_filter_select for more");
        invalid(n == 0 || sel_out != nullptr, "null selection output");
        const Pred P = make_pred(cols, ncols, terms, nterms);
        uint8_t* base = noalloc_scratch(ctx->device);
        *count_out = 0;
        if (n == 0) return QE_GPU_OK;
        const uint64_t tiles = (n + kPredTile - 1) / kPredTile;
        uint64_t* mask = reinterpret_cast<uint64_t*>(base + kNoAllocMaskOff);
        uint32_t* cnt = reinterpret_cast<uint32_t*>(base + kNoAllocCntOff);
        uint64_t* off = reinterpret_cast<uint64_t*>(base + kNoAllocTileOff);
        uint64_t* sums = reinterpret_cast<uint64_t*>(base + kNoAllocSumsOff);
        const uint64_t nsums = scan_sums_words(tiles);
        check(nsums * 8 <= kNoAllocSumsBytes, QE_GPU_UNSUPPORTED, "scan partials exceed their scratch part");
        pred_mask_kernel<<<tiles, kPredBlock, 0, ctx->stream>>>(P, n, mask, cnt);
        QE_HIP(hipGetLastError());
        exclusive_scan_u32_dev(ctx, cnt, off, tiles, sums);
        uint64_t total = 0;  // sums[nsums - 1] holds the scan's total
        QE_HIP(hipMemcpyAsync(&total, sums + nsums - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
        if (total > 0) {
            compact_kernel<<<tiles, kPredBlock, 0, ctx->stream>>>(mask, n, off, total, CompactCols{}, sel_out,
                                                                  nullptr);
            QE_HIP(hipGetLastError());
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *count_out = total;
    });
}

}  // extern "C"
