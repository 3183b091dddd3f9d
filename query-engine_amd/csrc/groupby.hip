// Hash group-by aggregation: qe_gpu_aggregate replaces qe::Aggregate
// (ops/aggregate.h:110-111; drain_and_build ops/aggregate.cpp:243-437).
//
// Strategy, chosen per call from the data (not assumed):
//  * zero keys      -> one slot; per-block LDS partials, block-order merge;
//  * one integer-like key (I32/I64/TS/BOOL/STR code) whose value range R is
//    small -> direct slots (key - min, plus one NULL slot): LDS partials when
//    (R+1) slots fit the LDS budget, HBM atomics otherwise;
//  * anything else  -> a device open-addressing hash table (canonical key words,
//    NULLs group together as in NullPolicy::kEqual) with HBM atomics.
// Output: key columns ++ aggregate columns, one row per group, unordered.
#include "agg_common.hpp"
#include "hash_common.hpp"

#include <algorithm>

using namespace qe_gpu;

namespace qe_gpu {

AggProg make_prog(const qe_agg_spec* aggs, uint32_t naggs, const ColView* inputs,
                  const int* input_types) {
    invalid(naggs >= 1, "aggregate needs >=1 aggregate");
    check(naggs <= kMaxAggs, QE_GPU_UNSUPPORTED, "too many aggregates for the device path");
    AggProg p{};
    p.n = static_cast<int32_t>(naggs);
    for (uint32_t a = 0; a < naggs; ++a) {
        const int f = aggs[a].func;
        invalid(f >= QE_AGG_COUNT_STAR && f <= QE_AGG_AVG, "unknown aggregate function");
        p.func[a] = f;
        if (f == QE_AGG_COUNT_STAR) {
            p.kind[a] = K_COUNT_STAR;
            p.in[a] = ColView{nullptr, nullptr, QE_I64, 0};
            continue;
        }
        const int t = input_types[a];
        p.in[a] = inputs[a];
        if (f == QE_AGG_COUNT) {
            p.kind[a] = K_COUNT;
            continue;
        }
        if (t == QE_STR) {
            // ops/aggregate.cpp:84-90: SUM/AVG(STR) are type errors.
            invalid(f != QE_AGG_SUM && f != QE_AGG_AVG,
                    f == QE_AGG_SUM ? "SUM is not defined on STR (VARCHAR)"
                                    : "AVG is not defined on STR (VARCHAR)");
            throw Error{QE_GPU_UNSUPPORTED, "MIN/MAX of STR compare by dictionary value; not on the device path"};
        }
        const bool fl = (t == QE_F64);
        switch (f) {
            case QE_AGG_SUM:
            case QE_AGG_AVG:
                p.kind[a] = fl ? K_SUM_F : K_SUM_I;
                break;
            case QE_AGG_MIN:
                p.kind[a] = fl ? K_MIN_F : K_MIN_I;
                break;
            case QE_AGG_MAX:
                p.kind[a] = fl ? K_MAX_F : K_MAX_I;
                break;
        }
    }
    p.nacc = p.ncnt = 0;
    auto same_col = [&](uint32_t a, uint32_t b) {
        return p.in[a].data == p.in[b].data && p.in[a].validity == p.in[b].validity && p.in[a].type == p.in[b].type;
    };
    for (uint32_t a = 0; a < naggs; ++a) {
        const int k = p.kind[a];
        p.acc_of[a] = -1;
        p.cnt_of[a] = -1;
        p.own_acc[a] = 0;
        p.own_cnt[a] = 0;
        if (k != K_COUNT_STAR && k != K_COUNT) {
            for (uint32_t b = 0; b < a && p.acc_of[a] < 0; ++b)
                if (p.kind[b] == k && p.acc_of[b] >= 0 && same_col(a, b)) p.acc_of[a] = p.acc_of[b];
            if (p.acc_of[a] < 0) {
                p.acc_kind[p.nacc] = k;
                p.acc_of[a] = p.nacc++;
                p.own_acc[a] = 1;
            }
        }
        if (k != K_COUNT_STAR && p.in[a].validity != nullptr) {
            for (uint32_t b = 0; b < a && p.cnt_of[a] < 0; ++b)
                if (p.cnt_of[b] >= 0 && same_col(a, b)) p.cnt_of[a] = p.cnt_of[b];
            if (p.cnt_of[a] < 0) {
                p.cnt_of[a] = p.ncnt++;
                p.own_cnt[a] = 1;
            }
        }
    }
    p.nfs = 0;
    for (int i = 0; i < kMaxAggs; ++i) p.fhi_of[i] = -1;
    for (int i = 0; i < p.nacc; ++i)
        if (p.acc_kind[i] == K_SUM_F) p.fhi_of[i] = p.nfs++;
    return p;
}

// ---- exact F64 sums: the scales ------------------------------------------------------------
// Columns summed in F64 and the accumulator each feeds.
struct F64Cols {
    ColView col[kMaxAggs];
    int32_t acc[kMaxAggs];
    int32_t n;
};

// Largest biased exponent of the finite, non-NULL values among rows 0, stride,
// 2*stride, ... of column blockIdx.y: one word per (column, block). The first
// block also clears the overflow word and the outlier count (ovf[0..1]).
__global__ void f64_exp_kernel(F64Cols fc, uint64_t n, uint64_t stride, uint32_t* part, unsigned int* ovf) {
    __shared__ uint32_t bm;
    if (threadIdx.x == 0) bm = 0;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2) ovf[threadIdx.x] = 0;
    __syncthreads();
    const ColView c = fc.col[blockIdx.y];
    const uint64_t ns = stride ? (n + stride - 1) / stride : 0;
    uint32_t mx = 0;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < ns; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t r = i * stride;
        if (!is_valid(c, r)) continue;
        const uint32_t be = static_cast<uint32_t>((static_cast<const uint64_t*>(c.data)[r] >> 52) & 0x7FF);
        if (be != 0x7FF && be > mx) mx = be;
    }
    atomicMax(&bm, mx);
    __syncthreads();
    if (threadIdx.x == 0) part[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = bm;
}

// Per column: the exponent bound e (|x| < 2^e: the largest exponent seen, +2
// bits of headroom when sampled) and the scale s = 63 - e, so every term is
// below 2^63: a 96-bit LDS partial holds 2^32 of them, and the 64-bit low
// word wraps into the high word only now and then (an add whose carry or
// borrow changes nothing skips the high word).
__global__ void f64_scale_kernel(const uint32_t* part, uint32_t nblk, F64Cols fc, int32_t headroom, int32_t* fsc) {
    const int j = threadIdx.x;
    if (j >= fc.n) return;
    uint32_t mx = 0;
    for (uint32_t b = 0; b < nblk; ++b) mx = max(mx, part[size_t(j) * nblk + b]);
    const int e = (mx == 0 ? -1022 : static_cast<int>(mx) - 1022) + headroom;
    fsc[fc.acc[j]] = 63 - e;
    fsc[kMaxAggs + fc.acc[j]] = e;
}

void prepare_f64(qe_gpu_ctx* ctx, AggProg& p, uint64_t* st, uint64_t S, uint64_t n_rows, bool exact, Scratch& buf) {
    prepare_f64_scales(ctx, p, n_rows, exact, buf);
    prepare_f64_state(ctx, p, st, S);
}

void prepare_f64_state(qe_gpu_ctx* ctx, AggProg& p, uint64_t* st, uint64_t S) {
    if (p.nfs == 0) return;
    if (st) memset_async(ctx, st + state_flags(S, p, 0), 0, size_t(2 * p.nfs) * S * 8);  // flags, outlier sums
    p.fflag = st ? st + state_flags(S, p, 0) : nullptr;
    p.fflag_stride = S;
}

void prepare_f64_scales(qe_gpu_ctx* ctx, AggProg& p, uint64_t n_rows, bool exact, Scratch& buf) {
    if (p.nfs == 0) return;
    F64Cols fc{};
    for (int a = 0; a < p.n; ++a)
        if (p.kind[a] == K_SUM_F && p.own_acc[a]) {
            fc.col[fc.n] = p.in[a];
            fc.acc[fc.n] = p.acc_of[a];
            ++fc.n;
        }
    const uint32_t nblk = exact ? std::max<uint32_t>(1, std::min<uint32_t>(ctx->num_cus * 4, (n_rows + 255) / 256)) : 64;
    // fsc | overflow word, outlier count | outliers | per-block exponents
    buf = Scratch(ctx, size_t(2 * kMaxAggs) * 4 + 8 + size_t(kMaxOutliers) * 16 + size_t(fc.n) * nblk * 4);
    int32_t* fsc = buf.as<int32_t>();
    unsigned int* ovf = reinterpret_cast<unsigned int*>(fsc + 2 * kMaxAggs);
    uint64_t* fout = reinterpret_cast<uint64_t*>(ovf + 2);
    uint32_t* part = reinterpret_cast<uint32_t*>(fout + 2 * size_t(kMaxOutliers));
    const uint64_t stride = exact ? 1 : std::max<uint64_t>(1, n_rows >> 16);
    f64_exp_kernel<<<dim3(nblk, fc.n), 256, 0, ctx->stream>>>(fc, n_rows, n_rows ? stride : 0, part, ovf);
    f64_scale_kernel<<<1, 64, 0, ctx->stream>>>(part, nblk, fc, exact ? 0 : 2, fsc);
    QE_HIP(hipGetLastError());
    p.fsc = fsc;
    p.fovf = ovf;
    p.fout = fout;
    p.fout_n = ovf + 1;
}

// The outliers' sums per (F64 sum, slot): one block sorts the list by (sum,
// slot, value bits) in LDS and each run is added in that order, so the sums
// do not depend on the order the values arrived in. They land in the state's
// osum words; the list is left empty.
__global__ void __launch_bounds__(1024) outlier_fold_kernel(AggProg p, uint64_t* st, uint64_t S) {
    __shared__ uint64_t key[kMaxOutliers], val[kMaxOutliers];
    const uint32_t n = min(*p.fout_n, kMaxOutliers);
    if (n == 0) return;
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        key[i] = i < n ? p.fout[2 * i] : ~uint64_t(0);
        val[i] = i < n ? p.fout[2 * i + 1] : 0;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1) {  // bitonic sort on (key, val)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l <= i) continue;
                const bool up = (i & k) == 0;
                const bool gt = key[i] > key[l] || (key[i] == key[l] && val[i] > val[l]);
                if (gt == up) {
                    const uint64_t tk = key[i], tv = val[i];
                    key[i] = key[l];
                    val[i] = val[l];
                    key[l] = tk;
                    val[l] = tv;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (i > 0 && key[i] == key[i - 1]) continue;
        double sum = 0.0;
        for (uint32_t j = i; j < n && key[j] == key[i]; ++j) sum += __longlong_as_double(static_cast<long long>(val[j]));
        const int fi = static_cast<int>(key[i] >> 48);
        const uint64_t slot = key[i] & ((uint64_t(1) << 48) - 1);
        st[state_osum(S, p, fi) + slot] = static_cast<uint64_t>(__double_as_longlong(sum));
    }
    __syncthreads();
    if (threadIdx.x == 0) *p.fout_n = 0;
}

void fold_outliers(qe_gpu_ctx* ctx, const AggProg& p, uint64_t* st, uint64_t S) {
    if (p.nfs == 0 || p.fout_n == nullptr) return;
    outlier_fold_kernel<<<1, 1024, 0, ctx->stream>>>(p, st, S);
    QE_HIP(hipGetLastError());
}

__global__ void init_state_kernel(uint64_t* st, uint64_t S, AggProg p) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
         s += uint64_t(gridDim.x) * blockDim.x) {
        st[s] = 0;
        for (int i = 0; i < p.nacc; ++i) st[(1 + size_t(i)) * S + s] = acc_identity(p.acc_kind[i]);
        for (int c = 0; c < p.ncnt; ++c) st[(1 + size_t(p.nacc) + c) * S + s] = 0;
        for (int j = 0; j < 3 * p.nfs; ++j) st[state_fhi(S, p, 0) + size_t(j) * S + s] = 0;  // fhi, flags, osum
    }
}

void init_state(qe_gpu_ctx* ctx, uint64_t* st, uint64_t S, const AggProg& p) {
    init_state_kernel<<<grid_for(S, 256, 8192), 256, 0, ctx->stream>>>(st, S, p);
    QE_HIP(hipGetLastError());
}

__device__ inline uint64_t combine(int kind, uint64_t a, uint64_t b) {
    switch (kind) {
        case K_SUM_I:
            return a + b;
        case K_MIN_I:
            return static_cast<int64_t>(a) < static_cast<int64_t>(b) ? a : b;
        case K_MAX_I:
            return static_cast<int64_t>(a) > static_cast<int64_t>(b) ? a : b;
        case K_MIN_F:
            return a < b ? a : b;
        case K_MAX_F:
            return a > b ? a : b;
        default:
            return 0;
    }
}

// One thread per (field, slot) and chunk of `chunk` consecutive blocks: the
// chunk's blocks fold in block order, then the chunks. Every sum is an integer
// sum (F64 sums as 128-bit fixed point: the thread of the low word folds the
// high word with its carries; the high-word field's own thread does nothing),
// MIN/MAX are order-free, so the result does not depend on the order. With
// `accumulate` the result folds into what `out` holds instead of replacing it.
__global__ void merge_partials_kernel(const uint64_t* part, uint32_t nblocks, uint32_t chunk, uint64_t* out,
                                      uint64_t S, AggProg p, uint64_t out_S, uint64_t out_lo, int accumulate) {
    const uint64_t W = merge_words(S, p);
    const uint64_t Wout = merge_words(out_S, p);
    const uint32_t nchunks = nblocks ? (nblocks + chunk - 1) / chunk : 1;  // 0 blocks: identities
    const uint64_t fhi0 = 1 + uint64_t(p.nacc) + p.ncnt;  // first high-word field
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < W * nchunks;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t i = t % W;
        const uint32_t c = static_cast<uint32_t>(t / W);
        const uint32_t b0 = c * chunk, b1 = min(nblocks, b0 + chunk);
        const uint64_t field = i / S;
        const uint64_t s = i - field * S;
        const size_t o = size_t(c) * Wout + field * out_S + out_lo + s;
        if (field >= fhi0) continue;  // high words: folded with their low words
        if (field == 0 || field > uint64_t(p.nacc)) {  // rows or cnt: integer sums
            uint64_t v = accumulate ? out[o] : 0;
            for (uint32_t b = b0; b < b1; ++b) v += part[size_t(b) * W + i];
            out[o] = v;
            continue;
        }
        const int ai = static_cast<int>(field - 1);
        const int kind = p.acc_kind[ai];
        if (kind == K_SUM_F) {
            const uint64_t hf = fhi0 + p.fhi_of[ai];
            const size_t oh = size_t(c) * Wout + hf * out_S + out_lo + s;
            uint64_t lo = accumulate ? out[o] : 0, hi = accumulate ? out[oh] : 0;
            for (uint32_t b = b0; b < b1; ++b) {
                const uint64_t l = part[size_t(b) * W + i];
                lo += l;
                hi += part[size_t(b) * W + hf * S + s] + (lo < l ? 1 : 0);
            }
            out[o] = lo;
            out[oh] = hi;
            continue;
        }
        uint64_t v = accumulate ? out[o] : acc_identity(kind);
        for (uint32_t b = b0; b < b1; ++b) v = combine(kind, v, part[size_t(b) * W + i]);
        out[o] = v;
    }
}

void merge_partials(qe_gpu_ctx* ctx, const uint64_t* part, uint32_t nblocks, uint64_t* st,
                    uint64_t S, const AggProg& p, uint64_t out_S, uint64_t out_lo, bool accumulate) {
    const int acc = accumulate ? 1 : 0;
    const uint64_t W = merge_words(S, p);
    if (out_S == 0) out_S = S;
    constexpr uint32_t kChunk = 16;
    if (nblocks > 2 * kChunk) {
        const uint32_t nchunks = (nblocks + kChunk - 1) / kChunk;
        Scratch mid(ctx, size_t(nchunks) * W * 8);
        merge_partials_kernel<<<grid_for(W * nchunks, 256, 16384), 256, 0, ctx->stream>>>(
            part, nblocks, kChunk, mid.as<uint64_t>(), S, p, S, 0, 0);
        merge_partials_kernel<<<grid_for(W, 256, 16384), 256, 0, ctx->stream>>>(
            mid.as<uint64_t>(), nchunks, nchunks, st, S, p, out_S, out_lo, acc);
    } else {
        merge_partials_kernel<<<grid_for(W, 256, 16384), 256, 0, ctx->stream>>>(
            part, nblocks, nblocks ? nblocks : 1, st, S, p, out_S, out_lo, acc);
    }
    QE_HIP(hipGetLastError());
}

__global__ void group_flags_kernel(const uint64_t* rows, uint64_t S, uint32_t* flags, int all_slots) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
         s += uint64_t(gridDim.x) * blockDim.x)
        flags[s] = (all_slots || rows[s] > 0) ? 1u : 0u;
}

struct OutCols {
    void* data[kMaxKeys + kMaxAggs];
    uint8_t* valid[kMaxKeys + kMaxAggs];
    int32_t type[kMaxKeys + kMaxAggs];
};

__device__ inline void store_typed(void* base, int type, uint64_t o, uint64_t bits) {
    switch (type) {
        case QE_I32:
        case QE_STR:
            static_cast<int32_t*>(base)[o] = static_cast<int32_t>(bits);
            break;
        case QE_BOOL:
            static_cast<uint8_t*>(base)[o] = static_cast<uint8_t>(bits & 1);
            break;
        default:
            static_cast<uint64_t*>(base)[o] = bits;
    }
}

// Output row o from slot s: keys then finalised aggregates; bit c of the
// result is column c's validity.
__device__ inline uint32_t emit_row(const uint64_t* st, uint64_t S, const AggProg& p, const KeyEmit& ke,
                                    const OutCols& oc, uint64_t s, uint64_t o) {
    uint32_t vbits = 0;
    for (int k = 0; k < ke.nkeys; ++k) {
        bool isnull;
        uint64_t w;
        if (ke.mode == 0) {
            isnull = (s == ke.null_slot);
            w = static_cast<uint64_t>(ke.kmin + static_cast<int64_t>(s));
        } else {
            isnull = (ke.nullmask[s] >> k) & 1ull;
            w = ke.words[size_t(k) * S + s];
        }
        vbits |= (isnull ? 0u : 1u) << k;
        store_typed(oc.data[k], ke.types[k], o, isnull ? 0 : w);
    }
    const uint64_t rows = st[s];
    for (int a = 0; a < p.n; ++a) {
        const int c = ke.nkeys + a;
        const uint64_t acc = p.acc_of[a] >= 0 ? st[(1 + size_t(p.acc_of[a])) * S + s] : 0;
        const uint64_t cnt = p.cnt_of[a] >= 0 ? st[(1 + size_t(p.nacc) + p.cnt_of[a]) * S + s] : rows;
        const int f = p.func[a];
        const int k = p.kind[a];
        uint64_t out = 0;
        bool valid = true;
        if (f == QE_AGG_COUNT_STAR) {
            out = rows;
        } else if (f == QE_AGG_COUNT) {
            out = cnt;
        } else if (cnt == 0) {
            valid = false;  // no non-NULL input: SUM/MIN/MAX/AVG are NULL
        } else if (k == K_SUM_F) {
            const int j = p.fhi_of[p.acc_of[a]];
            double sum = fixed_sum_value(acc, static_cast<int64_t>(st[state_fhi(S, p, j) + s]), p.fsc[p.acc_of[a]],
                                         st[state_flags(S, p, j) + s]);
            const uint64_t ob = st[state_osum(S, p, j) + s];
            if (ob) sum += __longlong_as_double(static_cast<long long>(ob));
            if (f == QE_AGG_AVG) sum /= static_cast<double>(cnt);
            memcpy(&out, &sum, 8);
        } else if (f == QE_AGG_AVG) {
            const double avg = static_cast<double>(static_cast<int64_t>(acc)) / static_cast<double>(cnt);
            memcpy(&out, &avg, 8);
        } else if (k == K_MIN_F || k == K_MAX_F) {
            const double d = f64_from_order_key(acc);
            memcpy(&out, &d, 8);
        } else {
            out = acc;
        }
        vbits |= (valid ? 1u : 0u) << c;
        store_typed(oc.data[c], oc.type[c], o, valid ? out : 0);
    }
    return vbits;
}

__global__ void emit_kernel(const uint64_t* st, uint64_t S, AggProg p, KeyEmit ke,
                            const uint32_t* flags, const uint64_t* offs, OutCols oc) {
    const int ncol = ke.nkeys + p.n;
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
         s += uint64_t(gridDim.x) * blockDim.x) {
        if (!flags[s]) continue;
        const uint64_t o = offs[s];
        const uint32_t vb = emit_row(st, S, p, ke, oc, s, o);
        for (int c = 0; c < ncol; ++c) oc.valid[c][o] = (vb >> c) & 1u;
    }
}

// Slot counts up to kEmitSmall: one 1024-thread block lists the groups (a
// block scan of the slot flags: slot_of[g] = slot of output row g, out[0] =
// the group count), then one thread per output row emits it in order, each
// wave writing its 64 rows' validity word per column straight from a ballot
// and counting NULLs into out[1 + c]; the host reads the counts back together.
constexpr uint32_t kEmitSmall = 32768;
// out[0] = the group count; out[1 .. nzero) are cleared for the emission's
// NULL counts; out[nzero], out[nzero + 1] receive *flag and *ovf (0 for a null
// pointer) so that one copy brings every scalar back.
__global__ void __launch_bounds__(1024) group_rows_kernel(const uint64_t* st, uint32_t S, int all_slots,
                                                          uint32_t* slot_of, unsigned long long* out, uint32_t nzero,
                                                          const unsigned int* flag, const unsigned int* ovf) {
    __shared__ uint32_t ws[kMaxWaves + 1];
    const uint32_t t = threadIdx.x, lane = lane_id(), w = t / kWave, nw = blockDim.x / kWave;
    for (uint32_t i = 1 + t; i < nzero; i += blockDim.x) out[i] = 0;
    if (t == 0 && nzero) {
        out[nzero] = flag ? *flag : 0u;
        out[nzero + 1] = ovf ? *ovf : 0u;
    }
    // Thread t owns the k consecutive slots from t * k (k <= 32 for S <=
    // kEmitSmall): their row counters are read eight at a time, the owned
    // groups counted, and one block scan of the counts places every thread's
    // groups, in slot order.
    const uint32_t k = (S + blockDim.x - 1) / blockDim.x;
    const uint32_t s0 = t * k;
    uint32_t bits = 0;
    for (uint32_t j0 = 0; j0 < k; j0 += 8) {
        uint64_t r[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t s = s0 + j0 + u;
            r[u] = st[j0 + u < k && s < S ? s : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t s = s0 + j0 + u;
            if (j0 + u < k && s < S && (all_slots || r[u] > 0)) bits |= 1u << (j0 + u);
        }
    }
    const uint32_t cnt = static_cast<uint32_t>(__popc(bits));
    uint32_t x = cnt;  // inclusive wave scan of the counts
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<unsigned>(off)) x += y;
    }
    if (lane == kWave - 1) ws[w] = x;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (uint32_t q = 0; q < nw; ++q) {
            const uint32_t c = ws[q];
            ws[q] = run;
            run += c;
        }
        ws[kMaxWaves] = run;
    }
    __syncthreads();
    uint32_t at = ws[w] + x - cnt;
    while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1;
        slot_of[at++] = s0 + j;
    }
    if (t == 0) out[0] = ws[kMaxWaves];
}

// Output validity words of each column (kernel argument: no upload).
struct OutWords {
    uint64_t* w[kMaxKeys + kMaxAggs];
};

__global__ void __launch_bounds__(256) emit_rows_kernel(const uint64_t* st, uint32_t S, AggProg p, KeyEmit ke,
                                                        const uint32_t* slot_of, OutCols oc, OutWords vwords,
                                                        unsigned long long* out) {
    const uint64_t G = out[0];
    if (uint64_t(blockIdx.x) * blockDim.x >= G) return;  // uniform across the block
    const unsigned lane = lane_id();
    const int ncol = ke.nkeys + p.n;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = g < G;
    const uint32_t vb = in ? emit_row(st, S, p, ke, oc, slot_of[g], g) : 0u;
    const uint64_t inmask = __ballot(in);
    for (int c = 0; c < ncol; ++c) {
        const uint64_t bits = __ballot(!in || ((vb >> c) & 1u));
        if (lane == 0 && (inmask & 1ull)) {
            vwords.w[c][g >> 6] = bits;
            const unsigned long long z = static_cast<unsigned long long>(__popcll(inmask & ~bits));
            if (z) atomicAdd(&out[1 + c], z);
        }
    }
}

static int emit_type(const AggProg& p, const KeyEmit& ke, int c) {
    if (c < ke.nkeys) return ke.types[c];
    const int a = c - ke.nkeys;
    return agg_result_type(p.func[a], p.kind[a] == K_COUNT_STAR ? QE_I64 : p.in[a].type);
}

// Small slot counts: columns allocated for every slot, one kernel, one read-back.
static qe_dev_table* emit_groups_small(qe_gpu_ctx* ctx, const uint64_t* st, uint32_t S, const AggProg& p,
                                       const KeyEmit& ke, bool all_slots, const unsigned int* flag,
                                       unsigned int* flag_out) {
    const int ncol = ke.nkeys + p.n;
    std::vector<qe_dev_col*> cols;
    try {
        OutCols oc{};
        OutWords vw{};
        for (int c = 0; c < ncol; ++c) {
            const int type = emit_type(p, ke, c);
            cols.push_back(new_col(ctx, type, S, true));
            oc.data[c] = cols.back()->data;
            oc.type[c] = type;
            vw.w[c] = cols.back()->validity;
        }
        // res: group count | NULL count per column | *flag | *fovf
        Scratch res(ctx, size_t(3 + ncol) * 8), slot_of(ctx, size_t(S) * 4);
        group_rows_kernel<<<1, 1024, 0, ctx->stream>>>(st, S, all_slots ? 1 : 0, slot_of.as<uint32_t>(),
                                                       res.as<unsigned long long>(), static_cast<uint32_t>(1 + ncol),
                                                       flag, p.fovf);
        emit_rows_kernel<<<(S + 255) / 256, 256, 0, ctx->stream>>>(st, S, p, ke, slot_of.as<uint32_t>(), oc, vw,
                                                                   res.as<unsigned long long>());
        QE_HIP(hipGetLastError());
        uint64_t* h = static_cast<uint64_t*>(pinned(ctx, size_t(3 + ncol) * 8));
        QE_HIP(hipMemcpyAsync(h, res.p, size_t(3 + ncol) * 8, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
        if (flag_out) *flag_out = static_cast<unsigned int>(h[1 + ncol]);
        if (h[2 + ncol]) throw ScaleRetry();
        const uint64_t G = h[0];
        qe_dev_table* t = new_table(G);
        for (int c = 0; c < ncol; ++c) {
            qe_dev_col* col = cols[c];
            col->len = G;
            if (h[1 + c] == 0) {  // no NULL: drop the bitmap (all_valid)
                dev_free(ctx, col->validity);
                col->validity = nullptr;
                col->all_valid = true;
            }
            t->cols.push_back(col);
        }
        return t;
    } catch (...) {
        for (auto* c : cols) free_col(ctx, c);
        throw;
    }
}

qe_dev_table* emit_groups(qe_gpu_ctx* ctx, const uint64_t* st, uint64_t S, const AggProg& p,
                          const KeyEmit& ke, bool all_slots, const unsigned int* flag, unsigned int* flag_out) {
    fold_outliers(ctx, p, const_cast<uint64_t*>(st), S);
    if (S > 0 && S <= kEmitSmall)
        return emit_groups_small(ctx, st, static_cast<uint32_t>(S), p, ke, all_slots, flag, flag_out);
    if (flag_out) *flag_out = flag ? read_scalar(ctx, flag) : 0u;
    if (p.fovf && read_scalar(ctx, p.fovf)) throw ScaleRetry();
    Scratch flags(ctx, S * 4), offs(ctx, S * 8);
    group_flags_kernel<<<grid_for(S, 256, 8192), 256, 0, ctx->stream>>>(st, S, flags.as<uint32_t>(),
                                                                         all_slots ? 1 : 0);
    QE_HIP(hipGetLastError());
    const uint64_t G = exclusive_scan_u32(ctx, flags.as<uint32_t>(), offs.as<uint64_t>(), S);
    qe_dev_table* t = new_table(G);
    const int ncol = ke.nkeys + p.n;
    OutCols oc{};
    std::vector<Scratch> vb;
    vb.reserve(ncol);
    try {
        for (int c = 0; c < ncol; ++c) {
            int type;
            if (c < ke.nkeys) {
                type = ke.types[c];
            } else {
                const int a = c - ke.nkeys;
                type = agg_result_type(p.func[a], p.kind[a] == K_COUNT_STAR ? QE_I64 : p.in[a].type);
            }
            qe_dev_col* col = new_col(ctx, type, G, true);
            t->cols.push_back(col);
            vb.emplace_back(ctx, std::max<uint64_t>(G, 1));
            oc.data[c] = col->data;
            oc.valid[c] = vb.back().as<uint8_t>();
            oc.type[c] = type;
        }
        if (G > 0) {
            emit_kernel<<<grid_for(S, 256, 8192), 256, 0, ctx->stream>>>(st, S, p, ke, flags.as<uint32_t>(),
                                                                          offs.as<uint64_t>(), oc);
            QE_HIP(hipGetLastError());
        }
        std::vector<const uint8_t*> vbp(ncol);
        for (int c = 0; c < ncol; ++c) vbp[c] = vb[c].as<uint8_t>();
        pack_validity_many(ctx, t->cols.data(), vbp.data(), ncol);
    } catch (...) {
        qe_gpu_table_free(ctx, t);
        throw;
    }
    return t;
}

// ---- direct-slot kernels -----------------------------------------------------------
struct DirectSlot {
    ColView key;
    int64_t kmin;
    uint32_t null_slot;
    __device__ uint32_t operator()(uint64_t r) const {
        if (!is_valid(key, r)) return null_slot;
        return static_cast<uint32_t>(static_cast<uint64_t>(int_value(key, r) - kmin));
    }
};
struct GlobalSlot {
    __device__ uint32_t operator()(uint64_t) const { return 0; }
};

template <class SlotFn>
__global__ void __launch_bounds__(512) agg_lds_kernel(uint64_t n, SlotFn slot_of, uint32_t S, AggProg p,
                                                      uint64_t* part) {
    extern __shared__ __align__(16) char smem[];
    lds_init(smem, S, p);
    __syncthreads();
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        lds_update(smem, S, p, slot_of(r), vrow);
    }
    __syncthreads();
    lds_flush(smem, S, p, part);
}

template <class SlotFn>
__global__ void agg_global_kernel(uint64_t n, SlotFn slot_of, uint64_t S, AggProg p, uint64_t* st) {
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        global_update(st, S, p, slot_of(r), vrow);
    }
}

template <class SlotFn>
static void run_slots(qe_gpu_ctx* ctx, uint64_t n, SlotFn fn, uint32_t S, AggProg& p, uint64_t* st,
                      const char* timer, bool exact, Scratch& fbuf) {
    const size_t lb = lds_bytes(S, p);
    if (lb <= kLdsBudget) {
        const unsigned block = 512;
        const unsigned per_cu = std::max<unsigned>(1, static_cast<unsigned>((160 * 1024) / std::max<size_t>(lb, 1)));
        // At least kMinRowsPerBlock rows per block: each block writes a full
        // set of partials (S slots) for the merge, which on a small input
        // would cost more than its rows. On 2^20 rows and 1024 groups, 4096
        // rows per block (256 blocks) ran 0.120-0.123 ms per query against
        // 0.122-0.127 at 2048, 0.125-0.132 at 8192 and 0.145-0.151 at 16384.
        constexpr uint64_t kMinRowsPerBlock = 4096;
        unsigned blocks = static_cast<unsigned>(std::min<uint64_t>(
            uint64_t(ctx->num_cus) * std::min(per_cu, 4u), (n + kMinRowsPerBlock - 1) / kMinRowsPerBlock));
        if (blocks == 0) blocks = 1;
        prepare_f64(ctx, p, st, S, n, exact, fbuf);
        Scratch part(ctx, size_t(blocks) * part_words(S, p) * 8);
        {
            TimedRegion tr(ctx, timer);
            agg_lds_kernel<SlotFn><<<blocks, block, lb, ctx->stream>>>(n, fn, S, p, part.as<uint64_t>());
            QE_HIP(hipGetLastError());
        }
        merge_partials(ctx, part.as<uint64_t>(), blocks, st, S, p);
    } else {
        init_state(ctx, st, S, p);
        prepare_f64(ctx, p, st, S, n, exact, fbuf);
        TimedRegion tr(ctx, timer);
        agg_global_kernel<SlotFn><<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(n, fn, S, p, st);
        QE_HIP(hipGetLastError());
    }
}

struct HashSlot {
    HashKeys hk;
    HashTableView t;
    __device__ uint32_t operator()(uint64_t r) const {
        uint64_t w[kMaxKeys];
        uint64_t nm, h;
        normalize_row(hk, r, w, nm, h);
        return ht_insert(t, hk.nkeys, w, nm, h);
    }
};

__global__ void agg_hash_kernel(uint64_t n, HashSlot fn, uint64_t S, AggProg p, uint64_t* st) {
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = fn(r);
        if (s == kNoSlot) continue;
        for (int a = 0; a < p.n; ++a) vrow[a] = r;
        global_update(st, S, p, s, vrow);
    }
}

bool is_int_like(int t) { return t == QE_I32 || t == QE_I64 || t == QE_TS || t == QE_BOOL || t == QE_STR; }

}  // namespace qe_gpu

// Direct-slot limit: at most this many slots before falling back to hashing.
static constexpr uint64_t kDirectMaxSlots = uint64_t(1) << 24;

qe_gpu_status qe_gpu::aggregate_codes(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                      const uint32_t* key_cols, uint32_t nkeys, const qe_agg_spec* aggs,
                                      uint32_t naggs, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        invalid(ncols >= 1, "aggregate needs an input with >=1 column");
        invalid(aggs != nullptr || naggs == 0, "null aggregate list");
        check(nkeys <= kMaxKeys, QE_GPU_UNSUPPORTED, "too many group keys for the device path");
        const uint64_t n = common_len(cols, ncols);
        for (uint32_t k = 0; k < nkeys; ++k) invalid(key_cols[k] < ncols, "group key column out of range");
        ColView ins[kMaxAggs];
        int types[kMaxAggs];
        for (uint32_t a = 0; a < naggs && a < uint32_t(kMaxAggs); ++a) {
            if (aggs[a].func == QE_AGG_COUNT_STAR) {
                ins[a] = ColView{nullptr, nullptr, QE_I64, 0};
                types[a] = QE_I64;
                continue;
            }
            invalid(aggs[a].input_col < ncols, "aggregate input column out of range");
            ins[a] = view(cols[aggs[a].input_col]);
            types[a] = cols[aggs[a].input_col]->type;
        }
        const AggProg p0 = make_prog(aggs, naggs, ins, types);
        KeyEmit ke{};
        ke.nkeys = static_cast<int>(nkeys);
        for (uint32_t k = 0; k < nkeys; ++k) ke.types[k] = cols[key_cols[k]]->type;
        // F64 sums scaled from a sample; a value past the sampled bound makes
        // the emission throw ScaleRetry and the query runs again with bounds
        // taken from every row.
        auto run = [&](bool exact) -> qe_dev_table* {
            AggProg p = p0;
            Scratch fbuf;
            if (nkeys == 0) {
                // Global aggregate: exactly one output row (ops/aggregate.h:20-41).
                Scratch st(ctx, state_words(1, p) * 8);
                run_slots(ctx, n, GlobalSlot{}, 1, p, st.as<uint64_t>(), "agg_global", exact, fbuf);
                ke.mode = 0;
                ke.null_slot = kNoSlot;
                return emit_groups(ctx, st.as<uint64_t>(), 1, p, ke, true);
            }
            const qe_dev_col* k0 = cols[key_cols[0]];
            if (nkeys == 1 && is_int_like(k0->type)) {
                const MinMax mm = column_minmax(ctx, k0, n);
                const uint64_t range = mm.valid ? static_cast<uint64_t>(mm.mx) - static_cast<uint64_t>(mm.mn) + 1 : 0;
                if (range < kDirectMaxSlots && range <= 4 * n + 1024) {
                    const uint32_t S = static_cast<uint32_t>(range + 1);  // + the NULL slot
                    Scratch st(ctx, state_words(S, p) * 8);
                    DirectSlot fn{view(k0), mm.valid ? mm.mn : 0, S - 1};
                    run_slots(ctx, n, fn, S, p, st.as<uint64_t>(), "agg_direct", exact, fbuf);
                    ke.mode = 0;
                    ke.kmin = mm.valid ? mm.mn : 0;
                    ke.null_slot = S - 1;
                    return emit_groups(ctx, st.as<uint64_t>(), S, p, ke, false);
                }
            }
            // Hash path.
            uint64_t cap = 1024;
            while (cap < 2 * n) cap <<= 1;
            check(cap <= (uint64_t(1) << 31), QE_GPU_UNSUPPORTED, "group-by input too large for one device table");
            HashSlot fn{};
            fn.hk.nkeys = static_cast<int>(nkeys);
            for (uint32_t k = 0; k < nkeys; ++k) fn.hk.key[k] = view(cols[key_cols[k]]);
            Scratch state(ctx, cap * 4), hash(ctx, cap * 8), words(ctx, cap * 8 * nkeys), nmask(ctx, cap * 8),
                ovf(ctx, 8), st(ctx, state_words(cap, p) * 8);
            memset_async(ctx, state.p, 0, cap * 4);
            memset_async(ctx, ovf.p, 0, 8);
            fn.t = HashTableView{state.as<uint32_t>(), hash.as<uint64_t>(), words.as<uint64_t>(),
                                 nmask.as<uint64_t>(), cap, ovf.as<unsigned int>()};
            init_state(ctx, st.as<uint64_t>(), cap, p);
            prepare_f64(ctx, p, st.as<uint64_t>(), cap, n, exact, fbuf);
            {
                TimedRegion tr(ctx, "agg_hash");
                agg_hash_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(n, fn, cap, p,
                                                                                               st.as<uint64_t>());
                QE_HIP(hipGetLastError());
            }
            check(read_scalar(ctx, ovf.as<unsigned int>()) == 0, QE_GPU_OVERFLOW, "device group table overflow");
            ke.mode = 1;
            ke.words = words.as<uint64_t>();
            ke.nullmask = nmask.as<uint64_t>();
            return emit_groups(ctx, st.as<uint64_t>(), cap, p, ke, false);
        };
        try {
            *out = run(false);
        } catch (const ScaleRetry&) {
            *out = run(true);
        }
    });
}

// STR group keys and MIN/MAX over STR compare by value (strings.hip), then the
// group-by proper.
extern "C" qe_gpu_status qe_gpu_aggregate(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                          const uint32_t* key_cols, uint32_t nkeys,
                                          const qe_agg_spec* aggs, uint32_t naggs, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        invalid(cols != nullptr && ncols >= 1, "aggregate needs an input with >=1 column");
        invalid(aggs != nullptr || naggs == 0, "null aggregate list");
        invalid(key_cols != nullptr || nkeys == 0, "null key list");
        OwnedCols own(ctx);
        std::vector<const qe_dev_col*> c2(cols, cols + ncols);
        std::vector<qe_agg_spec> a2(aggs, aggs + naggs);
        StrAggFix fix;
        const bool any = str_agg_prepare(ctx, c2, key_cols, nkeys, a2, fix, own);
        qe_dev_table* t = nullptr;
        const qe_gpu_status st = aggregate_codes(ctx, c2.data(), static_cast<uint32_t>(c2.size()), key_cols, nkeys,
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
