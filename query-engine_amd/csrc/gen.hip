// Seeded synthetic workloads generated directly in HBM. Every value is a pure
// function of (seed, stream, row) through a counter-based splitmix64 hash, so
// a rerun with the same seed reproduces the same tables bit for bit.
//
//  * group-by: key uniform in [0, key_range), value uniform [0,1000) with
//    Bernoulli(null_frac) NULLs (the 1M-row CPU group-by shape);
//  * join build: unique keys = a Feistel permutation of [0, n) (cycle walking
//    for non power-of-two n), group column uniform in [0, groups);
//  * join probe: with probability miss_frac a key in [n, 2n) (never matches),
//    else Zipf(s) over the build domain (value k-1 for rank k), payload uniform
//    [0,1000);
//  * as-of: per-symbol row counts proportional to Zipf(s) over nsym symbols;
//    the right side stored symbol-major with strictly increasing microsecond
//    timestamps (exponential gaps, mean 1 ms, blocks of 512 events at 10x rate
//    covering about 10% of the time), price uniform [1,200); left rows draw a
//    symbol from Zipf(s) and a time uniform over that symbol's span +-5%.
// Zipf variates use rejection-inversion (W. Hormann, G. Derflinger, "Rejection-
// inversion to generate variates from monotone discrete distributions", ACM
// TOMACS 6(3), 1996), O(1) expected work per sample.
#include "internal.hpp"

#include <cmath>
#include <vector>

using namespace qe_gpu;

namespace qe_gpu {

__host__ __device__ inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t row) {
    return mix64(mix64(seed ^ (stream * 0x9E3779B97F4A7C15ull)) ^ (row * 0xD1B54A32D192ED03ull + stream));
}
__host__ __device__ inline double unif(uint64_t seed, uint64_t stream, uint64_t row) {
    return static_cast<double>(rnd(seed, stream, row) >> 11) * (1.0 / 9007199254740992.0);
}

// ---- Zipf (rejection-inversion) ---------------------------------------------------
struct Zipf {
    double s, n, h_x1, h_n, sconst;
};
__host__ __device__ inline double zh_helper1(double x) {  // log1p(x)/x
    return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__host__ __device__ inline double zh_helper2(double x) {  // expm1(x)/x
    return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__host__ __device__ inline double zH(double s, double x) {  // integral of x^-s
    const double lx = log(x);
    return zh_helper2((1.0 - s) * lx) * lx;
}
__host__ __device__ inline double zHinv(double s, double x) {
    double t = x * (1.0 - s);
    if (t < -1.0) t = -1.0;  // guard rounding at the domain edge
    return exp(zh_helper1(t) * x);
}
__host__ __device__ inline double zh(double s, double x) { return exp(-s * log(x)); }

inline Zipf make_zipf(double s, uint64_t n) {
    Zipf z;
    z.s = s;
    z.n = static_cast<double>(n);
    z.h_x1 = zH(s, 1.5) - 1.0;
    z.h_n = zH(s, z.n + 0.5);
    z.sconst = 2.0 - zHinv(s, zH(s, 2.5) - zh(s, 2.0));
    return z;
}

// Rank in [1, n]. The rejection loop accepts with probability > 0.9 per try;
// after 64 tries the last candidate is kept (probability < 1e-60).
__device__ inline uint64_t zipf_sample(const Zipf& z, uint64_t seed, uint64_t stream, uint64_t row) {
    double k = 1.0;
    for (int it = 0; it < 64; ++it) {
        const double u = z.h_n + unif(seed, stream, row * 64 + it) * (z.h_x1 - z.h_n);
        const double x = zHinv(z.s, u);
        k = floor(x + 0.5);
        if (k < 1.0) k = 1.0;
        else if (k > z.n) k = z.n;
        if (k - x <= z.sconst || u >= zH(z.s, k + 0.5) - zh(z.s, k)) break;
    }
    return static_cast<uint64_t>(k);
}

// ---- Feistel permutation of [0, n) --------------------------------------------------
struct Feistel {
    uint64_t n;
    unsigned half;  // bits per half; domain 2^(2*half) >= n
    uint64_t key;
};
__device__ inline uint64_t feistel_once(const Feistel& f, uint64_t x) {
    const uint64_t m = (uint64_t(1) << f.half) - 1;
    uint64_t l = x >> f.half, r = x & m;
    for (int round = 0; round < 4; ++round) {
        const uint64_t nl = r;
        r = l ^ (mix64(r ^ (f.key + round * 0x632BE59BD9B4E019ull)) & m);
        l = nl;
    }
    return (l << f.half) | r;
}
// Cycle walking: x < n, so iterating the permutation of [0, 2^(2*half))
// returns to x's cycle member below n in at most the cycle length.
__device__ inline uint64_t feistel_perm(const Feistel& f, uint64_t x) {
    do {
        x = feistel_once(f, x);
    } while (x >= f.n);
    return x;
}

__global__ void gen_groupby_kernel(uint64_t seed, uint64_t n, uint64_t key_range, int64_t* key, double* val) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        key[r] = static_cast<int64_t>(rnd(seed, 1, r) % key_range);
        val[r] = unif(seed, 2, r) * 1000.0;
    }
}
__global__ void gen_validity_kernel(uint64_t seed, uint64_t stream, uint64_t n, double null_frac,
                                    uint64_t* words) {
    const uint64_t nw = words_for(n);
    for (uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; w < nw;
         w += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t bits = 0;
        for (unsigned b = 0; b < 64; ++b) {
            const uint64_t r = w * 64 + b;
            const bool v = r >= n || unif(seed, stream, r) >= null_frac;
            bits |= uint64_t(v) << b;
        }
        words[w] = bits;
    }
}
__global__ void gen_build_kernel(uint64_t seed, uint64_t n, uint32_t groups, Feistel f, int64_t* key,
                                 int32_t* grp) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        key[r] = static_cast<int64_t>(feistel_perm(f, r));
        grp[r] = static_cast<int32_t>(rnd(seed, 5, r) % groups);
    }
}
__global__ void gen_probe_kernel(uint64_t seed, uint64_t n, uint64_t build_n, Zipf z, double miss_frac,
                                 int64_t* key, double* pay) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        int64_t k;
        if (unif(seed, 6, r) < miss_frac)
            k = static_cast<int64_t>(build_n + rnd(seed, 7, r) % build_n);
        else
            k = static_cast<int64_t>(zipf_sample(z, seed, 8, r) - 1);
        key[r] = k;
        pay[r] = unif(seed, 9, r) * 1000.0;
    }
}

// ---- as-of ----------------------------------------------------------------------------
__device__ inline uint32_t sym_of_row(const uint64_t* offs, uint32_t nsym, uint64_t i) {
    uint32_t lo = 0, hi = nsym;  // offs[lo] <= i < offs[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (offs[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}
__global__ void gen_gaps_kernel(uint64_t seed, uint64_t n, const uint64_t* offs, uint32_t nsym, uint64_t* gap) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = sym_of_row(offs, nsym, i);
        const uint64_t k = i - offs[s];
        const bool burst = unif(seed, 11, (uint64_t(s) << 32) + k / 512) < 0.527;
        const double mean = burst ? 100.0 : 1000.0;
        const double g = -mean * log1p(-unif(seed, 12, i));
        uint64_t gi = static_cast<uint64_t>(g + 0.5);
        gap[i] = gi < 1 ? 1 : gi;
    }
}
__global__ void gen_right_kernel(uint64_t seed, uint64_t n, const uint64_t* offs, uint32_t nsym,
                                 const uint64_t* gap, const uint64_t* excl, int32_t* sym, int64_t* ts, double* px) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = sym_of_row(offs, nsym, i);
        const int64_t t0 = 1700000000000000ll + static_cast<int64_t>(rnd(seed, 13, s) % 3600000000ull);
        sym[i] = static_cast<int32_t>(s);
        ts[i] = t0 + static_cast<int64_t>(excl[i] + gap[i] - excl[offs[s]]);
        px[i] = 1.0 + 199.0 * unif(seed, 14, i);
    }
}
// Session-shaped times: each symbol's rows spread over `sessions` trading days
// (6.5 h of each 24 h) with a U-shaped intraday intensity, density
// proportional to 1 + 4 (2s - 1)^2 on the session fraction s (inverse CDF by
// bisection); + k keeps the times strictly increasing.
__device__ inline double ushape_inv(double u) {
    double lo = 0.0, hi = 1.0;
    for (int i = 0; i < 40; ++i) {
        const double m = 0.5 * (lo + hi), d = 2.0 * m - 1.0;
        const double F = (m + (2.0 / 3.0) * (d * d * d + 1.0)) * (3.0 / 7.0);
        if (F < u) lo = m;
        else hi = m;
    }
    return 0.5 * (lo + hi);
}
__global__ void gen_right_sessions_kernel(uint64_t seed, uint64_t n, const uint64_t* offs, uint32_t nsym,
                                          uint32_t sessions, int32_t* sym, int64_t* ts, double* px) {
    constexpr int64_t kDay = 86400000000ll, kSession = 23400000000ll;  // microseconds
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = sym_of_row(offs, nsym, i);
        const uint64_t k = i - offs[s], ns = offs[s + 1] - offs[s];
        const uint64_t per = (ns + sessions - 1) / sessions;
        const uint64_t d = k / per, kk = k - d * per;
        const double u = (static_cast<double>(kk) + unif(seed, 17, i)) / static_cast<double>(per);
        const int64_t t0 = 1700000000000000ll + static_cast<int64_t>(rnd(seed, 13, s) % 3600000000ull);
        sym[i] = static_cast<int32_t>(s);
        ts[i] = t0 + static_cast<int64_t>(d) * kDay + static_cast<int64_t>(ushape_inv(u) * kSession) +
                static_cast<int64_t>(kk);
        px[i] = 1.0 + 199.0 * unif(seed, 14, i);
    }
}
__global__ void gen_left_kernel(uint64_t seed, uint64_t n, Zipf z, const uint64_t* offs, const int64_t* rts,
                                int32_t* sym, int64_t* ts) {
    for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < n;
         j += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t s = static_cast<uint32_t>(zipf_sample(z, seed, 15, j) - 1);
        sym[j] = static_cast<int32_t>(s);
        const uint64_t b = offs[s], e = offs[s + 1];
        int64_t first, last;
        if (e > b) {
            first = rts[b];
            last = rts[e - 1];
        } else {
            first = 1700000000000000ll;
            last = first + 1000000;
        }
        const double len = static_cast<double>(last - first);
        ts[j] = first + static_cast<int64_t>(floor(-0.05 * len + unif(seed, 16, j) * 1.1 * len));
    }
}

}  // namespace qe_gpu

static void require(const qe_dev_col* c, int type, uint64_t n, const char* what) {
    invalid(c != nullptr, std::string("null column: ") + what);
    invalid(c->type == type, std::string("wrong column type: ") + what);
    invalid(c->len == n, std::string("wrong column length: ") + what);
}

extern "C" {

qe_gpu_status qe_gpu_gen_groupby(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint64_t key_range, double null_frac,
                                 qe_dev_col* key, qe_dev_col* val) {
    QE_ENTRY(ctx, {
        require(key, QE_I64, n, "key");
        require(val, QE_F64, n, "value");
        invalid(key_range >= 1, "key_range must be >= 1");
        if (n == 0) return QE_GPU_OK;
        gen_groupby_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
            seed, n, key_range, static_cast<int64_t*>(key->data), static_cast<double*>(val->data));
        if (null_frac > 0) {
            invalid(val->validity != nullptr, "value column must be nullable");
            gen_validity_kernel<<<grid_for(words_for(n), 256, 4096), 256, 0, ctx->stream>>>(seed, 3, n, null_frac,
                                                                                            val->validity);
            val->all_valid = false;
        }
        QE_HIP(hipGetLastError());
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_gen_join_build(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint32_t groups, qe_dev_col* key,
                                    qe_dev_col* grp) {
    QE_ENTRY(ctx, {
        require(key, QE_I64, n, "build key");
        require(grp, QE_I32, n, "build group");
        invalid(groups >= 1, "groups must be >= 1");
        if (n == 0) return QE_GPU_OK;
        unsigned bits = 2;
        while ((uint64_t(1) << bits) < n) ++bits;
        if (bits & 1) ++bits;
        Feistel f{n, bits / 2, mix64(seed ^ 0xFE157E1ull)};
        gen_build_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
            seed, n, groups, f, static_cast<int64_t*>(key->data), static_cast<int32_t*>(grp->data));
        QE_HIP(hipGetLastError());
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_gen_join_probe(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint64_t build_n, double zipf_s,
                                    double miss_frac, qe_dev_col* key, qe_dev_col* pay) {
    QE_ENTRY(ctx, {
        require(key, QE_I64, n, "probe key");
        require(pay, QE_F64, n, "probe payload");
        invalid(build_n >= 1 && zipf_s > 0, "bad generator parameters");
        if (n == 0) return QE_GPU_OK;
        const Zipf z = make_zipf(zipf_s, build_n);
        gen_probe_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
            seed, n, build_n, z, miss_frac, static_cast<int64_t*>(key->data), static_cast<double*>(pay->data));
        QE_HIP(hipGetLastError());
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_gen_asof(qe_gpu_ctx* ctx, uint64_t seed, uint64_t right_n, uint64_t left_n, uint32_t nsym,
                              double zipf_s, uint32_t sessions, qe_dev_col* r_sym, qe_dev_col* r_ts, qe_dev_col* r_px,
                              qe_dev_col* l_sym, qe_dev_col* l_ts) {
    QE_ENTRY(ctx, {
        require(r_sym, QE_I32, right_n, "right sym");
        require(r_ts, QE_I64, right_n, "right ts");
        require(r_px, QE_F64, right_n, "right px");
        require(l_sym, QE_I32, left_n, "left sym");
        require(l_ts, QE_I64, left_n, "left ts");
        invalid(nsym >= 1 && zipf_s > 0, "bad generator parameters");
        // Per-symbol row counts proportional to the Zipf pmf (remainder to the
        // head symbols), so the right side holds exactly right_n rows.
        std::vector<double> w(nsym);
        double tot = 0;
        for (uint32_t k = 0; k < nsym; ++k) tot += (w[k] = std::pow(double(k + 1), -zipf_s));
        std::vector<uint64_t> cnt(nsym), offs(nsym + 1, 0);
        uint64_t assigned = 0;
        for (uint32_t k = 0; k < nsym; ++k) assigned += (cnt[k] = static_cast<uint64_t>(right_n * (w[k] / tot)));
        for (uint64_t k = 0; assigned < right_n; k = (k + 1) % nsym, ++assigned) ++cnt[k];
        for (uint32_t k = 0; k < nsym; ++k) offs[k + 1] = offs[k] + cnt[k];
        Scratch doffs(ctx, (nsym + 1) * 8), gap(ctx, std::max<uint64_t>(right_n, 1) * 8),
            excl(ctx, std::max<uint64_t>(right_n, 1) * 8);
        QE_HIP(hipMemcpyAsync(doffs.p, offs.data(), (nsym + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
        if (right_n && sessions) {
            gen_right_sessions_kernel<<<grid_for(right_n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                seed, right_n, doffs.as<uint64_t>(), nsym, sessions, static_cast<int32_t*>(r_sym->data),
                static_cast<int64_t*>(r_ts->data), static_cast<double*>(r_px->data));
            QE_HIP(hipGetLastError());
        } else if (right_n) {
            gen_gaps_kernel<<<grid_for(right_n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                seed, right_n, doffs.as<uint64_t>(), nsym, gap.as<uint64_t>());
            QE_HIP(hipGetLastError());
            exclusive_scan_u64(ctx, gap.as<uint64_t>(), excl.as<uint64_t>(), right_n);
            gen_right_kernel<<<grid_for(right_n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                seed, right_n, doffs.as<uint64_t>(), nsym, gap.as<uint64_t>(), excl.as<uint64_t>(),
                static_cast<int32_t*>(r_sym->data), static_cast<int64_t*>(r_ts->data), static_cast<double*>(r_px->data));
            QE_HIP(hipGetLastError());
        }
        if (left_n) {
            const Zipf z = make_zipf(zipf_s, nsym);
            gen_left_kernel<<<grid_for(left_n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                seed, left_n, z, doffs.as<uint64_t>(), static_cast<const int64_t*>(r_ts->data),
                static_cast<int32_t*>(l_sym->data), static_cast<int64_t*>(l_ts->data));
            QE_HIP(hipGetLastError());
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
