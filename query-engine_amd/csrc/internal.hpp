// Shared internals of libqe_gpu: context, device columns, caching allocator,
// error plumbing, device helpers (hash, key canonicalisation, wave64 ops) and
// the device-wide scan. gfx950 only: wave64, 160 KiB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "qe_gpu.h"

namespace qe_gpu {

constexpr int kWave = 64;
constexpr int kMaxWaves = 16;  // waves in a 1024-thread block
constexpr uint32_t kNoRow = 0xFFFFFFFFu;
constexpr uint64_t kNullWord = 0x9E3779B97F4A7C15ull;  // NULL lane word (never compared alone)
constexpr uint64_t kCanonNaN = 0x7ff8000000000000ull;
constexpr uint64_t kHashSeed = 0x51E57A11AB1E5EEDull;

// ---------------------------------------------------------------------------
// Errors: every entry point returns a status; the message lives in the ctx.
struct Error {
    qe_gpu_status status;
    std::string msg;
};

#define QE_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::qe_gpu::Error{e_ == hipErrorOutOfMemory ? QE_GPU_OUT_OF_MEMORY    \
                                                            : QE_GPU_HIP_ERROR,       \
                                  std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

inline void check(bool ok, qe_gpu_status st, const std::string& msg) {
    if (!ok) throw Error{st, msg};
}
inline void invalid(bool ok, const std::string& msg) { check(ok, QE_GPU_INVALID_ARGUMENT, msg); }

inline size_t type_width(int t) {
    switch (t) {
        case QE_I32:
        case QE_STR:
            return 4;
        case QE_BOOL:
            return 1;
        default:
            return 8;
    }
}
inline bool valid_type(int t) { return t >= QE_I32 && t <= QE_STR; }
__host__ __device__ inline uint64_t words_for(uint64_t n) { return (n + 63) / 64; }

}  // namespace qe_gpu

// ---------------------------------------------------------------------------
// Opaque handle bodies (C ABI names).
struct qe_gpu_dict;

struct qe_dev_col {
    int32_t type = QE_I64;
    uint64_t len = 0;
    void* data = nullptr;
    uint64_t* validity = nullptr;  // nullptr <=> all_valid
    bool all_valid = true;
    bool owns = true;  // false for slices
    const qe_gpu_dict* dict = nullptr;  // STR: the dictionary its codes index (not owned)
};

// A string dictionary (core/string_dict.h): string i = bytes[off[i], off[i+1]).
// Derived tables are built on first use and kept with it.
struct qe_gpu_dict {
    std::vector<char> bytes;
    std::vector<uint32_t> off;  // n + 1
    uint32_t n = 0;
    bool unique = true;           // every value appears under one code
    std::vector<int32_t> canon;   // code -> first code with the same value
    std::unordered_map<std::string_view, int32_t> index;  // value -> first code
    std::vector<int32_t> rank;     // code -> rank of its value in byte order
    std::vector<int32_t> by_rank;  // rank -> first code with that value
    std::string_view at(uint32_t c) const { return std::string_view(bytes.data() + off[c], off[c + 1] - off[c]); }
};

struct qe_dev_table {
    uint64_t rows = 0;
    std::vector<qe_dev_col*> cols;
};

struct qe_dev_pairs {
    uint64_t count = 0;
    uint32_t* probe = nullptr;
    uint32_t* build = nullptr;
};

struct qe_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    uint64_t mem_limit = 0;
    uint64_t mem_in_use = 0;  // held from the device: live + cached blocks
    uint64_t live_bytes = 0;  // live blocks only
    std::multimap<size_t, void*> free_blocks;  // size -> block (cached)
    std::map<void*, size_t> live_blocks;       // block -> size
    bool timing = false;
    struct Timing {
        double ms = 0;
        uint64_t launches = 0;
    };
    std::map<std::string, Timing> timings;
    // Timed regions whose end event may still be pending (resolved when the
    // timings are read, so a region adds no synchronisation of its own).
    struct PendingRegion {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<PendingRegion> pending_regions;
    int num_cus = 256;
    // Tuning knobs (qe_gpu_set_tuning).
    int tune_join_pairs = 4;  // row pairs per thread per tile in the streaming probe
    int tune_asof_bucket_knots = 1;  // knots per learned-router bucket (set at train time)
    int tune_asof_data_lines = 1;    // learned index: bucket rows inline where they fit (set at train time)
    uint32_t tune_join_spill_cap = 0;  // spill-list records per block (0 = sized from the probe)
    int tune_join_blocks_per_cu = 1;   // streaming-probe blocks per CU (as many as LDS allows)
    // Small pinned host staging buffer for scalars read back after kernels.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // Recorded after a read-back copy that later work on the stream need not
    // delay (column_minmax_start / _finish).
    hipEvent_t readback_event = nullptr;
};

namespace qe_gpu {

// ---- caching allocator (per context, stream-ordered use only) ---------------
void* dev_alloc(qe_gpu_ctx* ctx, size_t bytes);
void dev_free(qe_gpu_ctx* ctx, void* p);
// Whether `bytes` more of device memory can be taken now: within the context's
// limit, and free on the device (counting this context's cached blocks) with
// 256 MiB to spare. For optional working sets that have a leaner fallback.
bool device_room(qe_gpu_ctx* ctx, uint64_t bytes);

// RAII scratch buffer.
struct Scratch {
    qe_gpu_ctx* ctx = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    Scratch() = default;
    Scratch(qe_gpu_ctx* c, size_t b) : ctx(c), p(b ? dev_alloc(c, b) : nullptr), bytes(b) {}
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    Scratch(Scratch&& o) noexcept : ctx(o.ctx), p(o.p), bytes(o.bytes) { o.p = nullptr; }
    Scratch& operator=(Scratch&& o) noexcept {
        release();
        ctx = o.ctx;
        p = o.p;
        bytes = o.bytes;
        o.p = nullptr;
        return *this;
    }
    ~Scratch() { release(); }
    void release() {
        if (p) dev_free(ctx, p);
        p = nullptr;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

qe_dev_col* new_col(qe_gpu_ctx* ctx, int type, uint64_t len, bool nullable);
void free_col(qe_gpu_ctx* ctx, qe_dev_col* c);
void* pinned(qe_gpu_ctx* ctx, size_t bytes);  // ctx-owned pinned staging
void memset_async(qe_gpu_ctx* ctx, void* p, int v, size_t bytes);  // via fill kernel
template <typename T>
T read_scalar(qe_gpu_ctx* ctx, const T* dev) {
    T* h = static_cast<T*>(pinned(ctx, sizeof(T)));
    QE_HIP(hipMemcpyAsync(h, dev, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    QE_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
}

// Kernel timing bracket (events on the ctx stream); the elapsed time is
// taken when the timings are read (resolve_timings), so a region does not
// synchronise the stream.
void resolve_timings(qe_gpu_ctx* ctx);

struct TimedRegion {
    qe_gpu_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedRegion(qe_gpu_ctx* c, const char* n);
    ~TimedRegion();
};

// ---- column views passed to kernels ------------------------------------------
struct ColView {
    const void* data;
    const uint64_t* validity;  // nullptr when all_valid
    int32_t type;
    int32_t pad;
};
inline ColView view(const qe_dev_col* c) {
    ColView v;
    v.data = c->data;
    v.validity = c->all_valid ? nullptr : c->validity;
    v.type = c->type;
    v.pad = 0;
    return v;
}

// ---- device helpers ------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t x) {
    // splitmix64 finalizer (ops/hash_kernels.h:7-12 documents the same mixer).
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
__host__ __device__ inline uint64_t hash_combine(uint64_t acc, uint64_t w) {
    return mix64(acc ^ mix64(w));
}

__device__ inline bool is_valid(const ColView& c, uint64_t r) {
    return c.validity == nullptr || ((c.validity[r >> 6] >> (r & 63)) & 1ull);
}

// Canonical key word (same normalisation as ops/hash_internal.cpp:29-60).
__device__ inline uint64_t canon_f64_bits(uint64_t b) {
    const uint64_t mag = b & 0x7fffffffffffffffull;
    if (mag > 0x7ff0000000000000ull) return kCanonNaN;  // any NaN
    if (mag == 0) return 0;                             // +-0
    return b;
}
__device__ inline uint64_t key_word(const ColView& c, uint64_t r) {
    switch (c.type) {
        case QE_I32:
        case QE_STR:
            return static_cast<uint32_t>(static_cast<const int32_t*>(c.data)[r]);
        case QE_BOOL:
            return static_cast<const uint8_t*>(c.data)[r];
        case QE_F64:
            return canon_f64_bits(static_cast<const uint64_t*>(c.data)[r]);
        default:
            return static_cast<const uint64_t*>(c.data)[r];
    }
}
// Integer value (I32 sign-extended) for SUM/MIN/MAX of integer families.
__device__ inline int64_t int_value(const ColView& c, uint64_t r) {
    switch (c.type) {
        case QE_I32:
        case QE_STR:
            return static_cast<const int32_t*>(c.data)[r];
        case QE_BOOL:
            return static_cast<const uint8_t*>(c.data)[r];
        default:
            return static_cast<const int64_t*>(c.data)[r];
    }
}
// int_value for a column read once, in order (a probe side): non-temporal
// loads, which leave the caches to data that is reused.
__device__ inline int64_t int_value_stream(const ColView& c, uint64_t r) {
    switch (c.type) {
        case QE_I32:
        case QE_STR:
            return __builtin_nontemporal_load(static_cast<const int32_t*>(c.data) + r);
        case QE_BOOL:
            return __builtin_nontemporal_load(static_cast<const uint8_t*>(c.data) + r);
        default:
            return __builtin_nontemporal_load(static_cast<const int64_t*>(c.data) + r);
    }
}
__device__ inline double f64_value(const ColView& c, uint64_t r) {
    return static_cast<const double*>(c.data)[r];
}

// Order-preserving u64 image of a double under the NaN-greatest total order
// (ops/agg_internal.h:34-40): NaN is canonicalised to +qNaN, which lands
// above +inf. -0.0 maps just below +0.0 (numerically equal values).
__host__ __device__ inline uint64_t f64_order_key(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    if ((b & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) b = kCanonNaN;
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double f64_from_order_key(uint64_t k) {
    const uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

__device__ inline int lane_id() { return __lane_id(); }

// ---- typed column access -----------------------------------------------------------
// Streaming kernels specialise on element width and issue every load of a row
// group before using any of them: unconditional loads at a row index clamped to
// the range, through global-address-space pointers (generic loads would also
// hold up the LDS wait counter). A column without NULLs reads one all-valid
// word instead of a bitmap (index mask 0).
template <typename T>
using gptr = const T __attribute__((address_space(1)))*;

template <typename T>
__device__ inline gptr<T> gcast(const void* p) {
    return (gptr<T>)(p);
}

static __device__ const uint64_t kAllValidWord = ~0ull;

struct BitsView {
    gptr<uint64_t> w;
    uint64_t mask;  // word index mask: ~0, or 0 for kAllValidWord
};
__device__ inline BitsView bits_view(const uint64_t* v) {
    return v ? BitsView{gcast<uint64_t>(v), ~0ull} : BitsView{gcast<uint64_t>(&kAllValidWord), 0ull};
}
__device__ inline uint32_t bit_at(const BitsView& b, uint64_t r) {
    return static_cast<uint32_t>((b.w[(r >> 6) & b.mask] >> (r & 63)) & 1ull);
}

// 64-bit pattern of a W-byte element: BOOL zero-extended, I32/STR sign-extended,
// I64/TS/F64 raw (the integer reading of int_value, F64 bits unchanged).
template <int W>
__device__ inline uint64_t elem_bits(const void* base, uint64_t r) {
    if constexpr (W == 8) return gcast<uint64_t>(base)[r];
    else if constexpr (W == 4) return static_cast<uint64_t>(static_cast<int64_t>(gcast<int32_t>(base)[r]));
    else return gcast<uint8_t>(base)[r];
}


// In-place exclusive scan of a[0, n) (n <= PER * blockDim.x) by one block;
// returns the total. `ws` holds kMaxWaves + 1 words of scratch.
template <int PER = 4>
__device__ inline uint32_t block_exclusive_scan(uint32_t* a, uint32_t n, uint32_t* ws) {
    const uint32_t t = threadIdx.x, lane = lane_id(), w = t / kWave, nw = blockDim.x / kWave;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t i = PER * t + k;
        v[k] = i < n ? a[i] : 0u;
        sum += v[k];
    }
    uint32_t x = sum;  // inclusive wave scan of the per-thread sums
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= static_cast<unsigned>(off)) x += y;
    }
    if (lane == kWave - 1) ws[w] = x;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < nw; ++k) {
            const uint32_t c = ws[k];
            ws[k] = run;
            run += c;
        }
        ws[kMaxWaves] = run;
    }
    __syncthreads();
    uint32_t base = ws[w] + x - sum;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t i = PER * t + k;
        if (i < n) a[i] = base;
        base += v[k];
    }
    const uint32_t total = ws[kMaxWaves];
    __syncthreads();
    return total;
}

// ---- launch geometry ----------------------------------------------------------
inline unsigned grid_for(uint64_t n, unsigned block, unsigned max_blocks = 65535u * 16) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return static_cast<unsigned>(g);
}

// ---- device-wide exclusive scan of u32 flags/counts into u64 offsets ---------
// out[i] = sum_{j<i} in[j]; returns the total (synchronises the stream).
uint64_t exclusive_scan_u32(qe_gpu_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n);

// ---- dictionary-encoded strings (strings.hip) ---------------------------------
// out[r] = table[codes[r]] as `out_type` (I32 or STR), NULL rows stay NULL; a
// valid code outside the table is an invalid_argument error.
qe_dev_col* remap_col(qe_gpu_ctx* ctx, const qe_dev_col* col, const std::vector<int32_t>& table, int out_type,
                      const qe_gpu_dict* out_dict);
const qe_gpu_dict* need_dict(const qe_dev_col* c);  // a STR column's dictionary, or invalid_argument
// Join-key id tables: build codes -> canonical codes; probe codes -> the build
// dictionary's code of the same value, or build.n + code when the build side
// lacks the value (never matches).
std::vector<int32_t> probe_ids(const qe_gpu_dict* build, const qe_gpu_dict* probe);

// Columns owned for the duration of one entry point.
struct OwnedCols {
    qe_gpu_ctx* ctx;
    std::vector<qe_dev_col*> cols;
    explicit OwnedCols(qe_gpu_ctx* c) : ctx(c) {}
    qe_dev_col* keep(qe_dev_col* c) {
        cols.push_back(c);
        return c;
    }
    ~OwnedCols();
};

// STR join keys made comparable by value: returns the columns to hash in place
// of (probe, build) key j (either unchanged or remapped id columns owned by `own`).
void str_join_keys(qe_gpu_ctx* ctx, const qe_dev_col* const* pk, const qe_dev_col* const* bk, uint32_t nkeys,
                   std::vector<const qe_dev_col*>& pk_out, std::vector<const qe_dev_col*>& bk_out, OwnedCols& own);

// An aggregate's STR handling: group keys on a dictionary with repeated values
// read canonical codes; MIN/MAX over STR read value ranks. After the aggregate,
// str_agg_finish gives STR key outputs their dictionary and maps rank outputs
// back to codes.
struct StrAggFix {
    std::vector<const qe_gpu_dict*> key_dict;  // per group key: STR dictionary or null
    std::vector<const qe_gpu_dict*> rank_dict; // per aggregate: STR MIN/MAX dictionary or null
};
// Rewrites `cols` (rank columns appended) and `aggs` for the STR rules above;
// returns whether any applied.
bool str_agg_prepare(qe_gpu_ctx* ctx, std::vector<const qe_dev_col*>& cols, const uint32_t* keys, uint32_t nkeys,
                     std::vector<qe_agg_spec>& aggs, StrAggFix& fix, OwnedCols& own);
void str_agg_finish(qe_gpu_ctx* ctx, qe_dev_table* t, const StrAggFix& fix);
// The group-by proper over prepared columns (no STR rewriting).
qe_gpu_status aggregate_codes(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols, const uint32_t* key_cols,
                              uint32_t nkeys, const qe_agg_spec* aggs, uint32_t naggs, qe_dev_table** out);
qe_dev_col* gather_col(qe_gpu_ctx* ctx, const qe_dev_col* src, const uint32_t* idx, uint64_t m);
// Join pairs from per-probe matches (kNoRow = no match), in row order: every
// row when keep_all, else the matched ones; probe/build are dev_alloc'ed.
uint64_t select_pairs(qe_gpu_ctx* ctx, const uint32_t* match, uint64_t n, bool keep_all, uint32_t** probe,
                      uint32_t** build);
// Column c of the result = srcs[c] gathered by idxs[c] (m rows each), with one
// host read-back for all of their NULL counts.
std::vector<qe_dev_col*> gather_cols(qe_gpu_ctx* ctx, const qe_dev_col* const* srcs, const uint32_t* const* idxs,
                                     size_t n, uint64_t m);
// Device-only variant (no allocation, no host read; safe under graph
// capture): `sums` holds scan_sums_words(n) words.
uint64_t scan_sums_words(uint64_t n);
void exclusive_scan_u32_dev(qe_gpu_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* sums);
// Resolves, once per device and with that device current, the device address
// of the no-allocation filter's static scratch array (called at context
// creation).
void noalloc_setup(int device);
// Same for u64 inputs (in-place allowed).
uint64_t exclusive_scan_u64(qe_gpu_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t n);

// ---- min/max of an integer-like column over valid rows ------------------------
struct MinMax {
    int64_t mn, mx;
    uint64_t valid;  // number of valid rows
};
MinMax column_minmax(qe_gpu_ctx* ctx, const qe_dev_col* c, uint64_t n);
// Several columns of length n, one read-back.
void column_minmax_n(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncol, uint64_t n, MinMax* out);
// The same in two steps, so the caller can queue device work that does not
// depend on the ranges between them: start() queues the kernels and the copy
// to pinned memory, finish() waits for the copy alone (an event recorded
// after it) and fills out.
struct MinMaxReadback {
    uint32_t ncol = 0;
    uint64_t n = 0;
    int64_t* h = nullptr;
};
MinMaxReadback column_minmax_start(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncol, uint64_t n);
void column_minmax_finish(qe_gpu_ctx* ctx, const MinMaxReadback& rb, MinMax* out);

// Stable LSD radix sort of (u64 key, u32 value) pairs on bits [begin_bit,
// end_bit), in place (sort.hip).
void radix_sort_pairs(qe_gpu_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, unsigned begin_bit,
                      unsigned end_bit);

// Column length helper with validation.
// Length shared by the columns of one input; rejects 2^32-1 rows or more
// (check_rows32).
uint64_t common_len(const qe_dev_col* const* cols, uint32_t n);
void check_rows32(uint64_t n);

// Gather rows idx[0, m) of src into a new column (idx == kNoRow gives NULL).

// Result column construction helpers (host-side allocation of result tables).
qe_dev_table* new_table(uint64_t rows);

// Pack per-row valid bytes (1 = valid) into the column's u64 validity words and
// apply the all_valid rule (core/column.h:42-54): a column with no NULL drops
// its bitmap. The column must have been allocated nullable.
void pack_validity(qe_gpu_ctx* ctx, qe_dev_col* col, const uint8_t* valid_bytes);
// The same for up to kMaxPackCols columns of one length: one launch, one read-back.
constexpr int kMaxPackCols = 32;
void pack_validity_many(qe_gpu_ctx* ctx, qe_dev_col* const* cols, const uint8_t* const* valid_bytes, int ncol);

}  // namespace qe_gpu

// Entry-point wrapper: converts exceptions into statuses + ctx messages.
#define QE_ENTRY(ctx, ...)                                             \
    do {                                                               \
        if (!(ctx)) return QE_GPU_INVALID_ARGUMENT;                    \
        try {                                                          \
            __VA_ARGS__;                                               \
            return QE_GPU_OK;                                          \
        } catch (const ::qe_gpu::Error& e_) {                          \
            (ctx)->last_error = e_.msg;                                \
            return e_.status;                                          \
        } catch (const std::bad_alloc&) {                              \
            (ctx)->last_error = "host allocation failed";              \
            return QE_GPU_OUT_OF_MEMORY;                               \
        } catch (const std::exception& e_) {                           \
            (ctx)->last_error = e_.what();                             \
            return QE_GPU_HIP_ERROR;                                   \
        }                                                              \
    } while (0)
