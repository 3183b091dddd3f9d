// Context, device columns, caching allocator, result tables, scans and the
// small C-ABI plumbing shared by every operator.
#include "internal.hpp"

#include <algorithm>
#include <cstdio>

using namespace qe_gpu;

namespace qe_gpu {

// Drop every cached block (after the stream drains, since cached blocks may
// still be read by queued work).
static void release_cache(qe_gpu_ctx* ctx) {
    QE_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->free_blocks) {
        hipFree(kv.second);
        ctx->mem_in_use -= kv.first;
    }
    ctx->free_blocks.clear();
}

void* dev_alloc(qe_gpu_ctx* ctx, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    // The context memory limit bounds the bytes in use (cached blocks are
    // released before a fresh allocation would push the total past it).
    if (ctx->mem_limit && ctx->live_bytes + bytes > ctx->mem_limit)
        throw Error{QE_GPU_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) +
                                              " bytes exceeds the context memory limit"};
    // Best fit among cached blocks no more than 2x the request.
    auto it = ctx->free_blocks.lower_bound(bytes);
    if (it != ctx->free_blocks.end() && it->first <= bytes * 2) {
        void* p = it->second;
        ctx->live_blocks[p] = it->first;
        ctx->live_bytes += it->first;
        ctx->free_blocks.erase(it);
        return p;
    }
    if (ctx->mem_limit && ctx->mem_in_use + bytes > ctx->mem_limit) release_cache(ctx);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_cache(ctx);
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{QE_GPU_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed"};
        }
    }
    ctx->mem_in_use += bytes;
    ctx->live_blocks[p] = bytes;
    ctx->live_bytes += bytes;
    return p;
}

bool device_room(qe_gpu_ctx* ctx, uint64_t bytes) {
    if (ctx->mem_limit && ctx->live_bytes + bytes > ctx->mem_limit) return false;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const uint64_t cached = ctx->mem_in_use - ctx->live_bytes;  // released on demand by dev_alloc
    return uint64_t(free_b) + cached >= bytes + (uint64_t(256) << 20);
}

void dev_free(qe_gpu_ctx* ctx, void* p) {
    if (!p) return;
    auto it = ctx->live_blocks.find(p);
    if (it == ctx->live_blocks.end()) return;
    // Stream-ordered reuse: blocks go back to the cache; later work on the same
    // stream is ordered after the work that used them.
    ctx->live_bytes -= it->second;
    ctx->free_blocks.emplace(it->second, p);
    ctx->live_blocks.erase(it);
}

void* pinned(qe_gpu_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) QE_HIP(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        size_t b = std::max<size_t>(bytes, 4096);
        QE_HIP(hipHostMalloc(&ctx->pinned, b));
        ctx->pinned_bytes = b;
    }
    return ctx->pinned;
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void fill_u8_kernel(uint8_t* p, uint8_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

void memset_async(qe_gpu_ctx* ctx, void* p, int v, size_t bytes) {
    if (bytes == 0) return;
    // A fill kernel rather than hipMemsetAsync keeps this graph-capture safe.
    if ((reinterpret_cast<uintptr_t>(p) & 3) == 0 && (bytes & 3) == 0) {
        const uint32_t b = static_cast<uint8_t>(v);
        const uint32_t w = b | (b << 8) | (b << 16) | (b << 24);
        fill_u32_kernel<<<grid_for(bytes / 4, 256, 4096), 256, 0, ctx->stream>>>(
            static_cast<uint32_t*>(p), w, bytes / 4);
    } else {
        fill_u8_kernel<<<grid_for(bytes, 256, 4096), 256, 0, ctx->stream>>>(
            static_cast<uint8_t*>(p), static_cast<uint8_t>(v), bytes);
    }
    QE_HIP(hipGetLastError());
}

qe_dev_col* new_col(qe_gpu_ctx* ctx, int type, uint64_t len, bool nullable) {
    auto* c = new qe_dev_col();
    c->type = type;
    c->len = len;
    c->data = dev_alloc(ctx, std::max<size_t>(len * type_width(type), 16));
    if (nullable) {
        c->validity = static_cast<uint64_t*>(dev_alloc(ctx, std::max<size_t>(words_for(len) * 8, 8)));
        c->all_valid = false;
    }
    return c;
}

void free_col(qe_gpu_ctx* ctx, qe_dev_col* c) {
    if (!c) return;
    if (c->owns) {
        dev_free(ctx, c->data);
        dev_free(ctx, c->validity);
    }
    delete c;
}

qe_dev_table* new_table(uint64_t rows) {
    auto* t = new qe_dev_table();
    t->rows = rows;
    return t;
}

// One wave per validity word: lane l reads row 64w + l's byte, the word is the
// wave's ballot (padding bits past n set: the all-valid padding convention),
// and lane 0 counts the word's NULLs.
__device__ inline uint64_t pack_word(const uint8_t* vb, uint64_t n, uint64_t w, unsigned long long& nulls) {
    const uint64_t r = w * 64 + lane_id();
    const bool in = r < n;
    const bool v = !in || vb[r] != 0;
    const uint64_t bits = __ballot(v);
    const uint64_t inmask = __ballot(in);
    nulls += static_cast<unsigned long long>(__popcll(inmask & ~bits));
    return bits;
}

__global__ void pack_validity_kernel(const uint8_t* vb, uint64_t n, uint64_t* words,
                                     unsigned long long* nulls) {
    const uint64_t nw = words_for(n);
    const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x / kWave);
    unsigned long long local = 0;
    for (uint64_t w = blockIdx.x * uint64_t(blockDim.x / kWave) + threadIdx.x / kWave; w < nw; w += waves) {
        const uint64_t bits = pack_word(vb, n, w, local);
        if (lane_id() == 0) words[w] = bits;
    }
    if (lane_id() == 0 && local) atomicAdd(nulls, local);
}

// Several columns at once: one launch, one read-back of the NULL counts.
struct PackSet {
    int n;
    uint64_t len;
    const uint8_t* vb[kMaxPackCols];
    uint64_t* words[kMaxPackCols];
};

__global__ void pack_validity_many_kernel(PackSet ps, unsigned long long* nulls) {
    const uint64_t nw = words_for(ps.len);
    const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x / kWave);
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x / kWave) + threadIdx.x / kWave; t < nw * ps.n; t += waves) {
        const int c = static_cast<int>(t / nw);
        const uint64_t w = t - uint64_t(c) * nw;
        unsigned long long local = 0;
        const uint64_t bits = pack_word(ps.vb[c], ps.len, w, local);
        if (lane_id() == 0) {
            ps.words[c][w] = bits;
            if (local) atomicAdd(nulls + c, local);
        }
    }
}

void pack_validity_many(qe_gpu_ctx* ctx, qe_dev_col* const* cols, const uint8_t* const* valid_bytes, int ncol) {
    check(ncol <= kMaxPackCols, QE_GPU_UNSUPPORTED, "too many columns to pack");
    const uint64_t len = ncol ? cols[0]->len : 0;
    if (len == 0) {
        for (int c = 0; c < ncol; ++c) pack_validity(ctx, cols[c], valid_bytes[c]);
        return;
    }
    PackSet ps{};
    ps.n = ncol;
    ps.len = len;
    for (int c = 0; c < ncol; ++c) {
        invalid(cols[c]->len == len && cols[c]->validity != nullptr, "pack needs equal-length nullable columns");
        ps.vb[c] = valid_bytes[c];
        ps.words[c] = cols[c]->validity;
    }
    Scratch cnt(ctx, size_t(ncol) * 8);
    memset_async(ctx, cnt.p, 0, size_t(ncol) * 8);
    pack_validity_many_kernel<<<grid_for(words_for(len) * ncol * kWave, 256, 2048), 256, 0, ctx->stream>>>(
        ps, cnt.as<unsigned long long>());
    QE_HIP(hipGetLastError());
    uint64_t* h = static_cast<uint64_t*>(pinned(ctx, size_t(ncol) * 8));
    QE_HIP(hipMemcpyAsync(h, cnt.p, size_t(ncol) * 8, hipMemcpyDeviceToHost, ctx->stream));
    QE_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < ncol; ++c) {
        if (h[c] == 0) {  // no NULL: drop the bitmap (all_valid)
            dev_free(ctx, cols[c]->validity);
            cols[c]->validity = nullptr;
            cols[c]->all_valid = true;
        }
    }
}

void pack_validity(qe_gpu_ctx* ctx, qe_dev_col* col, const uint8_t* valid_bytes) {
    if (col->validity == nullptr) return;
    if (col->len == 0) {
        dev_free(ctx, col->validity);
        col->validity = nullptr;
        col->all_valid = true;
        return;
    }
    Scratch cnt(ctx, 8);
    memset_async(ctx, cnt.p, 0, 8);
    pack_validity_kernel<<<grid_for(words_for(col->len) * kWave, 256, 2048), 256, 0, ctx->stream>>>(
        valid_bytes, col->len, col->validity, cnt.as<unsigned long long>());
    QE_HIP(hipGetLastError());
    const uint64_t nulls = read_scalar(ctx, cnt.as<uint64_t>());
    if (nulls == 0) {
        dev_free(ctx, col->validity);
        col->validity = nullptr;
        col->all_valid = true;
    } else {
        col->all_valid = false;
    }
}

TimedRegion::TimedRegion(qe_gpu_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    QE_HIP(hipEventCreate(&a));
    QE_HIP(hipEventCreate(&b));
    QE_HIP(hipEventRecord(a, ctx->stream));
}
TimedRegion::~TimedRegion() {
    if (!ctx->timing || !a) return;
    hipEventRecord(b, ctx->stream);
    ctx->pending_regions.push_back({name, a, b});
    if (ctx->pending_regions.size() >= 4096) resolve_timings(ctx);
}

void resolve_timings(qe_gpu_ctx* ctx) {
    for (auto& r : ctx->pending_regions) {
        hipEventSynchronize(r.b);
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& t = ctx->timings[r.name];
        t.ms += ms;
        t.launches += 1;
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    ctx->pending_regions.clear();
}

uint64_t common_len(const qe_dev_col* const* cols, uint32_t n) {
    invalid(n == 0 || cols != nullptr, "null column array");
    uint64_t len = 0;
    for (uint32_t i = 0; i < n; ++i) {
        invalid(cols[i] != nullptr, "null column");
        if (i == 0) len = cols[i]->len;
        invalid(cols[i]->len == len, "columns of one input must have equal length");
    }
    check_rows32(len);
    return len;
}

void check_rows32(uint64_t n) {
    // Row ids, join pairs, permutations and first-seen ids are 32-bit with
    // 0xFFFFFFFF as the no-row sentinel: inputs of 2^32 - 1 rows or more
    // would alias it.
    check(n < kNoRow, QE_GPU_UNSUPPORTED, "inputs of 2^32-1 rows or more are not supported (row ids are 32-bit)");
}

// ---- scan ------------------------------------------------------------------------
constexpr unsigned kScanBlock = 256;
constexpr unsigned kScanItems = 16;
constexpr uint64_t kScanTile = uint64_t(kScanBlock) * kScanItems;

template <typename T>
__global__ void scan_tile_sums(const T* in, uint64_t n, uint64_t* sums) {
    __shared__ uint64_t red[kScanBlock / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    uint64_t s = 0;
    for (unsigned k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + uint64_t(k) * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane_id() == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (unsigned w = 0; w < kScanBlock / kWave; ++w) t += red[w];
        sums[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of sums[0..m) in place; total to sums[m].
__global__ void scan_sums_single(uint64_t* sums, uint64_t m) {
    __shared__ uint64_t buf[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < m; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t v = (i < m) ? sums[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (unsigned off = 1; off < 1024; off <<= 1) {
            uint64_t t = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < m) sums[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[m] = carry;
}

template <typename T>
__global__ void scan_tile_apply(const T* in, uint64_t* out, uint64_t n, const uint64_t* sums) {
    // Blocked arrangement: thread t owns items [t*K, t*K+K) of the tile.
    __shared__ uint64_t wsum[kScanBlock / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
    uint64_t v[kScanItems];
    uint64_t local = 0;
    for (unsigned k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k;
        v[k] = (i < n) ? static_cast<uint64_t>(in[i]) : 0;
        local += v[k];
    }
    // Exclusive scan of `local` across the block.
    uint64_t incl = local;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint64_t t = __shfl_up(incl, off);
        if (lane_id() >= off) incl += t;
    }
    const unsigned w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint64_t wbase = 0;
    for (unsigned j = 0; j < w; ++j) wbase += wsum[j];
    uint64_t run = sums[blockIdx.x] + wbase + incl - local;
    for (unsigned k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

template <typename T>
static uint64_t exclusive_scan_impl(qe_gpu_ctx* ctx, const T* in, uint64_t* out, uint64_t n) {
    if (n == 0) return 0;
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    Scratch sums(ctx, (tiles + 1) * sizeof(uint64_t));
    scan_tile_sums<T><<<tiles, kScanBlock, 0, ctx->stream>>>(in, n, sums.as<uint64_t>());
    scan_sums_single<<<1, 1024, 0, ctx->stream>>>(sums.as<uint64_t>(), tiles);
    scan_tile_apply<T><<<tiles, kScanBlock, 0, ctx->stream>>>(in, out, n, sums.as<uint64_t>());
    QE_HIP(hipGetLastError());
    return read_scalar(ctx, sums.as<uint64_t>() + tiles);
}

uint64_t exclusive_scan_u32(qe_gpu_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n) {
    return exclusive_scan_impl<uint32_t>(ctx, in, out, n);
}
uint64_t scan_sums_words(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }
void exclusive_scan_u32_dev(qe_gpu_ctx* ctx, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* sums) {
    if (n == 0) return;
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    scan_tile_sums<uint32_t><<<tiles, kScanBlock, 0, ctx->stream>>>(in, n, sums);
    scan_sums_single<<<1, 1024, 0, ctx->stream>>>(sums, tiles);
    scan_tile_apply<uint32_t><<<tiles, kScanBlock, 0, ctx->stream>>>(in, out, n, sums);
    QE_HIP(hipGetLastError());
}
uint64_t exclusive_scan_u64(qe_gpu_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t n) {
    if (in == out && n > 0) {
        // In place: copy first (the apply kernel reads `in` while writing `out`).
        Scratch tmp(ctx, n * 8);
        QE_HIP(hipMemcpyAsync(tmp.p, in, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        return exclusive_scan_impl<uint64_t>(ctx, tmp.as<uint64_t>(), out, n);
    }
    return exclusive_scan_impl<uint64_t>(ctx, in, out, n);
}

// ---- min/max over valid rows -----------------------------------------------------
// One launch for up to kMaxMinMaxCols columns (grid.y = column): per-block (min,
// max, count) of the valid values, with eight independent typed loads in flight
// per thread, then one block per column folds the partials (no contended
// atomics: 2048 blocks hitting one address serialise).
constexpr uint32_t kMaxMinMaxCols = 4;
struct MinMaxCols {
    ColView c[kMaxMinMaxCols];
    int32_t width[kMaxMinMaxCols];
};

template <int W>
__device__ inline void minmax_rows(const ColView& c, uint64_t n, int64_t& mn, int64_t& mx, uint64_t& cnt) {
    const BitsView vb = bits_view(c.validity);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += 8 * stride) {
        int64_t v[8];
        uint32_t ok[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t r = i + k * stride;
            const uint64_t rc = r < n ? r : n - 1;
            v[k] = static_cast<int64_t>(elem_bits<W>(c.data, rc));
            ok[k] = bit_at(vb, rc) & (r < n ? 1u : 0u);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!ok[k]) continue;
            mn = v[k] < mn ? v[k] : mn;
            mx = v[k] > mx ? v[k] : mx;
            ++cnt;
        }
    }
}

// Block-wide (min, max, count) into out[0..2] by thread 0.
__device__ inline void minmax_block_store(int64_t mn, int64_t mx, uint64_t cnt, int64_t* out) {
    __shared__ int64_t smn[kMaxWaves], smx[kMaxWaves];
    __shared__ uint64_t scnt[kMaxWaves];
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t a = __shfl_down(mn, off), b = __shfl_down(mx, off);
        const uint64_t d = __shfl_down(cnt, off);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
        cnt += d;
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        smn[w] = mn;
        smx[w] = mx;
        scnt[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < static_cast<int>(blockDim.x / kWave); ++k) {
            mn = smn[k] < mn ? smn[k] : mn;
            mx = smx[k] > mx ? smx[k] : mx;
            cnt += scnt[k];
        }
        out[0] = mn;
        out[1] = mx;
        out[2] = static_cast<int64_t>(cnt);
    }
}

__global__ void __launch_bounds__(256) minmax_kernel(MinMaxCols C, uint64_t n, int64_t* part) {
    const uint32_t y = blockIdx.y;
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    uint64_t cnt = 0;
    switch (C.width[y]) {
        case 1:
            minmax_rows<1>(C.c[y], n, mn, mx, cnt);
            break;
        case 4:
            minmax_rows<4>(C.c[y], n, mn, mx, cnt);
            break;
        default:
            minmax_rows<8>(C.c[y], n, mn, mx, cnt);
    }
    minmax_block_store(mn, mx, cnt, part + 3 * (size_t(y) * gridDim.x + blockIdx.x));
}
// Block y folds column y's nparts partials into out[3y..3y+2].
__global__ void __launch_bounds__(256) minmax_final(const int64_t* part, uint32_t nparts, int64_t* out) {
    const int64_t* p = part + 3 * size_t(blockIdx.x) * nparts;
    int64_t mn = INT64_MAX, mx = INT64_MIN;
    uint64_t cnt = 0;
    for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x) {
        mn = p[3 * b] < mn ? p[3 * b] : mn;
        mx = p[3 * b + 1] > mx ? p[3 * b + 1] : mx;
        cnt += static_cast<uint64_t>(p[3 * b + 2]);
    }
    minmax_block_store(mn, mx, cnt, out + 3 * blockIdx.x);
}

// XCD (accelerator complex die) that runs each block: bits [3:0] of the
// XCC_ID hardware register (s_getreg; id 20, offset 0, size 4).
__global__ void xcd_of_block_kernel(uint32_t* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = static_cast<uint32_t>(__builtin_amdgcn_s_getreg((3 << 11) | 20));
}

MinMax column_minmax(qe_gpu_ctx* ctx, const qe_dev_col* c, uint64_t n) {
    MinMax mm;
    column_minmax_n(ctx, &c, 1, n, &mm);
    return mm;
}

void column_minmax_n(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncol, uint64_t n, MinMax* out) {
    column_minmax_finish(ctx, column_minmax_start(ctx, cols, ncol, n), out);
}

MinMaxReadback column_minmax_start(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncol, uint64_t n) {
    check(ncol >= 1 && ncol <= kMaxMinMaxCols, QE_GPU_UNSUPPORTED, "min/max of too many columns at once");
    MinMaxReadback rb;
    rb.ncol = ncol;
    rb.n = n;
    if (n == 0) return rb;
    // About 8-16 rows per thread: one or two batches of eight independent loads.
    const uint32_t blocks = static_cast<uint32_t>(grid_for((n + 7) / 8, 256, ctx->num_cus * 16));
    Scratch part(ctx, size_t(blocks) * 3 * 8 * ncol), s(ctx, size_t(3) * 8 * ncol);
    MinMaxCols C{};
    for (uint32_t c = 0; c < ncol; ++c) {
        C.c[c] = view(cols[c]);
        C.width[c] = static_cast<int32_t>(type_width(cols[c]->type));
    }
    minmax_kernel<<<dim3(blocks, ncol), 256, 0, ctx->stream>>>(C, n, part.as<int64_t>());
    minmax_final<<<ncol, 256, 0, ctx->stream>>>(part.as<int64_t>(), blocks, s.as<int64_t>());
    QE_HIP(hipGetLastError());
    rb.h = static_cast<int64_t*>(pinned(ctx, size_t(24) * ncol));
    QE_HIP(hipMemcpyAsync(rb.h, s.p, size_t(24) * ncol, hipMemcpyDeviceToHost, ctx->stream));
    QE_HIP(hipEventRecord(ctx->readback_event, ctx->stream));
    return rb;
}

void column_minmax_finish(qe_gpu_ctx* ctx, const MinMaxReadback& rb, MinMax* out) {
    if (rb.n == 0) {
        for (uint32_t c = 0; c < rb.ncol; ++c) out[c] = MinMax{INT64_MAX, INT64_MIN, 0};
        return;
    }
    QE_HIP(hipEventSynchronize(ctx->readback_event));
    const int64_t* h = rb.h;
    for (uint32_t c = 0; c < rb.ncol; ++c) out[c] = MinMax{h[3 * c], h[3 * c + 1], static_cast<uint64_t>(h[3 * c + 2])};
}

}  // namespace qe_gpu

// =================================================================================
// C ABI
extern "C" {

qe_gpu_status qe_gpu_ctx_create(int device, qe_gpu_ctx** out) {
    if (!out) return QE_GPU_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return QE_GPU_HIP_ERROR;
    }
    if (device < 0 || device >= n) return QE_GPU_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return QE_GPU_HIP_ERROR;
    auto* ctx = new qe_gpu_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return QE_GPU_HIP_ERROR;
    }
    if (hipEventCreateWithFlags(&ctx->readback_event, hipEventDisableTiming) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return QE_GPU_HIP_ERROR;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    qe_gpu::noalloc_setup(device);
    *out = ctx;
    return QE_GPU_OK;
}

void qe_gpu_ctx_destroy(qe_gpu_ctx* ctx) {
    if (!ctx) return;
    hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->pending_regions) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto& kv : ctx->free_blocks) hipFree(kv.second);
    for (auto& kv : ctx->live_blocks) hipFree(kv.first);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    hipEventDestroy(ctx->readback_event);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

qe_gpu_status qe_gpu_device_info(qe_gpu_ctx* ctx, qe_gpu_device_info_t* out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "device info needs an output");
        hipDeviceProp_t prop;
        QE_HIP(hipGetDeviceProperties(&prop, ctx->device));
        *out = qe_gpu_device_info_t{};
        snprintf(out->name, sizeof(out->name), "%s", prop.name);
        if (out->name[0] == '\0' && hipDeviceGetName(out->name, sizeof(out->name), ctx->device) != hipSuccess)
            (void)hipGetLastError();
        snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
        out->compute_units = prop.multiProcessorCount;
        out->hbm_bytes = prop.totalGlobalMem;
        out->lds_bytes_per_block = prop.sharedMemPerBlock;
        out->l2_bytes = prop.l2CacheSize;
        out->clock_khz = prop.clockRate;
        out->memory_clock_khz = prop.memoryClockRate;
        out->memory_bus_bits = prop.memoryBusWidth;
        int v = 0;
        if (hipRuntimeGetVersion(&v) == hipSuccess) out->hip_runtime_version = v;
        if (hipDriverGetVersion(&v) == hipSuccess) out->hip_driver_version = v;
        out->device = ctx->device;
    });
}

const char* qe_gpu_last_error(const qe_gpu_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

qe_gpu_status qe_gpu_sync(qe_gpu_ctx* ctx) {
    QE_ENTRY(ctx, QE_HIP(hipStreamSynchronize(ctx->stream)));
}

qe_gpu_status qe_gpu_mem_usage(qe_gpu_ctx* ctx, uint64_t* live_bytes, uint64_t* held_bytes) {
    QE_ENTRY(ctx, {
        if (live_bytes) *live_bytes = ctx->live_bytes;
        if (held_bytes) *held_bytes = ctx->mem_in_use;
    });
}

qe_gpu_status qe_gpu_set_mem_limit(qe_gpu_ctx* ctx, uint64_t bytes) {
    QE_ENTRY(ctx, ctx->mem_limit = bytes);
}

qe_gpu_status qe_gpu_set_timing(qe_gpu_ctx* ctx, int enabled) {
    QE_ENTRY(ctx, ctx->timing = enabled != 0);
}

qe_gpu_status qe_gpu_xcd_of_blocks(qe_gpu_ctx* ctx, uint32_t nblocks, uint32_t* out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && nblocks >= 1 && nblocks <= 65536, "bad argument");
        Scratch d(ctx, size_t(nblocks) * 4);
        xcd_of_block_kernel<<<nblocks, 64, 0, ctx->stream>>>(d.as<uint32_t>());
        QE_HIP(hipGetLastError());
        QE_HIP(hipMemcpyAsync(out, d.p, size_t(nblocks) * 4, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_set_tuning(qe_gpu_ctx* ctx, int32_t key, int64_t value) {
    QE_ENTRY(ctx, {
        switch (key) {
            case QE_TUNE_JOIN_PAIRS:
                invalid(value == 2 || value == 4, "QE_TUNE_JOIN_PAIRS must be 2 or 4");
                ctx->tune_join_pairs = static_cast<int>(value);
                break;
            case QE_TUNE_ASOF_BUCKET_KNOTS:
                invalid(value >= 1 && value <= 64, "QE_TUNE_ASOF_BUCKET_KNOTS must be in [1, 64]");
                ctx->tune_asof_bucket_knots = static_cast<int>(value);
                break;
            case QE_TUNE_JOIN_SPILL_CAP:
                invalid(value >= 0 && value <= (int64_t(1) << 30), "QE_TUNE_JOIN_SPILL_CAP must be in [0, 2^30]");
                ctx->tune_join_spill_cap = static_cast<uint32_t>(value);
                break;
            case QE_TUNE_JOIN_BLOCKS_PER_CU:
                invalid(value >= 1 && value <= 8, "QE_TUNE_JOIN_BLOCKS_PER_CU must be in [1, 8]");
                ctx->tune_join_blocks_per_cu = static_cast<int>(value);
                break;
            case QE_TUNE_ASOF_DATA_LINES:
                invalid(value == 0 || value == 1, "QE_TUNE_ASOF_DATA_LINES must be 0 or 1");
                ctx->tune_asof_data_lines = static_cast<int>(value);
                break;
            default:
                invalid(false, "unknown tuning key");
        }
    });
}

qe_gpu_status qe_gpu_timing(qe_gpu_ctx* ctx, const char* kernel, double* ms, uint64_t* launches) {
    QE_ENTRY(ctx, {
        invalid(kernel != nullptr, "null kernel name");
        resolve_timings(ctx);
        auto it = ctx->timings.find(kernel);
        if (ms) *ms = (it == ctx->timings.end()) ? 0.0 : it->second.ms;
        if (launches) *launches = (it == ctx->timings.end()) ? 0 : it->second.launches;
        if (it != ctx->timings.end()) ctx->timings.erase(it);
    });
}

qe_gpu_status qe_gpu_column_alloc(qe_gpu_ctx* ctx, int32_t type, uint64_t len, int32_t nullable,
                                  qe_dev_col** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        invalid(valid_type(type), "unknown column type");
        *out = new_col(ctx, type, len, nullable != 0);
        if (nullable) memset_async(ctx, (*out)->validity, 0xFF, words_for(len) * 8);
    });
}

qe_gpu_status qe_gpu_column_upload(qe_gpu_ctx* ctx, const qe_col_desc* d, qe_dev_col** out) {
    QE_ENTRY(ctx, {
        invalid(d != nullptr && out != nullptr, "null descriptor");
        invalid(valid_type(d->type), "unknown column type");
        invalid(d->len == 0 || d->data != nullptr, "null data with nonzero length");
        // core/column.h:42-54: (all_valid == false, validity == nullptr) is invalid.
        invalid(d->all_valid || d->validity != nullptr,
                "column claims nulls but has no validity bitmap");
        const bool nullable = !d->all_valid;
        qe_dev_col* c = new_col(ctx, d->type, d->len, nullable);
        const size_t bytes = d->len * type_width(d->type);
        if (bytes) QE_HIP(hipMemcpyAsync(c->data, d->data, bytes, hipMemcpyHostToDevice, ctx->stream));
        if (nullable && d->len) {
            QE_HIP(hipMemcpyAsync(c->validity, d->validity, words_for(d->len) * 8,
                                  hipMemcpyHostToDevice, ctx->stream));
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *out = c;
    });
}

qe_gpu_status qe_gpu_column_download(qe_gpu_ctx* ctx, const qe_dev_col* c, void* data,
                                     uint64_t* validity, int32_t* all_valid) {
    QE_ENTRY(ctx, {
        invalid(c != nullptr, "null column");
        const size_t bytes = c->len * type_width(c->type);
        if (data && bytes)
            QE_HIP(hipMemcpyAsync(data, c->data, bytes, hipMemcpyDeviceToHost, ctx->stream));
        if (validity && c->len) {
            if (c->all_valid) {
                QE_HIP(hipStreamSynchronize(ctx->stream));
                for (uint64_t w = 0; w < words_for(c->len); ++w) validity[w] = ~0ull;
            } else {
                QE_HIP(hipMemcpyAsync(validity, c->validity, words_for(c->len) * 8,
                                      hipMemcpyDeviceToHost, ctx->stream));
            }
        }
        if (all_valid) *all_valid = c->all_valid ? 1 : 0;
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_column_download_range(qe_gpu_ctx* ctx, const qe_dev_col* c, uint64_t offset,
                                           uint64_t count, void* data, uint8_t* valid_bytes) {
    QE_ENTRY(ctx, {
        invalid(c != nullptr, "null column");
        invalid(offset + count <= c->len, "range out of bounds");
        const size_t w = type_width(c->type);
        if (data && count)
            QE_HIP(hipMemcpyAsync(data, static_cast<const char*>(c->data) + offset * w, count * w,
                                  hipMemcpyDeviceToHost, ctx->stream));
        std::vector<uint64_t> words;
        const uint64_t w0 = offset / 64, w1 = words_for(offset + count);
        if (valid_bytes && count && !c->all_valid) {
            words.resize(w1 - w0);
            QE_HIP(hipMemcpyAsync(words.data(), c->validity + w0, (w1 - w0) * 8,
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        if (valid_bytes) {
            for (uint64_t r = 0; r < count; ++r) {
                if (c->all_valid) {
                    valid_bytes[r] = 1;
                } else {
                    const uint64_t g = offset + r;
                    valid_bytes[r] = (words[(g >> 6) - w0] >> (g & 63)) & 1ull;
                }
            }
        }
    });
}

void qe_gpu_column_free(qe_gpu_ctx* ctx, qe_dev_col* c) {
    if (ctx) free_col(ctx, c);
}
uint64_t qe_gpu_column_len(const qe_dev_col* c) { return c ? c->len : 0; }
int32_t qe_gpu_column_type(const qe_dev_col* c) { return c ? c->type : -1; }
int32_t qe_gpu_column_all_valid(const qe_dev_col* c) { return c ? (c->all_valid ? 1 : 0) : 0; }

qe_gpu_status qe_gpu_column_slice(qe_gpu_ctx* ctx, const qe_dev_col* c, uint64_t offset, uint64_t len,
                                  qe_dev_col** out) {
    QE_ENTRY(ctx, {
        invalid(c != nullptr && out != nullptr, "null column");
        invalid(offset % 64 == 0, "slice offset must be a multiple of 64");
        invalid(offset + len <= c->len, "slice out of bounds");
        auto* s = new qe_dev_col();
        s->type = c->type;
        s->len = len;
        s->data = static_cast<char*>(c->data) + offset * type_width(c->type);
        s->all_valid = c->all_valid;
        s->validity = c->all_valid ? nullptr : c->validity + offset / 64;
        s->owns = false;
        s->dict = c->dict;
        *out = s;
    });
}

uint64_t qe_gpu_table_rows(const qe_dev_table* t) { return t ? t->rows : 0; }
uint32_t qe_gpu_table_cols(const qe_dev_table* t) {
    return t ? static_cast<uint32_t>(t->cols.size()) : 0;
}
const qe_dev_col* qe_gpu_table_col(const qe_dev_table* t, uint32_t i) {
    return (t && i < t->cols.size()) ? t->cols[i] : nullptr;
}
void qe_gpu_table_free(qe_gpu_ctx* ctx, qe_dev_table* t) {
    if (!ctx || !t) return;
    for (auto* c : t->cols) free_col(ctx, c);
    delete t;
}

uint64_t qe_gpu_pairs_count(const qe_dev_pairs* p) { return p ? p->count : 0; }

qe_gpu_status qe_gpu_pairs_download(qe_gpu_ctx* ctx, const qe_dev_pairs* p, uint64_t offset,
                                    uint64_t count, uint32_t* probe_rows, uint32_t* build_rows) {
    QE_ENTRY(ctx, {
        invalid(p != nullptr, "null pairs");
        invalid(offset + count <= p->count, "pair range out of bounds");
        if (count && probe_rows)
            QE_HIP(hipMemcpyAsync(probe_rows, p->probe + offset, count * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (count && build_rows)
            QE_HIP(hipMemcpyAsync(build_rows, p->build + offset, count * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void qe_gpu_pairs_free(qe_gpu_ctx* ctx, qe_dev_pairs* p) {
    if (!ctx || !p) return;
    dev_free(ctx, p->probe);
    dev_free(ctx, p->build);
    delete p;
}

qe_gpu_status qe_gpu_malloc(qe_gpu_ctx* ctx, uint64_t bytes, void** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr, "null out");
        *out = dev_alloc(ctx, bytes);
    });
}
qe_gpu_status qe_gpu_free(qe_gpu_ctx* ctx, void* p) { QE_ENTRY(ctx, dev_free(ctx, p)); }
qe_gpu_status qe_gpu_memcpy_d2h(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    QE_ENTRY(ctx, {
        if (bytes) QE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) stream_copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                           uint64_t n) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = __builtin_nontemporal_load(src + i);
}

qe_gpu_status qe_gpu_stream_copy(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes, double* ms) {
    QE_ENTRY(ctx, {
        invalid(dst && src && bytes % 16 == 0, "stream copy needs 16-byte multiples");
        hipEvent_t a, b;
        QE_HIP(hipEventCreate(&a));
        QE_HIP(hipEventCreate(&b));
        const uint64_t n = bytes / 16;
        QE_HIP(hipEventRecord(a, ctx->stream));
        stream_copy_kernel<<<grid_for(n, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
            static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), n);
        QE_HIP(hipGetLastError());
        QE_HIP(hipEventRecord(b, ctx->stream));
        QE_HIP(hipEventSynchronize(b));
        float t = 0;
        QE_HIP(hipEventElapsedTime(&t, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
        if (ms) *ms = t;
    });
}

// Read-only stream: four independent 16-byte nontemporal loads in flight per
// thread per iteration, folded by xor. The store fires only when every bit of
// the fold is set, so it moves no bytes in practice while keeping the loads
// live; sink holds 64 u32x4.
__global__ void __launch_bounds__(256) stream_read_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ sink,
                                                          uint64_t n) {
    const uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    u32x4 acc = {0, 0, 0, 0};
    for (; i + 3 * G < n; i += 4 * G) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        const u32x4 b = __builtin_nontemporal_load(src + i + G);
        const u32x4 c = __builtin_nontemporal_load(src + i + 2 * G);
        const u32x4 d = __builtin_nontemporal_load(src + i + 3 * G);
        acc ^= a ^ b ^ c ^ d;
    }
    for (; i < n; i += G) acc ^= __builtin_nontemporal_load(src + i);
    if ((acc.x & acc.y & acc.z & acc.w) == 0xFFFFFFFFu) sink[threadIdx.x & 63] = acc;
}

qe_gpu_status qe_gpu_stream_read(qe_gpu_ctx* ctx, const void* src, uint64_t bytes, void* sink, double* ms) {
    QE_ENTRY(ctx, {
        invalid(src && sink && bytes % 16 == 0, "stream read needs 16-byte multiples and a sink");
        hipEvent_t a, b;
        QE_HIP(hipEventCreate(&a));
        QE_HIP(hipEventCreate(&b));
        const uint64_t n = bytes / 16;
        QE_HIP(hipEventRecord(a, ctx->stream));
        stream_read_kernel<<<grid_for(n, 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
            static_cast<const u32x4*>(src), static_cast<u32x4*>(sink), n);
        QE_HIP(hipGetLastError());
        QE_HIP(hipEventRecord(b, ctx->stream));
        QE_HIP(hipEventSynchronize(b));
        float t = 0;
        QE_HIP(hipEventElapsedTime(&t, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
        if (ms) *ms = t;
    });
}

// The map lookups of the fused join probe alone: 16-byte nontemporal loads of
// key pairs (four pairs in flight per thread), each key's offset clamped into
// [0, range) and its 16-bit map entry gathered, the entries xor-folded. No
// payload stream and no aggregation: the time this takes over a probe column
// and a map of the probe's size is the floor the fused probe cannot beat.
__global__ void __launch_bounds__(256) map_gather_kernel(const int64_t* __restrict__ keys, uint64_t n, int64_t kmin,
                                                         uint64_t range, const uint16_t* __restrict__ map,
                                                         u32x4* __restrict__ sink) {
    typedef long long i64x2 __attribute__((ext_vector_type(2)));
    const i64x2* kp = reinterpret_cast<const i64x2*>(keys);
    const uint64_t np = n / 2;
    const uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    uint32_t acc = 0;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < np; i += 4 * G) {
        i64x2 k[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t r = i + j * G;
            k[j] = __builtin_nontemporal_load(kp + (r < np ? r : np - 1));
        }
        uint16_t v[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t o0 = static_cast<uint64_t>(k[j].x - kmin), o1 = static_cast<uint64_t>(k[j].y - kmin);
            v[2 * j] = map[o0 < range ? o0 : range - 1];
            v[2 * j + 1] = map[o1 < range ? o1 : range - 1];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc = (acc << 1 | acc >> 31) ^ v[j];
    }
    if (n & 1) {
        const uint64_t o = static_cast<uint64_t>(keys[n - 1] - kmin);
        acc ^= map[o < range ? o : range - 1];
    }
    if (acc == 0xFFFFFFFFu) sink[threadIdx.x & 63] = u32x4{acc, acc, acc, acc};
}

qe_gpu_status qe_gpu_map_gather(qe_gpu_ctx* ctx, const qe_dev_col* key, int64_t kmin, uint64_t range,
                                const uint16_t* map, void* sink, double* ms) {
    QE_ENTRY(ctx, {
        invalid(key && map && sink && range > 0, "map gather needs a key column, a map and a sink");
        invalid(key->type == QE_I64 && reinterpret_cast<uintptr_t>(key->data) % 16 == 0,
                "map gather reads 16-byte aligned I64 keys");
        const int64_t* keys = static_cast<const int64_t*>(key->data);
        const uint64_t n = key->len;
        hipEvent_t a, b;
        QE_HIP(hipEventCreate(&a));
        QE_HIP(hipEventCreate(&b));
        QE_HIP(hipEventRecord(a, ctx->stream));
        map_gather_kernel<<<grid_for(std::max<uint64_t>(n / 2, 1), 256, ctx->num_cus * 8), 256, 0, ctx->stream>>>(
            keys, n, kmin, range, map, static_cast<u32x4*>(sink));
        QE_HIP(hipGetLastError());
        QE_HIP(hipEventRecord(b, ctx->stream));
        QE_HIP(hipEventSynchronize(b));
        float t = 0;
        QE_HIP(hipEventElapsedTime(&t, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
        if (ms) *ms = t;
    });
}

qe_gpu_status qe_gpu_memcpy_h2d(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    QE_ENTRY(ctx, {
        if (bytes) QE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
