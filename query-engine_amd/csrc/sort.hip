// Stable LSD radix sort of (u64 key, u32 value) pairs — the device Sort that
// orders the as-of build side by (partition, timestamp), replacing the
// reference's byte-radix path (ops/sort.cpp:125-188) for the key shapes the
// as-of join needs (tsx/asof.cpp:163-166).
//
// One pass per 8-bit digit over bits [begin_bit, end_bit):
//   1. tile histograms (tile = 256 threads x 16 keys) -> hist[digit][tile];
//   2. exclusive scan of hist in digit-major order -> global offsets;
//   3. scatter: each tile ranks its keys stably within their digit (wave64
//      ballot "match" per 64 keys, waves folded in order through LDS), stages
//      the tile in LDS in digit order, then writes each digit run contiguously
//      (coalesced) to its global offset.
// Stability across tiles comes from the digit-major scan; within a tile from
// the (iteration, wave, lane) position order used everywhere below.
#include "internal.hpp"
#include "tile_rank.hpp"

#include <algorithm>

using namespace qe_gpu;

namespace qe_gpu {

constexpr unsigned kSortBlock = kRankBlock;
constexpr unsigned kSortItems = 16;
constexpr unsigned kSortTile = kSortBlock * kSortItems;  // 4096 keys
constexpr unsigned kRadix = 256;

// Position p of the tile is item i = p / 256, wave w = (p % 256) / 64, lane.
__device__ inline uint64_t tile_pos(unsigned i) { return uint64_t(i) * kSortBlock + threadIdx.x; }

__global__ void __launch_bounds__(kSortBlock) radix_hist_kernel(const uint64_t* keys, uint64_t n, unsigned shift,
                                                                uint32_t* hist, uint64_t ntiles) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = uint64_t(blockIdx.x) * kSortTile;
    for (unsigned i = 0; i < kSortItems; ++i) {
        const uint64_t p = base + tile_pos(i);
        if (p < n) atomicAdd(&h[(keys[p] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[uint64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kSortBlock) radix_scatter_kernel(const uint64_t* keys, const uint32_t* vals,
                                                                   uint64_t n, unsigned shift,
                                                                   const uint64_t* offs, uint64_t ntiles,
                                                                   uint64_t* keys_out, uint32_t* vals_out) {
    __shared__ TileRankLds L;
    __shared__ uint64_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    __shared__ uint16_t srank[kSortTile];  // rank of each key within its digit in the tile
    __shared__ uint64_t doff[kRadix];      // this tile's output offset of each digit

    const uint64_t tile = die_tile(ntiles);
    if (tile >= ntiles) return;
    const uint64_t base = tile * kSortTile;
    doff[threadIdx.x] = offs[uint64_t(threadIdx.x) * ntiles + tile];  // read before the ranking's barriers
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t pos[kSortItems];
    auto live = [&](unsigned i) { return base + tile_row<kSortItems>(i) < n; };
    for (unsigned i = 0; i < kSortItems; ++i) {
        const uint64_t p = base + tile_row<kSortItems>(i);
        k[i] = live(i) ? keys[p] : ~0ull;
        v[i] = live(i) ? vals[p] : 0u;
    }
    auto digit = [&](unsigned i) { return static_cast<uint32_t>((k[i] >> shift) & 0xFF); };
    tile_digit_ranks<kSortItems>(L, digit, live, pos);
    for (unsigned i = 0; i < kSortItems; ++i) {
        if (!live(i)) continue;
        const unsigned s = pos[i];
        skeys[s] = k[i];
        svals[s] = v[i];
        srank[s] = static_cast<uint16_t>(s - L.start[digit(i)]);
    }
    __syncthreads();
    const uint64_t live_n = (n - base < kSortTile) ? (n - base) : kSortTile;
    for (unsigned s = threadIdx.x; s < live_n; s += kSortBlock) {
        const uint64_t key = skeys[s];
        const unsigned dg = static_cast<unsigned>((key >> shift) & 0xFF);
        const uint64_t dst = doff[dg] + srank[s];
        keys_out[dst] = key;
        vals_out[dst] = svals[s];
    }
}

// Sort pairs by bits [begin_bit, end_bit) of the keys. Uses (and may return)
// the alternate buffers; the sorted result is written back to keys/vals.
void radix_sort_pairs(qe_gpu_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, unsigned begin_bit,
                      unsigned end_bit) {
    if (n <= 1 || end_bit <= begin_bit) return;
    const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
    Scratch k2(ctx, n * 8), v2(ctx, n * 4), hist(ctx, ntiles * kRadix * 4), offs(ctx, ntiles * kRadix * 8);
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = k2.as<uint64_t>();
    uint32_t* vout = v2.as<uint32_t>();
    unsigned passes = 0;
    for (unsigned shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
        radix_hist_kernel<<<ntiles, kSortBlock, 0, ctx->stream>>>(kin, n, shift, hist.as<uint32_t>(), ntiles);
        QE_HIP(hipGetLastError());
        exclusive_scan_u32(ctx, hist.as<uint32_t>(), offs.as<uint64_t>(), ntiles * kRadix);
        radix_scatter_kernel<<<tiles_per_die(ntiles) * kTileDies, kSortBlock, 0, ctx->stream>>>(
            kin, vin, n, shift, offs.as<uint64_t>(), ntiles, kout, vout);
        QE_HIP(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        QE_HIP(hipMemcpyAsync(keys, kin, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        QE_HIP(hipMemcpyAsync(vals, vin, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    QE_HIP(hipStreamSynchronize(ctx->stream));
}

// ---- general stable sort permutation (Sort's radix path) ------------------------------
__global__ void iota_kernel(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = static_cast<uint32_t>(i);
}

// Order-preserving unsigned image of key column c at row perm[i] (sign bit
// flipped; complemented for DESC), minus `base` so fewer digits suffice. NULL
// rows get 0 here and are placed by a separate flag pass.
__global__ void encode_key_kernel(ColView c, const uint32_t* perm, uint64_t n, int desc, uint64_t base,
                                  uint64_t* out) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t r = perm[i];
        uint64_t u = 0;
        if (is_valid(c, r)) {
            if (c.type == QE_F64) {
                // NaN-greatest total order with all NaNs tied and -0 == +0
                // (ops/sort_internal.h:100-104).
                double d = f64_value(c, r);
                if (d == 0.0) d = 0.0;
                u = f64_order_key(d);
            } else {
                u = static_cast<uint64_t>(int_value(c, r)) ^ 0x8000000000000000ull;
            }
            if (desc) u = ~u;
            u -= base;
        }
        out[i] = u;
    }
}
__global__ void null_flag_kernel(ColView c, const uint32_t* perm, uint64_t n, int nulls_first, uint64_t* out) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const bool isnull = !is_valid(c, perm[i]);
        out[i] = (isnull == static_cast<bool>(nulls_first)) ? 0u : 1u;
    }
}

__global__ void check_index_kernel(const uint32_t* idx, uint64_t m, uint64_t n, unsigned int* bad) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
        if (idx[i] >= n) atomicOr(bad, 1u);
}

}  // namespace qe_gpu

// Stable sort permutation, keys[0] most significant, per-key ASC/DESC and NULLS
// FIRST/LAST (null placement is absolute, ops/sort.h:35-40). Integer-family
// keys take the radix encoding directly; F64 keys use the NaN-greatest total
// order of the comparison path (ops/sort.cpp:66-100); STR keys sort by value
// through their dictionary's rank table (ops/sort.cpp:82-86). Replaces Sort::build_perm_radix
// (ops/sort.cpp:125-188): LSD over keys from the last to the first, each key
// a value pass then a NULL-placement pass, all stable.
extern "C" qe_gpu_status qe_gpu_sort_perm(qe_gpu_ctx* ctx, const qe_dev_col* const* keys_in, uint32_t nkeys,
                                          const int32_t* desc, const int32_t* nulls_first, uint32_t* perm_out) {
    QE_ENTRY(ctx, {
        invalid(nkeys >= 1 && keys_in != nullptr && perm_out != nullptr, "sort needs >=1 key");
        const uint64_t n = common_len(keys_in, nkeys);
        OwnedCols own(ctx);
        std::vector<const qe_dev_col*> keys(keys_in, keys_in + nkeys);
        for (uint32_t j = 0; j < nkeys; ++j)
            if (keys[j]->type == QE_STR) keys[j] = own.keep(remap_col(ctx, keys[j], need_dict(keys[j])->rank, QE_I32, nullptr));
        check(n < kNoRow, QE_GPU_UNSUPPORTED, "row indices are 32-bit");
        if (n == 0) return QE_GPU_OK;
        const unsigned g = grid_for(n, 256, ctx->num_cus * 16);
        iota_kernel<<<g, 256, 0, ctx->stream>>>(perm_out, n);
        Scratch k64(ctx, n * 8);
        for (uint32_t jj = nkeys; jj-- > 0;) {
            const qe_dev_col* c = keys[jj];
            const bool fkey = c->type == QE_F64;
            const MinMax mm = fkey ? MinMax{} : column_minmax(ctx, c, n);
            const int d = desc ? desc[jj] : 0;
            uint64_t lo = 0, hi = fkey ? ~uint64_t(0) : 0;
            if (mm.valid) {
                uint64_t a = static_cast<uint64_t>(mm.mn) ^ 0x8000000000000000ull;
                uint64_t b = static_cast<uint64_t>(mm.mx) ^ 0x8000000000000000ull;
                if (d) {
                    a = ~a;
                    b = ~b;
                }
                lo = std::min(a, b);
                hi = std::max(a, b);
            }
            unsigned bits = 0;
            while (bits < 64 && ((hi - lo) >> bits) != 0) ++bits;
            encode_key_kernel<<<g, 256, 0, ctx->stream>>>(view(c), perm_out, n, d, lo, k64.as<uint64_t>());
            QE_HIP(hipGetLastError());
            radix_sort_pairs(ctx, k64.as<uint64_t>(), perm_out, n, 0, (bits + 7) / 8 * 8);
            if (!c->all_valid && (fkey || mm.valid < n)) {
                null_flag_kernel<<<g, 256, 0, ctx->stream>>>(view(c), perm_out, n, nulls_first ? nulls_first[jj] : 0,
                                                             k64.as<uint64_t>());
                QE_HIP(hipGetLastError());
                radix_sort_pairs(ctx, k64.as<uint64_t>(), perm_out, n, 0, 8);
            }
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" qe_gpu_status qe_gpu_take(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                     const uint32_t* idx, uint64_t m, qe_dev_table** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && cols != nullptr && ncols >= 1, "take needs >=1 column");
        invalid(idx != nullptr || m == 0, "null index array");
        const uint64_t n = common_len(cols, ncols);
        if (m > 0) {  // every index must address a row
            Scratch bad(ctx, 4);
            memset_async(ctx, bad.p, 0, 4);
            check_index_kernel<<<grid_for(m, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(idx, m, n,
                                                                                             bad.as<unsigned int>());
            QE_HIP(hipGetLastError());
            invalid(read_scalar(ctx, bad.as<unsigned int>()) == 0, "take index out of range");
        }
        qe_dev_table* t = new_table(m);
        try {
            std::vector<const uint32_t*> idxs(ncols, idx);
            t->cols = gather_cols(ctx, cols, idxs.data(), ncols, m);
        } catch (...) {
            qe_gpu_table_free(ctx, t);
            throw;
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
        *out = t;
    });
}
