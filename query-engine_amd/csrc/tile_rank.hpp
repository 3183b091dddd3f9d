// Stable ranking of a tile's rows by an 8-bit digit: the step shared by the
// radix sort's scatter passes and the as-of probe clustering.
#pragma once

#include "internal.hpp"

namespace qe_gpu {

// A tile is ITEMS rows per thread of a 256-thread block, laid out wave-major:
// wave w owns the tile's rows [w * ITEMS * 64, (w + 1) * ITEMS * 64), and its
// item i is the 64 consecutive rows from w * ITEMS * 64 + i * 64 (one
// coalesced load per item). Position order in the tile is then (wave, item,
// lane) order.
constexpr unsigned kRankBlock = 256;
constexpr unsigned kRankWaves = kRankBlock / kWave;
constexpr unsigned kRankDigits = 256;

template <unsigned ITEMS>
__device__ inline uint32_t tile_row(unsigned i) {
    return (threadIdx.x / kWave) * (ITEMS * kWave) + i * kWave + lane_id();
}

// Tile of block b when blocks are dealt round-robin to kTileDies dies (MI355X
// places consecutive blocks on consecutive XCDs): die x takes the contiguous
// tiles [x * per, (x + 1) * per) in order, so tiles written next to each other
// (a digit's runs from neighbouring tiles are adjacent in the output) go
// through the same L2 at about the same time and complete each other's partial
// lines there. The grid has kTileDies * per blocks; a tile >= ntiles is none.
constexpr uint64_t kTileDies = 8;
__host__ __device__ inline uint64_t tiles_per_die(uint64_t ntiles) { return (ntiles + kTileDies - 1) / kTileDies; }
__device__ inline uint64_t die_tile(uint64_t ntiles) {
    return (blockIdx.x % kTileDies) * tiles_per_die(ntiles) + blockIdx.x / kTileDies;
}

struct TileRankLds {
    uint32_t wrun[kRankWaves][kRankDigits];  // per-wave digit counts, then per-wave digit starts
    uint32_t start[kRankDigits];             // digit start within the tile (staging order)
    uint32_t wsum[kRankWaves];
};

// Each live row (live(i) for item i) gets pos[i], its staging position: rows
// ordered by digit (digit(i), 0..255), and within a digit in tile position
// order, so a scatter by digit is stable. L.start[d] receives the staging
// offset of digit d; pos[i] - L.start[digit(i)] is the row's rank within its
// digit. Each wave ranks its own rows without block barriers: lanes holding
// the same digit are found with eight ballots, and the wave's running digit
// counts live in its own LDS row (LDS operations of one wave complete in
// order). One block scan then turns the per-wave counts into starts. Called by
// all 256 threads of the block; ends with the block synchronised.
template <unsigned ITEMS, class DigitF, class LiveF>
__device__ inline void tile_digit_ranks(TileRankLds& L, const DigitF& digit, const LiveF& live,
                                        uint32_t (&pos)[ITEMS]) {
    const unsigned lane = lane_id();
    const unsigned w = threadIdx.x / kWave;
    uint32_t* run = L.wrun[w];
    for (unsigned j = lane; j < kRankDigits; j += kWave) run[j] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (unsigned i = 0; i < ITEMS; ++i) {
        const bool li = live(i);
        const uint32_t di = digit(i);
        uint64_t peers = __ballot(li);
        for (unsigned b = 0; b < 8; ++b) {
            const uint64_t m = __ballot(li && ((di >> b) & 1u));
            peers &= ((di >> b) & 1u) ? m : ~m;
        }
        const unsigned in_wave = __popcll(peers & lt);
        const uint32_t before = li ? run[di] : 0u;
        pos[i] = before + in_wave;  // rank within the wave's rows of this digit, for now
        __builtin_amdgcn_wave_barrier();
        if (li && in_wave == 0) run[di] = before + static_cast<uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Thread d: the tile's count of digit d, its exclusive scan over digits,
    // and each wave's start within the digit.
    const unsigned d = threadIdx.x;
    uint32_t c = 0;
    for (unsigned ww = 0; ww < kRankWaves; ++ww) c += L.wrun[ww][d];
    uint32_t incl = c;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= static_cast<unsigned>(off)) incl += t;
    }
    if (lane == kWave - 1) L.wsum[w] = incl;
    __syncthreads();
    uint32_t wb = 0;
    for (unsigned ww = 0; ww < w; ++ww) wb += L.wsum[ww];
    const uint32_t s0 = wb + incl - c;
    L.start[d] = s0;
    uint32_t acc = s0;
    for (unsigned ww = 0; ww < kRankWaves; ++ww) {
        const uint32_t cw = L.wrun[ww][d];
        L.wrun[ww][d] = acc;
        acc += cw;
    }
    __syncthreads();
    for (unsigned i = 0; i < ITEMS; ++i)
        if (live(i)) pos[i] += run[digit(i)];
}

}  // namespace qe_gpu
