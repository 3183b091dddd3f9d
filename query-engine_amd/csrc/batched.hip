// Batched group-by: many independent GROUP BY queries, each over its own
// morsels of uneven size, in one persistent launch (the engine's morsel
// driver, exec/parallel.cpp:452-480, fans morsels of one query over CPU
// threads; here every query's morsels share the GPU at once).
//
// Work = tiles of `tile_rows` rows cut from every morsel of every query, in a
// device queue. A persistent grid (blocks_per_cu blocks per CU) pulls tiles
// with one atomic per tile; every block exits when the queue is drained. Each
// row probes its query's device hash table (the shared open-addressing table,
// NULL keys group together) and folds into that slot's state; a per-block LDS
// cache of `lds_slots` slot entries absorbs repeated groups (small-cardinality
// queries) and is flushed with one HBM atomic per cached field at tile end.
//
// Modes: persistent (one launch), per-batch (one launch per 2048-row batch,
// the launch granularity of the pull engine's batches, as the baseline), and
// graph (the persistent pipeline captured once in a hipGraph and replayed).
// The autotuner times tile size x blocks per CU x LDS cache size on the actual
// batch and keeps the fastest.
#include "agg_common.hpp"
#include "hash_common.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

using namespace qe_gpu;

namespace qe_gpu {

constexpr int kBatchMaxAcc = 6;
constexpr uint32_t kEmptyTag = 0xFFFFFFFFu;

struct QueryDev {
    HashKeys hk;
    HashTableView t;
    uint64_t* st;  // state, compact layout, S = t.cap slots
    AggProg p;
};

struct Tile {
    uint32_t q;
    uint32_t pad;
    uint64_t begin, end;
};

// LDS layout for lds_slots = C: tag u32[C] | rows u32[C] | acc u64[kBatchMaxAcc][C] | cnt u32[kBatchMaxAcc][C]
__host__ __device__ inline size_t cache_bytes(uint32_t C) {
    return size_t(C) * (4 + 4 + 8 * kBatchMaxAcc + 4 * kBatchMaxAcc);
}

__device__ inline void cache_reset(char* smem, uint32_t C, const AggProg& p) {
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
        tag[i] = kEmptyTag;
        rows[i] = 0;
        for (int a = 0; a < p.nacc; ++a) acc[size_t(a) * C + i] = acc_identity(p.acc_kind[a]);
        for (int c = 0; c < p.ncnt; ++c) cnt[size_t(c) * C + i] = 0;
    }
}

// Fold one row into the LDS cache entry for global slot s if the entry is (or
// becomes) s's; returns false when the entry holds another slot.
__device__ inline bool cache_update(char* smem, uint32_t C, const AggProg& p, uint32_t s, uint64_t r) {
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    const uint32_t e = s & (C - 1);
    const uint32_t old = atomicCAS(&tag[e], kEmptyTag, s);
    if (old != kEmptyTag && old != s) return false;
    atomicAdd(&rows[e], 1u);
    for (int a = 0; a < p.n; ++a) {
        const int k = p.kind[a];
        if (k == K_COUNT_STAR || !is_valid(p.in[a], r)) continue;
        if (p.cnt_of[a] >= 0) atomicAdd(&cnt[size_t(p.cnt_of[a]) * C + e], 1u);
        if (p.acc_of[a] < 0) continue;
        uint64_t* ap = &acc[size_t(p.acc_of[a]) * C + e];
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
    return true;
}

__device__ inline void combine_atomic(int kind, uint64_t* dst, uint64_t v) {
    switch (kind) {
        case K_SUM_I:
            atomicAdd(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        case K_SUM_F: {
            double d;
            memcpy(&d, &v, 8);
            unsafeAtomicAdd(reinterpret_cast<double*>(dst), d);
            break;
        }
        case K_MIN_I:
            atomicMin(reinterpret_cast<long long*>(dst), static_cast<long long>(v));
            break;
        case K_MAX_I:
            atomicMax(reinterpret_cast<long long*>(dst), static_cast<long long>(v));
            break;
        case K_MIN_F:
            atomicMin(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        case K_MAX_F:
            atomicMax(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(v));
            break;
        default:
            break;
    }
}

__device__ inline void cache_flush(char* smem, uint32_t C, const QueryDev& q) {
    const AggProg& p = q.p;
    const uint64_t S = q.t.cap;
    uint32_t* tag = reinterpret_cast<uint32_t*>(smem);
    uint32_t* rows = tag + C;
    uint64_t* acc = reinterpret_cast<uint64_t*>(smem + size_t(8) * C);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem + size_t(8) * C + size_t(8) * kBatchMaxAcc * C);
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
        const uint32_t s = tag[i];
        if (s == kEmptyTag) continue;
        atomicAdd(reinterpret_cast<unsigned long long*>(&q.st[s]), static_cast<unsigned long long>(rows[i]));
        for (int a = 0; a < p.nacc; ++a) combine_atomic(p.acc_kind[a], &q.st[(1 + size_t(a)) * S + s], acc[size_t(a) * C + i]);
        for (int c = 0; c < p.ncnt; ++c)
            atomicAdd(reinterpret_cast<unsigned long long*>(&q.st[(1 + size_t(p.nacc) + c) * S + s]),
                      static_cast<unsigned long long>(cnt[size_t(c) * C + i]));
    }
}

__device__ inline void process_tile(const QueryDev& q, const Tile& tile, char* smem, uint32_t C) {
    uint64_t vrow[kMaxAggs];
    for (uint64_t r = tile.begin + threadIdx.x; r < tile.end; r += blockDim.x) {
        uint64_t w[kMaxKeys], nm, h;
        normalize_row(q.hk, r, w, nm, h);
        const uint32_t s = ht_insert(q.t, q.hk.nkeys, w, nm, h);
        if (s == kNoSlot) continue;  // overflow reported through the table flag
        if (C && cache_update(smem, C, q.p, s, r)) continue;
        for (int a = 0; a < q.p.n; ++a) vrow[a] = r;
        global_update(q.st, q.t.cap, q.p, s, vrow);
    }
}

// Persistent: blocks pull tiles until the queue is empty (every block reaches
// the exit test each iteration).
__global__ void __launch_bounds__(256) batch_persistent_kernel(const QueryDev* qs, const Tile* tiles, uint64_t ntiles,
                                                               unsigned long long* next, uint32_t C) {
    extern __shared__ __align__(16) char smem[];
    __shared__ uint64_t cur;
    while (true) {
        if (threadIdx.x == 0) cur = atomicAdd(next, 1ull);
        __syncthreads();
        const uint64_t i = cur;
        if (i >= ntiles) break;
        const Tile tile = tiles[i];
        const QueryDev& q = qs[tile.q];
        if (C) {
            cache_reset(smem, C, q.p);
            __syncthreads();
        }
        process_tile(q, tile, smem, C);
        if (C) {
            __syncthreads();
            cache_flush(smem, C, q);
        }
        __syncthreads();
    }
}

// One launch = one batch of <= 2048 rows of one query (the per-batch baseline).
__global__ void __launch_bounds__(256) batch_one_kernel(const QueryDev* qs, uint32_t qi, uint64_t begin, uint64_t end,
                                                        uint32_t C) {
    extern __shared__ __align__(16) char smem[];
    const QueryDev& q = qs[qi];
    // Each block takes a contiguous slice of the batch.
    const uint64_t per = (end - begin + gridDim.x - 1) / gridDim.x;
    Tile t{qi, 0, begin + per * blockIdx.x, min(end, begin + per * (blockIdx.x + 1))};
    if (C) {
        cache_reset(smem, C, q.p);
        __syncthreads();
    }
    if (t.begin < t.end) process_tile(q, t, smem, C);
    if (C) {
        __syncthreads();
        cache_flush(smem, C, q);
    }
}

__global__ void batch_init_kernel(const QueryDev* qs, uint32_t nq) {
    // grid-stride over (query, slot); slot counts differ per query.
    for (uint32_t qi = 0; qi < nq; ++qi) {
        const QueryDev& q = qs[qi];
        const uint64_t S = q.t.cap;
        for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < S;
             s += uint64_t(gridDim.x) * blockDim.x) {
            q.t.state[s] = 0;
            q.st[s] = 0;
            for (int a = 0; a < q.p.nacc; ++a) q.st[(1 + size_t(a)) * S + s] = acc_identity(q.p.acc_kind[a]);
            for (int c = 0; c < q.p.ncnt; ++c) q.st[(1 + size_t(q.p.nacc) + c) * S + s] = 0;
        }
    }
}

__global__ void batch_reset_kernel(unsigned long long* next, unsigned int* ovf) {
    *next = 0;
    *ovf = 0;
}

}  // namespace qe_gpu

// ---- host side -----------------------------------------------------------------------------
struct BatchPlan {
    uint32_t nq = 0;
    std::vector<QueryDev> hq;
    std::vector<Scratch> tables;  // per query: state, hash, words, nullmask, st
    std::vector<uint64_t> caps;
    Scratch dq, tiles, next, ovf;
    uint64_t ntiles = 0;
    uint32_t tile_rows = 0;
    std::vector<std::vector<uint64_t>> morsels;  // per query offsets
    std::vector<uint64_t> rows;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_C = 0, graph_blocks = 0;
    ~BatchPlan() {
        if (graph) hipGraphExecDestroy(graph);
    }
};

static void build_batch_plan(qe_gpu_ctx* ctx, BatchPlan& bp, const qe_batch_query* qs, uint32_t nq) {
    bp.nq = nq;
    bp.hq.resize(nq);
    bp.caps.resize(nq);
    bp.morsels.resize(nq);
    bp.rows.resize(nq);
    bp.ovf = Scratch(ctx, 8);
    for (uint32_t i = 0; i < nq; ++i) {
        const qe_batch_query& q = qs[i];
        invalid(q.key != nullptr && q.value != nullptr, "batched query needs key and value columns");
        invalid(q.key->len == q.value->len, "key and value columns differ in length");
        invalid(q.morsel_offsets != nullptr && q.nmorsels >= 1, "batched query needs >=1 morsel");
        std::vector<uint64_t> offs(q.morsel_offsets, q.morsel_offsets + q.nmorsels + 1);
        invalid(offs.front() == 0 && offs.back() == q.key->len, "morsel offsets must cover the columns");
        for (size_t m = 0; m + 1 < offs.size(); ++m) invalid(offs[m] <= offs[m + 1], "morsel offsets must ascend");
        bp.morsels[i] = offs;
        bp.rows[i] = q.key->len;
        ColView ins[kMaxAggs];
        int types[kMaxAggs];
        invalid(q.naggs >= 1 && q.aggs != nullptr, "batched query needs >=1 aggregate");
        for (uint32_t a = 0; a < q.naggs && a < uint32_t(kMaxAggs); ++a) {
            invalid(q.aggs[a].func == QE_AGG_COUNT_STAR || q.aggs[a].input_col <= 1,
                    "batched aggregate input must be 0 (key) or 1 (value)");
            const qe_dev_col* c = q.aggs[a].input_col == 0 ? q.key : q.value;
            ins[a] = view(c);
            types[a] = c->type;
        }
        QueryDev d{};
        d.p = make_prog(q.aggs, q.naggs, ins, types);
        check(d.p.nacc <= kBatchMaxAcc && d.p.ncnt <= kBatchMaxAcc, QE_GPU_UNSUPPORTED,
              "too many aggregates for the batched path");
        d.hk.nkeys = 1;
        d.hk.key[0] = view(q.key);
        // Capacity from the distinct-group bound the caller supplies (or the
        // row count), load <= 0.5.
        const uint64_t bound = q.group_bound ? std::min<uint64_t>(q.group_bound, q.key->len) : q.key->len;
        uint64_t cap = 1024;
        while (cap < 2 * bound + 2) cap <<= 1;
        check(cap <= (uint64_t(1) << 31), QE_GPU_UNSUPPORTED, "batched query too large");
        bp.caps[i] = cap;
        Scratch state(ctx, cap * 4), hash(ctx, cap * 8), words(ctx, cap * 8), nmask(ctx, cap * 8),
            st(ctx, state_words(cap, d.p) * 8);
        d.t = HashTableView{state.as<uint32_t>(), hash.as<uint64_t>(), words.as<uint64_t>(), nmask.as<uint64_t>(),
                            cap, bp.ovf.as<unsigned int>()};
        d.st = st.as<uint64_t>();
        bp.tables.push_back(std::move(state));
        bp.tables.push_back(std::move(hash));
        bp.tables.push_back(std::move(words));
        bp.tables.push_back(std::move(nmask));
        bp.tables.push_back(std::move(st));
        bp.hq[i] = d;
    }
    bp.dq = Scratch(ctx, nq * sizeof(QueryDev));
    QE_HIP(hipMemcpyAsync(bp.dq.p, bp.hq.data(), nq * sizeof(QueryDev), hipMemcpyHostToDevice, ctx->stream));
    bp.next = Scratch(ctx, 8);
    QE_HIP(hipStreamSynchronize(ctx->stream));
}

static void make_tiles(qe_gpu_ctx* ctx, BatchPlan& bp, uint32_t tile_rows) {
    if (bp.tile_rows == tile_rows && bp.ntiles) return;
    std::vector<Tile> tiles;
    // Interleave queries so concurrent blocks work on different tables.
    std::vector<std::vector<Tile>> per(bp.nq);
    for (uint32_t q = 0; q < bp.nq; ++q)
        for (size_t m = 0; m + 1 < bp.morsels[q].size(); ++m)
            for (uint64_t b = bp.morsels[q][m]; b < bp.morsels[q][m + 1]; b += tile_rows)
                per[q].push_back(Tile{q, 0, b, std::min<uint64_t>(bp.morsels[q][m + 1], b + tile_rows)});
    for (size_t k = 0;; ++k) {
        bool any = false;
        for (uint32_t q = 0; q < bp.nq; ++q)
            if (k < per[q].size()) {
                tiles.push_back(per[q][k]);
                any = true;
            }
        if (!any) break;
    }
    bp.ntiles = tiles.size();
    bp.tile_rows = tile_rows;
    bp.tiles = Scratch(ctx, std::max<uint64_t>(bp.ntiles, 1) * sizeof(Tile));
    if (bp.ntiles)
        QE_HIP(hipMemcpyAsync(bp.tiles.p, tiles.data(), bp.ntiles * sizeof(Tile), hipMemcpyHostToDevice, ctx->stream));
    QE_HIP(hipStreamSynchronize(ctx->stream));
}

static void enqueue_persistent(qe_gpu_ctx* ctx, BatchPlan& bp, uint32_t blocks, uint32_t C) {
    batch_reset_kernel<<<1, 1, 0, ctx->stream>>>(bp.next.as<unsigned long long>(), bp.ovf.as<unsigned int>());
    batch_init_kernel<<<ctx->num_cus * 4, 256, 0, ctx->stream>>>(bp.dq.as<QueryDev>(), bp.nq);
    batch_persistent_kernel<<<blocks, 256, cache_bytes(C), ctx->stream>>>(
        bp.dq.as<QueryDev>(), bp.tiles.as<Tile>(), bp.ntiles, bp.next.as<unsigned long long>(), C);
}

static void run_batch(qe_gpu_ctx* ctx, BatchPlan& bp, int mode, const qe_batch_tuning& tn) {
    const uint32_t C = tn.lds_slots;
    check(C == 0 || (C & (C - 1)) == 0, QE_GPU_INVALID_ARGUMENT, "lds_slots must be 0 or a power of two");
    check(cache_bytes(C) <= 64 * 1024, QE_GPU_INVALID_ARGUMENT, "lds_slots too large");
    const uint32_t blocks = std::max<uint32_t>(1, tn.blocks_per_cu) * ctx->num_cus;
    if (mode == QE_BATCH_PER_BATCH) {
        batch_reset_kernel<<<1, 1, 0, ctx->stream>>>(bp.next.as<unsigned long long>(), bp.ovf.as<unsigned int>());
        batch_init_kernel<<<ctx->num_cus * 4, 256, 0, ctx->stream>>>(bp.dq.as<QueryDev>(), bp.nq);
        TimedRegion tr(ctx, "batch_per_batch");
        for (uint32_t q = 0; q < bp.nq; ++q)
            for (size_t m = 0; m + 1 < bp.morsels[q].size(); ++m)
                for (uint64_t b = bp.morsels[q][m]; b < bp.morsels[q][m + 1]; b += 2048) {
                    const uint64_t e = std::min<uint64_t>(bp.morsels[q][m + 1], b + 2048);
                    batch_one_kernel<<<8, 256, cache_bytes(C), ctx->stream>>>(bp.dq.as<QueryDev>(), q, b, e, C);
                }
        QE_HIP(hipGetLastError());
        return;
    }
    make_tiles(ctx, bp, std::max<uint32_t>(256, tn.tile_rows));
    if (mode == QE_BATCH_GRAPH) {
        if (!bp.graph || bp.graph_C != C || bp.graph_blocks != blocks) {
            if (bp.graph) QE_HIP(hipGraphExecDestroy(bp.graph));
            bp.graph = nullptr;
            hipStream_t cs;
            QE_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t g;
            hipStream_t saved = ctx->stream;
            ctx->stream = cs;
            QE_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            enqueue_persistent(ctx, bp, blocks, C);
            QE_HIP(hipStreamEndCapture(cs, &g));
            ctx->stream = saved;
            QE_HIP(hipGraphInstantiate(&bp.graph, g, nullptr, nullptr, 0));
            QE_HIP(hipGraphDestroy(g));
            QE_HIP(hipStreamDestroy(cs));
            bp.graph_C = C;
            bp.graph_blocks = blocks;
        }
        TimedRegion tr(ctx, "batch_graph");
        QE_HIP(hipGraphLaunch(bp.graph, ctx->stream));
        return;
    }
    TimedRegion tr(ctx, "batch_persistent");
    enqueue_persistent(ctx, bp, blocks, C);
    QE_HIP(hipGetLastError());
}

static void emit_batch(qe_gpu_ctx* ctx, BatchPlan& bp, const qe_batch_query* qs, qe_dev_table** outs) {
    check(read_scalar(ctx, bp.ovf.as<unsigned int>()) == 0, QE_GPU_OVERFLOW,
          "batched group table overflow (group_bound too small?)");
    for (uint32_t i = 0; i < bp.nq; ++i) {
        KeyEmit ke{};
        ke.mode = 1;
        ke.nkeys = 1;
        ke.types[0] = qs[i].key->type;
        ke.words = bp.hq[i].t.words;
        ke.nullmask = bp.hq[i].t.nullmask;
        outs[i] = emit_groups(ctx, bp.hq[i].st, bp.caps[i], bp.hq[i].p, ke, false);
    }
}

struct qe_batch_plan {
    BatchPlan bp;
};

extern "C" {

qe_gpu_status qe_gpu_batch_prepare(qe_gpu_ctx* ctx, const qe_batch_query* queries, uint32_t nq, qe_batch_plan** out) {
    QE_ENTRY(ctx, {
        invalid(queries != nullptr && out != nullptr && nq >= 1, "batched group-by needs >=1 query");
        auto* p = new qe_batch_plan();
        try {
            build_batch_plan(ctx, p->bp, queries, nq);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

qe_gpu_status qe_gpu_batch_run(qe_gpu_ctx* ctx, qe_batch_plan* plan, const qe_batch_query* queries, int32_t mode,
                               const qe_batch_tuning* tuning, qe_dev_table** outs) {
    QE_ENTRY(ctx, {
        invalid(plan != nullptr && queries != nullptr, "null argument");
        invalid(mode >= QE_BATCH_PERSISTENT && mode <= QE_BATCH_GRAPH, "unknown batch mode");
        qe_batch_tuning tn = tuning ? *tuning : qe_batch_tuning{8192, 4, 256};
        run_batch(ctx, plan->bp, mode, tn);
        if (outs) emit_batch(ctx, plan->bp, queries, outs);
        else QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qe_gpu_status qe_gpu_batch_autotune(qe_gpu_ctx* ctx, qe_batch_plan* plan, qe_batch_tuning* out) {
    QE_ENTRY(ctx, {
        invalid(plan != nullptr && out != nullptr, "null argument");
        double best = 1e30;
        qe_batch_tuning bt{8192, 4, 256};
        for (uint32_t tile : {4096u, 16384u, 65536u})
            for (uint32_t bpc : {2u, 4u, 8u})
                for (uint32_t C : {0u, 256u, 512u}) {
                    qe_batch_tuning tn{tile, bpc, C};
                    run_batch(ctx, plan->bp, QE_BATCH_PERSISTENT, tn);  // warm
                    QE_HIP(hipStreamSynchronize(ctx->stream));
                    const auto t0 = std::chrono::steady_clock::now();
                    run_batch(ctx, plan->bp, QE_BATCH_PERSISTENT, tn);
                    QE_HIP(hipStreamSynchronize(ctx->stream));
                    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (s < best) {
                        best = s;
                        bt = tn;
                    }
                }
        *out = bt;
    });
}

void qe_gpu_batch_free(qe_gpu_ctx* ctx, qe_batch_plan* plan) {
    if (!ctx || !plan) return;
    hipStreamSynchronize(ctx->stream);
    delete plan;
}

__global__ void gen_uniform_kernel(uint64_t seed, uint64_t stream, uint64_t n, int type, double lo, double hi,
                                   void* out) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n;
         r += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t h = mix64(mix64(seed ^ (stream * 0x9E3779B97F4A7C15ull)) ^ (r * 0xD1B54A32D192ED03ull));
        const double u = static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
        const double v = lo + u * (hi - lo);
        switch (type) {
            case QE_I32:
                static_cast<int32_t*>(out)[r] = static_cast<int32_t>(floor(v));
                break;
            case QE_F64:
                static_cast<double*>(out)[r] = v;
                break;
            default:
                static_cast<int64_t*>(out)[r] = static_cast<int64_t>(floor(v));
        }
    }
}
__global__ void gen_nulls_kernel(uint64_t seed, uint64_t stream, uint64_t n, double p, uint64_t* words) {
    const uint64_t nw = words_for(n);
    for (uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; w < nw;
         w += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t bits = 0;
        for (unsigned b = 0; b < 64; ++b) {
            const uint64_t r = w * 64 + b;
            const uint64_t h = mix64(mix64(seed ^ ((stream + 1) * 0x9E3779B97F4A7C15ull)) ^ (r * 0xD1B54A32D192ED03ull));
            const double u = static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
            bits |= uint64_t(r >= n || u >= p) << b;
        }
        words[w] = bits;
    }
}

qe_gpu_status qe_gpu_gen_uniform(qe_gpu_ctx* ctx, uint64_t seed, uint64_t stream, qe_dev_col* col, double lo,
                                 double hi, double null_frac) {
    QE_ENTRY(ctx, {
        invalid(col != nullptr, "null column");
        invalid(col->type == QE_I32 || col->type == QE_I64 || col->type == QE_TS || col->type == QE_F64,
                "uniform generator supports I32/I64/TS/F64");
        const uint64_t n = col->len;
        if (n) {
            gen_uniform_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(seed, stream, n, col->type,
                                                                                             lo, hi, col->data);
            if (null_frac > 0) {
                invalid(col->validity != nullptr, "column must be nullable for NULLs");
                gen_nulls_kernel<<<grid_for(words_for(n), 256, 4096), 256, 0, ctx->stream>>>(seed, stream, n, null_frac,
                                                                                            col->validity);
                col->all_valid = false;
            }
            QE_HIP(hipGetLastError());
        }
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
