// Dictionary-encoded VARCHAR (Type::STR). A STR column holds int32 codes into a
// string dictionary (core/string_dict.h:21-23 layout: strings concatenated plus
// n + 1 offsets). Values compare by their bytes, never by code
// (core/string_dict.h:9-14), so every entry point that groups, joins or orders
// on STR first maps codes to something whose equality or order is the value's:
//  * group keys: codes are already value ids within one dictionary whose values
//    are unique; a dictionary holding a value under several codes is first
//    mapped to the value's first code (ops/aggregate.cpp:164-187);
//  * join keys: the build side's codes (canonical as above) are the ids; each
//    probe code maps to the build dictionary's code of the same value, or to a
//    fresh id that no build row has (ops/join.cpp:48-71 interns both sides into
//    one id space; only equality matters);
//  * MIN/MAX and ordering: codes map to the rank of their value in byte order
//    (ops/aggregate.cpp:189-205, std::string_view ordering), and ranks map back
//    to codes on output.
// The tables are built on the host from the dictionary (O(dictionary)) and
// applied on the device by one gather kernel.
#include <algorithm>
#include <numeric>

#include "internal.hpp"

using namespace qe_gpu;

namespace qe_gpu {

__global__ void remap_kernel(const int32_t* __restrict__ codes, const uint64_t* __restrict__ valid, uint64_t n,
                             const int32_t* __restrict__ table, uint32_t tlen, int32_t* __restrict__ out,
                             unsigned int* bad) {
    for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < n; r += uint64_t(gridDim.x) * blockDim.x) {
        const bool v = valid == nullptr || ((valid[r >> 6] >> (r & 63)) & 1ull);
        int32_t o = 0;
        if (v) {
            const int32_t c = codes[r];
            if (c >= 0 && static_cast<uint32_t>(c) < tlen) o = table[c];
            else atomicOr(bad, 1u);
        }
        out[r] = o;
    }
}

qe_dev_col* remap_col(qe_gpu_ctx* ctx, const qe_dev_col* col, const std::vector<int32_t>& table, int out_type,
                      const qe_gpu_dict* out_dict) {
    invalid(col->type == QE_STR || col->type == QE_I32, "remap needs a 4-byte code column");
    const uint64_t n = col->len;
    const bool nullable = !col->all_valid && col->validity != nullptr;
    qe_dev_col* out = new_col(ctx, out_type, n, nullable);
    try {
        out->dict = out_dict;
        if (nullable) QE_HIP(hipMemcpyAsync(out->validity, col->validity, words_for(n) * 8, hipMemcpyDeviceToDevice,
                                            ctx->stream));
        if (n > 0) {
            Scratch tab(ctx, std::max<size_t>(table.size(), 1) * 4), bad(ctx, 4);
            if (!table.empty())
                QE_HIP(hipMemcpyAsync(tab.p, table.data(), table.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            memset_async(ctx, bad.p, 0, 4);
            remap_kernel<<<grid_for(n, 256, ctx->num_cus * 16), 256, 0, ctx->stream>>>(
                static_cast<const int32_t*>(col->data), nullable ? col->validity : nullptr, n, tab.as<int32_t>(),
                static_cast<uint32_t>(table.size()), static_cast<int32_t*>(out->data), bad.as<unsigned int>());
            QE_HIP(hipGetLastError());
            invalid(read_scalar(ctx, bad.as<unsigned int>()) == 0, "STR code outside its dictionary");
        }
    } catch (...) {
        free_col(ctx, out);
        throw;
    }
    return out;
}

const qe_gpu_dict* need_dict(const qe_dev_col* c) {
    invalid(c->dict != nullptr, "STR column has no dictionary (qe_gpu_column_set_dict)");
    return c->dict;
}

void build_ranks(qe_gpu_dict* d) {
    std::vector<int32_t> firsts;  // one code per distinct value
    for (uint32_t c = 0; c < d->n; ++c)
        if (d->canon[c] == static_cast<int32_t>(c)) firsts.push_back(static_cast<int32_t>(c));
    std::sort(firsts.begin(), firsts.end(), [d](int32_t a, int32_t b) { return d->at(a) < d->at(b); });
    std::vector<int32_t> rank_of_first(d->n, 0);
    for (size_t r = 0; r < firsts.size(); ++r) rank_of_first[firsts[r]] = static_cast<int32_t>(r);
    d->rank.resize(d->n);
    for (uint32_t c = 0; c < d->n; ++c) d->rank[c] = rank_of_first[d->canon[c]];
    d->by_rank = std::move(firsts);
}

std::vector<int32_t> probe_ids(const qe_gpu_dict* build, const qe_gpu_dict* probe) {
    invalid(uint64_t(build->n) + probe->n < uint64_t(INT32_MAX), "dictionaries too large for int32 join ids");
    std::vector<int32_t> t(probe->n);
    for (uint32_t c = 0; c < probe->n; ++c) {
        auto it = build->index.find(probe->at(c));
        t[c] = it != build->index.end() ? it->second : static_cast<int32_t>(build->n + c);
    }
    return t;
}

OwnedCols::~OwnedCols() {
    for (auto* c : cols) free_col(ctx, c);
}

// The same codes read as I32 ids (a handle over the column's memory).
static qe_dev_col* i32_alias(const qe_dev_col* c) {
    auto* a = new qe_dev_col(*c);
    a->type = QE_I32;
    a->dict = nullptr;
    a->owns = false;
    return a;
}

void str_join_keys(qe_gpu_ctx* ctx, const qe_dev_col* const* pk, const qe_dev_col* const* bk, uint32_t nkeys,
                   std::vector<const qe_dev_col*>& pk_out, std::vector<const qe_dev_col*>& bk_out, OwnedCols& own) {
    pk_out.assign(pk, pk + nkeys);
    bk_out.assign(bk, bk + nkeys);
    for (uint32_t j = 0; j < nkeys; ++j) {
        if (pk[j] == nullptr || bk[j] == nullptr || pk[j]->type != QE_STR || bk[j]->type != QE_STR) continue;
        const qe_gpu_dict* bd = need_dict(bk[j]);
        const qe_gpu_dict* pd = need_dict(pk[j]);
        if (pd == bd && bd->unique) continue;  // codes are value ids on both sides
        // Both sides become I32 value ids: canonical build codes, probe codes
        // mapped through the build dictionary.
        bk_out[j] = own.keep(bd->unique ? i32_alias(bk[j]) : remap_col(ctx, bk[j], bd->canon, QE_I32, nullptr));
        pk_out[j] = own.keep(remap_col(ctx, pk[j], pd == bd ? bd->canon : probe_ids(bd, pd), QE_I32, nullptr));
    }
}

bool str_agg_prepare(qe_gpu_ctx* ctx, std::vector<const qe_dev_col*>& cols, const uint32_t* keys, uint32_t nkeys,
                     std::vector<qe_agg_spec>& aggs, StrAggFix& fix, OwnedCols& own) {
    bool any = false;
    fix.key_dict.assign(nkeys, nullptr);
    fix.rank_dict.assign(aggs.size(), nullptr);
    for (uint32_t k = 0; k < nkeys; ++k) {
        invalid(keys[k] < cols.size(), "group key column out of range");
        const qe_dev_col* c = cols[keys[k]];
        if (c == nullptr || c->type != QE_STR) continue;
        const qe_gpu_dict* d = need_dict(c);
        fix.key_dict[k] = d;
        any = true;
        if (!d->unique) cols[keys[k]] = own.keep(remap_col(ctx, c, d->canon, QE_STR, d));
    }
    for (size_t a = 0; a < aggs.size(); ++a) {
        if (aggs[a].func != QE_AGG_MIN && aggs[a].func != QE_AGG_MAX) continue;
        invalid(aggs[a].input_col < cols.size(), "aggregate input column out of range");
        const qe_dev_col* c = cols[aggs[a].input_col];
        if (c == nullptr || c->type != QE_STR) continue;
        const qe_gpu_dict* d = need_dict(c);
        fix.rank_dict[a] = d;
        any = true;
        cols.push_back(own.keep(remap_col(ctx, c, d->rank, QE_I32, nullptr)));
        aggs[a].input_col = static_cast<uint32_t>(cols.size() - 1);
    }
    return any;
}

void str_agg_finish(qe_gpu_ctx* ctx, qe_dev_table* t, const StrAggFix& fix) {
    for (size_t k = 0; k < fix.key_dict.size(); ++k) {
        if (!fix.key_dict[k]) continue;
        t->cols[k]->type = QE_STR;
        t->cols[k]->dict = fix.key_dict[k];
    }
    const size_t nk = fix.key_dict.size();
    for (size_t a = 0; a < fix.rank_dict.size(); ++a) {
        const qe_gpu_dict* d = fix.rank_dict[a];
        if (!d) continue;
        qe_dev_col* ranks = t->cols[nk + a];
        t->cols[nk + a] = remap_col(ctx, ranks, d->by_rank, QE_STR, d);
        free_col(ctx, ranks);
    }
}

}  // namespace qe_gpu

extern "C" {

qe_gpu_status qe_gpu_dict_create(qe_gpu_ctx* ctx, const char* bytes, const uint32_t* offsets, uint32_t n,
                                 qe_gpu_dict** out) {
    QE_ENTRY(ctx, {
        invalid(out != nullptr && offsets != nullptr, "null argument");
        invalid(offsets[0] == 0, "dictionary offsets must start at 0");
        for (uint32_t i = 0; i < n; ++i) invalid(offsets[i] <= offsets[i + 1], "dictionary offsets must not decrease");
        invalid(n < uint32_t(INT32_MAX), "dictionary too large for int32 codes");
        invalid(bytes != nullptr || offsets[n] == 0, "null dictionary bytes");
        auto* d = new qe_gpu_dict();
        try {
            d->n = n;
            d->bytes.assign(bytes, bytes + offsets[n]);
            d->off.assign(offsets, offsets + n + 1);
            d->canon.resize(n);
            d->index.reserve(n);
            for (uint32_t c = 0; c < n; ++c) {
                auto ins = d->index.emplace(d->at(c), static_cast<int32_t>(c));
                d->canon[c] = ins.first->second;
                if (!ins.second) d->unique = false;
            }
            build_ranks(d);
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}

void qe_gpu_dict_free(qe_gpu_ctx* ctx, qe_gpu_dict* d) {
    (void)ctx;
    delete d;
}

uint32_t qe_gpu_dict_size(const qe_gpu_dict* d) { return d ? d->n : 0; }

qe_gpu_status qe_gpu_column_set_dict(qe_gpu_ctx* ctx, qe_dev_col* col, const qe_gpu_dict* d) {
    QE_ENTRY(ctx, {
        invalid(col != nullptr, "null column");
        invalid(col->type == QE_STR, "only STR columns carry a dictionary");
        col->dict = d;
    });
}

const qe_gpu_dict* qe_gpu_column_dict(const qe_dev_col* col) { return col ? col->dict : nullptr; }

qe_gpu_status qe_gpu_column_remap(qe_gpu_ctx* ctx, const qe_dev_col* col, const int32_t* table, uint32_t table_len,
                                  int32_t out_type, qe_dev_col** out) {
    QE_ENTRY(ctx, {
        invalid(col != nullptr && out != nullptr, "null argument");
        invalid(table != nullptr || table_len == 0, "null table");
        invalid(out_type == QE_I32 || out_type == QE_STR, "remap produces I32 or STR");
        std::vector<int32_t> t(table, table + table_len);
        *out = remap_col(ctx, col, t, out_type, nullptr);
        QE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
