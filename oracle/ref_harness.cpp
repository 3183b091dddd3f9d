// ref_harness — a thin extern "C" shim over the COMPILED REFERENCE operators.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) against the
// reference sources where they lie under the reference checkout; the output
// library goes to oracle/_ref/ and is never committed. It lets the tests pin the
// CPU restatement (oracle/qe_oracle.cpp) against the reference's own
// HashTable (ops/hashtable.h:108-190), Aggregate (ops/aggregate.h:102-144),
// HashJoin (ops/join.h:69-123), AsofJoin (tsx/asof.h:86-150) and Filter
// (ops/filter.h) over expr::cmp / expr::logic_and (expr/expr.h:160-180), driven
// through Scan (ops/scan.h:37) exactly as the reference's tests drive them, and
// it serves as the "reference" CPU baseline.
//
// Row identity for joins is recovered by appending a hidden I64 row-id column
// to each input table and reading it back from the operator output.

#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <vector>

#include "core/owned_batch.h"
#include "core/string_dict.h"
#include "core/validity.h"
#include "expr/expr.h"
#include "ops/aggregate.h"
#include "ops/filter.h"
#include "ops/hashtable.h"
#include "ops/join.h"
#include "ops/scan.h"
#include "ops/sort.h"
#include "ops/table.h"
#include "tsx/asof.h"

extern "C" {
typedef struct {
    int32_t type;
    int32_t all_valid;
    uint64_t len;
    const void* data;
    const uint64_t* validity;
    const char* dict_bytes;  // STR: the dictionary's strings, concatenated
    const uint32_t* dict_off;  // STR: dict_n + 1 offsets
    uint64_t dict_n;
} orc_col;
}

namespace {

using namespace qe;

Type ty(int32_t t) { return static_cast<Type>(t); }

// A reference Column view over a caller buffer (core/column.h:55-67).
Column view_of(const orc_col& c) {
    Column v;
    v.type = ty(c.type);
    v.len = c.len;
    v.data = static_cast<const std::byte*>(c.data);
    v.all_valid = c.all_valid || c.validity == nullptr;
    v.validity = v.all_valid ? nullptr : c.validity;
    v.dict = nullptr;
    return v;
}

OwnedColumn owned_of(const orc_col& c, uint64_t n) {
    if (c.type == static_cast<int32_t>(Type::STR)) {
        // A fresh StringDict interning the given strings in order; the column's
        // codes index that list, so they are translated through intern().
        auto dict = std::make_shared<StringDict>();
        std::vector<int32_t> code_of(c.dict_n);
        for (uint64_t i = 0; i < c.dict_n; ++i)
            code_of[i] = dict->intern(std::string_view(c.dict_bytes + c.dict_off[i], c.dict_off[i + 1] - c.dict_off[i]));
        OwnedColumn o = OwnedColumn::make_str(n, dict);
        auto* codes = reinterpret_cast<int32_t*>(o.mutable_data());
        const auto* src = static_cast<const int32_t*>(c.data);
        for (uint64_t r = 0; r < n; ++r) {
            const bool valid = c.all_valid || c.validity == nullptr || validity::get_bit(c.validity, r);
            if (valid) codes[r] = code_of[src[r]];
            else o.set_null(r);
        }
        return o;
    }
    OwnedColumn o = OwnedColumn::make(ty(c.type), n);
    std::memcpy(o.mutable_data(), c.data, n * byte_width(ty(c.type)));
    if (!(c.all_valid || c.validity == nullptr))
        for (uint64_t r = 0; r < n; ++r)
            if (!validity::get_bit(c.validity, r)) o.set_null(r);
    return o;
}

std::unique_ptr<Table> make_table(const orc_col* cols, uint32_t ncols, uint64_t n,
                                  bool with_rowid) {
    Schema s;
    std::vector<OwnedColumn> oc;
    for (uint32_t c = 0; c < ncols; ++c) {
        s.fields.emplace_back("c" + std::to_string(c), ty(cols[c].type));
        oc.push_back(owned_of(cols[c], n));
    }
    if (with_rowid) {
        s.fields.emplace_back("rowid", Type::I64);
        OwnedColumn r = OwnedColumn::make(Type::I64, n);
        auto* d = reinterpret_cast<int64_t*>(r.mutable_data());
        for (uint64_t i = 0; i < n; ++i) d[i] = static_cast<int64_t>(i);
        oc.push_back(std::move(r));
    }
    return std::make_unique<Table>(std::move(s), std::move(oc));
}

// Drained operator output, one typed column at a time.
struct ResultRows {
    std::vector<int32_t> types;
    std::vector<std::vector<int64_t>> iv;
    std::vector<std::vector<double>> dv;
    std::vector<std::vector<uint8_t>> nulls;
    std::vector<std::vector<std::string>> sv;  // STR columns: resolved values
    uint64_t nrows = 0;
};

ResultRows* drain(Operator& op) {
    auto* rs = new ResultRows();
    const Schema sch = op.output_schema();
    for (const auto& f : sch.fields) rs->types.push_back(static_cast<int32_t>(f.second));
    const size_t nc = sch.fields.size();
    rs->iv.resize(nc);
    rs->dv.resize(nc);
    rs->nulls.resize(nc);
    rs->sv.resize(nc);
    op.open();
    while (std::optional<Batch> b = op.next()) {
        for (size_t k = 0; k < b->row_count; ++k) {
            const uint32_t p = b->sel ? b->sel->idx[k] : static_cast<uint32_t>(k);
            for (size_t c = 0; c < nc; ++c) {
                const Column& col = b->cols[c];
                const bool valid = col.all_valid || validity::get_bit(col.validity, p);
                rs->nulls[c].push_back(valid ? 0 : 1);
                int64_t i = 0;
                double d = 0;
                if (valid) {
                    switch (col.type) {
                        case Type::I32:
                            i = reinterpret_cast<const int32_t*>(col.data)[p];
                            break;
                        case Type::STR:
                            i = reinterpret_cast<const int32_t*>(col.data)[p];
                            rs->sv[c].resize(rs->iv[c].size());
                            rs->sv[c].emplace_back(col.dict->at(static_cast<int32_t>(i)));
                            break;
                        case Type::BOOL:
                            i = reinterpret_cast<const uint8_t*>(col.data)[p];
                            break;
                        case Type::F64:
                            d = reinterpret_cast<const double*>(col.data)[p];
                            break;
                        default:
                            i = reinterpret_cast<const int64_t*>(col.data)[p];
                    }
                }
                rs->iv[c].push_back(i);
                rs->dv[c].push_back(d);
                if (col.type == Type::STR) rs->sv[c].resize(rs->iv[c].size());
            }
        }
        rs->nrows += b->row_count;
    }
    op.close();
    return rs;
}

}  // namespace

extern "C" {

// --- HashTable ---------------------------------------------------------------
void* ref_ht_new(const int32_t* types, uint32_t ncols, int32_t policy, uint64_t init_cap,
                 double lf, uint64_t seed) {
    std::vector<Type> kt;
    for (uint32_t j = 0; j < ncols; ++j) kt.push_back(ty(types[j]));
    HashTableConfig cfg;
    cfg.initial_capacity = init_cap;
    cfg.max_load_factor = lf;
    return new HashTable(kt, policy ? NullPolicy::kNeverMatch : NullPolicy::kEqual, cfg,
                         seed);
}
void ref_ht_free(void* h) { delete static_cast<HashTable*>(h); }
int ref_ht_insert_or_find(void* h, const orc_col* cols, const uint32_t* sel, uint64_t n,
                          uint32_t* out) {
    auto* t = static_cast<HashTable*>(h);
    std::vector<Column> v;
    for (size_t j = 0; j < t->key_types().size(); ++j) v.push_back(view_of(cols[j]));
    SelectionVector sv{sel, n};
    KeyColumns kc{v.data(), v.size(), sel ? &sv : nullptr};
    try {
        t->insert_or_find(kc, n, out);
    } catch (...) {
        return 1;
    }
    return 0;
}
int ref_ht_find(void* h, const orc_col* cols, const uint32_t* sel, uint64_t n,
                uint32_t* out) {
    auto* t = static_cast<HashTable*>(h);
    std::vector<Column> v;
    for (size_t j = 0; j < t->key_types().size(); ++j) v.push_back(view_of(cols[j]));
    SelectionVector sv{sel, n};
    KeyColumns kc{v.data(), v.size(), sel ? &sv : nullptr};
    t->find(kc, n, out);
    return 0;
}
uint64_t ref_ht_num_groups(void* h) { return static_cast<HashTable*>(h)->num_groups(); }
uint64_t ref_ht_key_word(void* h, uint32_t g, uint32_t col) {
    return static_cast<HashTable*>(h)->group_key_word(g, col);
}
int32_t ref_ht_key_is_null(void* h, uint32_t g, uint32_t col) {
    return static_cast<HashTable*>(h)->group_is_null(g, col);
}

// --- result sets -------------------------------------------------------------
uint64_t ref_rs_nrows(void* h) { return static_cast<ResultRows*>(h)->nrows; }
uint32_t ref_rs_ncols(void* h) {
    return static_cast<uint32_t>(static_cast<ResultRows*>(h)->types.size());
}
int32_t ref_rs_type(void* h, uint32_t c) { return static_cast<ResultRows*>(h)->types[c]; }
void ref_rs_col(void* h, uint32_t c, int64_t* iv, double* dv, uint8_t* nulls) {
    auto* r = static_cast<ResultRows*>(h);
    std::memcpy(iv, r->iv[c].data(), r->nrows * 8);
    std::memcpy(dv, r->dv[c].data(), r->nrows * 8);
    std::memcpy(nulls, r->nulls[c].data(), r->nrows);
}
void ref_rs_free(void* h) { delete static_cast<ResultRows*>(h); }
// STR column c: total bytes of its values; then the values concatenated with
// nrows + 1 offsets (a NULL row is an empty string; ref_rs_col has its NULL flag).
uint64_t ref_rs_str_bytes(void* h, uint32_t c) {
    uint64_t b = 0;
    for (const auto& x : static_cast<ResultRows*>(h)->sv[c]) b += x.size();
    return b;
}
void ref_rs_str_get(void* h, uint32_t c, char* bytes, uint64_t* offs) {
    const auto& v = static_cast<ResultRows*>(h)->sv[c];
    const uint64_t n = static_cast<ResultRows*>(h)->nrows;
    uint64_t o = 0;
    for (uint64_t r = 0; r < n; ++r) {
        offs[r] = o;
        if (r < v.size()) {
            std::memcpy(bytes + o, v[r].data(), v[r].size());
            o += v[r].size();
        }
    }
    offs[n] = o;
}

static std::vector<AggSpec> specs_of(const int32_t* funcs, const uint32_t* inputs,
                                     uint32_t naggs) {
    std::vector<AggSpec> s;
    for (uint32_t a = 0; a < naggs; ++a)
        s.push_back(AggSpec{static_cast<AggFunc>(funcs[a]), inputs[a],
                            "a" + std::to_string(a)});
    return s;
}

// Aggregate(Scan(t)) drained; rows come out in first-seen group order.
void* ref_agg_run(const orc_col* cols, uint32_t ncols, uint64_t nrows,
                  const uint32_t* key_cols, uint32_t nkeys, const int32_t* funcs,
                  const uint32_t* inputs, uint32_t naggs, uint64_t batch_size) {
    auto t = make_table(cols, ncols, nrows, false);
    Aggregate op(std::make_unique<Scan>(*t, batch_size),
                 std::vector<uint32_t>(key_cols, key_cols + nkeys),
                 specs_of(funcs, inputs, naggs));
    return drain(op);
}

// HashJoin(Scan(p), Scan(b)); the last two output columns' row ids are in the
// result (probe rowid at npc, build rowid at npc+nbc+1).
void* ref_join_run(const orc_col* pcols, uint32_t npc, uint64_t np, const orc_col* bcols,
                   uint32_t nbc, uint64_t nb, const uint32_t* pkeys, const uint32_t* bkeys,
                   uint32_t nkeys, int32_t type, uint64_t batch_size) {
    auto pt = make_table(pcols, npc, np, true);
    auto bt = make_table(bcols, nbc, nb, true);
    HashJoin op(std::make_unique<Scan>(*pt, batch_size),
                std::make_unique<Scan>(*bt, batch_size),
                std::vector<uint32_t>(pkeys, pkeys + nkeys),
                std::vector<uint32_t>(bkeys, bkeys + nkeys),
                type ? JoinType::Left : JoinType::Inner);
    return drain(op);
}

// Aggregate(HashJoin(Scan(p), Scan(b))): group keys / agg inputs index the join
// output (probe cols ++ build cols, no row ids).
void* ref_join_agg_run(const orc_col* pcols, uint32_t npc, uint64_t np,
                       const orc_col* bcols, uint32_t nbc, uint64_t nb,
                       const uint32_t* pkeys, const uint32_t* bkeys, uint32_t nkeys,
                       int32_t type, const uint32_t* gkeys, uint32_t ngk,
                       const int32_t* funcs, const uint32_t* inputs, uint32_t naggs,
                       uint64_t batch_size) {
    auto pt = make_table(pcols, npc, np, false);
    auto bt = make_table(bcols, nbc, nb, false);
    auto join = std::make_unique<HashJoin>(
        std::make_unique<Scan>(*pt, batch_size), std::make_unique<Scan>(*bt, batch_size),
        std::vector<uint32_t>(pkeys, pkeys + nkeys),
        std::vector<uint32_t>(bkeys, bkeys + nkeys), type ? JoinType::Left : JoinType::Inner);
    Aggregate op(std::move(join), std::vector<uint32_t>(gkeys, gkeys + ngk),
                 specs_of(funcs, inputs, naggs));
    return drain(op);
}

// tsx::AsofJoin(Scan(p), Scan(b)) with row-id columns appended to both sides.
void* ref_asof_run(const orc_col* pcols, uint32_t npc, uint64_t np, const orc_col* bcols,
                   uint32_t nbc, uint64_t nb, const uint32_t* pkeys, const uint32_t* bkeys,
                   uint32_t nkeys, uint32_t ptime, uint32_t btime, int32_t type,
                   int32_t has_tol, int64_t tol, uint64_t batch_size) {
    auto pt = make_table(pcols, npc, np, true);
    auto bt = make_table(bcols, nbc, nb, true);
    std::optional<int64_t> t;
    if (has_tol) t = tol;
    tsx::AsofJoin op(std::make_unique<Scan>(*pt, batch_size),
                     std::make_unique<Scan>(*bt, batch_size),
                     std::vector<uint32_t>(pkeys, pkeys + nkeys),
                     std::vector<uint32_t>(bkeys, bkeys + nkeys), ptime, btime,
                     type ? tsx::AsofType::Left : tsx::AsofType::Inner, t);
    return drain(op);
}

// Filter(Scan(t), cmp_0 AND cmp_1 AND ...) with a row-id column appended; the
// result's last column holds the passing rows' input indices.
typedef struct {
    uint32_t col;
    int32_t op;
    int32_t lit_type;
    int32_t lit_is_null;
    int64_t lit_i;
    double lit_f;
} ref_pred;

void* ref_filter_run(const orc_col* cols, uint32_t ncols, uint64_t n, const ref_pred* terms,
                     uint32_t nterms, uint64_t batch_size) {
    auto t = make_table(cols, ncols, n, true);
    expr::Expr pred;
    for (uint32_t j = 0; j < nterms; ++j) {
        const ref_pred& p = terms[j];
        expr::Scalar s;
        s.type = ty(p.lit_type);
        s.is_null = p.lit_is_null != 0;
        s.i = p.lit_i;
        s.f = p.lit_f;
        expr::Expr c = expr::cmp(static_cast<expr::CmpOp>(p.op),
                                 expr::col(ty(cols[p.col].type), p.col), expr::lit(s));
        pred = j == 0 ? c : expr::logic_and(pred, c);
    }
    Filter op(std::make_unique<Scan>(*t, batch_size), pred);
    return drain(op);
}

// Sort(Scan(t), keys) with a row-id column appended: the last output column is
// the permutation (ops/sort.h:84).
void* ref_sort_run(const orc_col* cols, uint32_t ncols, uint64_t n, const uint32_t* keys, const int32_t* desc,
                   const int32_t* nulls_first, uint32_t nk, uint64_t batch_size) {
    auto t = make_table(cols, ncols, n, true);
    std::vector<SortKey> sk;
    for (uint32_t j = 0; j < nk; ++j)
        sk.push_back(SortKey{keys[j], desc[j] ? SortDir::Desc : SortDir::Asc,
                             nulls_first[j] ? NullOrder::First : NullOrder::Last});
    Sort op(std::make_unique<Scan>(*t, batch_size), sk);
    return drain(op);
}

}  // extern "C"
