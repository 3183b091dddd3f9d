// Host drop-in operators over the C ABI. See qe_host.h.
#include "qe_host.h"
#include "qe_expr.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <new>

namespace qe {

// ---- owning storage --------------------------------------------------------------------
OwnedColumn OwnedColumn::make(Type t, size_t len) {
    OwnedColumn c;
    c.type_ = t;
    c.len_ = len;
    c.data_.assign(std::max<size_t>(len * byte_width(t), 1), std::byte{0});
    return c;
}

void OwnedColumn::set_null(size_t i) {
    if (validity_.empty()) validity_.assign((len_ + 63) / 64 + 1, ~0ull);
    validity_[i >> 6] &= ~(1ull << (i & 63));
    all_valid_ = false;
}

OwnedColumn OwnedColumn::make_str(size_t len, std::shared_ptr<const StringDict> dict) {
    OwnedColumn c = make(Type::STR, len);
    c.dict_ = std::move(dict);
    return c;
}

Column OwnedColumn::view() const {
    return Column{type_, len_, data_.data(), all_valid_ ? nullptr : validity_.data(), all_valid_, dict_.get()};
}

int32_t StringDict::intern(std::string_view s) {
    std::string key(s);
    auto it = index_.find(key);
    if (it != index_.end()) return it->second;
    const auto code = static_cast<int32_t>(offsets_.size() - 1);
    bytes_.insert(bytes_.end(), s.begin(), s.end());
    offsets_.push_back(static_cast<uint32_t>(bytes_.size()));
    index_.emplace(std::move(key), code);
    return code;
}

Batch OwnedBatch::view() const {
    Batch b;
    b.sel = nullptr;
    b.row_count = cols_.empty() ? 0 : cols_[0].len();
    for (const auto& c : cols_) b.cols.push_back(c.view());
    return b;
}

Table::Table(Schema schema, std::vector<OwnedColumn> cols) : schema_(std::move(schema)), cols_(std::move(cols)) {
    if (cols_.size() != schema_.fields.size()) throw std::invalid_argument("Table: schema/column count mismatch");
    rows_ = cols_.empty() ? 0 : cols_[0].len();
    for (const auto& c : cols_)
        if (c.len() != rows_) throw std::invalid_argument("Table: columns differ in length");
}

// ---- device plumbing -----------------------------------------------------------------------
void throw_status(qe_gpu_status st, qe_gpu_ctx* ctx) {
    if (st == QE_GPU_OK) return;
    const std::string msg = ctx ? qe_gpu_last_error(ctx) : "no device context";
    switch (st) {
        case QE_GPU_INVALID_ARGUMENT:
            throw std::invalid_argument(msg);
        case QE_GPU_OVERFLOW:
            throw std::overflow_error(msg);
        case QE_GPU_OUT_OF_MEMORY:
            throw std::bad_alloc();
        default:
            throw std::runtime_error("device: " + msg);
    }
}

qe_gpu_ctx* device_context() {
    static qe_gpu_ctx* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const qe_gpu_status st = qe_gpu_ctx_create(0, &ctx);
        if (st != QE_GPU_OK) throw std::runtime_error("no MI355X device context (status " + std::to_string(st) + ")");
    });
    return ctx;
}

#define QE_CALL(expr)                                        \
    do {                                                     \
        const qe_gpu_status st_ = (expr);                    \
        if (st_ != QE_GPU_OK) throw_status(st_, device_context()); \
    } while (0)

// Device dictionary -> the host dictionary it was made from (kept alive while
// registered), so STR results come back with a host dictionary.
namespace {
std::mutex g_dict_mu;
std::map<const qe_gpu_dict*, std::shared_ptr<const StringDict>> g_dicts;

qe_gpu_dict* register_dict(std::shared_ptr<const StringDict> host) {
    std::vector<char> bytes;
    std::vector<uint32_t> off{0};
    for (size_t i = 0; i < host->size(); ++i) {
        const std::string_view v = host->at(static_cast<int32_t>(i));
        bytes.insert(bytes.end(), v.begin(), v.end());
        off.push_back(static_cast<uint32_t>(bytes.size()));
    }
    qe_gpu_dict* d = nullptr;
    QE_CALL(qe_gpu_dict_create(device_context(), bytes.data(), off.data(), static_cast<uint32_t>(host->size()), &d));
    std::lock_guard<std::mutex> g(g_dict_mu);
    g_dicts[d] = std::move(host);
    return d;
}

void unregister_dict(qe_gpu_dict* d) {
    {
        std::lock_guard<std::mutex> g(g_dict_mu);
        g_dicts.erase(d);
    }
    qe_gpu_dict_free(device_context(), d);
}
}  // namespace

std::shared_ptr<const StringDict> host_dict_of(const qe_dev_col* c) {
    const qe_gpu_dict* d = qe_gpu_column_dict(c);
    std::lock_guard<std::mutex> g(g_dict_mu);
    auto it = g_dicts.find(d);
    if (it == g_dicts.end()) throw std::logic_error("device STR column without a registered host dictionary");
    return it->second;
}

DeviceColumns::~DeviceColumns() {
    if (owned)
        for (auto* c : cols) qe_gpu_column_free(device_context(), c);
    for (auto* d : dicts) unregister_dict(d);
}

// One column to HBM; a STR column's dictionary goes with it (one device copy
// per host dictionary per upload).
static qe_dev_col* upload_view(const Column& c, DeviceColumns& dc, std::shared_ptr<const StringDict> owner,
                               std::map<const StringDict*, qe_gpu_dict*>& made) {
    qe_col_desc d{static_cast<int32_t>(c.type), c.all_valid ? 1 : 0, c.len, c.data, c.validity};
    qe_dev_col* out = nullptr;
    QE_CALL(qe_gpu_column_upload(device_context(), &d, &out));
    if (c.type == Type::STR) {
        if (!c.dict || !owner) {
            qe_gpu_column_free(device_context(), out);
            throw std::invalid_argument("STR column without a dictionary");
        }
        qe_gpu_dict*& gd = made[c.dict];
        if (!gd) {
            gd = register_dict(std::move(owner));
            dc.dicts.push_back(gd);
        }
        QE_CALL(qe_gpu_column_set_dict(device_context(), out, gd));
    }
    return out;
}

std::shared_ptr<DeviceColumns> upload_table(const Table& t) {
    if (t.device_) return t.device_;
    auto dc = std::make_shared<DeviceColumns>();
    dc->rows = t.num_rows();
    std::map<const StringDict*, qe_gpu_dict*> made;
    for (size_t i = 0; i < t.schema().fields.size(); ++i)
        dc->cols.push_back(upload_view(t.column(i).view(), *dc, t.column(i).dict_shared(), made));
    t.device_ = dc;  // uploaded once per Table (inputs stay resident in HBM)
    return dc;
}

void drop_device_copy(const Table& t) { t.device_.reset(); }

// Drain any operator to host columns (honouring selection vectors), then upload.
std::shared_ptr<DeviceColumns> device_input(Operator& op) {
    if (auto* ds = dynamic_cast<DeviceSource*>(&op)) return ds->device_columns();
    const Schema s = op.output_schema();
    std::vector<OwnedColumn> acc;
    std::vector<std::vector<std::byte>> data(s.fields.size());
    std::vector<std::vector<uint8_t>> valid(s.fields.size());
    // STR values are re-interned into one dictionary per column: batches may
    // carry different dictionaries (the BuildStore rule, ops/join.cpp:150-172).
    std::vector<std::shared_ptr<StringDict>> dicts(s.fields.size());
    for (size_t c = 0; c < s.fields.size(); ++c)
        if (s.fields[c].second == Type::STR) dicts[c] = std::make_shared<StringDict>();
    size_t rows = 0;
    op.open();
    while (auto b = op.next()) {
        for (size_t c = 0; c < s.fields.size(); ++c) {
            const Column& col = b->cols[c];
            const size_t w = byte_width(col.type);
            for (size_t k = 0; k < b->row_count; ++k) {
                const size_t p = b->sel ? b->sel->idx[k] : k;
                const bool v = col.all_valid || ((col.validity[p >> 6] >> (p & 63)) & 1ull);
                if (dicts[c]) {
                    const int32_t code =
                        v ? dicts[c]->intern(col.dict->at(reinterpret_cast<const int32_t*>(col.data)[p])) : 0;
                    const auto* cb = reinterpret_cast<const std::byte*>(&code);
                    data[c].insert(data[c].end(), cb, cb + 4);
                } else {
                    const std::byte* src = col.data + p * w;
                    data[c].insert(data[c].end(), src, src + w);
                }
                valid[c].push_back(v);
            }
        }
        rows += b->row_count;
    }
    op.close();
    auto dc = std::make_shared<DeviceColumns>();
    dc->rows = rows;
    std::map<const StringDict*, qe_gpu_dict*> made;
    for (size_t c = 0; c < s.fields.size(); ++c) {
        OwnedColumn oc = dicts[c] ? OwnedColumn::make_str(rows, dicts[c]) : OwnedColumn::make(s.fields[c].second, rows);
        if (rows) std::memcpy(oc.mutable_data(), data[c].data(), data[c].size());
        for (size_t r = 0; r < rows; ++r)
            if (!valid[c][r]) oc.set_null(r);
        dc->cols.push_back(upload_view(oc.view(), *dc, dicts[c], made));
    }
    return dc;
}

// ---- Scan ----------------------------------------------------------------------------------
Scan::Scan(const Table& t, size_t batch_size) : table_(t), batch_size_(batch_size) {
    if (batch_size_ == 0 || batch_size_ % 64 != 0)
        throw std::invalid_argument("Scan: batch_size must be a positive multiple of 64");
}

std::optional<Batch> Scan::next() {
    if (cursor_ >= table_.num_rows()) return std::nullopt;
    const size_t m = std::min(batch_size_, table_.num_rows() - cursor_);
    Batch b;
    b.sel = nullptr;
    b.row_count = m;
    for (size_t c = 0; c < table_.schema().fields.size(); ++c) {
        Column v = table_.column(c).view();
        v.data += cursor_ * byte_width(v.type);
        if (!v.all_valid) v.validity += cursor_ / 64;
        v.len = m;
        b.cols.push_back(v);
    }
    cursor_ += m;
    return b;
}

// The materialised join output stays on the device for a device parent: the
// columns are whole-column views into a gathered table that lives as long as
// the returned object.
struct TableColumns : DeviceColumns {
    qe_dev_table* table = nullptr;
    ~TableColumns() {
        for (auto* c : cols) qe_gpu_column_free(device_context(), c);  // views: frees the handle only
        cols.clear();
        if (table) qe_gpu_table_free(device_context(), table);
    }
};

std::shared_ptr<DeviceColumns> table_columns(qe_dev_table* t, std::vector<std::shared_ptr<const DeviceColumns>> sources) {
    auto keep = std::make_shared<TableColumns>();
    keep->table = t;
    keep->sources = std::move(sources);
    keep->rows = qe_gpu_table_rows(t);
    for (uint32_t i = 0; i < qe_gpu_table_cols(t); ++i) {
        qe_dev_col* v = nullptr;
        QE_CALL(qe_gpu_column_slice(device_context(), qe_gpu_table_col(t, i), 0, keep->rows, &v));
        keep->cols.push_back(v);
    }
    return keep;
}

// ---- result emission ----------------------------------------------------------------------
void ResultEmitter::reset(qe_dev_table* t, std::vector<Type> types,
                          std::vector<std::shared_ptr<const DeviceColumns>> sources) {
    if (table_) qe_gpu_table_free(device_context(), table_);
    table_ = t;
    types_ = std::move(types);
    sources_ = std::move(sources);
    cursor_ = 0;
}

ResultEmitter::~ResultEmitter() {
    if (table_) qe_gpu_table_free(device_context(), table_);
}

std::optional<Batch> ResultEmitter::next() {
    if (!table_) return std::nullopt;
    const size_t rows = qe_gpu_table_rows(table_);
    if (cursor_ >= rows) return std::nullopt;
    const size_t m = std::min<size_t>(2048, rows - cursor_);
    OwnedBatch out;
    std::vector<uint8_t> vb(m);
    for (uint32_t c = 0; c < qe_gpu_table_cols(table_); ++c) {
        OwnedColumn oc = OwnedColumn::make(types_[c], m);
        if (types_[c] == Type::STR) oc.set_dict(host_dict_of(qe_gpu_table_col(table_, c)));
        QE_CALL(qe_gpu_column_download_range(device_context(), qe_gpu_table_col(table_, c), cursor_, m,
                                             oc.mutable_data(), vb.data()));
        for (size_t r = 0; r < m; ++r)
            if (!vb[r]) oc.set_null(r);
        out.add_column(std::move(oc));
    }
    cursor_ += m;
    current_ = std::move(out);
    return current_.view();
}

Type agg_result_type(AggFunc f, Type input) {
    switch (f) {
        case AggFunc::CountStar:
        case AggFunc::Count:
            return Type::I64;
        case AggFunc::Sum:
            if (input == Type::STR) throw std::invalid_argument("SUM is not defined on STR (VARCHAR)");
            return input == Type::F64 ? Type::F64 : Type::I64;
        case AggFunc::Avg:
            if (input == Type::STR) throw std::invalid_argument("AVG is not defined on STR (VARCHAR)");
            return Type::F64;
        default:
            return input;
    }
}

static std::vector<qe_agg_spec> abi_aggs(const std::vector<AggSpec>& aggs) {
    std::vector<qe_agg_spec> out;
    for (const auto& a : aggs) out.push_back(qe_agg_spec{static_cast<int32_t>(a.func), a.input_col});
    return out;
}

static Schema agg_schema(const Schema& in, const std::vector<uint32_t>& keys, const std::vector<AggSpec>& aggs) {
    Schema s;
    for (uint32_t k : keys) s.fields.push_back(in.fields.at(k));
    for (const auto& a : aggs) {
        const Type t = a.func == AggFunc::CountStar ? Type::I64 : in.fields.at(a.input_col).second;
        s.fields.emplace_back(a.out_name, agg_result_type(a.func, t));
    }
    return s;
}

static std::vector<Type> types_of(const Schema& s) {
    std::vector<Type> t;
    for (const auto& f : s.fields) t.push_back(f.second);
    return t;
}

// ---- GpuAggregate ----------------------------------------------------------------------------
GpuAggregate::GpuAggregate(std::unique_ptr<Operator> child, std::vector<uint32_t> key_cols,
                           std::vector<AggSpec> aggs)
    : child_(std::move(child)), key_cols_(std::move(key_cols)), aggs_(std::move(aggs)) {
    if (aggs_.empty()) throw std::invalid_argument("Aggregate needs >=1 aggregate");
    child_schema_ = child_->output_schema();
    (void)agg_schema(child_schema_, key_cols_, aggs_);  // validates types / indices
}

Schema GpuAggregate::output_schema() const { return agg_schema(child_schema_, key_cols_, aggs_); }

qe_dev_table* GpuAggregate::run(std::shared_ptr<DeviceColumns>& in) {
    in = device_input(*child_);
    auto specs = abi_aggs(aggs_);
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_aggregate(device_context(), in->cols.data(), static_cast<uint32_t>(in->cols.size()),
                             key_cols_.data(), static_cast<uint32_t>(key_cols_.size()), specs.data(),
                             static_cast<uint32_t>(specs.size()), &t));
    return t;
}

std::shared_ptr<DeviceColumns> GpuAggregate::device_columns() {
    std::shared_ptr<DeviceColumns> in;
    qe_dev_table* t = run(in);
    return table_columns(t, {in});
}

void GpuAggregate::open() {
    std::shared_ptr<DeviceColumns> in;
    qe_dev_table* t = run(in);
    emit_.reset(t, types_of(output_schema()), {in});
}

void GpuAggregate::close() { emit_.reset(nullptr, {}); }

// ---- GpuHashJoin -------------------------------------------------------------------------------
static Schema concat(const Schema& a, const Schema& b) {
    Schema s = a;
    for (const auto& f : b.fields) s.fields.push_back(f);
    return s;
}

static void check_keys(const Schema& ps, const Schema& bs, const std::vector<uint32_t>& pk,
                       const std::vector<uint32_t>& bk, const char* who) {
    if (pk.size() != bk.size()) throw std::invalid_argument(std::string(who) + ": key lists differ in length");
    for (size_t i = 0; i < pk.size(); ++i)
        if (ps.fields.at(pk[i]).second != bs.fields.at(bk[i]).second)
            throw std::invalid_argument(std::string(who) + ": join key column types must match positionally");
}

GpuHashJoin::GpuHashJoin(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                         std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys, JoinType type)
    : probe_(std::move(probe)), build_(std::move(build)), probe_keys_(std::move(probe_keys)),
      build_keys_(std::move(build_keys)), type_(type) {
    probe_schema_ = probe_->output_schema();
    build_schema_ = build_->output_schema();
    if (probe_keys_.empty()) throw std::invalid_argument("HashJoin: equi-join needs >=1 key");
    check_keys(probe_schema_, build_schema_, probe_keys_, build_keys_, "HashJoin");
}

Schema GpuHashJoin::output_schema() const { return concat(probe_schema_, build_schema_); }



std::shared_ptr<DeviceColumns> GpuHashJoin::device_columns() {
    open();
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_pairs_gather(device_context(), pairs_, 0, qe_gpu_pairs_count(pairs_), pcols_->cols.data(),
                                static_cast<uint32_t>(pcols_->cols.size()), bcols_->cols.data(),
                                static_cast<uint32_t>(bcols_->cols.size()), &t));
    auto keep = table_columns(t, {pcols_, bcols_});
    close();
    return keep;
}

void GpuHashJoin::open() {
    pcols_ = device_input(*probe_);
    bcols_ = device_input(*build_);
    std::vector<qe_dev_col*> pk, bk;
    for (uint32_t k : probe_keys_) pk.push_back(pcols_->cols.at(k));
    for (uint32_t k : build_keys_) bk.push_back(bcols_->cols.at(k));
    if (pairs_) qe_gpu_pairs_free(device_context(), pairs_);
    pairs_ = nullptr;
    QE_CALL(qe_gpu_hash_join(device_context(), pk.data(), bk.data(), static_cast<uint32_t>(pk.size()),
                             type_ == JoinType::Left ? QE_JOIN_LEFT : QE_JOIN_INNER, &pairs_));
    cursor_ = 0;
}

std::optional<Batch> GpuHashJoin::next() {
    // Drain the current chunk, then gather the next <= 2048 pairs.
    if (auto b = chunk_.next()) return b;
    const size_t total = pairs_ ? qe_gpu_pairs_count(pairs_) : 0;
    if (cursor_ >= total) return std::nullopt;
    const size_t m = std::min<size_t>(kOutBatch, total - cursor_);
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_pairs_gather(device_context(), pairs_, cursor_, m, pcols_->cols.data(),
                                static_cast<uint32_t>(pcols_->cols.size()), bcols_->cols.data(),
                                static_cast<uint32_t>(bcols_->cols.size()), &t));
    cursor_ += m;
    chunk_.reset(t, types_of(output_schema()));
    return chunk_.next();
}

void GpuHashJoin::close() {
    chunk_.reset(nullptr, {});
    if (pairs_) qe_gpu_pairs_free(device_context(), pairs_);
    pairs_ = nullptr;
    pcols_.reset();
    bcols_.reset();
}

// ---- GpuJoinAggregate -----------------------------------------------------------------------------
GpuJoinAggregate::GpuJoinAggregate(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                                   std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys,
                                   JoinType type, std::vector<uint32_t> group_keys, std::vector<AggSpec> aggs)
    : probe_(std::move(probe)), build_(std::move(build)), probe_keys_(std::move(probe_keys)),
      build_keys_(std::move(build_keys)), type_(type), group_keys_(std::move(group_keys)), aggs_(std::move(aggs)) {
    const Schema ps = probe_->output_schema(), bs = build_->output_schema();
    if (probe_keys_.empty()) throw std::invalid_argument("HashJoin: equi-join needs >=1 key");
    check_keys(ps, bs, probe_keys_, build_keys_, "HashJoin");
    join_schema_ = concat(ps, bs);
    if (aggs_.empty()) throw std::invalid_argument("Aggregate needs >=1 aggregate");
    (void)agg_schema(join_schema_, group_keys_, aggs_);
}

Schema GpuJoinAggregate::output_schema() const { return agg_schema(join_schema_, group_keys_, aggs_); }

std::shared_ptr<DeviceColumns> GpuJoinAggregate::device_columns() {
    std::shared_ptr<DeviceColumns> pc, bc;
    qe_dev_table* t = run(pc, bc);
    return table_columns(t, {pc, bc});
}

void GpuJoinAggregate::open() {
    std::shared_ptr<DeviceColumns> pc, bc;
    qe_dev_table* t = run(pc, bc);
    emit_.reset(t, types_of(output_schema()), {pc, bc});
}

qe_dev_table* GpuJoinAggregate::run(std::shared_ptr<DeviceColumns>& pc, std::shared_ptr<DeviceColumns>& bc) {
    pc = device_input(*probe_);
    bc = device_input(*build_);
    auto specs = abi_aggs(aggs_);
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_join_aggregate(device_context(), pc->cols.data(), static_cast<uint32_t>(pc->cols.size()),
                                  bc->cols.data(), static_cast<uint32_t>(bc->cols.size()), probe_keys_.data(),
                                  build_keys_.data(), static_cast<uint32_t>(probe_keys_.size()),
                                  type_ == JoinType::Left ? QE_JOIN_LEFT : QE_JOIN_INNER, group_keys_.data(),
                                  static_cast<uint32_t>(group_keys_.size()), specs.data(),
                                  static_cast<uint32_t>(specs.size()), &t));
    return t;
}

void GpuJoinAggregate::close() { emit_.reset(nullptr, {}); }

// ---- GpuAsofJoin ------------------------------------------------------------------------------------
GpuAsofJoin::GpuAsofJoin(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                         std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys, uint32_t probe_time,
                         uint32_t build_time, tsx::AsofType type, std::optional<int64_t> tolerance,
                         tsx::AsofMethod method)
    : probe_(std::move(probe)), build_(std::move(build)), probe_keys_(std::move(probe_keys)),
      build_keys_(std::move(build_keys)), probe_time_(probe_time), build_time_(build_time), type_(type),
      tolerance_(tolerance), method_(method) {
    probe_schema_ = probe_->output_schema();
    build_schema_ = build_->output_schema();
    auto ts_ok = [](Type t) { return t == Type::I32 || t == Type::I64 || t == Type::TS; };
    if (!ts_ok(probe_schema_.fields.at(probe_time_).second) || !ts_ok(build_schema_.fields.at(build_time_).second))
        throw std::invalid_argument("AsofJoin: time columns must be I32/I64/TS");
    check_keys(probe_schema_, build_schema_, probe_keys_, build_keys_, "AsofJoin");
}

Schema GpuAsofJoin::output_schema() const { return concat(probe_schema_, build_schema_); }

static std::atomic<size_t> g_cluster_min_rows{size_t(1) << 22};
static std::atomic<uint64_t> g_cluster_min_line_bytes{uint64_t(1) << 31};

void GpuAsofJoin::set_cluster_threshold(size_t min_rows, uint64_t min_line_bytes) {
    g_cluster_min_rows = min_rows;
    g_cluster_min_line_bytes = min_line_bytes;
}

void GpuAsofJoin::open() {
    pcols_ = device_input(*probe_);
    bcols_ = device_input(*build_);
    std::vector<qe_dev_col*> pk, bk;
    for (uint32_t k : probe_keys_) pk.push_back(pcols_->cols.at(k));
    for (uint32_t k : build_keys_) bk.push_back(bcols_->cols.at(k));
    qe_gpu_ctx* ctx = device_context();
    // The index (sorted build side, and the learned router) is kept with the
    // build columns, so a plan re-run over the same Table reuses it.
    std::string sig = "asof";
    for (uint32_t k : build_keys_) sig += ":" + std::to_string(k);
    sig += "|" + std::to_string(build_time_) + (method_ == tsx::AsofMethod::Learned ? "|learned" : "|binary");
    std::shared_ptr<void>& slot = bcols_->derived[sig];
    if (!slot) {
        qe_dev_asof* built = nullptr;
        QE_CALL(qe_gpu_asof_build(ctx, bk.data(), static_cast<uint32_t>(bk.size()), bcols_->cols.at(build_time_),
                                  &built));
        std::shared_ptr<void> owner(built, [](void* p) { qe_gpu_asof_free(device_context(), static_cast<qe_dev_asof*>(p)); });
        // One knot per router bucket (the context default), so most buckets
        // hold their rows inline in one 128-byte line: pieces of 32 rows, or
        // 64 above 2^30 rows, where the probe's random lines would otherwise
        // span more than the GPU's address translation covers.
        if (method_ == tsx::AsofMethod::Learned) {
            qe_asof_stats st{};
            QE_CALL(qe_gpu_asof_stats(ctx, built, &st));
            QE_CALL(qe_gpu_asof_train(ctx, built, st.build_rows > (uint64_t(1) << 30) ? 64 : 32, 300, 0.003f));
        }
        slot = std::move(owner);
    }
    qe_dev_asof* idx = static_cast<qe_dev_asof*>(slot.get());
    // A learned probe over an index whose bucket lines outgrow what the GPU's
    // address translation covers (about 2 GiB) first groups its probe rows by
    // the index region they read (qe_gpu_asof_cluster): the join then emits in
    // that order, input order within a group.
    if (method_ == tsx::AsofMethod::Learned && pcols_->rows >= g_cluster_min_rows.load()) {
        qe_asof_stats st{};
        QE_CALL(qe_gpu_asof_stats(ctx, idx, &st));
        if (st.router_buckets * 128 >= g_cluster_min_line_bytes.load()) {
            qe_dev_table* t = nullptr;
            QE_CALL(qe_gpu_asof_cluster(ctx, idx, pk.data(), static_cast<uint32_t>(pk.size()),
                                        pcols_->cols.at(probe_time_), pcols_->cols.data(),
                                        static_cast<uint32_t>(pcols_->cols.size()), &t));
            pcols_ = table_columns(t, {pcols_});
            pk.clear();
            for (uint32_t k : probe_keys_) pk.push_back(pcols_->cols.at(k));
        }
    }
    const size_t n = pcols_->rows;
    void* match = nullptr;
    try {
        const int32_t mode = method_ == tsx::AsofMethod::Learned ? QE_ASOF_LEARNED : QE_ASOF_BINARY;
        QE_CALL(qe_gpu_malloc(ctx, std::max<size_t>(n, 1) * 4, &match));
        QE_CALL(qe_gpu_asof_probe(ctx, idx, pk.data(), static_cast<uint32_t>(pk.size()), pcols_->cols.at(probe_time_),
                                  tolerance_ ? 1 : 0, tolerance_ ? *tolerance_ : 0, mode,
                                  static_cast<uint32_t*>(match)));
        if (pairs_) qe_gpu_pairs_free(ctx, pairs_);
        pairs_ = nullptr;
        QE_CALL(qe_gpu_asof_pairs(ctx, static_cast<uint32_t*>(match), n,
                                  type_ == tsx::AsofType::Left ? QE_JOIN_LEFT : QE_JOIN_INNER, &pairs_));
    } catch (...) {
        if (match) qe_gpu_free(ctx, match);
        throw;
    }
    qe_gpu_free(ctx, match);
    cursor_ = 0;
}

std::shared_ptr<DeviceColumns> GpuAsofJoin::device_columns() {
    open();
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_pairs_gather(device_context(), pairs_, 0, qe_gpu_pairs_count(pairs_), pcols_->cols.data(),
                                static_cast<uint32_t>(pcols_->cols.size()), bcols_->cols.data(),
                                static_cast<uint32_t>(bcols_->cols.size()), &t));
    auto keep = table_columns(t, {pcols_, bcols_});
    close();
    return keep;
}

std::optional<Batch> GpuAsofJoin::next() {
    if (auto b = chunk_.next()) return b;
    const size_t total = pairs_ ? qe_gpu_pairs_count(pairs_) : 0;
    if (cursor_ >= total) return std::nullopt;
    const size_t m = std::min<size_t>(kOutBatch, total - cursor_);
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_pairs_gather(device_context(), pairs_, cursor_, m, pcols_->cols.data(),
                                static_cast<uint32_t>(pcols_->cols.size()), bcols_->cols.data(),
                                static_cast<uint32_t>(bcols_->cols.size()), &t));
    cursor_ += m;
    chunk_.reset(t, types_of(output_schema()));
    return chunk_.next();
}

void GpuAsofJoin::close() {
    chunk_.reset(nullptr, {});
    if (pairs_) qe_gpu_pairs_free(device_context(), pairs_);
    pairs_ = nullptr;
    pcols_.reset();
    bcols_.reset();
}

// ---- GpuSort ---------------------------------------------------------------------------------------
GpuSort::GpuSort(std::unique_ptr<Operator> child, std::vector<SortKey> keys)
    : child_(std::move(child)), keys_(std::move(keys)) {
    if (keys_.empty()) throw std::invalid_argument("Sort needs >=1 key");
    child_schema_ = child_->output_schema();
    for (const SortKey& k : keys_)
        if (k.col >= child_schema_.fields.size()) throw std::invalid_argument("Sort: key column out of range");
}

qe_dev_table* GpuSort::run(std::shared_ptr<DeviceColumns>& in) {
    in = device_input(*child_);
    qe_gpu_ctx* ctx = device_context();
    const size_t n = in->rows;
    std::vector<qe_dev_col*> kc;
    std::vector<int32_t> desc, nf;
    for (const SortKey& k : keys_) {
        kc.push_back(in->cols.at(k.col));
        desc.push_back(k.dir == SortDir::Desc ? 1 : 0);
        nf.push_back(k.nulls == NullOrder::First ? 1 : 0);
    }
    void* perm = nullptr;
    qe_dev_table* t = nullptr;
    try {
        QE_CALL(qe_gpu_malloc(ctx, std::max<size_t>(n, 1) * 4, &perm));
        if (n) QE_CALL(qe_gpu_sort_perm(ctx, kc.data(), static_cast<uint32_t>(kc.size()), desc.data(), nf.data(),
                                        static_cast<uint32_t*>(perm)));
        QE_CALL(qe_gpu_take(ctx, in->cols.data(), static_cast<uint32_t>(in->cols.size()),
                            static_cast<uint32_t*>(perm), n, &t));
    } catch (...) {
        if (perm) qe_gpu_free(ctx, perm);
        throw;
    }
    qe_gpu_free(ctx, perm);
    return t;
}

void GpuSort::open() {
    std::shared_ptr<DeviceColumns> in;
    qe_dev_table* t = run(in);
    emit_.reset(t, types_of(output_schema()), {in});
}

void GpuSort::close() { emit_.reset(nullptr, {}); }

std::shared_ptr<DeviceColumns> GpuSort::device_columns() {
    std::shared_ptr<DeviceColumns> in;
    qe_dev_table* t = run(in);
    return table_columns(t, {in});
}

// ---- GpuHashTable ----------------------------------------------------------------------------------
GpuHashTable::GpuHashTable(std::vector<Type> key_types, NullPolicy p, HashTableConfig cfg, uint64_t seed)
    : types_(std::move(key_types)) {
    std::vector<int32_t> t;
    for (Type k : types_) t.push_back(static_cast<int32_t>(k));
    QE_CALL(qe_gpu_ht_create(device_context(), t.data(), static_cast<uint32_t>(t.size()),
                             p == NullPolicy::kNeverMatch ? QE_NULL_NEVER_MATCH : QE_NULL_EQUAL,
                             cfg.initial_capacity, cfg.max_load_factor, seed, &ht_));
}

GpuHashTable::~GpuHashTable() { qe_gpu_ht_free(device_context(), ht_); }

void GpuHashTable::run(const KeyColumns& keys, size_t n, uint32_t* out, bool insert) const {
    qe_gpu_ctx* ctx = device_context();
    std::vector<qe_dev_col*> cols;
    void* sel = nullptr;
    void* dout = nullptr;
    try {
        // The HashTable hashes STR keys by raw code (ops/hash_internal.cpp:29-60);
        // callers canonicalise values first, as Aggregate and HashJoin do.
        for (size_t j = 0; j < keys.num_cols; ++j) {
            const Column& c = keys.cols[j];
            qe_col_desc d{static_cast<int32_t>(c.type), c.all_valid ? 1 : 0, c.len, c.data, c.validity};
            qe_dev_col* dc = nullptr;
            QE_CALL(qe_gpu_column_upload(ctx, &d, &dc));
            cols.push_back(dc);
        }
        if (keys.sel) {
            QE_CALL(qe_gpu_malloc(ctx, std::max<size_t>(n, 1) * 4, &sel));
            QE_CALL(qe_gpu_memcpy_h2d(ctx, sel, keys.sel->idx, n * 4));
        }
        QE_CALL(qe_gpu_malloc(ctx, std::max<size_t>(n, 1) * 4, &dout));
        if (insert)
            QE_CALL(qe_gpu_ht_insert_or_find(ctx, ht_, cols.data(), static_cast<uint32_t*>(sel), n,
                                             static_cast<uint32_t*>(dout)));
        else
            QE_CALL(qe_gpu_ht_find(ctx, ht_, cols.data(), static_cast<uint32_t*>(sel), n, static_cast<uint32_t*>(dout)));
        QE_CALL(qe_gpu_memcpy_d2h(ctx, out, dout, n * 4));
    } catch (...) {
        for (auto* c : cols) qe_gpu_column_free(ctx, c);
        if (sel) qe_gpu_free(ctx, sel);
        if (dout) qe_gpu_free(ctx, dout);
        throw;
    }
    for (auto* c : cols) qe_gpu_column_free(ctx, c);
    if (sel) qe_gpu_free(ctx, sel);
    qe_gpu_free(ctx, dout);
}

void GpuHashTable::insert_or_find(const KeyColumns& keys, size_t n, uint32_t* out, HashPath) {
    run(keys, n, out, true);
}
void GpuHashTable::find(const KeyColumns& keys, size_t n, uint32_t* out, HashPath) const { run(keys, n, out, false); }
size_t GpuHashTable::num_groups() const { return qe_gpu_ht_num_groups(ht_); }

bool GpuHashTable::group_is_null(uint32_t g, size_t col) const {
    const size_t n = num_groups();
    std::vector<uint64_t> w(std::max<size_t>(n, 1));
    std::vector<uint8_t> nl(std::max<size_t>(n, 1));
    QE_CALL(qe_gpu_ht_key_words(device_context(), ht_, static_cast<uint32_t>(col), w.data(), nl.data()));
    return nl.at(g) != 0;
}

uint64_t GpuHashTable::group_key_word(uint32_t g, size_t col) const {
    const size_t n = num_groups();
    std::vector<uint64_t> w(std::max<size_t>(n, 1));
    std::vector<uint8_t> nl(std::max<size_t>(n, 1));
    QE_CALL(qe_gpu_ht_key_words(device_context(), ht_, static_cast<uint32_t>(col), w.data(), nl.data()));
    return w.at(g);
}

// ---- GpuFilter ----------------------------------------------------------------------------------
namespace {

struct FlatPred {
    std::vector<qe_pred_term> terms;
    bool never = false;  // a FALSE or NULL conjunct: no row passes
};

[[noreturn]] void filter_shape(const char* what) {
    throw std::runtime_error(std::string("GpuFilter: ") + what +
                             " is not on the device path (conjunctions of column-versus-literal comparisons)");
}

expr::CmpOp mirrored(expr::CmpOp op) {  // a < c  <=>  c > a
    switch (op) {
        case expr::CmpOp::Lt:
            return expr::CmpOp::Gt;
        case expr::CmpOp::Le:
            return expr::CmpOp::Ge;
        case expr::CmpOp::Gt:
            return expr::CmpOp::Lt;
        case expr::CmpOp::Ge:
            return expr::CmpOp::Le;
        default:
            return op;
    }
}

void check_column(const expr::Node& n, const Schema& s) {  // plan/plan.cpp:38-54
    if (n.col_index >= s.fields.size())
        throw std::invalid_argument("Filter: predicate column " + std::to_string(n.col_index) + " out of range");
    if (s.fields[n.col_index].second != n.type)
        throw std::invalid_argument("Filter: predicate column " + std::to_string(n.col_index) + " type mismatch");
}

void flatten(const expr::Expr& e, const Schema& s, FlatPred& fp) {
    const expr::Node& n = e.node();
    switch (n.kind) {
        case expr::NodeKind::Logic:
            if (n.logic_op != expr::LogicOp::And) filter_shape("OR / NOT");
            for (const expr::Expr& c : n.children) flatten(c, s, fp);
            return;
        case expr::NodeKind::Literal:  // a BOOL literal conjunct
            if (n.literal.is_null || n.literal.i == 0) fp.never = true;
            return;
        case expr::NodeKind::Column: {  // a BOOL column conjunct: TRUE iff valid and != 0
            check_column(n, s);
            qe_pred_term t{};
            t.col = n.col_index;
            t.op = QE_CMP_NE;
            t.lit_type = QE_BOOL;
            fp.terms.push_back(t);
            return;
        }
        case expr::NodeKind::Cmp: {
            const expr::Node* c = &n.children[0].node();
            const expr::Node* l = &n.children[1].node();
            expr::CmpOp op = n.cmp_op;
            if (c->kind == expr::NodeKind::Literal && l->kind == expr::NodeKind::Column) {
                std::swap(c, l);
                op = mirrored(op);
            }
            if (c->kind != expr::NodeKind::Column || l->kind != expr::NodeKind::Literal)
                filter_shape("a comparison that is not column-versus-literal");
            if (n.cmp_type == Type::STR) filter_shape("a STR comparison");
            check_column(*c, s);
            qe_pred_term t{};
            t.col = c->col_index;
            t.op = static_cast<int32_t>(op);
            t.lit_type = static_cast<int32_t>(l->literal.type);
            t.lit_is_null = l->literal.is_null ? 1 : 0;
            t.lit_i = l->literal.i;
            t.lit_f = l->literal.f;
            fp.terms.push_back(t);
            return;
        }
    }
}

}  // namespace

GpuFilter::GpuFilter(std::unique_ptr<Operator> child, expr::Expr predicate)
    : child_(std::move(child)), predicate_(std::move(predicate)) {}

qe_dev_table* GpuFilter::run() {
    if (!predicate_ || predicate_.type() != Type::BOOL)  // ops/filter.cpp:18-20
        throw std::invalid_argument("Filter: predicate must be BOOL");
    const Schema s = child_->output_schema();
    FlatPred fp;
    flatten(predicate_, s, fp);
    if (fp.never) {  // no row passes: one NULL-literal term yields the empty result
        fp.terms.assign(1, qe_pred_term{});
        fp.terms[0].lit_type = static_cast<int32_t>(s.fields.at(0).second);
        fp.terms[0].op = QE_CMP_EQ;
        fp.terms[0].lit_is_null = 1;
    }
    if (fp.terms.empty()) return nullptr;
    in_ = device_input(*child_);
    qe_dev_table* t = nullptr;
    QE_CALL(qe_gpu_filter(device_context(), in_->cols.data(), static_cast<uint32_t>(in_->cols.size()),
                          fp.terms.data(), static_cast<uint32_t>(fp.terms.size()), &t));
    return t;
}

void GpuFilter::open() {
    qe_dev_table* t = run();
    pass_through_ = t == nullptr;
    if (pass_through_) child_->open();
    emit_.reset(t, types_of(output_schema()));
}

std::optional<Batch> GpuFilter::next() {
    if (!pass_through_) return emit_.next();
    while (auto b = child_->next())  // TRUE for every row: the child's non-empty batches
        if (b->row_count) return b;
    return std::nullopt;
}

void GpuFilter::close() {
    if (pass_through_) child_->close();
    pass_through_ = false;
    emit_.reset(nullptr, {});
    in_.reset();
}

std::shared_ptr<DeviceColumns> GpuFilter::device_columns() {
    qe_dev_table* t = run();
    if (!t) return device_input(*child_);
    auto keep = table_columns(t, {in_});
    in_.reset();
    return keep;
}

}  // namespace qe
