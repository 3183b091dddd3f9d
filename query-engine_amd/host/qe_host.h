// Host-side drop-in layer over the C ABI (include/qe_gpu.h).
//
// The columnar view types and the pull-based Operator contract are restated
// field for field from the reference (core/types.h:18-48, core/column.h:55-90,
// ops/operator.h:40-47) so a plan can lower to the device operators below in
// place of the reference's HashJoin / Aggregate / AsofJoin, with the same
// constructor shapes (ops/join.h:79-81, ops/aggregate.h:110-111,
// tsx/asof.h:101-105), the same output schemas and the same exception types
// (std::invalid_argument, std::overflow_error, std::bad_alloc).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <utility>
#include <vector>

#include "qe_gpu.h"

namespace qe {

// ---- columnar format (core/types.h, core/column.h) -------------------------------
enum class Type { I32, I64, F64, BOOL, TS, STR };

inline constexpr std::size_t byte_width(Type t) {
    switch (t) {
        case Type::I32:
        case Type::STR:
            return 4;
        case Type::BOOL:
            return 1;
        default:
            return 8;
    }
}

// The code -> bytes table behind a STR column (core/string_dict.h:34-72):
// intern() dedups so equal strings share one code within a dictionary; codes
// are assigned 0, 1, 2, ... in first-seen order. STR values compare by bytes,
// never by code.
class StringDict {
   public:
    StringDict() { offsets_.push_back(0); }
    int32_t intern(std::string_view s);
    std::string_view at(int32_t code) const {
        return std::string_view(bytes_.data() + offsets_[code], offsets_[code + 1] - offsets_[code]);
    }
    size_t size() const noexcept { return offsets_.size() - 1; }

   private:
    std::vector<char> bytes_;
    std::vector<uint32_t> offsets_;
    std::unordered_map<std::string, int32_t> index_;
};

struct Column {
    Type type;
    size_t len;
    const std::byte* data;
    const uint64_t* validity;  // nullptr => no NULLs (implies all_valid)
    bool all_valid;            // authoritative; validity ignored when true
    const StringDict* dict;
};

struct SelectionVector {
    const uint32_t* idx;
    size_t len;
};

struct Batch {
    std::vector<Column> cols;
    const SelectionVector* sel;  // nullptr => dense
    size_t row_count;
};

struct Schema {
    std::vector<std::pair<std::string, Type>> fields;
};

// ---- the pull contract (ops/operator.h:40-47) -------------------------------------------
class Operator {
   public:
    virtual void open() = 0;
    virtual std::optional<Batch> next() = 0;  // view valid until the next next()/close()
    virtual void close() = 0;
    virtual Schema output_schema() const = 0;
    virtual ~Operator() = default;
};

// ---- owning storage --------------------------------------------------------------------
class OwnedColumn {
   public:
    static OwnedColumn make(Type t, size_t len);
    // A STR column of `len` codes into `dict` (core/owned_batch.h:37-41).
    static OwnedColumn make_str(size_t len, std::shared_ptr<const StringDict> dict);
    void set_dict(std::shared_ptr<const StringDict> d) { dict_ = std::move(d); }
    const StringDict* dict() const noexcept { return dict_.get(); }
    const std::shared_ptr<const StringDict>& dict_shared() const noexcept { return dict_; }
    Type type() const { return type_; }
    size_t len() const { return len_; }
    std::byte* mutable_data() { return data_.data(); }
    const std::byte* data() const { return data_.data(); }
    void set_null(size_t i);
    bool all_valid() const { return all_valid_; }
    const uint64_t* validity() const { return all_valid_ ? nullptr : validity_.data(); }
    Column view() const;

   private:
    Type type_ = Type::I32;
    size_t len_ = 0;
    std::vector<std::byte> data_;
    std::vector<uint64_t> validity_;
    bool all_valid_ = true;
    std::shared_ptr<const StringDict> dict_;  // STR only
};

class OwnedBatch {
   public:
    void add_column(OwnedColumn c) { cols_.push_back(std::move(c)); }
    Batch view() const;
    size_t num_columns() const { return cols_.size(); }

   private:
    std::vector<OwnedColumn> cols_;
};

// In-memory table (ops/table.h:24-45): owns equal-length columns.
struct DeviceColumns;

class Table {
   public:
    Table(Schema schema, std::vector<OwnedColumn> cols);
    Table(const Table&) = delete;
    Table& operator=(const Table&) = delete;
    const Schema& schema() const { return schema_; }
    size_t num_rows() const { return rows_; }
    const OwnedColumn& column(size_t i) const { return cols_[i]; }

   private:
    friend std::shared_ptr<DeviceColumns> upload_table(const Table& t);
// The host dictionary a device STR column's codes index (uploads register it).
std::shared_ptr<const StringDict> host_dict_of(const qe_dev_col* c);
    friend void drop_device_copy(const Table& t);
    Schema schema_;
    std::vector<OwnedColumn> cols_;
    size_t rows_ = 0;
    // HBM copy of the columns, made on first device use and released with the
    // Table (or by drop_device_copy); operators holding it keep it alive.
    mutable std::shared_ptr<DeviceColumns> device_;
};

// ---- device plumbing -------------------------------------------------------------------
// Throws the reference exception type that matches a C-ABI status.
void throw_status(qe_gpu_status st, qe_gpu_ctx* ctx);

// One device context per process (device 0), created on first use.
qe_gpu_ctx* device_context();

// Device-resident copies of host columns; a Table's columns are uploaded once
// and kept with the Table.
struct DeviceColumns {
    std::vector<qe_dev_col*> cols;
    size_t rows = 0;
    bool owned = true;
    // Device copies of the STR columns' dictionaries, registered so that device
    // results map back to the host dictionary (released with these columns).
    std::vector<qe_gpu_dict*> dicts;
    // Inputs whose dictionaries these columns' STR codes still index.
    std::vector<std::shared_ptr<const DeviceColumns>> sources;
    // Indexes built over these columns and kept with them (e.g. a Table's
    // as-of index), by a signature of what was built.
    mutable std::map<std::string, std::shared_ptr<void>> derived;
    ~DeviceColumns();
};
std::shared_ptr<DeviceColumns> upload_table(const Table& t);
// The host dictionary a device STR column's codes index (uploads register it).
std::shared_ptr<const StringDict> host_dict_of(const qe_dev_col* c);
void drop_device_copy(const Table& t);

// Operators that can hand their whole output to a device parent without a
// host round trip.
class DeviceSource {
   public:
    virtual std::shared_ptr<DeviceColumns> device_columns() = 0;
    virtual ~DeviceSource() = default;
};

// Scan (ops/scan.h:37): zero-copy <=batch_size windows of a Table; batch_size
// must be a positive multiple of 64 (ops/scan.cpp:15-18). Also a DeviceSource.
class Scan : public Operator, public DeviceSource {
   public:
    static constexpr size_t kDefaultBatchSize = 2048;
    explicit Scan(const Table& t, size_t batch_size = kDefaultBatchSize);
    void open() override { cursor_ = 0; }
    std::optional<Batch> next() override;
    void close() override {}
    Schema output_schema() const override { return table_.schema(); }
    std::shared_ptr<DeviceColumns> device_columns() override { return upload_table(table_); }

   private:
    const Table& table_;
    size_t batch_size_;
    size_t cursor_ = 0;
};

// ---- aggregate vocabulary (ops/aggregate.h:61-94) -----------------------------------------
enum class AggFunc { CountStar, Count, Sum, Min, Max, Avg };
struct AggSpec {
    AggFunc func = AggFunc::CountStar;
    uint32_t input_col = 0;
    std::string out_name;
    static AggSpec count_star(std::string n) { return {AggFunc::CountStar, 0, std::move(n)}; }
    static AggSpec count(uint32_t c, std::string n) { return {AggFunc::Count, c, std::move(n)}; }
    static AggSpec sum(uint32_t c, std::string n) { return {AggFunc::Sum, c, std::move(n)}; }
    static AggSpec min(uint32_t c, std::string n) { return {AggFunc::Min, c, std::move(n)}; }
    static AggSpec max(uint32_t c, std::string n) { return {AggFunc::Max, c, std::move(n)}; }
    static AggSpec avg(uint32_t c, std::string n) { return {AggFunc::Avg, c, std::move(n)}; }
};
Type agg_result_type(AggFunc f, Type input);  // ops/aggregate.cpp:78-97

enum class JoinType { Inner, Left };

namespace tsx {
enum class AsofType { Inner, Left };
enum class AsofMethod { Binary, Learned };
}  // namespace tsx

// Kernel-path selectors of the reference's scalar==vector test seams
// (ops/hashtable.h:77, ops/join.h:67, ops/aggregate.h:100, ops/sort.h:83). The
// device has one kernel per operator, so the device operators accept and
// record them (the same results for every value) and tests written against
// the seams run unchanged.
enum class HashPath { kVector, kScalar };
enum class GatherPath { kVector, kScalar };
enum class AggKernelPath { kVector, kScalar };

// Device result drained into host batches of <= 2048 rows.
class ResultEmitter {
   public:
    // `sources`: the inputs whose dictionaries the table's STR columns index.
    void reset(qe_dev_table* t, std::vector<Type> types,
               std::vector<std::shared_ptr<const DeviceColumns>> sources = {});
    std::optional<Batch> next();
    ~ResultEmitter();

   private:
    qe_dev_table* table_ = nullptr;
    std::vector<Type> types_;
    std::vector<std::shared_ptr<const DeviceColumns>> sources_;
    size_t cursor_ = 0;
    OwnedBatch current_;
};

// ---- device operators -------------------------------------------------------------------
// Device result tables handed to a device parent as columns.
std::shared_ptr<DeviceColumns> table_columns(qe_dev_table* t, std::vector<std::shared_ptr<const DeviceColumns>> sources);

// Aggregate (ops/aggregate.h:110-111): GROUP BY key_cols with aggs; output =
// key columns ++ aggregate columns; zero keys -> exactly one row.
class GpuAggregate : public Operator, public DeviceSource {
   public:
    static constexpr size_t kOutBatch = 2048;
    GpuAggregate(std::unique_ptr<Operator> child, std::vector<uint32_t> key_cols, std::vector<AggSpec> aggs);
    // Test seam (ops/aggregate.h:116-119), call before open(): accepted; the
    // device path is the same for every value.
    void set_paths(HashPath hash_path, AggKernelPath kernel_path) {
        hash_path_ = hash_path;
        kernel_path_ = kernel_path;
    }
    void open() override;
    std::optional<Batch> next() override { return emit_.next(); }
    void close() override;
    Schema output_schema() const override;
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    qe_dev_table* run(std::shared_ptr<DeviceColumns>& in);
    std::unique_ptr<Operator> child_;
    std::vector<uint32_t> key_cols_;
    std::vector<AggSpec> aggs_;
    Schema child_schema_;
    ResultEmitter emit_;
    HashPath hash_path_ = HashPath::kVector;
    AggKernelPath kernel_path_ = AggKernelPath::kVector;
};

// HashJoin (ops/join.h:79-81): output = probe columns ++ build columns.
class GpuHashJoin : public Operator, public DeviceSource {
   public:
    static constexpr size_t kOutBatch = 2048;
    GpuHashJoin(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys, JoinType type);
    // Test seam (ops/join.h:86-89), call before open(): accepted; the device path is
    // the same for every value.
    void set_paths(HashPath hash_path, GatherPath gather_path) {
        hash_path_ = hash_path;
        gather_path_ = gather_path;
    }
    // A learned probe of at least min_rows rows over at least min_line_bytes
    // of router bucket lines first groups its rows by the index region they
    // read (qe_gpu_asof_cluster). Defaults: 2^22 rows, 2 GiB. Process-wide,
    // read at open(); no counterpart in the reference.
    static void set_cluster_threshold(size_t min_rows, uint64_t min_line_bytes);
    void open() override;
    std::optional<Batch> next() override;
    void close() override;
    Schema output_schema() const override;
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    std::unique_ptr<Operator> probe_, build_;
    std::vector<uint32_t> probe_keys_, build_keys_;
    JoinType type_;
    Schema probe_schema_, build_schema_;
    std::shared_ptr<DeviceColumns> pcols_, bcols_;
    qe_dev_pairs* pairs_ = nullptr;
    size_t cursor_ = 0;
    ResultEmitter chunk_;
    HashPath hash_path_ = HashPath::kVector;
    GatherPath gather_path_ = GatherPath::kVector;
};

// Aggregate(HashJoin(probe, build)) fused on the device (one streaming probe).
class GpuJoinAggregate : public Operator, public DeviceSource {
   public:
    GpuJoinAggregate(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                     std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys, JoinType type,
                     std::vector<uint32_t> group_keys, std::vector<AggSpec> aggs);
    void open() override;
    std::optional<Batch> next() override { return emit_.next(); }
    void close() override;
    Schema output_schema() const override;
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    qe_dev_table* run(std::shared_ptr<DeviceColumns>& pc, std::shared_ptr<DeviceColumns>& bc);
    std::unique_ptr<Operator> probe_, build_;
    std::vector<uint32_t> probe_keys_, build_keys_;
    JoinType type_;
    std::vector<uint32_t> group_keys_;
    std::vector<AggSpec> aggs_;
    Schema join_schema_;
    ResultEmitter emit_;
};

// tsx::AsofJoin (tsx/asof.h:101-105): backward as-of; output = probe columns ++
// build columns; the method selects binary search or the learned index.
class GpuAsofJoin : public Operator, public DeviceSource {
   public:
    static constexpr size_t kOutBatch = 2048;
    GpuAsofJoin(std::unique_ptr<Operator> probe, std::unique_ptr<Operator> build,
                std::vector<uint32_t> probe_keys, std::vector<uint32_t> build_keys, uint32_t probe_time,
                uint32_t build_time, tsx::AsofType type, std::optional<int64_t> tolerance = std::nullopt,
                tsx::AsofMethod method = tsx::AsofMethod::Learned);
    // Test seam (tsx/asof.h:110-113), call before open(): accepted; the device path is
    // the same for every value.
    void set_paths(HashPath hash_path, GatherPath gather_path) {
        hash_path_ = hash_path;
        gather_path_ = gather_path;
    }
    // A learned probe of at least min_rows rows over at least min_line_bytes
    // of router bucket lines first groups its rows by the index region they
    // read (qe_gpu_asof_cluster). Defaults: 2^22 rows, 2 GiB. Process-wide,
    // read at open(); no counterpart in the reference.
    static void set_cluster_threshold(size_t min_rows, uint64_t min_line_bytes);
    void open() override;
    std::optional<Batch> next() override;
    void close() override;
    Schema output_schema() const override;
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    std::unique_ptr<Operator> probe_, build_;
    std::vector<uint32_t> probe_keys_, build_keys_;
    uint32_t probe_time_, build_time_;
    tsx::AsofType type_;
    std::optional<int64_t> tolerance_;
    tsx::AsofMethod method_;
    Schema probe_schema_, build_schema_;
    std::shared_ptr<DeviceColumns> pcols_, bcols_;
    qe_dev_pairs* pairs_ = nullptr;
    size_t cursor_ = 0;
    ResultEmitter chunk_;
    HashPath hash_path_ = HashPath::kVector;
    GatherPath gather_path_ = GatherPath::kVector;
};

// ---- sort vocabulary (ops/sort.h:33-42) ---------------------------------------------------
enum class SortDir { Asc, Desc };
enum class NullOrder { First, Last };
struct SortKey {
    uint32_t col = 0;
    SortDir dir = SortDir::Asc;
    NullOrder nulls = NullOrder::Last;
};

// Sort (ops/sort.h:84): a stable sort of the child's rows by `keys` (ORDER BY
// precedence), every column gathered by the permutation; NULL placement is
// absolute. Integer-family keys, F64 (NaN greatest, -0 == +0) and STR (by
// value) on the device. Output batches of <= 2048 rows.
class GpuSort : public Operator, public DeviceSource {
   public:
    static constexpr size_t kOutBatch = 2048;
    GpuSort(std::unique_ptr<Operator> child, std::vector<SortKey> keys);
    // Test seam (ops/sort.h:90), call before open(): accepted; one device path.
    void set_gather_path(GatherPath g) { gather_path_ = g; }
    void open() override;
    std::optional<Batch> next() override { return emit_.next(); }
    void close() override;
    Schema output_schema() const override { return child_schema_; }
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    qe_dev_table* run(std::shared_ptr<DeviceColumns>& in);
    std::unique_ptr<Operator> child_;
    std::vector<SortKey> keys_;
    Schema child_schema_;
    ResultEmitter emit_;
    GatherPath gather_path_ = GatherPath::kVector;
};

// Device HashTable with the reference's API (ops/hashtable.h:108-190).
enum class NullPolicy { kEqual, kNeverMatch };
inline constexpr uint32_t kNoGroup = UINT32_MAX;
struct HashTableConfig {
    size_t initial_capacity = 1024;
    double max_load_factor = 0.7;
};
struct KeyColumns {
    const Column* cols;
    size_t num_cols;
    const SelectionVector* sel;
};
class GpuHashTable {
   public:
    explicit GpuHashTable(std::vector<Type> key_types, NullPolicy p = NullPolicy::kEqual,
                          HashTableConfig cfg = {}, uint64_t seed = 0x51E57A11AB1E5EEDull);
    ~GpuHashTable();
    // `path` is the reference's scalar/vector seam (ops/hashtable.h:125-132):
    // accepted, one device path.
    void insert_or_find(const KeyColumns& keys, size_t n, uint32_t* out_groups,
                        HashPath path = HashPath::kVector);
    void find(const KeyColumns& keys, size_t n, uint32_t* out_groups, HashPath path = HashPath::kVector) const;
    size_t num_groups() const;
    bool group_is_null(uint32_t g, size_t col) const;
    uint64_t group_key_word(uint32_t g, size_t col) const;

   private:
    void run(const KeyColumns& keys, size_t n, uint32_t* out, bool insert) const;
    std::vector<Type> types_;
    qe_dev_ht* ht_ = nullptr;
};

// Every device operator can drain any child to device columns.
std::shared_ptr<DeviceColumns> device_input(Operator& op);

}  // namespace qe
