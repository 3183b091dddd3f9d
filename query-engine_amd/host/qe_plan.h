// Dataframe plan builder for the device path: the scan / filter / join /
// aggregate / asof_join / sort subset of the reference PlanBuilder
// (plan/plan.h:205-227), with
// the same validation (key lists resolved by index or name, positional key
// type equality, time columns I32/I64/TS) and the same output schemas, and a
// lowering that produces the device operators (plan/plan.cpp:570-615 lowers
// the same nodes to the CPU operators). Aggregate over Join lowers to the
// fused GpuJoinAggregate.
#pragma once

#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "qe_expr.h"
#include "qe_host.h"

namespace qe::plan {

enum class PlanKind { Scan, Filter, Aggregate, Join, AsofJoin, Sort };

struct ColRef {
    bool by_name = false;
    uint32_t index = 0;
    std::string name;
    ColRef(int i) : index(static_cast<uint32_t>(i)) {}
    ColRef(unsigned i) : index(i) {}
    ColRef(const char* n) : by_name(true), name(n) {}
    ColRef(std::string n) : by_name(true), name(std::move(n)) {}
    uint32_t resolve(const Schema& s) const;
};

// A sort key by name or index (plan/plan.h:101-105).
struct SortBy {
    ColRef col;
    SortDir dir = SortDir::Asc;
    NullOrder nulls = NullOrder::Last;
};

class Plan;
struct PlanNode {
    PlanKind kind;
    Schema out_schema;
    std::vector<Plan> children;
    const Table* table = nullptr;
    expr::Expr predicate;  // Filter (BOOL)
    std::vector<uint32_t> group_keys;
    std::vector<AggSpec> aggs;
    std::vector<uint32_t> left_keys, right_keys;
    JoinType join_type = JoinType::Inner;
    uint32_t left_time = 0, right_time = 0;
    tsx::AsofType asof_type = tsx::AsofType::Inner;
    std::optional<int64_t> tolerance;
    std::vector<SortKey> sort_keys;  // Sort
};

struct LowerOptions {
    size_t batch_size = Scan::kDefaultBatchSize;
    bool fuse_join_aggregate = true;
    tsx::AsofMethod asof_method = tsx::AsofMethod::Learned;
};

class Plan {
   public:
    Plan() = default;
    explicit Plan(std::shared_ptr<const PlanNode> n) : node_(std::move(n)) {}
    PlanKind kind() const { return node_->kind; }
    const Schema& output_schema() const { return node_->out_schema; }
    const PlanNode& node() const { return *node_; }
    std::unique_ptr<Operator> lower(const LowerOptions& opt = {}) const;
    std::string to_string() const;

   private:
    std::shared_ptr<const PlanNode> node_;
};

class PlanBuilder {
   public:
    explicit PlanBuilder(Plan p) : plan_(std::move(p)) {}
    PlanBuilder filter(expr::Expr predicate) const;
    PlanBuilder sort(std::vector<SortKey> keys) const;  // index path
    PlanBuilder sort(std::vector<SortBy> keys) const;   // name-friendly path
    PlanBuilder aggregate(std::vector<ColRef> keys, std::vector<AggSpec> aggs) const;
    PlanBuilder join(const Plan& build, std::vector<ColRef> left_keys, std::vector<ColRef> right_keys,
                     JoinType type) const;
    PlanBuilder asof_join(const Plan& build, std::vector<ColRef> left_keys, std::vector<ColRef> right_keys,
                          ColRef left_time, ColRef right_time, tsx::AsofType type,
                          std::optional<int64_t> tolerance = std::nullopt) const;
    const Plan& plan() const { return plan_; }
    operator const Plan&() const { return plan_; }
    const Schema& schema() const { return plan_.output_schema(); }
    std::unique_ptr<Operator> build(const LowerOptions& opt = {}) const { return plan_.lower(opt); }

   private:
    Plan plan_;
};

PlanBuilder scan(const Table& table);

}  // namespace qe::plan
