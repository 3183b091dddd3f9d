// Plan builder + device lowering. See qe_plan.h.
#include "qe_plan.h"

#include <sstream>

namespace qe::plan {

namespace {

const char* type_name(Type t) {
    switch (t) {
        case Type::I32:
            return "I32";
        case Type::I64:
            return "I64";
        case Type::F64:
            return "F64";
        case Type::BOOL:
            return "BOOL";
        case Type::TS:
            return "TS";
        case Type::STR:
            return "STR";
    }
    return "?";
}

std::shared_ptr<PlanNode> make(PlanKind k) {
    auto n = std::make_shared<PlanNode>();
    n->kind = k;
    return n;
}

Schema concat(const Schema& a, const Schema& b) {
    Schema s = a;
    for (const auto& f : b.fields) s.fields.push_back(f);
    return s;
}

}  // namespace

uint32_t ColRef::resolve(const Schema& s) const {
    if (!by_name) {
        if (index >= s.fields.size())
            throw std::invalid_argument("plan: column index " + std::to_string(index) + " out of range");
        return index;
    }
    for (uint32_t i = 0; i < s.fields.size(); ++i)
        if (s.fields[i].first == name) return i;
    throw std::invalid_argument("plan: unknown column name '" + name + "'");
}

PlanBuilder scan(const Table& table) {
    auto n = make(PlanKind::Scan);
    n->table = &table;
    n->out_schema = table.schema();
    return PlanBuilder(Plan(std::move(n)));
}

// plan/plan.cpp:38-54: every column reference in range and of the schema's type.
static void validate_expr(const expr::Expr& e, const Schema& s) {
    const expr::Node& n = e.node();
    if (n.kind == expr::NodeKind::Column) {
        if (n.col_index >= s.fields.size())
            throw std::invalid_argument("plan: expression column index " + std::to_string(n.col_index) +
                                        " out of range (schema has " + std::to_string(s.fields.size()) +
                                        " columns)");
        if (s.fields[n.col_index].second != n.type)
            throw std::invalid_argument("plan: expression column " + std::to_string(n.col_index) +
                                        " type mismatch (expr says " + type_name(n.type) + ", schema says " +
                                        type_name(s.fields[n.col_index].second) + ")");
    }
    for (const expr::Expr& c : n.children) validate_expr(c, s);
}

// plan/plan.cpp:307-321: a non-empty BOOL predicate; the schema is unchanged.
PlanBuilder PlanBuilder::filter(expr::Expr predicate) const {
    const Schema& cs = plan_.output_schema();
    if (!predicate) throw std::invalid_argument("plan: filter predicate is empty");
    if (predicate.type() != Type::BOOL)
        throw std::invalid_argument(std::string("plan: filter predicate must be BOOL, got ") +
                                    type_name(predicate.type()));
    validate_expr(predicate, cs);
    auto n = make(PlanKind::Filter);
    n->children.push_back(plan_);
    n->predicate = std::move(predicate);
    n->out_schema = cs;
    return PlanBuilder(Plan(std::move(n)));
}

// plan/plan.cpp:548-566: >=1 key, each in range; the schema is unchanged.
PlanBuilder PlanBuilder::sort(std::vector<SortKey> keys) const {
    const Schema& cs = plan_.output_schema();
    if (keys.empty()) throw std::invalid_argument("plan: sort needs >=1 key");
    for (const SortKey& k : keys)
        if (k.col >= cs.fields.size())
            throw std::invalid_argument("plan: sort key index " + std::to_string(k.col) + " out of range");
    auto n = make(PlanKind::Sort);
    n->children.push_back(plan_);
    n->sort_keys = std::move(keys);
    n->out_schema = cs;
    return PlanBuilder(Plan(std::move(n)));
}

PlanBuilder PlanBuilder::sort(std::vector<SortBy> keys) const {
    const Schema& cs = plan_.output_schema();
    std::vector<SortKey> resolved;
    for (const SortBy& k : keys) resolved.push_back(SortKey{k.col.resolve(cs), k.dir, k.nulls});
    return sort(std::move(resolved));
}

PlanBuilder PlanBuilder::aggregate(std::vector<ColRef> keys, std::vector<AggSpec> aggs) const {
    const Schema& cs = plan_.output_schema();
    if (aggs.empty()) throw std::invalid_argument("plan: aggregate needs >=1 aggregate");
    auto n = make(PlanKind::Aggregate);
    for (const ColRef& k : keys) {
        const uint32_t i = k.resolve(cs);
        n->group_keys.push_back(i);
        n->out_schema.fields.push_back(cs.fields[i]);
    }
    for (const AggSpec& a : aggs) {
        Type in = Type::I64;
        if (a.func != AggFunc::CountStar) {
            if (a.input_col >= cs.fields.size())
                throw std::invalid_argument("plan: aggregate input column out of range");
            in = cs.fields[a.input_col].second;
        }
        n->out_schema.fields.emplace_back(a.out_name, agg_result_type(a.func, in));
    }
    n->aggs = std::move(aggs);
    n->children.push_back(plan_);
    return PlanBuilder(Plan(std::move(n)));
}

PlanBuilder PlanBuilder::join(const Plan& build, std::vector<ColRef> left_keys, std::vector<ColRef> right_keys,
                              JoinType type) const {
    const Schema& ls = plan_.output_schema();
    const Schema& rs = build.output_schema();
    if (left_keys.size() != right_keys.size() || left_keys.empty())
        throw std::invalid_argument("plan: join needs an equal, non-empty number of left/right keys");
    auto n = make(PlanKind::Join);
    for (size_t i = 0; i < left_keys.size(); ++i) {
        const uint32_t li = left_keys[i].resolve(ls), ri = right_keys[i].resolve(rs);
        if (ls.fields[li].second != rs.fields[ri].second)
            throw std::invalid_argument("plan: join key " + std::to_string(i) + " type mismatch (left " +
                                        type_name(ls.fields[li].second) + " vs right " +
                                        type_name(rs.fields[ri].second) + ")");
        n->left_keys.push_back(li);
        n->right_keys.push_back(ri);
    }
    n->join_type = type;
    n->out_schema = concat(ls, rs);
    n->children = {plan_, build};
    return PlanBuilder(Plan(std::move(n)));
}

PlanBuilder PlanBuilder::asof_join(const Plan& build, std::vector<ColRef> left_keys, std::vector<ColRef> right_keys,
                                   ColRef left_time, ColRef right_time, tsx::AsofType type,
                                   std::optional<int64_t> tolerance) const {
    const Schema& ls = plan_.output_schema();
    const Schema& rs = build.output_schema();
    if (left_keys.size() != right_keys.size())
        throw std::invalid_argument("plan: asof_join needs an equal number of left/right keys");
    auto n = make(PlanKind::AsofJoin);
    for (size_t i = 0; i < left_keys.size(); ++i) {
        const uint32_t li = left_keys[i].resolve(ls), ri = right_keys[i].resolve(rs);
        if (ls.fields[li].second != rs.fields[ri].second)
            throw std::invalid_argument("plan: asof_join key " + std::to_string(i) + " type mismatch");
        n->left_keys.push_back(li);
        n->right_keys.push_back(ri);
    }
    n->left_time = left_time.resolve(ls);
    n->right_time = right_time.resolve(rs);
    auto is_time = [](Type t) { return t == Type::TS || t == Type::I32 || t == Type::I64; };
    if (!is_time(ls.fields[n->left_time].second) || !is_time(rs.fields[n->right_time].second))
        throw std::invalid_argument("plan: asof_join timestamp columns must be TS/I32/I64");
    n->asof_type = type;
    n->tolerance = tolerance;
    n->out_schema = concat(ls, rs);
    n->children = {plan_, build};
    return PlanBuilder(Plan(std::move(n)));
}

std::unique_ptr<Operator> Plan::lower(const LowerOptions& opt) const {
    const PlanNode& n = *node_;
    switch (n.kind) {
        case PlanKind::Scan:
            return std::make_unique<Scan>(*n.table, opt.batch_size);
        case PlanKind::Filter:
            return std::make_unique<GpuFilter>(n.children[0].lower(opt), n.predicate);
        case PlanKind::Sort:
            return std::make_unique<GpuSort>(n.children[0].lower(opt), n.sort_keys);
        case PlanKind::Aggregate: {
            const Plan& c = n.children[0];
            if (opt.fuse_join_aggregate && c.kind() == PlanKind::Join) {
                const PlanNode& j = c.node();
                return std::make_unique<GpuJoinAggregate>(j.children[0].lower(opt), j.children[1].lower(opt),
                                                          j.left_keys, j.right_keys, j.join_type, n.group_keys,
                                                          n.aggs);
            }
            return std::make_unique<GpuAggregate>(c.lower(opt), n.group_keys, n.aggs);
        }
        case PlanKind::Join:
            return std::make_unique<GpuHashJoin>(n.children[0].lower(opt), n.children[1].lower(opt), n.left_keys,
                                                 n.right_keys, n.join_type);
        case PlanKind::AsofJoin:
            return std::make_unique<GpuAsofJoin>(n.children[0].lower(opt), n.children[1].lower(opt), n.left_keys,
                                                 n.right_keys, n.left_time, n.right_time, n.asof_type, n.tolerance,
                                                 opt.asof_method);
    }
    throw std::logic_error("plan: unhandled node kind");
}

namespace {
void print(const Plan& p, int depth, std::ostringstream& os) {
    const PlanNode& n = p.node();
    os << std::string(depth * 2, ' ');
    switch (n.kind) {
        case PlanKind::Scan:
            os << "Scan(" << n.out_schema.fields.size() << " cols)";
            break;
        case PlanKind::Filter:
            os << "Filter";
            break;
        case PlanKind::Sort:
            os << "Sort(keys:" << n.sort_keys.size() << ")";
            break;
        case PlanKind::Aggregate:
            os << "Aggregate(keys:" << n.group_keys.size() << " aggs:" << n.aggs.size() << ")";
            break;
        case PlanKind::Join:
            os << "Join(" << (n.join_type == JoinType::Inner ? "INNER" : "LEFT") << " keys:" << n.left_keys.size()
               << ")";
            break;
        case PlanKind::AsofJoin:
            os << "AsofJoin(" << (n.asof_type == tsx::AsofType::Inner ? "INNER" : "LEFT")
               << " keys:" << n.left_keys.size() << ")";
            break;
    }
    os << "\n";
    for (const Plan& c : n.children) print(c, depth + 1, os);
}
}  // namespace

std::string Plan::to_string() const {
    std::ostringstream os;
    print(*this, 0, os);
    return os.str();
}

}  // namespace qe::plan
