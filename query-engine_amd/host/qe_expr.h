// Expression trees for device predicates: the subset of the reference's
// expression IR (expr/expr.h:79-180) that a Filter hands to the device —
// column references, typed literals, comparisons and three-valued logic —
// with the same builders, the same build-time type inference and the same
// std::invalid_argument on a type error (expr/expr.cpp:100-151).
//
// Numeric promotion is implicit here: cmp() records the promoted type
// (F64 > I64 > I32) instead of inserting cast nodes, and the device compares
// in that type. BOOL compares BOOL, TS compares TS, STR compares STR.
#pragma once

#include <cstdint>
#include <memory>
#include <string_view>
#include <vector>

#include "qe_host.h"

namespace qe::expr {

enum class CmpOp { Lt, Le, Gt, Ge, Eq, Ne };   // expr/expr.h:79
enum class LogicOp { And, Or, Not };            // expr/expr.h:80

// One typed literal (expr/expr.h:84-104): integer-family payloads in `i`
// (I32 in int32 range, BOOL as 0/1, TS as int64 ns), F64 in `f`.
struct Scalar {
    Type type = Type::I32;
    bool is_null = false;
    std::int64_t i = 0;
    double f = 0.0;
    static Scalar i32(std::int32_t v) { return {Type::I32, false, v, 0.0}; }
    static Scalar i64(std::int64_t v) { return {Type::I64, false, v, 0.0}; }
    static Scalar f64(double v) { return {Type::F64, false, 0, v}; }
    static Scalar boolean(bool v) { return {Type::BOOL, false, v ? 1 : 0, 0.0}; }
    static Scalar ts(std::int64_t ns) { return {Type::TS, false, ns, 0.0}; }
    static Scalar null(Type t) { return {t, true, 0, 0.0}; }
};

struct Node;

class Expr {
   public:
    Expr() = default;
    explicit Expr(std::shared_ptr<const Node> n) : node_(std::move(n)) {}
    Type type() const;
    const Node& node() const { return *node_; }
    explicit operator bool() const { return static_cast<bool>(node_); }

   private:
    std::shared_ptr<const Node> node_;
};

enum class NodeKind { Column, Literal, Cmp, Logic };

struct Node {
    NodeKind kind;
    Type type;                    // inferred result type
    std::uint32_t col_index = 0;  // Column
    Scalar literal{};             // Literal
    CmpOp cmp_op{};               // Cmp
    Type cmp_type = Type::I32;    // Cmp: the promoted operand type
    LogicOp logic_op{};           // Logic
    std::vector<Expr> children;
};

inline Type Expr::type() const { return node_->type; }

Expr col(Type type, std::uint32_t index);
Expr col(const Schema& schema, std::string_view name);
Expr lit(Scalar s);

Expr cmp(CmpOp op, Expr a, Expr b);
Expr lt(Expr a, Expr b);
Expr le(Expr a, Expr b);
Expr gt(Expr a, Expr b);
Expr ge(Expr a, Expr b);
Expr eq(Expr a, Expr b);
Expr ne(Expr a, Expr b);

Expr logic_and(Expr a, Expr b);
Expr logic_or(Expr a, Expr b);
Expr logic_not(Expr a);

}  // namespace qe::expr

namespace qe {

// Filter (ops/filter.h; ops/filter.cpp:15-66): rows whose predicate is TRUE, in
// input order, every column compacted with its validity; a BOOL predicate is
// checked at open(). On the device the predicate must be a conjunction
// (logic_and) of column-versus-literal comparisons, BOOL columns and BOOL
// literals; other shapes (OR, NOT, column-versus-column, STR) throw
// std::runtime_error at open(). Output batches hold <= 2048 rows and are never
// empty. Also a DeviceSource, so Aggregate / HashJoin above it stay on the
// device.
class GpuFilter : public Operator, public DeviceSource {
   public:
    static constexpr size_t kOutBatch = 2048;
    GpuFilter(std::unique_ptr<Operator> child, expr::Expr predicate);
    void open() override;
    std::optional<Batch> next() override;
    void close() override;
    Schema output_schema() const override { return child_->output_schema(); }
    std::shared_ptr<DeviceColumns> device_columns() override;

   private:
    // Runs the device filter over the child's columns; nullptr when the
    // predicate is TRUE for every row (then the child's rows pass as they are).
    qe_dev_table* run();
    std::unique_ptr<Operator> child_;
    expr::Expr predicate_;
    bool pass_through_ = false;
    std::shared_ptr<DeviceColumns> in_;  // the filtered input (its dictionaries back STR results)
    ResultEmitter emit_;
};

}  // namespace qe
