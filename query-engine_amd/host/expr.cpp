// Expression builders for device predicates. See qe_expr.h.
#include <string>

#include "qe_expr.h"

namespace qe::expr {

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

bool is_numeric(Type t) { return t == Type::I32 || t == Type::I64 || t == Type::F64; }

// Promotion rank F64 > I64 > I32 (expr/expr.cpp:21-39).
int rank(Type t) { return t == Type::F64 ? 2 : (t == Type::I64 ? 1 : 0); }

std::shared_ptr<Node> make(NodeKind k, Type t) {
    auto n = std::make_shared<Node>();
    n->kind = k;
    n->type = t;
    return n;
}

void need(const Expr& e, const char* who) {
    if (!e) throw std::invalid_argument(std::string(who) + ": empty operand");
}

}  // namespace

Expr col(Type type, std::uint32_t index) {
    auto n = make(NodeKind::Column, type);
    n->col_index = index;
    return Expr(std::move(n));
}

Expr col(const Schema& schema, std::string_view name) {
    for (std::uint32_t i = 0; i < schema.fields.size(); ++i)
        if (schema.fields[i].first == name) return col(schema.fields[i].second, i);
    throw std::invalid_argument("col: unknown column name '" + std::string(name) + "'");
}

Expr lit(Scalar s) {
    if (!s.is_null && s.type == Type::I32 && (s.i < INT32_MIN || s.i > INT32_MAX))
        throw std::invalid_argument("lit: I32 payload out of range");
    if (!s.is_null && s.type == Type::BOOL && s.i != 0 && s.i != 1)
        throw std::invalid_argument("lit: BOOL payload must be 0 or 1");
    auto n = make(NodeKind::Literal, s.type);
    n->literal = s;
    return Expr(std::move(n));
}

// expr/expr.cpp:133-151: numerics promote; otherwise the types must match.
Expr cmp(CmpOp op, Expr a, Expr b) {
    need(a, "cmp");
    need(b, "cmp");
    Type common;
    if (is_numeric(a.type()) && is_numeric(b.type())) {
        common = rank(a.type()) >= rank(b.type()) ? a.type() : b.type();
    } else if (a.type() == b.type()) {
        common = a.type();
    } else {
        throw std::invalid_argument(std::string("cmp: incomparable operand types ") + type_name(a.type()) +
                                    " and " + type_name(b.type()));
    }
    auto n = make(NodeKind::Cmp, Type::BOOL);
    n->cmp_op = op;
    n->cmp_type = common;
    n->children = {std::move(a), std::move(b)};
    return Expr(std::move(n));
}

Expr lt(Expr a, Expr b) { return cmp(CmpOp::Lt, std::move(a), std::move(b)); }
Expr le(Expr a, Expr b) { return cmp(CmpOp::Le, std::move(a), std::move(b)); }
Expr gt(Expr a, Expr b) { return cmp(CmpOp::Gt, std::move(a), std::move(b)); }
Expr ge(Expr a, Expr b) { return cmp(CmpOp::Ge, std::move(a), std::move(b)); }
Expr eq(Expr a, Expr b) { return cmp(CmpOp::Eq, std::move(a), std::move(b)); }
Expr ne(Expr a, Expr b) { return cmp(CmpOp::Ne, std::move(a), std::move(b)); }

static Expr logic(LogicOp op, std::vector<Expr> ch) {
    for (const Expr& e : ch) {
        need(e, "logic");
        if (e.type() != Type::BOOL)
            throw std::invalid_argument(std::string("logic: operand must be BOOL, got ") + type_name(e.type()));
    }
    auto n = make(NodeKind::Logic, Type::BOOL);
    n->logic_op = op;
    n->children = std::move(ch);
    return Expr(std::move(n));
}

Expr logic_and(Expr a, Expr b) { return logic(LogicOp::And, {std::move(a), std::move(b)}); }
Expr logic_or(Expr a, Expr b) { return logic(LogicOp::Or, {std::move(a), std::move(b)}); }
Expr logic_not(Expr a) { return logic(LogicOp::Not, {std::move(a)}); }

}  // namespace qe::expr
