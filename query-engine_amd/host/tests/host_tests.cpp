// Drop-in tests of the host operator layer: the device operators behind the
// reference's Operator contract and plan builder, driven the way the
// reference's own suites drive HashJoin / Aggregate / AsofJoin / HashTable
// (tests/join_differential_test.cpp, tests/aggregate_differential_test.cpp,
// tests/asof_differential_test.cpp, tests/hashtable_test.cpp, the Filter cases
// of tests/ops_test.cpp), with
// hand-derived expected results. Needs a GPU; run from tests/test_gpu_host.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "qe_expr.h"
#include "qe_host.h"
#include "qe_plan.h"

using namespace qe;

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        if (cond) {                                                            \
            ++g_pass;                                                          \
        } else {                                                               \
            ++g_fail;                                                          \
            std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
        }                                                                      \
    } while (0)

template <typename T>
OwnedColumn mk(Type t, const std::vector<T>& v, std::vector<size_t> nulls = {}) {
    OwnedColumn c = OwnedColumn::make(t, v.size());
    for (size_t i = 0; i < v.size(); ++i) std::memcpy(c.mutable_data() + i * byte_width(t), &v[i], sizeof(T));
    for (size_t i : nulls) c.set_null(i);
    return c;
}

// Drain an operator into rows of optional doubles (integers widened exactly for
// the small values used here).
using Row = std::vector<std::optional<double>>;
std::vector<Row> drain(Operator& op, size_t* batches = nullptr, bool sorted = true) {
    std::vector<Row> rows;
    size_t nb = 0;
    op.open();
    while (auto b = op.next()) {
        ++nb;
        if (b->row_count > 2048) ++g_fail;  // <= 2048-row output batches
        for (size_t k = 0; k < b->row_count; ++k) {
            Row r;
            for (const Column& c : b->cols) {
                const bool valid = c.all_valid || ((c.validity[k >> 6] >> (k & 63)) & 1ull);
                if (!valid) {
                    r.push_back(std::nullopt);
                    continue;
                }
                double v = 0;
                switch (c.type) {
                    case Type::I32:
                    case Type::STR:
                        v = reinterpret_cast<const int32_t*>(c.data)[k];
                        break;
                    case Type::BOOL:
                        v = reinterpret_cast<const uint8_t*>(c.data)[k];
                        break;
                    case Type::F64:
                        v = reinterpret_cast<const double*>(c.data)[k];
                        break;
                    default:
                        v = static_cast<double>(reinterpret_cast<const int64_t*>(c.data)[k]);
                }
                r.push_back(v);
            }
            rows.push_back(r);
        }
    }
    op.close();
    if (batches) *batches = nb;
    if (sorted) std::sort(rows.begin(), rows.end());
    return rows;
}

Schema sch(std::vector<std::pair<std::string, Type>> f) { return Schema{std::move(f)}; }

void test_join_groupby_plan() {
    // build: k {1,2,3,4} -> g {10,10,20,20}; probe: k {1,1,2,3,5} pay {1,2,3,4,5}
    Table build(sch({{"k", Type::I64}, {"g", Type::I32}}),
                [] {
                    std::vector<OwnedColumn> c;
                    c.push_back(mk<int64_t>(Type::I64, {1, 2, 3, 4}));
                    c.push_back(mk<int32_t>(Type::I32, {10, 10, 20, 20}));
                    return c;
                }());
    Table probe(sch({{"k", Type::I64}, {"pay", Type::F64}}),
                [] {
                    std::vector<OwnedColumn> c;
                    c.push_back(mk<int64_t>(Type::I64, {1, 1, 2, 3, 5}));
                    c.push_back(mk<double>(Type::F64, {1.0, 2.0, 3.0, 4.0, 5.0}));
                    return c;
                }());
    for (bool fuse : {true, false}) {
        plan::LowerOptions opt;
        opt.fuse_join_aggregate = fuse;
        auto op = plan::scan(probe)
                      .join(plan::scan(build), {"k"}, {"k"}, JoinType::Inner)
                      .aggregate({"g"}, {AggSpec::sum(1, "s"), AggSpec::count_star("n")})
                      .build(opt);
        const auto rows = drain(*op);
        // g=10: pays 1,2,3 -> 6, 3 rows; g=20: pay 4 -> 4, 1 row.
        CHECK(rows.size() == 2);
        CHECK(rows.size() == 2 && rows[0] == (Row{10.0, 6.0, 3.0}) && rows[1] == (Row{20.0, 4.0, 1.0}));
    }
}

void test_join_fanout_crosses_output_batches() {
    std::vector<int32_t> bk(3000, 7);
    std::vector<int64_t> bv(3000);
    for (int i = 0; i < 3000; ++i) bv[i] = i;
    Table build(sch({{"c0", Type::I32}, {"c1", Type::I64}}), [&] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, bk));
        c.push_back(mk<int64_t>(Type::I64, bv));
        return c;
    }());
    Table probe(sch({{"c0", Type::I32}, {"c1", Type::I64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {7, 9, 7}));
        c.push_back(mk<int64_t>(Type::I64, {100, 200, 300}));
        return c;
    }());
    for (JoinType jt : {JoinType::Inner, JoinType::Left}) {
        GpuHashJoin j(std::make_unique<Scan>(probe), std::make_unique<Scan>(build), {0}, {0}, jt);
        size_t nb = 0;
        const auto rows = drain(j, &nb);
        CHECK(rows.size() == (jt == JoinType::Inner ? 6000u : 6001u));
        CHECK(nb >= 3);
        CHECK(j.output_schema().fields.size() == 4);
        if (jt == JoinType::Left) {
            // The unmatched probe row (9) carries NULL build columns.
            const auto it = std::find_if(rows.begin(), rows.end(), [](const Row& r) { return r[0] == 9.0; });
            CHECK(it != rows.end() && !(*it)[2].has_value() && !(*it)[3].has_value());
        }
    }
}

void test_aggregate_edges() {
    Table t(sch({{"c0", Type::I32}, {"c1", Type::I32}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 1, 2, 2}));
        c.push_back(mk<int32_t>(Type::I32, {0, 0, 5, 6}, {0, 1}));
        return c;
    }());
    GpuAggregate a(std::make_unique<Scan>(t), {0},
                   {AggSpec::count_star("cs"), AggSpec::count(1, "c"), AggSpec::sum(1, "s"), AggSpec::avg(1, "a")});
    const auto rows = drain(a);
    CHECK(rows.size() == 2);
    CHECK(rows.size() == 2 && rows[0] == (Row{1.0, 2.0, 0.0, std::nullopt, std::nullopt}));
    CHECK(rows.size() == 2 && rows[1] == (Row{2.0, 2.0, 2.0, 11.0, 5.5}));
    Table e(sch({{"c0", Type::I32}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {}));
        return c;
    }());
    GpuAggregate g(std::make_unique<Scan>(e), {}, {AggSpec::count_star("cs"), AggSpec::sum(0, "s")});
    const auto gr = drain(g);
    CHECK(gr.size() == 1 && gr[0] == (Row{0.0, std::nullopt}));
    GpuAggregate k(std::make_unique<Scan>(e), {0}, {AggSpec::count_star("cs")});
    CHECK(drain(k).empty());
}

void test_asof_plan() {
    Table build(sch({{"c0", Type::I32}, {"c1", Type::TS}, {"c2", Type::I64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 1, 1, 2}));
        c.push_back(mk<int64_t>(Type::TS, {10, 20, 30, 15}));
        c.push_back(mk<int64_t>(Type::I64, {100, 200, 300, 150}));
        return c;
    }());
    Table probe(sch({{"c0", Type::I32}, {"c1", Type::TS}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 1, 1, 2, 3, 1}));
        c.push_back(mk<int64_t>(Type::TS, {5, 10, 25, 15, 99, 50}));
        return c;
    }());
    // Learned twice: as configured, and with the probe clustering forced on.
    for (int pass = 0; pass < 3; ++pass) {
        const auto method = pass == 0 ? tsx::AsofMethod::Binary : tsx::AsofMethod::Learned;
        if (pass == 2) GpuAsofJoin::set_cluster_threshold(1, 0);
        plan::LowerOptions opt;
        opt.asof_method = method;
        auto op = plan::scan(probe)
                      .asof_join(plan::scan(build), {0}, {0}, 1, 1, tsx::AsofType::Inner)
                      .build(opt);
        const auto rows = drain(*op);
        CHECK(rows.size() == 4);
        // (1,10)->100, (1,25)->200, (2,15)->150, (1,50)->300
        std::vector<double> pay;
        for (const auto& r : rows) pay.push_back(*r[4]);
        std::sort(pay.begin(), pay.end());
        CHECK((pay == std::vector<double>{100, 150, 200, 300}));
    }
    GpuAsofJoin::set_cluster_threshold(size_t(1) << 22, uint64_t(1) << 31);
}

void test_exceptions() {
    Table a(sch({{"k", Type::I32}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1}));
        return c;
    }());
    Table b(sch({{"k", Type::I64}, {"f", Type::F64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int64_t>(Type::I64, {1}));
        c.push_back(mk<double>(Type::F64, {1.0}));
        return c;
    }());
    bool threw = false;
    try {
        GpuHashJoin j(std::make_unique<Scan>(a), std::make_unique<Scan>(b), {0}, {0}, JoinType::Inner);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        Scan s(a, 100);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        plan::scan(b).asof_join(plan::scan(b), {}, {}, "f", "f", tsx::AsofType::Inner);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        plan::scan(a).join(plan::scan(b), {"k"}, {"k"}, JoinType::Inner);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
}

void test_hash_table_api() {
    GpuHashTable t({Type::I32});
    OwnedColumn c = mk<int32_t>(Type::I32, {7, 7, 3, 7, 3, 100});
    Column v = c.view();
    KeyColumns kc{&v, 1, nullptr};
    uint32_t g[6];
    t.insert_or_find(kc, 6, g);
    CHECK(g[0] == 0 && g[1] == 0 && g[2] == 1 && g[3] == 0 && g[4] == 1 && g[5] == 2);
    CHECK(t.num_groups() == 3);
    OwnedColumn p = mk<int32_t>(Type::I32, {3, 99});
    Column pv = p.view();
    KeyColumns pk{&pv, 1, nullptr};
    uint32_t f[2];
    t.find(pk, 2, f);
    CHECK(f[0] == 1 && f[1] == kNoGroup);
    CHECK(static_cast<int32_t>(t.group_key_word(2, 0)) == 100);
    GpuHashTable n({Type::I32}, NullPolicy::kNeverMatch);
    OwnedColumn nc = mk<int32_t>(Type::I32, {0, 5, 0, 5}, {0, 2});
    Column nv = nc.view();
    KeyColumns nk{&nv, 1, nullptr};
    uint32_t ng[4];
    n.insert_or_find(nk, 4, ng);
    CHECK(ng[0] == 0 && ng[1] == 1 && ng[2] == 2 && ng[3] == 1);
}

// The reference's scalar==vector seams (ops/join.h:86-89, ops/aggregate.h:116-119,
// tsx/asof.h:110-113, ops/hashtable.h:125-132): every path selection gives the
// same result through the device operators.
void test_path_seams() {
    Table build(sch({{"c0", Type::I32}, {"c1", Type::TS}, {"c2", Type::I64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 1, 1, 2}));
        c.push_back(mk<int64_t>(Type::TS, {10, 20, 30, 15}));
        c.push_back(mk<int64_t>(Type::I64, {100, 200, 300, 150}));
        return c;
    }());
    Table probe(sch({{"c0", Type::I32}, {"c1", Type::TS}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 1, 1, 2, 3, 1}));
        c.push_back(mk<int64_t>(Type::TS, {5, 10, 25, 15, 99, 50}));
        return c;
    }());
    auto sorted = [](std::vector<Row> r) {
        std::sort(r.begin(), r.end());
        return r;
    };
    std::vector<std::vector<Row>> joins, aggs, asofs;
    for (HashPath hp : {HashPath::kVector, HashPath::kScalar}) {
        for (int k = 0; k < 2; ++k) {
            const GatherPath gp = k ? GatherPath::kScalar : GatherPath::kVector;
            GpuHashJoin j(std::make_unique<Scan>(probe), std::make_unique<Scan>(build), {0}, {0}, JoinType::Left);
            j.set_paths(hp, gp);
            joins.push_back(sorted(drain(j)));
            GpuAggregate a(std::make_unique<Scan>(build), {0}, {AggSpec::count_star("cs"), AggSpec::sum(2, "s")});
            a.set_paths(hp, k ? AggKernelPath::kScalar : AggKernelPath::kVector);
            aggs.push_back(sorted(drain(a)));
            GpuAsofJoin s(std::make_unique<Scan>(probe), std::make_unique<Scan>(build), {0}, {0}, 1, 1,
                          tsx::AsofType::Inner);
            s.set_paths(hp, gp);
            asofs.push_back(sorted(drain(s)));
        }
    }
    for (size_t i = 1; i < joins.size(); ++i) {
        CHECK(joins[i] == joins[0]);
        CHECK(aggs[i] == aggs[0]);
        CHECK(asofs[i] == asofs[0]);
    }
    CHECK(joins[0].size() == 14 && aggs[0].size() == 2 && asofs[0].size() == 4);  // 4x3 + 1 + one NULL-extended
    GpuHashTable t({Type::I32});
    OwnedColumn c = mk<int32_t>(Type::I32, {7, 7, 3});
    Column v = c.view();
    KeyColumns kc{&v, 1, nullptr};
    uint32_t g[3], f[3];
    t.insert_or_find(kc, 3, g, HashPath::kScalar);
    t.find(kc, 3, f, HashPath::kScalar);
    CHECK(g[0] == 0 && g[1] == 0 && g[2] == 1 && f[0] == 0 && f[2] == 1);
}

// tests/ops_test.cpp:42-56: c0 = 0..n-1 (NULL where i % 7 == 0), c1 = 1.5 * c0.
Table seq_table(size_t n) {
    std::vector<int32_t> a(n);
    std::vector<double> b(n);
    std::vector<size_t> nulls;
    for (size_t i = 0; i < n; ++i) {
        a[i] = static_cast<int32_t>(i);
        b[i] = 1.5 * static_cast<double>(i);
        if (i % 7 == 0) nulls.push_back(i);
    }
    std::vector<OwnedColumn> c;
    c.push_back(mk<int32_t>(Type::I32, a, nulls));
    c.push_back(mk<double>(Type::F64, b));
    return Table(sch({{"c0", Type::I32}, {"c1", Type::F64}}), std::move(c));
}

void test_filter_reference_cases() {
    using namespace qe::expr;
    Table t = seq_table(300);
    {  // TRUE keeps all, FALSE keeps none (tests/ops_test.cpp:99-113)
        GpuFilter f(std::make_unique<Scan>(t, 64), lit(Scalar::boolean(true)));
        CHECK(drain(f).size() == 300);
        GpuFilter g(std::make_unique<Scan>(t, 64), lit(Scalar::boolean(false)));
        CHECK(drain(g).empty());
        GpuFilter h(std::make_unique<Scan>(t, 64), lit(Scalar::null(Type::BOOL)));
        CHECK(drain(h).empty());
    }
    {  // c0 >= 100 keeps rows 100..299 minus NULL c0, in input order
        GpuFilter f(std::make_unique<Scan>(t, 64), ge(col(Type::I32, 0), lit(Scalar::i32(100))));
        size_t nb = 0;
        const auto rows = drain(f, &nb, false);
        std::vector<Row> want;
        for (int i = 100; i < 300; ++i)
            if (i % 7) want.push_back(Row{double(i), 1.5 * i});
        CHECK(rows == want);
        CHECK(nb >= 1);
    }
    {  // literal on the left mirrors the operator: 100 <= c0
        GpuFilter f(std::make_unique<Scan>(t, 64), le(lit(Scalar::i32(100)), col(Type::I32, 0)));
        CHECK(drain(f).size() == 200 - 28);  // 100..299 minus the 28 multiples of 7 (NULL c0)
    }
    {  // conjunction with promotion: c0 > 10.5 AND c1 < 30 (c1 = 1.5 * c0)
        GpuFilter f(std::make_unique<Scan>(t),
                    logic_and(gt(col(Type::I32, 0), lit(Scalar::f64(10.5))), lt(col(Type::F64, 1), lit(Scalar::i64(30)))));
        const auto rows = drain(f, nullptr, false);
        std::vector<Row> want;
        for (int i = 11; i < 20; ++i)
            if (i % 7) want.push_back(Row{double(i), 1.5 * i});
        CHECK(rows == want);
    }
}

void test_filter_plan_and_device_chaining() {
    using namespace qe::expr;
    Table t = seq_table(5000);
    // Aggregate(Filter(Scan)): the filtered columns stay on the device.
    auto op = plan::scan(t)
                  .filter(gt(col(t.schema(), "c1"), lit(Scalar::f64(100.0))))
                  .aggregate({}, {AggSpec::count_star("n"), AggSpec::count(0, "c"), AggSpec::sum(0, "s")})
                  .build();
    const auto rows = drain(*op);
    double n = 0, c = 0, s = 0;
    for (int i = 0; i < 5000; ++i)
        if (1.5 * i > 100.0) {
            ++n;
            if (i % 7) {
                ++c;
                s += i;
            }
        }
    CHECK(rows.size() == 1 && rows[0] == (Row{n, c, s}));
    // Join(Filter(Scan), Scan) and output batches of <= 2048 rows.
    Table dim(sch({{"k", Type::I32}, {"g", Type::I64}}), [] {
        std::vector<OwnedColumn> c;
        std::vector<int32_t> k(5000);
        std::vector<int64_t> g(5000);
        for (int i = 0; i < 5000; ++i) {
            k[i] = i;
            g[i] = i % 3;
        }
        c.push_back(mk<int32_t>(Type::I32, k));
        c.push_back(mk<int64_t>(Type::I64, g));
        return c;
    }());
    auto j = plan::scan(t)
                 .filter(logic_and(ge(col(Type::I32, 0), lit(Scalar::i32(1000))), lt(col(Type::I32, 0), lit(Scalar::i64(4000)))))
                 .join(plan::scan(dim), {"c0"}, {"k"}, JoinType::Inner)
                 .aggregate({"g"}, {AggSpec::count_star("n")})
                 .build();
    const auto jr = drain(*j);
    double cnt[3] = {0, 0, 0};
    for (int i = 1000; i < 4000; ++i)
        if (i % 7) cnt[i % 3] += 1;
    CHECK(jr.size() == 3 && jr[0] == (Row{0.0, cnt[0]}) && jr[1] == (Row{1.0, cnt[1]}) && jr[2] == (Row{2.0, cnt[2]}));
    CHECK(plan::scan(t).filter(gt(col(Type::I32, 0), lit(Scalar::i32(0)))).plan().to_string().rfind("Filter", 0) == 0);
}

void test_filter_exceptions() {
    using namespace qe::expr;
    Table t = seq_table(10);
    auto expect = [](auto fn, bool runtime) {
        bool threw = false;
        try {
            fn();
        } catch (const std::invalid_argument&) {
            threw = !runtime;
        } catch (const std::runtime_error&) {
            threw = runtime;
        }
        CHECK(threw);
    };
    // type errors at build time (expr/expr.cpp:133-151) and in the plan (plan/plan.cpp:307-321)
    expect([&] { lt(col(Type::I32, 0), lit(Scalar::boolean(true))); }, false);
    expect([&] { logic_and(col(Type::I32, 0), lit(Scalar::boolean(true))); }, false);
    expect([&] { plan::scan(t).filter(col(Type::I32, 0)); }, false);
    expect([&] { plan::scan(t).filter(gt(col(Type::I64, 0), lit(Scalar::i64(0)))); }, false);
    expect([&] { plan::scan(t).filter(gt(col(Type::I32, 5), lit(Scalar::i32(0)))); }, false);
    // Filter::open rejects a non-BOOL predicate (ops/filter.cpp:18-20)
    expect([&] {
        GpuFilter f(std::make_unique<Scan>(t), col(Type::I32, 0));
        f.open();
    }, false);
    // shapes outside the device path
    expect([&] {
        GpuFilter f(std::make_unique<Scan>(t), logic_or(gt(col(Type::I32, 0), lit(Scalar::i32(1))),
                                                       lt(col(Type::I32, 0), lit(Scalar::i32(0)))));
        f.open();
    }, true);
    expect([&] {
        GpuFilter f(std::make_unique<Scan>(t), lt(col(Type::I32, 0), col(Type::F64, 1)));
        f.open();
    }, true);
}

// ---- dictionary-encoded VARCHAR ---------------------------------------------------------------
OwnedColumn str_col(const std::vector<std::optional<std::string>>& vals) {
    auto dict = std::make_shared<StringDict>();
    OwnedColumn c = OwnedColumn::make_str(vals.size(), dict);
    auto* codes = reinterpret_cast<int32_t*>(c.mutable_data());
    for (size_t i = 0; i < vals.size(); ++i) {
        if (vals[i]) codes[i] = dict->intern(*vals[i]);
        else c.set_null(i);
    }
    return c;
}

// Every cell as text ("NULL", integers, doubles with %.6g, strings resolved
// through the column's dictionary), rows sorted.
std::vector<std::vector<std::string>> drain_text(Operator& op) {
    std::vector<std::vector<std::string>> rows;
    op.open();
    while (auto b = op.next()) {
        for (size_t k = 0; k < b->row_count; ++k) {
            const size_t p = b->sel ? b->sel->idx[k] : k;
            std::vector<std::string> r;
            for (const Column& c : b->cols) {
                const bool valid = c.all_valid || ((c.validity[p >> 6] >> (p & 63)) & 1ull);
                if (!valid) {
                    r.push_back("NULL");
                    continue;
                }
                char buf[64];
                switch (c.type) {
                    case Type::STR:
                        r.emplace_back(c.dict->at(reinterpret_cast<const int32_t*>(c.data)[p]));
                        continue;
                    case Type::I32:
                        std::snprintf(buf, sizeof buf, "%d", reinterpret_cast<const int32_t*>(c.data)[p]);
                        break;
                    case Type::F64:
                        std::snprintf(buf, sizeof buf, "%.6g", reinterpret_cast<const double*>(c.data)[p]);
                        break;
                    case Type::BOOL:
                        std::snprintf(buf, sizeof buf, "%d", reinterpret_cast<const uint8_t*>(c.data)[p]);
                        break;
                    default:
                        std::snprintf(buf, sizeof buf, "%lld",
                                      static_cast<long long>(reinterpret_cast<const int64_t*>(c.data)[p]));
                }
                r.emplace_back(buf);
            }
            rows.push_back(std::move(r));
        }
    }
    op.close();
    std::sort(rows.begin(), rows.end());
    return rows;
}

void test_str_join_by_value() {
    // tests/wp7b_strings_differential_test.cpp:249-268: build "AAPL" is code 0,
    // the probe's "AAPL" is code 1 in its own dictionary.
    Table build(sch({{"c0", Type::STR}, {"c1", Type::I64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(str_col({"AAPL", "MSFT"}));
        c.push_back(mk<int64_t>(Type::I64, {10, 20}));
        return c;
    }());
    Table probe(sch({{"c0", Type::STR}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(str_col({"ZZZ", "AAPL", std::nullopt}));
        return c;
    }());
    GpuHashJoin j(std::make_unique<Scan>(probe), std::make_unique<Scan>(build), {0}, {0}, JoinType::Inner);
    const auto rows = drain_text(j);
    CHECK(rows == (std::vector<std::vector<std::string>>{{"AAPL", "AAPL", "10"}}));
    GpuHashJoin l(std::make_unique<Scan>(probe), std::make_unique<Scan>(build), {0}, {0}, JoinType::Left);
    const auto lrows = drain_text(l);
    CHECK(lrows == (std::vector<std::vector<std::string>>{
                       {"AAPL", "AAPL", "10"}, {"NULL", "NULL", "NULL"}, {"ZZZ", "NULL", "NULL"}}));
}

void test_str_group_by_name_plan() {
    // The tick pipeline's tail: trades (sym id, px) join the symbol table on
    // the id, GROUP BY the dictionary-encoded name with COUNT, AVG(px) and
    // MIN/MAX(name); two ids share the name "ACME".
    Table syms(sch({{"id", Type::I32}, {"name", Type::STR}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {0, 1, 2, 3}));
        c.push_back(str_col({"ZETA", "ACME", "beta", "ACME"}));
        return c;
    }());
    Table trades(sch({{"sym", Type::I32}, {"px", Type::F64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {1, 3, 0, 2, 2, 9, 1}));
        c.push_back(mk<double>(Type::F64, {10, 20, 30, 40, 50, 60, 70}));
        return c;
    }());
    for (bool fuse : {true, false}) {
        plan::LowerOptions opt;
        opt.fuse_join_aggregate = fuse;
        auto op = plan::scan(trades)
                      .join(plan::scan(syms), {"sym"}, {"id"}, JoinType::Inner)
                      .aggregate({"name"}, {AggSpec::count_star("n"), AggSpec::avg(1, "avg_px"),
                                            AggSpec::min(3, "lo"), AggSpec::max(3, "hi")})
                      .build(opt);
        const auto rows = drain_text(*op);
        CHECK(rows == (std::vector<std::vector<std::string>>{{"ACME", "3", "33.3333", "ACME", "ACME"},
                                                              {"ZETA", "1", "30", "ZETA", "ZETA"},
                                                              {"beta", "2", "45", "beta", "beta"}}));
    }
    // MIN/MAX over names by byte value: "ACME" < "ZETA" < "beta"
    auto g = plan::scan(syms).aggregate({}, {AggSpec::min(1, "lo"), AggSpec::max(1, "hi"), AggSpec::count(1, "n")}).build();
    CHECK(drain_text(*g) == (std::vector<std::vector<std::string>>{{"ACME", "beta", "4"}}));
    // Filter keeps the STR payload (and its dictionary)
    using namespace qe::expr;
    auto f = plan::scan(syms).filter(ge(col(Type::I32, 0), lit(Scalar::i32(2)))).build();
    CHECK(drain_text(*f) == (std::vector<std::vector<std::string>>{{"2", "beta"}, {"3", "ACME"}}));
}

void test_sort_operator() {
    // ops/sort.h:84: stable; DESC; NULL placement absolute; F64 NaN greatest; STR by value.
    Table t(sch({{"k", Type::I64}, {"f", Type::F64}, {"s", Type::STR}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int64_t>(Type::I64, {3, 1, 3, 2, 1}, {3}));
        c.push_back(mk<double>(Type::F64, {0.5, std::nan(""), -1.0, 2.0, 0.5}));
        c.push_back(str_col({"b", "a", "C", std::nullopt, "a"}));
        return c;
    }());
    auto text_rows = [](Operator& op) {  // in output order
        std::vector<std::vector<std::string>> rows;
        op.open();
        while (auto b = op.next()) {
            for (size_t k = 0; k < b->row_count; ++k) {
                std::vector<std::string> r;
                for (const Column& c : b->cols) {
                    const bool valid = c.all_valid || ((c.validity[k >> 6] >> (k & 63)) & 1ull);
                    if (!valid) {
                        r.push_back("NULL");
                    } else if (c.type == Type::STR) {
                        r.emplace_back(c.dict->at(reinterpret_cast<const int32_t*>(c.data)[k]));
                    } else if (c.type == Type::F64) {
                        char buf[32];
                        std::snprintf(buf, sizeof buf, "%g", reinterpret_cast<const double*>(c.data)[k]);
                        r.emplace_back(buf);
                    } else {
                        r.push_back(std::to_string(reinterpret_cast<const int64_t*>(c.data)[k]));
                    }
                }
                rows.push_back(r);
            }
        }
        op.close();
        return rows;
    };
    {
        auto op = plan::scan(t).sort({plan::SortBy{"k", SortDir::Desc, NullOrder::Last}}).build();
        const auto rows = text_rows(*op);
        CHECK(rows == (std::vector<std::vector<std::string>>{{"3", "0.5", "b"}, {"3", "-1", "C"},
                                                              {"1", "nan", "a"}, {"1", "0.5", "a"},
                                                              {"NULL", "2", "NULL"}}));
    }
    {
        GpuSort op(std::make_unique<Scan>(t), {SortKey{1, SortDir::Asc, NullOrder::First}});
        const auto rows = text_rows(op);
        CHECK(rows.size() == 5 && rows[0][1] == "-1" && rows[1][1] == "0.5" && rows[1][0] == "3" &&
              rows[2][1] == "0.5" && rows[2][0] == "1" && rows[3][1] == "2" && rows[4][1] == "nan");
    }
    {
        GpuSort op(std::make_unique<Scan>(t), {SortKey{2, SortDir::Asc, NullOrder::First}});
        const auto rows = text_rows(op);  // NULL, then byte order "C" < "a" < "b", stable
        CHECK(rows.size() == 5 && rows[0][2] == "NULL" && rows[1][2] == "C" && rows[2][2] == "a" &&
              rows[2][0] == "1" && rows[2][1] == "nan" && rows[3][2] == "a" && rows[4][2] == "b");
    }
    bool threw = false;
    try {
        plan::scan(t).sort(std::vector<SortKey>{});
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
}

void test_tick_pipeline_plan() {
    // The tick pipeline end to end on the device: trades WHERE px > 100, as-of
    // joined to quotes (the latest quote at or before the trade, same symbol),
    // joined to the symbol table on the id, GROUP BY name COUNT(*) / AVG(px),
    // ORDER BY the count descending.
    Table trades(sch({{"sym", Type::I32}, {"ts", Type::TS}, {"px", Type::F64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {0, 0, 1, 1, 2, 2, 2, 0, 1}));
        c.push_back(mk<int64_t>(Type::TS, {10, 20, 5, 30, 40, 50, 60, 1, 31}));
        c.push_back(mk<double>(Type::F64, {150, 50, 120, 130, 101, 102, 99, 200, 140}));
        return c;
    }());
    Table quotes(sch({{"sym", Type::I32}, {"ts", Type::TS}, {"bid", Type::F64}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {0, 1, 2, 1}));
        c.push_back(mk<int64_t>(Type::TS, {5, 25, 45, 3}));
        c.push_back(mk<double>(Type::F64, {1, 2, 3, 4}));
        return c;
    }());
    Table syms(sch({{"id", Type::I32}, {"name", Type::STR}}), [] {
        std::vector<OwnedColumn> c;
        c.push_back(mk<int32_t>(Type::I32, {2, 0, 1}));
        c.push_back(str_col({"CCC", "AAA", "BB"}));
        return c;
    }());
    using namespace qe::expr;
    // Binary, learned, and learned with the probe clustering forced on.
    for (int pass = 0; pass < 3; ++pass) {
        const auto method = pass == 0 ? tsx::AsofMethod::Binary : tsx::AsofMethod::Learned;
        GpuAsofJoin::set_cluster_threshold(pass == 2 ? 1 : size_t(1) << 22, pass == 2 ? 0 : uint64_t(1) << 31);
        plan::LowerOptions opt;
        opt.asof_method = method;
        auto op = plan::scan(trades)
                      .filter(gt(col(Type::F64, 2), lit(Scalar::f64(100.0))))
                      .asof_join(plan::scan(quotes), {"sym"}, {"sym"}, "ts", "ts", tsx::AsofType::Inner)
                      .join(plan::scan(syms), {0}, {"id"}, JoinType::Inner)
                      .aggregate({"name"}, {AggSpec::count_star("n"), AggSpec::avg(2, "avg_px")})
                      .sort({plan::SortBy{"n", SortDir::Desc, NullOrder::Last}})
                      .build(opt);
        // px > 100: (0,10,150) (1,5,120) (1,30,130) (2,40,101) (2,50,102) (0,1,200) (1,31,140).
        // As-of (quote ts <= trade ts, same sym): sym0 quote@5 matches ts 10 (not ts 1);
        // sym1 quotes @3,@25 match ts 5, 30, 31; sym2 quote@45 matches ts 50 (not 40).
        // -> BB: 3 rows avg 130; AAA: 1 row 150; CCC: 1 row 102.
        std::vector<std::vector<std::string>> rows;
        op->open();
        while (auto b = op->next()) {
            for (size_t k = 0; k < b->row_count; ++k) {
                const Column& name = b->cols[0];
                char buf[64];
                std::snprintf(buf, sizeof buf, "%lld|%g",
                              static_cast<long long>(reinterpret_cast<const int64_t*>(b->cols[1].data)[k]),
                              reinterpret_cast<const double*>(b->cols[2].data)[k]);
                rows.push_back({std::string(name.dict->at(reinterpret_cast<const int32_t*>(name.data)[k])), buf});
            }
        }
        op->close();
        CHECK(rows.size() == 3 && rows[0] == (std::vector<std::string>{"BB", "3|130"}));
        std::sort(rows.begin() + 1, rows.end());
        CHECK(rows.size() == 3 && rows[1] == (std::vector<std::string>{"AAA", "1|150"}) &&
              rows[2] == (std::vector<std::string>{"CCC", "1|102"}));
    }
    GpuAsofJoin::set_cluster_threshold(size_t(1) << 22, uint64_t(1) << 31);
}

int main() {
    const std::vector<std::pair<const char*, std::function<void()>>> tests = {
        {"join_groupby_plan", test_join_groupby_plan},
        {"join_fanout_crosses_output_batches", test_join_fanout_crosses_output_batches},
        {"aggregate_edges", test_aggregate_edges},
        {"asof_plan", test_asof_plan},
        {"exceptions", test_exceptions},
        {"hash_table_api", test_hash_table_api},
        {"path_seams", test_path_seams},
        {"filter_reference_cases", test_filter_reference_cases},
        {"filter_plan_and_device_chaining", test_filter_plan_and_device_chaining},
        {"filter_exceptions", test_filter_exceptions},
        {"str_join_by_value", test_str_join_by_value},
        {"str_group_by_name_plan", test_str_group_by_name_plan},
        {"sort_operator", test_sort_operator},
        {"tick_pipeline_plan", test_tick_pipeline_plan},
    };
    for (const auto& [name, fn] : tests) {
        std::printf("[host] %s\n", name);
        std::fflush(stdout);
        try {
            fn();
        } catch (const std::exception& e) {
            ++g_fail;
            std::printf("  FAIL exception: %s\n", e.what());
        }
    }
    std::printf("[host] %d checks passed, %d failed\n", g_pass, g_fail);
    return g_fail == 0 ? 0 : 1;
}
