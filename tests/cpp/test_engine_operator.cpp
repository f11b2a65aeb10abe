// test_engine_operator.cpp -- PhysicalGpuHashAggregate as a subclass of the
// reference engine's own PhysicalOperator, fed the engine's own DataChunk /
// Vector / string_t (compiled from the reference's include/ and
// src/common/types/*.cpp by oracle/Makefile's `engine` target; built only
// where the reference tree is present).
//
//   test_engine_operator --cpu   planning, CapturePredicates, GpuAggregateSupports (no GPU)
//   test_engine_operator         everything (needs the GPU)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../slothdb_amd/host/engine/physical_gpu_hash_aggregate.hpp"

using namespace slothdb;
using gpu_aggregate::Aggregate;

static int g_failed = 0, g_checks = 0;
#define CHECK(cond)                                                                       \
    do {                                                                                  \
        g_checks++;                                                                       \
        if (!(cond)) {                                                                    \
            g_failed++;                                                                   \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
        }                                                                                 \
    } while (0)

// A child operator emitting prepared DataChunks (the scan below the aggregate).
class ChunkSource : public PhysicalOperator {
public:
    explicit ChunkSource(std::vector<LogicalType> types) : PhysicalOperator(PhysicalOperatorType::DUMMY_SCAN, types) {}
    std::vector<std::vector<Value>> rows;  // all rows; emitted VECTOR_SIZE at a time
    void Init() override { pos_ = 0; }
    bool GetData(DataChunk &out) override {
        if (pos_ >= rows.size()) return false;
        out.Initialize(GetTypes());
        idx_t n = 0;
        for (; pos_ < rows.size() && n < VECTOR_SIZE; pos_++, n++)
            for (idx_t c = 0; c < rows[pos_].size(); c++) out.SetValue(c, n, rows[pos_][c]);
        out.SetCardinality(n);
        return true;
    }

private:
    size_t pos_ = 0;
};

// The planner's compiled predicate, as physical_planner.cpp:110-135 declares
// it (file-local there; CapturePredicates takes the engine's own type).
enum class SimpleCmpOp { LT, LE, GT, GE, EQ, NE };
struct SimplePredicate {
    idx_t col_idx;
    SimpleCmpOp op;
    double dval;
    int64_t ival = 0;
    std::string sval;
    bool str_form = false;
    bool like_contains = false;
    bool like_negated = false;
    bool int_in_form = false;
    std::vector<int64_t> in_set;
};

// (name VARCHAR, dept VARCHAR, salary INTEGER): test/unit/execution/test_aggregation.cpp:7-14
static std::unique_ptr<ChunkSource> emp(const char *const *depts = nullptr) {
    static const char *const d0[] = {"Engineering", "Engineering", "Sales", "Sales", "Marketing"};
    if (!depts) depts = d0;
    const char *names[] = {"Alice", "Bob", "Charlie", "Diana", "Eve"};
    const int32_t sal[] = {100, 120, 80, 90, 95};
    auto src = std::make_unique<ChunkSource>(
        std::vector<LogicalType>{LogicalType::VARCHAR(), LogicalType::VARCHAR(), LogicalType::INTEGER()});
    for (int i = 0; i < 5; i++) src->rows.push_back({Value(names[i]), Value(depts[i]), Value::INTEGER(sal[i])});
    return src;
}

static std::vector<LogicalType> emp_types() {
    return {LogicalType::VARCHAR(), LogicalType::VARCHAR(), LogicalType::INTEGER()};
}

// Every row of the operator's output, as Values read back through the engine's own GetValue.
static std::vector<std::vector<Value>> drain(PhysicalOperator &op, size_t *chunks = nullptr) {
    std::vector<std::vector<Value>> out;
    op.Init();
    DataChunk c;
    size_t nc = 0;
    while (op.GetData(c)) {
        nc++;
        CHECK(c.size() <= VECTOR_SIZE);
        for (idx_t r = 0; r < c.size(); r++) {
            std::vector<Value> row;
            for (idx_t k = 0; k < c.ColumnCount(); k++) row.push_back(c.GetValue(k, r));
            out.push_back(row);
        }
        c = DataChunk();
    }
    if (chunks) *chunks = nc;
    return out;
}

static void test_planning_cpu() {
    // GpuAggregateSupports: path types and aggregate kinds
    CHECK(GpuAggregateSupports(emp_types(), {1}, {{"count_star"}, {"sum", 2}, {"avg", 2}, {"min", 2}, {"max", 1}}));
    CHECK(GpuAggregateSupports(emp_types(), {}, {{"count", 0, true}}));
    CHECK(!GpuAggregateSupports({LogicalType::BOOLEAN(), LogicalType::INTEGER()}, {0}, {{"sum", 1}}));
    CHECK(!GpuAggregateSupports(emp_types(), {1}, {{"stddev", 2}}));
    CHECK(!GpuAggregateSupports(emp_types(), {1}, {{"sum", 2, true}}));
    CHECK(!GpuAggregateSupports(emp_types(), {}, {}));

    // CapturePredicates: every SimplePredicate form
    std::vector<SimplePredicate> sp(5);
    sp[0].col_idx = 2; sp[0].op = SimpleCmpOp::LT; sp[0].ival = 30; sp[0].dval = 30.0;
    sp[1].col_idx = 2; sp[1].op = SimpleCmpOp::EQ; sp[1].int_in_form = true; sp[1].in_set = {3, 7, 435090932899640449LL};
    sp[2].col_idx = 1; sp[2].op = SimpleCmpOp::EQ; sp[2].str_form = true; sp[2].sval = "Sales";
    sp[3].col_idx = 0; sp[3].op = SimpleCmpOp::EQ; sp[3].like_contains = true; sp[3].like_negated = true; sp[3].sval = "li";
    sp[4].col_idx = 2; sp[4].op = SimpleCmpOp::GE; sp[4].ival = 435090932899640449LL; sp[4].dval = 4.3509093289964045e17;
    GpuPredicates gp;
    CHECK(CapturePredicates(sp, gp));
    CHECK(gp.preds.size() == 5);
    CHECK(gp.preds[0].column == 2 && gp.preds[0].op == GBAGG_CMP_LT && gp.preds[0].ival == 30);
    CHECK(gp.preds[1].op == GBAGG_CMP_IN && gp.preds[1].in_count == 3 && gp.preds[1].in_set[2] == 435090932899640449LL);
    CHECK(gp.preds[2].op == GBAGG_CMP_EQ && gp.preds[2].slen == 5 && std::memcmp(gp.preds[2].sval, "Sales", 5) == 0);
    CHECK(gp.preds[3].op == GBAGG_CMP_NOT_LIKE && gp.preds[3].slen == 2);
    CHECK(gp.preds[4].op == GBAGG_CMP_GE && gp.preds[4].ival == 435090932899640449LL);  // exact, not via dval
    std::vector<SimplePredicate> bad(1);
    bad[0].col_idx = 1; bad[0].op = SimpleCmpOp::LT; bad[0].str_form = true; bad[0].sval = "x";
    CHECK(!CapturePredicates(bad, gp));  // VARCHAR < has no GPU form

    // plans are created on the host (no device calls); bad result types throw
    PhysicalGpuHashAggregate ok(emp_types(), {1}, {{"count_star"}, {"sum", 2}},
                                {LogicalType::VARCHAR(), LogicalType::BIGINT(), LogicalType::BIGINT()});
    bool threw = false;
    try {
        PhysicalGpuHashAggregate wrong(emp_types(), {1}, {{"count_star"}}, {LogicalType::VARCHAR()});
    } catch (const GpuAggregateError &e) {
        threw = e.status == GBAGG_INVALID;
    }
    CHECK(threw);
    threw = false;
    try {
        PhysicalGpuHashAggregate sv(emp_types(), {}, {{"sum", 1}}, {LogicalType::BIGINT()});
        sv.Init();  // SUM over VARCHAR: refused by the plan
    } catch (const GpuAggregateError &e) {
        threw = e.status == GBAGG_INVALID;
    }
    CHECK(threw);
}

static void test_known_answers_gpu() {
    // SELECT COUNT(*), SUM(salary), AVG(salary), MIN(salary), MAX(salary) FROM emp
    {
        PhysicalGpuHashAggregate op(emp_types(), {}, {{"count_star"}, {"sum", 2}, {"avg", 2}, {"min", 2}, {"max", 2}},
                                    {LogicalType::BIGINT(), LogicalType::BIGINT(), LogicalType::DOUBLE(),
                                     LogicalType::INTEGER(), LogicalType::INTEGER()});
        op.children.push_back(emp());
        auto rows = drain(op);
        CHECK(rows.size() == 1);
        if (rows.size() == 1) {
            CHECK(rows[0][0].GetValue<int64_t>() == 5);
            CHECK(rows[0][1].GetValue<int64_t>() == 485);
            CHECK(std::fabs(rows[0][2].GetValue<double>() - 97.0) < 1e-12);
            CHECK(rows[0][3].GetValue<int32_t>() == 80);
            CHECK(rows[0][4].GetValue<int32_t>() == 120);
        }
    }
    // SELECT dept, COUNT(*), SUM(salary), AVG(salary) FROM emp GROUP BY dept (:62-108),
    // once with the test's names and once with names longer than string_t's 12 inline bytes
    static const char *const longd[] = {"Engineering and Research", "Engineering and Research", "Sales (EMEA)",
                                        "Sales (EMEA)", "Marketing & Communications"};
    for (int variant = 0; variant < 2; variant++) {
        PhysicalGpuHashAggregate op(emp_types(), {1}, {{"count_star"}, {"sum", 2}, {"avg", 2}},
                                    {LogicalType::VARCHAR(), LogicalType::BIGINT(), LogicalType::BIGINT(),
                                     LogicalType::DOUBLE()});
        op.children.push_back(emp(variant ? longd : nullptr));
        auto rows = drain(op);
        CHECK(rows.size() == 3);
        std::map<std::string, std::vector<Value>> by;
        for (auto &r : rows) by[r[0].GetValue<std::string>()] = r;
        const std::string eng = variant ? longd[0] : "Engineering", sal = variant ? longd[2] : "Sales",
                          mkt = variant ? longd[4] : "Marketing";
        CHECK(by.count(eng) && by[eng][1].GetValue<int64_t>() == 2 && by[eng][2].GetValue<int64_t>() == 220 &&
              by[eng][3].GetValue<double>() == 110.0);
        CHECK(by.count(sal) && by[sal][1].GetValue<int64_t>() == 2 && by[sal][2].GetValue<int64_t>() == 170 &&
              by[sal][3].GetValue<double>() == 85.0);
        CHECK(by.count(mkt) && by[mkt][1].GetValue<int64_t>() == 1 && by[mkt][2].GetValue<int64_t>() == 95);
    }
    // 5000 rows over three chunks, NULL keys and NULL values, a captured
    // predicate (salary >= 0), and the result spanning several output chunks
    {
        auto src = std::make_unique<ChunkSource>(emp_types());
        std::map<std::string, std::pair<int64_t, int64_t>> exp;  // key -> (count, sum)
        int64_t null_cnt = 0;
        for (int i = 0; i < 5000; i++) {
            const int32_t s = (i * 37) % 1000 - 100;
            const bool null_key = i % 97 == 0, null_val = i % 13 == 0;
            const std::string dept = "dept_number_" + std::to_string(i % 2500);
            src->rows.push_back({Value("n"), null_key ? Value() : Value(dept), null_val ? Value() : Value::INTEGER(s)});
            if (null_val || s < 0) continue;  // the predicate drops NULL and negative salaries
            if (null_key) null_cnt++;
            else { exp[dept].first++; exp[dept].second += s; }
        }
        std::vector<SimplePredicate> sp(1);
        sp[0].col_idx = 2; sp[0].op = SimpleCmpOp::GE; sp[0].ival = 0; sp[0].dval = 0.0;
        GpuPredicates gp;
        CHECK(CapturePredicates(sp, gp));
        PhysicalGpuHashAggregate op(emp_types(), {1}, {{"count_star"}, {"sum", 2}},
                                    {LogicalType::VARCHAR(), LogicalType::BIGINT(), LogicalType::BIGINT()}, std::move(gp));
        op.children.push_back(std::move(src));
        size_t chunks = 0;
        auto rows = drain(op, &chunks);
        CHECK(rows.size() == exp.size() + (null_cnt ? 1 : 0));
        CHECK(chunks == (rows.size() + VECTOR_SIZE - 1) / VECTOR_SIZE && chunks > 1);
        size_t seen = 0;
        for (auto &r : rows) {
            if (r[0].IsNull()) {
                CHECK(r[1].GetValue<int64_t>() == null_cnt);
                continue;
            }
            auto it = exp.find(r[0].GetValue<std::string>());
            CHECK(it != exp.end());
            if (it == exp.end()) continue;
            seen++;
            CHECK(r[1].GetValue<int64_t>() == it->second.first && r[2].GetValue<int64_t>() == it->second.second);
        }
        CHECK(seen == exp.size());
    }
    // ORDER BY COUNT(*) DESC LIMIT 1 through SetTopNHint (the aggregate's output slot 1)
    {
        PhysicalGpuHashAggregate op(emp_types(), {1}, {{"count_star"}, {"max", 2}},
                                    {LogicalType::VARCHAR(), LogicalType::BIGINT(), LogicalType::INTEGER()});
        op.SetTopNHint(1, false, 1);
        op.children.push_back(emp());
        auto rows = drain(op);
        CHECK(rows.size() == 1);
        if (rows.size() == 1) CHECK(rows[0][1].GetValue<int64_t>() == 2);
    }
}

int main(int argc, char **argv) {
    const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    try {
        test_planning_cpu();
        if (!cpu_only) test_known_answers_gpu();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    std::printf("%d checks, %d failed\n", g_checks, g_failed);
    return g_failed ? 1 : 0;
}
