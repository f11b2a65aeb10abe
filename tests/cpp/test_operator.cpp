// test_operator.cpp -- PhysicalGpuHashAggregate through its C++ surface
// (Init / ConsumeChunk / ConsumeRowGroup / GetData / SetTopNHint /
// SetRowLimit), with the reference's own known answers where they exist.
//
//   test_operator --cpu   planning, result types and error paths (no GPU)
//   test_operator         everything (needs the GPU)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../slothdb_amd/host/gpu_hash_aggregate.hpp"

using namespace slothdb_gpu;

static int g_failed = 0, g_checks = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        g_checks++;                                                                   \
        if (!(cond)) {                                                                \
            g_failed++;                                                               \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
        }                                                                             \
    } while (0)

using LT = LogicalTypeId;

// (dept VARCHAR, salary INTEGER) of test/unit/execution/test_aggregation.cpp:7-14
static DataChunk emp_chunk() {
    DataChunk c;
    c.Initialize({LT::VARCHAR, LT::INTEGER});
    const char *dept[] = {"Engineering", "Engineering", "Sales", "Sales", "Marketing"};
    const int32_t sal[] = {100, 120, 80, 90, 95};
    for (int i = 0; i < 5; i++) {
        c.data[0].Strings()[i] = dept[i];
        c.data[1].GetData<int32_t>()[i] = sal[i];
    }
    c.SetCardinality(5);
    return c;
}

static std::vector<AggregateSpec> emp_aggs() {
    return {{"count_star", -1}, {"sum", 1}, {"avg", 1}, {"min", 1}, {"max", 1}};
}

// Drains GetData into rows of strings (easy to compare, NULL as "NULL").
static std::vector<std::vector<std::string>> drain(PhysicalGpuHashAggregate &op, size_t *chunks = nullptr) {
    std::vector<std::vector<std::string>> rows;
    DataChunk out;
    size_t nchunks = 0;
    while (op.GetData(out)) {
        nchunks++;
        for (idx_t r = 0; r < out.size(); r++) {
            std::vector<std::string> row;
            for (auto &v : out.data) {
                if (!v.Validity().RowIsValid(r)) {
                    row.push_back("NULL");
                    continue;
                }
                char buf[64];
                switch (v.GetType()) {
                case LT::INTEGER: std::snprintf(buf, sizeof buf, "%d", v.GetData<int32_t>()[r]); break;
                case LT::BIGINT: std::snprintf(buf, sizeof buf, "%lld", (long long)v.GetData<int64_t>()[r]); break;
                case LT::DOUBLE: std::snprintf(buf, sizeof buf, "%.6f", v.GetData<double>()[r]); break;
                case LT::VARCHAR: row.push_back(v.Strings()[r]); continue;
                }
                row.push_back(buf);
            }
            rows.push_back(row);
        }
    }
    if (chunks) *chunks = nchunks;
    return rows;
}

// ---- no GPU needed ---------------------------------------------------------

static void test_result_types() {
    // binder.cpp:1345-1392: COUNT BIGINT, SUM(int) BIGINT, SUM(double) DOUBLE,
    // AVG DOUBLE, MIN/MAX keep the argument type; groups keep their type
    PhysicalGpuHashAggregate op({LT::INTEGER, LT::BIGINT, LT::DOUBLE, LT::VARCHAR}, {3, 0},
                                {{"count_star", -1}, {"COUNT", 1}, {"sum", 0}, {"sum", 2}, {"avg", 0},
                                 {"min", 0}, {"max", 2}});
    auto t = op.GetTypes();
    std::vector<LT> want = {LT::VARCHAR, LT::INTEGER, LT::BIGINT, LT::BIGINT, LT::BIGINT,
                            LT::DOUBLE,  LT::DOUBLE,  LT::INTEGER, LT::DOUBLE};
    CHECK(t == want);
}

static void test_errors() {
    bool threw = false;
    try {
        PhysicalGpuHashAggregate op({LT::INTEGER}, {}, {{"median", 0}});
    } catch (const GpuAggregateException &e) {
        threw = e.status == GBAGG_INVALID;
    }
    CHECK(threw);
    threw = false;
    try {  // SUM over VARCHAR is not a plan the operator accepts: INVALID at planning
        PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER}, {1}, {{"sum", 0}});
        op.Init();
    } catch (const GpuAggregateException &e) {
        threw = e.status == GBAGG_INVALID && std::strlen(e.what()) > 0;
    }
    CHECK(threw);
    threw = false;
    try {
        PhysicalGpuHashAggregate op({LT::INTEGER}, {5}, {{"count_star", -1}});
    } catch (const GpuAggregateException &e) {
        threw = true;
    }
    CHECK(threw);
}

// ---- GPU ---------------------------------------------------------------------

static void test_emp_scalar() {
    PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER}, {}, emp_aggs(), {}, {false, false});
    op.Init();
    op.ConsumeChunk(emp_chunk());
    auto rows = drain(op);
    // test_aggregation.cpp:16-56: COUNT 5, SUM 485, AVG 97, MIN 80, MAX 120
    CHECK(rows.size() == 1);
    if (rows.size() == 1)
        CHECK((rows[0] == std::vector<std::string>{"5", "485", "97.000000", "80", "120"}));
}

static void test_emp_group_by() {
    PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER}, {0}, emp_aggs(), {}, {false, false});
    op.Init();
    op.ConsumeChunk(emp_chunk());
    std::map<std::string, std::vector<std::string>> got;
    for (auto &r : drain(op)) got[r[0]] = std::vector<std::string>(r.begin() + 1, r.end());
    // test_aggregation.cpp:58-120
    CHECK(got.size() == 3);
    CHECK((got["Engineering"] == std::vector<std::string>{"2", "220", "110.000000", "100", "120"}));
    CHECK((got["Sales"] == std::vector<std::string>{"2", "170", "85.000000", "80", "90"}));
    CHECK((got["Marketing"] == std::vector<std::string>{"1", "95", "95.000000", "95", "95"}));
}

static void test_emp_two_chunks() {
    // the same chunk twice: VARCHAR keys of both chunks meet in one group each
    PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER}, {0}, emp_aggs(), {}, {false, false});
    op.Init();
    op.ConsumeChunk(emp_chunk());
    op.ConsumeChunk(emp_chunk());
    std::map<std::string, std::vector<std::string>> got;
    for (auto &r : drain(op)) got[r[0]] = std::vector<std::string>(r.begin() + 1, r.end());
    CHECK(got.size() == 3);
    CHECK((got["Engineering"] == std::vector<std::string>{"4", "440", "110.000000", "100", "120"}));
    CHECK((got["Sales"] == std::vector<std::string>{"4", "340", "85.000000", "80", "90"}));
    CHECK((got["Marketing"] == std::vector<std::string>{"2", "190", "95.000000", "95", "95"}));
}

static void test_three_column_key_chunks() {
    // GROUP BY (dept VARCHAR, grade INTEGER, bonus DOUBLE): a key of three
    // columns (interned row tuples), strings from two chunks, a NULL grade
    PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER, LT::DOUBLE, LT::BIGINT}, {0, 1, 2},
                                {{"count_star", -1}, {"sum", 3}}, {}, {false, true, false, false});
    op.Init();
    const char *dept[] = {"Eng", "Eng", "Ops", "Eng", "Ops", "Eng"};
    const int32_t grade[] = {1, 1, 2, 1, 2, 3};
    const double bonus[] = {0.5, 0.5, 0.0, -0.0, 0.0, 0.5};
    for (int rep = 0; rep < 2; rep++) {
        DataChunk c;
        c.Initialize({LT::VARCHAR, LT::INTEGER, LT::DOUBLE, LT::BIGINT});
        for (int i = 0; i < 6; i++) {
            c.data[0].Strings()[i] = dept[i];
            c.data[1].GetData<int32_t>()[i] = grade[i];
            c.data[2].GetData<double>()[i] = bonus[i];
            c.data[3].GetData<int64_t>()[i] = 10 * i + rep;
        }
        c.data[1].Validity().SetInvalid(5);  // (Eng, NULL, 0.5)
        c.SetCardinality(6);
        op.ConsumeChunk(c);
    }
    std::map<std::string, std::vector<std::string>> got;
    for (auto &r : drain(op)) got[r[0] + "|" + r[1] + "|" + r[2]] = {r[3], r[4]};
    // (Eng,1,0.5): rows 0,1 twice; (Ops,2,0): rows 2,4 twice; (Eng,1,-0.0 = 0): row 3 twice
    CHECK(got.size() == 4);
    CHECK((got["Eng|1|0.500000"] == std::vector<std::string>{"4", "22"}));
    CHECK((got["Ops|2|0.000000"] == std::vector<std::string>{"4", "122"}));
    CHECK((got["Eng|1|0.000000"] == std::vector<std::string>{"2", "61"}));
    CHECK((got["Eng|NULL|0.500000"] == std::vector<std::string>{"2", "101"}));
}

static void test_emp_topn() {
    PhysicalGpuHashAggregate op({LT::VARCHAR, LT::INTEGER}, {0}, {{"count_star", -1}}, {}, {false, false});
    op.SetTopNHint(1, false, 3);  // ORDER BY cnt DESC
    op.Init();
    op.ConsumeChunk(emp_chunk());
    auto rows = drain(op);
    // test_aggregation.cpp:196-212: cnt DESC, dept ASC
    CHECK(rows.size() == 3);
    if (rows.size() == 3) {
        CHECK(rows[2][0] == "Marketing" && rows[2][1] == "1");
        CHECK(rows[0][1] == "2" && rows[1][1] == "2");
    }
}

static void test_sum_1_to_10000_rowgroup() {
    std::vector<int32_t> x(10000);
    for (int i = 0; i < 10000; i++) x[i] = i + 1;
    PhysicalGpuHashAggregate op({LT::INTEGER}, {}, {{"sum", 0}}, {}, {false});
    op.Init();
    ColumnBatch b;
    b.num_rows = x.size();
    b.columns.push_back(ColumnData{LT::INTEGER, x.data()});
    op.ConsumeRowGroup(b);
    auto rows = drain(op);
    CHECK(rows.size() == 1 && rows[0][0] == "50005000");  // test_parallel.cpp:105-123
}

static void test_50k_five_depts() {
    // test_parallel.cpp:129-160, fed as row groups of dictionary codes
    const char *names[] = {"Eng", "Sales", "Marketing", "HR", "Finance"};
    std::string bytes;
    std::vector<uint32_t> offs{0};
    for (auto n : names) {
        bytes += n;
        offs.push_back((uint32_t)bytes.size());
    }
    std::vector<int32_t> id(50000), salary(50000);
    std::vector<uint32_t> dept(50000);
    for (int i = 0; i < 50000; i++) {
        id[i] = i;
        dept[i] = i % 5;
        salary[i] = 50000 + (i % 100) * 1000;
    }
    PhysicalGpuHashAggregate op({LT::INTEGER, LT::VARCHAR, LT::INTEGER}, {1}, {{"count_star", -1}}, {},
                                {false, false, false});
    op.Init();
    for (int s = 0; s < 50000; s += 12288) {  // several row groups, one dictionary
        const int n = std::min(12288, 50000 - s);
        ColumnBatch b;
        b.num_rows = n;
        b.columns.push_back(ColumnData{LT::INTEGER, id.data() + s});
        ColumnData d{LT::VARCHAR, dept.data() + s};
        d.dict_offsets = offs.data();
        d.dict_bytes = bytes.data();
        d.dict_size = 5;
        d.dict_unique = true;
        b.columns.push_back(d);
        b.columns.push_back(ColumnData{LT::INTEGER, salary.data() + s});
        op.ConsumeRowGroup(b);
    }
    std::map<std::string, std::string> got;
    for (auto &r : drain(op)) got[r[0]] = r[1];
    CHECK(got.size() == 5);
    for (auto n : names) CHECK(got[n] == "10000");
}

static void test_getdata_chunks_and_limit() {
    // 5000 groups leave through ceil(5000 / VECTOR_SIZE) = 3 chunks
    const int N = 20000;
    std::vector<int64_t> k(N), v(N);
    for (int i = 0; i < N; i++) {
        k[i] = (int64_t)(i % 5000) * 7919 - 100000;
        v[i] = i;
    }
    ColumnBatch b;
    b.num_rows = N;
    b.columns.push_back(ColumnData{LT::BIGINT, k.data()});
    b.columns.push_back(ColumnData{LT::BIGINT, v.data()});
    PhysicalGpuHashAggregate op({LT::BIGINT, LT::BIGINT}, {0}, {{"count_star", -1}, {"sum", 1}, {"max", 1}}, {},
                                {false, false});
    op.Init();
    op.ConsumeRowGroup(b);
    size_t chunks = 0;
    auto rows = drain(op, &chunks);
    CHECK(rows.size() == 5000);
    CHECK(chunks == 3);
    std::map<long long, std::tuple<long long, long long, long long>> exp;
    for (int i = 0; i < N; i++) {
        auto &e = exp[k[i]];
        std::get<0>(e)++;
        std::get<1>(e) += v[i];
        std::get<2>(e) = std::max<long long>(std::get<2>(e), v[i]);
    }
    bool all = rows.size() == exp.size();
    for (auto &r : rows) {
        auto it = exp.find(std::stoll(r[0]));
        all = all && it != exp.end() && std::stoll(r[1]) == std::get<0>(it->second) &&
              std::stoll(r[2]) == std::get<1>(it->second) && std::stoll(r[3]) == std::get<2>(it->second);
    }
    CHECK(all);
    // re-Init replays the plan; a row limit caps the result (SetRowLimit)
    op.SetRowLimit(10);
    op.Init();
    op.ConsumeRowGroup(b);
    CHECK(drain(op).size() == 10);
}

static void test_filter_and_nulls() {
    // WHERE c < 30 on a nullable column; NULL keys form one group
    DataChunk c;
    c.Initialize({LT::INTEGER, LT::INTEGER});
    const int n = 1000;
    long long exp_cnt[4] = {0, 0, 0, 0}, exp_null = 0;
    for (int i = 0; i < n; i++) {
        c.data[0].GetData<int32_t>()[i] = i % 4;
        c.data[1].GetData<int32_t>()[i] = i % 100;
        const bool null_key = i % 50 == 7, null_c = i % 33 == 5;
        if (null_key) c.data[0].Validity().SetInvalid(i);
        if (null_c) c.data[1].Validity().SetInvalid(i);
        if (!null_c && i % 100 < 30) (null_key ? exp_null : exp_cnt[i % 4])++;
    }
    c.SetCardinality(n);
    PredicateSpec lt30;
    lt30.column = 1;
    lt30.op = GBAGG_CMP_LT;
    lt30.ival = 30;
    lt30.dval = 30.0;
    PhysicalGpuHashAggregate op({LT::INTEGER, LT::INTEGER}, {0}, {{"count_star", -1}}, {lt30});
    op.Init();
    op.ConsumeChunk(c);
    std::map<std::string, std::string> got;
    for (auto &r : drain(op)) got[r[0]] = r[1];
    CHECK(got.size() == 5);
    for (int k = 0; k < 4; k++) CHECK(got[std::to_string(k)] == std::to_string(exp_cnt[k]));
    CHECK(got["NULL"] == std::to_string(exp_null));
}

int main(int argc, char **argv) {
    const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    test_result_types();
    test_errors();
    if (!cpu_only) {
        test_emp_scalar();
        test_emp_group_by();
        test_emp_two_chunks();
        test_three_column_key_chunks();
        test_emp_topn();
        test_sum_1_to_10000_rowgroup();
        test_50k_five_depts();
        test_getdata_chunks_and_limit();
        test_filter_and_nulls();
    }
    std::printf("test_operator: %d checks, %d failed%s\n", g_checks, g_failed, cpu_only ? " (cpu only)" : "");
    return g_failed ? 1 : 0;
}
