// physical_gpu_hash_aggregate.hpp -- the GPU GROUP BY operator as a subclass of
// the engine's own PhysicalOperator, for compilation inside SlothDB (its
// include/ directory on the include path, next to include/slothdb_gpu.h).
//
//   PhysicalGpuHashAggregate   : slothdb::PhysicalOperator
//       (include/slothdb/execution/physical_operator.hpp:23-65), the drop-in
//       for PhysicalHashAggregate (src/execution/physical_planner.cpp:6009-6060)
//       that PhysicalPlanner::PlanAggregate (physical_planner.cpp:14356-14362)
//       constructs when GpuAggregateSupports() accepts the bound query. It
//       pulls its child's DataChunks (data_chunk.hpp:10-55) and reads their
//       Vectors (vector.hpp:65-147), ValidityMasks (validity_mask.hpp:9-69)
//       and 16-byte string_t cells (string_type.hpp:19-96) in place: numeric
//       columns are passed by pointer, strings are packed once into offsets +
//       bytes for the C ABI (no std::string per cell). Results are written
//       into the caller's DataChunk as typed arrays, validity bits and
//       string_t cells backed by the vector's own string buffer (no Value
//       boxing).
//   CapturePredicates          the planner's compiled SimplePredicate list
//       (physical_planner.cpp:110-135, filled by TryCompileSimplePredicate,
//       :180-327) as gbagg_predicate conjuncts, evaluated inside the GPU
//       kernels (the Filter->Scan peek-through, :7138-7161, :10303-10316).
//   GpuAggregateSupports       the planner hook's test: types and aggregate
//       kinds the operator takes.
//
// Everything below the operator is the C ABI of include/slothdb_gpu.h.
#pragma once

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "slothdb/execution/physical_operator.hpp"
#include "slothdb_gpu.h"

namespace slothdb {

// The operator's failures: status and message of the C ABI (the engine would
// rethrow as InternalException, include/slothdb/common/exception.hpp:90-180).
class GpuAggregateError : public std::runtime_error {
public:
    GpuAggregateError(gbagg_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}
    gbagg_status status;
};

namespace gpu_aggregate {

inline void Check(gbagg_status st) {
    if (st != GBAGG_OK) throw GpuAggregateError(st, gbagg_last_error());
}

// LogicalTypeId -> the C ABI's type ids. Those are slothdb_type's
// (include/slothdb/api/slothdb.h:31-39), not LogicalTypeId's: DOUBLE is 12
// in slothdb_type and 13 in LogicalTypeId, so the ids are mapped by name.
inline bool ToGpuType(const LogicalType &t, gbagg_type &out) {
    switch (t.id()) {
    case LogicalTypeId::INTEGER: out = GBAGG_TYPE_INTEGER; return true;
    case LogicalTypeId::BIGINT: out = GBAGG_TYPE_BIGINT; return true;
    case LogicalTypeId::DOUBLE: out = GBAGG_TYPE_DOUBLE; return true;
    case LogicalTypeId::VARCHAR: out = GBAGG_TYPE_VARCHAR; return true;
    default: return false;
    }
}

// One bound aggregate (BoundFunction{function_name, arguments, is_distinct}).
struct Aggregate {
    std::string function_name;  // "count_star", "count", "sum", "avg", "min", "max"
    int32_t column = -1;        // input column; -1 for COUNT(*)
    bool is_distinct = false;   // with "count": COUNT(DISTINCT column)
};

inline bool ToGpuAggregate(const Aggregate &a, gbagg_agg &out) {
    const std::string &f = a.function_name;
    out.column = a.column;
    if (f == "count_star" || (f == "count" && a.column < 0)) out.kind = GBAGG_AGG_COUNT_STAR;
    else if (f == "count") out.kind = a.is_distinct ? GBAGG_AGG_COUNT_DISTINCT : GBAGG_AGG_COUNT;
    else if (a.is_distinct) return false;
    else if (f == "sum") out.kind = GBAGG_AGG_SUM;
    else if (f == "avg") out.kind = GBAGG_AGG_AVG;
    else if (f == "min") out.kind = GBAGG_AGG_MIN;
    else if (f == "max") out.kind = GBAGG_AGG_MAX;
    else return false;
    return true;
}

}  // namespace gpu_aggregate

// The compiled WHERE conjuncts in the C ABI's form; the vectors own what the
// gbagg_predicate pointers point at.
struct GpuPredicates {
    std::vector<gbagg_predicate> preds;
    std::vector<std::vector<int64_t>> in_sets;
    std::vector<std::string> svals;
};

// SimplePredicate (physical_planner.cpp:110-135) -> gbagg_predicate, field by
// field: int_in_form -> IN over in_set; like_contains -> LIKE / NOT LIKE
// '%sval%'; str_form -> VARCHAR = / <> sval; otherwise the comparison op with
// the exact int64 literal (ival, read for INTEGER/BIGINT columns) and its
// double form (dval, read for DOUBLE columns). The template takes the
// engine's own SimplePredicate type (file-local to physical_planner.cpp, so
// the planner instantiates it there); its SimpleCmpOp lists LT, LE, GT, GE,
// EQ, NE in the order of gbagg_cmp_op. Returns false for a conjunct the GPU
// kernels cannot evaluate (the planner then keeps the CPU operator).
template <class SimplePredicateT>
bool CapturePredicates(const std::vector<SimplePredicateT> &src, GpuPredicates &out) {
    out = GpuPredicates{};
    if (src.size() > GBAGG_MAX_PREDICATES) return false;
    out.in_sets.reserve(src.size());
    out.svals.reserve(src.size());
    for (const auto &p : src) {
        gbagg_predicate g{};
        g.column = (int32_t)p.col_idx;
        g.ival = p.ival;
        g.dval = p.dval;
        const int op = (int)p.op;
        if (op < 0 || op > (int)GBAGG_CMP_NE) return false;
        if (p.int_in_form) {
            if (p.in_set.size() > GBAGG_MAX_IN_SET) return false;
            g.op = GBAGG_CMP_IN;
            out.in_sets.push_back(p.in_set);
        } else if (p.like_contains) {
            g.op = p.like_negated ? GBAGG_CMP_NOT_LIKE : GBAGG_CMP_LIKE;
            out.svals.push_back(p.sval);
        } else if (p.str_form) {
            if (op != (int)GBAGG_CMP_EQ && op != (int)GBAGG_CMP_NE) return false;
            g.op = (gbagg_cmp_op)op;
            out.svals.push_back(p.sval);
        } else {
            g.op = (gbagg_cmp_op)op;
        }
        out.preds.push_back(g);
    }
    // pointers only once the owning vectors no longer grow
    size_t si = 0, ii = 0;
    for (size_t i = 0; i < src.size(); i++) {
        gbagg_predicate &g = out.preds[i];
        if (g.op == GBAGG_CMP_IN) {
            const auto &s = out.in_sets[ii++];
            g.in_set = s.data();
            g.in_count = (uint32_t)s.size();
        } else if (src[i].like_contains || src[i].str_form) {
            const auto &s = out.svals[si++];
            g.sval = s.data();
            g.slen = (uint32_t)s.size();
        }
    }
    return true;
}

// The planner hook's test (PlanAggregate, physical_planner.cpp:14356-14362):
// every input column the query reads has a type on the path, 0-8 group
// columns, and every aggregate is one of COUNT(*), COUNT, SUM, AVG, MIN, MAX,
// COUNT(DISTINCT). Finer limits (columns per plan, COUNT(DISTINCT) next to
// MIN/MAX over VARCHAR) are checked when the plan is created.
inline bool GpuAggregateSupports(const std::vector<LogicalType> &input_types, const std::vector<idx_t> &groups,
                                 const std::vector<gpu_aggregate::Aggregate> &aggs) {
    if (groups.size() > GBAGG_MAX_GROUP_COLS || aggs.size() > GBAGG_MAX_AGGS) return false;
    if (groups.empty() && aggs.empty()) return false;
    gbagg_type t;
    for (idx_t g : groups)
        if (g >= input_types.size() || !gpu_aggregate::ToGpuType(input_types[g], t)) return false;
    for (const auto &a : aggs) {
        gbagg_agg ga;
        if (!gpu_aggregate::ToGpuAggregate(a, ga)) return false;
        if (ga.column >= 0 && ((size_t)ga.column >= input_types.size() ||
                               !gpu_aggregate::ToGpuType(input_types[ga.column], t)))
            return false;
    }
    return true;
}

class PhysicalGpuHashAggregate : public PhysicalOperator {
public:
    // input_types: the child's output types; groups: group column indexes
    // into them; result_types: groups, then aggregates (the binder's types,
    // src/binder/binder.cpp:1345-1392).
    PhysicalGpuHashAggregate(std::vector<LogicalType> input_types, std::vector<idx_t> groups,
                             std::vector<gpu_aggregate::Aggregate> aggregates, std::vector<LogicalType> result_types,
                             GpuPredicates predicates = {}, std::vector<bool> input_nullable = {})
        : PhysicalOperator(PhysicalOperatorType::TABLE_SCAN, std::move(result_types)),
          input_types_(std::move(input_types)), groups_(std::move(groups)), aggregates_(std::move(aggregates)),
          predicates_(std::move(predicates)), input_nullable_(std::move(input_nullable)) {
        if (!GpuAggregateSupports(input_types_, groups_, aggregates_))
            throw GpuAggregateError(GBAGG_INVALID, "query shape not supported by the GPU aggregate");
        if (GetTypes().size() != groups_.size() + aggregates_.size())
            throw GpuAggregateError(GBAGG_INVALID, "result types must list the groups, then the aggregates");
    }
    ~PhysicalGpuHashAggregate() override {
        if (handle_) gbagg_destroy(handle_);
    }

    void Init() override {
        for (auto &child : children) child->Init();
        if (!handle_) Plan();
        gpu_aggregate::Check(gbagg_reset(handle_));
        finalized_ = false;
        emit_pos_ = 0;
    }

    void SetRowLimit(idx_t n) override { row_limit_ = n; }
    void SetTopNHint(idx_t col_idx, bool ascending, idx_t limit) override {
        topn_col_ = (int64_t)col_idx;
        topn_ascending_ = ascending;
        topn_limit_ = limit;
    }

    // Pulls every chunk of the child (the generic chunk loop,
    // physical_planner.cpp:12291-12516), aggregates on the GPU, then emits
    // VECTOR_SIZE rows per call.
    bool GetData(DataChunk &result) override {
        if (!handle_) Init();
        if (!finalized_) {
            for (auto &child : children) {
                DataChunk chunk;
                while (child->GetData(chunk)) {
                    ConsumeChunk(chunk);
                    chunk = DataChunk();
                }
            }
            gpu_aggregate::Check(gbagg_finalize(handle_));
            gpu_aggregate::Check(gbagg_result(handle_, &view_));
            finalized_ = true;
        }
        if (emit_pos_ >= view_.num_rows) return false;
        const auto &types = GetTypes();
        result.Initialize(types);
        const idx_t n = std::min<idx_t>(VECTOR_SIZE, view_.num_rows - emit_pos_);
        for (idx_t c = 0; c < types.size(); c++) EmitColumn(result.GetVector(c), view_.columns[c], emit_pos_, n);
        result.SetCardinality(n);
        emit_pos_ += n;
        return true;
    }

    // One chunk of the child's rows (FLAT or not; flattened here).
    void ConsumeChunk(DataChunk &chunk) {
        if (!handle_) Init();
        const idx_t n = chunk.size();
        if (!n) return;
        if (chunk.ColumnCount() != input_types_.size())
            throw GpuAggregateError(GBAGG_INVALID, "chunk column count differs from the input types");
        std::vector<gbagg_column_view> views(input_types_.size());
        valid_.resize(input_types_.size());
        offsets_.resize(input_types_.size());
        bytes_.resize(input_types_.size());
        for (idx_t c = 0; c < input_types_.size(); c++) {
            Vector &v = chunk.GetVector(c);
            if (v.GetVectorType() != VectorType::FLAT) v.Flatten(n);
            gbagg_column_view &w = views[c];
            w = gbagg_column_view{};
            gpu_aggregate::ToGpuType(input_types_[c], w.type);
            const ValidityMask &m = v.GetValidity();
            if (!m.AllValid()) {
                valid_[c].resize(n);
                for (idx_t r = 0; r < n; r++) valid_[c][r] = m.RowIsValid(r) ? 1 : 0;
                w.validity = valid_[c].data();
            }
            if (w.type == GBAGG_TYPE_VARCHAR) {
                // string_t cells packed once into offsets + bytes (GBAGG_STR_RAW)
                const string_t *s = v.GetData<string_t>();
                std::vector<uint32_t> &off = offsets_[c];
                std::vector<char> &by = bytes_[c];
                off.resize(n + 1);
                size_t total = 0;
                for (idx_t r = 0; r < n; r++) total += m.RowIsValid(r) ? s[r].GetSize() : 0;
                if (total > 0xFFFFFFFFull) throw GpuAggregateError(GBAGG_INVALID, "chunk strings exceed 4 GiB");
                by.resize(std::max<size_t>(total, 1));
                uint32_t pos = 0;
                for (idx_t r = 0; r < n; r++) {
                    off[r] = pos;
                    if (!m.RowIsValid(r)) continue;
                    const uint32_t len = s[r].GetSize();
                    std::memcpy(by.data() + pos, s[r].GetData(), len);
                    pos += len;
                }
                off[n] = pos;
                w.data = off.data();
                w.dict_bytes = by.data();
                w.flags = GBAGG_STR_RAW;
            } else {
                w.data = v.GetData();
            }
        }
        gpu_aggregate::Check(gbagg_consume(handle_, views.data(), n));
    }

    gbagg_plan_info PlanInfo() const {
        gbagg_plan_info info{};
        if (handle_) gpu_aggregate::Check(gbagg_plan_info_get(handle_, &info));
        return info;
    }

private:
    void Plan() {
        std::vector<gbagg_type> types(input_types_.size());
        std::vector<uint8_t> nullable(input_types_.size(), 1);
        for (size_t c = 0; c < types.size(); c++) {
            gpu_aggregate::ToGpuType(input_types_[c], types[c]);
            if (c < input_nullable_.size()) nullable[c] = input_nullable_[c] ? 1 : 0;
        }
        std::vector<uint32_t> groups(groups_.begin(), groups_.end());
        std::vector<gbagg_agg> aggs(aggregates_.size());
        for (size_t a = 0; a < aggs.size(); a++) gpu_aggregate::ToGpuAggregate(aggregates_[a], aggs[a]);
        gbagg_query_desc q{};
        q.num_columns = (uint32_t)types.size();
        q.column_types = types.data();
        q.column_nullable = nullable.data();
        q.num_group_cols = (uint32_t)groups.size();
        q.group_cols = groups.data();
        q.num_aggs = (uint32_t)aggs.size();
        q.aggs = aggs.data();
        q.num_predicates = (uint32_t)predicates_.preds.size();
        q.predicates = predicates_.preds.data();
        q.topn_col = topn_limit_ ? (int32_t)topn_col_ : -1;
        q.topn_ascending = topn_ascending_ ? 1 : 0;
        q.topn_limit = topn_limit_;
        q.row_limit = row_limit_;
        q.drop_null_keys = 0;  // SQL: NULL keys form one group (the generic path, physical_planner.cpp:6187)
        gpu_aggregate::Check(gbagg_plan_create(&q, &handle_));
        // the binder's result types against the operator's
        for (size_t c = 0; c < GetTypes().size(); c++) {
            gbagg_type want;
            if (!gpu_aggregate::ToGpuType(GetTypes()[c], want))
                throw GpuAggregateError(GBAGG_INVALID, "result type not on the GPU path");
            result_types_.push_back(want);
        }
    }

    // Rows [first, first + n) of one result column into an engine Vector.
    static void EmitColumn(Vector &out, const gbagg_result_column &col, idx_t first, idx_t n) {
        ValidityMask &m = out.GetValidity();
        for (idx_t r = 0; r < n; r++)
            if (!col.validity[first + r]) m.SetInvalid(r);
        switch (col.type) {
        case GBAGG_TYPE_INTEGER:
            std::memcpy(out.GetData<int32_t>(), (const int32_t *)col.data + first, n * 4);
            break;
        case GBAGG_TYPE_BIGINT:
            std::memcpy(out.GetData<int64_t>(), (const int64_t *)col.data + first, n * 8);
            break;
        case GBAGG_TYPE_DOUBLE:
            std::memcpy(out.GetData<double>(), (const double *)col.data + first, n * 8);
            break;
        default: {  // VARCHAR: string_t cells over the vector's own string buffer
            string_t *dst = out.GetData<string_t>();
            VectorStringBuffer &heap = out.GetStringBuffer();
            for (idx_t r = 0; r < n; r++) {
                if (!col.validity[first + r]) {
                    dst[r] = string_t();
                    continue;
                }
                const uint32_t a = col.str_offsets[first + r], b = col.str_offsets[first + r + 1];
                const char *p = heap.AddString(col.str_bytes + a, b - a);
                dst[r] = string_t(p, b - a);
            }
            break;
        }
        }
    }

    std::vector<LogicalType> input_types_;
    std::vector<idx_t> groups_;
    std::vector<gpu_aggregate::Aggregate> aggregates_;
    GpuPredicates predicates_;
    std::vector<bool> input_nullable_;
    std::vector<gbagg_type> result_types_;
    int64_t topn_col_ = -1;
    bool topn_ascending_ = false;
    idx_t topn_limit_ = 0;
    idx_t row_limit_ = 0;

    gbagg_handle *handle_ = nullptr;
    bool finalized_ = false;
    gbagg_result_view view_{};
    idx_t emit_pos_ = 0;
    // per-column staging of the chunk being consumed (validity bytes, strings)
    std::vector<std::vector<uint8_t>> valid_;
    std::vector<std::vector<uint32_t>> offsets_;
    std::vector<std::vector<char>> bytes_;
};

}  // namespace slothdb
