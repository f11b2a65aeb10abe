// gpu_hash_aggregate.cpp -- PhysicalGpuHashAggregate over the gbagg C ABI.
#include "gpu_hash_aggregate.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>

namespace slothdb_gpu {

namespace {

void check(gbagg_status st) {
    if (st != GBAGG_OK) throw GpuAggregateException(st, gbagg_last_error());
}

size_t type_width(LogicalTypeId t) {
    switch (t) {
    case LogicalTypeId::INTEGER: return 4;
    case LogicalTypeId::VARCHAR: return 0;
    default: return 8;
    }
}

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

gbagg_agg_kind agg_kind(const AggregateSpec &a) {
    const std::string f = lower(a.function_name);
    if (f == "count_star" || (f == "count" && a.arg < 0)) return GBAGG_AGG_COUNT_STAR;
    if (f == "count") return a.is_distinct ? GBAGG_AGG_COUNT_DISTINCT : GBAGG_AGG_COUNT;
    if (a.is_distinct) throw GpuAggregateException(GBAGG_INVALID, "DISTINCT is supported for COUNT only");
    if (f == "sum") return GBAGG_AGG_SUM;
    if (f == "avg" || f == "mean") return GBAGG_AGG_AVG;
    if (f == "min") return GBAGG_AGG_MIN;
    if (f == "max") return GBAGG_AGG_MAX;
    throw GpuAggregateException(GBAGG_INVALID, "aggregate function not on the GPU path: " + a.function_name);
}

}  // namespace

// ---- columnar types ---------------------------------------------------------

void ValidityMask::SetInvalid(idx_t row) {
    if (words_.empty()) words_.assign((std::max(capacity_, row + 1) + 63) / 64, ~0ULL);
    if (row / 64 >= words_.size()) words_.resize(row / 64 + 1, ~0ULL);
    words_[row / 64] &= ~(1ULL << (row % 64));
}

void ValidityMask::SetValid(idx_t row) {
    if (words_.empty()) return;
    if (row / 64 >= words_.size()) words_.resize(row / 64 + 1, ~0ULL);
    words_[row / 64] |= 1ULL << (row % 64);
}

Vector::Vector(LogicalTypeId type, idx_t capacity) : type_(type) {
    if (type == LogicalTypeId::VARCHAR) strings_.resize(capacity);
    else data_.resize(capacity * type_width(type));
    validity_.Resize(capacity);
}

void DataChunk::Initialize(const std::vector<LogicalTypeId> &types, idx_t capacity) {
    data.clear();
    for (auto t : types) data.emplace_back(t, capacity);
    count_ = 0;
}

void DataChunk::Reset() {
    for (auto &v : data) v.Validity().Reset();
    count_ = 0;
}

// ---- operator -----------------------------------------------------------------

PhysicalGpuHashAggregate::PhysicalGpuHashAggregate(std::vector<LogicalTypeId> input_types, std::vector<idx_t> groups,
                                                   std::vector<AggregateSpec> aggregates,
                                                   std::vector<PredicateSpec> predicates,
                                                   std::vector<bool> input_nullable)
    : input_types_(std::move(input_types)), groups_(std::move(groups)), aggregates_(std::move(aggregates)),
      predicates_(std::move(predicates)) {
    nullable_.assign(input_types_.size(), 1);
    for (size_t i = 0; i < input_nullable.size() && i < nullable_.size(); i++) nullable_[i] = input_nullable[i];
    for (auto g : groups_)
        if (g >= input_types_.size()) throw GpuAggregateException(GBAGG_INVALID, "group column out of range");
    for (auto &a : aggregates_) agg_kind(a);  // reject unsupported functions at construction
}

PhysicalGpuHashAggregate::~PhysicalGpuHashAggregate() {
    if (handle_) gbagg_destroy(handle_);
}

std::vector<LogicalTypeId> PhysicalGpuHashAggregate::GetTypes() const {
    // group columns keep their type; aggregate types per binder.cpp:1345-1392
    std::vector<LogicalTypeId> out;
    for (auto g : groups_) out.push_back(input_types_[g]);
    for (auto &a : aggregates_) {
        switch (agg_kind(a)) {
        case GBAGG_AGG_COUNT_STAR:
        case GBAGG_AGG_COUNT:
        case GBAGG_AGG_COUNT_DISTINCT: out.push_back(LogicalTypeId::BIGINT); break;
        case GBAGG_AGG_SUM:
            out.push_back(input_types_.at(a.arg) == LogicalTypeId::DOUBLE ? LogicalTypeId::DOUBLE
                                                                           : LogicalTypeId::BIGINT);
            break;
        case GBAGG_AGG_AVG: out.push_back(LogicalTypeId::DOUBLE); break;
        default: out.push_back(input_types_.at(a.arg)); break;
        }
    }
    return out;
}

void PhysicalGpuHashAggregate::SetTopNHint(idx_t col_idx, bool ascending, idx_t limit) {
    topn_col_ = (int64_t)col_idx;
    topn_ascending_ = ascending;
    topn_limit_ = limit;
    if (handle_) {  // hints arrive from the parent's Init: re-plan at the next Init
        gbagg_destroy(handle_);
        handle_ = nullptr;
    }
}

void PhysicalGpuHashAggregate::SetRowLimit(idx_t limit) {
    row_limit_ = limit;
    if (handle_) {
        gbagg_destroy(handle_);
        handle_ = nullptr;
    }
}

void PhysicalGpuHashAggregate::SetNeededOutputs(const std::vector<bool> &) {
    // every output column is produced; a pruned column is simply not read
}

void PhysicalGpuHashAggregate::SetParam(const std::string &name, int64_t value) {
    params_.emplace_back(name, value);
    if (handle_) check(gbagg_set_param(handle_, name.c_str(), value));
}

void PhysicalGpuHashAggregate::Plan() {
    std::vector<gbagg_type> types;
    for (auto t : input_types_) types.push_back((gbagg_type)t);
    std::vector<uint32_t> groups(groups_.begin(), groups_.end());
    std::vector<gbagg_agg> aggs;
    for (auto &a : aggregates_) aggs.push_back(gbagg_agg{agg_kind(a), a.arg});
    std::vector<gbagg_predicate> preds;
    for (auto &p : predicates_) {
        gbagg_predicate g{};
        g.column = p.column;
        g.op = p.op;
        g.ival = p.ival;
        g.dval = p.dval;
        g.in_set = p.in_set.empty() ? nullptr : p.in_set.data();
        g.in_count = (uint32_t)p.in_set.size();
        g.sval = p.sval.empty() ? nullptr : p.sval.data();
        g.slen = (uint32_t)p.sval.size();
        preds.push_back(g);
    }
    gbagg_query_desc q{};
    q.num_columns = (uint32_t)types.size();
    q.column_types = types.data();
    q.column_nullable = nullable_.data();
    q.num_group_cols = (uint32_t)groups.size();
    q.group_cols = groups.data();
    q.num_aggs = (uint32_t)aggs.size();
    q.aggs = aggs.data();
    q.num_predicates = (uint32_t)preds.size();
    q.predicates = preds.data();
    q.topn_col = (int32_t)topn_col_;
    q.topn_ascending = topn_ascending_ ? 1 : 0;
    q.topn_limit = topn_col_ >= 0 ? topn_limit_ : 0;
    q.row_limit = row_limit_;
    q.drop_null_keys = drop_null_keys_ ? 1 : 0;
    q.strategy = strategy_;
    gbagg_handle *h = nullptr;
    check(gbagg_plan_create(&q, &h));
    handle_ = h;
    for (auto &[name, value] : params_) check(gbagg_set_param(handle_, name.c_str(), value));
}

void PhysicalGpuHashAggregate::Init() {
    if (!handle_) Plan();
    else check(gbagg_reset(handle_));
    finalized_ = false;
    emit_pos_ = 0;
    result_ = gbagg_result_view{};
}

void PhysicalGpuHashAggregate::ConsumeRowGroup(const ColumnBatch &batch) {
    if (!handle_) Init();
    if (batch.columns.size() != input_types_.size())
        throw GpuAggregateException(GBAGG_INVALID, "row group has the wrong number of columns");
    std::vector<gbagg_column_view> views(batch.columns.size());
    bool device = false;
    for (size_t c = 0; c < batch.columns.size(); c++) {
        const ColumnData &d = batch.columns[c];
        if (c == 0) device = d.on_device;
        else if (d.on_device != device)
            throw GpuAggregateException(GBAGG_INVALID, "row group mixes host and device columns");
        views[c] = gbagg_column_view{(gbagg_type)d.type, d.values, d.validity, d.dict_offsets, d.dict_bytes,
                                     d.dict_size, d.dict_unique ? GBAGG_DICT_UNIQUE : 0u};
    }
    if (device) check(gbagg_consume_device(handle_, views.data(), batch.num_rows));
    else check(gbagg_consume(handle_, views.data(), batch.num_rows));
}

void PhysicalGpuHashAggregate::ConsumeChunk(const DataChunk &chunk) {
    if (!handle_) Init();
    if (chunk.ColumnCount() != input_types_.size())
        throw GpuAggregateException(GBAGG_INVALID, "chunk has the wrong number of columns");
    const idx_t n = chunk.size();
    if (n == 0) return;
    // flat typed arrays pass as they are; validity bits become bytes; VARCHAR
    // strings are flattened to offsets + bytes, which the GPU interns
    std::vector<std::vector<uint8_t>> valid(chunk.ColumnCount());
    std::vector<std::vector<uint32_t>> offsets(chunk.ColumnCount());
    std::vector<std::string> bytes(chunk.ColumnCount());
    std::vector<gbagg_column_view> views(chunk.ColumnCount());
    for (size_t c = 0; c < chunk.ColumnCount(); c++) {
        const Vector &v = chunk.data[c];
        if (v.GetType() != input_types_[c])
            throw GpuAggregateException(GBAGG_INVALID, "chunk column type differs from the plan");
        gbagg_column_view &g = views[c];
        g.type = (gbagg_type)v.GetType();
        if (!v.Validity().AllValid()) {
            if (!nullable_[c]) throw GpuAggregateException(GBAGG_INVALID, "NULL in a column planned as not nullable");
            valid[c].resize(n);
            for (idx_t r = 0; r < n; r++) valid[c][r] = v.Validity().RowIsValid(r) ? 1 : 0;
            g.validity = valid[c].data();
        }
        if (v.GetType() == LogicalTypeId::VARCHAR) {
            offsets[c].resize(n + 1);
            offsets[c][0] = 0;
            for (idx_t r = 0; r < n; r++) {
                if (v.Validity().RowIsValid(r)) bytes[c].append(v.Strings()[r]);
                offsets[c][r + 1] = (uint32_t)bytes[c].size();
            }
            g.data = offsets[c].data();
            g.dict_bytes = bytes[c].data();
            g.flags = GBAGG_STR_RAW;
        } else {
            g.data = v.GetData<uint8_t>();
        }
    }
    check(gbagg_consume(handle_, views.data(), n));
}

void PhysicalGpuHashAggregate::Finalize() {
    check(gbagg_finalize(handle_));
    check(gbagg_result(handle_, &result_));
    finalized_ = true;
    emit_pos_ = 0;
}

bool PhysicalGpuHashAggregate::GetData(DataChunk &chunk) {
    if (!handle_) Init();
    if (!finalized_) Finalize();
    const auto types = GetTypes();
    if (chunk.ColumnCount() != types.size()) chunk.Initialize(types);
    chunk.Reset();
    if (emit_pos_ >= result_.num_rows) return false;
    const idx_t n = std::min<idx_t>(VECTOR_SIZE, result_.num_rows - emit_pos_);
    for (size_t c = 0; c < types.size(); c++) {
        const gbagg_result_column &rc = result_.columns[c];
        Vector &v = chunk.data[c];
        for (idx_t r = 0; r < n; r++)
            if (!rc.validity[emit_pos_ + r]) v.Validity().SetInvalid(r);
        switch (types[c]) {
        case LogicalTypeId::INTEGER:
            std::memcpy(v.GetData<int32_t>(), (const int32_t *)rc.data + emit_pos_, n * 4);
            break;
        case LogicalTypeId::VARCHAR:
            for (idx_t r = 0; r < n; r++) {
                const uint32_t a = rc.str_offsets[emit_pos_ + r], b = rc.str_offsets[emit_pos_ + r + 1];
                v.Strings()[r].assign(rc.str_bytes + a, b - a);
            }
            break;
        default:
            std::memcpy(v.GetData<uint8_t>(), (const uint8_t *)rc.data + emit_pos_ * 8, n * 8);
            break;
        }
    }
    chunk.SetCardinality(n);
    emit_pos_ += n;
    return true;
}

gbagg_plan_info PhysicalGpuHashAggregate::PlanInfo() const {
    gbagg_plan_info info{};
    if (handle_) check(gbagg_plan_info_get(handle_, &info));
    return info;
}

}  // namespace slothdb_gpu
