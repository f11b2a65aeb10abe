// gpu_hash_aggregate.hpp -- C++20 host side of the MI355X GROUP BY operator.
//
// PhysicalGpuHashAggregate has the member surface of the reference's
// aggregate operator (class PhysicalOperator,
// include/slothdb/execution/physical_operator.hpp:23-65; the aggregate
// PhysicalHashAggregate, src/execution/physical_planner.cpp:6009-6060):
//   Init(), GetData(DataChunk&), SetTopNHint(col, asc, limit), SetRowLimit(n),
//   SetNeededOutputs(mask), GetTypes()
// and the two ways the engine feeds an aggregate:
//   ConsumeRowGroup(ColumnBatch)  the fused-path RGConsumerFn callback
//                                 (physical_planner.cpp:1060-1062) over
//                                 ParquetColumnData-like typed arrays
//   ConsumeChunk(DataChunk)       the generic child-pull chunk loop
//                                 (physical_planner.cpp:12291-12516)
// It calls the HIP kernels only through the C ABI in include/slothdb_gpu.h.
//
// The columnar types below (ValidityMask, Vector, DataChunk) follow the
// reference's contract (include/slothdb/common/types/data_chunk.hpp:10-55,
// vector.hpp:65-147, validity_mask.hpp:9-69): FLAT vectors of at most
// VECTOR_SIZE rows, validity as 64-bit words LSB-first with "no mask" meaning
// all valid. In the engine, the operator would use the engine's own types;
// only the conversion helpers at the bottom of the .cpp change.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/slothdb_gpu.h"

namespace slothdb_gpu {

using idx_t = uint64_t;
constexpr idx_t VECTOR_SIZE = 2048;  // include/slothdb/common/constants.hpp

// Same values as slothdb's LogicalTypeId for the types on the path.
enum class LogicalTypeId : uint8_t { INTEGER = 5, BIGINT = 6, DOUBLE = 12, VARCHAR = 15 };

// Thrown for every failure below the operator (the engine's InternalException
// role, include/slothdb/common/exception.hpp:90-180).
class GpuAggregateException : public std::runtime_error {
public:
    GpuAggregateException(gbagg_status status, const std::string &msg)
        : std::runtime_error(msg), status(status) {}
    gbagg_status status;
};

class ValidityMask {
public:
    bool AllValid() const { return words_.empty(); }
    bool RowIsValid(idx_t row) const { return words_.empty() || ((words_[row / 64] >> (row % 64)) & 1); }
    void SetInvalid(idx_t row);
    void SetValid(idx_t row);
    void Reset() { words_.clear(); }
    void Resize(idx_t rows) { capacity_ = rows; }

private:
    std::vector<uint64_t> words_;  // empty: every row valid
    idx_t capacity_ = VECTOR_SIZE;
};

// A FLAT vector: typed array of `capacity` rows (VARCHAR as std::string).
class Vector {
public:
    explicit Vector(LogicalTypeId type = LogicalTypeId::BIGINT, idx_t capacity = VECTOR_SIZE);
    LogicalTypeId GetType() const { return type_; }
    template <class T>
    T *GetData() { return reinterpret_cast<T *>(data_.data()); }
    template <class T>
    const T *GetData() const { return reinterpret_cast<const T *>(data_.data()); }
    std::vector<std::string> &Strings() { return strings_; }
    const std::vector<std::string> &Strings() const { return strings_; }
    ValidityMask &Validity() { return validity_; }
    const ValidityMask &Validity() const { return validity_; }

private:
    LogicalTypeId type_;
    std::vector<uint8_t> data_;
    std::vector<std::string> strings_;
    ValidityMask validity_;
};

class DataChunk {
public:
    void Initialize(const std::vector<LogicalTypeId> &types, idx_t capacity = VECTOR_SIZE);
    void Reset();
    idx_t size() const { return count_; }
    void SetCardinality(idx_t n) { count_ = n; }
    idx_t ColumnCount() const { return data.size(); }
    std::vector<Vector> data;

private:
    idx_t count_ = 0;
};

// One row group of the fused scan path (ParquetColumnData,
// include/slothdb/storage/parquet.hpp:147-207): typed arrays, a validity byte
// per row or none, VARCHAR as dictionary codes plus the dictionary.
struct ColumnData {
    LogicalTypeId type = LogicalTypeId::BIGINT;
    const void *values = nullptr;        // int32 / int64 / double / uint32 dictionary codes
    const uint8_t *validity = nullptr;   // nullptr = all valid
    const uint32_t *dict_offsets = nullptr;  // VARCHAR: dict_size + 1 offsets into dict_bytes
    const char *dict_bytes = nullptr;
    uint32_t dict_size = 0;
    bool dict_unique = false;            // entries pairwise distinct
    bool on_device = false;              // values/validity already in HBM
};
struct ColumnBatch {
    std::vector<ColumnData> columns;
    idx_t num_rows = 0;
};

// Bound aggregate (BoundFunction{function_name, arguments, is_distinct}):
// "count_star", "count", "sum", "avg", "min", "max"; arg = input column, -1
// for COUNT(*); is_distinct with "count" is COUNT(DISTINCT arg).
struct AggregateSpec {
    std::string function_name;
    int32_t arg = -1;
    bool is_distinct = false;
};

// One compiled WHERE conjunct (SimplePredicate, physical_planner.cpp:110-135).
struct PredicateSpec {
    int32_t column = 0;
    gbagg_cmp_op op = GBAGG_CMP_EQ;
    int64_t ival = 0;
    double dval = 0.0;
    std::vector<int64_t> in_set;
    std::string sval;
};

class PhysicalGpuHashAggregate {
public:
    PhysicalGpuHashAggregate(std::vector<LogicalTypeId> input_types, std::vector<idx_t> groups,
                             std::vector<AggregateSpec> aggregates, std::vector<PredicateSpec> predicates = {},
                             std::vector<bool> input_nullable = {});
    ~PhysicalGpuHashAggregate();
    PhysicalGpuHashAggregate(const PhysicalGpuHashAggregate &) = delete;
    PhysicalGpuHashAggregate &operator=(const PhysicalGpuHashAggregate &) = delete;

    // PhysicalOperator surface
    void Init();
    bool GetData(DataChunk &chunk);
    void SetTopNHint(idx_t col_idx, bool ascending, idx_t limit);
    void SetRowLimit(idx_t limit);
    void SetNeededOutputs(const std::vector<bool> &needed);
    std::vector<LogicalTypeId> GetTypes() const;  // groups, then aggregates

    // Inputs
    void ConsumeRowGroup(const ColumnBatch &batch);
    void ConsumeChunk(const DataChunk &chunk);

    // Planner knobs (gbagg_set_param) and what the last run chose.
    void SetParam(const std::string &name, int64_t value);
    gbagg_plan_info PlanInfo() const;
    void SetStrategy(gbagg_strategy s) { strategy_ = s; }
    void SetDropNullKeys(bool drop) { drop_null_keys_ = drop; }

private:
    void Plan();
    void Finalize();

    std::vector<LogicalTypeId> input_types_;
    std::vector<idx_t> groups_;
    std::vector<AggregateSpec> aggregates_;
    std::vector<PredicateSpec> predicates_;
    std::vector<uint8_t> nullable_;
    int64_t topn_col_ = -1;
    bool topn_ascending_ = false;
    idx_t topn_limit_ = 0;
    idx_t row_limit_ = 0;
    bool drop_null_keys_ = false;
    gbagg_strategy strategy_ = GBAGG_STRATEGY_AUTO;
    std::vector<std::pair<std::string, int64_t>> params_;

    gbagg_handle *handle_ = nullptr;
    bool finalized_ = false;
    gbagg_result_view result_{};
    idx_t emit_pos_ = 0;
};

}  // namespace slothdb_gpu
