/*
 * slothdb_gpu.h -- C ABI of the MI355X GROUP BY hash-aggregate operator.
 *
 * This is the drop-in boundary for SlothDB's GROUP BY hot path: the
 * PhysicalHashAggregate operator (src/execution/physical_planner.cpp:6009-12627)
 * together with the fused WHERE filter that feeds it (SimplePredicate,
 * physical_planner.cpp:110-135, 180-327, 396-496, 505-720) and the TopN pushdown
 * that consumes it (PhysicalOperator::SetTopNHint / SetRowLimit,
 * include/slothdb/execution/physical_operator.hpp:44-58).
 *
 * All signatures use plain pointers and sizes. Host-side callers (a C++
 * PhysicalOperator subclass, a ctypes/cgo/JNI binding) own their input
 * buffers; the library owns its device buffers and the host result buffers
 * it hands back, which stay valid until the next gbagg_finalize/gbagg_reset
 * or gbagg_destroy on the same handle.
 *
 * Threading: a handle is single-owner (one HIP stream per handle), matching
 * the reference's single-threaded GetData contract. Different handles may be
 * used from different threads.
 *
 * Errors: every call returns a gbagg_status; on failure the message is kept
 * in thread-local storage and returned by gbagg_last_error(). This mirrors
 * the C API's slothdb_status {OK, ERROR, INVALID} + result->error convention
 * (include/slothdb/api/slothdb.h:24-28; src/api/slothdb_api.cpp:88-94).
 */
#ifndef SLOTHDB_GPU_H
#define SLOTHDB_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBAGG_ABI_VERSION 1

/* Limits of one plan. */
#define GBAGG_MAX_COLUMNS 16
#define GBAGG_MAX_GROUP_COLS 8
#define GBAGG_MAX_AGGS 16
#define GBAGG_MAX_PREDICATES 8
#define GBAGG_MAX_IN_SET 64
#define GBAGG_MAX_TOPN 1024

/* Status codes: same values as slothdb_status (include/slothdb/api/slothdb.h:24-28). */
typedef enum {
    GBAGG_OK = 0,
    GBAGG_ERROR = 1,
    GBAGG_INVALID = 2,
} gbagg_status;

/* Column types: same values as slothdb_type (include/slothdb/api/slothdb.h:31-39).
 * slothdb_type says it matches LogicalTypeId, but the engine's LogicalTypeId
 * numbers FLOAT 12 and DOUBLE 13 (include/slothdb/common/types/logical_type.hpp),
 * so engine-side callers map LogicalTypeId to these by name, not by value
 * (slothdb_amd/host/engine/physical_gpu_hash_aggregate.hpp). VARCHAR columns arrive dictionary
 * encoded, as ParquetColumnData::str_dict_indices + str_dict_values
 * (include/slothdb/storage/parquet.hpp:166-168). */
typedef enum {
    GBAGG_TYPE_INTEGER = 5,
    GBAGG_TYPE_BIGINT = 6,
    GBAGG_TYPE_DOUBLE = 12,
    GBAGG_TYPE_VARCHAR = 15,
} gbagg_type;

/* One input column of one batch of rows. Mirrors the typed-array contract of
 * ParquetColumnData (include/slothdb/storage/parquet.hpp:147-207):
 *   data      INTEGER int32[n], BIGINT int64[n], DOUBLE double[n],
 *             VARCHAR uint32 dictionary codes[n]
 *   validity  one byte per row, 0 = NULL (ParquetColumnData::validity);
 *             a null pointer means all rows are valid (all_valid = true)
 * For gbagg_consume these are host pointers; for gbagg_consume_device,
 * data and validity are device pointers. The dictionary (VARCHAR only) is
 * always host memory: dict_offsets has dict_size + 1 entries into dict_bytes.
 * Codes must be < dict_size. */
typedef struct {
    gbagg_type type;
    const void *data;
    const uint8_t *validity;
    const uint32_t *dict_offsets;
    const char *dict_bytes;
    uint32_t dict_size;
    uint32_t flags; /* GBAGG_DICT_* */
} gbagg_column_view;

/* The dictionary's entries are pairwise distinct, so equal codes mean equal
 * strings and codes can be grouped without canonicalisation. When every batch
 * passes the same dictionary (pointer and size) with this flag, the operator
 * skips the per-batch host pass over the dictionary. */
#define GBAGG_DICT_UNIQUE 1u

/* A VARCHAR column given as its strings, one per row, instead of dictionary
 * codes: the string_t vectors of a DataChunk
 * (include/slothdb/common/types/string_type.hpp:19-96) flattened to
 *   data        uint32 offsets[n + 1], ascending, into dict_bytes
 *   dict_bytes  the strings' bytes; row r is [offsets[r], offsets[r + 1])
 *   dict_offsets = NULL, dict_size = 0
 * data and dict_bytes follow the batch's pointer kind (host for
 * gbagg_consume, device for gbagg_consume_device). The operator interns every
 * distinct string of the column on the GPU (exact byte comparison) and groups,
 * filters and counts by its code, as for a dictionary-encoded column; this
 * replaces the reference's non-dictionary VARCHAR group-key path
 * (RadixHashCountStr, src/execution/radix_count_agg.cpp:390-761). Batches of
 * one column may mix the two forms (row groups whose dictionary fitted next to
 * row groups that fell back to PLAIN strings): a group column is then interned
 * over all its batches, the coded ones read through their dictionaries. */
#define GBAGG_STR_RAW 2u

/* Aggregate functions on the path (agg_emit_helpers.cpp:8-98; return types
 * per src/binder/binder.cpp:1345-1392):
 *   COUNT(*)  BIGINT, never NULL
 *   COUNT(c)  BIGINT, non-NULL inputs
 *   SUM(c)    BIGINT for INTEGER/BIGINT c, DOUBLE for DOUBLE c; NULL if no input
 *   AVG(c)    DOUBLE; NULL if no input
 *   MIN/MAX   type of c; NULL if no input. Over VARCHAR, byte order of the
 *             strings (the generic path's Value comparison); the column is
 *             interned at finalize with codes relabelled in that order.
 *             Not next to COUNT(DISTINCT).
 *   COUNT(DISTINCT c)  BIGINT, distinct non-NULL values of c, of any type
 *             (the generic path's distinct_set,
 *             physical_planner.cpp:12325-12330). Runs in two phases: GROUP BY
 *             (group columns, c) with partial states of the other aggregates
 *             (rows with a NULL c form their own groups there), then GROUP BY
 *             the group columns over those rows, counting non-NULL c; a group
 *             whose c is always NULL counts 0. Each DISTINCT argument column
 *             of a query runs its own phase 1 (the other aggregates' partial
 *             states ride on the first) and phase 2 counts each over its own
 *             phase 1's groups (phase 2 loads the group key, one column per
 *             argument and one per partial state: at most 8 in all, as any
 *             plan); at most 7 group columns, of any type, with
 *             NULL keys grouped or dropped as for other queries. DOUBLE values
 *             are distinct by value (-0.0 equals 0.0; all NaNs are one value),
 *             where the reference compares their shortest decimal text.
 * Integer SUM is exact two's-complement int64 (the reference accumulates in
 * double and casts; the two agree while partial sums stay below 2^53). */
typedef enum {
    GBAGG_AGG_COUNT_STAR = 0,
    GBAGG_AGG_COUNT = 1,
    GBAGG_AGG_SUM = 2,
    GBAGG_AGG_AVG = 3,
    GBAGG_AGG_MIN = 4,
    GBAGG_AGG_MAX = 5,
    GBAGG_AGG_COUNT_DISTINCT = 6,
} gbagg_agg_kind;

typedef struct {
    gbagg_agg_kind kind;
    int32_t column; /* input column index; -1 for COUNT(*) */
} gbagg_agg;

/* Compiled simple predicate, one conjunct of the WHERE clause. Same fields
 * and meaning as SimplePredicate (physical_planner.cpp:110-135), filled the
 * way TryCompileSimplePredicateImpl does (:180-312):
 *   INTEGER/BIGINT column, LT..NE   compare the int64 value with ival
 *   DOUBLE column, LT..NE           compare the double value with dval
 *   INTEGER/BIGINT column, IN       value is in in_set[0..in_count)
 *   VARCHAR column, EQ/NE           string (not) equal to sval[0..slen)
 *   VARCHAR column, LIKE/NOT_LIKE   LIKE '%sval%' (case-sensitive contains)
 * A NULL in a predicate column drops the row (EvalSimplePredicates, :396-403). */
typedef enum {
    GBAGG_CMP_LT = 0,
    GBAGG_CMP_LE = 1,
    GBAGG_CMP_GT = 2,
    GBAGG_CMP_GE = 3,
    GBAGG_CMP_EQ = 4,
    GBAGG_CMP_NE = 5,
    GBAGG_CMP_IN = 6,
    GBAGG_CMP_LIKE = 7,
    GBAGG_CMP_NOT_LIKE = 8,
} gbagg_cmp_op;

typedef struct {
    int32_t column;
    gbagg_cmp_op op;
    int64_t ival;
    double dval;
    const int64_t *in_set;
    uint32_t in_count;
    const char *sval;
    uint32_t slen;
} gbagg_predicate;

/* Execution strategies. AUTO lets the planner (HyperLogLog estimate, or the
 * learned selector when one is loaded) pick. */
typedef enum {
    GBAGG_STRATEGY_AUTO = 0,
    GBAGG_STRATEGY_DIRECT = 1,      /* one-column key over a narrow exact range: one record
                                       per key value, no hashing (IntGroupAgg's dense
                                       slot_of[key], int_group_agg.cpp:69-277); a range
                                       wider than 2^22 runs the AUTO plan instead */
    GBAGG_STRATEGY_LDS_HASH = 2,    /* per-workgroup LDS hash table + global merge */
    GBAGG_STRATEGY_GLOBAL_HASH = 3, /* global hash table only */
    GBAGG_STRATEGY_RADIX = 4,       /* radix partition + per-partition LDS tables */
} gbagg_strategy;

/* The planned aggregate: what PhysicalPlanner::PlanAggregate
 * (physical_planner.cpp:14356-14362) hands PhysicalHashAggregate, after the
 * planner's Filter->Scan peek-through has compiled the WHERE clause
 * (:7138-7161) and the OrderBy/Limit above it has forwarded its hints
 * (:1958-1975, :4026-4031).
 * Output columns are the group columns in order, then the aggregates in order
 * (the TopN col_idx space of SetTopNHint, physical_planner.cpp:6017-6027). */
typedef struct {
    uint32_t num_columns;
    const gbagg_type *column_types;
    /* Per input column: 0 = the caller guarantees no NULLs in any batch
     * (validity must then be a null pointer), 1 = may contain NULLs.
     * A null pointer here means every column may contain NULLs. */
    const uint8_t *column_nullable;
    /* 0-8 group columns. One INTEGER/BIGINT/VARCHAR column is the key word;
     * two columns with a BIGINT among them form a 16-byte key
     * (InlineRowAggBigKey / RadixMultiAggBigKey, inline_row_agg.cpp:269-517,
     * radix_multi_agg.cpp:218-225); two 32-bit columns pack into 8 bytes.
     * Keys those forms cannot hold (more than two columns, a DOUBLE column,
     * or two columns whose NULLs form groups; the generic paths' row keys,
     * physical_planner.cpp:12291-12516, count_group_agg.cpp:32-227) are
     * packed per row and interned on the GPU into one code; such a plan takes
     * at most 15 columns (the code is one more) and no TopN hint on a group
     * column. DOUBLE keys group -0.0 with 0.0 and all NaNs together. */
    uint32_t num_group_cols;
    const uint32_t *group_cols;
    uint32_t num_aggs;
    const gbagg_agg *aggs;
    uint32_t num_predicates;
    const gbagg_predicate *predicates;
    /* SetTopNHint(col_idx, ascending, limit): topn_col < 0 disables. */
    int32_t topn_col;
    int32_t topn_ascending;
    uint64_t topn_limit;
    /* SetRowLimit(n): bare LIMIT without ORDER BY; 0 disables. */
    uint64_t row_limit;
    /* Expected input rows (ReserveExpectedRows analogue); 0 = unknown. */
    uint64_t expected_rows;
    /* 0: rows with a NULL group key form one NULL group (SQL; the generic
     *    path, physical_planner.cpp:6187);
     * 1: rows with a NULL group key are dropped (the fused radix/inline
     *    paths, physical_planner.cpp:7556-7557, 10480-10481). */
    int32_t drop_null_keys;
    gbagg_strategy strategy;
} gbagg_query_desc;

/* One result column, host memory owned by the handle. */
typedef struct {
    gbagg_type type;
    const void *data;        /* INTEGER int32, BIGINT int64, DOUBLE double; VARCHAR: unused */
    const uint8_t *validity; /* one byte per row, 0 = NULL; never a null pointer */
    const uint32_t *str_offsets; /* VARCHAR: num_rows + 1 offsets into str_bytes */
    const char *str_bytes;
} gbagg_result_column;

typedef struct {
    uint64_t num_rows;
    uint32_t num_columns;
    const gbagg_result_column *columns;
} gbagg_result_view;

/* What the planner chose and what the last run measured (HIP events on the
 * handle's stream). */
typedef struct {
    gbagg_strategy strategy;    /* strategy the last finalize ran */
    uint64_t estimated_groups;  /* HyperLogLog estimate (radix strategy) */
    uint64_t groups;            /* result rows of the last finalize */
    uint64_t table_capacity;    /* global (fallback) table slots */
    uint32_t lds_slots;         /* per-workgroup LDS table slots */
    uint32_t radix_bits;        /* partition fanout = 2^radix_bits (radix strategy) */
    uint32_t workgroups;        /* aggregate / scatter grid */
    uint32_t retries;           /* re-runs after a table overflow */
    uint64_t rows_consumed;
    uint64_t heavy_keys;        /* heavy-hitter keys aggregated in LDS (radix strategy) */
    uint64_t cold_rows;         /* rows written to partitions (radix strategy) */
    uint64_t sample_rows;       /* sampled rows that passed the filter */
    uint64_t sample_distinct;   /* distinct keys among them */
    uint64_t sample_f1;         /* keys seen exactly once among them */
    double ms_aggregate;        /* the input pass: k_aggregate or k_rp_scatter */
    double ms_total;            /* first kernel to result ready */
    double ms_sample;           /* heavy-hitter sample and selection */
    double ms_count;            /* radix count pass + HyperLogLog */
    double ms_partition;        /* per-partition aggregation + heavy fold */
    double ms_emit;             /* result compaction / TopN */
    char kernel_shape[32];      /* compiled query shape the row kernels ran, or "dynamic" */
    double selector_distinct;   /* learned selector's distinct-group prediction (0 = not run) */
    int32_t selector_used;      /* 1: the plan followed the selector's path; 0: default plan (or no selector) */
    int32_t selector_path;      /* the selector's path: 0 LDS_HASH, 1 GLOBAL_HASH, 2 RADIX; -1 = not run */
    double ms_strings;          /* interning of GBAGG_STR_RAW columns, before the aggregate (0 = none) */
    uint64_t strings_distinct;  /* distinct strings over those columns */
    uint32_t sum_reruns;        /* re-runs for double SUM/AVG: exact bound, or IEEE atomics */
    int32_t sum_ieee;           /* 1: double SUM/AVG ran with IEEE double atomics (order-dependent);
                                   0: exact fixed-point sums, identical for any row order */
} gbagg_plan_info;

typedef struct gbagg_handle gbagg_handle;

/* Library / device information. */
int32_t gbagg_abi_version(void);
const char *gbagg_last_error(void);
gbagg_status gbagg_device_name(char *buf, uint32_t buflen);

/* PhysicalHashAggregate lifecycle (physical_planner.cpp:6009-6060):
 *   plan_create  ~ constructor + SetTopNHint/SetRowLimit/Init
 *   consume      ~ the RGConsumerFn callback of the fused Parquet paths
 *                  (physical_planner.cpp:1060-1062), one call per batch
 *   finalize     ~ ComputeAggregates' merge + emit (result_rows_)
 *   result       ~ GetData's chunks, as whole columns */
gbagg_status gbagg_plan_create(const gbagg_query_desc *query, gbagg_handle **out);
gbagg_status gbagg_consume(gbagg_handle *h, const gbagg_column_view *columns, uint64_t num_rows);
gbagg_status gbagg_consume_device(gbagg_handle *h, const gbagg_column_view *columns,
                                  uint64_t num_rows);
gbagg_status gbagg_finalize(gbagg_handle *h);
gbagg_status gbagg_result(gbagg_handle *h, gbagg_result_view *out);
gbagg_status gbagg_reset(gbagg_handle *h);
gbagg_status gbagg_plan_info_get(gbagg_handle *h, gbagg_plan_info *out);
void gbagg_destroy(gbagg_handle *h);

/* Tuning knobs of one handle (no counterpart in the reference, whose
 * strategy constants are fixed: N_RADIX = 16, radix_count_agg.hpp:304-305).
 * Set by the autotuner and the learned strategy selector; results never
 * depend on them.
 *   "strategy"      gbagg_strategy to run (AUTO decides per query)
 *   "heavy_max"     heavy-hitter keys aggregated in LDS by the radix strategy (0-4096)
 *   "heavy_min"     minimum sample count of a heavy key (1-64)
 *   "part_bits"     radix fanout override, -1 = from the HyperLogLog estimate
 *   "msplit_fanout" largest radix fanout whose scatter takes write offsets by
 *                   wave multisplit (ballot ranks per shared partition) rather
 *                   than one LDS atomic per row (0-8192, default 0: off)
 *   "radix_ahead_rows" inputs of at least this many rows run the radix count
 *                   pass before the plan is read back (default 4M rows)
 *   "lds_kb"        LDS budget of a per-partition table (8-160; default 78, 156 for TopN)
 *   "fallback_cap"  initial slots of the radix strategy's global fallback table
 *   "wide_seed"     first seed of the 16-byte key mixing (reruns take the next
 *                   seeds while two groups share a key word; results never change)
 *   "static_shapes" 1 = use compile-time specialised row kernels when the query
 *                   shape matches one (default), 0 = always the generic kernels
 *   "str_hash_bits" bits of the string hash kept when interning GBAGG_STR_RAW
 *                   columns (1-64, default 64; a row's key holds at most 48 of
 *                   them beside the string's length); fewer bits only force
 *                   strings to share keys, which the byte comparison resolves
 *   "str_codes"     first code capacity and slot count of that interning
 *                   (0 = from the distinct-count estimate, default); smaller
 *                   values only make the tables grow, results never change
 *   "selector"      1 = AUTO plans consult the built-in trained selector
 *                   (default), 0 = the default plan only */
gbagg_status gbagg_set_param(gbagg_handle *h, const char *name, int64_t value);

/* GPUEngine::HashAggregate (include/slothdb/execution/gpu/gpu_engine.hpp:79-85):
 * GROUP BY int32 key, SUM and COUNT of int32 values. Host buffers in and out;
 * out_* hold `capacity` entries; *out_groups receives the group count (an
 * error is returned if it exceeds capacity). */
gbagg_status gbagg_hash_aggregate_i32(const int32_t *keys, const int32_t *values, uint64_t count,
                                      int32_t *out_keys, int64_t *out_sums, int64_t *out_counts,
                                      uint64_t capacity, uint64_t *out_groups);

/* gbagg_hash_aggregate_i32 without any device allocation. It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Every pointer is device memory; the calls are
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting. On completion out_keys/out_sums/out_counts hold
 * min(groups, capacity) groups in unspecified order and *out_groups the group
 * count (larger than capacity means rows were left out).
 * Limit: count <= GBAGG_NOALLOC_MAX_ROWS (524288), else GBAGG_INVALID. The
 * hash table is a fixed 24 MiB scratch array that each GPU the library runs
 * on holds until the process ends; every call on one GPU shares it and
 * clears it first, so calls that overlap on one GPU race. */
#define GBAGG_NOALLOC_MAX_ROWS 524288u
gbagg_status gbagg_hash_aggregate_i32_noalloc(const int32_t *keys, const int32_t *values, uint64_t count,
                                              int32_t *out_keys, int64_t *out_sums, int64_t *out_counts,
                                              uint64_t capacity, uint64_t *out_groups, void *stream);

/* ---- Batches: many GROUP BY queries in one submission --------------------
 * For callers with many small queries, which would otherwise run the
 * aggregate operator (physical_planner.cpp:6009-12627) once per query.
 * Queries of the shape
 *   one INTEGER/BIGINT group column, 1-4 aggregates among COUNT(*),
 *   COUNT/SUM/MIN/MAX over INTEGER/BIGINT columns, columns declared
 *   not nullable, no predicates, no TopN, no row limit
 * run together: a per-query sketch sizes each hash table, then one persistent
 * work-queue kernel aggregates every query's row tiles, all captured once in a
 * HIP graph and replayed while the batch keeps the same queries and inputs.
 * Any other query runs on the single-query operator inside the same call.
 * Columns are device pointers (gbagg_consume_device rules) and must stay valid
 * until the batch is destroyed or run on other inputs. */
typedef struct gbagg_batch gbagg_batch;

typedef struct {
    const gbagg_query_desc *query;
    const gbagg_column_view *columns; /* query->num_columns device columns */
    uint64_t num_rows;
} gbagg_batch_item;

/* Launch shape of the batched kernels; gbagg_batch_autotune picks one. */
typedef struct {
    uint32_t tile_rows;  /* rows per work item */
    uint32_t lds_slots;  /* LDS hash-table slots per workgroup */
    uint32_t block;      /* threads per aggregation workgroup */
    uint32_t wgs_per_cu; /* resident aggregation workgroups per CU */
} gbagg_batch_tune;

typedef struct {
    uint32_t queries;        /* items of the last run */
    uint32_t batched;        /* of which ran in the batched kernels */
    uint32_t single;         /* of which ran on the single-query operator */
    uint32_t graph_captures; /* graphs built so far */
    uint32_t graph_replays;  /* graph launches so far */
    double ms_device;        /* last run: batched graph, launch to completion (HIP events) */
    double ms_single;        /* last run: single-query operator runs (host clock) */
    gbagg_batch_tune tune;
    uint32_t selector_queries; /* batched queries whose mode the learned selector chose */
} gbagg_batch_info;

gbagg_status gbagg_batch_create(gbagg_batch **out);
gbagg_status gbagg_batch_set_tune(gbagg_batch *b, const gbagg_batch_tune *tune);
/* "cap_pct": hash-table slots per estimated distinct key, in percent
 * (default 150; a query whose table overflows re-runs on the single-query
 * operator).
 * "mode_rule": which queries aggregate each row tile in an LDS table before
 * their global table: 0 = those whose distinct estimate fits it, 1 = all,
 * 2 (default) = those, and those whose keys repeat within a tile (first
 * tile's estimate at most half its rows); 3 = none: every query whose keys
 * 1024 partitions hold takes radix partitions. Results never depend on it.
 * "radix": 1 (default) = a query that is not LDS-first writes its rows as
 * records to radix partitions of its key hash (up to 1024, sized from its
 * HyperLogLog estimate), each then aggregated in one LDS table, instead of
 * updating a global table row by row; an LDS-first query writes the rows
 * that find no room in a tile's LDS table, and each tile's partial states,
 * to such partitions instead of a global table; 0 = the global table.
 * "part_slots": LDS table slots per radix partition (default 4096, 256-8192,
 * a multiple of 64; fewer when the LDS holds fewer for the batch's widest
 * query). Results never depend on either.
 * "selector": 1 (default) = each query's mode follows the learned selector's
 * path for its input (DIRECT or LDS_HASH: LDS-first; GLOBAL_HASH or RADIX:
 * not LDS-first; planned on the host once per input, from the same row sample
 * and features as the single-query operator's AUTO plan) and its predicted
 * distinct count may enlarge the table, which never gets smaller than the
 * HyperLogLog estimate sizes it; mode_rule applies to queries whose
 * prediction is more than 4x from their sample's own estimate. 0 = mode_rule
 * only. Results never depend on it. */
gbagg_status gbagg_batch_set_param(gbagg_batch *b, const char *name, int64_t value);
/* Times each candidate launch shape on these items and keeps the fastest. */
gbagg_status gbagg_batch_autotune(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count);
gbagg_status gbagg_batch_run(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count);
/* Result of item `index` of the last run; valid until the next run or destroy. */
gbagg_status gbagg_batch_result(gbagg_batch *b, uint32_t index, gbagg_result_view *out);
gbagg_status gbagg_batch_info_get(gbagg_batch *b, gbagg_batch_info *out);
void gbagg_batch_destroy(gbagg_batch *b);

/* ---- Learned strategy selector ----------------------------------------------
 * An MLP (32 -> 64 -> 64 -> 64 -> 5, FP32, ReLU) over features of the AUTO
 * plan's row sample predicts the distinct-group count, as the correction
 * log2((distinct + 1) / (default + 1)) / 4 to the sample's own estimate
 * default = seen + f1 / sampled * rows, and the fastest path (DIRECT,
 * LDS_HASH, GLOBAL_HASH, RADIX). An AUTO plan takes the predicted path; the
 * predicted count sizes an LDS_HASH / GLOBAL_HASH table when it is within 4x
 * of the sample's own estimate, and an LDS_HASH / GLOBAL_HASH prediction
 * farther off than that leaves the plan to the default rules (RADIX sizes
 * itself from its HyperLogLog count pass, DIRECT from the exact key range,
 * so those paths are taken whatever the count). DIRECT falls back to LDS_HASH
 * when the exact key range is too wide. Results never depend on it.
 * Every handle starts with the built-in selector: the trained parameters
 * shipped with the library (slothdb_amd/selector_weights.json, compiled in).
 * It is consulted for the query shapes it was trained on (one or two
 * INTEGER/BIGINT group columns, no predicates, no TopN hint or row limit);
 * other AUTO plans follow the default rules. gbagg_selector_set replaces it
 * (an attached selector is consulted for every AUTO plan);
 * gbagg_selector_set(h, NULL, 0) or gbagg_set_param(h, "selector", 0) leaves
 * AUTO plans to the default rules.
 * Parameter vector: GBAGG_SELECTOR_PARAMS floats = weights and biases of the
 * four layers in order (W row-major [out][in], then b), then the input
 * standardisation mean[32] and scale[32]. */
#define GBAGG_SELECTOR_FEATURES 32
#define GBAGG_SELECTOR_OUTPUTS 5
#define GBAGG_SELECTOR_PARAMS 10821

/* Per handle; params = NULL removes the selector (the built-in one too). */
gbagg_status gbagg_selector_set(gbagg_handle *h, const float *params, uint32_t num_params);
/* Features (32 floats) and, with a selector, its raw outputs (5 floats) of
 * the last finalize; features are kept when a selector is set or after
 * gbagg_set_param(h, "features", 1). Either pointer may be NULL. */
gbagg_status gbagg_plan_features(gbagg_handle *h, float *features, float *outputs);
/* Trains on the GPU: minibatch (64) Adam over `epochs` passes in row order.
 * x: n x 32 features; y: n x 2 (target correction as above, path class
 * 0..3). params: initial parameters in (including the standardisation), the
 * trained ones out. epoch_loss: `epochs` floats of mean loss, or NULL. */
gbagg_status gbagg_selector_train(const float *x, const float *y, uint32_t n, uint32_t epochs, float lr,
                                  float lambda, float *params, uint32_t num_params, float *epoch_loss);
/* Forward pass on the GPU: out = n x 5 outputs for n x 32 features. */
gbagg_status gbagg_selector_infer(const float *params, uint32_t num_params, const float *x, uint32_t n,
                                  float *out);

/* ---- Parquet row-group decode on the GPU -----------------------------------
 * The fused scan's column source: ParquetReader (include/slothdb/storage/
 * parquet.hpp:213-284) opened on a file image, and ReadColumnInto
 * (parquet.hpp:263-265), which fills ParquetColumnData (parquet.hpp:147-207)
 * per row group and column. Here the decoded columns land in HBM as
 * gbagg_column_views ready for gbagg_consume_device: INTEGER / BIGINT / DOUBLE
 * values, VARCHAR as dictionary codes with the row group's dictionary (host
 * memory) when all its pages are dictionary-encoded, else (a PLAIN page, as
 * writers emit once a dictionary grows too large) as its strings
 * (GBAGG_STR_RAW: device offsets and bytes), validity bytes for OPTIONAL
 * columns.
 * Supported: flat schemas; uncompressed column chunks; INT32, INT64, DOUBLE,
 * BYTE_ARRAY; data pages V1 and V2; PLAIN, PLAIN_DICTIONARY and
 * RLE_DICTIONARY. Anything else is
 * GBAGG_INVALID with a message, and so is malformed data: a page outside its
 * chunk, truncated values, levels or index streams, a dictionary index
 * outside its dictionary, a string length past its page (the kernels check
 * what they read and report through error bits; the call then fails). The file image must outlive the handle; every
 * decoded column stays in HBM until close, so one operator can consume many
 * row groups before its finalize. */
typedef struct gbagg_parquet gbagg_parquet;
gbagg_status gbagg_parquet_open(const uint8_t *file, uint64_t size, gbagg_parquet **out);
/* NumRowGroups (parquet.hpp:284), column count, NumRows (parquet.hpp:234). */
gbagg_status gbagg_parquet_shape(gbagg_parquet *p, uint32_t *num_row_groups, uint32_t *num_columns,
                                 uint64_t *num_rows);
/* GetColumnTypes / GetColumnNames (parquet.hpp:220-224); nullable = OPTIONAL. */
gbagg_status gbagg_parquet_column(gbagg_parquet *p, uint32_t column, gbagg_type *type, uint8_t *nullable,
                                  const char **name);
gbagg_status gbagg_parquet_row_group_rows(gbagg_parquet *p, uint32_t row_group, uint64_t *rows);
/* Decodes `num_columns` columns of one row group; out[i] describes cols[i]. */
gbagg_status gbagg_parquet_decode(gbagg_parquet *p, uint32_t row_group, const uint32_t *cols, uint32_t num_columns,
                                  gbagg_column_view *out, uint64_t *rows);
/* The same over row groups [first, last), concatenated into one batch (VARCHAR
 * dictionaries merged into one of distinct strings): one set of kernels for
 * the whole range (no counterpart in the reference, which decodes row group
 * by row group on its worker threads). */
gbagg_status gbagg_parquet_decode_range(gbagg_parquet *p, uint32_t first, uint32_t last, const uint32_t *cols,
                                        uint32_t num_columns, gbagg_column_view *out, uint64_t *rows);
/* The row-group consumer's form (RGConsumerFn, src/execution/
 * physical_planner.cpp:1060-1062, called once per row group from
 * RunParallelRGs, :1267-1310): plans row group `row_group` as
 * gbagg_parquet_decode does and returns its views and row count at once, but
 * queues its kernels instead of running them and waits for nothing. Queued
 * row groups are launched together once `flush_rows` rows are queued
 * (gbagg_parquet_set_flush_rows; 4M by default) and by gbagg_parquet_sync.
 * The views' device buffers hold the values only after gbagg_parquet_sync
 * returns GBAGG_OK: they may be passed to gbagg_consume_device before that
 * (it reads no device data), but gbagg_parquet_sync must return before the
 * operator's gbagg_finalize. Errors the host sees (unsupported encodings, a
 * page outside its chunk) fail this call and leave the queue as it was;
 * errors the kernels find (an index outside its dictionary, a truncated
 * string) are reported by gbagg_parquet_sync. Calls on one handle must not
 * overlap. gbagg_parquet_decode / _decode_range first run and wait for what
 * is queued; gbagg_parquet_reset drops what is queued. */
gbagg_status gbagg_parquet_decode_async(gbagg_parquet *p, uint32_t row_group, const uint32_t *cols,
                                        uint32_t num_columns, gbagg_column_view *out, uint64_t *rows);
/* Launches the queued row groups, waits for every decode of this handle, and
 * returns the first error their kernels found (GBAGG_INVALID). */
gbagg_status gbagg_parquet_sync(gbagg_parquet *p);
/* Queued rows that trigger a launch (>= 1): fewer overlap decode with the
 * caller's planning of later row groups, more launch fewer, larger kernels. */
gbagg_status gbagg_parquet_set_flush_rows(gbagg_parquet *p, uint64_t rows);
/* Forgets every decoded column (their views become invalid); later decodes
 * reuse the device memory. */
gbagg_status gbagg_parquet_reset(gbagg_parquet *p);
/* Copies the whole file image to HBM once; later decodes read their chunks
 * there instead of uploading them (no counterpart in the reference, which
 * decodes from host memory). */
gbagg_status gbagg_parquet_upload(gbagg_parquet *p);
void gbagg_parquet_close(gbagg_parquet *p);

/* Device memory helpers for callers without their own allocator. */
gbagg_status gbagg_device_alloc(uint64_t bytes, void **out);
gbagg_status gbagg_device_free(void *ptr);
gbagg_status gbagg_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
gbagg_status gbagg_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
gbagg_status gbagg_device_synchronize(void);

/* Hash of the reference's hash maps for a 64-bit key: ankerl::unordered_dense
 * hash<uint64_t> = wyhash::mix(x, 0x9E3779B97F4A7C15)
 * (include/third_party/unordered_dense.h:158-191, 266-267). Host-side, for
 * callers that need the operator's partition bits. */
uint64_t gbagg_hash_u64(uint64_t x);

#ifdef __cplusplus
}
#endif

#endif /* SLOTHDB_GPU_H */
