// gbagg.hip -- host side of the MI355X GROUP BY operator and its C ABI
// (include/slothdb_gpu.h).
//
// A handle holds one planned aggregate (PhysicalHashAggregate's groups,
// aggregates, compiled predicates and TopN hint), its HIP stream, and its
// device buffers. Batches handed to consume/consume_device are queued; the
// whole pipeline runs on the handle's stream in finalize:
//
//   k_hll (per batch) -> k_hll_finish       table sized on the device
//   k_table_init
//   k_aggregate (per batch)                 filter + hash + LDS table + global merge
//   k_occ_count -> k_scan_blocks -> k_compact_write    (full result; k_compact_small
//                                                       for tables of <= 4096 slots), or
//   k_topn_table -> k_topn_level1 -> k_topn_final      (TopN hint, K <= 32)
//
// One device->host copy of the small plan block ends the run; a probe
// overflow (estimate far too low) reruns with a 4x table.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/slothdb_gpu.h"
#include "gbagg_radix.hpp"
#include "gbagg_selector.hpp"
#include "gbagg_strings.hpp"

using namespace gbagg;

static_assert(SEL_NPARAM == GBAGG_SELECTOR_PARAMS, "selector parameter layout");
// The trained selector (slothdb_amd/selector_weights.json, generated at build time).
#include "selector_builtin.inc"
static_assert(sizeof(kSelectorBuiltin) / sizeof(float) == SEL_NPARAM, "built-in selector parameter count");
static_assert(SEL_IN == GBAGG_SELECTOR_FEATURES && SEL_OUT == GBAGG_SELECTOR_OUTPUTS, "selector shape");

namespace {

// ---- compiled query shapes -------------------------------------------------
// Queries whose shape matches one of these run straight-line specialisations
// of the row kernels; every other query runs the same kernels over DynShape.
using KAgg = void (*)(const DevQuery *, BatchRows, DevTable, uint32_t, uint32_t);
using KCount = void (*)(const DevCtx *, BatchRows, const uint32_t *);
using KScatter = void (*)(const DevCtx *, BatchRows, DevTable);
using KPart = void (*)(const DevCtx *, DevTable, uint32_t, uint32_t);
using KRange = void (*)(const DevQuery *, BatchRows, const DevPlan *, int64_t *, uint64_t);
using KDirect = void (*)(const DevQuery *, BatchRows, DevTable, int64_t, uint32_t, uint32_t);

struct ShapeKernels {
    const char *name;
    bool dynamic;
    ShapeSig sig;
    KAgg agg;
    KCount count;
    KScatter scatter;
    KPart part_full, part16, part32;
    KRange range;
    KDirect direct_lds, direct_glb;
};

template <class S>
ShapeKernels shape_kernels(const char *name) {
    ShapeKernels k{};
    k.name = name;
    if constexpr (S::kStatic) k.sig = S::sig();
    k.dynamic = !S::kStatic;
    k.agg = &k_aggregate<S>;
    k.count = &k_rp_count<S>;
    k.scatter = &k_rp_scatter<S>;
    k.part_full = &k_rp_agg<S>;
    k.part16 = &k_rp_agg_topn<S, 16>;
    k.part32 = &k_rp_agg_topn<S, 32>;
    k.range = &k_key_range<S>;
    k.direct_lds = &k_direct<S, true>;
    k.direct_glb = &k_direct<S, false>;
    return k;
}

// slot kinds: 2 bits per slot; accumulator ops / input slots: 4 bits each
// (input slot 15 = COUNT(*)); predicate: slot | kind << 3 | op << 5.
// GROUP BY int32 k: COUNT(*), SUM/MIN/MAX(int64 v)                     (plain aggregate)
using ShPlain = StaticShape<2, 1, 0x4, 0, KM_SINGLE, 0, 0, 4, 0x5420, 0x111F, 0, 0>;
// GROUP BY int32 k: COUNT(*), SUM(int64 v)                              (mixed-query batches)
using ShCountSum = StaticShape<2, 1, 0x4, 0, KM_SINGLE, 0, 0, 2, 0x20, 0x1F, 0, 0>;
// GROUP BY int64 k: COUNT(*), SUM(int64), MIN/MAX(int32), SUM/AVG(double) (high cardinality)
using ShHighcard = StaticShape<4, 1, 0x85, 0, KM_SINGLE, 0, 0, 6, 0x835420, 0x33221F, 0, 0>;
// WHERE c < lit GROUP BY int32 a, VARCHAR b: COUNT(*), SUM(int32 c)   (composite key + filter)
using ShComposite = StaticShape<3, 3, 0xC, 0, KM_PACK32, 0, 1, 2, 0x20, 0x2F, 1, 2>;
// GROUP BY int64 k: COUNT(*)                                          (RadixCountAgg shape)
using ShCount64 = StaticShape<1, 1, 0x1, 0, KM_SINGLE, 0, 0, 1, 0x0, 0xF, 0, 0>;

const std::vector<ShapeKernels> &shape_table() {
    static const std::vector<ShapeKernels> t = {
        shape_kernels<ShPlain>("plain_i32_i64"),
        shape_kernels<ShCountSum>("count_sum_i32_i64"),
        shape_kernels<ShHighcard>("highcard_i64"),
        shape_kernels<ShComposite>("composite_i32_dict_filter"),
        shape_kernels<ShCount64>("count_i64"),
    };
    return t;
}

const ShapeKernels &dynamic_kernels() {
    static const ShapeKernels k = shape_kernels<DynShape>("dynamic");
    return k;
}

thread_local std::string g_last_error;
thread_local const char *g_na_error = nullptr;  // gbagg_hash_aggregate_i32_noalloc's errors (static text)

struct ApiError {
    gbagg_status status;
    std::string msg;
};

[[noreturn]] void fail(gbagg_status st, const std::string &msg) { throw ApiError{st, msg}; }

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            fail(GBAGG_ERROR, std::string(#expr) + ": " + hipGetErrorString(_e));          \
    } while (0)

template <typename Fn>
gbagg_status guarded(Fn &&fn) {
    try {
        fn();
        return GBAGG_OK;
    } catch (const ApiError &e) {
        g_last_error = e.msg;
        g_na_error = nullptr;
        return e.status;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        g_na_error = nullptr;
        return GBAGG_ERROR;
    }
}

uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

struct DeviceInfo {
    int cus = 256;
    size_t lds_per_cu = 160 * 1024;
    size_t lds_per_block = 64 * 1024;
    std::string name;
};

DeviceInfo &device_info() {
    static DeviceInfo info;
    static bool init = false;
    if (!init) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIPCHK(hipGetDeviceProperties(&p, dev));
        info.cus = p.multiProcessorCount;
        info.lds_per_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor
                                                             : 65536;
        info.lds_per_block = p.sharedMemPerBlock;
        info.name = p.gcnArchName;
        init = true;
    }
    return info;
}

// Device buffer that only grows.
// Inputs from this many rows run the radix count pass ahead of the plan
// (default of the "radix_ahead_rows" knob).
constexpr uint64_t RADIX_SPEC_ROWS = 4u << 20;

// Bytes of the parameter block ahead of the plan block in one allocation.
constexpr size_t CTX_BYTES = (sizeof(DevCtx) + 255) & ~(size_t)255;

// Launch shape of the radix count and scatter passes.
struct RadixLaunch {
    uint32_t W1 = 0;      // count workgroups (0 without rows)
    uint64_t per1 = 0;    // rows per count workgroup
    uint64_t per3 = 0;    // rows per scatter workgroup
    size_t k1_lds = 0;    // count pass LDS bytes
};

// A device block owned elsewhere.
struct PlanRef {
    void *ptr = nullptr;
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
};

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(bytes, o.bytes);
        return *this;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (ptr) HIPCHK(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        HIPCHK(hipMalloc(&ptr, n));
        bytes = n;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
};

int col_kind(gbagg_type t) {
    switch (t) {
    case GBAGG_TYPE_INTEGER: return CK_I32;
    case GBAGG_TYPE_BIGINT: return CK_I64;
    case GBAGG_TYPE_DOUBLE: return CK_F64;
    default: return CK_U32;
    }
}

size_t type_width(gbagg_type t) {
    return (t == GBAGG_TYPE_INTEGER || t == GBAGG_TYPE_VARCHAR) ? 4 : 8;
}

bool valid_type(int t) {
    return t == GBAGG_TYPE_INTEGER || t == GBAGG_TYPE_BIGINT || t == GBAGG_TYPE_DOUBLE ||
           t == GBAGG_TYPE_VARCHAR;
}

struct StrCol;

// Host string dictionary for one VARCHAR group column: codes of every batch
// map to one id space so groups match across batches.
struct KeyDict {
    int mode = 0;  // 0 unset, 1 identity (unique dictionary shared by all batches), 2 map,
                   // 4 interned strings, dictionary still on the device (dev)
    const StrCol *dev = nullptr;
    const uint32_t *offsets = nullptr;
    const char *bytes = nullptr;
    uint32_t size = 0;
    std::vector<uint32_t> own_offsets;  // identity mode over a host-consumed batch
    std::vector<char> own_bytes;
    std::unordered_map<std::string, uint32_t> map;
    std::vector<std::string> strings;

    std::string lookup(uint32_t id) const {
        if (mode == 1) {
            const uint32_t *o = own_offsets.empty() ? offsets : own_offsets.data();
            const char *b = own_bytes.empty() ? bytes : own_bytes.data();
            return std::string(b + o[id], o[id + 1] - o[id]);
        }
        return strings[id];
    }
};

// Interning state of one GBAGG_STR_RAW column (device buffers reused across
// finalizes of the handle).
struct StrCol {
    DevBuf segs, hash, pos, codes, slot, code, arena, ctr, hllp, hllm, plan, blk, list, lcnt;
    DevBuf coff, cbytes;                    // group column: its dictionary batches' strings (canonical ids)
    std::vector<uint32_t> coff_h;
    std::vector<char> cbytes_h;
    std::vector<DevBuf> lit, match;  // per predicate
    StrTab tab{};                    // the dictionary: arena and code records
    uint64_t ncodes = 0, words = 0;  // codes, arena words in use
};

struct OutCol {
    int kind;
    gbagg_type type;
    int val_acc;
    int cnt_acc;
    int key_type;
    int lo_acc = -1;  // double SUM/AVG: the ACC_SUM_FL accumulator (val_acc: ACC_SUM_FH)
};

struct Batch {
    DevBatch db{};
    uint64_t nrows = 0;
    const uint32_t *str_off[MAXU] = {};  // GBAGG_STR_RAW columns: device offsets and bytes
    const uint8_t *str_bytes[MAXU] = {};
    std::vector<uint32_t> odoff[MAXU];   // ordered VARCHAR columns given as codes: the batch's dictionary
    std::vector<char> odbytes[MAXU];
    // multi-column keys: the batch's group columns (device), and for VARCHAR
    // columns given as codes a host copy of the batch's dictionary
    struct TupleIn {
        const void *data = nullptr;
        const uint8_t *valid = nullptr;
        const uint32_t *str_off = nullptr;
        const uint8_t *str_bytes = nullptr;
        std::vector<uint32_t> doff;
        std::vector<char> dbytes;
    };
    std::vector<TupleIn> tin;
    std::vector<std::unique_ptr<DevBuf>> owned;  // device copies / remaps / match tables
};

}  // namespace

struct gbagg_handle {
    // plan
    std::vector<gbagg_type> col_types;
    std::vector<uint8_t> nullable;
    std::vector<uint32_t> group_cols;
    std::vector<gbagg_agg> aggs;
    std::vector<gbagg_predicate> preds;
    std::vector<std::vector<int64_t>> in_sets;
    std::vector<std::string> pred_str;
    int32_t topn_col = -1, topn_asc = 0;
    uint64_t topn_limit = 0, row_limit = 0, expected_rows = 0;
    int32_t drop_null_keys = 0;
    gbagg_strategy strategy = GBAGG_STRATEGY_AUTO;

    int nu = 0;
    int ucol[MAXU];
    uint32_t umask_all = 0, umask_hll = 0;
    DevQuery dq{};
    const ShapeKernels *kern = nullptr;
    bool allow_static = true;
    std::vector<OutCol> outs;
    KeyDict kdict[2];
    bool key_varchar[2] = {false, false};
    int32_t wide_acc[2] = {0, 0};  // KM_WIDE: MIN / MAX of the kept key component
    // COUNT(DISTINCT): phase 1 groups by (group column, argument) with partial
    // states of the other aggregates; this handle is phase 2 over its result
    std::unique_ptr<gbagg_handle> phase1;
    // COUNT(DISTINCT) over further argument columns: one more phase 1 each,
    // grouping by (group columns, that argument), no aggregates
    std::vector<std::unique_ptr<gbagg_handle>> phase1_more;
    uint32_t distinct_args = 0;                      // argument columns (phase 1 handles)
    uint32_t distinct_ng = 0;                        // the query's group columns
    std::vector<std::vector<uint32_t>> p1_dict_off;  // per (phase 1, column): VARCHAR group key
    std::vector<std::vector<char>> p1_dict_bytes;    // dictionaries, as phase 2 input
    DevBuf p2_null;                                  // zeros: data and validity of the NULL columns
    uint64_t wide_seed0 = 0;       // KM_WIDE: first seed of wide_mix
    // VARCHAR columns given as strings (GBAGG_STR_RAW), interned at finalize
    std::vector<uint8_t> str_mode;  // per column: 0 no batch yet, 1 dictionary codes, 2 strings
    std::vector<uint8_t> str_ordered;  // per column: MIN/MAX over it, so its codes follow byte order
    std::vector<std::unique_ptr<StrCol>> strs;
    uint64_t str_key_mask = ~0ull;  // stored string-hash bits ("str_hash_bits")
    uint64_t str_codes0 = 0;        // first code capacity and slot count, 0 = from the estimate ("str_codes")
    // multi-column keys (more than two columns, a DOUBLE column, or a nullable
    // pair without drop_null_keys): the plan groups by one hidden column, the
    // code of each row's interned key tuple; results unpack the tuples
    bool tuple = false;
    uint32_t ext_cols = 0;              // the caller's columns; the hidden one is column ext_cols
    std::vector<uint32_t> tcols;        // the caller's group columns
    std::vector<gbagg_type> ttypes;
    std::vector<uint32_t> twordof;      // first word of each column's value in a packed row
    uint32_t twords = 0;                // words per packed row
    int32_t tdrop = 0;                  // the caller's drop_null_keys
    std::vector<std::unique_ptr<StrCol>> tstr;  // VARCHAR group columns' interning
    std::unique_ptr<StrCol> tsc;        // the tuples' interning
    DevBuf tpack, tkeep, tcomp, touts;
    std::vector<DevBuf> tout_data, tout_valid;  // group columns unpacked on the device
    // double sums (fixed-point words, see AccOp): per accumulator the scale s
    std::vector<std::pair<int, int>> fx_pairs;  // (ACC_SUM_FH, ACC_SUM_FL) accumulators per DOUBLE column
    bool fx_ieee = false;          // after OVF_FX_PREC: IEEE double atomics instead
    bool fx_exact = false;         // scales from the columns' exact maxima (after OVF_FX)
    bool fx_known = false;         // fx_exp holds those maxima
    uint32_t fx_exp[MAXU] = {};    // exact largest biased exponents per loaded slot
    uint32_t fx_reruns = 0;
    // TopN on the device
    bool topn_device = false;
    DevTopn tspec{};
    uint32_t topn_k = 0;

    // runtime
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::vector<Batch> batches;
    std::vector<DevBuf> in_set_bufs;
    DevBuf table, hll, blkcnt, cand_rec, cand_null, topn_lvl;
    PlanRef dplan;  // the device plan block, inside the ctx allocation
    DevBuf ghist, hllm, skeys, scnt, hkeys, hidx, hdense, hfinal, hist, hpart, base, ptotal, pstart, recs;
    std::vector<DevBuf> out_data, out_valid;
    DevPlan *hplan = nullptr;  // pinned
    DevCtx *hctx = nullptr;    // pinned parameter blocks: hctx, then hplan at CTX_BYTES
    bool ctx_stale = false;    // hctx changed since its last upload
    RadixLaunch rl;            // radix passes of the current finalize
    bool range_ahead = false;  // the exact key range of this finalize came back with the sample's plan
    DevBatch *hbatches = nullptr;
    size_t hbatch_cap = 0;
    DevBuf ctx, dbatches;
    uint64_t cap_max = 0;
    uint64_t out_capacity = 0;
    uint32_t cand_cap = 0;
    uint64_t rows_total = 0;
    bool finalized = false;
    uint64_t result_rows = 0;
    // tuning (autotuner / selector may override)
    uint32_t heavy_max = 1024;
    uint32_t heavy_min = 4;
    uint64_t radix_fallback_cap = 65536;
    int32_t force_part_bits = -1;   // radix fanout override (-1 = from the estimate)
    uint32_t msplit_fanout = MSPLIT_MAX_FANOUT;  // radix scatter offsets by wave multisplit up to this fanout
    uint64_t radix_ahead_rows = RADIX_SPEC_ROWS;  // inputs from this many rows count ahead of the plan
    uint32_t k4_lds_kb = 0;         // per-partition LDS table budget in KiB; 0: 78 (two workgroups per
                                    // CU), 156 for TopN plans (one workgroup per CU, larger tables)
    // learned selector
    std::vector<float> sel_params;  // empty: no selector attached
    bool sel_builtin = true;        // AUTO plans without an attached selector use the built-in one
    bool collect_features = false;
    DevBuf dsel, dfeat, dpred;
    float feat[SEL_IN] = {}, pred[SEL_OUT] = {};
    bool feat_valid = false;
    double sel_distinct = 0;
    int32_t sel_used = 0, sel_path = -1;
    bool sel_count = false;  // the selector's count is within 4x of the sample's estimate
    // last run
    gbagg_strategy used_strategy = GBAGG_STRATEGY_AUTO;
    uint32_t lds_slots = 0, agg_grid = 0, retries = 0, radix_bits = 0;
    double ms_agg = 0, ms_total = 0, ms_sample = 0, ms_count = 0, ms_partition = 0, ms_emit = 0;
    double ms_strings = 0;
    uint64_t strings_distinct = 0;

    // host result
    std::vector<std::vector<uint8_t>> res_data, res_valid;
    std::vector<std::vector<uint32_t>> res_str_off;
    std::vector<std::vector<char>> res_str_bytes;
    std::vector<gbagg_result_column> res_cols;
    bool result_ready = false;

    // host batches (gbagg_consume): their bytes are gathered into pinned
    // staging, two buffers that alternate (each reused once the event
    // recorded after its copies has passed), and copied asynchronously into
    // device memory bump-allocated from arenas kept until reset: a consume
    // returns without waiting for its copies
    uint8_t *hstage[2] = {nullptr, nullptr};
    size_t hstage_bytes[2] = {0, 0};
    hipEvent_t hstage_done[2] = {nullptr, nullptr};
    int hstage_next = 0;
    std::vector<DevBuf> barena;
    size_t barena_i = 0, barena_used = 0;
    uint8_t *batch_alloc(size_t n) {
        n = (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
        while (barena_i < barena.size() && barena_used + n > barena[barena_i].bytes) {
            barena_i++;
            barena_used = 0;
        }
        if (barena_i == barena.size()) {
            barena.emplace_back();
            barena.back().ensure(std::max<size_t>(n, (size_t)64 << 20));
            barena_used = 0;
        }
        uint8_t *p = barena[barena_i].as<uint8_t>() + barena_used;
        barena_used += n;
        return p;
    }

    ~gbagg_handle() {
        if (stream) (void)hipStreamSynchronize(stream);  // copies out of the staging buffers
        if (hctx) (void)hipHostFree(hctx);  // hplan lives in the same block
        if (hbatches) (void)hipHostFree(hbatches);
        for (int k = 0; k < 2; k++) {
            if (hstage_done[k]) (void)hipEventDestroy(hstage_done[k]);
            if (hstage[k]) (void)hipHostFree(hstage[k]);
        }
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int add_u(gbagg_handle *h, uint32_t col) {
    for (int u = 0; u < h->nu; u++)
        if ((uint32_t)h->ucol[u] == col) return u;
    if (h->nu >= MAXU) fail(GBAGG_INVALID, "a plan may reference at most 8 distinct input columns");
    h->ucol[h->nu] = (int)col;
    h->dq.ukind[h->nu] = col_kind(h->col_types[col]);
    return h->nu++;
}

int add_acc(gbagg_handle *h, int op, int u) {
    for (int a = 0; a < h->dq.nacc; a++)
        if (h->dq.acc_op[a] == op && h->dq.acc_u[a] == u) return a;
    if (h->dq.nacc >= MAX_ACC) fail(GBAGG_INVALID, "too many accumulators in one plan");
    h->dq.acc_op[h->dq.nacc] = op;
    h->dq.acc_u[h->dq.nacc] = u;
    return h->dq.nacc++;
}

void build_distinct(gbagg_handle *h, const gbagg_query_desc *q);
void pick_kernels(gbagg_handle *h);

// Group keys that the one-word and 16-byte keys cannot hold: more than two
// columns, a DOUBLE column, or two columns with NULLs that form groups.
bool needs_tuple(const gbagg_query_desc *q) {
    if (!q->column_types || !q->group_cols || q->num_group_cols == 0 || q->num_columns > GBAGG_MAX_COLUMNS)
        return false;
    if (q->num_group_cols > 2) return true;
    bool nullable = false;
    for (uint32_t g = 0; g < q->num_group_cols; g++) {
        const uint32_t c = q->group_cols[g];
        if (c >= q->num_columns) return false;
        if (q->column_types[c] == GBAGG_TYPE_DOUBLE) return true;
        nullable |= q->column_nullable ? q->column_nullable[c] != 0 : true;
    }
    return q->num_group_cols == 2 && nullable && !q->drop_null_keys;
}

void build_tuple(gbagg_handle *h, const gbagg_query_desc *q);

void build_plan(gbagg_handle *h, const gbagg_query_desc *q) {
    if (!q) fail(GBAGG_INVALID, "null query");
    if (q->aggs && q->num_aggs <= GBAGG_MAX_AGGS)
        for (uint32_t a = 0; a < q->num_aggs; a++)
            if (q->aggs[a].kind == GBAGG_AGG_COUNT_DISTINCT) return build_distinct(h, q);
    if (needs_tuple(q)) return build_tuple(h, q);
    if (q->num_columns == 0 || q->num_columns > GBAGG_MAX_COLUMNS || !q->column_types)
        fail(GBAGG_INVALID, "num_columns must be 1..16 with column_types");
    for (uint32_t c = 0; c < q->num_columns; c++) {
        if (!valid_type(q->column_types[c])) fail(GBAGG_INVALID, "unsupported column type");
        h->col_types.push_back(q->column_types[c]);
        h->nullable.push_back(q->column_nullable ? q->column_nullable[c] : 1);
    }
    if (q->num_group_cols > GBAGG_MAX_GROUP_COLS) fail(GBAGG_INVALID, "at most 8 group columns");
    if (q->num_group_cols && !q->group_cols) fail(GBAGG_INVALID, "null group_cols");
    if (q->num_group_cols > 2) fail(GBAGG_INVALID, "group columns out of range");  // keys of 3+ columns are tuples
    if (q->num_aggs > GBAGG_MAX_AGGS) fail(GBAGG_INVALID, "at most 16 aggregates");
    if (q->num_predicates > GBAGG_MAX_PREDICATES) fail(GBAGG_INVALID, "at most 8 predicates");
    if (q->num_aggs == 0 && q->num_group_cols == 0) fail(GBAGG_INVALID, "nothing to aggregate");
    h->drop_null_keys = q->drop_null_keys;
    h->strategy = q->strategy;
    h->topn_col = q->topn_col;
    h->topn_asc = q->topn_ascending;
    h->topn_limit = q->topn_limit;
    h->row_limit = q->row_limit;
    h->expected_rows = q->expected_rows;
    h->dq.drop_null_keys = q->drop_null_keys;

    // group keys
    for (uint32_t g = 0; g < q->num_group_cols; g++) {
        uint32_t c = q->group_cols[g];
        if (c >= q->num_columns) fail(GBAGG_INVALID, "group column out of range");
        gbagg_type t = h->col_types[c];
        if (t == GBAGG_TYPE_DOUBLE) fail(GBAGG_INVALID, "DOUBLE group keys are not supported");
        h->group_cols.push_back(c);
        h->key_varchar[g] = (t == GBAGG_TYPE_VARCHAR);
    }
    if (q->num_group_cols == 0) {
        h->dq.key_mode = KM_NONE;
    } else if (q->num_group_cols == 1) {
        h->dq.key_mode = KM_SINGLE;
        h->dq.key_u[0] = add_u(h, h->group_cols[0]);
        h->dq.key_signed32[0] = h->col_types[h->group_cols[0]] != GBAGG_TYPE_VARCHAR;
    } else {
        bool any_null = h->nullable[h->group_cols[0]] || h->nullable[h->group_cols[1]];
        if (any_null && !q->drop_null_keys)
            fail(GBAGG_INVALID, "two-column keys over nullable columns require drop_null_keys");
        // two 32-bit components pack into the key word; a BIGINT component
        // makes a 16-byte key (KM_WIDE), keeping a non-VARCHAR component in
        // hidden accumulators
        const bool big0 = h->col_types[h->group_cols[0]] == GBAGG_TYPE_BIGINT;
        const bool big1 = h->col_types[h->group_cols[1]] == GBAGG_TYPE_BIGINT;
        h->dq.key_mode = big0 || big1 ? KM_WIDE : KM_PACK32;
        for (int g = 0; g < 2; g++) {
            h->dq.key_u[g] = add_u(h, h->group_cols[g]);
            h->dq.key_signed32[g] = h->col_types[h->group_cols[g]] == GBAGG_TYPE_INTEGER;
        }
        h->dq.wide_comp = h->key_varchar[1] ? 0 : 1;
        h->dq.wide_seed = 0;
    }

    // predicates
    for (uint32_t i = 0; i < q->num_predicates; i++) {
        gbagg_predicate p = q->predicates[i];
        if (p.column < 0 || (uint32_t)p.column >= q->num_columns) fail(GBAGG_INVALID, "predicate column out of range");
        gbagg_type t = h->col_types[p.column];
        DevPred &dp = h->dq.pred[i];
        dp.u = add_u(h, (uint32_t)p.column);
        dp.op = p.op;
        dp.ival = p.ival;
        dp.dval = p.dval;
        h->in_sets.emplace_back();
        h->pred_str.emplace_back();
        if (t == GBAGG_TYPE_VARCHAR) {
            if (p.op != GBAGG_CMP_EQ && p.op != GBAGG_CMP_NE && p.op != GBAGG_CMP_LIKE &&
                p.op != GBAGG_CMP_NOT_LIKE)
                fail(GBAGG_INVALID, "VARCHAR predicates support =, <>, LIKE '%x%', NOT LIKE");
            dp.kind = PK_DICT;
            h->pred_str.back().assign(p.sval ? p.sval : "", p.sval ? p.slen : 0);
        } else if (p.op == GBAGG_CMP_IN) {
            if (t == GBAGG_TYPE_DOUBLE) fail(GBAGG_INVALID, "IN is supported on INTEGER/BIGINT");
            if (p.in_count > GBAGG_MAX_IN_SET || (p.in_count && !p.in_set))
                fail(GBAGG_INVALID, "IN set too large or missing");
            dp.kind = PK_IN;
            h->in_sets.back().assign(p.in_set, p.in_set + p.in_count);
        } else if (p.op <= GBAGG_CMP_NE) {
            dp.kind = (t == GBAGG_TYPE_DOUBLE) ? PK_F64 : PK_INT;
        } else {
            fail(GBAGG_INVALID, "LIKE is supported on VARCHAR columns only");
        }
        h->preds.push_back(p);
    }
    h->dq.npred = (int32_t)q->num_predicates;
    h->dq.nkp = h->nu;
    h->umask_hll = (1u << h->nu) - 1;  // keys + predicate columns

    // aggregates -> accumulators (acc 0 is always COUNT(*))
    add_acc(h, ACC_CNT_STAR, -1);
    if (h->dq.key_mode == KM_WIDE) {
        const int u = h->dq.key_u[h->dq.wide_comp];
        h->wide_acc[0] = add_acc(h, ACC_MIN_I, u);
        h->wide_acc[1] = add_acc(h, ACC_MAX_I, u);
    }
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        gbagg_agg ag = q->aggs[a];
        OutCol oc{};
        oc.cnt_acc = -1;
        if (ag.kind == GBAGG_AGG_COUNT_STAR) {
            oc.kind = OUT_COUNT;
            oc.type = GBAGG_TYPE_BIGINT;
            oc.val_acc = 0;
        } else {
            if (ag.column < 0 || (uint32_t)ag.column >= q->num_columns)
                fail(GBAGG_INVALID, "aggregate column out of range");
            gbagg_type t = h->col_types[ag.column];
            int u = add_u(h, (uint32_t)ag.column);
            int cnt = h->nullable[ag.column] ? add_acc(h, ACC_CNT_VALID, u) : 0;
            switch (ag.kind) {
            case GBAGG_AGG_COUNT:
                oc.kind = OUT_COUNT;
                oc.type = GBAGG_TYPE_BIGINT;
                oc.val_acc = cnt;
                break;
            case GBAGG_AGG_SUM:
            case GBAGG_AGG_AVG:
                if (t == GBAGG_TYPE_VARCHAR) fail(GBAGG_INVALID, "SUM/AVG require a numeric argument");
                oc.cnt_acc = cnt;
                if (t == GBAGG_TYPE_DOUBLE) {
                    // exact fixed-point words; fx_mode() swaps in IEEE atomics when needed
                    oc.val_acc = add_acc(h, ACC_SUM_FH, u);
                    oc.lo_acc = add_acc(h, ACC_SUM_FL, u);
                    const std::pair<int, int> pr(oc.val_acc, oc.lo_acc);
                    if (std::find(h->fx_pairs.begin(), h->fx_pairs.end(), pr) == h->fx_pairs.end())
                        h->fx_pairs.push_back(pr);
                    oc.kind = ag.kind == GBAGG_AGG_SUM ? OUT_SUM_F : OUT_AVG_F;
                    oc.type = GBAGG_TYPE_DOUBLE;
                } else {
                    oc.val_acc = add_acc(h, ACC_SUM_I, u);
                    oc.kind = ag.kind == GBAGG_AGG_SUM ? OUT_SUM_I : OUT_AVG_I;
                    oc.type = ag.kind == GBAGG_AGG_SUM ? GBAGG_TYPE_BIGINT : GBAGG_TYPE_DOUBLE;
                }
                break;
            case GBAGG_AGG_MIN:
            case GBAGG_AGG_MAX: {
                bool mn = ag.kind == GBAGG_AGG_MIN;
                oc.cnt_acc = cnt;
                oc.type = t;
                if (t == GBAGG_TYPE_VARCHAR) {
                    // the column is interned with codes in byte order: MIN/MAX of codes
                    if (h->str_ordered.size() != h->col_types.size()) h->str_ordered.assign(h->col_types.size(), 0);
                    h->str_ordered[ag.column] = 1;
                    oc.val_acc = add_acc(h, mn ? ACC_MIN_I : ACC_MAX_I, u);
                    oc.kind = OUT_MM_I32;
                } else if (t == GBAGG_TYPE_DOUBLE) {
                    oc.val_acc = add_acc(h, mn ? ACC_MIN_F : ACC_MAX_F, u);
                    oc.kind = OUT_MM_F;
                } else {
                    oc.val_acc = add_acc(h, mn ? ACC_MIN_I : ACC_MAX_I, u);
                    oc.kind = t == GBAGG_TYPE_INTEGER ? OUT_MM_I32 : OUT_MM_I64;
                }
                break;
            }
            default: fail(GBAGG_INVALID, "unknown aggregate kind");
            }
            // count of non-NULL inputs decides NULL results; COUNT(*) when non-nullable
            if (oc.kind != OUT_COUNT && !h->nullable[ag.column]) oc.cnt_acc = 0;
        }
        h->aggs.push_back(ag);
        h->outs.push_back(oc);
    }
    // key output columns go first
    std::vector<OutCol> keys;
    for (uint32_t g = 0; g < h->group_cols.size(); g++) {
        OutCol oc{};
        oc.kind = OUT_KEY;
        oc.type = h->col_types[h->group_cols[g]];
        oc.val_acc = (int)g;
        oc.cnt_acc = -1;
        oc.key_type = col_kind(oc.type);
        keys.push_back(oc);
    }
    h->outs.insert(h->outs.begin(), keys.begin(), keys.end());
    h->dq.nu = h->nu;
    h->umask_all = (1u << h->nu) - 1;

    // TopN on the device: order column must never be NULL, K <= 32, and a
    // group-column order only for single numeric keys.
    if (h->topn_col >= 0 && h->topn_limit > 0) {
        if ((uint32_t)h->topn_col >= h->outs.size()) fail(GBAGG_INVALID, "topn_col out of range");
        const OutCol &oc = h->outs[h->topn_col];
        DevTopn &t = h->tspec;
        t.ascending = h->topn_asc;
        t.key_mode = h->dq.key_mode;
        t.key_signed[0] = h->dq.key_signed32[0];
        t.key_signed[1] = h->dq.key_signed32[1];
        t.acc = oc.val_acc;
        t.lo_acc = oc.lo_acc < 0 ? 0 : oc.lo_acc;
        t.cnt_acc = oc.cnt_acc < 0 ? 0 : oc.cnt_acc;
        bool ok = h->topn_limit <= 32;
        bool never_null = oc.cnt_acc <= 0;
        switch (oc.kind) {
        case OUT_KEY:
            t.kind = TK_KEY;
            ok = ok && h->dq.key_mode == KM_SINGLE && !h->key_varchar[0] &&
                 (h->drop_null_keys || !h->nullable[h->group_cols[0]]);
            break;
        case OUT_COUNT: t.kind = TK_ACC_I; break;
        case OUT_SUM_I:
        case OUT_MM_I32:
        case OUT_MM_I64: t.kind = TK_ACC_I; ok = ok && never_null; break;
        case OUT_MM_F: t.kind = TK_ACC_O; ok = ok && never_null; break;
        case OUT_SUM_F: t.kind = TK_SUM_F; ok = ok && never_null; break;
        case OUT_AVG_I: t.kind = TK_AVG_I; ok = ok && never_null; break;
        default: t.kind = TK_AVG_F; ok = ok && never_null; break;
        }
        h->topn_device = ok;
        h->topn_k = (uint32_t)std::min<uint64_t>(h->topn_limit, 32);
        t.k = h->topn_k;
    }
}

// Multi-column keys: the same plan over one more column, the tuple code,
// grouped by that column alone (rows dropped through its validity).
void build_tuple(gbagg_handle *h, const gbagg_query_desc *q) {
    const uint32_t G = q->num_group_cols, nc = q->num_columns;
    if (G > GBAGG_MAX_GROUP_COLS) fail(GBAGG_INVALID, "at most 8 group columns");
    if (nc + 1 > GBAGG_MAX_COLUMNS) fail(GBAGG_INVALID, "multi-column keys take at most 15 columns");
    if (q->topn_limit && q->topn_col >= 0 && (uint32_t)q->topn_col < G)
        fail(GBAGG_INVALID, "TopN on a group column of a multi-column key is not supported");
    std::vector<gbagg_type> types(q->column_types, q->column_types + nc);
    for (gbagg_type t : types)
        if (!valid_type(t)) fail(GBAGG_INVALID, "unsupported column type");
    types.push_back(GBAGG_TYPE_INTEGER);
    std::vector<uint8_t> nullable(nc + 1, 1);
    for (uint32_t c = 0; c < nc; c++) nullable[c] = q->column_nullable ? q->column_nullable[c] : 1;
    uint32_t code_col = nc;
    gbagg_query_desc q2 = *q;
    q2.num_columns = nc + 1;
    q2.column_types = types.data();
    q2.column_nullable = nullable.data();
    q2.num_group_cols = 1;
    q2.group_cols = &code_col;
    q2.drop_null_keys = 1;
    if (q->topn_limit && q->topn_col >= 0) q2.topn_col = q->topn_col - (int32_t)G + 1;
    build_plan(h, &q2);
    h->tuple = true;
    h->ext_cols = nc;
    h->tdrop = q->drop_null_keys;
    uint32_t word = (G + 3) / 4;
    for (uint32_t g = 0; g < G; g++) {
        const uint32_t c = q->group_cols[g];
        if (c >= nc) fail(GBAGG_INVALID, "group column out of range");
        h->tcols.push_back(c);
        h->ttypes.push_back(types[c]);
        h->twordof.push_back(word);
        word += types[c] == GBAGG_TYPE_BIGINT || types[c] == GBAGG_TYPE_DOUBLE ? 2 : 1;
    }
    h->twords = word;
    h->tstr.resize(G);
}

// COUNT(DISTINCT x) in two phases (the generic path's per-group distinct_set,
// physical_planner.cpp:12325-12330, as a decomposed aggregate):
//   phase 1  GROUP BY (k..., x) with partial states of the other aggregates
//            (COUNT, SUM, MIN, MAX; AVG as SUM and COUNT); a NULL x is a key
//            value like any other
//   phase 2  GROUP BY k... over phase 1's device-resident result: COUNT(x) is
//            the distinct count, the partial states combine (SUM of COUNT and
//            SUM, MIN of MIN, MAX of MAX, AVG = SUM of sums / SUM of counts)
// k is zero to seven columns of any type; x is one column of any type (DOUBLE
// arguments are distinct by value: -0.0 equals 0.0, and all NaNs are one).
// Several argument columns x_0 .. x_{m-1} (each aggregate keeps its own
// distinct set, physical_planner.cpp:6076, 12325-12349): phase 1 runs once
// per argument, GROUP BY (k..., x_j), the partial states only in the first;
// phase 2 takes every phase 1 result as a batch of its own and counts each
// x_j over its own batch's rows (NULL in the others).
void build_distinct(gbagg_handle *h, const gbagg_query_desc *q) {
    if (q->num_columns == 0 || q->num_columns > GBAGG_MAX_COLUMNS || !q->column_types)
        fail(GBAGG_INVALID, "num_columns must be 1..16 with column_types");
    if (q->num_group_cols >= GBAGG_MAX_GROUP_COLS)
        fail(GBAGG_INVALID, "COUNT(DISTINCT) supports at most 7 group columns");
    if (q->num_group_cols && !q->group_cols) fail(GBAGG_INVALID, "null group_cols");
    for (uint32_t g = 0; g < q->num_group_cols; g++)
        if (q->group_cols[g] >= q->num_columns) fail(GBAGG_INVALID, "group column out of range");
    auto nullable = [&](uint32_t c) { return q->column_nullable ? q->column_nullable[c] != 0 : true; };
    std::vector<int32_t> xs;  // the argument columns, in first-use order
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        const gbagg_agg &g = q->aggs[a];
        if (g.kind != GBAGG_AGG_COUNT_STAR && (g.column < 0 || (uint32_t)g.column >= q->num_columns))
            fail(GBAGG_INVALID, "aggregate column out of range");
        if (g.kind == GBAGG_AGG_COUNT_DISTINCT && std::find(xs.begin(), xs.end(), g.column) == xs.end())
            xs.push_back(g.column);
    }
    const int32_t x = xs[0];
    const uint32_t m = (uint32_t)xs.size();
    for (uint32_t a = 0; a < q->num_aggs; a++)
        if ((q->aggs[a].kind == GBAGG_AGG_MIN || q->aggs[a].kind == GBAGG_AGG_MAX) &&
            q->column_types[q->aggs[a].column] == GBAGG_TYPE_VARCHAR)
            fail(GBAGG_INVALID, "MIN/MAX over VARCHAR next to COUNT(DISTINCT) is not supported");
    for (int32_t xc : xs)
        if (!valid_type(q->column_types[xc])) fail(GBAGG_INVALID, "unsupported column type");
    const uint32_t ng = q->num_group_cols;
    h->distinct_args = m;
    h->distinct_ng = ng;
    // NULL arguments stay in phase 1 as a group of their own (its partial
    // states reach the other aggregates, and a group whose arguments are all
    // NULL still appears, with COUNT(DISTINCT) 0); phase 2 counts non-NULL
    // arguments. NULL keys are kept by phase 1 unless the query drops them
    // and the argument has no NULLs (then phase 1 drops them itself).
    const bool keep_null_x = nullable((uint32_t)x);
    bool null_k = false;
    for (uint32_t g = 0; g < ng; g++) null_k |= nullable(q->group_cols[g]);
    const bool p1_drop = !keep_null_x && (q->drop_null_keys || !null_k);

    // phase 1: partial states, deduplicated
    std::vector<gbagg_agg> a1;
    auto p1_agg = [&](gbagg_agg_kind k, int32_t c) {
        for (size_t i = 0; i < a1.size(); i++)
            if (a1[i].kind == k && a1[i].column == c) return (uint32_t)i;
        a1.push_back(gbagg_agg{k, c});
        return (uint32_t)(a1.size() - 1);
    };
    std::vector<uint32_t> src(q->num_aggs, 0), src_cnt(q->num_aggs, 0);  // phase-1 agg feeding user agg a
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        const gbagg_agg &g = q->aggs[a];
        switch (g.kind) {
        case GBAGG_AGG_COUNT_DISTINCT: break;
        case GBAGG_AGG_COUNT_STAR: src[a] = p1_agg(GBAGG_AGG_COUNT_STAR, -1); break;
        case GBAGG_AGG_COUNT: src[a] = p1_agg(GBAGG_AGG_COUNT, g.column); break;
        case GBAGG_AGG_SUM: src[a] = p1_agg(GBAGG_AGG_SUM, g.column); break;
        case GBAGG_AGG_AVG:
            src[a] = p1_agg(GBAGG_AGG_SUM, g.column);
            src_cnt[a] = p1_agg(GBAGG_AGG_COUNT, g.column);
            break;
        case GBAGG_AGG_MIN:
        case GBAGG_AGG_MAX: src[a] = p1_agg(g.kind, g.column); break;
        default: fail(GBAGG_INVALID, "unknown aggregate kind");
        }
    }
    uint32_t gc1[GBAGG_MAX_GROUP_COLS];
    for (uint32_t g = 0; g < ng; g++) gc1[g] = q->group_cols[g];
    gc1[ng] = (uint32_t)x;
    gbagg_query_desc d1 = *q;
    d1.num_group_cols = ng + 1;
    d1.group_cols = gc1;
    d1.num_aggs = (uint32_t)a1.size();
    d1.aggs = a1.empty() ? nullptr : a1.data();
    d1.topn_col = -1;
    d1.topn_limit = 0;
    d1.row_limit = 0;
    d1.drop_null_keys = p1_drop ? 1 : 0;
    h->phase1 = std::make_unique<gbagg_handle>();
    build_plan(h->phase1.get(), &d1);
    pick_kernels(h->phase1.get());
    // the further arguments: GROUP BY (group columns, x_j), nothing else
    for (uint32_t j = 1; j < m; j++) {
        uint32_t gcj[GBAGG_MAX_GROUP_COLS];
        for (uint32_t g = 0; g < ng; g++) gcj[g] = q->group_cols[g];
        gcj[ng] = (uint32_t)xs[j];
        gbagg_query_desc dj = *q;
        dj.num_group_cols = ng + 1;
        dj.group_cols = gcj;
        dj.num_aggs = 0;
        dj.aggs = nullptr;
        dj.topn_col = -1;
        dj.topn_limit = 0;
        dj.row_limit = 0;
        // a row with a NULL x_j adds nothing to its distinct count, so these
        // drop as the query drops NULL keys (all groups come from phase 1 of x_0)
        dj.drop_null_keys = q->drop_null_keys || !null_k ? 1 : 0;
        h->phase1_more.push_back(std::make_unique<gbagg_handle>());
        build_plan(h->phase1_more.back().get(), &dj);
        pick_kernels(h->phase1_more.back().get());
    }

    // phase 2 over the phase 1 outputs: [k,] x_0 .. x_{m-1}, partial states.
    // Each phase 1 result is one batch; its x_j fills column ng + j and the
    // other arguments and (for j > 0) the partial states are NULL. Phase 2
    // counts only whether an argument is NULL, so with several arguments they
    // are declared by width (INTEGER for 4-byte values and codes, BIGINT for
    // 8-byte values and DOUBLE bits).
    const gbagg_handle *p1 = h->phase1.get();
    std::vector<gbagg_type> t2;
    std::vector<uint8_t> n2;
    for (uint32_t c = 0; c < ng; c++) {
        t2.push_back(q->column_types[gc1[c]]);
        n2.push_back(!p1_drop && nullable(gc1[c]) ? 1 : 0);
    }
    for (uint32_t j = 0; j < m; j++) {
        const gbagg_type xt = q->column_types[xs[j]];
        if (m == 1) {
            t2.push_back(xt);
            n2.push_back(!p1_drop && nullable((uint32_t)x) ? 1 : 0);
        } else {
            t2.push_back(xt == GBAGG_TYPE_BIGINT || xt == GBAGG_TYPE_DOUBLE ? GBAGG_TYPE_BIGINT : GBAGG_TYPE_INTEGER);
            n2.push_back(1);
        }
    }
    for (size_t i = 0; i < a1.size(); i++) {
        const size_t oc = p1->tuple ? i + 1 : ng + 1 + i;  // phase 1's output column of partial state i
        const bool cnt = p1->outs[oc].kind == OUT_COUNT;
        const int32_t arg = a1[i].column;
        t2.push_back(p1->outs[oc].type);
        n2.push_back(m > 1 ? 1 : cnt ? 0 : (arg >= 0 && nullable((uint32_t)arg)) ? 1 : 0);
    }
    const uint32_t base = ng + m;  // first partial-state column
    std::vector<gbagg_agg> a2;
    std::vector<int> avg_hidden(q->num_aggs, -1);
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        const gbagg_agg &g = q->aggs[a];
        const int32_t col = (int32_t)(base + src[a]);
        switch (g.kind) {
        case GBAGG_AGG_COUNT_DISTINCT:
            a2.push_back({GBAGG_AGG_COUNT,
                          (int32_t)(ng + (std::find(xs.begin(), xs.end(), g.column) - xs.begin()))});
            break;
        case GBAGG_AGG_MIN:
        case GBAGG_AGG_MAX: a2.push_back({g.kind, col}); break;
        default: a2.push_back({GBAGG_AGG_SUM, col}); break;  // COUNT(*), COUNT, SUM; AVG's sum
        }
    }
    for (uint32_t a = 0; a < q->num_aggs; a++)
        if (q->aggs[a].kind == GBAGG_AGG_AVG) {  // AVG's count: a hidden output, removed below
            avg_hidden[a] = (int)a2.size();
            a2.push_back({GBAGG_AGG_SUM, (int32_t)(base + src_cnt[a])});
        }
    uint32_t gc2[GBAGG_MAX_GROUP_COLS];
    for (uint32_t g = 0; g < ng; g++) gc2[g] = g;
    gbagg_query_desc d2{};
    d2.num_columns = (uint32_t)t2.size();
    d2.column_types = t2.data();
    d2.column_nullable = n2.data();
    d2.num_group_cols = ng;
    d2.group_cols = gc2;
    d2.num_aggs = (uint32_t)a2.size();
    d2.aggs = a2.data();
    d2.topn_col = q->topn_col;
    d2.topn_ascending = q->topn_ascending;
    d2.topn_limit = q->topn_limit;
    d2.row_limit = q->row_limit;
    d2.drop_null_keys = q->drop_null_keys;
    d2.strategy = GBAGG_STRATEGY_AUTO;
    build_plan(h, &d2);
    // user-visible semantics of the phase-2 outputs (after the key columns,
    // or after the hidden tuple code when phase 2 groups by several columns)
    const uint32_t ko = h->tuple ? 1 : ng;
    bool topn_avg = false;
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        OutCol &oc = h->outs[ko + a];
        const gbagg_agg_kind k = q->aggs[a].kind;
        if (k == GBAGG_AGG_COUNT_STAR || k == GBAGG_AGG_COUNT) {
            oc.kind = OUT_COUNT;  // sums of counts: never NULL
            oc.cnt_acc = -1;
        } else if (k == GBAGG_AGG_AVG) {
            const OutCol &hid = h->outs[ko + avg_hidden[a]];
            oc.kind = oc.kind == OUT_SUM_F ? OUT_AVG_F : OUT_AVG_I;
            oc.type = GBAGG_TYPE_DOUBLE;
            oc.cnt_acc = hid.val_acc;
            topn_avg = topn_avg || (int32_t)(ng + a) == q->topn_col;
        }
    }
    h->outs.resize(ko + q->num_aggs);
    if (topn_avg) h->topn_device = false;  // AVG order: the full result is sorted on the host
    h->strategy = q->strategy == GBAGG_STRATEGY_DIRECT ? GBAGG_STRATEGY_AUTO : q->strategy;
}

ShapeSig plan_sig(const gbagg_handle *h) {
    ShapeSig g{};
    const DevQuery &q = h->dq;
    g.nu = q.nu;
    g.nkp = q.nkp;
    for (int u = 0; u < q.nu; u++) {
        g.kinds |= (uint32_t)q.ukind[u] << (2 * u);
        if (h->nullable[h->ucol[u]]) g.nulls |= 1u << u;
    }
    g.km = q.key_mode;
    g.ku0 = q.key_mode != KM_NONE ? q.key_u[0] : 0;
    g.ku1 = q.key_mode == KM_PACK32 || q.key_mode == KM_WIDE ? q.key_u[1] : 0;
    g.na = q.nacc;
    for (int a = 0; a < q.nacc; a++) {
        g.ops |= (uint64_t)q.acc_op[a] << (4 * a);
        g.aus |= (uint64_t)(q.acc_u[a] < 0 ? 15 : q.acc_u[a]) << (4 * a);
    }
    g.np = q.npred;
    for (int p = 0; p < q.npred; p++)
        g.pred |= (uint32_t)(q.pred[p].u | (q.pred[p].kind << 3) | (q.pred[p].op << 5)) << (9 * p);
    return g;
}

void pick_kernels(gbagg_handle *h) {
    h->kern = &dynamic_kernels();
    if (!h->allow_static || h->dq.npred > 2) return;
    const ShapeSig g = plan_sig(h);
    for (const auto &k : shape_table())
        if (k.sig == g) { h->kern = &k; return; }
}

void ensure_runtime(gbagg_handle *h) {
    if (h->stream) return;
    device_info();
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
    // the parameter block and the plan block side by side, host and device,
    // so one copy can upload both
    HIPCHK(hipHostMalloc((void **)&h->hctx, CTX_BYTES + sizeof(DevPlan), hipHostMallocDefault));
    std::memset((void *)h->hctx, 0, CTX_BYTES + sizeof(DevPlan));
    h->hplan = (DevPlan *)((char *)h->hctx + CTX_BYTES);
    h->ctx.ensure(CTX_BYTES + sizeof(DevPlan));
    h->dplan.ptr = (char *)h->ctx.ptr + CTX_BYTES;
    h->in_set_bufs.resize(h->preds.size());
    for (size_t i = 0; i < h->preds.size(); i++) {
        if (h->in_sets[i].empty()) continue;
        h->in_set_bufs[i].ensure(h->in_sets[i].size() * 8);
        HIPCHK(hipMemcpy(h->in_set_bufs[i].ptr, h->in_sets[i].data(), h->in_sets[i].size() * 8,
                         hipMemcpyHostToDevice));
    }
}

bool like_contains(const char *s, uint32_t l, const std::string &nd) {
    if (nd.empty()) return true;
    if (l < nd.size()) return false;
    return std::search(s, s + l, nd.begin(), nd.end()) != s + l;
}

// One batch's host-to-device copies: each piece gets its device memory when
// it is registered; commit gathers the bytes into a pinned staging buffer and
// copies them on the handle's stream, one copy when the pieces' device
// memory is contiguous. Host data that would not outlive the call is kept.
struct Stager {
    gbagg_handle *h;
    struct Piece {
        const void *src;
        size_t bytes;
        uint8_t *dst;
    };
    std::vector<Piece> pieces;
    std::vector<std::vector<uint8_t>> keep;
    void *to_dev(const void *src, size_t bytes, size_t pad = 0) {
        uint8_t *d = h->batch_alloc(bytes + pad);
        if (bytes) pieces.push_back(Piece{src, bytes, d});
        return d;
    }
    void *to_dev_kept(const void *src, size_t bytes, size_t pad = 0) {
        keep.emplace_back((const uint8_t *)src, (const uint8_t *)src + bytes);
        return to_dev(keep.back().data(), bytes, pad);
    }
    void commit() {
        if (pieces.empty()) return;
        size_t sum = 0;
        bool one = true;
        for (size_t i = 0; i < pieces.size(); i++) {
            sum += pieces[i].bytes;
            if (i && pieces[i].dst < pieces[i - 1].dst + pieces[i - 1].bytes) one = false;
        }
        const size_t span = (size_t)(pieces.back().dst + pieces.back().bytes - pieces.front().dst);
        if (one && span > sum + pieces.size() * 256) one = false;  // the arena moved on mid-batch
        size_t need = 0;
        if (one) {
            need = span;
        } else {
            for (const Piece &pc : pieces) need += (pc.bytes + 15) & ~(size_t)15;
        }
        const int k = h->hstage_next;
        h->hstage_next ^= 1;
        if (h->hstage_done[k]) HIPCHK(hipEventSynchronize(h->hstage_done[k]));  // its copies have read it
        else HIPCHK(hipEventCreateWithFlags(&h->hstage_done[k], hipEventDisableTiming));
        if (h->hstage_bytes[k] < need) {
            if (h->hstage[k]) HIPCHK(hipHostFree(h->hstage[k]));
            h->hstage[k] = nullptr;
            h->hstage_bytes[k] = 0;
            const size_t cap = std::max<size_t>(need, (size_t)4 << 20);
            HIPCHK(hipHostMalloc((void **)&h->hstage[k], cap, hipHostMallocDefault));
            h->hstage_bytes[k] = cap;
        }
        uint8_t *st = h->hstage[k];
        if (one) {
            for (const Piece &pc : pieces) std::memcpy(st + (pc.dst - pieces.front().dst), pc.src, pc.bytes);
            HIPCHK(hipMemcpyAsync(pieces.front().dst, st, span, hipMemcpyHostToDevice, h->stream));
        } else {
            size_t off = 0;
            for (const Piece &pc : pieces) {
                std::memcpy(st + off, pc.src, pc.bytes);
                HIPCHK(hipMemcpyAsync(pc.dst, st + off, pc.bytes, hipMemcpyHostToDevice, h->stream));
                off += (pc.bytes + 15) & ~(size_t)15;
            }
        }
        HIPCHK(hipEventRecord(h->hstage_done[k], h->stream));
    }
};

// Registers one batch; `device` tells whether data/validity are device pointers.
void add_batch(gbagg_handle *h, const gbagg_column_view *cols, uint64_t n, bool device) {
    if (h->finalized) fail(GBAGG_INVALID, "consume after finalize: call gbagg_reset first");
    if (!cols && n) fail(GBAGG_INVALID, "null columns");
    ensure_runtime(h);
    Batch b;
    b.nrows = n;
    b.db.nrows = n;
    Stager st{h};
    if (h->str_mode.size() != h->col_types.size()) h->str_mode.assign(h->col_types.size(), 0);
    std::vector<uint8_t> modes(h->str_mode);
    for (int u = 0; u < h->nu; u++) {
        if (h->tuple && (uint32_t)h->ucol[u] == h->ext_cols) {  // the tuple code: written at finalize
            b.db.data[u] = nullptr;
            b.db.valid[u] = nullptr;
            continue;
        }
        const gbagg_column_view &c = cols[h->ucol[u]];
        if (c.type != h->col_types[h->ucol[u]]) fail(GBAGG_INVALID, "column type differs from the plan");
        if (n && !c.data) fail(GBAGG_INVALID, "null column data");
        if (c.validity && !h->nullable[h->ucol[u]])
            fail(GBAGG_INVALID, "validity given for a column declared non-nullable");
        const bool strs = (c.flags & GBAGG_STR_RAW) != 0;
        if (strs && c.type != GBAGG_TYPE_VARCHAR) fail(GBAGG_INVALID, "GBAGG_STR_RAW is for VARCHAR columns");
        if (strs && (c.dict_offsets || c.dict_size)) fail(GBAGG_INVALID, "GBAGG_STR_RAW columns carry no dictionary");
        const bool ordered = (size_t)h->ucol[u] < h->str_ordered.size() && h->str_ordered[h->ucol[u]];
        if (c.type == GBAGG_TYPE_VARCHAR) {
            // a column with a batch of strings, or under MIN/MAX, is interned at finalize
            uint8_t &m = modes[h->ucol[u]];
            m = strs || ordered ? 2 : std::max<uint8_t>(m, 1);
        }
        if (ordered && !strs && n) {
            if (c.dict_size && (!c.dict_offsets || !c.dict_bytes)) fail(GBAGG_INVALID, "missing dictionary");
            b.odoff[u].assign(1, 0);
            if (c.dict_size) {
                b.odoff[u].assign(c.dict_offsets, c.dict_offsets + c.dict_size + 1);
                b.odbytes[u].assign(c.dict_bytes, c.dict_bytes + c.dict_offsets[c.dict_size]);
            }
        }
        size_t w = type_width(c.type);
        if (strs) {
            // offsets and bytes; the codes are written when the column is interned
            if (device || n == 0) {
                b.str_off[u] = (const uint32_t *)c.data;
                b.str_bytes[u] = (const uint8_t *)c.dict_bytes;
            } else {
                const uint32_t *off = (const uint32_t *)c.data;
                for (uint64_t r = 0; r < n; r++)
                    if (off[r + 1] < off[r]) fail(GBAGG_INVALID, "VARCHAR string offsets must be ascending");
                if (off[n] && !c.dict_bytes) fail(GBAGG_INVALID, "null string bytes");
                b.str_off[u] = (const uint32_t *)st.to_dev(off, (n + 1) * 4);
                b.str_bytes[u] = (const uint8_t *)st.to_dev(c.dict_bytes, off[n], 4);
            }
            b.db.data[u] = nullptr;
            if (device || n == 0) b.db.valid[u] = c.validity;
        } else if (device || n == 0) {
            b.db.data[u] = c.data;
            b.db.valid[u] = c.validity;
        } else {
            b.db.data[u] = st.to_dev(c.data, n * w);
        }
        if (!device && n && c.validity) b.db.valid[u] = (const uint8_t *)st.to_dev(c.validity, n);
    }
    // multi-column keys: every group column of the batch on the device
    if (h->tuple) {
        auto to_dev = [&](const void *p, size_t bytes) -> const void * {
            if (device || !n || !p) return p;
            return st.to_dev(p, bytes, 4);
        };
        b.tin.resize(h->tcols.size());
        for (size_t j = 0; j < h->tcols.size(); j++) {
            const uint32_t col = h->tcols[j];
            const gbagg_column_view &c = cols[col];
            if (c.type != h->col_types[col]) fail(GBAGG_INVALID, "column type differs from the plan");
            if (n && !c.data) fail(GBAGG_INVALID, "null column data");
            if (c.validity && !h->nullable[col]) fail(GBAGG_INVALID, "validity given for a column declared non-nullable");
            Batch::TupleIn &ti = b.tin[j];
            ti.valid = (const uint8_t *)to_dev(c.validity, n);
            if (c.type == GBAGG_TYPE_VARCHAR && (c.flags & GBAGG_STR_RAW)) {
                if (c.dict_offsets || c.dict_size) fail(GBAGG_INVALID, "GBAGG_STR_RAW columns carry no dictionary");
                const uint32_t *off = (const uint32_t *)c.data;
                if (!device && n) {
                    for (uint64_t r = 0; r < n; r++)
                        if (off[r + 1] < off[r]) fail(GBAGG_INVALID, "VARCHAR string offsets must be ascending");
                    if (off[n] && !c.dict_bytes) fail(GBAGG_INVALID, "null string bytes");
                }
                ti.str_off = (const uint32_t *)to_dev(c.data, (n + 1) * 4);
                ti.str_bytes = (const uint8_t *)to_dev(c.dict_bytes, !device && n ? off[n] : 0);
            } else {
                ti.data = to_dev(c.data, n * type_width(c.type));
                if (c.type == GBAGG_TYPE_VARCHAR) {
                    if (c.dict_size && (!c.dict_offsets || !c.dict_bytes)) fail(GBAGG_INVALID, "missing dictionary");
                    ti.doff.assign(1, 0);
                    if (c.dict_size) {
                        ti.doff.assign(c.dict_offsets, c.dict_offsets + c.dict_size + 1);
                        ti.dbytes.assign(c.dict_bytes, c.dict_bytes + c.dict_offsets[c.dict_size]);
                    }
                }
            }
        }
    }
    // VARCHAR group keys: canonical ids across batches
    for (size_t g = 0; g < h->group_cols.size(); g++) {
        if (!h->key_varchar[g]) continue;
        const gbagg_column_view &c = cols[h->group_cols[g]];
        if (c.flags & GBAGG_STR_RAW) continue;  // codes of the interned column
        KeyDict &kd = h->kdict[g];
        if (c.dict_size && (!c.dict_offsets || !c.dict_bytes)) fail(GBAGG_INVALID, "missing dictionary");
        b.db.key_dict_size[g] = c.dict_size;
        bool unique = c.flags & GBAGG_DICT_UNIQUE;
        if (kd.mode == 0) kd.mode = unique ? 1 : 2;
        if (kd.mode == 1) {
            if (!unique) fail(GBAGG_INVALID, "all batches must keep GBAGG_DICT_UNIQUE once set");
            if (kd.offsets == nullptr && kd.own_offsets.empty()) {
                kd.size = c.dict_size;
                if (device) {
                    kd.offsets = c.dict_offsets;
                    kd.bytes = c.dict_bytes;
                } else {
                    kd.own_offsets.assign(c.dict_offsets, c.dict_offsets + c.dict_size + 1);
                    kd.own_bytes.assign(c.dict_bytes, c.dict_bytes + c.dict_offsets[c.dict_size]);
                    kd.offsets = c.dict_offsets;
                    kd.bytes = c.dict_bytes;
                }
            } else if (kd.offsets != c.dict_offsets || kd.size != c.dict_size) {
                fail(GBAGG_INVALID, "GBAGG_DICT_UNIQUE batches must share one dictionary");
            }
            b.db.remap[g] = nullptr;
        } else {
            std::vector<uint32_t> remap(c.dict_size);
            for (uint32_t d = 0; d < c.dict_size; d++) {
                std::string s(c.dict_bytes + c.dict_offsets[d], c.dict_offsets[d + 1] - c.dict_offsets[d]);
                auto it = kd.map.find(s);
                if (it == kd.map.end()) {
                    uint32_t id = (uint32_t)kd.strings.size();
                    kd.map.emplace(s, id);
                    kd.strings.push_back(std::move(s));
                    remap[d] = id;
                } else {
                    remap[d] = it->second;
                }
            }
            b.db.remap[g] = (const uint32_t *)st.to_dev_kept(remap.data(), remap.size() * 4, 4);
        }
    }
    // predicates: IN sets and per-dictionary-entry VARCHAR matches
    for (size_t i = 0; i < h->preds.size(); i++) {
        const gbagg_predicate &p = h->preds[i];
        if (h->dq.pred[i].kind == PK_IN) {
            b.db.pred_ptr[i] = h->in_set_bufs[i].ptr;
            b.db.pred_count[i] = (uint32_t)h->in_sets[i].size();
        } else if (h->dq.pred[i].kind == PK_DICT) {
            const gbagg_column_view &c = cols[p.column];
            if (c.flags & GBAGG_STR_RAW) continue;  // matched per distinct string when interned
            if (c.dict_size && (!c.dict_offsets || !c.dict_bytes)) fail(GBAGG_INVALID, "missing dictionary");
            std::vector<uint8_t> m(std::max<uint32_t>(1, c.dict_size));
            const std::string &nd = h->pred_str[i];
            for (uint32_t d = 0; d < c.dict_size; d++) {
                const char *s = c.dict_bytes + c.dict_offsets[d];
                uint32_t l = c.dict_offsets[d + 1] - c.dict_offsets[d];
                bool r;
                if (p.op == GBAGG_CMP_LIKE || p.op == GBAGG_CMP_NOT_LIKE) {
                    r = like_contains(s, l, nd);
                    if (p.op == GBAGG_CMP_NOT_LIKE) r = !r;
                } else {
                    bool eq = l == nd.size() && (l == 0 || std::memcmp(s, nd.data(), l) == 0);
                    r = p.op == GBAGG_CMP_EQ ? eq : !eq;
                }
                m[d] = r;
            }
            b.db.pred_ptr[i] = st.to_dev_kept(m.data(), m.size());
            b.db.pred_count[i] = c.dict_size;
        }
    }
    // the caller's buffers are valid only during the call: gathered into
    // pinned staging here, copied to the device without a wait
    st.commit();
    h->str_mode = std::move(modes);
    h->rows_total += n;
    h->batches.push_back(std::move(b));
}

size_t lds_table_bytes(uint32_t S, int nacc) { return (size_t)S * 8 * (1 + nacc) + 16; }

// Largest power-of-two LDS table within `budget` bytes.
uint32_t lds_slots_for(size_t budget, int nacc, uint32_t smax) {
    uint32_t S = smax;
    while (S > 64 && lds_table_bytes(S, nacc) > budget) S >>= 1;
    return S;
}

// Largest LDS table of any size (a multiple of 64 slots) within `budget` bytes.
uint32_t lds_slots_exact(size_t budget, int nacc, uint32_t smax) {
    const size_t per = (size_t)8 * (1 + nacc);
    size_t S = budget > 16 ? (budget - 16) / per : 0;
    S = std::min<size_t>(smax, S / 64 * 64);
    return (uint32_t)std::max<size_t>(64, S);
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        HIPCHK(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

struct Timer {
    gbagg_handle *h;
    std::vector<hipEvent_t> evs;
    void mark() {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, h->stream));
        evs.push_back(e);
    }
    void truncate(size_t n) {
        while (evs.size() > n) {
            (void)hipEventDestroy(evs.back());
            evs.pop_back();
        }
    }
    float between(size_t a, size_t b) {
        float ms = 0;
        if (a < evs.size() && b < evs.size()) HIPCHK(hipEventElapsedTime(&ms, evs[a], evs[b]));
        return ms;
    }
    ~Timer() {
        for (auto e : evs) (void)hipEventDestroy(e);
    }
};

DevTable make_table(gbagg_handle *h) {
    DevTable T;
    T.rec = h->table.as<uint64_t>();
    T.stride = 1 + h->dq.nacc;
    T.max_probe = 4096;
    T.plan = h->dplan.as<DevPlan>();
    return T;
}

// Device addresses of the parameter blocks (pointer arithmetic only).
DevCtx *dctx(gbagg_handle *h) { return h->ctx.as<DevCtx>(); }
const DevQuery *d_q(gbagg_handle *h) { return &dctx(h)->q; }
const DevOut *d_o(gbagg_handle *h) { return &dctx(h)->o; }
const DevTopn *d_t(gbagg_handle *h) { return &dctx(h)->t; }
const RadixParams *d_rp(gbagg_handle *h) { return &dctx(h)->rp; }
const DevBatch *d_b(gbagg_handle *h, size_t i) { return h->dbatches.as<DevBatch>() + i; }

// Fills the output / TopN blocks of the host context.
void fill_out_topn(gbagg_handle *h) {
    DevOut &o = h->hctx->o;
    o = DevOut{};
    o.ncols = (int32_t)h->outs.size();
    o.key_mode = h->dq.key_mode;
    o.capacity = h->out_capacity;
    o.wide_acc[0] = h->wide_acc[0];
    o.wide_acc[1] = h->wide_acc[1];
    o.wide_comp = h->dq.wide_comp;
    o.wide_seed = h->dq.wide_seed;
    o.flags = &h->dplan.as<DevPlan>()->overflow;
    for (size_t c = 0; c < h->outs.size(); c++) {
        o.data[c] = h->out_data[c].ptr;
        o.valid[c] = h->out_valid[c].as<uint8_t>();
        o.kind[c] = h->outs[c].kind;
        o.val_acc[c] = h->outs[c].val_acc;
        o.cnt_acc[c] = h->outs[c].cnt_acc;
        o.key_type[c] = h->outs[c].key_type;
        const int lo = h->outs[c].lo_acc;
        o.lo_acc[c] = lo >= 0 && h->dq.acc_op[lo] == ACC_SUM_FL ? lo : -1;
        o.fx_shift[c] = lo >= 0 ? h->dq.acc_fx[lo] : 0;
    }
    DevTopn &t = h->hctx->t;
    t = h->tspec;
    t.enabled = h->topn_device ? 1 : 0;
    t.rec_words = 1 + h->dq.nacc;
    t.cand_rec = h->cand_rec.as<uint64_t>();
    t.cand_null = h->cand_null.as<uint8_t>();
    t.cand_cap = h->cand_cap;
    t.wide_acc[0] = h->wide_acc[0];
    t.wide_acc[1] = h->wide_acc[1];
    t.flags = o.flags;
    if (t.kind == TK_SUM_F || t.kind == TK_AVG_F) {
        t.fx_shift = h->dq.acc_fx[t.lo_acc];
        if (h->dq.acc_op[t.lo_acc] != ACC_SUM_FL) t.lo_acc = -1;
    }
}

// The host context is pinned; callers modify it only after the stream has
// consumed the previous upload (every caller syncs through read_plan first).
void upload_ctx(gbagg_handle *h) {
    HIPCHK(hipMemcpyAsync(h->ctx.ptr, h->hctx, sizeof(DevCtx), hipMemcpyHostToDevice, h->stream));
    h->ctx_stale = false;
}


// The batches' DevBatch blocks, then their first rows in the concatenation
// (nb + 1 words, the last one the total), in one copy.
void upload_batches(gbagg_handle *h) {
    const size_t nb = h->batches.size();
    const size_t slots = std::max<size_t>(1, nb);
    const size_t bytes = slots * sizeof(DevBatch) + (nb + 1) * 8;
    if (h->hbatch_cap < bytes) {
        if (h->hbatches) HIPCHK(hipHostFree(h->hbatches));
        h->hbatches = nullptr;
        HIPCHK(hipHostMalloc((void **)&h->hbatches, bytes, hipHostMallocDefault));
        h->hbatch_cap = bytes;
    }
    uint64_t *start = (uint64_t *)(h->hbatches + slots);
    uint64_t r = 0;
    for (size_t i = 0; i < nb; i++) {
        h->hbatches[i] = h->batches[i].db;
        start[i] = r;
        r += h->batches[i].nrows;
    }
    start[nb] = r;
    h->dbatches.ensure(bytes);
    HIPCHK(hipMemcpyAsync(h->dbatches.ptr, h->hbatches, bytes, hipMemcpyHostToDevice, h->stream));
}

// Launch rows over every batch: workgroup w takes `per` rows of the batches'
// concatenation (after upload_batches).
BatchRows batch_rows(gbagg_handle *h, uint64_t per) {
    const size_t nb = h->batches.size();
    BatchRows br{};
    br.b = h->dbatches.as<DevBatch>();
    br.start = (const uint64_t *)(h->dbatches.as<DevBatch>() + std::max<size_t>(1, nb));
    br.nb = (uint32_t)nb;
    br.per = std::max<uint64_t>(per, 1);
    return br;
}

// Workgroups of `unit`-row tiles over all rows: at most `max_wg` of them,
// each a whole number of tiles; per = rows per workgroup.
uint32_t tiled_grid(uint64_t rows, uint64_t unit, uint64_t max_wg, uint64_t &per) {
    const uint64_t tiles = std::max<uint64_t>(1, (rows + unit - 1) / unit);
    const uint64_t g = std::max<uint64_t>(1, std::min<uint64_t>(tiles, max_wg));
    per = (tiles + g - 1) / g * unit;
    return (uint32_t)std::max<uint64_t>(1, (rows + per - 1) / per);
}

// Resets the run counters of the device plan block; the sample / heavy-hitter
// statistics of this run (already read back into hplan) are kept.
void prepare_plan(gbagg_handle *h, uint64_t cap) {
    DevPlan keep = *h->hplan;
    std::memset(h->hplan, 0, sizeof(DevPlan));
    h->hplan->sample_rows = keep.sample_rows;
    h->hplan->sample_distinct = keep.sample_distinct;
    h->hplan->sample_f1 = keep.sample_f1;
    h->hplan->heavy_count = keep.heavy_count;
    h->hplan->heavy_sample_rows = keep.heavy_sample_rows;
    h->hplan->heavy_threshold = keep.heavy_threshold;
    h->hplan->est_groups = keep.est_groups;  // a count pass run ahead of the plan
    h->hplan->cap = cap;
    h->hplan->cap_max = h->cap_max;
}

void upload_plan(gbagg_handle *h, uint64_t cap) {
    prepare_plan(h, cap);
    HIPCHK(hipMemcpyAsync(h->dplan.ptr, h->hplan, sizeof(DevPlan), hipMemcpyHostToDevice, h->stream));
}

// Parameter block and plan block (as upload_plan) in one copy.
void upload_ctx_plan(gbagg_handle *h, uint64_t cap) {
    prepare_plan(h, cap);
    HIPCHK(hipMemcpyAsync(h->ctx.ptr, h->hctx, CTX_BYTES + sizeof(DevPlan), hipMemcpyHostToDevice, h->stream));
    h->ctx_stale = false;
}

void read_plan(gbagg_handle *h) {
    HIPCHK(hipMemcpyAsync(h->hplan, h->dplan.ptr, sizeof(DevPlan), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
}

// Global-table groups -> result (append) or TopN candidates.
void emit_table(gbagg_handle *h, const DevTable &T, uint64_t cap) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    if (h->topn_device) {
        uint32_t grid = (uint32_t)std::min<uint64_t>((cap + 2 + 1023) / 1024, (uint64_t)di.cus * 4);
        if (h->topn_k <= 16) hipLaunchKernelGGL(k_topn_table<16>, dim3(grid), dim3(256), 0, s, T, d_t(h));
        else hipLaunchKernelGGL(k_topn_table<32>, dim3(grid), dim3(256), 0, s, T, d_t(h));
    } else if (cap <= COMPACT_SMALL) {
        hipLaunchKernelGGL(k_compact_small, dim3(1), dim3(1024), 0, s, T, d_o(h));
    } else {
        uint64_t nblk = (cap + 2 + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
        uint32_t grid = (uint32_t)std::min<uint64_t>(nblk, (uint64_t)di.cus * 4);
        uint64_t *bases = (uint64_t *)(h->blkcnt.as<uint8_t>() + ((nblk * 4 + 15) / 16) * 16);
        hipLaunchKernelGGL(k_occ_count, dim3(grid), dim3(256), 0, s, T, h->blkcnt.as<uint32_t>());
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, T, h->blkcnt.as<uint32_t>(), bases);
        hipLaunchKernelGGL(k_compact_write, dim3(grid), dim3(256), 0, s, T, (const uint64_t *)bases, d_o(h));
    }
}

void topn_final(gbagg_handle *h) {
    if (!h->topn_device) return;
    // level-1 blocks cover the whole candidate buffer; the level-2 block
    // holds their lists (nlvl * KMAX <= 1024 * TOPN_NI2)
    const uint32_t nlvl = (h->cand_cap + TOPN_L1 - 1) / TOPN_L1;
    const uint32_t kmax = h->topn_k <= 16 ? 16 : 32;
    if ((uint64_t)nlvl * kmax > 1024ull * TOPN_NI2) fail(GBAGG_ERROR, "TopN candidate buffer too large");
    h->topn_lvl.ensure((size_t)nlvl * kmax * 4);
    uint32_t *lvl = h->topn_lvl.as<uint32_t>();
    if (kmax == 16) {
        hipLaunchKernelGGL(k_topn_level1<16>, dim3(nlvl), dim3(1024), 0, h->stream, d_t(h), h->dplan.as<DevPlan>(), lvl);
        hipLaunchKernelGGL(k_topn_final<16>, dim3(1), dim3(1024), 0, h->stream, d_t(h), h->dplan.as<DevPlan>(), d_o(h),
                           (const uint32_t *)lvl, nlvl);
    } else {
        hipLaunchKernelGGL(k_topn_level1<32>, dim3(nlvl), dim3(1024), 0, h->stream, d_t(h), h->dplan.as<DevPlan>(), lvl);
        hipLaunchKernelGGL(k_topn_final<32>, dim3(1), dim3(1024), 0, h->stream, d_t(h), h->dplan.as<DevPlan>(), d_o(h),
                           (const uint32_t *)lvl, nlvl);
    }
}

void ensure_table(gbagg_handle *h, uint64_t cap) {
    h->table.ensure((cap + 2) * (1 + (size_t)h->dq.nacc) * 8);
    uint64_t nblk = (cap + 2 + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
    h->blkcnt.ensure(((nblk * 4 + 15) / 16) * 16 + nblk * 8 + 64);
}

// Partition record layout: key, then each loaded slot an accumulator reads,
// then the validity word when an input is nullable.
void record_layout(gbagg_handle *h) {
    RadixParams &rp = h->hctx->rp;
    uint32_t words = 1;
    bool nullable_inputs = false;
    std::memset(rp.rec_pos, 0, sizeof(rp.rec_pos));
    for (int a = 0; a < h->dq.nacc; a++) {
        int u = h->dq.acc_u[a];
        if (u < 0 || rp.rec_pos[u]) continue;
        rp.rec_pos[u] = (int32_t)words++;
        if (h->nullable[h->ucol[u]]) nullable_inputs = true;
    }
    rp.vmask_word = nullable_inputs ? words++ : 0;
    rp.rec_words = words;
}

// Heavy-hitter sample over all batches (strided), then the heavy table.
void run_sample(gbagg_handle *h) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    RadixParams &rp = h->hctx->rp;
    const uint32_t S_TOTAL = 131072;
    rp.scap = 1u << 18;
    rp.hmax = h->heavy_max;
    // load factor <= 1/4: the longest probe among a wave's 64 lanes sets the
    // wave's lookup time
    rp.hslots = (uint32_t)next_pow2(4ull * std::max<uint32_t>(rp.hmax, 1));
    rp.heavy_min = h->heavy_min;
    // the heavy table must fit the scatter kernel's LDS next to its cursors
    // (8192 partitions at most) and its smallest record ring, and the count
    // kernel's LDS
    const size_t lds_max = di.lds_per_block;
    record_layout(h);
    auto fits = [&] {
        size_t k3 = scatter_lds_bytes(rp.hslots, h->dq.nacc, rp.hmax, 8192, 256, rp.rec_words);
        size_t k1 = (size_t)rp.hslots * 8 + 8192 * 4 + HLL_M * 4;
        return k3 <= lds_max && k1 <= lds_max;
    };
    while (rp.hmax > 0 && !fits()) {
        rp.hmax /= 2;
        rp.hslots = (uint32_t)next_pow2(4ull * std::max<uint32_t>(rp.hmax, 1));
    }
    h->skeys.ensure((size_t)rp.scap * 8);
    h->scnt.ensure((size_t)rp.scap * 4);
    h->hkeys.ensure((size_t)rp.hslots * 8);
    h->hidx.ensure((size_t)rp.hslots * 4);
    h->hdense.ensure((size_t)std::max<uint32_t>(rp.hmax, 1) * 8);
    h->hfinal.ensure((size_t)std::max<uint32_t>(rp.hmax, 1) * (1 + h->dq.nacc) * 8);
    rp.skeys = h->skeys.as<uint64_t>();
    rp.scnt = h->scnt.as<uint32_t>();
    rp.hkeys = h->hkeys.as<uint64_t>();
    rp.hidx = h->hidx.as<uint32_t>();
    rp.hdense = h->hdense.as<uint64_t>();
    rp.hfinal = h->hfinal.as<uint64_t>();
    upload_ctx(h);
    h->ghist.ensure(65 * 4);
    hipLaunchKernelGGL(k_sample_init, dim3(256), dim3(256), 0, s, d_rp(h), h->dplan.as<DevPlan>(),
                       h->ghist.as<uint32_t>());
    if (h->rows_total) {  // evenly strided over the batches' concatenation
        const uint64_t want = std::min<uint64_t>(S_TOTAL, h->rows_total);
        const uint64_t stride = std::max<uint64_t>(1, h->rows_total / want);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((want + 255) / 256, (uint64_t)di.cus * 2);
        hipLaunchKernelGGL(k_sample, dim3(grid), dim3(256), 0, s, (const DevCtx *)dctx(h), batch_rows(h, 1),
                           (uint32_t)want, stride, h->dplan.as<DevPlan>());
    }
    // the histogram runs in few blocks (each ends with device-scope atomics on
    // the same plan words); the insert takes one atomic per wave
    const uint32_t hgrid = (uint32_t)std::min<uint64_t>((rp.scap + 255) / 256, 128);
    const uint32_t sgrid = (uint32_t)std::min<uint64_t>((rp.scap + 255) / 256, 64);
    hipLaunchKernelGGL(k_heavy_hist, dim3(hgrid), dim3(256), 0, s, d_rp(h), h->dplan.as<DevPlan>(), h->ghist.as<uint32_t>());
    hipLaunchKernelGGL(k_heavy_insert, dim3(sgrid), dim3(256), 0, s, d_rp(h), h->dplan.as<DevPlan>(),
                       (const uint32_t *)h->ghist.as<uint32_t>(), (uint32_t)(h->strategy == GBAGG_STRATEGY_RADIX));
}

// Selector features of the sample just taken, and with a selector its
// outputs; one small copy back.
void run_features(gbagg_handle *h, bool infer) {
    uint32_t key_kind = 0;
    if (h->dq.key_mode == KM_WIDE) key_kind = 3;
    else if (h->dq.key_mode == KM_PACK32) key_kind = 2;
    else if (h->dq.key_mode == KM_SINGLE && h->dq.ukind[h->dq.key_u[0]] == CK_I64) key_kind = 1;
    hipLaunchKernelGGL(k_sel_plan, dim3(1), dim3(SEL_H), 0, h->stream, h->dplan.as<DevPlan>(), h->rows_total, key_kind,
                       infer ? (const float *)h->dsel.as<float>() : (const float *)nullptr);
}

// After read_plan: the features and outputs run_features left in the plan block.
void take_features(gbagg_handle *h) {
    std::memcpy(h->feat, h->hplan->feat, sizeof(h->feat));
    std::memcpy(h->pred, h->hplan->pred, sizeof(h->pred));
    h->feat_valid = true;
}

// Query shapes the built-in selector was trained on (one or two INTEGER /
// BIGINT key columns, no predicates, no TopN or row limit); for others an AUTO
// plan follows the default rules unless a selector was attached.
bool selector_in_domain(const gbagg_handle *h) {
    if (h->dq.npred || h->topn_device || h->topn_col >= 0 || h->row_limit || h->tuple || h->phase1) return false;
    if (h->dq.key_mode == KM_NONE) return false;
    for (uint32_t g = 0; g < h->group_cols.size(); g++)
        if (h->key_varchar[g]) return false;
    return true;
}

// ---- double sums ------------------------------------------------------------
// Fixed-point words (exact, order-independent) or, after a group flagged
// possible precision loss, IEEE double atomics (see AccOp).
void fx_mode(gbagg_handle *h, bool ieee) {
    if (h->fx_pairs.empty() || h->fx_ieee == ieee) return;
    for (const auto &pr : h->fx_pairs) {
        h->dq.acc_op[pr.first] = ieee ? ACC_SUM_F : ACC_SUM_FH;
        h->dq.acc_op[pr.second] = ieee ? ACC_CNT_STAR : ACC_SUM_FL;  // unused under atomics
    }
    h->fx_ieee = ieee;
    pick_kernels(h);
}

constexpr uint32_t FX_MARGIN = 4;  // bound over the sampled maximum: 16x

// The scale s of each fixed-point double SUM: |v| < 2^(63 - s) for every input,
// from the sampled maximum with a margin (a row beyond it sets OVF_FX and the
// query re-runs), or from the exact maximum over all rows (k_fx_range) when
// the sample saw no value of the column or after such a re-run. Updates and
// uploads the device context.
void fx_prepare(gbagg_handle *h, bool sampled) {
    if (h->fx_pairs.empty() || h->fx_ieee) return;
    const DeviceInfo &di = device_info();
    bool exact = h->fx_exact || h->fx_known || !sampled;
    for (const auto &pr : h->fx_pairs)
        if (sampled && h->hplan->fx_exp[h->dq.acc_u[pr.first]] == 0) exact = true;
    if (exact && !h->fx_known) {
        if (h->rows_total) {
            uint64_t per;
            const uint32_t g = tiled_grid(h->rows_total, 256, (uint64_t)di.cus * 8, per);
            hipLaunchKernelGGL(k_fx_range, dim3(g), dim3(256), 0, h->stream, d_q(h), batch_rows(h, per),
                               h->dplan.as<DevPlan>());
        }
        read_plan(h);
        std::memcpy(h->fx_exp, h->hplan->fx_exp, sizeof(h->fx_exp));
        h->fx_known = true;
    }
    for (const auto &pr : h->fx_pairs) {
        const int u = h->dq.acc_u[pr.first];
        const uint32_t e = exact ? h->fx_exp[u] : h->hplan->fx_exp[u] + FX_MARGIN;
        const int sc = e ? std::max(-1100, std::min(FX_SHIFT_MAX, 1085 - (int)e)) : 0;
        h->dq.acc_fx[pr.first] = h->dq.acc_fx[pr.second] = sc;
        h->hctx->q.acc_fx[pr.first] = h->hctx->q.acc_fx[pr.second] = sc;
    }
    fill_out_topn(h);
    h->ctx_stale = true;  // uploaded by the strategy that runs next
}

// The selector's outputs of the last sample (h->pred) as a distinct count and
// a path (sel_distinct, sel_path); true when the count is within 4x of the
// sample's own estimate, so the plan may follow it. Output 0 is
// log2((distinct + 1) / (default + 1)) / 4, where default is the sample's own
// estimate (keys seen plus the unseen share of the singletons).
bool selector_decision(gbagg_handle *h) {
    const DevPlan &p = *h->hplan;
    const double seen = (double)p.sample_distinct;
    const double rest = p.sample_rows ? (double)p.sample_f1 / p.sample_rows * (double)h->rows_total : 0;
    const double dflt = std::max(1.0, std::min((double)h->rows_total, seen + rest));
    const double pred = std::max(1.0, (dflt + 1.0) * std::exp2(4.0 * (double)h->pred[0]) - 1.0);
    int path = 0;
    for (int k = 1; k < SEL_CLS; k++)
        if (h->pred[1 + k] > h->pred[1 + path]) path = k;
    h->sel_distinct = pred;
    h->sel_path = path;
    return pred <= 4.0 * dflt && pred * 4.0 >= dflt;
}

void run_lds_hash(gbagg_handle *h, Timer &tm, bool global_only, uint64_t est = 0) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    // table size from the sample: distinct keys seen, plus a Good-Turing style
    // allowance for unseen keys (f1 / sample rows of the remaining rows)
    const DevPlan &hp = *h->hplan;
    double seen = (double)hp.sample_distinct;
    double rest = hp.sample_rows ? (double)hp.sample_f1 / hp.sample_rows * (double)h->rows_total : 0;
    uint64_t want = (uint64_t)(2.0 * (seen + rest)) + 1024;
    if (h->sel_used && h->sel_count) want = (uint64_t)(2.0 * h->sel_distinct) + 1024;  // the learned count sizes the table
    if (est) want = 2 * est + 1024;  // a HyperLogLog count over all rows
    if (h->dq.key_mode == KM_NONE) want = 1024;
    uint64_t cap = std::min<uint64_t>(h->cap_max, next_pow2(want));
    for (;;) {
        ensure_table(h, cap);
        DevTable T = make_table(h);
        if (h->ctx_stale) upload_ctx_plan(h, cap);
        else upload_plan(h, cap);
        uint32_t grid = (uint32_t)std::min<uint64_t>((cap + 2 + 255) / 256, (uint64_t)di.cus * 8);
        hipLaunchKernelGGL(k_table_init, dim3(grid), dim3(256), 0, s, T, d_q(h));
        tm.mark();  // aggregate start
        size_t lds_bytes = 0;
        uint32_t S = 0;
        if (!global_only) {
            size_t budget = std::min<size_t>(di.lds_per_cu / 2, di.lds_per_block);
            S = lds_slots_for(budget, h->dq.nacc, 4096);
            lds_bytes = lds_table_bytes(S, h->dq.nacc);
            allow_lds(h->kern->agg, lds_bytes);
        }
        h->lds_slots = S;
        if (h->rows_total) {  // one launch over every batch
            const uint64_t n = h->rows_total;
            const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
            uint64_t per;
            if (!global_only) {
                // few workgroups for small inputs: each flushes its LDS table
                uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / lds_bytes));
                uint64_t by_rows = std::max<uint64_t>(1, n / 16384);
                uint32_t g = tiled_grid(n, tile, std::min<uint64_t>((uint64_t)di.cus * per_cu, by_rows), per);
                h->agg_grid = g;
                hipLaunchKernelGGL(h->kern->agg, dim3(g), dim3(AGG_BLOCK), lds_bytes, s, d_q(h), batch_rows(h, per), T,
                                   S, S * 3 / 4);
            } else {
                uint32_t g = tiled_grid(n, tile, (uint64_t)di.cus * 8, per);
                h->agg_grid = g;
                hipLaunchKernelGGL(k_aggregate_global, dim3(g), dim3(AGG_BLOCK), 0, s, d_q(h), batch_rows(h, per), T);
            }
        }
        HIPCHK(hipGetLastError());
        tm.mark();  // aggregate end
        emit_table(h, T, cap);
        topn_final(h);
        tm.mark();  // emit end
        read_plan(h);
        if (h->hplan->overflow & 2) fail(GBAGG_ERROR, "TopN candidate buffer overflow");
        if (!(h->hplan->overflow & 1)) break;
        if (cap >= h->cap_max) fail(GBAGG_ERROR, "global hash table overflow at maximum size");
        cap = std::min<uint64_t>(h->cap_max, cap * 4);
        h->retries++;
        tm.truncate(2);
    }
}

// Returns false, with nothing aggregated, when the HyperLogLog count of cold
// groups exceeds what the largest fanout's LDS tables hold (est_out: the
// count); the caller then runs the global table.
// Launch shape of the radix count and scatter passes, and their buffers
// (parameter block fields; uploaded with it).
void radix_setup(gbagg_handle *h) {
    const DeviceInfo &di = device_info();
    const int nacc = h->dq.nacc;
    RadixParams &rp = h->hctx->rp;
    RadixLaunch &rl = h->rl;
    rp.bucket_bits = 13;
    const uint32_t nb = 1u << rp.bucket_bits;
    record_layout(h);
    // workgroup slots for count + scatter
    const size_t k3_fixed = scatter_lds_bytes(rp.hslots, nacc, rp.hmax, 0, 256, rp.rec_words);
    rl.k1_lds = (size_t)rp.hslots * 8 + (size_t)nb * 4 + HLL_M * 4;
    const size_t k3_lds_max = k3_fixed + (size_t)nb * 4;
    // The scatter pass runs as many workgroups per CU as its LDS allows; the
    // lighter count pass runs g1 times as many, each over 1/g1 of a scatter
    // workgroup's rows, so its histogram rows stay contiguous per scatter
    // workgroup.
    const uint32_t per_cu3 = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / k3_lds_max));
    const uint32_t per_cu1 = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / rl.k1_lds));
    const uint32_t g1 = std::max<uint32_t>(1, per_cu1 / per_cu3);
    // one launch per pass over every batch: scatter workgroup k takes rows
    // [k * per3, (k + 1) * per3) of the batches' concatenation, count
    // workgroup w rows [w * per1, (w + 1) * per1), per3 = g1 * per1
    const uint64_t tile = (uint64_t)RP_BLOCK * AGG_ROWS * 2 * g1;
    (void)tiled_grid(h->rows_total, tile, (uint64_t)di.cus * per_cu3, rl.per3);
    rl.per1 = rl.per3 / g1;
    rl.W1 = h->rows_total ? (uint32_t)((h->rows_total + rl.per1 - 1) / rl.per1) : 0;
    const uint32_t W = std::max<uint32_t>(rl.W1, 1);
    const uint32_t W3 = std::max<uint32_t>((rl.W1 + g1 - 1) / g1, 1);
    rp.W = W;
    rp.W3 = W3;
    rp.g1 = g1;
    h->hist.ensure((size_t)W * nb * 4);
    h->hll.ensure((size_t)W * HLL_M);
    h->hpart.ensure((size_t)W3 * nacc * std::max<uint32_t>(rp.hmax, 1) * 8);
    rp.hist = h->hist.as<uint32_t>();
    rp.hll = h->hll.as<uint8_t>();
    rp.hpart = h->hpart.as<uint64_t>();
    // every count workgroup writes its whole sketch row; with no rows the one
    // row the merge reads is cleared here
    if (h->rows_total == 0) HIPCHK(hipMemsetAsync(rp.hll, 0, (size_t)W * HLL_M, h->stream));
}

// The count pass and the HyperLogLog estimate of its keys (into the plan
// block). gate: see k_rp_count.
void radix_count(gbagg_handle *h, const uint32_t *gate) {
    hipStream_t s = h->stream;
    const RadixParams &rp = h->hctx->rp;
    const RadixLaunch &rl = h->rl;
    allow_lds(h->kern->count, rl.k1_lds);
    if (rl.W1)
        hipLaunchKernelGGL(h->kern->count, dim3(rl.W1), dim3(RP_BLOCK), rl.k1_lds, s, (const DevCtx *)dctx(h),
                           batch_rows(h, rl.per1), gate);
    const uint32_t hslices = std::min<uint32_t>(32, rp.W);
    h->hllm.ensure((size_t)HLL_M * hslices);
    hipLaunchKernelGGL(k_hll_merge, dim3(HLL_M / 256, hslices), dim3(256), 0, s, (const uint8_t *)rp.hll, rp.W,
                       h->hllm.as<uint8_t>(), gate);
    hipLaunchKernelGGL(k_hll_finish, dim3(1), dim3(1024), 0, s, (const uint8_t *)h->hllm.as<uint8_t>(), hslices,
                       h->dplan.as<DevPlan>(), gate);
}

// counted: radix_setup and radix_count already ran for this finalize and the
// plan block read back holds their estimate.
bool run_radix(gbagg_handle *h, Timer &tm, uint64_t &est_out, bool counted) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    const int nacc = h->dq.nacc;
    RadixParams &rp = h->hctx->rp;
    if (!counted) radix_setup(h);
    const RadixLaunch &rl = h->rl;
    const uint32_t W1 = rl.W1;
    const uint64_t per3 = rl.per3;
    const uint32_t W = rp.W, W3 = rp.W3;

    // fallback global table (small; NULL / SENTINEL keys and LDS overflow)
    uint64_t cap = h->radix_fallback_cap;
    for (;;) {
        ensure_table(h, cap);
        DevTable T = make_table(h);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((cap + 2 + 255) / 256, (uint64_t)di.cus * 8);
        const bool fresh = !counted;
        counted = false;  // a retry counts again
        if (fresh) {
            upload_ctx_plan(h, cap);
            hipLaunchKernelGGL(k_table_init, dim3(grid), dim3(256), 0, s, T, d_q(h));
            tm.mark();  // count start
            radix_count(h, nullptr);
            tm.mark();  // count end
            read_plan(h);  // HLL estimate and heavy count pick the fanout
        }
        // the count pass sketches the cold keys; heavy keys are counted exactly
        const uint64_t heavy = std::min<uint64_t>(h->hplan->heavy_count, h->hctx->rp.hmax);
        const uint64_t cold = std::max<uint64_t>(1, h->hplan->est_groups);
        const uint64_t est = cold + heavy;
        h->hplan->est_groups = est;  // reported as the estimate of all groups
        // TopN partitions: one workgroup per CU with a larger table measured
        // faster than two with smaller ones (composite 50M-row TopN: partition
        // pass 0.34 against 0.37 ms)
        const uint32_t lds_kb = h->k4_lds_kb ? h->k4_lds_kb : h->topn_device ? 156u : 78u;
        // the partition kernel's own static LDS comes out of the block's limit
        const void *kpart = !h->topn_device ? (const void *)h->kern->part_full
                            : h->topn_k <= 16 ? (const void *)h->kern->part16 : (const void *)h->kern->part32;
        hipFuncAttributes fa{};
        HIPCHK(hipFuncGetAttributes(&fa, kpart));
        const size_t k4_budget = std::min<size_t>(di.lds_per_block - std::min<size_t>(di.lds_per_block, fa.sharedSizeBytes),
                                                  (size_t)lds_kb * 1024);
        const uint32_t S = lds_slots_exact(k4_budget, nacc, h->topn_device ? RA_TOPN_SMAX : RA_SMAX);
        if (h->force_part_bits < 0 && (double)cold > 0.75 * (double)S * (double)(1u << rp.bucket_bits)) {
            est_out = est;
            return false;
        }
        // fanout: LDS tables at most 4/5 full on average (fewer partitions mean
        // fewer partially written record lines in the scatter; the per-partition
        // fill limit is 7/8, several standard deviations of a partition's group
        // count above the mean at these sizes), and at most 64K rows per
        // partition workgroup (few groups over many rows still spread over the CUs)
        uint32_t pbits = 0;
        while (pbits < rp.bucket_bits && ((double)cold / (double)(1u << pbits) > 0.8 * S ||
                                          (double)h->rows_total / (double)(1u << pbits) > 65536.0))
            pbits++;
        if (h->force_part_bits >= 0) pbits = std::min<uint32_t>((uint32_t)h->force_part_bits, rp.bucket_bits);
        rp.part_bits = pbits;
        const uint32_t P = 1u << pbits;
        h->radix_bits = pbits;
        h->lds_slots = S;
        h->base.ensure((size_t)W * P * 4);
        h->ptotal.ensure((size_t)P * 4);
        h->pstart.ensure((size_t)(P + 1) * 8);
        h->recs.ensure(std::max<uint64_t>(1, h->rows_total) * rp.rec_words * 8);
        rp.base = h->base.as<uint32_t>();
        rp.ptotal = h->ptotal.as<uint32_t>();
        rp.pstart = h->pstart.as<uint64_t>();
        rp.recs = h->recs.as<uint64_t>();
        // record ring: as deep as the LDS left next to the cursors allows
        uint32_t ring = 2048;
        while (ring > 256 && scatter_lds_bytes(rp.hslots, nacc, rp.hmax, P, ring, rp.rec_words) > di.lds_per_block)
            ring >>= 1;
        rp.ring = ring;
        rp.msplit_max = h->msplit_fanout;
        const size_t k3_lds = scatter_lds_bytes(rp.hslots, nacc, rp.hmax, P, ring, rp.rec_words);
        if (k3_lds > di.lds_per_block) fail(GBAGG_ERROR, "scatter kernel LDS exceeds the device limit");
        if (fresh) {
            upload_ctx(h);
        } else {  // the count ran ahead: parameters, plan and table set-up in one go
            upload_ctx_plan(h, cap);
            hipLaunchKernelGGL(k_table_init, dim3(grid), dim3(256), 0, s, T, d_q(h));
        }
        hipLaunchKernelGGL(k_rp_scan_w, dim3((P + SCAN_PARTS - 1) / SCAN_PARTS), dim3(SCAN_PARTS * SCAN_CHUNKS), 0, s,
                           d_rp(h));
        hipLaunchKernelGGL(k_rp_scan_p, dim3(1), dim3(1024), 0, s, d_rp(h), T.plan);
        allow_lds(h->kern->scatter, k3_lds);
        tm.mark();  // scatter start
        if (W1)
            hipLaunchKernelGGL(h->kern->scatter, dim3(W3), dim3(RP_BLOCK), k3_lds, s, (const DevCtx *)dctx(h),
                               batch_rows(h, per3), T);
        h->agg_grid = W3;
        tm.mark();  // scatter end
        const size_t k4_lds = lds_table_bytes(S, nacc);
        const uint32_t hgrid = h->topn_device ? 1 : std::max<uint32_t>(1, (rp.hmax + 255) / 256);
        const dim3 fgrid(std::max<uint32_t>(1, (rp.hmax + 63) / 64), (uint32_t)nacc);
        if (rp.hmax) hipLaunchKernelGGL(k_heavy_fold, fgrid, dim3(256), 0, s, (const DevCtx *)dctx(h), T);
        // partitions own the first P * K candidates; other producers append after them
        if (h->topn_device)
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, &T.plan->cand_count, P * h->topn_k);
        if (!h->topn_device) {
            allow_lds(h->kern->part_full, k4_lds);
            hipLaunchKernelGGL(h->kern->part_full, dim3(P), dim3(RA_BLOCK), k4_lds, s, (const DevCtx *)dctx(h), T, S, S * 7 / 8);
            if (rp.hmax) hipLaunchKernelGGL(k_heavy_emit<16>, dim3(hgrid), dim3(256), 0, s, (const DevCtx *)dctx(h), T);
        } else if (h->topn_k <= 16) {
            allow_lds(h->kern->part16, k4_lds);
            hipLaunchKernelGGL(h->kern->part16, dim3(P), dim3(RA_BLOCK), k4_lds, s, (const DevCtx *)dctx(h), T, S, S * 7 / 8);
            if (rp.hmax) hipLaunchKernelGGL(k_heavy_emit<16>, dim3(hgrid), dim3(256), 0, s, (const DevCtx *)dctx(h), T);
        } else {
            allow_lds(h->kern->part32, k4_lds);
            hipLaunchKernelGGL(h->kern->part32, dim3(P), dim3(RA_BLOCK), k4_lds, s, (const DevCtx *)dctx(h), T, S, S * 7 / 8);
            if (rp.hmax) hipLaunchKernelGGL(k_heavy_emit<32>, dim3(hgrid), dim3(256), 0, s, (const DevCtx *)dctx(h), T);
        }
        HIPCHK(hipGetLastError());
        tm.mark();  // partition aggregate end
        emit_table(h, T, cap);
        topn_final(h);
        tm.mark();  // emit end
        read_plan(h);
        if (h->hplan->overflow & OVF_RING) fail(GBAGG_ERROR, "scatter record ring wait bound exceeded");
        if (h->hplan->overflow & 2) fail(GBAGG_ERROR, "TopN candidate buffer overflow");
        if (!(h->hplan->overflow & 1)) break;
        if (cap >= h->cap_max) fail(GBAGG_ERROR, "fallback hash table overflow at maximum size");
        cap = std::min<uint64_t>(h->cap_max, cap * 16);
        h->retries++;
        tm.truncate(2);
    }
    return true;
}

// ---- direct-mapped strategy ---------------------------------------------------
constexpr uint64_t DIRECT_MAX_SPAN = 1ull << 22;

size_t direct_lds_budget() {
    const DeviceInfo &di = device_info();
    return std::min<size_t>(di.lds_per_cu / 2, di.lds_per_block);
}

// Widest key range whose dense table [nacc][span] fits one workgroup's LDS.
uint64_t direct_lds_span(const gbagg_handle *h) { return direct_lds_budget() / (8 * (size_t)h->dq.nacc); }

// Exact key range of the filtered rows (one pass over the key and filter
// columns): false for multi-column keys or a range wider than DIRECT_MAX_SPAN.
// The exact key range pass over every batch into out[0..1] (see k_key_range).
void launch_key_range(gbagg_handle *h, int64_t *out, uint64_t gate_span) {
    if (!h->rows_total) return;
    const DeviceInfo &di = device_info();
    uint64_t per;
    const uint32_t g = tiled_grid(h->rows_total, (uint64_t)AGG_BLOCK * AGG_ROWS, (uint64_t)di.cus * 2, per);
    hipLaunchKernelGGL(h->kern->range, dim3(g), dim3(AGG_BLOCK), 0, h->stream, d_q(h), batch_rows(h, per),
                       (const DevPlan *)h->dplan.as<DevPlan>(), out, gate_span);
}

bool direct_range(gbagg_handle *h, uint64_t &span, int64_t &kmin) {
    if (h->dq.key_mode != KM_SINGLE) return false;
    if (h->range_ahead) {
        // the pass ran with the sample (range_ahead_ready): its result came back with the plan
        h->hplan->key_min = h->hplan->rng_min;
        h->hplan->key_max = h->hplan->rng_max;
    } else {
        hipLaunchKernelGGL(k_key_range_init, dim3(1), dim3(1), 0, h->stream, h->dplan.as<DevPlan>());
        launch_key_range(h, &h->dplan.as<DevPlan>()->key_min, 0);
        HIPCHK(hipGetLastError());
        read_plan(h);
    }
    const DevPlan &p = *h->hplan;
    if (p.key_min > p.key_max) {  // no key: every row NULL-keyed or filtered out
        kmin = 0;
        span = 1;
        return true;
    }
    const uint64_t d = (uint64_t)p.key_max - (uint64_t)p.key_min;
    if (d >= DIRECT_MAX_SPAN) return false;
    kmin = p.key_min;
    span = d + 1;
    return true;
}

// Key word w owns table record w - kmin; NULL and SENTINEL keys keep their
// side records. A workgroup-private LDS table when [nacc][span] fits.
void run_direct(gbagg_handle *h, Timer &tm, uint64_t span, int64_t kmin) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    const uint64_t cap = std::max<uint64_t>(1024, next_pow2(span));
    ensure_table(h, cap);
    DevTable T = make_table(h);
    if (h->ctx_stale) upload_ctx_plan(h, cap);
    else upload_plan(h, cap);
    const uint32_t igrid = (uint32_t)std::min<uint64_t>((cap + 2 + 255) / 256, (uint64_t)di.cus * 8);
    hipLaunchKernelGGL(k_table_init, dim3(igrid), dim3(256), 0, s, T, d_q(h));
    tm.mark();  // aggregate start
    const size_t lds_bytes = (size_t)span * h->dq.nacc * 8;
    const bool lds = span < direct_lds_span(h);
    const uint32_t cache = lds ? 0 : lds_slots_for(direct_lds_budget(), h->dq.nacc, 4096);
    const size_t cache_bytes = lds ? 0 : lds_table_bytes(cache, h->dq.nacc);
    if (lds) allow_lds(h->kern->direct_lds, lds_bytes);
    else allow_lds(h->kern->direct_glb, cache_bytes);
    h->lds_slots = lds ? (uint32_t)span : 0;
    if (h->rows_total) {  // one launch over every batch
        const uint64_t n = h->rows_total;
        const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
        uint64_t per;
        uint32_t g;
        if (lds) {
            // each workgroup folds its whole table at the end: few of them for small inputs
            const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / std::max<size_t>(lds_bytes, 1)));
            const uint64_t by_rows = std::max<uint64_t>(1, n / std::max<uint64_t>(16384, span * 4));
            g = tiled_grid(n, tile, std::min<uint64_t>((uint64_t)di.cus * per_cu, by_rows), per);
            hipLaunchKernelGGL(h->kern->direct_lds, dim3(g), dim3(AGG_BLOCK), lds_bytes, s, d_q(h), batch_rows(h, per),
                               T, kmin, (uint32_t)span, 0u);
        } else {
            // an LDS hash cache for the frequent keys in front of the dense records
            const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / cache_bytes));
            const uint64_t by_rows = std::max<uint64_t>(1, n / 16384);
            g = tiled_grid(n, tile, std::min<uint64_t>((uint64_t)di.cus * per_cu, by_rows), per);
            hipLaunchKernelGGL(h->kern->direct_glb, dim3(g), dim3(AGG_BLOCK), cache_bytes, s, d_q(h),
                               batch_rows(h, per), T, kmin, (uint32_t)span, cache);
        }
        h->agg_grid = g;
    }
    HIPCHK(hipGetLastError());
    tm.mark();  // aggregate end
    emit_table(h, T, cap);
    topn_final(h);
    tm.mark();  // emit end
    read_plan(h);
    if (h->hplan->overflow & OVF_DIRECT) fail(GBAGG_ERROR, "direct-mapped strategy: a key outside its measured range");
    if (h->hplan->overflow & 2) fail(GBAGG_ERROR, "TopN candidate buffer overflow");
}

void finalize_once(gbagg_handle *h);

void run_phase1(gbagg_handle *h);

// Interns one GBAGG_STR_RAW column over all its batches: codes per row (the
// batches' column data), the per-code match tables of its VARCHAR predicates
// and, for a group column, its dictionary on the host.
// Interns rows [0, total) given by segs: codes per row in sc.codes, the
// dictionary in sc (tab, ncodes, words).
void intern_run(gbagg_handle *h, StrCol &sc, const std::vector<StrSeg> &segs, uint64_t total) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    uint64_t ctr[SC_N] = {};
    StrTab t{};
    sc.ncodes = 0;
    sc.words = 0;
    if (!total) {
        sc.tab = t;
        return;
    }
    // rows are 32-bit ids in the slots and lists, below the flag values
    if (total >= STR_HOME) fail(GBAGG_ERROR, "a GBAGG_STR_RAW column holds at most 2^32 - 3 rows");
    sc.segs.ensure(segs.size() * sizeof(StrSeg));
    HIPCHK(hipMemcpyAsync(sc.segs.ptr, segs.data(), segs.size() * sizeof(StrSeg), hipMemcpyHostToDevice, s));
    sc.hash.ensure(total * 8);
    sc.pos.ensure(total * 4);
    sc.codes.ensure(total * 4);
    sc.ctr.ensure(SC_N * 8);
    sc.plan.ensure(sizeof(DevPlan));
    t.ctr = sc.ctr.as<uint64_t>();
    t.key_mask = h->str_key_mask;
    t.seed = 0x5157C0DE5EEDull;
    const StrSeg *dsegs = sc.segs.as<StrSeg>();
    const uint32_t nseg = (uint32_t)segs.size();
    // 8 workgroups per CU (their sketches take 128 KB of a CU's LDS): 4 per
    // CU left k_str_hash's loads too few waves to overlap
    const uint32_t W = (uint32_t)std::min<uint64_t>((total + 255) / 256, (uint64_t)di.cus * 8);
    sc.hllp.ensure((size_t)W * HLL_M);
    const uint32_t hslices = std::min<uint32_t>(32, W);
    sc.hllm.ensure((size_t)HLL_M * hslices);
    HIPCHK(hipMemsetAsync(t.ctr, 0, SC_N * 8, s));
    HIPCHK(hipMemsetAsync(sc.plan.ptr, 0, sizeof(DevPlan), s));
    hipLaunchKernelGGL(k_str_hash, dim3(W), dim3(256), 0, s, dsegs, nseg, total, t, sc.hash.as<uint64_t>(),
                       sc.pos.as<uint32_t>(), sc.codes.as<uint32_t>(), sc.hllp.as<uint8_t>());
    hipLaunchKernelGGL(k_hll_merge, dim3(HLL_M / 256, hslices), dim3(256), 0, s, (const uint8_t *)sc.hllp.ptr, W,
                       sc.hllm.as<uint8_t>(), (const uint32_t *)nullptr);
    hipLaunchKernelGGL(k_hll_finish, dim3(1), dim3(1024), 0, s, (const uint8_t *)sc.hllm.ptr, hslices,
                       sc.plan.as<DevPlan>(), (const uint32_t *)nullptr);
    DevPlan pl{};
    HIPCHK(hipMemcpyAsync(ctr, t.ctr, SC_N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&pl, sc.plan.ptr, sizeof(DevPlan), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ctr[SC_FLAGS] & STR_BAD_OFFSETS) fail(GBAGG_INVALID, "VARCHAR string offsets must be ascending");
    const uint64_t rows = ctr[SC_ROWS], bytes = ctr[SC_BYTES];
    const uint64_t est = std::max<uint64_t>(1, pl.est_groups);
    // codes: the estimate with a margin (at most one per valid row); slots
    // at least twice the codes; arena words from the mean string length
    uint64_t code_cap = std::max<uint64_t>(1, std::min<uint64_t>(rows, est + est / 4 + 1024));
    uint64_t slots = next_pow2(std::max<uint64_t>(1024, 2 * code_cap));
    const uint64_t words_max = bytes / 4 + rows + 1;
    const uint64_t mean_words = rows ? bytes / rows / 4 + 2 : 1;
    uint64_t arena_words = std::min<uint64_t>(words_max, code_cap * mean_words * 3 / 2 + 1024);
    if (h->str_codes0) {  // small first tables: the resizes below run
        code_cap = std::min<uint64_t>(std::max<uint64_t>(1, rows), h->str_codes0);
        slots = next_pow2(h->str_codes0);
        arena_words = std::min<uint64_t>(words_max, code_cap * mean_words);
    }
    // the probe's workgroups all resident at once (no partial last wave of
    // workgroups): as many per CU as its registers allow
    static const int probe_per_cu = [] {
        int n = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_str_probe, 256, 0));
        return std::max(1, std::min(8, n));
    }();
    // the probe's row lists for k_str_check: per workgroup, room for the
    // rows it visits (256 per grid stride)
    const uint64_t rg_max = std::min<uint64_t>((total + 255) / 256, (uint64_t)di.cus * probe_per_cu);
    sc.list.ensure((total + rg_max * 256) * 4);
    sc.lcnt.ensure(rg_max * 4);
    for (uint32_t attempt = 0; rows; attempt++) {
        if (slots > (1ull << 30)) fail(GBAGG_ERROR, "string table larger than 2^30 slots");
        if (arena_words > 0xFFFFFFFFull) fail(GBAGG_ERROR, "distinct strings exceed 16 GiB");
        sc.slot.ensure(slots * sizeof(StrSlot));
        sc.arena.ensure(arena_words * 4);
        sc.code.ensure(code_cap * sizeof(StrCode));
        t.slot = sc.slot.as<StrSlot>();
        t.code = sc.code.as<StrCode>();
        t.arena = sc.arena.as<uint32_t>();
        t.mask = slots - 1;
        t.arena_words = arena_words;
        t.code_cap = code_cap;
        const uint32_t ig = (uint32_t)std::min<uint64_t>((slots + 255) / 256, (uint64_t)di.cus * 8);
        hipLaunchKernelGGL(k_str_init, dim3(ig), dim3(256), 0, s, t);
        sc.blk.ensure(4096 * 16);
        // k_str_hash left every valid row at its home slot; a restart puts them back
        if (attempt) {
            const uint32_t rg_all = (uint32_t)std::min<uint64_t>((total + 255) / 256, (uint64_t)di.cus * 8);
            hipLaunchKernelGGL(k_str_rehome, dim3(rg_all), dim3(256), 0, s, total, (const uint64_t *)sc.hash.ptr,
                               sc.pos.as<uint32_t>());
        }
        // a prefix of the rows first: its strings (the frequent ones above
        // all) then have codes when the other rows probe, which settle in
        // k_str_probe without the representative and check kernels
        const uint64_t prefix = std::max<uint64_t>(1u << 16, total / 16);
        std::vector<uint64_t> lims;
        if (prefix < total) lims.push_back(prefix);
        lims.push_back(total);
        for (size_t li = 0; li < lims.size(); li++) {
            const uint64_t lim = lims[li];
            const uint32_t rg = (uint32_t)std::min<uint64_t>((lim + 255) / 256, (uint64_t)di.cus * probe_per_cu);
            const uint64_t per_block = 256 * ((lim + (uint64_t)rg * 256 - 1) / ((uint64_t)rg * 256));
            // code assignment: blocks of contiguous rows, at most 4096 of them
            const uint64_t chunk = std::max<uint64_t>(256, ((lim + 4095) / 4096 + 255) / 256 * 256);
            const uint32_t nblk = (uint32_t)((lim + chunk - 1) / chunk);
            for (uint32_t round = 0;; round++) {
                hipLaunchKernelGGL(k_str_probe, dim3(rg), dim3(256), 0, s, dsegs, nseg, lim, t,
                                   (const uint64_t *)sc.hash.ptr, sc.pos.as<uint32_t>(), sc.codes.as<uint32_t>(),
                                   sc.list.as<uint32_t>(), sc.lcnt.as<uint32_t>(), per_block);
                hipLaunchKernelGGL(k_str_reps, dim3(nblk), dim3(256), 0, s, dsegs, nseg, lim, t,
                                   sc.pos.as<uint32_t>(), chunk, sc.blk.as<uint64_t>());
                hipLaunchKernelGGL(k_str_scan, dim3(1), dim3(1024), 0, s, sc.blk.as<uint64_t>(), nblk, t.ctr);
                hipLaunchKernelGGL(k_str_assign, dim3(nblk), dim3(256), 0, s, dsegs, nseg, lim, t,
                                   sc.pos.as<uint32_t>(), sc.codes.as<uint32_t>(), chunk,
                                   (const uint64_t *)sc.blk.ptr);
                hipLaunchKernelGGL(k_str_check, dim3(rg), dim3(256), 0, s, dsegs, nseg, t, sc.pos.as<uint32_t>(),
                                   sc.codes.as<uint32_t>(), (const uint32_t *)sc.list.ptr,
                                   (const uint32_t *)sc.lcnt.ptr, per_block, rg);
                // the prefix's rows left unresolved probe on in the next stage,
                // and a full table shows there too: no host round trip here
                if (li + 1 < lims.size()) break;
                HIPCHK(hipMemcpyAsync(ctr, t.ctr, SC_N * 8, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if ((ctr[SC_FLAGS] & (STR_FULL_TABLE | STR_FULL_DICT)) || !ctr[SC_UNRESOLVED]) break;
                // rows whose string differs from the code at their slot: probe on
                if (round > (1u << 20)) fail(GBAGG_ERROR, "string interning did not converge");
                HIPCHK(hipMemsetAsync(t.ctr + SC_UNRESOLVED, 0, 8, s));
            }
        }
        // ctr: as the last stage's last round left it
        if (!(ctr[SC_FLAGS] & (STR_FULL_TABLE | STR_FULL_DICT))) break;
        if (attempt == 16) fail(GBAGG_ERROR, "string interning: tables still full after 16 resizes");
        if (ctr[SC_FLAGS] & STR_FULL_TABLE) slots *= 2;
        if (ctr[SC_FLAGS] & STR_FULL_DICT) {
            code_cap = std::min<uint64_t>(rows, code_cap * 2);
            arena_words = std::min<uint64_t>(words_max, arena_words * 2);
            slots = std::max<uint64_t>(slots, next_pow2(2 * code_cap));
        }
    }
    sc.ncodes = rows ? ctr[SC_CODES] : 0;
    sc.words = rows ? ctr[SC_WORDS] : 0;
    sc.tab = t;
}

void strcol_host(gbagg_handle *h, const StrCol &sc, std::vector<StrCode> &cd, std::vector<uint32_t> &words);

void intern_column(gbagg_handle *h, int u) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    const uint32_t col = (uint32_t)h->ucol[u];
    if (h->strs.size() != h->col_types.size()) h->strs.resize(h->col_types.size());
    if (!h->strs[col]) h->strs[col] = std::make_unique<StrCol>();
    StrCol &sc = *h->strs[col];
    // batches of strings, and for a group column also its batches of
    // dictionary codes (read through the canonical dictionary), so that every
    // batch ends with codes of one dictionary
    int key_g = -1;
    for (size_t g = 0; g < h->group_cols.size() && key_g < 0; g++)
        if (h->group_cols[g] == col) key_g = (int)g;
    bool dict_batches = false;
    for (const Batch &b : h->batches) dict_batches |= b.nrows && !b.str_off[u];
    if (key_g >= 0 && dict_batches) {
        const KeyDict &kd = h->kdict[key_g];
        sc.coff_h.assign(1, 0);
        sc.cbytes_h.clear();
        if (kd.mode == 1) {
            const uint32_t *o = kd.own_offsets.empty() ? kd.offsets : kd.own_offsets.data();
            const char *by = kd.own_bytes.empty() ? kd.bytes : kd.own_bytes.data();
            sc.coff_h.assign(o, o + kd.size + 1);
            sc.cbytes_h.assign(by, by + o[kd.size]);
        } else if (kd.mode == 2) {
            for (const std::string &str : kd.strings) {
                sc.cbytes_h.insert(sc.cbytes_h.end(), str.begin(), str.end());
                if (sc.cbytes_h.size() > 0xFFFFFFFFull) fail(GBAGG_ERROR, "dictionary larger than 4 GiB");
                sc.coff_h.push_back((uint32_t)sc.cbytes_h.size());
            }
        }
        sc.coff.ensure(sc.coff_h.size() * 4);
        sc.cbytes.ensure(sc.cbytes_h.size() + 16);
        HIPCHK(hipMemcpyAsync(sc.coff.ptr, sc.coff_h.data(), sc.coff_h.size() * 4, hipMemcpyHostToDevice, s));
        if (!sc.cbytes_h.empty())
            HIPCHK(hipMemcpyAsync(sc.cbytes.ptr, sc.cbytes_h.data(), sc.cbytes_h.size(), hipMemcpyHostToDevice, s));
    }
    std::vector<StrSeg> segs;
    std::vector<int64_t> row0s(h->batches.size(), -1);
    uint64_t total = 0;
    for (size_t i = 0; i < h->batches.size(); i++) {
        const Batch &b = h->batches[i];
        if (!b.nrows) continue;
        const bool ordered = col < h->str_ordered.size() && h->str_ordered[col];
        if (b.str_off[u]) {
            segs.push_back(StrSeg{b.str_off[u], b.str_bytes[u], b.db.valid[u], nullptr, nullptr, total, b.nrows, 0, 0});
        } else if (key_g < 0 && ordered) {
            // codes under MIN/MAX: through the batch's own dictionary
            Batch &bm = h->batches[i];
            auto off = std::make_unique<DevBuf>(), by = std::make_unique<DevBuf>();
            off->ensure(bm.odoff[u].size() * 4);
            by->ensure(bm.odbytes[u].size() + 16);
            HIPCHK(hipMemcpyAsync(off->ptr, bm.odoff[u].data(), bm.odoff[u].size() * 4, hipMemcpyHostToDevice, s));
            if (!bm.odbytes[u].empty())
                HIPCHK(hipMemcpyAsync(by->ptr, bm.odbytes[u].data(), bm.odbytes[u].size(), hipMemcpyHostToDevice, s));
            segs.push_back(StrSeg{off->as<uint32_t>(), by->as<uint8_t>(), b.db.valid[u], (const uint32_t *)b.db.data[u],
                                  nullptr, total, b.nrows, (uint32_t)(bm.odoff[u].size() - 1), 0});
            bm.owned.push_back(std::move(off));
            bm.owned.push_back(std::move(by));
        } else if (key_g >= 0) {
            segs.push_back(StrSeg{sc.coff.as<uint32_t>(), sc.cbytes.as<uint8_t>(), b.db.valid[u],
                                  (const uint32_t *)b.db.data[u], b.db.remap[key_g], total, b.nrows,
                                  b.db.key_dict_size[key_g], 0});
        } else {
            continue;  // dictionary codes of a column that is not grouped by keep their own codes
        }
        row0s[i] = (int64_t)total;
        total += b.nrows;
    }
    if (total > (1ull << 30)) fail(GBAGG_INVALID, "more than 2^30 rows in one aggregate");
    sc.ncodes = 0;
    intern_run(h, sc, segs, total);
    if (col < h->str_ordered.size() && h->str_ordered[col] && sc.ncodes) {
        // codes in byte order of their strings: rank on the host, relabel rows
        // and code records on the device
        std::vector<StrCode> cd;
        std::vector<uint32_t> words;
        strcol_host(h, sc, cd, words);
        const char *a0 = (const char *)words.data();
        std::vector<uint32_t> order(cd.size()), rank(cd.size());
        for (uint32_t c = 0; c < order.size(); c++) order[c] = c;
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return std::string_view(a0 + (uint64_t)cd[x].word * 4, cd[x].len) <
                   std::string_view(a0 + (uint64_t)cd[y].word * 4, cd[y].len);
        });
        std::vector<StrCode> sorted(cd.size());
        for (uint32_t r = 0; r < order.size(); r++) {
            rank[order[r]] = r;
            sorted[r] = cd[order[r]];
        }
        auto drank = std::make_unique<DevBuf>();
        drank->ensure(rank.size() * 4);
        HIPCHK(hipMemcpyAsync(drank->ptr, rank.data(), rank.size() * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(sc.code.ptr, sorted.data(), sorted.size() * sizeof(StrCode), hipMemcpyHostToDevice, s));
        const uint32_t grid = (uint32_t)std::min<uint64_t>((total + 255) / 256, (uint64_t)di.cus * 8);
        hipLaunchKernelGGL(k_str_relabel, dim3(grid), dim3(256), 0, s, sc.codes.as<uint32_t>(), total,
                           (const uint32_t *)drank->ptr, sc.ncodes);
        HIPCHK(hipStreamSynchronize(s));  // rank and sorted are host vectors
    }
    const StrTab t = sc.tab;
    const uint32_t nc = (uint32_t)sc.ncodes;
    // predicates on this column: one match per distinct string
    sc.lit.resize(h->preds.size());
    sc.match.resize(h->preds.size());
    for (size_t i = 0; i < h->preds.size(); i++) {
        if (h->preds[i].column != col || h->dq.pred[i].kind != PK_DICT) continue;
        const std::string &lit = h->pred_str[i];
        const gbagg_cmp_op op = h->preds[i].op;
        const int mop = op == GBAGG_CMP_EQ ? 0 : op == GBAGG_CMP_NE ? 1 : op == GBAGG_CMP_LIKE ? 2 : 3;
        sc.lit[i].ensure(lit.size() + 1);
        if (!lit.empty()) HIPCHK(hipMemcpyAsync(sc.lit[i].ptr, lit.data(), lit.size(), hipMemcpyHostToDevice, s));
        sc.match[i].ensure(std::max<uint32_t>(1, nc));
        if (nc)
            hipLaunchKernelGGL(k_str_match, dim3((uint32_t)std::min<uint64_t>((nc + 255) / 256, (uint64_t)di.cus * 4)),
                               dim3(256), 0, s, t, (uint64_t)nc, (const uint8_t *)sc.lit[i].ptr,
                               (uint32_t)lit.size(), mop, sc.match[i].as<uint8_t>());
        for (size_t k = 0; k < h->batches.size(); k++)
            if (row0s[k] >= 0) {
                h->batches[k].db.pred_ptr[i] = sc.match[i].ptr;
                h->batches[k].db.pred_count[i] = nc;
            }
    }
    for (size_t i = 0; i < h->batches.size(); i++)
        if (row0s[i] >= 0) h->batches[i].db.data[u] = sc.codes.as<uint32_t>() + row0s[i];
    // group column: codes are the ids; the strings stay on the device until
    // result rows (or a second aggregation phase) need them
    for (size_t g = 0; g < h->group_cols.size(); g++) {
        if (h->group_cols[g] != col) continue;
        KeyDict &kd = h->kdict[g];
        kd = KeyDict();
        kd.mode = 4;
        kd.size = nc;
        kd.dev = &sc;
        for (Batch &b : h->batches) {
            b.db.key_dict_size[g] = std::max<uint32_t>(nc, 1);  // dropped rows carry STR_DROP
            b.db.remap[g] = nullptr;
        }
    }
}

// The whole dictionary of an interned group column, in code order, on the host.
void host_dict(gbagg_handle *h, KeyDict &kd) {
    if (kd.mode != 4) return;
    const StrCol &sc = *kd.dev;
    const uint64_t nc = sc.ncodes;
    hipStream_t s = h->stream;
    std::vector<StrCode> cd(nc);
    std::vector<uint32_t> words(std::max<uint64_t>(1, sc.words));
    if (nc) {
        HIPCHK(hipMemcpyAsync(cd.data(), sc.code.ptr, (size_t)nc * sizeof(StrCode), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(words.data(), sc.arena.ptr, (size_t)sc.words * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    kd.own_offsets.resize((size_t)nc + 1);
    kd.own_offsets[0] = 0;
    uint64_t nb = 0;
    for (uint64_t c = 0; c < nc; c++) {
        nb += cd[c].len;
        if (nb > 0xFFFFFFFFull) fail(GBAGG_ERROR, "distinct strings exceed 4 GiB");
        kd.own_offsets[c + 1] = (uint32_t)nb;
    }
    kd.own_bytes.resize(std::max<uint64_t>(1, nb));
    const char *src = (const char *)words.data();
    for (uint64_t c = 0; c < nc; c++)
        std::memcpy(kd.own_bytes.data() + kd.own_offsets[c], src + (uint64_t)cd[c].word * 4, cd[c].len);
    kd.offsets = kd.own_offsets.data();
    kd.bytes = kd.own_bytes.data();
    kd.mode = 1;
    kd.dev = nullptr;
}

// Strings of result rows whose cells are codes of an interned column,
// gathered on the device: only the rows' own strings cross to the host.
void gather_strings(gbagg_handle *h, const StrCol &sc, const uint32_t *ids, const uint8_t *valid, uint64_t n,
                    std::vector<uint32_t> &off, std::vector<char> &bytes) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    off.assign(n + 1, 0);
    bytes.clear();
    if (!n || !sc.ncodes) return;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)di.cus * 8);
    DevBuf dlen, doff, dbytes;
    dlen.ensure(n * 4);
    hipLaunchKernelGGL(k_str_lengths, dim3(grid), dim3(256), 0, s, sc.tab, ids, valid, n, dlen.as<uint32_t>());
    std::vector<uint32_t> len(n);
    HIPCHK(hipMemcpyAsync(len.data(), dlen.ptr, n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<uint64_t> o64(n);
    uint64_t tot = 0;
    for (uint64_t r = 0; r < n; r++) {
        o64[r] = tot;
        off[r] = (uint32_t)tot;
        tot += len[r];
        if (tot > 0xFFFFFFFFull) fail(GBAGG_ERROR, "result strings exceed 4 GiB");
    }
    off[n] = (uint32_t)tot;
    bytes.resize(tot);
    if (!tot) return;
    doff.ensure(n * 8);
    dbytes.ensure(tot);
    HIPCHK(hipMemcpyAsync(doff.ptr, o64.data(), n * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_str_gather, dim3(grid), dim3(256), 0, s, sc.tab, ids, valid, n,
                       (const uint64_t *)doff.ptr, dbytes.as<uint8_t>());
    HIPCHK(hipMemcpyAsync(bytes.data(), dbytes.ptr, tot, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

void gather_result_strings(gbagg_handle *h, size_t g, uint64_t n) {
    gather_strings(h, *h->kdict[g].dev, h->out_data[g].as<uint32_t>(), h->out_valid[g].as<uint8_t>(), n,
                   h->res_str_off[g], h->res_str_bytes[g]);
}

// Multi-column keys: VARCHAR group columns interned (each over all its
// batches, coded batches through their own dictionaries), every row's group
// columns packed, the packed rows interned; their codes fill the hidden
// key column of every batch.
void intern_tuples(gbagg_handle *h) {
    const DeviceInfo &di = device_info();
    hipStream_t s = h->stream;
    const size_t G = h->tcols.size();
    std::vector<uint64_t> row0(h->batches.size());
    uint64_t total = 0;
    for (size_t i = 0; i < h->batches.size(); i++) {
        row0[i] = total;
        total += h->batches[i].nrows;
    }
    if (total > (1ull << 30)) fail(GBAGG_INVALID, "more than 2^30 rows in one aggregate");
    for (size_t j = 0; j < G; j++) {
        if (h->ttypes[j] != GBAGG_TYPE_VARCHAR) continue;
        if (!h->tstr[j]) h->tstr[j] = std::make_unique<StrCol>();
        std::vector<StrSeg> segs;
        for (size_t i = 0; i < h->batches.size(); i++) {
            Batch &b = h->batches[i];
            if (!b.nrows) continue;
            Batch::TupleIn &ti = b.tin[j];
            if (ti.str_off) {
                segs.push_back(StrSeg{ti.str_off, ti.str_bytes, ti.valid, nullptr, nullptr, row0[i], b.nrows, 0, 0});
            } else {
                auto off = std::make_unique<DevBuf>(), by = std::make_unique<DevBuf>();
                off->ensure(ti.doff.size() * 4);
                by->ensure(ti.dbytes.size() + 16);
                HIPCHK(hipMemcpyAsync(off->ptr, ti.doff.data(), ti.doff.size() * 4, hipMemcpyHostToDevice, s));
                if (!ti.dbytes.empty())
                    HIPCHK(hipMemcpyAsync(by->ptr, ti.dbytes.data(), ti.dbytes.size(), hipMemcpyHostToDevice, s));
                segs.push_back(StrSeg{off->as<uint32_t>(), by->as<uint8_t>(), ti.valid, (const uint32_t *)ti.data,
                                      nullptr, row0[i], b.nrows, (uint32_t)(ti.doff.size() - 1), 0});
                b.owned.push_back(std::move(off));
                b.owned.push_back(std::move(by));
            }
        }
        intern_run(h, *h->tstr[j], segs, total);
    }
    // pack every row's group columns
    const uint32_t W = h->twords;
    h->tpack.ensure(std::max<uint64_t>(1, total) * W * 4);
    h->tkeep.ensure(std::max<uint64_t>(1, total));
    std::vector<TupleComp> comps(h->batches.size() * G);
    for (size_t i = 0; i < h->batches.size(); i++)
        for (size_t j = 0; j < G; j++) {
            const Batch::TupleIn &ti = h->batches[i].tin[j];
            TupleComp &c = comps[i * G + j];
            const gbagg_type t = h->ttypes[j];
            c.valid = ti.valid;
            c.word = h->twordof[j];
            if (t == GBAGG_TYPE_VARCHAR) {
                c.kind = TK_CODE;
                c.data = h->tstr[j]->codes.as<uint32_t>() + row0[i];
            } else {
                c.kind = t == GBAGG_TYPE_INTEGER ? TK_I32 : t == GBAGG_TYPE_BIGINT ? TK_I64 : TK_F64;
                c.data = ti.data;
            }
        }
    h->tcomp.ensure(std::max<size_t>(1, comps.size()) * sizeof(TupleComp));
    if (!comps.empty())
        HIPCHK(hipMemcpyAsync(h->tcomp.ptr, comps.data(), comps.size() * sizeof(TupleComp), hipMemcpyHostToDevice, s));
    std::vector<StrSeg> segs;
    for (size_t i = 0; i < h->batches.size(); i++) {
        const uint64_t n = h->batches[i].nrows;
        if (!n) continue;
        uint32_t *out = h->tpack.as<uint32_t>() + row0[i] * W;
        uint8_t *keep = h->tkeep.as<uint8_t>() + row0[i];
        const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)di.cus * 8);
        hipLaunchKernelGGL(k_tuple_pack, dim3(grid), dim3(256), 0, s, (const TupleComp *)h->tcomp.as<TupleComp>() + i * G,
                           (uint32_t)G, n, W, h->tdrop, out, keep);
        segs.push_back(StrSeg{nullptr, (const uint8_t *)out, keep, nullptr, nullptr, row0[i], n, 0, W * 4});
    }
    if (!h->tsc) h->tsc = std::make_unique<StrCol>();
    intern_run(h, *h->tsc, segs, total);
    // the hidden key column: tuple codes, rows dropped where keep is 0
    int ut = -1;
    for (int u = 0; u < h->nu; u++)
        if ((uint32_t)h->ucol[u] == h->ext_cols) ut = u;
    if (ut < 0) fail(GBAGG_ERROR, "multi-column key: hidden key column missing from the plan");
    for (size_t i = 0; i < h->batches.size(); i++) {
        if (!h->batches[i].nrows) continue;
        h->batches[i].db.data[ut] = h->tsc->codes.as<uint32_t>() + row0[i];
        h->batches[i].db.valid[ut] = h->tkeep.as<uint8_t>() + row0[i];
    }
}

void intern_strings(gbagg_handle *h) {
    h->ms_strings = 0;
    h->strings_distinct = 0;
    bool any = false;
    for (int u = 0; u < h->nu; u++) any |= (size_t)h->ucol[u] < h->str_mode.size() && h->str_mode[h->ucol[u]] == 2;
    if (!any) return;
    Timer tm{h, {}};
    tm.mark();
    for (int u = 0; u < h->nu; u++)
        if ((size_t)h->ucol[u] < h->str_mode.size() && h->str_mode[h->ucol[u]] == 2) {
            intern_column(h, u);
            h->strings_distinct += h->strs[h->ucol[u]]->ncodes;
        }
    tm.mark();
    HIPCHK(hipStreamSynchronize(h->stream));
    h->ms_strings = tm.between(0, 1);
}

void finalize(gbagg_handle *h) {
    if (h->finalized) fail(GBAGG_INVALID, "already finalized: call gbagg_reset first");
    if (h->phase1) run_phase1(h);
    ensure_runtime(h);
    intern_strings(h);
    if (h->tuple) {
        Timer tm{h, {}};
        tm.mark();
        intern_tuples(h);
        tm.mark();
        HIPCHK(hipStreamSynchronize(h->stream));
        h->ms_strings += tm.between(0, 1);
    }
    uint64_t rows = std::max<uint64_t>(h->rows_total, h->expected_rows);
    if (rows > (1ULL << 30)) fail(GBAGG_INVALID, "more than 2^30 rows in one aggregate");
    h->cap_max = std::max<uint64_t>(h->cap_max, next_pow2(std::max<uint64_t>(1024, rows + rows / 4 + 2)));
    // result buffers: every row may be its own group; TopN keeps K rows
    h->out_capacity = h->topn_device ? h->topn_k : h->rows_total + 2;
    h->out_data.resize(h->outs.size());
    h->out_valid.resize(h->outs.size());
    for (size_t c = 0; c < h->outs.size(); c++) {
        h->out_data[c].ensure(std::max<uint64_t>(1, h->out_capacity) * 8);
        h->out_valid[c].ensure(std::max<uint64_t>(1, h->out_capacity));
    }
    h->retries = 0;
    h->sel_used = 0;
    h->sel_count = false;
    h->sel_path = -1;
    h->sel_distinct = 0;
    h->feat_valid = false;
    h->radix_bits = 0;
    fx_mode(h, false);
    h->fx_exact = h->fx_known = false;
    h->fx_reruns = 0;
    // 16-byte keys rerun with the next seed while two groups share a key word;
    // double sums rerun with the exact bound (a value beyond the sampled one)
    // or with IEEE atomics (a group that may have lost precision)
    for (uint32_t attempt = 0, seed = 0;; attempt++) {
        h->dq.wide_seed = h->wide_seed0 + seed;
        finalize_once(h);
        read_plan(h);
        const uint32_t ov = h->hplan->overflow;
        if ((ov & OVF_FX) && !h->fx_ieee && !h->fx_exact) {
            h->fx_exact = true;
            h->fx_reruns++;
            continue;
        }
        if (ov & OVF_FX) fail(GBAGG_ERROR, "double SUM: an input beyond the exact fixed-point bound");
        if ((ov & OVF_FX_PREC) && !h->fx_ieee) {
            fx_mode(h, true);
            h->fx_reruns++;
            continue;
        }
        if (h->dq.key_mode != KM_WIDE || !(ov & OVF_WIDE)) break;
        if (++seed == 8) fail(GBAGG_ERROR, "16-byte group keys: key words still shared after 8 seeds");
        h->retries++;
    }
    h->result_rows = h->hplan->groups;
    if (h->result_rows > h->out_capacity) fail(GBAGG_ERROR, "result larger than its buffer");
    h->finalized = true;
    h->result_ready = false;
}

// COUNT(DISTINCT): phase 1 aggregates the consumed rows; its device-resident
// result columns become this handle's only input batch.
void strcol_host(gbagg_handle *h, const StrCol &sc, std::vector<StrCode> &cd, std::vector<uint32_t> &words);

// Multi-column keys: the result's group columns unpacked from the tuple
// codes into device columns (tout_data / tout_valid).
void unpack_tuples_device(gbagg_handle *h) {
    const size_t G = h->tcols.size();
    const uint64_t rows = h->result_rows;
    h->tout_data.resize(G);
    h->tout_valid.resize(G);
    std::vector<TupleOut> outs(G);
    for (size_t j = 0; j < G; j++) {
        const uint32_t words = h->ttypes[j] == GBAGG_TYPE_BIGINT || h->ttypes[j] == GBAGG_TYPE_DOUBLE ? 2 : 1;
        h->tout_data[j].ensure(std::max<uint64_t>(1, rows) * words * 4);
        h->tout_valid[j].ensure(std::max<uint64_t>(1, rows));
        outs[j] = TupleOut{h->tout_data[j].as<uint32_t>(), h->tout_valid[j].as<uint8_t>(), h->twordof[j], words};
    }
    if (!rows) return;
    h->touts.ensure(G * sizeof(TupleOut));
    HIPCHK(hipMemcpyAsync(h->touts.ptr, outs.data(), G * sizeof(TupleOut), hipMemcpyHostToDevice, h->stream));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((rows + 255) / 256, (uint64_t)device_info().cus * 8);
    hipLaunchKernelGGL(k_tuple_unpack, dim3(grid), dim3(256), 0, h->stream, (const uint32_t *)h->out_data[0].ptr, rows,
                       h->tsc->tab, (const TupleOut *)h->touts.ptr, (uint32_t)G);
    HIPCHK(hipStreamSynchronize(h->stream));  // the table of outputs is on the host stack
}

// Phase 1 result j as one phase-2 batch (see build_distinct): its group
// columns (unpacked on the device when they form a multi-column key) and, for
// j = 0, its partial states; phase-2 columns it does not fill are NULL.
void add_phase1_batch(gbagg_handle *h, gbagg_handle *p1, uint32_t j, const void *zeros) {
    const uint32_t ng = h->distinct_ng, m = h->distinct_args;
    const size_t G = p1->tuple ? p1->tcols.size() : 0;
    const size_t nc = p1->tuple ? G + p1->outs.size() - 1 : p1->outs.size();  // phase 1's columns
    if (p1->tuple) unpack_tuples_device(p1);
    const size_t ng2 = h->tuple ? h->tcols.size() : h->group_cols.size();
    const size_t n2 = h->tuple ? h->ext_cols : h->col_types.size();  // phase 2's columns
    auto type2 = [&](size_t c) { return h->col_types[c]; };
    std::vector<gbagg_column_view> v(n2);
    for (size_t c = 0; c < n2; c++) {  // NULL unless filled below
        v[c] = gbagg_column_view{};
        v[c].type = type2(c);
        v[c].data = zeros;
        v[c].validity = (const uint8_t *)zeros;
    }
    for (size_t c = 0; c < nc; c++) {
        // phase 1 column c: a group column, x_j (c == ng), or partial state c - ng - 1
        const size_t d = c < ng ? c : c == ng ? ng + j : c + (m - 1);
        if (d >= n2) continue;
        std::vector<uint32_t> &doff = h->p1_dict_off[(size_t)j * n2 + d];
        std::vector<char> &dbytes = h->p1_dict_bytes[(size_t)j * n2 + d];
        gbagg_column_view &w = v[d];
        w.type = type2(d);
        if (c < G) {
            w.data = p1->tout_data[c].ptr;
            w.validity = h->nullable[d] ? p1->tout_valid[c].as<uint8_t>() : nullptr;
            if (p1->ttypes[c] == GBAGG_TYPE_VARCHAR && d < ng2) {
                // the column's interned codes: its interned dictionary, as a unique one
                std::vector<StrCode> cd;
                std::vector<uint32_t> words;
                strcol_host(p1, *p1->tstr[c], cd, words);
                doff.assign(1, 0);
                const char *src = (const char *)words.data();
                for (const StrCode &e : cd) {
                    dbytes.insert(dbytes.end(), src + (uint64_t)e.word * 4, src + (uint64_t)e.word * 4 + e.len);
                    doff.push_back((uint32_t)dbytes.size());
                }
                if (dbytes.empty()) dbytes.push_back(0);
                w.dict_offsets = doff.data();
                w.dict_bytes = dbytes.data();
                w.dict_size = (uint32_t)cd.size();
                w.flags = m == 1 ? GBAGG_DICT_UNIQUE : 0;  // several phase 1 results: several dictionaries
            }
            continue;
        }
        const size_t ci = p1->tuple ? c - G + 1 : c;
        w.data = p1->out_data[ci].ptr;
        w.validity = h->nullable[d] ? p1->out_valid[ci].as<uint8_t>() : nullptr;
        if (p1->outs[ci].type == GBAGG_TYPE_VARCHAR && d < ng2) {
            // phase 1's canonical ids index its key dictionary: pass it on as a unique one
            KeyDict &kd = p1->kdict[ci];
            host_dict(p1, kd);
            const uint32_t n = kd.mode == 1 ? kd.size : (uint32_t)kd.strings.size();
            doff.assign(1, 0);
            for (uint32_t i = 0; i < n; i++) {
                const std::string str = kd.lookup(i);
                dbytes.insert(dbytes.end(), str.begin(), str.end());
                doff.push_back((uint32_t)dbytes.size());
            }
            if (dbytes.empty()) dbytes.push_back(0);
            w.dict_offsets = doff.data();
            w.dict_bytes = dbytes.data();
            w.dict_size = n;
            w.flags = m == 1 ? GBAGG_DICT_UNIQUE : 0;
        }
    }
    if (p1->result_rows) add_batch(h, v.data(), p1->result_rows, true);
}

// COUNT(DISTINCT): the phase 1 handles aggregate the consumed rows; their
// device-resident results become this handle's input batches.
void run_phase1(gbagg_handle *h) {
    std::vector<gbagg_handle *> p1s{h->phase1.get()};
    for (auto &p : h->phase1_more) p1s.push_back(p.get());
    uint64_t rows = 1;
    for (gbagg_handle *p1 : p1s) {
        finalize(p1);
        rows = std::max<uint64_t>(rows, p1->result_rows);
    }
    const size_t n2 = h->tuple ? h->ext_cols : h->col_types.size();
    h->p1_dict_off.assign(p1s.size() * n2, {});
    h->p1_dict_bytes.assign(p1s.size() * n2, {});
    ensure_runtime(h);
    const void *zeros = nullptr;
    if (p1s.size() > 1) {  // NULL columns: zero validity bytes, zero data
        h->p2_null.ensure(rows * 8);
        HIPCHK(hipMemsetAsync(h->p2_null.ptr, 0, rows * 8, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        zeros = h->p2_null.ptr;
    }
    h->batches.clear();
    h->rows_total = 0;
    for (auto &kd : h->kdict) kd = KeyDict();
    for (uint32_t j = 0; j < p1s.size(); j++) add_phase1_batch(h, p1s[j], j, zeros);
}

void finalize_once(gbagg_handle *h) {
    Timer tm{h, {}};
    tm.mark();  // start
    gbagg_strategy st = h->strategy;
    if (h->dq.key_mode == KM_NONE) st = GBAGG_STRATEGY_GLOBAL_HASH;
    if (h->topn_device) {
        // producers of K candidates each: every radix partition (<= 8192), the
        // heavy-key block, and the global-table blocks (<= 4 per CU)
        const uint64_t producers = 8192 + 1 + (uint64_t)device_info().cus * 4;
        h->cand_cap = (uint32_t)(producers * h->topn_k + 4096);
        h->cand_rec.ensure((size_t)h->cand_cap * (1 + h->dq.nacc) * 8);
        h->cand_null.ensure(h->cand_cap);
    }
    std::memset(h->hplan, 0, sizeof(DevPlan));
    std::memset(h->hctx, 0, sizeof(DevCtx));
    h->range_ahead = false;
    h->dq.fx_flags = &h->dplan.as<DevPlan>()->overflow;
    h->hctx->q = h->dq;
    fill_out_topn(h);
    upload_batches(h);
    upload_ctx_plan(h, 1024);
    uint64_t span = 0;
    int64_t kmin = 0;
    if (st == GBAGG_STRATEGY_DIRECT || (st == GBAGG_STRATEGY_GLOBAL_HASH && h->dq.key_mode == KM_NONE))
        fx_prepare(h, false);  // no sample on these paths
    if (st == GBAGG_STRATEGY_DIRECT) {
        tm.mark();  // no sample
        if (!direct_range(h, span, kmin)) {
            st = GBAGG_STRATEGY_AUTO;
            tm.truncate(1);
        }
    }
    if (st == GBAGG_STRATEGY_DIRECT) {
        run_direct(h, tm, span, kmin);
    } else if (st == GBAGG_STRATEGY_GLOBAL_HASH && h->dq.key_mode == KM_NONE) {
        tm.mark();
        run_lds_hash(h, tm, true);
    } else {
        run_sample(h);
        // an attached selector, or for AUTO plans of the shapes it was trained
        // on the built-in one
        if (h->sel_params.empty() && h->sel_builtin && st == GBAGG_STRATEGY_AUTO && selector_in_domain(h)) {
            h->sel_params.assign(kSelectorBuiltin, kSelectorBuiltin + SEL_NPARAM);
            h->dsel.ensure((size_t)SEL_NPARAM * 4);
            HIPCHK(hipMemcpyAsync(h->dsel.ptr, h->sel_params.data(), (size_t)SEL_NPARAM * 4, hipMemcpyHostToDevice,
                                  h->stream));
        }
        const bool sel = !h->sel_params.empty();
        if (sel || h->collect_features) run_features(h, sel);
        tm.mark();  // sample end
        // Inputs of many rows: the radix count pass runs before the plan is
        // read back, so the host's round trip overlaps it and the count's
        // estimate returns with the sample's statistics (no second round trip).
        // Under AUTO it is gated on the device by the default plan's radix test
        // (DevPlan::radix_gate); it is wasted only when the selector then picks
        // another path.
        const bool spec = (st == GBAGG_STRATEGY_AUTO || st == GBAGG_STRATEGY_RADIX) &&
                          h->rows_total >= h->radix_ahead_rows && h->dq.key_mode != KM_NONE;
        // Likewise the exact key range of a one-column key, which a
        // low-cardinality plan needs to go direct-mapped: gated on the device
        // by the sample (few distinct keys, sampled range within a
        // workgroup's dense table), so it runs when that plan is likely.
        const bool spec_range = st == GBAGG_STRATEGY_AUTO && h->dq.key_mode == KM_SINGLE && h->rows_total > 0;
        if (spec_range) launch_key_range(h, &h->dplan.as<DevPlan>()->rng_min, direct_lds_span(h));
        if (spec) {
            radix_setup(h);
            upload_ctx(h);
            tm.mark();  // count start
            radix_count(h, st == GBAGG_STRATEGY_RADIX ? nullptr : &h->dplan.as<DevPlan>()->radix_gate);
            tm.mark();  // count end
        }
        read_plan(h);
        const bool counted = spec && (st == GBAGG_STRATEGY_RADIX || h->hplan->radix_gate);
        {
            const DevPlan &p = *h->hplan;  // the same test as the range pass's gate
            h->range_ahead = spec_range && !p.radix_gate && p.key_min <= p.key_max &&
                             (uint64_t)p.key_max - (uint64_t)p.key_min < direct_lds_span(h);
        }
        if (sel || h->collect_features) take_features(h);
        fx_prepare(h, true);
        if (st == GBAGG_STRATEGY_AUTO) {
            const DevPlan &p = *h->hplan;
            bool low_card = p.sample_distinct <= 4096 && p.sample_f1 * 8 <= p.sample_distinct + 8;
            st = (low_card || h->rows_total <= 65536) ? GBAGG_STRATEGY_LDS_HASH : GBAGG_STRATEGY_RADIX;
            // a low-cardinality one-column key whose sampled range fits a
            // workgroup's LDS as a dense table goes direct-mapped
            if (st == GBAGG_STRATEGY_LDS_HASH && h->dq.key_mode == KM_SINGLE && p.key_min <= p.key_max &&
                (uint64_t)p.key_max - (uint64_t)p.key_min < direct_lds_span(h))
                st = GBAGG_STRATEGY_DIRECT;
            if (sel) {
                // the selector's path and count, unless they disagree with the
                // sample's own estimate by more than 4x (then the default plan)
                // RADIX sizes itself from the count pass and DIRECT from the
                // exact key range: their path is taken whatever the count
                const bool count_ok = selector_decision(h);
                h->sel_count = count_ok;
                if (count_ok || h->sel_path == 0 || h->sel_path == 3) {
                    const int path = h->sel_path;
                    h->sel_used = 1;
                    static const gbagg_strategy paths[SEL_CLS] = {GBAGG_STRATEGY_DIRECT, GBAGG_STRATEGY_LDS_HASH,
                                                                  GBAGG_STRATEGY_GLOBAL_HASH, GBAGG_STRATEGY_RADIX};
                    st = paths[path];
                }
            }
            // direct-mapped needs the exact range: too wide, and the
            // low-cardinality hash path runs instead
            if (st == GBAGG_STRATEGY_DIRECT && !direct_range(h, span, kmin)) st = GBAGG_STRATEGY_LDS_HASH;
        }
        uint64_t est = 0;
        // the count's marks stay only for a radix plan that uses its pass
        if (spec && !(st == GBAGG_STRATEGY_RADIX && counted)) tm.truncate(2);
        if (st == GBAGG_STRATEGY_DIRECT) {
            run_direct(h, tm, span, kmin);
        } else if (st == GBAGG_STRATEGY_RADIX && !run_radix(h, tm, est, counted)) {
            // more groups than the partitions' LDS tables hold: one global table
            st = GBAGG_STRATEGY_GLOBAL_HASH;
            run_lds_hash(h, tm, true, est);
        } else if (st != GBAGG_STRATEGY_RADIX) {
            run_lds_hash(h, tm, st == GBAGG_STRATEGY_GLOBAL_HASH);
        }
    }
    h->used_strategy = st;
    // phase times from the recorded events
    const size_t ne = tm.evs.size();
    h->ms_total = tm.between(0, ne - 1);
    h->ms_sample = tm.between(0, 1);
    if (st == GBAGG_STRATEGY_RADIX && ne >= 8) {
        // start, sample, count start, count end, scatter start, scatter end, partition end, emit end
        h->ms_count = tm.between(2, 3);
        h->ms_agg = tm.between(4, 5);
        h->ms_partition = tm.between(5, 6);
        h->ms_emit = tm.between(6, 7);
    } else if (ne >= 5) {
        h->ms_count = 0;
        h->ms_agg = tm.between(ne - 3, ne - 2);
        h->ms_partition = 0;
        h->ms_emit = tm.between(ne - 2, ne - 1);
    }
}

// The learned selector's plan for the consumed rows without aggregating them:
// the AUTO plan's sample, its features and the MLP outputs. Returns false when
// the prediction is more than 4x from the sample's own estimate.
bool selector_plan(gbagg_handle *h, double &distinct, int &path) {
    ensure_runtime(h);
    std::memset(h->hplan, 0, sizeof(DevPlan));
    std::memset(h->hctx, 0, sizeof(DevCtx));
    h->dq.fx_flags = &h->dplan.as<DevPlan>()->overflow;
    h->hctx->q = h->dq;
    upload_batches(h);
    upload_ctx_plan(h, 1024);
    run_sample(h);
    if (h->sel_params.empty()) {
        h->sel_params.assign(kSelectorBuiltin, kSelectorBuiltin + SEL_NPARAM);
        h->dsel.ensure((size_t)SEL_NPARAM * 4);
        HIPCHK(hipMemcpyAsync(h->dsel.ptr, h->sel_params.data(), (size_t)SEL_NPARAM * 4, hipMemcpyHostToDevice,
                              h->stream));
    }
    run_features(h, true);
    read_plan(h);
    take_features(h);
    const bool ok = selector_decision(h);
    distinct = h->sel_distinct;
    path = h->sel_path;
    return ok;
}

// A StrCol's dictionary on the host: per code its first arena word and length,
// and the arena.
void strcol_host(gbagg_handle *h, const StrCol &sc, std::vector<StrCode> &cd, std::vector<uint32_t> &words) {
    cd.resize(sc.ncodes);
    words.assign(std::max<uint64_t>(1, sc.words), 0);
    if (!sc.ncodes) return;
    HIPCHK(hipMemcpyAsync(cd.data(), sc.code.ptr, sc.ncodes * sizeof(StrCode), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(words.data(), sc.arena.ptr, sc.words * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
}

// Multi-column keys: the result's first column holds tuple codes; replace it
// by the group columns unpacked from the tuples.
void expand_tuples(gbagg_handle *h, uint64_t rows) {
    const size_t G = h->tcols.size(), nin = h->outs.size(), nout = G + nin - 1;
    std::vector<StrCode> tcd;
    std::vector<uint32_t> tw;
    strcol_host(h, *h->tsc, tcd, tw);
    std::vector<std::vector<StrCode>> scd(G);
    std::vector<std::vector<uint32_t>> sw(G);
    for (size_t j = 0; j < G; j++)
        if (h->ttypes[j] == GBAGG_TYPE_VARCHAR) strcol_host(h, *h->tstr[j], scd[j], sw[j]);
    std::vector<std::vector<uint8_t>> data(nout), valid(nout);
    std::vector<std::vector<uint32_t>> soff(nout);
    std::vector<std::vector<char>> sbytes(nout);
    const uint32_t *codes = (const uint32_t *)h->res_data[0].data();
    for (size_t j = 0; j < G; j++) {
        const gbagg_type t = h->ttypes[j];
        const size_t w = type_width(t);
        data[j].assign(std::max<uint64_t>(1, rows) * w, 0);
        valid[j].assign(std::max<uint64_t>(1, rows), 0);
        if (t == GBAGG_TYPE_VARCHAR) soff[j].assign(rows + 1, 0);
        for (uint64_t r = 0; r < rows; r++) {
            const uint32_t c = codes[r];
            if (c >= tcd.size()) fail(GBAGG_ERROR, "multi-column key: tuple code out of range");
            const uint32_t *tup = tw.data() + tcd[c].word;
            const bool v = (tup[j / 4] >> (8 * (j % 4))) & 0xFF;
            valid[j][r] = v;
            const uint32_t *val = tup + h->twordof[j];
            if (t == GBAGG_TYPE_VARCHAR) {
                soff[j][r] = (uint32_t)sbytes[j].size();
                if (v) {
                    const uint32_t sc = val[0];
                    if (sc >= scd[j].size()) fail(GBAGG_ERROR, "multi-column key: string code out of range");
                    const char *p = (const char *)(sw[j].data() + scd[j][sc].word);
                    sbytes[j].insert(sbytes[j].end(), p, p + scd[j][sc].len);
                }
                std::memcpy(&data[j][r * 4], val, 4);
            } else {
                std::memcpy(&data[j][r * w], val, w);
            }
        }
        if (t == GBAGG_TYPE_VARCHAR) soff[j][rows] = (uint32_t)sbytes[j].size();
    }
    std::vector<gbagg_type> types(nout);
    for (size_t j = 0; j < G; j++) types[j] = h->ttypes[j];
    for (size_t c = 1; c < nin; c++) {
        data[G + c - 1] = std::move(h->res_data[c]);
        valid[G + c - 1] = std::move(h->res_valid[c]);
        soff[G + c - 1] = std::move(h->res_str_off[c]);
        sbytes[G + c - 1] = std::move(h->res_str_bytes[c]);
        types[G + c - 1] = h->outs[c].type;
    }
    h->res_data = std::move(data);
    h->res_valid = std::move(valid);
    h->res_str_off = std::move(soff);
    h->res_str_bytes = std::move(sbytes);
    h->res_cols.assign(nout, {});
    for (size_t c = 0; c < nout; c++) {
        gbagg_result_column &rc = h->res_cols[c];
        rc.type = types[c];
        rc.data = h->res_data[c].data();
        rc.validity = h->res_valid[c].data();
        if (rc.type == GBAGG_TYPE_VARCHAR) {
            if (h->res_str_bytes[c].empty()) h->res_str_bytes[c].push_back(0);
            rc.str_offsets = h->res_str_off[c].data();
            rc.str_bytes = h->res_str_bytes[c].data();
        }
    }
}

void download_result(gbagg_handle *h) {
    if (!h->finalized) fail(GBAGG_INVALID, "gbagg_result before gbagg_finalize");
    if (h->result_ready) return;
    uint64_t n = h->result_rows;
    const size_t nc = h->outs.size();
    bool scalar_empty = h->dq.key_mode == KM_NONE && n == 0;
    uint64_t rows = scalar_empty ? 1 : n;
    h->res_data.assign(nc, {});
    h->res_valid.assign(nc, {});
    h->res_str_off.assign(nc, {});
    h->res_str_bytes.assign(nc, {});
    for (size_t c = 0; c < nc; c++) {
        size_t w = (h->outs[c].type == GBAGG_TYPE_INTEGER || h->outs[c].type == GBAGG_TYPE_VARCHAR) ? 4 : 8;
        h->res_data[c].assign(rows * w, 0);
        h->res_valid[c].assign(rows, 0);
        if (n) {
            HIPCHK(hipMemcpyAsync(h->res_data[c].data(), h->out_data[c].ptr, n * w, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipMemcpyAsync(h->res_valid[c].data(), h->out_valid[c].ptr, n, hipMemcpyDeviceToHost, h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (scalar_empty) {  // SQL: an aggregate without GROUP BY over no rows is one row
        for (size_t c = 0; c < nc; c++) h->res_valid[c][0] = h->outs[c].kind == OUT_COUNT;
    }
    // VARCHAR keys: ids -> strings
    for (size_t g = 0; g < h->group_cols.size(); g++) {
        if (!h->key_varchar[g]) continue;
        if (h->kdict[g].mode == 4) {
            gather_result_strings(h, g, n);
            if (scalar_empty) h->res_str_off[g].assign(rows + 1, 0);
            continue;
        }
        const uint32_t *ids = (const uint32_t *)h->res_data[g].data();
        auto &off = h->res_str_off[g];
        auto &bytes = h->res_str_bytes[g];
        off.assign(rows + 1, 0);
        for (uint64_t r = 0; r < rows; r++) {
            off[r] = (uint32_t)bytes.size();
            if (h->res_valid[g][r]) {
                std::string s = h->kdict[g].lookup(ids[r]);
                bytes.insert(bytes.end(), s.begin(), s.end());
            }
        }
        off[rows] = (uint32_t)bytes.size();
    }
    // MIN/MAX over VARCHAR: codes of the argument's interned column
    for (size_t c = h->group_cols.size(); c < nc; c++) {
        if (h->outs[c].type != GBAGG_TYPE_VARCHAR) continue;
        const uint32_t col = (uint32_t)h->aggs[c - h->group_cols.size()].column;
        if (col >= h->strs.size() || !h->strs[col]) fail(GBAGG_ERROR, "VARCHAR MIN/MAX: column not interned");
        gather_strings(h, *h->strs[col], h->out_data[c].as<uint32_t>(), h->out_valid[c].as<uint8_t>(), n,
                       h->res_str_off[c], h->res_str_bytes[c]);
        if (scalar_empty) h->res_str_off[c].assign(rows + 1, 0);
    }
    // TopN that could not run on the device: sort the full result here.
    if (h->topn_col >= 0 && h->topn_limit > 0 && !h->topn_device && rows > 0) {
        std::vector<uint64_t> idx(rows);
        for (uint64_t i = 0; i < rows; i++) idx[i] = i;
        auto cell_cmp = [&](size_t c, uint64_t a, uint64_t b) -> int {
            bool va = h->res_valid[c][a], vb = h->res_valid[c][b];
            if (va != vb) return va ? -1 : 1;  // NULLs last
            if (!va) return 0;
            const OutCol &oc = h->outs[c];
            if (oc.type == GBAGG_TYPE_VARCHAR) {
                const auto &off = h->res_str_off[c];
                const char *B = h->res_str_bytes[c].data();
                std::string_view x(B + off[a], off[a + 1] - off[a]), y(B + off[b], off[b + 1] - off[b]);
                return x < y ? -1 : (x > y ? 1 : 0);
            }
            if (oc.type == GBAGG_TYPE_DOUBLE) {
                double x, y;
                std::memcpy(&x, &h->res_data[c][a * 8], 8);
                std::memcpy(&y, &h->res_data[c][b * 8], 8);
                return x < y ? -1 : (x > y ? 1 : 0);
            }
            int64_t x, y;
            if (oc.type == GBAGG_TYPE_INTEGER) {
                int32_t x32, y32;
                std::memcpy(&x32, &h->res_data[c][a * 4], 4);
                std::memcpy(&y32, &h->res_data[c][b * 4], 4);
                x = x32;
                y = y32;
            } else {
                std::memcpy(&x, &h->res_data[c][a * 8], 8);
                std::memcpy(&y, &h->res_data[c][b * 8], 8);
            }
            return x < y ? -1 : (x > y ? 1 : 0);
        };
        size_t oc = (size_t)h->topn_col;
        bool asc = h->topn_asc;
        auto less = [&](uint64_t a, uint64_t b) {
            int c = cell_cmp(oc, a, b);
            if (c) return asc ? c < 0 : c > 0;
            for (size_t g = 0; g < h->group_cols.size(); g++) {
                c = cell_cmp(g, a, b);
                if (c) return c < 0;
            }
            return false;
        };
        uint64_t k = std::min<uint64_t>(rows, h->topn_limit);
        std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), less);
        idx.resize(k);
        auto permute = [&](std::vector<uint8_t> &v, size_t w) {
            std::vector<uint8_t> out(k * w);
            for (uint64_t i = 0; i < k; i++) std::memcpy(&out[i * w], &v[idx[i] * w], w);
            v.swap(out);
        };
        for (size_t c = 0; c < nc; c++) {
            size_t w = (h->outs[c].type == GBAGG_TYPE_INTEGER || h->outs[c].type == GBAGG_TYPE_VARCHAR) ? 4 : 8;
            if (h->outs[c].type == GBAGG_TYPE_VARCHAR) {
                std::vector<uint32_t> no(k + 1, 0);
                std::vector<char> nb;
                for (uint64_t i = 0; i < k; i++) {
                    no[i] = (uint32_t)nb.size();
                    const auto &off = h->res_str_off[c];
                    nb.insert(nb.end(), h->res_str_bytes[c].begin() + off[idx[i]],
                              h->res_str_bytes[c].begin() + off[idx[i] + 1]);
                }
                no[k] = (uint32_t)nb.size();
                h->res_str_off[c].swap(no);
                h->res_str_bytes[c].swap(nb);
            }
            permute(h->res_data[c], w);
            permute(h->res_valid[c], 1);
        }
        rows = k;
    } else if (h->row_limit > 0 && rows > h->row_limit && !(h->topn_col >= 0 && h->topn_limit > 0)) {
        rows = h->row_limit;  // bare LIMIT: any `row_limit` groups
    }
    if (h->tuple) {
        expand_tuples(h, rows);
        h->result_rows = rows;
        h->result_ready = true;
        return;
    }
    h->res_cols.assign(nc, {});
    for (size_t c = 0; c < nc; c++) {
        gbagg_result_column &rc = h->res_cols[c];
        rc.type = h->outs[c].type;
        rc.data = h->res_data[c].data();
        rc.validity = h->res_valid[c].data();
        if (rc.type == GBAGG_TYPE_VARCHAR) {
            rc.str_offsets = h->res_str_off[c].data();
            rc.str_bytes = h->res_str_bytes[c].data();
        }
    }
    h->result_rows = rows;
    h->result_ready = true;
}

}  // namespace

extern "C" {

int32_t gbagg_abi_version(void) { return GBAGG_ABI_VERSION; }

const char *gbagg_last_error(void) { return g_na_error ? g_na_error : g_last_error.c_str(); }

gbagg_status gbagg_device_name(char *buf, uint32_t buflen) {
    return guarded([&] {
        if (!buf || !buflen) fail(GBAGG_INVALID, "null buffer");
        std::string n = device_info().name;
        std::snprintf(buf, buflen, "%s", n.c_str());
    });
}

gbagg_status gbagg_plan_create(const gbagg_query_desc *query, gbagg_handle **out) {
    return guarded([&] {
        if (!out) fail(GBAGG_INVALID, "null out");
        *out = nullptr;
        auto h = std::make_unique<gbagg_handle>();
        build_plan(h.get(), query);
        pick_kernels(h.get());
        *out = h.release();
    });
}

gbagg_status gbagg_consume(gbagg_handle *h, const gbagg_column_view *columns, uint64_t num_rows) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        if (!h->phase1) add_batch(h, columns, num_rows, false);
        else {
            add_batch(h->phase1.get(), columns, num_rows, false);
            for (auto &p : h->phase1_more) add_batch(p.get(), columns, num_rows, false);
        }
    });
}

gbagg_status gbagg_consume_device(gbagg_handle *h, const gbagg_column_view *columns,
                                  uint64_t num_rows) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        if (!h->phase1) add_batch(h, columns, num_rows, true);
        else {
            add_batch(h->phase1.get(), columns, num_rows, true);
            for (auto &p : h->phase1_more) add_batch(p.get(), columns, num_rows, true);
        }
    });
}

gbagg_status gbagg_finalize(gbagg_handle *h) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        finalize(h);
    });
}

gbagg_status gbagg_result(gbagg_handle *h, gbagg_result_view *out) {
    return guarded([&] {
        if (!h || !out) fail(GBAGG_INVALID, "null argument");
        download_result(h);
        out->num_rows = h->result_rows;
        out->num_columns = (uint32_t)h->res_cols.size();
        out->columns = h->res_cols.data();
    });
}

void reset_handle(gbagg_handle *h) {
    if (h->stream) HIPCHK(hipStreamSynchronize(h->stream));
    h->batches.clear();
    h->barena_i = 0;  // batch copies: the stream has drained
    h->barena_used = 0;
    h->rows_total = 0;
    h->finalized = false;
    h->result_ready = false;
    h->result_rows = 0;
    for (auto &kd : h->kdict) kd = KeyDict();
    std::fill(h->str_mode.begin(), h->str_mode.end(), 0);
    if (h->phase1) reset_handle(h->phase1.get());
    for (auto &p : h->phase1_more) reset_handle(p.get());
}

gbagg_status gbagg_reset(gbagg_handle *h) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        reset_handle(h);
    });
}

gbagg_status gbagg_plan_info_get(gbagg_handle *h, gbagg_plan_info *out) {
    return guarded([&] {
        if (!h || !out) fail(GBAGG_INVALID, "null argument");
        std::memset(out, 0, sizeof(*out));
        out->strategy = h->used_strategy;
        if (h->hplan) {
            out->estimated_groups = h->hplan->est_groups;
            out->table_capacity = h->hplan->cap;
            out->groups = h->hplan->groups;
            out->heavy_keys = h->hplan->heavy_count;
            out->cold_rows = h->hplan->cold_rows;
            out->sample_rows = h->hplan->sample_rows;
            out->sample_distinct = h->hplan->sample_distinct;
            out->sample_f1 = h->hplan->sample_f1;
        }
        out->lds_slots = h->lds_slots;
        out->radix_bits = h->radix_bits;
        out->workgroups = h->agg_grid;
        out->retries = h->retries;
        out->rows_consumed = h->rows_total;
        out->ms_aggregate = h->ms_agg;
        out->ms_total = h->ms_total;
        out->ms_sample = h->ms_sample;
        out->ms_count = h->ms_count;
        out->ms_partition = h->ms_partition;
        out->ms_emit = h->ms_emit;
        std::snprintf(out->kernel_shape, sizeof(out->kernel_shape), "%s", h->kern ? h->kern->name : "");
        out->selector_distinct = h->sel_distinct;
        out->selector_used = h->sel_used;
        out->selector_path = h->sel_path;
        out->ms_strings = h->ms_strings;
        out->strings_distinct = h->strings_distinct;
        out->sum_reruns = h->fx_reruns;
        out->sum_ieee = h->fx_ieee ? 1 : 0;
    });
}

gbagg_status gbagg_set_param(gbagg_handle *h, const char *name, int64_t value) {
    return guarded([&] {
        if (!h || !name) fail(GBAGG_INVALID, "null argument");
        std::string n(name);
        auto in = [&](int64_t lo, int64_t hi) {
            if (value < lo || value > hi) fail(GBAGG_INVALID, n + " out of range");
        };
        if (n == "heavy_max") { in(0, 4096); h->heavy_max = (uint32_t)value; }
        else if (n == "heavy_min") { in(1, 64); h->heavy_min = (uint32_t)value; }
        else if (n == "fallback_cap") { in(1024, 1LL << 30); h->radix_fallback_cap = next_pow2((uint64_t)value); }
        else if (n == "part_bits") { in(-1, 13); h->force_part_bits = (int32_t)value; }
        else if (n == "msplit_fanout") { in(0, 8192); h->msplit_fanout = (uint32_t)value; }
        else if (n == "radix_ahead_rows") { in(0, 1LL << 40); h->radix_ahead_rows = (uint64_t)value; }
        else if (n == "lds_kb") { in(8, 160); h->k4_lds_kb = (uint32_t)value; }
        else if (n == "features") { in(0, 1); h->collect_features = value != 0; }
        else if (n == "wide_seed") { in(0, 1LL << 40); h->wide_seed0 = (uint64_t)value; }
        else if (n == "str_hash_bits") { in(1, 64); h->str_key_mask = value == 64 ? ~0ull : (1ull << value) - 1; }
        else if (n == "str_codes") { in(0, 1LL << 30); h->str_codes0 = (uint64_t)value; }
        else if (n == "strategy") { in(0, 4); h->strategy = (gbagg_strategy)value; }
        else if (n == "static_shapes") { in(0, 1); h->allow_static = value != 0; pick_kernels(h); }
        else if (n == "selector") {  // 0: no selector; 1: the built-in trained one (default)
            in(0, 1);
            h->sel_builtin = value != 0;
            if (!value) h->sel_params.clear();
        }
        else fail(GBAGG_INVALID, "unknown parameter " + n);
    });
}

void gbagg_destroy(gbagg_handle *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

gbagg_status gbagg_hash_aggregate_i32(const int32_t *keys, const int32_t *values, uint64_t count,
                                      int32_t *out_keys, int64_t *out_sums, int64_t *out_counts,
                                      uint64_t capacity, uint64_t *out_groups) {
    gbagg_handle *h = nullptr;
    gbagg_status st = guarded([&] {
        if ((!keys || !values) && count) fail(GBAGG_INVALID, "null input");
        if (!out_groups) fail(GBAGG_INVALID, "null out_groups");
        gbagg_type types[2] = {GBAGG_TYPE_INTEGER, GBAGG_TYPE_INTEGER};
        uint8_t nullable[2] = {0, 0};
        uint32_t gcol = 0;
        gbagg_agg aggs[2] = {{GBAGG_AGG_SUM, 1}, {GBAGG_AGG_COUNT_STAR, -1}};
        gbagg_query_desc q{};
        q.num_columns = 2;
        q.column_types = types;
        q.column_nullable = nullable;
        q.num_group_cols = 1;
        q.group_cols = &gcol;
        q.num_aggs = 2;
        q.aggs = aggs;
        q.topn_col = -1;
        auto hp = std::make_unique<gbagg_handle>();
        build_plan(hp.get(), &q);
        pick_kernels(hp.get());
        h = hp.release();
        gbagg_column_view cols[2] = {};
        cols[0].type = GBAGG_TYPE_INTEGER;
        cols[0].data = keys;
        cols[1].type = GBAGG_TYPE_INTEGER;
        cols[1].data = values;
        add_batch(h, cols, count, false);
        finalize(h);
        download_result(h);
        uint64_t n = h->result_rows;
        *out_groups = n;
        if (n > capacity) fail(GBAGG_INVALID, "output capacity too small");
        std::memcpy(out_keys, h->res_data[0].data(), n * 4);
        std::memcpy(out_sums, h->res_data[1].data(), n * 8);
        std::memcpy(out_counts, h->res_data[2].data(), n * 8);
    });
    gbagg_destroy(h);
    return st;
}

gbagg_status gbagg_selector_set(gbagg_handle *h, const float *params, uint32_t num_params) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        if (!params) {  // no selector at all, not even the built-in one
            h->sel_params.clear();
            h->sel_builtin = false;
            return;
        }
        if (num_params != (uint32_t)SEL_NPARAM) fail(GBAGG_INVALID, "selector parameter count differs from GBAGG_SELECTOR_PARAMS");
        for (uint32_t i = 0; i < num_params; i++)
            if (!std::isfinite(params[i])) fail(GBAGG_INVALID, "selector parameters are not finite");
        for (int i = 0; i < SEL_IN; i++)
            if (!(params[SEL_OFF_SIG + i] > 0)) fail(GBAGG_INVALID, "selector feature scale must be positive");
        ensure_runtime(h);
        h->sel_params.assign(params, params + num_params);
        h->dsel.ensure((size_t)num_params * 4);
        HIPCHK(hipMemcpy(h->dsel.ptr, params, (size_t)num_params * 4, hipMemcpyHostToDevice));
    });
}

gbagg_status gbagg_plan_features(gbagg_handle *h, float *features, float *outputs) {
    return guarded([&] {
        if (!h) fail(GBAGG_INVALID, "null handle");
        if (!h->feat_valid) fail(GBAGG_INVALID, "no features: finalize an AUTO plan with a selector or \"features\" set");
        if (features) std::memcpy(features, h->feat, sizeof(h->feat));
        if (outputs) {
            if (h->sel_params.empty()) fail(GBAGG_INVALID, "no selector outputs without a selector");
            std::memcpy(outputs, h->pred, sizeof(h->pred));
        }
    });
}

gbagg_status gbagg_selector_train(const float *x, const float *y, uint32_t n, uint32_t epochs, float lr, float lambda,
                                  float *params, uint32_t num_params, float *epoch_loss) {
    return guarded([&] {
        if (!x || !y || !params || n == 0 || epochs == 0) fail(GBAGG_INVALID, "bad training arguments");
        if (num_params != (uint32_t)SEL_NPARAM) fail(GBAGG_INVALID, "selector parameter count differs");
        DevBuf dx, dy, dp, dm, dv, dl;
        dx.ensure((size_t)n * SEL_IN * 4);
        dy.ensure((size_t)n * 2 * 4);
        dp.ensure((size_t)SEL_NPARAM * 4);
        dm.ensure((size_t)SEL_NW * 4);
        dv.ensure((size_t)SEL_NW * 4);
        dl.ensure((size_t)epochs * 4);
        HIPCHK(hipMemcpy(dx.ptr, x, (size_t)n * SEL_IN * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dy.ptr, y, (size_t)n * 2 * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp.ptr, params, (size_t)SEL_NPARAM * 4, hipMemcpyHostToDevice));
        std::vector<float> zero(SEL_NW, 0.f);
        HIPCHK(hipMemcpy(dm.ptr, zero.data(), (size_t)SEL_NW * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dv.ptr, zero.data(), (size_t)SEL_NW * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_sel_train, dim3(1), dim3(SEL_TT), 0, 0, dp.as<float>(), dm.as<float>(), dv.as<float>(),
                           (const float *)dx.as<float>(), (const float *)dy.as<float>(), n, epochs, lr, lambda,
                           dl.as<float>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(params, dp.ptr, (size_t)SEL_NPARAM * 4, hipMemcpyDeviceToHost));
        if (epoch_loss) HIPCHK(hipMemcpy(epoch_loss, dl.ptr, (size_t)epochs * 4, hipMemcpyDeviceToHost));
    });
}

gbagg_status gbagg_selector_infer(const float *params, uint32_t num_params, const float *x, uint32_t n, float *out) {
    return guarded([&] {
        if (!params || !x || !out) fail(GBAGG_INVALID, "null argument");
        if (num_params != (uint32_t)SEL_NPARAM) fail(GBAGG_INVALID, "selector parameter count differs");
        if (n == 0) return;
        DevBuf dp, dx, dy;
        dp.ensure((size_t)SEL_NPARAM * 4);
        dx.ensure((size_t)n * SEL_IN * 4);
        dy.ensure((size_t)n * SEL_OUT * 4);
        HIPCHK(hipMemcpy(dp.ptr, params, (size_t)SEL_NPARAM * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dx.ptr, x, (size_t)n * SEL_IN * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_sel_infer, dim3(n), dim3(SEL_H), 0, 0, (const float *)dp.as<float>(),
                           (const float *)dx.as<float>(), dy.as<float>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dy.ptr, (size_t)n * SEL_OUT * 4, hipMemcpyDeviceToHost));
    });
}

gbagg_status gbagg_device_alloc(uint64_t bytes, void **out) {
    return guarded([&] {
        if (!out) fail(GBAGG_INVALID, "null out");
        HIPCHK(hipMalloc(out, std::max<uint64_t>(bytes, 1)));
    });
}

gbagg_status gbagg_device_free(void *ptr) {
    return guarded([&] { HIPCHK(hipFree(ptr)); });
}

gbagg_status gbagg_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
    return guarded([&] { HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

gbagg_status gbagg_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
    return guarded([&] { HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

gbagg_status gbagg_device_synchronize(void) {
    return guarded([&] { HIPCHK(hipDeviceSynchronize()); });
}

uint64_t gbagg_hash_u64(uint64_t x) {
    unsigned __int128 r = (unsigned __int128)x * HASH_C;
    return (uint64_t)r ^ (uint64_t)(r >> 64);
}

}  // extern "C"

// =============================================================================
// Batches (gbagg_batch_*): many queries, one captured graph of six kernels
// =============================================================================
#include "gbagg_batch.hpp"

struct gbagg_batch {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BatchTune tune{65536, 2048, 512, 2};
    uint32_t cap_pct = 150;  // table slots per estimated key (percent)
    uint32_t mode_rule = 2;  // BatchCtl::mode_rule (2: LDS-first also for keys that repeat within a tile)
    uint32_t radix = 1;      // BatchCtl::radix: queries not LDS-first take radix partitions
    uint32_t part_slots = 4096;  // k_b_part LDS table slots (fewer when the LDS holds fewer)
    uint32_t use_selector = 1;  // per query: the learned selector's path picks the mode
    // selector plans per input (key column, rows, query shape): (distinct, mode) or mode ~0u when not usable
    std::unordered_map<uint64_t, std::pair<uint64_t, uint32_t>> sel_cache;
    uint32_t sel_queries = 0;   // batched queries of the last run whose mode the selector chose
    // the batched queries of the last run
    std::vector<uint32_t> item_of;        // batched query j -> item index
    std::vector<BatchQuery> hq;           // uploaded description
    std::vector<BatchQuery> dq_back;      // device state read back after the run
    std::vector<uint2> items;
    std::vector<std::vector<gbagg_type>> out_types;  // per batched query: key, aggs
    uint32_t nacc_max = 1;
    DevBuf dq, ditems, dctl, keys, acc, out_key, out_acc, hll, bhist, pstart, bfill, recs;
    uint64_t pool_slots = 0;
    // graph of the current shape
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t graph_sig = 0;
    uint32_t captures = 0, replays = 0;
    // items on the single-query operator
    std::vector<gbagg_handle *> single;   // per item: handle or null
    // last run
    uint32_t n_items = 0;
    double ms_device = 0, ms_single = 0;
    // host results of batched queries, per item (filled on demand)
    std::vector<std::vector<std::vector<uint8_t>>> res_data;
    std::vector<std::vector<std::vector<uint8_t>>> res_valid;
    std::vector<std::vector<gbagg_result_column>> res_cols;
    std::vector<int> res_ready;
    std::vector<int32_t> batched_index;   // item -> batched query j, or -1

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_sig = 0;
    }
    void drop_single() {
        for (auto *h : single)
            if (h) gbagg_destroy(h);
        single.clear();
    }
    ~gbagg_batch() {
        drop_graph();
        drop_single();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// The query shape the batched kernels run; anything else goes to the
// single-query operator.
bool batchable(const gbagg_batch_item &it) {
    const gbagg_query_desc *q = it.query;
    if (!q || !it.columns || q->num_group_cols != 1 || q->num_predicates != 0 || q->topn_col >= 0 ||
        q->row_limit != 0 || q->num_aggs < 1 || q->num_aggs > (uint32_t)BQ_MAXACC || !q->column_nullable ||
        it.num_rows > 0x7FFFFFFFull)
        return false;
    auto intcol = [&](int32_t c) {
        if (c < 0 || (uint32_t)c >= q->num_columns) return false;
        const gbagg_type t = q->column_types[c];
        return (t == GBAGG_TYPE_INTEGER || t == GBAGG_TYPE_BIGINT) && q->column_nullable[c] == 0 &&
               it.columns[c].validity == nullptr && it.columns[c].type == t;
    };
    if (!intcol((int32_t)q->group_cols[0])) return false;
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        const gbagg_agg &g = q->aggs[a];
        if (g.kind == GBAGG_AGG_COUNT_STAR) continue;
        if (g.kind == GBAGG_AGG_AVG || g.kind == GBAGG_AGG_COUNT_DISTINCT || !intcol(g.column)) return false;
    }
    return true;
}

uint64_t batch_pool_bound(uint64_t rows, uint32_t cap_pct) {
    // the estimate kernel's capacity formula with the estimate at its cap (all rows distinct)
    uint64_t cap = (rows * cap_pct + 99) / 100 + 64;
    return (cap + BQ_CHUNK - 1) / BQ_CHUNK * BQ_CHUNK;
}

uint64_t mix_sig(uint64_t h, uint64_t v) { return (h ^ v) * 0x100000001B3ull + (h >> 29); }

// Builds the batched query list, buffers and (when the shape changed) the graph.
void batch_prepare(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count) {
    const DeviceInfo &di = device_info();
    if (!b->stream) {
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&b->ev0));
        HIPCHK(hipEventCreate(&b->ev1));
    }
    b->hq.clear();
    b->item_of.clear();
    b->out_types.clear();
    b->items.clear();
    b->batched_index.assign(count, -1);
    b->nacc_max = 1;
    uint64_t pool = 0, rec_words = 0, sig = 0xcbf29ce484222325ull;
    for (uint32_t i = 0; i < count; i++) {
        if (!batchable(items[i])) continue;
        const gbagg_query_desc *q = items[i].query;
        BatchQuery bq{};
        const uint32_t kc = q->group_cols[0];
        bq.key = items[i].columns[kc].data;
        bq.key64 = q->column_types[kc] == GBAGG_TYPE_BIGINT;
        bq.rows = items[i].num_rows;
        bq.nacc = q->num_aggs;
        std::vector<gbagg_type> types{q->column_types[kc]};
        bq.rec_words = 1;
        for (uint32_t a = 0; a < q->num_aggs; a++) {
            const gbagg_agg &g = q->aggs[a];
            int op = ACC_CNT_STAR;
            gbagg_type t = GBAGG_TYPE_BIGINT;
            if (g.kind == GBAGG_AGG_SUM) op = ACC_SUM_I;
            if (g.kind == GBAGG_AGG_MIN) op = ACC_MIN_I;
            if (g.kind == GBAGG_AGG_MAX) op = ACC_MAX_I;
            if (op != ACC_CNT_STAR) {
                bq.val[a] = items[i].columns[g.column].data;
                if (q->column_types[g.column] == GBAGG_TYPE_BIGINT) bq.val64 |= 1u << a;
                if (g.kind != GBAGG_AGG_SUM) t = q->column_types[g.column];
                bq.valmask |= 1u << a;
                bq.rec_words++;
            }
            bq.ops |= (uint32_t)op << (4 * a);
            types.push_back(t);
        }
        b->nacc_max = std::max<uint32_t>(b->nacc_max, bq.nacc);
        const uint32_t j = (uint32_t)b->hq.size();
        b->batched_index[i] = (int32_t)j;
        b->item_of.push_back(i);
        b->out_types.push_back(types);
        bq.item0 = (uint32_t)b->items.size();
        for (uint64_t t = 0; t * b->tune.tile_rows < bq.rows; t++) b->items.push_back(make_uint2(j, (uint32_t)t));
        bq.nitems = (uint32_t)b->items.size() - bq.item0;
        if (b->use_selector) {
            // the learned selector's plan of this input: its path picks the
            // mode (DIRECT / LDS_HASH: LDS-first tiles; GLOBAL_HASH / RADIX:
            // the global table), its count may enlarge the table
            uint64_t ck = mix_sig(mix_sig(mix_sig(0x5e1ec7, (uint64_t)(uintptr_t)bq.key), bq.rows), bq.ops);
            for (int a = 0; a < BQ_MAXACC; a++) ck = mix_sig(ck, (uint64_t)(uintptr_t)bq.val[a]);
            auto it = b->sel_cache.find(ck);
            if (it == b->sel_cache.end()) {
                gbagg_handle *sh = nullptr;
                std::pair<uint64_t, uint32_t> plan(0, ~0u);
                if (gbagg_plan_create(q, &sh) == GBAGG_OK &&
                    gbagg_consume_device(sh, items[i].columns, items[i].num_rows) == GBAGG_OK) {
                    double d = 0;
                    int path = -1;
                    if (selector_plan(sh, d, path) && path >= 0)
                        plan = {(uint64_t)std::min<double>(d, (double)bq.rows), path <= 1 ? (uint32_t)BM_LDS : (uint32_t)BM_GLOBAL};
                }
                if (sh) gbagg_destroy(sh);
                it = b->sel_cache.emplace(ck, plan).first;
            }
            if (it->second.second != ~0u) {
                bq.sel = 1;
                bq.sel_est = it->second.first;
                bq.sel_mode = it->second.second;
            }
        }
        pool += batch_pool_bound(bq.rows, b->cap_pct);
        // the wider record layout (one word per accumulator) bounds both
        rec_words += b->radix || b->mode_rule == 3 ? bq.rows * (1 + bq.nacc) : 0;
        sig = mix_sig(sig, (uint64_t)(uintptr_t)bq.key);
        sig = mix_sig(sig, bq.rows);
        sig = mix_sig(sig, bq.ops | (uint64_t)bq.val64 << 32);
        for (int a = 0; a < BQ_MAXACC; a++) sig = mix_sig(sig, (uint64_t)(uintptr_t)bq.val[a]);
        b->hq.push_back(bq);
    }
    const BatchTune &tn = b->tune;
    sig = mix_sig(sig, tn.tile_rows);
    sig = mix_sig(sig, tn.lds_slots);
    sig = mix_sig(sig, tn.block);
    sig = mix_sig(sig, tn.wgs_per_cu);
    sig = mix_sig(sig, b->cap_pct);
    sig = mix_sig(sig, b->mode_rule);
    sig = mix_sig(sig, b->radix);
    sig = mix_sig(sig, b->part_slots);
    sig = mix_sig(sig, b->use_selector);
    const uint32_t nq = (uint32_t)b->hq.size();
    if (nq == 0) {
        b->drop_graph();
        return;
    }
    if (sig == b->graph_sig && b->exec) return;  // same queries, inputs and shape: replay
    b->drop_graph();
    pool = std::max<uint64_t>(pool, BQ_CHUNK);
    b->pool_slots = pool;
    b->dq.ensure(nq * sizeof(BatchQuery));
    b->ditems.ensure(std::max<size_t>(1, b->items.size()) * sizeof(uint2));
    b->dctl.ensure(sizeof(BatchCtl));
    b->keys.ensure(pool * 8);
    b->acc.ensure(pool * 8 * BQ_MAXACC);
    b->out_key.ensure(pool * 8);
    b->out_acc.ensure(pool * 8 * BQ_MAXACC);
    b->hll.ensure(std::max<size_t>(1, b->items.size()) * BQ_HLL_M);
    b->bhist.ensure(std::max<size_t>(1, b->items.size()) * BQ_RBUCKETS * 4);
    b->pstart.ensure((size_t)nq * (BQ_RBUCKETS + 1) * 4);
    b->bfill.ensure(std::max<size_t>(1, b->items.size()) * BQ_RBUCKETS * 4);
    // records of every query that may take radix partitions (the modes are
    // picked on the device)
    b->recs.ensure(std::max<uint64_t>(1, rec_words) * 8);
    BatchCtl ctl{};
    ctl.pool_slots = pool;
    ctl.n_items = (uint32_t)b->items.size();
    ctl.lds_slots = tn.lds_slots;
    ctl.cap_pct = b->cap_pct;
    ctl.mode_rule = b->mode_rule;
    ctl.radix = b->radix;
    ctl.part_slots = (uint32_t)std::min<size_t>(
        b->part_slots, (di.lds_per_block - 4096) / (8 * (1 + (size_t)b->nacc_max)) / 64 * 64);
    ctl.rec_cap = rec_words;
    HIPCHK(hipMemcpyAsync(b->dq.ptr, b->hq.data(), nq * sizeof(BatchQuery), hipMemcpyHostToDevice, b->stream));
    if (!b->items.empty())
        HIPCHK(hipMemcpyAsync(b->ditems.ptr, b->items.data(), b->items.size() * sizeof(uint2), hipMemcpyHostToDevice,
                              b->stream));
    HIPCHK(hipMemcpyAsync(b->dctl.ptr, &ctl, sizeof(ctl), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));

    BatchPool bp{b->keys.as<uint64_t>(), b->acc.as<uint64_t>(), b->out_key.as<int64_t>(), b->out_acc.as<int64_t>(),
                 b->bhist.as<uint32_t>(), b->pstart.as<uint32_t>(), b->bfill.as<uint32_t>(), b->recs.as<uint64_t>()};
    // LDS table, then the partition cursors of a tile that writes records
    const size_t agg_lds = (size_t)tn.lds_slots * 8 * (1 + b->nacc_max) + BQ_RBUCKETS * 4;
    if (agg_lds > di.lds_per_block) fail(GBAGG_INVALID, "batch tune: LDS table larger than the device allows");
    allow_lds(k_b_agg, agg_lds);
    // the partition table: part_slots, or as many as the LDS holds next to the
    // kernel's 4 KiB of static LDS
    const uint32_t part_slots = (uint32_t)std::min<size_t>(
        b->part_slots, (di.lds_per_block - 4096) / (8 * (1 + (size_t)b->nacc_max)) / 64 * 64);
    const size_t part_lds = (size_t)part_slots * 8 * (1 + b->nacc_max);
    allow_lds(k_b_part, part_lds);
    const uint32_t part_per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(4, di.lds_per_cu / part_lds));
    const uint32_t chunks = (uint32_t)std::min<uint64_t>((pool + BQ_CHUNK - 1) / BQ_CHUNK, (uint64_t)di.cus * 16);
    BatchQuery *dqp = b->dq.as<BatchQuery>();
    BatchCtl *dctl = b->dctl.as<BatchCtl>();
    // capture: every node is a kernel (no memset nodes)
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const uint32_t hgrid = (uint32_t)std::min<size_t>(std::max<size_t>(1, b->items.size()), (size_t)di.cus * 8);
    hipLaunchKernelGGL(k_b_hll, dim3(hgrid), dim3(256), 0, b->stream, (const BatchQuery *)dqp, (const BatchCtl *)dctl,
                       (const uint2 *)b->ditems.ptr, b->hll.as<uint8_t>(), b->bhist.as<uint32_t>(), tn);
    hipLaunchKernelGGL(k_b_estimate, dim3(nq), dim3(256), 0, b->stream, dqp, (const BatchCtl *)dctl,
                       (const uint8_t *)b->hll.as<uint8_t>());
    hipLaunchKernelGGL(k_b_layout, dim3(1), dim3(1024), 0, b->stream, dqp, nq, dctl);
    hipLaunchKernelGGL(k_b_rscan, dim3(nq), dim3(1024), 0, b->stream, (const BatchQuery *)dqp, bp);
    hipLaunchKernelGGL(k_b_init, dim3(chunks), dim3(BQ_CHUNK), 0, b->stream, (const BatchQuery *)dqp, nq,
                       (const BatchCtl *)dctl, bp);
    hipLaunchKernelGGL(k_b_agg, dim3((uint32_t)di.cus * tn.wgs_per_cu), dim3(tn.block), agg_lds, b->stream, dqp, dctl,
                       (const uint2 *)b->ditems.ptr, bp, tn);
    hipLaunchKernelGGL(k_b_part, dim3((uint32_t)di.cus * part_per_cu), dim3(BQ_PART_BLOCK), part_lds, b->stream, dqp,
                       nq, dctl, bp);
    hipLaunchKernelGGL(k_b_compact, dim3(chunks), dim3(BQ_CHUNK), 0, b->stream, dqp, nq, (const BatchCtl *)dctl, bp);
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_sig = sig;
    b->captures++;
}

void batch_run(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count) {
    if (count && !items) fail(GBAGG_INVALID, "null items");
    for (uint32_t i = 0; i < count; i++)
        if (!items[i].query || (items[i].query->num_columns && !items[i].columns))
            fail(GBAGG_INVALID, "batch item without query or columns");
    b->drop_single();
    b->res_ready.assign(count, 0);
    b->res_data.assign(count, {});
    b->res_valid.assign(count, {});
    b->res_cols.assign(count, {});
    b->n_items = count;
    batch_prepare(b, items, count);
    b->single.assign(count, nullptr);
    b->ms_device = 0;
    const uint32_t nq = (uint32_t)b->hq.size();
    if (nq) {
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        b->replays++;
        b->dq_back.resize(nq);
        HIPCHK(hipMemcpyAsync(b->dq_back.data(), b->dq.ptr, nq * sizeof(BatchQuery), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->ms_device = ms;
    }
    // single-query operator: the queries outside the batched shape, and
    // batched queries whose table overflowed
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < count; i++) {
        const int32_t j = b->batched_index[i];
        if (j >= 0 && !b->dq_back[j].overflow) continue;
        gbagg_handle *h = nullptr;
        gbagg_status st = gbagg_plan_create(items[i].query, &h);
        if (st == GBAGG_OK) st = gbagg_consume_device(h, items[i].columns, items[i].num_rows);
        if (st == GBAGG_OK) st = gbagg_finalize(h);
        if (st != GBAGG_OK) {
            std::string msg = std::string("batch item ") + std::to_string(i) + ": " + gbagg_last_error();
            if (h) gbagg_destroy(h);
            fail(st, msg);
        }
        b->single[i] = h;
        if (j >= 0) b->batched_index[i] = -1;
    }
    b->ms_single = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void batch_result(gbagg_batch *b, uint32_t i, gbagg_result_view *out) {
    if (i >= b->n_items) fail(GBAGG_INVALID, "batch result index out of range");
    if (b->single[i]) {
        HIPCHK(hipSuccess);
        gbagg_status st = gbagg_result(b->single[i], out);
        if (st != GBAGG_OK) fail(st, gbagg_last_error());
        return;
    }
    const int32_t j = b->batched_index[i];
    if (j < 0) fail(GBAGG_ERROR, "batch item has no result");
    if (!b->res_ready[i]) {
        const BatchQuery &q = b->dq_back[j];
        const uint64_t n = q.groups;
        const auto &types = b->out_types[j];
        auto &data = b->res_data[i];
        auto &valid = b->res_valid[i];
        data.assign(types.size(), {});
        valid.assign(types.size(), std::vector<uint8_t>(n, 1));
        std::vector<int64_t> tmp(n);
        for (size_t c = 0; c < types.size(); c++) {
            const int64_t *src = c == 0 ? b->out_key.as<int64_t>() + q.tbl_off
                                        : b->out_acc.as<int64_t>() + (uint64_t)(c - 1) * b->pool_slots + q.tbl_off;
            if (n) HIPCHK(hipMemcpyAsync(tmp.data(), src, n * 8, hipMemcpyDeviceToHost, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            if (types[c] == GBAGG_TYPE_INTEGER) {
                data[c].resize(n * 4);
                int32_t *d = (int32_t *)data[c].data();
                for (uint64_t r = 0; r < n; r++) d[r] = (int32_t)tmp[r];
            } else {
                data[c].resize(n * 8);
                if (n) std::memcpy(data[c].data(), tmp.data(), n * 8);
            }
        }
        auto &cols = b->res_cols[i];
        cols.assign(types.size(), gbagg_result_column{});
        for (size_t c = 0; c < types.size(); c++) {
            cols[c].type = types[c];
            cols[c].data = data[c].data();
            cols[c].validity = valid[c].data();
        }
        b->res_ready[i] = 1;
    }
    out->num_rows = b->dq_back[j].groups;
    out->num_columns = (uint32_t)b->res_cols[i].size();
    out->columns = b->res_cols[i].data();
}

}  // namespace

extern "C" {

gbagg_status gbagg_batch_create(gbagg_batch **out) {
    return guarded([&] {
        if (!out) fail(GBAGG_INVALID, "null out");
        *out = new gbagg_batch();
    });
}

gbagg_status gbagg_batch_set_tune(gbagg_batch *b, const gbagg_batch_tune *t) {
    return guarded([&] {
        if (!b || !t) fail(GBAGG_INVALID, "null argument");
        const bool pow2 = t->lds_slots && !(t->lds_slots & (t->lds_slots - 1));
        if (t->tile_rows < 1024 || t->tile_rows > (1u << 24) || !pow2 || t->lds_slots < 256 || t->lds_slots > 8192 ||
            (t->block != 256 && t->block != 512 && t->block != 1024) || t->wgs_per_cu < 1 || t->wgs_per_cu > 8)
            fail(GBAGG_INVALID, "batch tune out of range");
        b->tune = BatchTune{t->tile_rows, t->lds_slots, t->block, t->wgs_per_cu};
        b->drop_graph();
    });
}

gbagg_status gbagg_batch_set_param(gbagg_batch *b, const char *name, int64_t value) {
    return guarded([&] {
        if (!b || !name) fail(GBAGG_INVALID, "null argument");
        const std::string n(name);
        if (n == "cap_pct") {
            if (value < 1 || value > 1000) fail(GBAGG_INVALID, "cap_pct out of range [1, 1000]");
            b->cap_pct = (uint32_t)value;
        } else if (n == "mode_rule") {
            if (value < 0 || value > 3) fail(GBAGG_INVALID, "mode_rule out of range [0, 3]");
            b->mode_rule = (uint32_t)value;
        } else if (n == "radix") {
            if (value < 0 || value > 1) fail(GBAGG_INVALID, "radix out of range [0, 1]");
            b->radix = (uint32_t)value;
        } else if (n == "part_slots") {
            if (value < 256 || value > 8192 || value % 64) fail(GBAGG_INVALID, "part_slots out of range [256, 8192] or not a multiple of 64");
            b->part_slots = (uint32_t)value;
        } else if (n == "selector") {
            if (value < 0 || value > 1) fail(GBAGG_INVALID, "selector out of range [0, 1]");
            b->use_selector = (uint32_t)value;
        } else {
            fail(GBAGG_INVALID, "unknown batch parameter: " + n);
        }
        b->drop_graph();
    });
}

gbagg_status gbagg_batch_autotune(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count) {
    return guarded([&] {
        if (!b) fail(GBAGG_INVALID, "null batch");
        const DeviceInfo &di = device_info();
        const BatchTune keep = b->tune;
        BatchTune best = keep;
        double best_ms = 1e30;
        uint32_t nacc_max = 1;
        for (uint32_t i = 0; i < count; i++)
            if (items && batchable(items[i])) nacc_max = std::max<uint32_t>(nacc_max, items[i].query->num_aggs);
        for (uint32_t tile : {32768u, 65536u, 131072u})
            for (uint32_t slots : {1024u, 2048u, 4096u})
                for (uint32_t block : {256u, 512u, 1024u})
                    for (uint32_t wpc : {1u, 2u, 4u}) {
                        if ((size_t)slots * 8 * (1 + nacc_max) + BQ_RBUCKETS * 4 > di.lds_per_block) continue;
                        if ((size_t)block * wpc > 2048) continue;  // 32 waves per CU
                        b->tune = BatchTune{tile, slots, block, wpc};
                        b->drop_graph();
                        batch_run(b, items, count);  // capture
                        double ms = 1e30;
                        for (int rep = 0; rep < 2; rep++) {
                            batch_run(b, items, count);  // replays
                            ms = std::min(ms, b->ms_device);
                        }
                        if (ms < best_ms) {
                            best_ms = ms;
                            best = b->tune;
                        }
                    }
        b->tune = best;
        b->drop_graph();
    });
}

gbagg_status gbagg_batch_run(gbagg_batch *b, const gbagg_batch_item *items, uint32_t count) {
    return guarded([&] {
        if (!b) fail(GBAGG_INVALID, "null batch");
        batch_run(b, items, count);
    });
}

gbagg_status gbagg_batch_result(gbagg_batch *b, uint32_t index, gbagg_result_view *out) {
    return guarded([&] {
        if (!b || !out) fail(GBAGG_INVALID, "null argument");
        batch_result(b, index, out);
    });
}

gbagg_status gbagg_batch_info_get(gbagg_batch *b, gbagg_batch_info *out) {
    return guarded([&] {
        if (!b || !out) fail(GBAGG_INVALID, "null argument");
        std::memset(out, 0, sizeof(*out));
        out->queries = b->n_items;
        for (uint32_t i = 0; i < b->n_items && i < b->single.size(); i++) {
            if (b->single[i]) out->single++;
            else if (b->batched_index[i] >= 0) out->batched++;
        }
        out->graph_captures = b->captures;
        out->graph_replays = b->replays;
        out->ms_device = b->ms_device;
        out->ms_single = b->ms_single;
        out->tune = gbagg_batch_tune{b->tune.tile_rows, b->tune.lds_slots, b->tune.block, b->tune.wgs_per_cu};
        for (const auto &q : b->hq) out->selector_queries += q.sel;
    });
}

void gbagg_batch_destroy(gbagg_batch *b) { delete b; }

}  // extern "C"

// =============================================================================
// Parquet row-group decode (gbagg_parquet_*): the fused path's column source
// =============================================================================
#include "gbagg_parquet.hpp"

namespace {

void parquet_stream(gbagg_parquet *p) {
    if (!p->stream) {
        (void)device_info();
        HIPCHK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    }
    if (!p->hbad) HIPCHK(hipHostMalloc((void **)&p->hbad, 4, hipHostMallocDefault));
}

// Clears the error bits before the first launch since the last wait.
void parquet_arm(gbagg_parquet *p) {
    if (p->armed) return;
    p->bad.ensure(4);
    HIPCHK(hipMemsetAsync(p->bad.ptr, 0, 4, p->stream));
    p->armed = true;
}

// Device addresses of column c's chunks of row groups [g0, g1): in place in
// the resident image, else copied (16 readable bytes past each) into `own`
// (a per-call buffer) or, when own is null, into arena memory kept until reset.
std::vector<const uint8_t *> parquet_chunks(gbagg_parquet *p, uint32_t g0, uint32_t g1, const uint32_t *cols,
                                            uint32_t num_cols, DevBuf *own) {
    const uint32_t ng = g1 - g0;
    std::vector<const uint8_t *> dbufs((size_t)num_cols * ng);
    std::vector<uint64_t> at((size_t)num_cols * ng, 0);
    uint64_t total = 0;
    for (uint32_t i = 0; i < num_cols; i++) {
        if (cols[i] >= p->cols.size()) fail(GBAGG_INVALID, "column out of range");
        for (uint32_t g = 0; g < ng; g++) {
            int64_t start, size;
            parquet_chunk_range(p, g0 + g, cols[i], start, size);
            at[(size_t)i * ng + g] = p->resident ? (uint64_t)start : total;
            total += ((uint64_t)size + 16 + 255) & ~255ull;
        }
    }
    const uint8_t *base = p->image.as<uint8_t>();
    if (!p->resident && num_cols) {
        uint8_t *dst;
        if (own) {
            own->ensure(total);
            dst = own->as<uint8_t>();
        } else {
            dst = p->alloc(total);
        }
        for (uint32_t i = 0; i < num_cols; i++)
            for (uint32_t g = 0; g < ng; g++) {
                int64_t start, size;
                parquet_chunk_range(p, g0 + g, cols[i], start, size);
                uint8_t *d = dst + at[(size_t)i * ng + g];
                HIPCHK(hipMemsetAsync(d + size, 0, 16, p->stream));
                if (size) HIPCHK(hipMemcpyAsync(d, p->file + start, (size_t)size, hipMemcpyHostToDevice, p->stream));
            }
        base = dst;
    }
    for (size_t k = 0; k < dbufs.size(); k++) dbufs[k] = base + at[k];
    return dbufs;
}

// Launches the queued row groups of gbagg_parquet_decode_async.
void parquet_flush(gbagg_parquet *p) {
    if (!p->pending || !p->pending_rows) return;
    std::unique_ptr<PqCall> call = std::move(p->pending);
    p->pending_rows = 0;
    parquet_launch(p, *call);
}

// Launches what is queued, waits for every launch, and reports their errors.
void parquet_wait(gbagg_parquet *p) {
    if (!p->stream) return;
    parquet_flush(p);
    if (p->armed) HIPCHK(hipMemcpyAsync(p->hbad, p->bad.ptr, 4, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (!p->armed) return;
    p->armed = false;
    const uint32_t bad = *p->hbad;
    if (bad & 1) fail(GBAGG_INVALID, "parquet: dictionary index outside the dictionary");
    if (bad & 2) fail(GBAGG_INVALID, "parquet: truncated BYTE_ARRAY value");
    if (bad & 4) fail(GBAGG_INVALID, "parquet: malformed dictionary index stream");
}

}  // namespace

extern "C" {

gbagg_status gbagg_parquet_open(const uint8_t *file, uint64_t size, gbagg_parquet **out) {
    return guarded([&] {
        if (!out) fail(GBAGG_INVALID, "null out");
        *out = nullptr;
        auto p = std::make_unique<gbagg_parquet>();
        parquet_open(p.get(), file, size);
        *out = p.release();
    });
}

gbagg_status gbagg_parquet_shape(gbagg_parquet *p, uint32_t *num_row_groups, uint32_t *num_columns,
                                 uint64_t *num_rows) {
    return guarded([&] {
        if (!p) fail(GBAGG_INVALID, "null handle");
        if (num_row_groups) *num_row_groups = (uint32_t)p->rgs.size();
        if (num_columns) *num_columns = (uint32_t)p->cols.size();
        if (num_rows) *num_rows = (uint64_t)p->num_rows;
    });
}

gbagg_status gbagg_parquet_column(gbagg_parquet *p, uint32_t col, gbagg_type *type, uint8_t *nullable,
                                  const char **name) {
    return guarded([&] {
        if (!p || col >= p->cols.size()) fail(GBAGG_INVALID, "column out of range");
        const PqSchemaCol &c = p->cols[col];
        if (type) *type = parquet_type(c.type);
        if (nullable) *nullable = c.repetition == 1;
        if (name) *name = c.name.c_str();
    });
}

gbagg_status gbagg_parquet_row_group_rows(gbagg_parquet *p, uint32_t rg, uint64_t *rows) {
    return guarded([&] {
        if (!p || rg >= p->rgs.size() || !rows) fail(GBAGG_INVALID, "row group out of range");
        *rows = (uint64_t)p->rgs[rg].num_rows;
    });
}

gbagg_status gbagg_parquet_decode_range(gbagg_parquet *p, uint32_t rg0, uint32_t rg1, const uint32_t *cols,
                                        uint32_t num_cols, gbagg_column_view *out, uint64_t *rows) {
    return guarded([&] {
        if (!p || rg0 >= rg1 || rg1 > p->rgs.size()) fail(GBAGG_INVALID, "row group range out of range");
        if (num_cols && (!cols || !out)) fail(GBAGG_INVALID, "null columns");
        parquet_stream(p);
        parquet_wait(p);  // queued row groups first, with their errors
        parquet_arm(p);
        const uint32_t ng = rg1 - rg0;
        const std::vector<const uint8_t *> dbufs = parquet_chunks(p, rg0, rg1, cols, num_cols, &p->chunks);
        PqCall call;
        call.dev_streams = (uint64_t)ng * num_cols >= 32;
        for (uint32_t i = 0; i < num_cols; i++) {
            p->outs.push_back(std::make_unique<gbagg_parquet::Out>());
            parquet_plan_column(p, rg0, rg1, cols[i], dbufs.data() + (size_t)i * ng, *p->outs.back(), out[i], call);
        }
        parquet_launch(p, call);
        parquet_wait(p);  // the chunk buffer is reused by the next call
        if (rows) {
            uint64_t n = 0;
            for (uint32_t g = rg0; g < rg1; g++) n += (uint64_t)p->rgs[g].num_rows;
            *rows = n;
        }
    });
}

gbagg_status gbagg_parquet_decode(gbagg_parquet *p, uint32_t rg, const uint32_t *cols, uint32_t num_cols,
                                  gbagg_column_view *out, uint64_t *rows) {
    return gbagg_parquet_decode_range(p, rg, rg + 1, cols, num_cols, out, rows);
}

gbagg_status gbagg_parquet_decode_async(gbagg_parquet *p, uint32_t rg, const uint32_t *cols, uint32_t num_cols,
                                        gbagg_column_view *out, uint64_t *rows) {
    return guarded([&] {
        if (!p || rg >= p->rgs.size()) fail(GBAGG_INVALID, "row group out of range");
        if (num_cols && (!cols || !out)) fail(GBAGG_INVALID, "null columns");
        parquet_stream(p);
        parquet_arm(p);
        const std::vector<const uint8_t *> dbufs = parquet_chunks(p, rg, rg + 1, cols, num_cols, nullptr);
        // planned on its own, then queued: a row group that fails to plan
        // leaves the queue as it was
        PqCall call;
        call.dev_streams = true;  // index run headers read on the device, many streams per launch
        std::vector<gbagg_column_view> views(num_cols);
        for (uint32_t i = 0; i < num_cols; i++) {
            p->outs.push_back(std::make_unique<gbagg_parquet::Out>());
            parquet_plan_column(p, rg, rg + 1, cols[i], dbufs.data() + i, *p->outs.back(), views[i], call);
        }
        if (!p->pending) p->pending = std::make_unique<PqCall>();
        call_append(*p->pending, std::move(call));
        const uint64_t n = (uint64_t)p->rgs[rg].num_rows;
        p->pending_rows += std::max<uint64_t>(n, 1);
        for (uint32_t i = 0; i < num_cols; i++) out[i] = views[i];
        if (rows) *rows = n;
        if (p->pending_rows >= p->flush_rows) parquet_flush(p);
    });
}

gbagg_status gbagg_parquet_sync(gbagg_parquet *p) {
    return guarded([&] {
        if (!p) fail(GBAGG_INVALID, "null handle");
        parquet_wait(p);
    });
}

gbagg_status gbagg_parquet_set_flush_rows(gbagg_parquet *p, uint64_t rows) {
    return guarded([&] {
        if (!p) fail(GBAGG_INVALID, "null handle");
        if (!rows) fail(GBAGG_INVALID, "flush_rows must be at least 1");
        p->flush_rows = rows;
    });
}

gbagg_status gbagg_parquet_reset(gbagg_parquet *p) {
    return guarded([&] {
        if (!p) fail(GBAGG_INVALID, "null handle");
        // queued row groups are dropped; launched ones finish first
        p->pending.reset();
        p->pending_rows = 0;
        if (p->stream) HIPCHK(hipStreamSynchronize(p->stream));
        p->armed = false;
        p->outs.clear();
        p->arena = 0;
        p->arena_used = 0;
    });
}

gbagg_status gbagg_parquet_upload(gbagg_parquet *p) {
    return guarded([&] {
        if (!p) fail(GBAGG_INVALID, "null handle");
        parquet_stream(p);
        parquet_wait(p);  // queued row groups read the chunks from where they are now
        p->image.ensure(p->size + 16);
        HIPCHK(hipMemsetAsync(p->image.as<uint8_t>() + p->size, 0, 16, p->stream));
        HIPCHK(hipMemcpyAsync(p->image.ptr, p->file, p->size, hipMemcpyHostToDevice, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        p->resident = true;
    });
}

void gbagg_parquet_close(gbagg_parquet *p) {
    if (!p) return;
    delete p;  // waits for launched work; queued row groups are dropped
}

}  // extern "C"

#include "gbagg_noalloc.hpp"
