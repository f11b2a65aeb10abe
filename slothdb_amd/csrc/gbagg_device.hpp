// gbagg_device.hpp -- device-side building blocks of the GROUP BY operator:
// the parameter blocks the host passes to kernels, the key hash, row loading,
// predicate evaluation, group-key construction and the accumulator updates
// shared by every aggregation strategy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbagg {

constexpr int MAXU = 8;          // distinct input columns one plan may reference
constexpr int MAX_ACC = 16;      // 64-bit accumulators per group
constexpr int MAX_PRED = 8;
constexpr uint64_t SENTINEL = 0x8000000000000000ULL;  // empty key slot
constexpr uint64_t HASH_C = 0x9E3779B97F4A7C15ULL;
constexpr int HLL_P = 12;
constexpr int HLL_M = 1 << HLL_P;

// Accumulator kinds. Every accumulator is one 64-bit word per group.
//
// A double SUM is exact and independent of the order of the additions: each
// finite input v becomes the integer X = round(v * 2^s) (|X| < 2^63; s per
// accumulator, chosen by the host from the largest magnitude of the column),
// split as X = hi * 2^31 + lo with 0 <= lo < 2^31. ACC_SUM_FH sums the hi
// parts as plain int64 words; ACC_SUM_FL sums the lo parts in bits 0..60 (at
// most 2^30 rows per aggregate keep that sum below 2^61) and ORs flags into
// bits 61 (an input was rounded to the 2^-s grid), 62 (+Inf seen) and 63
// (-Inf seen; NaN sets both). The group's sum is the 128-bit integer
// hi * 2^31 + lo, scaled by 2^-s and rounded once to double at emit; with
// specials it is the IEEE sum of them. Inputs of magnitude >= 2^(52 - s) are
// exact; smaller ones are rounded, so a group whose rounded sum is not large
// against its row count (fx_hazard) sets OVF_FX_PREC, and the host re-runs
// the query with IEEE double atomics (ACC_SUM_F). A value beyond the host's
// bound (|v| >= 2^(63 - s)) sets OVF_FX, and the host re-runs with the bound
// of the column's exact maximum.
enum AccOp : int32_t {
    ACC_CNT_STAR = 0,   // += 1 per row
    ACC_CNT_VALID = 1,  // += 1 per non-NULL input
    ACC_SUM_I = 2,      // int64 two's-complement sum
    ACC_SUM_FH = 3,     // double sum, fixed-point high words (int64 sum)
    ACC_MIN_I = 4,      // int64 min
    ACC_MAX_I = 5,      // int64 max
    ACC_MIN_F = 6,      // min of the order-preserving uint64 image of a double
    ACC_MAX_F = 7,      // max of the same
    ACC_SUM_FL = 8,     // double sum, fixed-point low words + flags
    ACC_SUM_F = 9,      // IEEE double sum with atomics (order-dependent; only after OVF_FX_PREC)
};

constexpr int FX_LO_BITS = 31;
constexpr uint64_t FX_INEXACT = 1ULL << 61;
constexpr uint64_t FX_POS_INF = 1ULL << 62;
constexpr uint64_t FX_NEG_INF = 1ULL << 63;
constexpr uint64_t FX_SPECIAL = FX_POS_INF | FX_NEG_INF;  // both: NaN
constexpr uint64_t FX_FLAGS = FX_INEXACT | FX_SPECIAL;
constexpr uint64_t FX_LOW = ~FX_FLAGS;
constexpr uint32_t OVF_FX = 32u;       // DevPlan::overflow bit: a double SUM input beyond the fixed-point bound
constexpr uint32_t OVF_FX_PREC = 64u;  // DevPlan::overflow bit: a group's fixed-point sum may be imprecise
constexpr int FX_SHIFT_MAX = 1084;     // largest scale: every subnormal stays within the bound

enum ColKind : int32_t { CK_I32 = 0, CK_I64 = 1, CK_F64 = 2, CK_U32 = 3 };
enum PredKind : int32_t { PK_INT = 0, PK_F64 = 1, PK_IN = 2, PK_DICT = 3 };
enum KeyMode : int32_t {
    KM_NONE = 0,    // no GROUP BY: one group, key word 0
    KM_SINGLE = 1,  // one INTEGER / BIGINT / VARCHAR-code column, key = int64 image
    KM_PACK32 = 2,  // two 32-bit columns: key = (u32)c0 << 32 | (u32)c1
    KM_WIDE = 3,    // two columns, one of them BIGINT: key = c0 ^ wide_mix(c1, seed)
};

struct DevPred {
    int32_t u;     // loaded-column slot
    int32_t kind;  // PredKind
    int32_t op;    // gbagg_cmp_op for PK_INT / PK_F64
    int32_t pad;
    int64_t ival;
    double dval;
};

// Per-query constants.
struct DevQuery {
    int32_t nu;
    int32_t nkp;               // leading slots holding keys and predicate columns
    int32_t ukind[MAXU];
    int32_t key_mode;
    int32_t key_u[2];
    int32_t key_signed32[2];  // KM_PACK32: component is a signed INTEGER (order only)
    int32_t drop_null_keys;
    int32_t npred;
    DevPred pred[MAX_PRED];
    int32_t nacc;
    int32_t acc_op[MAX_ACC];
    int32_t acc_u[MAX_ACC];  // -1 for COUNT(*)
    int32_t wide_comp;       // KM_WIDE: the component kept in hidden accumulators (never VARCHAR)
    int32_t pad;
    uint64_t wide_seed;      // KM_WIDE: seed of wide_mix
    int32_t acc_fx[MAX_ACC]; // ACC_SUM_FH / ACC_SUM_FL: the scale s of X = round(v * 2^s)
    uint32_t *fx_flags;      // DevPlan::overflow (OVF_FX)
};

// Per-batch device pointers.
struct DevBatch {
    const void *data[MAXU];
    const uint8_t *valid[MAXU];
    const uint32_t *remap[2];     // VARCHAR group keys: code -> canonical id (null = identity)
    uint32_t key_dict_size[2];    // VARCHAR group keys: codes >= size drop the row
    const void *pred_ptr[MAX_PRED];   // PK_IN: int64 set; PK_DICT: uint8 match per code
    uint32_t pred_count[MAX_PRED];    // PK_IN: set size; PK_DICT: dictionary size
    uint64_t nrows;
};

// Device-resident plan values and run statistics.
struct DevPlan {
    uint64_t est_groups;   // HyperLogLog estimate
    uint64_t cap;          // global table slots in use (power of two)
    uint64_t groups;       // result rows written so far
    uint32_t overflow;     // bit 0: table probe bound hit; bit 1: candidate buffer full
    uint32_t cand_count;   // TopN candidates pushed
    uint64_t cap_max;      // allocation limit of the global table
    uint64_t sample_rows;      // rows sampled for heavy hitters (after the filter)
    uint64_t sample_distinct;  // distinct keys in the sample
    uint64_t sample_f1;        // keys seen exactly once in the sample
    uint64_t heavy_count;      // keys in the heavy-hitter table
    uint64_t heavy_sample_rows;// sample rows whose key is heavy
    uint64_t cold_rows;        // rows scattered to partitions
    uint32_t heavy_threshold;  // minimum sample count of a heavy key
    uint32_t radix_gate;       // the sample is not of low cardinality (the default plan's radix test)
    uint64_t topn_lb;          // TopN: no group with a smaller order value can be in the top K
    int64_t key_min;           // direct-mapped strategy: exact key word range
    int64_t key_max;
    int64_t rng_min;           // the same, from a pass launched before the plan is read back
    int64_t rng_max;
    uint32_t fx_exp[MAXU];     // per loaded slot: largest biased exponent of its finite double
                               // values (sampled rows, or all rows for the exact bound)
    // sample frequency statistics for the learned selector's features
    uint64_t sample_f2, sample_f3;  // keys seen exactly twice / three times
    uint64_t sample_max;            // largest count of one key
    uint64_t sample_ent;            // sum over keys of c * log2(c), fixed point with 2^-20 units
    uint64_t sample_hist[16];       // keys with count in [2^j, 2^(j+1))
    float feat[32];                 // the selector's features of the sample
    float pred[8];                  // its outputs (5 used)
};

// The global table: (cap + 2) records of (1 + nacc) words: key, accumulators.
// Record cap is the group whose key equals SENTINEL, record cap + 1 the NULL
// key group; both are occupied iff their COUNT(*) accumulator is non-zero.
struct DevTable {
    uint64_t *rec;
    uint32_t stride;
    uint32_t max_probe;
    DevPlan *plan;
};

// Global-memory accesses through address space 1. Pointers read out of
// parameter blocks are generic, and a generic (flat) access is counted by
// both the vector-memory and the LDS wait counters, so every LDS operation
// would also wait for outstanding row loads; global_load/global_store are
// counted by the vector-memory counter alone.
#define GBAGG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T *p) {
    return *(const GBAGG_GLOBAL T *)p;
}
// Input columns are read once per pass: non-temporal loads keep them from
// displacing the partially written record lines the radix scatter relies on
// the caches to combine (C2 scatter: 1.67 -> 1.52 ms).
template <class T>
__device__ __forceinline__ T gld_stream(const T *p) {
    return __builtin_nontemporal_load((const GBAGG_GLOBAL T *)p);
}
template <class T>
__device__ __forceinline__ void gst(T *p, T v) {
    *(GBAGG_GLOBAL T *)p = v;
}
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));  // 16-byte global access

__device__ __forceinline__ uint64_t hash64(uint64_t x) {
    // wyhash mum/mix with the 64-bit golden ratio: lo ^ hi of x * C
    // (include/third_party/unordered_dense.h:158-191, 266-267).
    return (x * HASH_C) ^ __umul64hi(x, HASH_C);
}

// KM_WIDE keys (two columns, at least one of them 64-bit; 16 bytes in all).
// With c the component named by wide_comp and d the other one, the table key
// word is w = d ^ wide_mix(c, seed): for a fixed c it is one-to-one in d, so
// a group is fixed by (w, c), and d = w ^ wide_mix(c, seed). c travels in two
// hidden accumulators, MIN(c) and MAX(c): they differ only if two groups
// share a key word, which the emit detects (the operator then reruns with
// another seed). wide_mix is the SplitMix64 output function (Steele, Lea,
// Flood 2014) of c + seed * golden ratio.
__host__ __device__ __forceinline__ uint64_t wide_mix(uint64_t c1, uint64_t seed) {
    uint64_t z = c1 + seed * HASH_C;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// X = round-half-even(v * 2^s) for a finite double v given by its bits; ok is
// false when |v| >= 2^(63 - s), where X would not fit an int64; inexact tells
// whether X differs from v * 2^s.
__host__ __device__ __forceinline__ int64_t fx_scale(uint64_t bits, int s, bool &ok, bool &inexact) {
    const int e = (int)((bits >> 52) & 0x7FF);
    uint64_t m = bits & 0xFFFFFFFFFFFFFULL;
    int k;  // |v| * 2^s = m * 2^k
    if (e) {
        m |= 1ULL << 52;
        k = e - 1075 + s;
    } else {
        k = -1074 + s;
    }
    ok = (e ? e : 1) - 1022 + s <= 63;  // |v| < 2^(e - 1022)
    inexact = false;
    if (!ok) return 0;
    uint64_t x;
    if (k >= 0) {
        x = m << k;  // k <= 10 here
    } else if (k <= -55) {
        x = 0;  // below half of the unit
        inexact = m != 0;
    } else {
        const int r = -k;
        x = m >> r;
        const uint64_t rem = m & ((1ULL << r) - 1), half = 1ULL << (r - 1);
        inexact = rem != 0;
        if (rem > half || (rem == half && (x & 1))) x++;
    }
    return (bits >> 63) ? -(int64_t)x : (int64_t)x;
}

// The word a double input adds to ACC_SUM_FH (hi) or ACC_SUM_FL (lo).
__device__ __forceinline__ uint64_t fx_part(uint64_t bits, int s, bool hi, uint32_t *flags) {
    if (((bits >> 52) & 0x7FF) == 0x7FF) {  // Inf or NaN: flags in the low word only
        if (hi) return 0;
        if (bits & 0xFFFFFFFFFFFFFULL) return FX_SPECIAL;
        return (bits >> 63) ? FX_NEG_INF : FX_POS_INF;
    }
    bool ok, inexact;
    const int64_t x = fx_scale(bits, s, ok, inexact);
    if (!ok) {
        if (hi) __hip_atomic_fetch_or(flags, OVF_FX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (hi) return (uint64_t)(x >> FX_LO_BITS);
    return ((uint64_t)x & ((1ULL << FX_LO_BITS) - 1)) | (inexact ? FX_INEXACT : 0);
}

// Two ACC_SUM_FL words combined (partial states of one group).
__host__ __device__ __forceinline__ uint64_t fx_lo_combine(uint64_t x, uint64_t y) {
    return ((x & FX_LOW) + (y & FX_LOW)) | ((x | y) & FX_FLAGS);
}

// |hi * 2^31 + lo| as a 128-bit magnitude (th:tl); returns the sign.
__host__ __device__ __forceinline__ bool fx_magnitude(uint64_t hi, uint64_t lo, uint64_t &th, uint64_t &tl) {
    lo &= FX_LOW;
    tl = (hi << FX_LO_BITS) + lo;
    th = (uint64_t)((int64_t)hi >> (64 - FX_LO_BITS)) + (tl < lo ? 1 : 0);
    const bool neg = (int64_t)th < 0;
    if (neg) {
        tl = ~tl + 1;
        th = ~th + (tl == 0 ? 1 : 0);
    }
    return neg;
}

// The group's double sum: (hi * 2^31 + lo) * 2^-s rounded once to double, or
// the IEEE sum of the specials it saw.
__host__ __device__ __forceinline__ double fx_value(uint64_t hi, uint64_t lo, int s) {
    const uint64_t f = lo & FX_SPECIAL;
    if (f) {
        if (f == FX_SPECIAL) return __builtin_bit_cast(double, 0x7FF8000000000000ULL);
        return __builtin_bit_cast(double, f == FX_POS_INF ? 0x7FF0000000000000ULL : 0xFFF0000000000000ULL);
    }
    uint64_t th, tl;
    const bool neg = fx_magnitude(hi, lo, th, tl);
    double d;
    if (th == 0) {
        d = (double)tl;  // one rounding
    } else {
        // the top 64 bits, with every lower nonzero bit folded into bit 0 so
        // that the conversion rounds as the full value would
        const int lz = __builtin_clzll(th);
        const int sh = 64 - lz;
        const uint64_t top = lz ? (th << lz) | (tl >> sh) : th;
        const uint64_t sticky = (lz ? (tl << lz) : tl) != 0;
        d = ldexp((double)(top | sticky), sh);
    }
    d = ldexp(d, -s);
    return neg ? -d : d;
}

// A group whose inputs were rounded to the 2^-s grid (at most half a unit
// each, cnt of them) may be off by more than 2^-42 of its sum when the sum is
// below cnt * 2^41 units.
__host__ __device__ __forceinline__ bool fx_hazard(uint64_t hi, uint64_t lo, uint64_t cnt) {
    if (!(lo & FX_INEXACT) || (lo & FX_SPECIAL)) return false;
    uint64_t th, tl;
    fx_magnitude(hi, lo, th, tl);
    if (th >> 23) return false;  // |T| >= 2^87 > 2^30 * 2^41
    // compare T with cnt * 2^41 in 128 bits
    const uint64_t bh = cnt >> 23, bl = cnt << 41;
    return th < bh || (th == bh && tl < bl);
}

__host__ __device__ __forceinline__ uint64_t acc_identity(int op) {
    switch (op) {
    case ACC_MIN_I: return 0x7FFFFFFFFFFFFFFFULL;
    case ACC_MAX_I: return 0x8000000000000000ULL;
    case ACC_MIN_F: return ~0ULL;
    default: return 0ULL;
    }
}

// Loaded slots feeding a fixed-point double SUM.
__host__ __device__ __forceinline__ uint32_t fx_slots(const DevQuery &q) {
    uint32_t m = 0;
    for (int a = 0; a < q.nacc; a++)
        if (q.acc_op[a] == ACC_SUM_FH) m |= 1u << q.acc_u[a];
    return m;
}

// Biased exponent of a finite nonzero double (subnormals count as 1), 0 for
// zero and the specials: |v| < 2^(e - 1022).
__host__ __device__ __forceinline__ uint32_t fx_exp_of(uint64_t bits) {
    const uint32_t e = (uint32_t)(bits >> 52) & 0x7FF;
    if (e == 0x7FF) return 0;
    if (e == 0) return (bits << 1) ? 1u : 0u;
    return e;
}

// Order-preserving image of a double: x < y  <=>  ord(x) < ord(y) for
// non-NaN values; -0.0 maps to the image of +0.0.
__host__ __device__ __forceinline__ uint64_t f64_ord(uint64_t bits) {
    if (bits == 0x8000000000000000ULL) bits = 0;
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ULL);
}
__host__ __device__ __forceinline__ uint64_t f64_unord(uint64_t o) {
    return (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFULL) : ~o;
}
__host__ __device__ __forceinline__ bool f64_isnan(uint64_t bits) {
    return ((bits >> 52) & 0x7FF) == 0x7FF && (bits & 0xFFFFFFFFFFFFFULL);
}
// MIN/MAX over DOUBLE skip NaN unless every input of the group is NaN (then
// the result is NaN), whatever the NaN's sign or payload: a NaN's image sits
// next to the accumulator's identity, beyond every number (MIN: just below
// ~0, MAX: 1), so any number replaces it.
constexpr uint64_t ORD_NAN_MIN = ~0ULL - 1;
constexpr uint64_t ORD_NAN_MAX = 1;
__host__ __device__ __forceinline__ uint64_t f64_ord_mm(uint64_t bits, bool is_min) {
    if (f64_isnan(bits)) return is_min ? ORD_NAN_MIN : ORD_NAN_MAX;
    return f64_ord(bits);
}
// Inverse of f64_ord_mm (NaN comes back as the canonical quiet NaN).
__host__ __device__ __forceinline__ uint64_t f64_unord_mm(uint64_t o) {
    if (o == ORD_NAN_MIN || o == ORD_NAN_MAX) return 0x7FF8000000000000ULL;
    return f64_unord(o);
}

struct Row {
    uint64_t v[MAXU];
    uint32_t vm;  // bit u set: column slot u is valid in this row
};

// row.v[u] for a runtime (wave-uniform) u. The mask passes through an empty
// register constraint so the compiler cannot rebuild the select chain into a
// load through a computed address, which would move the whole Row to scratch.
__device__ __forceinline__ uint64_t sel(const Row &row, int u) {
    uint64_t x = 0;
#pragma unroll
    for (int i = 0; i < MAXU; i++) {
        uint64_t m = 0ULL - (uint64_t)(u == i);
        asm volatile("" : "+v"(m));
        x |= row.v[i] & m;
    }
    return x;
}

__device__ __forceinline__ void load_row(const DevQuery &q, const DevBatch &b, uint64_t r,
                                         uint32_t umask, Row &row) {
    row.vm = 0;
#pragma unroll
    for (int u = 0; u < MAXU; u++) {
        row.v[u] = 0;
        if (u < q.nu && ((umask >> u) & 1)) {
            uint64_t x;
            switch (q.ukind[u]) {
            case CK_I32: x = (uint64_t)(int64_t)gld_stream((const int32_t *)b.data[u] + r); break;
            case CK_U32: x = (uint64_t)gld_stream((const uint32_t *)b.data[u] + r); break;
            default: x = gld_stream((const uint64_t *)b.data[u] + r); break;
            }
            row.v[u] = x;
            uint32_t ok = b.valid[u] ? (gld_stream(b.valid[u] + r) != 0) : 1u;
            row.vm |= ok << u;
        }
    }
}

__device__ __forceinline__ bool cmp_i64(int64_t v, int64_t lit, int op) {
    switch (op) {
    case 0: return v < lit;
    case 1: return v <= lit;
    case 2: return v > lit;
    case 3: return v >= lit;
    case 4: return v == lit;
    case 5: return v != lit;
    default: return false;
    }
}

__device__ __forceinline__ bool cmp_f64(double v, double lit, int op) {
    switch (op) {
    case 0: return v < lit;
    case 1: return v <= lit;
    case 2: return v > lit;
    case 3: return v >= lit;
    case 4: return v == lit;
    case 5: return v != lit;
    default: return false;
    }
}

// EvalSimplePredicates semantics (physical_planner.cpp:396-496): a NULL in a
// predicate column drops the row; integer columns compare int64-exact.
__device__ __forceinline__ bool eval_preds(const DevQuery &q, const DevBatch &b, const Row &row) {
#pragma unroll
    for (int p = 0; p < MAX_PRED; p++) {  // unrolled: b.pred_ptr[p] stays a register
        if (p >= q.npred) break;
        const DevPred &P = q.pred[p];
        if (!((row.vm >> P.u) & 1)) return false;
        uint64_t x = sel(row, P.u);
        bool ok;
        switch (P.kind) {
        case PK_INT: ok = cmp_i64((int64_t)x, P.ival, P.op); break;
        case PK_F64: ok = cmp_f64(__longlong_as_double((long long)x), P.dval, P.op); break;
        case PK_IN: {
            const int64_t *s = (const int64_t *)b.pred_ptr[p];
            ok = false;
            for (uint32_t i = 0; i < b.pred_count[p]; i++) ok |= (s[i] == (int64_t)x);
            break;
        }
        default: {
            uint32_t c = (uint32_t)x;
            ok = c < b.pred_count[p] && ((const uint8_t *)b.pred_ptr[p])[c] != 0;
            break;
        }
        }
        if (!ok) return false;
    }
    return true;
}

enum KeyResult : int { KEY_DROP = 0, KEY_OK = 1, KEY_NULL = 2 };

__device__ __forceinline__ int make_key(const DevQuery &q, const DevBatch &b, const Row &row,
                                        uint64_t &w) {
    w = 0;
    if (q.key_mode == KM_NONE) return KEY_OK;
    uint64_t c0 = sel(row, q.key_u[0]);
    bool v0 = (row.vm >> q.key_u[0]) & 1;
    if (!v0) return q.drop_null_keys ? KEY_DROP : KEY_NULL;
    if (b.remap[0] || b.key_dict_size[0]) {
        uint32_t c = (uint32_t)c0;
        if (c >= b.key_dict_size[0]) return KEY_DROP;
        if (b.remap[0]) c0 = b.remap[0][c];
    }
    if (q.key_mode == KM_SINGLE) {
        w = c0;
        return KEY_OK;
    }
    uint64_t c1 = sel(row, q.key_u[1]);
    bool v1 = (row.vm >> q.key_u[1]) & 1;
    if (!v1) return KEY_DROP;  // packed keys are planned only when NULLs drop
    if (b.remap[1] || b.key_dict_size[1]) {
        uint32_t c = (uint32_t)c1;
        if (c >= b.key_dict_size[1]) return KEY_DROP;
        if (b.remap[1]) c1 = b.remap[1][c];
    }
    if (q.key_mode == KM_WIDE)
        w = q.wide_comp ? c0 ^ wide_mix(c1, q.wide_seed) : c1 ^ wide_mix(c0, q.wide_seed);
    else
        w = (c0 << 32) | (c1 & 0xFFFFFFFFULL);
    return KEY_OK;
}

// ---- accumulator updates ------------------------------------------------

template <int SCOPE>
__device__ __forceinline__ void acc_apply(uint64_t *p, int op, uint64_t x) {
    switch (op) {
    case ACC_CNT_STAR:
    case ACC_CNT_VALID:
    case ACC_SUM_I:
    case ACC_SUM_FH:
        __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, SCOPE);
        break;
    case ACC_SUM_FL:
        if (x & FX_LOW) __hip_atomic_fetch_add(p, x & FX_LOW, __ATOMIC_RELAXED, SCOPE);
        if (x & FX_FLAGS) __hip_atomic_fetch_or(p, x & FX_FLAGS, __ATOMIC_RELAXED, SCOPE);
        break;
    case ACC_SUM_F:
        if (SCOPE == __HIP_MEMORY_SCOPE_WORKGROUP)
            __hip_atomic_fetch_add((double *)p, __longlong_as_double((long long)x), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            unsafeAtomicAdd((double *)p, __longlong_as_double((long long)x));
        break;
    case ACC_MIN_I:
        __hip_atomic_fetch_min((int64_t *)p, (int64_t)x, __ATOMIC_RELAXED, SCOPE);
        break;
    case ACC_MAX_I:
        __hip_atomic_fetch_max((int64_t *)p, (int64_t)x, __ATOMIC_RELAXED, SCOPE);
        break;
    case ACC_MIN_F:
        __hip_atomic_fetch_min(p, x, __ATOMIC_RELAXED, SCOPE);
        break;
    default:
        __hip_atomic_fetch_max(p, x, __ATOMIC_RELAXED, SCOPE);
        break;
    }
}

// Value a row contributes to accumulator a, or false if it contributes nothing.
__device__ __forceinline__ bool acc_input(const DevQuery &q, int a, const Row &row, uint64_t &x) {
    int op = q.acc_op[a], u = q.acc_u[a];
    if (op == ACC_CNT_STAR) { x = 1; return true; }
    if (!((row.vm >> u) & 1)) return false;
    uint64_t v = sel(row, u);
    switch (op) {
    case ACC_CNT_VALID: x = 1; break;
    case ACC_MIN_F: x = f64_ord_mm(v, true); break;
    case ACC_MAX_F: x = f64_ord_mm(v, false); break;
    case ACC_SUM_FH: x = fx_part(v, q.acc_fx[a], true, q.fx_flags); break;
    case ACC_SUM_FL: x = fx_part(v, q.acc_fx[a], false, q.fx_flags); break;
    default: x = v; break;
    }
    return true;
}

// Two partial values of accumulator kind op combined (no atomics).
__host__ __device__ __forceinline__ uint64_t acc_combine(int op, uint64_t x, uint64_t y) {
    switch (op) {
    case ACC_MIN_I: return (int64_t)y < (int64_t)x ? y : x;
    case ACC_MAX_I: return (int64_t)y > (int64_t)x ? y : x;
    case ACC_MIN_F: return y < x ? y : x;
    case ACC_MAX_F: return y > x ? y : x;
    case ACC_SUM_FL: return fx_lo_combine(x, y);
    case ACC_SUM_F:
        return __builtin_bit_cast(uint64_t, __builtin_bit_cast(double, x) + __builtin_bit_cast(double, y));
    default: return x + y;  // counts, integer sums, fixed-point high words
    }
}

// Whether a partial accumulator value changes the merged result.
__device__ __forceinline__ bool acc_nontrivial(int op, uint64_t x) {
    return x != acc_identity(op);
}

// ------------------------------------------------------------------------
// Query shapes. The per-row code is written once against a shape: DynShape
// reads the plan at run time (any query); StaticShape fixes the column kinds,
// key layout, predicates and accumulator ops at compile time, so the row loop
// compiles to straight-line code for the query shapes the host registers.
// ------------------------------------------------------------------------
struct ShapeSig {
    int nu;          // loaded slots
    int nkp;         // leading slots holding keys and predicate columns
    uint32_t kinds;  // ColKind of slot u at bits 2u..2u+1
    uint32_t nulls;  // bit u: slot u may carry a validity array
    int km, ku0, ku1;
    int na;
    uint64_t ops;    // AccOp of accumulator a at bits 4a..4a+3
    uint64_t aus;    // input slot of accumulator a at bits 4a..4a+3 (15 = none)
    int np;
    uint32_t pred;   // predicate p: slot (3 bits), kind (2 bits), op (4 bits) at bit 9p
    bool operator==(const ShapeSig &o) const {
        return nu == o.nu && nkp == o.nkp && kinds == o.kinds && nulls == o.nulls && km == o.km &&
               ku0 == o.ku0 && ku1 == o.ku1 && na == o.na && ops == o.ops && aus == o.aus &&
               np == o.np && pred == o.pred;
    }
};

struct DynShape {
    static constexpr bool kStatic = false;
    // record width is a run-time value; 0 keeps the vector record path off
    __host__ __device__ static constexpr int rec_words() { return 1; }
    const DevQuery &q;
    __device__ explicit DynShape(const DevQuery &q_) : q(q_) {}
    __device__ int nu() const { return q.nu; }
    __device__ int nkp() const { return q.nkp; }
    __device__ int kind(int u) const { return q.ukind[u]; }
    __device__ bool nullable(int) const { return true; }
    __device__ int key_mode() const { return q.key_mode; }
    __device__ int key_u(int i) const { return q.key_u[i]; }
    __device__ int npred() const { return q.npred; }
    __device__ int pred_u(int p) const { return q.pred[p].u; }
    __device__ int pred_kind(int p) const { return q.pred[p].kind; }
    __device__ int pred_op(int p) const { return q.pred[p].op; }
    __device__ int nacc() const { return q.nacc; }
    __device__ int op(int a) const { return q.acc_op[a]; }
    __device__ int au(int a) const { return q.acc_u[a]; }
};

template <int NU, int NKP, uint32_t KINDS, uint32_t NULLS, int KM, int KU0, int KU1, int NA, uint64_t OPS,
          uint64_t AUS, int NP, uint32_t PRED>
struct StaticShape {
    static constexpr bool kStatic = true;
    const DevQuery &q;
    __device__ explicit StaticShape(const DevQuery &q_) : q(q_) {}
    __host__ __device__ static constexpr int nu() { return NU; }
    __host__ __device__ static constexpr int nkp() { return NKP; }
    __host__ __device__ static constexpr int kind(int u) { return (int)((KINDS >> (2 * u)) & 3u); }
    __host__ __device__ static constexpr bool nullable(int u) { return (NULLS >> u) & 1u; }
    __host__ __device__ static constexpr int key_mode() { return KM; }
    __host__ __device__ static constexpr int key_u(int i) { return i == 0 ? KU0 : KU1; }
    __host__ __device__ static constexpr int npred() { return NP; }
    __host__ __device__ static constexpr int pred_u(int p) { return (int)((PRED >> (9 * p)) & 7u); }
    __host__ __device__ static constexpr int pred_kind(int p) { return (int)((PRED >> (9 * p + 3)) & 3u); }
    __host__ __device__ static constexpr int pred_op(int p) { return (int)((PRED >> (9 * p + 5)) & 15u); }
    __host__ __device__ static constexpr int nacc() { return NA; }
    __host__ __device__ static constexpr int op(int a) { return (int)((OPS >> (4 * a)) & 15u); }
    __host__ __device__ static constexpr int au(int a) {
        return ((AUS >> (4 * a)) & 15u) == 15u ? -1 : (int)((AUS >> (4 * a)) & 15u);
    }
    // radix record word of slot u (key at word 0; inputs in first-use order)
    __host__ __device__ static constexpr int rec_pos(int u) {
        int w = 1;
        for (int a = 0; a < NA; a++) {
            int x = au(a);
            if (x < 0) continue;
            bool seen = false;
            for (int b = 0; b < a; b++) seen = seen || au(b) == x;
            if (seen) continue;
            if (x == u) return w;
            w++;
        }
        return 0;
    }
    __host__ __device__ static constexpr int rec_inputs() {
        int w = 0;
        for (int u = 0; u < 8; u++) w += rec_pos(u) ? 1 : 0;
        return w;
    }
    __host__ __device__ static constexpr int rec_words() {
        return 1 + rec_inputs() + (any_nullable_input() ? 1 : 0);
    }
    __host__ __device__ static constexpr bool any_nullable_input() {
        bool n = false;
        for (int a = 0; a < NA; a++) n = n || (au(a) >= 0 && nullable(au(a)));
        return n;
    }
    static constexpr ShapeSig sig() {
        return ShapeSig{NU, NKP, KINDS, NULLS, KM, KU0, KU1, NA, OPS, AUS, NP, PRED};
    }
};

// row.v[u]: a plain register for static shapes (u is a compile-time constant).
template <class S>
__device__ __forceinline__ uint64_t sel_s(const Row &row, int u) {
    if constexpr (S::kStatic) return row.v[u];
    else return sel(row, u);
}

// Loads the first `nload` slots of row r.
template <class S>
__device__ __forceinline__ void load_row_s(const S &sh, const DevBatch &b, uint64_t r, int nload, Row &row) {
    row.vm = 0;
#pragma unroll
    for (int u = 0; u < MAXU; u++) {
        row.v[u] = 0;
        if (u < nload) {
            uint64_t x;
            switch (sh.kind(u)) {
            case CK_I32: x = (uint64_t)(int64_t)gld_stream((const int32_t *)b.data[u] + r); break;
            case CK_U32: x = (uint64_t)gld_stream((const uint32_t *)b.data[u] + r); break;
            default: x = gld_stream((const uint64_t *)b.data[u] + r); break;
            }
            row.v[u] = x;
            uint32_t ok = 1u;
            if (sh.nullable(u) && b.valid[u]) ok = gld_stream(b.valid[u] + r) != 0;
            row.vm |= ok << u;
        }
    }
}

template <class S>
__device__ __forceinline__ bool eval_preds_s(const S &sh, const DevQuery &q, const DevBatch &b, const Row &row) {
#pragma unroll
    for (int p = 0; p < MAX_PRED; p++) {
        if (p >= sh.npred()) break;
        const int u = sh.pred_u(p);
        if (sh.nullable(u) && !((row.vm >> u) & 1)) return false;
        const uint64_t x = sel_s<S>(row, u);
        bool ok;
        switch (sh.pred_kind(p)) {
        case PK_INT: ok = cmp_i64((int64_t)x, q.pred[p].ival, sh.pred_op(p)); break;
        case PK_F64: ok = cmp_f64(__longlong_as_double((long long)x), q.pred[p].dval, sh.pred_op(p)); break;
        case PK_IN: {
            const int64_t *st = (const int64_t *)b.pred_ptr[p];
            ok = false;
            for (uint32_t i = 0; i < b.pred_count[p]; i++) ok |= (st[i] == (int64_t)x);
            break;
        }
        default: {
            uint32_t c = (uint32_t)x;
            ok = c < b.pred_count[p] && ((const uint8_t *)b.pred_ptr[p])[c] != 0;
            break;
        }
        }
        if (!ok) return false;
    }
    return true;
}

template <class S>
__device__ __forceinline__ int make_key_s(const S &sh, const DevQuery &q, const DevBatch &b, const Row &row,
                                          uint64_t &w) {
    w = 0;
    if (sh.key_mode() == KM_NONE) return KEY_OK;
    const int u0 = sh.key_u(0);
    uint64_t c0 = sel_s<S>(row, u0);
    if (sh.nullable(u0) && !((row.vm >> u0) & 1)) return q.drop_null_keys ? KEY_DROP : KEY_NULL;
    if (sh.kind(u0) == CK_U32 && (b.remap[0] || b.key_dict_size[0])) {
        uint32_t c = (uint32_t)c0;
        if (c >= b.key_dict_size[0]) return KEY_DROP;
        if (b.remap[0]) c0 = b.remap[0][c];
    }
    if (sh.key_mode() == KM_SINGLE) {
        w = c0;
        return KEY_OK;
    }
    const int u1 = sh.key_u(1);
    uint64_t c1 = sel_s<S>(row, u1);
    if (sh.nullable(u1) && !((row.vm >> u1) & 1)) return KEY_DROP;
    if (sh.kind(u1) == CK_U32 && (b.remap[1] || b.key_dict_size[1])) {
        uint32_t c = (uint32_t)c1;
        if (c >= b.key_dict_size[1]) return KEY_DROP;
        if (b.remap[1]) c1 = b.remap[1][c];
    }
    if (sh.key_mode() == KM_WIDE)
        w = q.wide_comp ? c0 ^ wide_mix(c1, q.wide_seed) : c1 ^ wide_mix(c0, q.wide_seed);
    else
        w = (c0 << 32) | (c1 & 0xFFFFFFFFULL);
    return KEY_OK;
}

template <class S>
__device__ __forceinline__ bool acc_input_s(const S &sh, int a, const Row &row, uint64_t &x) {
    const int op = sh.op(a), u = sh.au(a);
    if (op == ACC_CNT_STAR) { x = 1; return true; }
    if (sh.nullable(u) && !((row.vm >> u) & 1)) return false;
    const uint64_t v = sel_s<S>(row, u);
    switch (op) {
    case ACC_CNT_VALID: x = 1; break;
    case ACC_MIN_F: x = f64_ord_mm(v, true); break;
    case ACC_MAX_F: x = f64_ord_mm(v, false); break;
    case ACC_SUM_FH: x = fx_part(v, sh.q.acc_fx[a], true, sh.q.fx_flags); break;
    case ACC_SUM_FL: x = fx_part(v, sh.q.acc_fx[a], false, sh.q.fx_flags); break;
    default: x = v; break;
    }
    return true;
}

// acc a of one row into dst[a * stride] with atomics of the given scope.
template <int SCOPE, class S>
__device__ __forceinline__ void row_update_s(const S &sh, const Row &row, uint64_t *dst, uint64_t stride) {
#pragma unroll
    for (int a = 0; a < MAX_ACC; a++) {
        if (a >= sh.nacc()) break;
        uint64_t x;
        if (acc_input_s(sh, a, row, x)) acc_apply<SCOPE>(dst + (uint64_t)a * stride, sh.op(a), x);
    }
}

}  // namespace gbagg
