// gbagg_kernels.hpp -- the HIP kernels of the GROUP BY operator.
//
//   k_hll            HyperLogLog sketch of the (filtered) group keys, per block
//   k_hll_finish     merges the block sketches, estimates the group count and
//                    sizes the global table on the device (no host round trip)
//   k_table_init     global table: empty keys, identity accumulators
//   k_aggregate      fused filter + key hash + per-workgroup LDS hash table
//                    with atomic COUNT/SUM/MIN/MAX; rows whose key finds no LDS
//                    slot, and the LDS table at the end, merge into the global
//                    table with device-scope atomics
//   k_occ_count / k_scan_blocks / k_compact_write
//                    dense, columnar result in slot order
//   k_topn_local / k_topn_final / k_gather_write
//                    block-wise TopN straight off the table (SetTopNHint)
#pragma once

#include "gbagg_device.hpp"

namespace gbagg {

constexpr int AGG_BLOCK = 256;
constexpr int AGG_ROWS = 4;          // rows per thread per tile
constexpr int COMPACT_CHUNK = 4096;  // table slots per compaction block
constexpr uint32_t INVALID_SLOT = 0xFFFFFFFFu;

// ------------------------------------------------------------------------
// HyperLogLog (Flajolet et al. 2007), p = 12: register = top 12 hash bits,
// rank = leading zeros of the remaining bits + 1.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hll(DevQuery q, DevBatch b, uint32_t umask,
                                             uint8_t *partial) {
    __shared__ uint32_t reg[HLL_M];
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x) reg[i] = 0;
    __syncthreads();
    const uint64_t n = b.nrows;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (uint64_t)gridDim.x * blockDim.x) {
        Row row;
        load_row(q, b, r, umask, row);
        if (q.npred && !eval_preds(q, b, row)) continue;
        uint64_t w;
        int kr = make_key(q, b, row, w);
        if (kr == KEY_DROP) continue;
        uint64_t h = kr == KEY_NULL ? 0x5bd1e9955bd1e995ULL : hash64(w);
        uint32_t idx = (uint32_t)(h >> (64 - HLL_P));
        uint64_t rest = h << HLL_P;
        uint32_t rank = rest ? (uint32_t)__clzll((long long)rest) + 1 : (64 - HLL_P + 1);
        __hip_atomic_fetch_max(&reg[idx], rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    uint8_t *out = partial + (uint64_t)blockIdx.x * HLL_M;
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x) out[i] = (uint8_t)reg[i];
}

// One block of 1024 threads. growth: table slots per estimated group.
__global__ __launch_bounds__(1024) void k_hll_finish(const uint8_t *partial, uint32_t nparts,
                                                     DevPlan *plan, double growth,
                                                     uint64_t min_cap) {
    __shared__ double s_sum[1024];
    __shared__ uint32_t s_zero[1024];
    double sum = 0.0;
    uint32_t zeros = 0;
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x) {
        uint32_t m = 0;
        for (uint32_t p = 0; p < nparts; p++) m = max(m, (uint32_t)partial[(uint64_t)p * HLL_M + i]);
        sum += ldexp(1.0, -(int)m);
        zeros += (m == 0);
    }
    s_sum[threadIdx.x] = sum;
    s_zero[threadIdx.x] = zeros;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
            s_zero[threadIdx.x] += s_zero[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = (double)HLL_M;
        const double alpha = 0.7213 / (1.0 + 1.079 / m);
        double e = alpha * m * m / s_sum[0];
        if (e <= 2.5 * m && s_zero[0] > 0) e = m * log(m / (double)s_zero[0]);
        uint64_t est = (uint64_t)(e + 0.5);
        plan->est_groups = est;
        uint64_t want = plan->forced_cap ? plan->forced_cap : (uint64_t)((double)est * growth) + 1;
        uint64_t cap = min_cap;
        while (cap < want && cap < plan->cap_max) cap <<= 1;
        if (cap > plan->cap_max) cap = plan->cap_max;
        plan->cap = cap;
        plan->overflow = 0;
        plan->groups = 0;
        plan->topn_count = 0;
    }
}

// Host-sized variant (estimate skipped): cap comes from forced_cap.
__global__ void k_plan_fixed(DevPlan *plan) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint64_t cap = plan->forced_cap;
        if (cap > plan->cap_max) cap = plan->cap_max;
        plan->cap = cap;
        plan->overflow = 0;
        plan->groups = 0;
        plan->topn_count = 0;
    }
}

__global__ __launch_bounds__(256) void k_table_init(DevTable T, DevQuery q) {
    const uint64_t cap = T.plan->cap;
    const uint64_t nrec = cap + 2;
    uint64_t id[MAX_ACC];
#pragma unroll
    for (int a = 0; a < MAX_ACC; a++) id[a] = a < q.nacc ? acc_identity(q.acc_op[a]) : 0;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nrec;
         s += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t *r = T.rec + s * T.stride;
        r[0] = SENTINEL;
#pragma unroll
        for (int a = 0; a < MAX_ACC; a++)
            if (a < q.nacc) r[1 + a] = id[a];
    }
}

// ------------------------------------------------------------------------
// Global table: linear probing on the key word itself (claim = CAS from
// SENTINEL), so no separate occupancy word or publication step is needed.
// ------------------------------------------------------------------------
__device__ __forceinline__ int64_t g_find(const DevTable &T, uint64_t mask, uint64_t h,
                                          uint64_t k) {
    if (k == SENTINEL) return (int64_t)(mask + 1);
    uint64_t pos = h & mask;
    for (uint32_t i = 0; i < T.max_probe; i++) {
        uint64_t *kp = T.rec + pos * T.stride;
        uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == k) return (int64_t)pos;
        if (cur == SENTINEL) {
            uint64_t expected = SENTINEL;
            __hip_atomic_compare_exchange_strong(kp, &expected, k, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (expected == SENTINEL || expected == k) return (int64_t)pos;
        }
        pos = (pos + 1) & mask;
    }
    __hip_atomic_fetch_or(&T.plan->overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return -1;
}

__device__ __forceinline__ void g_update(const DevQuery &q, const DevTable &T, uint64_t g,
                                         const Row &row) {
    uint64_t *acc = T.rec + g * T.stride + 1;
    for (int a = 0; a < q.nacc; a++) {
        uint64_t x;
        if (acc_input(q, a, row, x)) acc_apply<__HIP_MEMORY_SCOPE_AGENT>(acc + a, q.acc_op[a], x);
    }
}

// Per-workgroup LDS table: keys[S] then acc[nacc][S] then a fill counter.
struct LdsTable {
    uint64_t *key;
    uint64_t *acc;
    uint32_t *fill;
    uint32_t S;
    uint32_t limit;
};

__device__ __forceinline__ int lds_find(const LdsTable &t, uint64_t h, uint64_t k) {
    const uint32_t mask = t.S - 1;
    uint32_t pos = (uint32_t)h & mask;
    for (int i = 0; i < 32; i++) {
        uint64_t cur = __hip_atomic_load(&t.key[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == k) return (int)pos;
        if (cur == SENTINEL) {
            if (__hip_atomic_load(t.fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= t.limit)
                return -1;
            uint64_t expected = SENTINEL;
            __hip_atomic_compare_exchange_strong(&t.key[pos], &expected, k, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (expected == SENTINEL) {
                __hip_atomic_fetch_add(t.fill, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                return (int)pos;
            }
            if (expected == k) return (int)pos;
        }
        pos = (pos + 1) & mask;
    }
    return -1;
}

// All 64 lanes carry the same key (scalar aggregates, clustered keys): reduce
// across the wave and let one lane apply. Returns true if handled.
__device__ __forceinline__ bool wave_uniform_apply(const DevQuery &q, const Row &row, uint64_t w,
                                                   uint64_t *dst, uint32_t dst_stride, bool lds) {
    uint64_t w0 = __shfl(w, 0);
    if (__ballot(w == w0) != ~0ULL) return false;
    const int lane = threadIdx.x & 63;
    for (int a = 0; a < q.nacc; a++) {
        int op = q.acc_op[a];
        uint64_t x;
        if (!acc_input(q, a, row, x)) x = acc_identity(op);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint64_t y = __shfl_xor(x, off);
            switch (op) {
            case ACC_SUM_F: x = (uint64_t)__double_as_longlong(
                                __longlong_as_double((long long)x) + __longlong_as_double((long long)y));
                break;
            case ACC_MIN_I: x = (int64_t)x < (int64_t)y ? x : y; break;
            case ACC_MAX_I: x = (int64_t)x > (int64_t)y ? x : y; break;
            case ACC_MIN_F: x = x < y ? x : y; break;
            case ACC_MAX_F: x = x > y ? x : y; break;
            default: x = x + y; break;
            }
        }
        if (lane == 0 && acc_nontrivial(op, x)) {
            if (lds) acc_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(dst + (uint64_t)a * dst_stride, op, x);
            else acc_apply<__HIP_MEMORY_SCOPE_AGENT>(dst + a, op, x);
        }
    }
    return true;
}

__global__ __launch_bounds__(AGG_BLOCK) void k_aggregate(DevQuery q, DevBatch b, DevTable T,
                                                         uint32_t S, uint32_t fill_limit,
                                                         uint32_t umask) {
    extern __shared__ uint64_t smem[];
    LdsTable t;
    t.S = S;
    t.limit = fill_limit;
    t.key = smem;
    t.acc = smem + S;
    t.fill = (uint32_t *)(t.acc + (uint64_t)S * q.nacc);
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) t.key[i] = SENTINEL;
    for (int a = 0; a < q.nacc; a++) {
        uint64_t id = acc_identity(q.acc_op[a]);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) t.acc[(uint64_t)a * S + i] = id;
    }
    if (threadIdx.x == 0) *t.fill = 0;
    __syncthreads();

    const uint64_t cap = T.plan->cap;
    const uint64_t mask = cap - 1;
    const uint64_t n = b.nrows;
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row(q, b, r, umask, rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            bool live = r < n;
            if (live && q.npred) live = eval_preds(q, b, rows[j]);
            uint64_t w = 0;
            int kr = KEY_DROP;
            if (live) kr = make_key(q, b, rows[j], w);
            if (kr == KEY_DROP) continue;
            if (kr == KEY_NULL) {
                g_update(q, T, cap + 1, rows[j]);
                continue;
            }
            if (w == SENTINEL) {
                g_update(q, T, cap, rows[j]);
                continue;
            }
            const uint64_t h = hash64(w);
            int slot = lds_find(t, h, w);
            if (slot >= 0) {
                for (int a = 0; a < q.nacc; a++) {
                    uint64_t x;
                    if (acc_input(q, a, rows[j], x))
                        acc_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(t.acc + (uint64_t)a * S + slot,
                                                                q.acc_op[a], x);
                }
            } else {
                int64_t g = g_find(T, mask, h, w);
                if (g >= 0) g_update(q, T, (uint64_t)g, rows[j]);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        uint64_t k = t.key[i];
        if (k == SENTINEL) continue;
        int64_t g = g_find(T, mask, hash64(k), k);
        if (g < 0) continue;
        uint64_t *acc = T.rec + (uint64_t)g * T.stride + 1;
        for (int a = 0; a < q.nacc; a++) {
            uint64_t x = t.acc[(uint64_t)a * S + i];
            if (acc_nontrivial(q.acc_op[a], x)) acc_apply<__HIP_MEMORY_SCOPE_AGENT>(acc + a, q.acc_op[a], x);
        }
    }
}

// Scalar aggregate / global-table-only variant (no LDS table).
__global__ __launch_bounds__(AGG_BLOCK) void k_aggregate_global(DevQuery q, DevBatch b, DevTable T,
                                                                uint32_t umask) {
    const uint64_t cap = T.plan->cap;
    const uint64_t mask = cap - 1;
    const uint64_t n = b.nrows;
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row(q, b, r, umask, rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            bool live = r < n;
            if (live && q.npred) live = eval_preds(q, b, rows[j]);
            uint64_t w = 0;
            int kr = KEY_DROP;
            if (live) kr = make_key(q, b, rows[j], w);
            // a full wave with one key (scalar aggregates) reduces first
            if (__ballot(kr == KEY_OK && w != SENTINEL) == ~0ULL) {
                const uint64_t w0 = __shfl(w, 0);
                if (__ballot(w == w0) == ~0ULL) {
                    int64_t g = -1;
                    if ((threadIdx.x & 63) == 0) g = g_find(T, mask, hash64(w0), w0);
                    g = __shfl(g, 0);
                    if (g >= 0) wave_uniform_apply(q, rows[j], w, T.rec + (uint64_t)g * T.stride + 1, 1, false);
                    continue;
                }
            }
            if (kr == KEY_DROP) continue;
            int64_t g;
            if (kr == KEY_NULL) g = (int64_t)(cap + 1);
            else g = g_find(T, mask, hash64(w), w);
            if (g >= 0) g_update(q, T, (uint64_t)g, rows[j]);
        }
    }
}

// ------------------------------------------------------------------------
// Result columns
// ------------------------------------------------------------------------
enum OutKind : int32_t {
    OUT_KEY = 0,     // group key component -> INTEGER / BIGINT / VARCHAR id
    OUT_COUNT = 1,   // BIGINT, never NULL
    OUT_SUM_I = 2,   // BIGINT, NULL when count == 0
    OUT_SUM_F = 3,   // DOUBLE
    OUT_AVG_I = 4,   // DOUBLE = (double)sum / count
    OUT_AVG_F = 5,   // DOUBLE = sum / count
    OUT_MM_I32 = 6,  // INTEGER
    OUT_MM_I64 = 7,  // BIGINT
    OUT_MM_F = 8,    // DOUBLE (from the ordered image)
};

constexpr int MAX_OUT = 2 + 16;

struct DevOut {
    void *data[MAX_OUT];
    uint8_t *valid[MAX_OUT];
    int32_t kind[MAX_OUT];
    int32_t val_acc[MAX_OUT];  // OUT_KEY: component index
    int32_t cnt_acc[MAX_OUT];
    int32_t key_type[MAX_OUT]; // OUT_KEY: CK_I32 / CK_I64 / CK_U32
    int32_t ncols;
    int32_t key_mode;
};

__device__ __forceinline__ uint64_t key_component(int key_mode, uint64_t key, int comp) {
    if (key_mode == KM_PACK32) return comp == 0 ? (key >> 32) : (key & 0xFFFFFFFFULL);
    return key;
}

__device__ __forceinline__ void write_row(const DevOut &o, const uint64_t *rec, bool null_key,
                                          uint64_t row) {
    const uint64_t key = rec[0];
    const uint64_t *acc = rec + 1;
    for (int c = 0; c < o.ncols; c++) {
        uint64_t cnt = o.cnt_acc[c] >= 0 ? acc[o.cnt_acc[c]] : 1;
        uint8_t ok = cnt != 0;
        switch (o.kind[c]) {
        case OUT_KEY: {
            uint64_t v = key_component(o.key_mode, key, o.val_acc[c]);
            ok = !null_key;
            if (o.key_type[c] == CK_I64) ((int64_t *)o.data[c])[row] = ok ? (int64_t)v : 0;
            else ((int32_t *)o.data[c])[row] = ok ? (int32_t)(uint32_t)v : 0;
            break;
        }
        case OUT_COUNT:
            ((int64_t *)o.data[c])[row] = (int64_t)acc[o.val_acc[c]];
            ok = 1;
            break;
        case OUT_SUM_I:
        case OUT_MM_I64:
            ((int64_t *)o.data[c])[row] = ok ? (int64_t)acc[o.val_acc[c]] : 0;
            break;
        case OUT_MM_I32:
            ((int32_t *)o.data[c])[row] = ok ? (int32_t)(int64_t)acc[o.val_acc[c]] : 0;
            break;
        case OUT_SUM_F:
            ((double *)o.data[c])[row] = ok ? __longlong_as_double((long long)acc[o.val_acc[c]]) : 0.0;
            break;
        case OUT_AVG_I:
            ((double *)o.data[c])[row] = ok ? (double)(int64_t)acc[o.val_acc[c]] / (double)cnt : 0.0;
            break;
        case OUT_AVG_F:
            ((double *)o.data[c])[row] =
                ok ? __longlong_as_double((long long)acc[o.val_acc[c]]) / (double)cnt : 0.0;
            break;
        default:  // OUT_MM_F
            ((double *)o.data[c])[row] =
                ok ? __longlong_as_double((long long)f64_unord(acc[o.val_acc[c]])) : 0.0;
            break;
        }
        o.valid[c][row] = ok;
    }
}

__device__ __forceinline__ bool slot_occupied(const uint64_t *rec, uint64_t s, uint64_t cap) {
    return s < cap ? rec[0] != SENTINEL : rec[1] != 0;  // side slots: COUNT(*) != 0
}

__global__ __launch_bounds__(256) void k_occ_count(DevTable T, uint32_t *blk_counts) {
    const uint64_t cap = T.plan->cap;
    const uint64_t nrec = cap + 2;
    const uint64_t nblk = (nrec + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
    __shared__ uint32_t s_cnt;
    for (uint64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        uint32_t c = 0;
        uint64_t end = min(nrec, (blk + 1) * COMPACT_CHUNK);
        for (uint64_t s = blk * COMPACT_CHUNK + threadIdx.x; s < end; s += blockDim.x)
            c += slot_occupied(T.rec + s * T.stride, s, cap);
        atomicAdd(&s_cnt, c);
        __syncthreads();
        if (threadIdx.x == 0) blk_counts[blk] = s_cnt;
        __syncthreads();
    }
}

// Exclusive scan of the block counts in place; total -> plan->groups.
__global__ __launch_bounds__(1024) void k_scan_blocks(DevTable T, uint32_t *blk_counts) {
    const uint64_t nrec = T.plan->cap + 2;
    const uint32_t nblk = (uint32_t)((nrec + COMPACT_CHUNK - 1) / COMPACT_CHUNK);
    __shared__ uint64_t s_part[1024];
    const uint32_t per = (nblk + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
    uint64_t local = 0;
    for (uint32_t i = b0; i < b1; i++) local += blk_counts[i];
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        uint64_t v = threadIdx.x >= (uint32_t)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t i = b0; i < b1; i++) {
        uint32_t c = blk_counts[i];
        blk_counts[i] = (uint32_t)run;
        run += c;
    }
    if (threadIdx.x == blockDim.x - 1) T.plan->groups = s_part[blockDim.x - 1];
}

__global__ __launch_bounds__(256) void k_compact_write(DevTable T, const uint32_t *blk_offsets,
                                                       DevOut o) {
    const uint64_t cap = T.plan->cap;
    const uint64_t nrec = cap + 2;
    const uint64_t nblk = (nrec + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        if (threadIdx.x == 0) s_base = blk_offsets[blk];
        __syncthreads();
        for (uint64_t s0 = blk * COMPACT_CHUNK; s0 < min(nrec, (blk + 1) * COMPACT_CHUNK); s0 += blockDim.x) {
            uint64_t s = s0 + threadIdx.x;
            const uint64_t *rec = T.rec + s * T.stride;
            bool occ = s < nrec && slot_occupied(rec, s, cap);
            uint64_t bal = __ballot(occ);
            if (lane == 0) s_wave[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0;
            for (int w = 0; w < wave; w++) before += s_wave[w];
            uint32_t rank = (uint32_t)__popcll(bal & ((1ULL << lane) - 1ULL));
            if (occ) write_row(o, rec, s == cap + 1, s_base + before + rank);
            __syncthreads();
            if (threadIdx.x == 0) s_base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------
// TopN (SetTopNHint): order value of a slot as an unsigned "larger is
// better" word, ties broken by the smaller group key.
// ------------------------------------------------------------------------
enum TopnKind : int32_t {
    TK_KEY = 0,    // the group key word (single-column keys)
    TK_ACC_I = 1,  // int64 accumulator (COUNT, integer SUM/MIN/MAX)
    TK_ACC_O = 2,  // ordered-double accumulator (MIN/MAX of DOUBLE)
    TK_SUM_F = 3,  // double SUM
    TK_AVG_I = 4,
    TK_AVG_F = 5,
};

struct DevTopn {
    int32_t kind;
    int32_t acc;
    int32_t cnt_acc;
    int32_t ascending;
    int32_t key_mode;
    int32_t key_signed[2];
    uint32_t k;
    uint32_t *cand_slot;  // [blocks][k]
    uint64_t *cand_ord;   // [blocks][k][2]
    uint32_t *final_slot; // [k]
};

__device__ __forceinline__ uint64_t key_order(const DevTopn &t, uint64_t key) {
    if (t.key_mode == KM_PACK32) {
        if (t.key_signed[0]) key ^= 0x8000000000000000ULL;
        if (t.key_signed[1]) key ^= 0x0000000080000000ULL;
        return key;
    }
    return t.key_signed[0] ? key ^ 0x8000000000000000ULL : key;
}

__device__ __forceinline__ void slot_order(const DevTopn &t, const uint64_t *rec, bool null_key,
                                           uint64_t &ord, uint64_t &tie) {
    const uint64_t *acc = rec + 1;
    uint64_t o;
    switch (t.kind) {
    case TK_KEY: o = null_key ? 0 : key_order(t, rec[0]) ; break;
    case TK_ACC_I: o = acc[t.acc] ^ 0x8000000000000000ULL; break;
    case TK_ACC_O: o = acc[t.acc]; break;
    case TK_SUM_F: o = f64_ord(acc[t.acc]); break;
    case TK_AVG_I: {
        double v = (double)(int64_t)acc[t.acc] / (double)acc[t.cnt_acc];
        o = f64_ord((uint64_t)__double_as_longlong(v));
        break;
    }
    default: {
        double v = __longlong_as_double((long long)acc[t.acc]) / (double)acc[t.cnt_acc];
        o = f64_ord((uint64_t)__double_as_longlong(v));
        break;
    }
    }
    ord = t.ascending ? ~o : o;
    // smaller key wins ties; the NULL key sorts first
    tie = null_key ? ~0ULL : ~key_order(t, rec[0]);
}

__device__ __forceinline__ bool topn_better(uint64_t oa, uint64_t ta, uint32_t sa, uint64_t ob,
                                            uint64_t tb, uint32_t sb) {
    if (sb == INVALID_SLOT) return sa != INVALID_SLOT;
    if (sa == INVALID_SLOT) return false;
    return oa > ob || (oa == ob && (ta > tb || (ta == tb && sa < sb)));
}

template <int KMAX>
struct TopList {
    uint64_t o[KMAX], t[KMAX];
    uint32_t s[KMAX];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < KMAX; i++) { o[i] = 0; t[i] = 0; s[i] = INVALID_SLOT; }
    }
    // Replace the worst entry if the candidate beats it (K <= KMAX entries used).
    __device__ __forceinline__ void offer(uint64_t oc, uint64_t tc, uint32_t sc, uint32_t K) {
        int w = 0;
        uint64_t wo = o[0], wt = t[0];
        uint32_t ws = s[0];
#pragma unroll
        for (int i = 1; i < KMAX; i++) {
            if ((uint32_t)i < K && topn_better(wo, wt, ws, o[i], t[i], s[i])) {
                w = i; wo = o[i]; wt = t[i]; ws = s[i];
            }
        }
        if (!topn_better(oc, tc, sc, wo, wt, ws)) return;
#pragma unroll
        for (int i = 0; i < KMAX; i++)
            if (i == w) { o[i] = oc; t[i] = tc; s[i] = sc; }
    }
};

// Block-wide selection of the best K of blockDim.x * KMAX candidates held in
// per-thread lists; writes them best-first to out_*.
template <int KMAX>
__device__ void block_select(TopList<KMAX> &L, uint32_t K, uint32_t *out_slot, uint64_t *out_ord) {
    __shared__ uint64_t s_o[1024 / 64], s_t[1024 / 64];
    __shared__ uint32_t s_s[1024 / 64], s_owner[1024 / 64];
    __shared__ uint32_t s_win_owner;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    uint32_t used = 0;  // bit i: list entry i already emitted
    for (uint32_t r = 0; r < K; r++) {
        // this thread's best unused entry
        uint64_t bo = 0, bt = 0;
        uint32_t bs = INVALID_SLOT;
        int bi = -1;
#pragma unroll
        for (int i = 0; i < KMAX; i++) {
            if ((uint32_t)i < K && !((used >> i) & 1) && topn_better(L.o[i], L.t[i], L.s[i], bo, bt, bs)) {
                bo = L.o[i]; bt = L.t[i]; bs = L.s[i]; bi = i;
            }
        }
        uint32_t owner = threadIdx.x;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint64_t oo = __shfl_xor(bo, off), ot = __shfl_xor(bt, off);
            uint32_t os = __shfl_xor(bs, off), ow = __shfl_xor(owner, off);
            if (topn_better(oo, ot, os, bo, bt, bs)) { bo = oo; bt = ot; bs = os; owner = ow; }
        }
        if (lane == 0) { s_o[wave] = bo; s_t[wave] = bt; s_s[wave] = bs; s_owner[wave] = owner; }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t wo = s_o[0], wt = s_t[0];
            uint32_t ws = s_s[0], wn = s_owner[0];
            for (int w = 1; w < nwaves; w++)
                if (topn_better(s_o[w], s_t[w], s_s[w], wo, wt, ws)) { wo = s_o[w]; wt = s_t[w]; ws = s_s[w]; wn = s_owner[w]; }
            out_slot[r] = ws;
            if (out_ord) { out_ord[2 * r] = wo; out_ord[2 * r + 1] = wt; }
            s_win_owner = ws == INVALID_SLOT ? 0xFFFFFFFFu : wn;
        }
        __syncthreads();
        if (s_win_owner == threadIdx.x && bi >= 0) used |= 1u << bi;
        __syncthreads();
    }
}

template <int KMAX>
__global__ __launch_bounds__(256) void k_topn_local(DevTable T, DevTopn t) {
    const uint64_t cap = T.plan->cap, nrec = cap + 2;
    TopList<KMAX> L;
    L.init();
    const uint64_t per = (nrec + gridDim.x - 1) / gridDim.x;
    const uint64_t s0 = (uint64_t)blockIdx.x * per, s1 = min(nrec, s0 + per);
    for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const uint64_t *rec = T.rec + s * T.stride;
        if (!slot_occupied(rec, s, cap)) continue;
        uint64_t o, ti;
        slot_order(t, rec, s == cap + 1, o, ti);
        L.offer(o, ti, (uint32_t)s, t.k);
    }
    block_select<KMAX>(L, t.k, t.cand_slot + (uint64_t)blockIdx.x * t.k,
                       t.cand_ord + (uint64_t)blockIdx.x * t.k * 2);
}

template <int KMAX>
__global__ __launch_bounds__(1024) void k_topn_final(DevTable T, DevTopn t, uint32_t nblocks) {
    TopList<KMAX> L;
    L.init();
    const uint64_t total = (uint64_t)nblocks * t.k;
    for (uint64_t i = threadIdx.x; i < total; i += blockDim.x) {
        uint32_t s = t.cand_slot[i];
        if (s == INVALID_SLOT) continue;
        L.offer(t.cand_ord[2 * i], t.cand_ord[2 * i + 1], s, t.k);
    }
    block_select<KMAX>(L, t.k, t.final_slot, nullptr);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t n = 0;
        for (uint32_t r = 0; r < t.k; r++) n += t.final_slot[r] != INVALID_SLOT;
        T.plan->topn_count = n;
        T.plan->groups = n;
    }
}

__global__ void k_gather_write(DevTable T, DevTopn t, DevOut o) {
    const uint64_t cap = T.plan->cap;
    for (uint32_t r = threadIdx.x; r < t.k; r += blockDim.x) {
        uint32_t s = t.final_slot[r];
        if (s == INVALID_SLOT) continue;
        write_row(o, T.rec + (uint64_t)s * T.stride, s == cap + 1, r);
    }
}

}  // namespace gbagg
