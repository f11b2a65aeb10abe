// gbagg_kernels.hpp -- kernels shared by every aggregation strategy, and the
// LDS-hash strategy for low and moderate group counts.
//
//   k_hll_finish      merges HyperLogLog partial sketches into an estimate
//   k_table_init      global hash table: empty keys, identity accumulators
//   k_aggregate       LDS-hash strategy: fused filter + key hash + per-workgroup
//                     LDS table with atomic COUNT/SUM/MIN/MAX; rows whose key
//                     finds no LDS slot, and the LDS table at the end, merge
//                     into the global table with device-scope atomics
//   k_aggregate_global  global table only (scalar aggregates, tests)
//   k_occ_count / k_scan_blocks / k_compact_write
//                     append the global table's groups to the dense result
//   k_topn_table      TopN candidates of the global table
//   k_topn_final      block-wise TopN over all producers' candidates
//
// Every group leaves the operator through write_group(): dense result
// columns in full mode, or one of K candidate records per producer block in
// TopN mode (SetTopNHint).
#pragma once

#include "gbagg_device.hpp"

namespace gbagg {

constexpr int AGG_BLOCK = 256;
constexpr int AGG_ROWS = 4;          // rows per thread per tile
constexpr int COMPACT_CHUNK = 1024;  // table slots per compaction block
constexpr uint32_t INVALID_SLOT = 0xFFFFFFFFu;

// ------------------------------------------------------------------------
// Every batch of a finalize in one launch: the batches' rows concatenated,
// start[i] the first row of batch i and start[nb] the total. Workgroup w
// takes rows [w * per, (w + 1) * per) of the concatenation; f(b, lo, hi)
// runs for each batch b overlapping them, [lo, hi) local to b. Batches of a
// few row groups each thus share workgroups instead of a launch apiece.
// ------------------------------------------------------------------------
struct BatchRows {
    const DevBatch *b;
    const uint64_t *start;
    uint32_t nb;
    uint32_t pad;
    uint64_t per;
};

// The batch holding concatenated row r (the last one starting at or before r).
__device__ __forceinline__ uint32_t batch_of(const BatchRows &br, uint64_t r) {
    uint32_t lo = 0, hi = br.nb - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (br.start[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class F>
__device__ __forceinline__ void for_batch_rows(const BatchRows &br, F &&f) {
    const uint64_t total = br.start[br.nb];
    const uint64_t R0 = (uint64_t)blockIdx.x * br.per;
    const uint64_t R1 = R0 + br.per < total ? R0 + br.per : total;
    if (R0 >= R1) return;
    for (uint32_t i = batch_of(br, R0); i < br.nb && br.start[i] < R1; i++) {
        const uint64_t a = R0 > br.start[i] ? R0 : br.start[i];
        const uint64_t e = R1 < br.start[i + 1] ? R1 : br.start[i + 1];
        if (a < e) f(br.b[i], a - br.start[i], e - br.start[i]);
    }
}

// ------------------------------------------------------------------------
// HyperLogLog (Flajolet et al. 2007), p = 12: register = top 12 hash bits,
// rank = leading zeros of the remaining bits + 1.
// ------------------------------------------------------------------------
__device__ __forceinline__ void hll_add(uint32_t *reg, uint64_t h) {
    uint32_t idx = (uint32_t)(h >> (64 - HLL_P));
    uint64_t rest = h << HLL_P;
    uint32_t rank = rest ? (uint32_t)__clzll((long long)rest) + 1 : (64 - HLL_P + 1);
    // registers only grow: once warm, almost every row's rank is no larger
    // than its register, and a plain LDS read replaces the atomic
    if (__hip_atomic_load(&reg[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < rank)
        __hip_atomic_fetch_max(&reg[idx], rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave multisplit over LDS counters: every active lane adds 1 to cnt[b] (its
// partition) and gets its exclusive offset. Lanes sharing
// the bucket of the first remaining lane take one atomic for the whole group
// and their ranks in it from the group's ballot; this repeats for up to
// MSPLIT_GROUPS groups while the groups hold two lanes or more, and the
// remaining lanes (all distinct buckets, in a high fanout) take one atomic
// each. Every active lane of the wave must call it (ballots and shuffles).
constexpr int MSPLIT_GROUPS = 4;
// Default of the radix scatter's RadixParams::msplit_max: 0, the multisplit
// off. Measured on that scatter against one LDS atomic per row, it was slower
// at every fanout tried (1 to 256 partitions: +2-13%; 8192: +1%), the scatter
// being bound by its memory traffic rather than by same-address LDS atomics.
constexpr uint32_t MSPLIT_MAX_FANOUT = 0;
template <bool RET>
__device__ __forceinline__ uint32_t wave_multisplit(uint32_t *cnt, uint32_t b) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ULL << lane) - 1;
    uint64_t todo = __ballot(1);
    uint32_t pos = 0;
    bool done = false;
    for (int g = 0; g < MSPLIT_GROUPS; g++) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lb = __shfl(b, leader);
        const uint64_t same = __ballot(!done && b == lb);
        uint32_t base = 0;
        if ((int)lane == leader)
            base = __hip_atomic_fetch_add(&cnt[lb], (uint32_t)__popcll(same), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
        if (RET) base = __shfl(base, leader);
        if (!done && b == lb) {
            pos = base + (uint32_t)__popcll(same & below);
            done = true;
        }
        todo &= ~same;
        if (!todo || __popcll(same) < 2) break;  // uniform: ballot results
    }
    if (!done) {
        if (RET) pos = __hip_atomic_fetch_add(&cnt[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else __hip_atomic_fetch_add(&cnt[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return pos;
}

// Register-wise max over `nparts` sketches: thread per register, blockIdx.y
// picks one of gridDim.y slices of the sketches; slice y's merge goes to
// merged[y * HLL_M ...] (k_hll_finish folds the slices).
__global__ __launch_bounds__(256) void k_hll_merge(const uint8_t *partial, uint32_t nparts, uint8_t *merged,
                                                   const uint32_t *gate) {
    if (gate && *gate == 0) return;  // gate: as k_rp_count's
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= HLL_M) return;
    const uint32_t per = (nparts + gridDim.y - 1) / gridDim.y;
    const uint32_t p0 = min(nparts, blockIdx.y * per), p1 = min(nparts, p0 + per);
    uint32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    uint32_t p = p0;
    for (; p + 4 <= p1; p += 4) {
        m0 = max(m0, (uint32_t)partial[(uint64_t)p * HLL_M + i]);
        m1 = max(m1, (uint32_t)partial[(uint64_t)(p + 1) * HLL_M + i]);
        m2 = max(m2, (uint32_t)partial[(uint64_t)(p + 2) * HLL_M + i]);
        m3 = max(m3, (uint32_t)partial[(uint64_t)(p + 3) * HLL_M + i]);
    }
    for (; p < p1; p++) m0 = max(m0, (uint32_t)partial[(uint64_t)p * HLL_M + i]);
    merged[(uint64_t)blockIdx.y * HLL_M + i] = (uint8_t)max(max(m0, m1), max(m2, m3));
}

// One block of 1024 threads: max-merge `nparts` sketches, write the estimate.
__global__ __launch_bounds__(1024) void k_hll_finish(const uint8_t *partial, uint32_t nparts,
                                                     DevPlan *plan, const uint32_t *gate) {
    if (gate && *gate == 0) return;
    __shared__ double s_sum[1024];
    __shared__ uint32_t s_zero[1024];
    double sum = 0.0;
    uint32_t zeros = 0;
    // thread t owns registers 4t .. 4t + 3 (HLL_M == 4 * 1024), read as one
    // word per sketch; the bytewise max of a word is four register maxima
    static_assert(HLL_M == 4 * 1024, "one word of registers per thread");
    uint32_t m4 = 0;
#pragma unroll 8
    for (uint32_t p = 0; p < nparts; p++) {
        const uint32_t x = ((const uint32_t *)(partial + (uint64_t)p * HLL_M))[threadIdx.x];
        m4 = max(m4 & 0xFFu, x & 0xFFu) | max(m4 & 0xFF00u, x & 0xFF00u) | max(m4 & 0xFF0000u, x & 0xFF0000u) |
             max(m4 & 0xFF000000u, x & 0xFF000000u);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t m = (m4 >> (8 * k)) & 0xFFu;
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
        plan->est_groups = (uint64_t)(e + 0.5);
    }
}

// *p = v, in stream order.
__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ __launch_bounds__(256) void k_table_init(DevTable T, const DevQuery *__restrict__ qp) {
    const DevQuery &q = *qp;
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
#pragma nounroll
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

// Lookup without insert: the record of k, or -1 when k is absent.
__device__ __forceinline__ int64_t g_lookup(const DevTable &T, uint64_t mask, uint64_t h, uint64_t k) {
    uint64_t pos = h & mask;
#pragma nounroll
    for (uint32_t i = 0; i < T.max_probe; i++) {
        const uint64_t cur = __hip_atomic_load(T.rec + pos * T.stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == k) return (int64_t)pos;
        if (cur == SENTINEL) return -1;
        pos = (pos + 1) & mask;
    }
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

// Merge one partial state (acc a at acc[a * stride]) into global record g.
__device__ __forceinline__ void g_merge(const DevQuery &q, const DevTable &T, uint64_t g,
                                        const uint64_t *acc, uint32_t stride) {
    uint64_t *dst = T.rec + g * T.stride + 1;
    for (int a = 0; a < q.nacc; a++) {
        uint64_t x = acc[(uint64_t)a * stride];
        if (acc_nontrivial(q.acc_op[a], x)) acc_apply<__HIP_MEMORY_SCOPE_AGENT>(dst + a, q.acc_op[a], x);
    }
}

// Rows with a NULL group key or the SENTINEL key go to the table's two side
// records (cap + 1, cap); returns false for rows the key logic dropped.
__device__ __forceinline__ bool side_row(const DevQuery &q, const DevTable &T, uint64_t cap, int kr,
                                         uint64_t w, const Row &row) {
    if (kr == KEY_NULL) { g_update(q, T, cap + 1, row); return true; }
    if (w == SENTINEL) { g_update(q, T, cap, row); return true; }
    return false;
}

// ------------------------------------------------------------------------
// Per-workgroup LDS hash table: keys[S], then acc[nacc][S], then a fill count.
// ------------------------------------------------------------------------
struct LdsTable {
    uint64_t *key;
    uint64_t *acc;
    uint32_t *fill;
    uint32_t S;
    uint32_t limit;
};

__device__ __forceinline__ LdsTable lds_table_setup(uint64_t *smem, const DevQuery &q, uint32_t S,
                                                    uint32_t limit) {
    LdsTable t;
    t.S = S;
    t.limit = limit;
    t.key = smem;
    t.acc = smem + S;
    t.fill = (uint32_t *)(t.acc + (uint64_t)S * q.nacc);
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) t.key[i] = SENTINEL;
    for (int a = 0; a < q.nacc; a++) {
        uint64_t id = acc_identity(q.acc_op[a]);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) t.acc[(uint64_t)a * S + i] = id;
    }
    if (threadIdx.x == 0) *t.fill = 0;
    return t;
}

// Any table size: the home slot is the low 32 hash bits scaled to [0, S)
// (Lemire's multiply-shift), so S need not be a power of two and the LDS can
// be filled exactly. Radix partitions use the top hash bits, which stay
// independent of these. The probe bound (128) is rarely reached even at the
// radix tables' 4/5 fill; a key that reaches it goes to the global table.
__device__ __forceinline__ int lds_find(const LdsTable &t, uint64_t h, uint64_t k) {
    uint32_t pos = (uint32_t)(((uint64_t)(uint32_t)h * t.S) >> 32);
    uint32_t probes = 1;
    uint64_t cur = __hip_atomic_load(&t.key[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
        // scan to k or to an empty slot (single exit: cheap divergent loop)
        while (cur != k && cur != SENTINEL && probes < 128) {
            pos = pos + 1 == t.S ? 0 : pos + 1;
            cur = __hip_atomic_load(&t.key[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            probes++;
        }
        if (cur == k) return (int)pos;
        if (cur != SENTINEL) return -1;  // probe bound
        if (__hip_atomic_load(t.fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= t.limit) return -1;
        uint64_t expected = SENTINEL;
        __hip_atomic_compare_exchange_strong(&t.key[pos], &expected, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        if (expected == SENTINEL) {
            __hip_atomic_fetch_add(t.fill, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return (int)pos;
        }
        cur = expected;  // taken meanwhile: by k itself (found) or by another key (keep probing)
    }
}

__device__ __forceinline__ void lds_update(const DevQuery &q, const LdsTable &t, int slot,
                                           const Row &row) {
    for (int a = 0; a < q.nacc; a++) {
        uint64_t x;
        if (acc_input(q, a, row, x))
            acc_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(t.acc + (uint64_t)a * t.S + slot, q.acc_op[a], x);
    }
}

// All 64 lanes carry the same key (scalar aggregates, clustered keys): reduce
// across the wave and let lane 0 apply. dst holds acc a at dst[a * stride].
__device__ __forceinline__ void wave_reduce_apply(const DevQuery &q, const Row &row, uint64_t *dst,
                                                  uint32_t stride, bool lds) {
    const int lane = threadIdx.x & 63;
    for (int a = 0; a < q.nacc; a++) {
        int op = q.acc_op[a];
        uint64_t x;
        if (!acc_input(q, a, row, x)) x = acc_identity(op);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x = acc_combine(op, x, __shfl_xor(x, off));
        if (lane == 0 && acc_nontrivial(op, x)) {
            if (lds) acc_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(dst + (uint64_t)a * stride, op, x);
            else acc_apply<__HIP_MEMORY_SCOPE_AGENT>(dst + (uint64_t)a * stride, op, x);
        }
    }
}

template <class S>
__global__ __launch_bounds__(AGG_BLOCK) void k_aggregate(const DevQuery *__restrict__ qp, BatchRows br, DevTable T,
                                                         uint32_t lds_slots, uint32_t fill_limit) {
    const DevQuery &q = *qp;
    const S sh(q);
    extern __shared__ uint64_t smem[];
    LdsTable t = lds_table_setup(smem, q, lds_slots, fill_limit);
    __syncthreads();
    const uint64_t cap = T.plan->cap;
    const uint64_t mask = cap - 1;
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t lo, uint64_t n) {
    const DevBatch b = bb;  // column pointers stay in registers
    for (uint64_t base = lo; base < n; base += tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row_s(sh, b, r, sh.nu(), rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            bool live = r < n;
            if (live && sh.npred()) live = eval_preds_s(sh, q, b, rows[j]);
            uint64_t w = 0;
            int kr = KEY_DROP;
            if (live) kr = make_key_s(sh, q, b, rows[j], w);
            if (kr == KEY_DROP) continue;
            if (side_row(q, T, cap, kr, w, rows[j])) continue;
            const uint64_t h = hash64(w);
            int slot = lds_find(t, h, w);
            if (slot >= 0) {
                row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, rows[j], t.acc + slot, t.S);
            } else {
                int64_t g = g_find(T, mask, h, w);
                if (g >= 0) row_update_s<__HIP_MEMORY_SCOPE_AGENT>(sh, rows[j], T.rec + (uint64_t)g * T.stride + 1, 1);
            }
        }
    }
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < t.S; i += blockDim.x) {
        uint64_t k = t.key[i];
        if (k == SENTINEL) continue;
        int64_t g = g_find(T, mask, hash64(k), k);
        if (g >= 0) g_merge(q, T, (uint64_t)g, t.acc + i, t.S);
    }
}

// ------------------------------------------------------------------------
// Direct-mapped strategy (IntGroupAgg's dense slot_of[key],
// src/execution/int_group_agg.cpp:69-277): with the exact key range
// [kmin, kmin + span) known, key word w owns global record w - kmin, so a row
// needs no hash and no probe.
// ------------------------------------------------------------------------
constexpr uint32_t OVF_DIRECT = 16u;  // DevPlan::overflow bit: a key outside the measured range

// Exact range of the key words of the rows that pass the filter (NULL, dropped
// and SENTINEL keys excluded) into out[0] (min) and out[1] (max): one atomic
// min / max per block. gate_span > 0: a pass launched before the plan is read
// back, which runs only for a sample of few distinct keys (radix_gate clear)
// whose sampled key range (plan->key_min / key_max) spans less than gate_span.
template <class S>
__global__ __launch_bounds__(AGG_BLOCK) void k_key_range(const DevQuery *__restrict__ qp, BatchRows br,
                                                         const DevPlan *plan, int64_t *out, uint64_t gate_span) {
    if (gate_span && (plan->radix_gate || plan->key_min > plan->key_max ||
                      (uint64_t)plan->key_max - (uint64_t)plan->key_min >= gate_span))
        return;
    const DevQuery &q = *qp;
    const S sh(q);
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t r0, uint64_t n) {
    const DevBatch b = bb;
    for (uint64_t base = r0; base < n; base += tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row_s(sh, b, r, sh.nkp(), rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            bool live = r < n;
            if (live && sh.npred()) live = eval_preds_s(sh, q, b, rows[j]);
            uint64_t w = 0;
            if (live && make_key_s(sh, q, b, rows[j], w) == KEY_OK && w != SENTINEL) {
                lo = min(lo, (int64_t)w);
                hi = max(hi, (int64_t)w);
            }
        }
    }
    });
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (int64_t)__shfl_xor(lo, off));
        hi = max(hi, (int64_t)__shfl_xor(hi, off));
    }
    // one device-scope update per block: updates of one word serialise
    __shared__ int64_t s_lo[AGG_BLOCK / 64], s_hi[AGG_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < AGG_BLOCK / 64; w++) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        if (lo <= hi) {
            __hip_atomic_fetch_min(&out[0], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&out[1], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void k_key_range_init(DevPlan *plan) {
    plan->key_min = INT64_MAX;
    plan->key_max = INT64_MIN;
}

// Exact largest exponents of every double SUM input column over all rows of
// a batch (valid rows; the filter is not applied, a superset is a bound too):
// one atomic max per wave and column.
__global__ __launch_bounds__(AGG_BLOCK) void k_fx_range(const DevQuery *__restrict__ qp, BatchRows br,
                                                        DevPlan *plan) {
    const DevQuery &q = *qp;
    const uint32_t fxm = fx_slots(q);
    uint32_t emax[MAXU] = {};
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t lo, uint64_t n) {
        const DevBatch b = bb;
        for (uint64_t r = lo + threadIdx.x; r < n; r += blockDim.x) {
#pragma unroll
            for (int u = 0; u < MAXU; u++) {
                if (!((fxm >> u) & 1)) continue;
                if (b.valid[u] && !gld_stream(b.valid[u] + r)) continue;
                emax[u] = max(emax[u], fx_exp_of(gld_stream((const uint64_t *)b.data[u] + r)));
            }
        }
    });
#pragma unroll
    for (int u = 0; u < MAXU; u++) {
        if (!((fxm >> u) & 1)) continue;
        uint32_t e = emax[u];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) e = max(e, (uint32_t)__shfl_xor(e, off));
        if ((threadIdx.x & 63) == 0 && e)
            __hip_atomic_fetch_max(&plan->fx_exp[u], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// LDS: one workgroup-private dense table [nacc][span], folded into the global
// records at the end (span * nacc words must fit the workgroup's LDS). Else a
// workgroup-private LDS hash table of `cache` slots absorbs the frequent keys
// (as in k_aggregate) and the other rows update their global record directly
// with device-scope atomics.
template <class S, bool LDS>
__global__ __launch_bounds__(AGG_BLOCK) void k_direct(const DevQuery *__restrict__ qp, BatchRows br, DevTable T,
                                                      int64_t kmin, uint32_t span, uint32_t cache) {
    const DevQuery &q = *qp;
    const S sh(q);
    extern __shared__ uint64_t smem[];
    LdsTable t{};
    if constexpr (LDS) {
        for (int a = 0; a < sh.nacc(); a++) {
            const uint64_t id = acc_identity(sh.op(a));
            for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) smem[(uint64_t)a * span + i] = id;
        }
    } else {
        t = lds_table_setup(smem, q, cache, cache * 3 / 4);
    }
    __syncthreads();
    const uint64_t cap = T.plan->cap;
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t lo, uint64_t n) {
    const DevBatch b = bb;
    for (uint64_t base = lo; base < n; base += tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row_s(sh, b, r, sh.nu(), rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            bool live = r < n;
            if (live && sh.npred()) live = eval_preds_s(sh, q, b, rows[j]);
            uint64_t w = 0;
            int kr = KEY_DROP;
            if (live) kr = make_key_s(sh, q, b, rows[j], w);
            if (kr == KEY_DROP) continue;
            if (side_row(q, T, cap, kr, w, rows[j])) continue;
            const uint64_t slot = w - (uint64_t)kmin;
            if (slot >= span) {  // cannot happen with the exact range; never written out of bounds
                __hip_atomic_fetch_or(&T.plan->overflow, OVF_DIRECT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                continue;
            }
            if constexpr (LDS) {
                row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, rows[j], smem + slot, span);
            } else {
                const int ls = lds_find(t, hash64(w), w);
                if (ls >= 0) {
                    row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, rows[j], t.acc + ls, t.S);
                } else {
                    uint64_t *rec = T.rec + slot * T.stride;
                    if (__hip_atomic_load(rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == SENTINEL)
                        __hip_atomic_store(rec, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    row_update_s<__HIP_MEMORY_SCOPE_AGENT>(sh, rows[j], rec + 1, 1);
                }
            }
        }
    }
    });
    __syncthreads();
    if constexpr (LDS) {
        for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) {
            if (smem[i] == 0) continue;  // COUNT(*) of the slot (accumulator 0)
            __hip_atomic_store(T.rec + (uint64_t)i * T.stride, (uint64_t)kmin + i, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            g_merge(q, T, i, smem + i, span);
        }
    } else {
        for (uint32_t i = threadIdx.x; i < t.S; i += blockDim.x) {
            const uint64_t k = t.key[i];
            if (k == SENTINEL) continue;
            const uint64_t slot = k - (uint64_t)kmin;  // within [0, span): only in-range keys enter the cache
            __hip_atomic_store(T.rec + slot * T.stride, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            g_merge(q, T, slot, t.acc + i, t.S);
        }
    }
}

// Global-table-only variant (no LDS table).
__global__ __launch_bounds__(AGG_BLOCK) void k_aggregate_global(const DevQuery *__restrict__ qp, BatchRows br,
                                                                DevTable T) {
    const DevQuery &q = *qp;
    const uint64_t cap = T.plan->cap;
    const uint64_t mask = cap - 1;
    const uint64_t tile = (uint64_t)AGG_BLOCK * AGG_ROWS;
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t lo, uint64_t n) {
    const DevBatch b = bb;  // column pointers stay in registers
    for (uint64_t base = lo; base < n; base += tile) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * AGG_BLOCK + threadIdx.x;
            if (r < n) load_row(q, b, r, (1u << q.nu) - 1, rows[j]);
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
                    if (g >= 0) wave_reduce_apply(q, rows[j], T.rec + (uint64_t)g * T.stride + 1, 1, false);
                    continue;
                }
            }
            if (kr == KEY_DROP) continue;
            if (side_row(q, T, cap, kr, w, rows[j])) continue;
            int64_t g = g_find(T, mask, hash64(w), w);
            if (g >= 0) g_update(q, T, (uint64_t)g, rows[j]);
        }
    }
    });
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
    uint64_t capacity;         // rows the columns hold
    int32_t wide_acc[2];       // KM_WIDE: the MIN and MAX accumulators of component wide_comp
    int32_t wide_comp;
    int32_t pad;
    uint64_t wide_seed;
    uint32_t *flags;           // DevPlan::overflow
    int32_t lo_acc[MAX_OUT];   // OUT_SUM_F / OUT_AVG_F: the ACC_SUM_FL accumulator (val_acc: ACC_SUM_FH)
    int32_t fx_shift[MAX_OUT]; // and their scale s
};

constexpr uint32_t OVF_WIDE = 8u;  // DevPlan::overflow bit: two KM_WIDE keys share a key word

// KM_WIDE: the group's component kept in accumulators (flags a shared key word).
__device__ __forceinline__ uint64_t wide_kept(const int32_t (&wa)[2], uint32_t *flags, const uint64_t *acc,
                                            uint32_t stride) {
    const uint64_t lo = acc[(uint64_t)wa[0] * stride], hi = acc[(uint64_t)wa[1] * stride];
    if (lo != hi) __hip_atomic_fetch_or(flags, OVF_WIDE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return lo;
}

__device__ __forceinline__ uint64_t key_component(const DevOut &o, uint64_t key, int comp, const uint64_t *acc,
                                                  uint32_t stride) {
    if (o.key_mode == KM_PACK32) return comp == 0 ? (key >> 32) : (key & 0xFFFFFFFFULL);
    if (o.key_mode == KM_WIDE) {
        const uint64_t c = wide_kept(o.wide_acc, o.flags, acc, stride);
        return comp == o.wide_comp ? c : key ^ wide_mix(c, o.wide_seed);
    }
    return key;
}

// Double SUM of output column c from its accumulators (fixed-point words, or
// an IEEE double after the host switched a query to atomics); flags groups whose
// fixed-point sum may have lost precision.
__device__ __forceinline__ double out_fsum(const DevOut &o, int c, uint64_t v, const uint64_t *acc, uint32_t stride,
                                           uint64_t cnt) {
    if (o.lo_acc[c] < 0) return __longlong_as_double((long long)v);
    const uint64_t lo = acc[(uint64_t)o.lo_acc[c] * stride];
    if (fx_hazard(v, lo, cnt)) __hip_atomic_fetch_or(o.flags, OVF_FX_PREC, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return fx_value(v, lo, o.fx_shift[c]);
}

// One result row from a group state: acc a at acc[a * stride].
__device__ __forceinline__ void write_group(const DevOut &o, uint64_t key, const uint64_t *acc,
                                            uint32_t stride, bool null_key, uint64_t row) {
    if (row >= o.capacity) return;
    for (int c = 0; c < o.ncols; c++) {
        const int va = o.val_acc[c];
        uint64_t cnt = o.cnt_acc[c] >= 0 ? acc[(uint64_t)o.cnt_acc[c] * stride] : 1;
        uint8_t ok = cnt != 0;
        uint64_t v = o.kind[c] == OUT_KEY ? 0 : acc[(uint64_t)va * stride];
        switch (o.kind[c]) {
        case OUT_KEY: {
            uint64_t kv = key_component(o, key, va, acc, stride);
            ok = !null_key;
            if (o.key_type[c] == CK_I64) gst((int64_t *)o.data[c] + row, ok ? (int64_t)kv : (int64_t)0);
            else gst((int32_t *)o.data[c] + row, ok ? (int32_t)(uint32_t)kv : (int32_t)0);
            break;
        }
        case OUT_COUNT:
            gst((int64_t *)o.data[c] + row, (int64_t)v);
            ok = 1;
            break;
        case OUT_SUM_I:
        case OUT_MM_I64:
            gst((int64_t *)o.data[c] + row, ok ? (int64_t)v : (int64_t)0);
            break;
        case OUT_MM_I32:
            gst((int32_t *)o.data[c] + row, ok ? (int32_t)(int64_t)v : (int32_t)0);
            break;
        case OUT_SUM_F:
            gst((double *)o.data[c] + row, ok ? out_fsum(o, c, v, acc, stride, cnt) : 0.0);
            break;
        case OUT_AVG_I:
            gst((double *)o.data[c] + row, ok ? (double)(int64_t)v / (double)cnt : 0.0);
            break;
        case OUT_AVG_F:
            gst((double *)o.data[c] + row, ok ? out_fsum(o, c, v, acc, stride, cnt) / (double)cnt : 0.0);
            break;
        default:  // OUT_MM_F
            gst((double *)o.data[c] + row, ok ? __longlong_as_double((long long)f64_unord_mm(v)) : 0.0);
            break;
        }
        gst(o.valid[c] + row, ok);
    }
}

// Block-wide exclusive prefix of `mine` plus one atomicAdd on *counter for the
// block total: returns this thread's first output row. All threads must call.
__device__ __forceinline__ uint64_t block_reserve(uint32_t mine, uint64_t *counter) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint64_t s_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < nw; w++) {
            uint32_t c = s_wave[w];
            s_wave[w] = run;
            run += c;
        }
        s_base = run ? __hip_atomic_fetch_add(counter, (uint64_t)run, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                     : 0;
    }
    __syncthreads();
    const uint64_t first = s_base + s_wave[wave] + incl - mine;
    __syncthreads();  // the shared words are reused by the next call
    return first;
}

// ------------------------------------------------------------------------
// TopN (SetTopNHint): order value of a group as an unsigned "larger is
// better" word; ties go to the smaller group key.
// ------------------------------------------------------------------------
enum TopnKind : int32_t {
    TK_KEY = 0,    // the group key word (single-column keys)
    TK_ACC_I = 1,  // int64 accumulator (COUNT, integer SUM/MIN/MAX)
    TK_ACC_O = 2,  // ordered-double accumulator (MIN/MAX of DOUBLE)
    TK_SUM_F = 3,  // double SUM (fixed-point words acc, lo_acc)
    TK_AVG_I = 4,
    TK_AVG_F = 5,
};

struct DevTopn {
    int32_t enabled;
    int32_t kind;
    int32_t acc;
    int32_t cnt_acc;
    int32_t ascending;
    int32_t key_mode;
    int32_t key_signed[2];
    uint32_t k;
    uint32_t rec_words;   // 1 + nacc
    uint64_t *cand_rec;   // candidate group states: key, accs
    uint8_t *cand_null;
    uint32_t cand_cap;
    int32_t wide_acc[2];  // KM_WIDE: the MIN and MAX accumulators of the kept component
    uint32_t *flags;      // DevPlan::overflow
    int32_t lo_acc;       // TK_SUM_F / TK_AVG_F: the ACC_SUM_FL accumulator
    int32_t fx_shift;     // and the scale s
};

__device__ __forceinline__ uint64_t key_order(const DevTopn &t, uint64_t key) {
    if (t.key_mode == KM_PACK32) {
        if (t.key_signed[0]) key ^= 0x8000000000000000ULL;
        if (t.key_signed[1]) key ^= 0x0000000080000000ULL;
        return key;
    }
    return t.key_signed[0] ? key ^ 0x8000000000000000ULL : key;
}

// The TopN order column's double sum (see out_fsum).
__device__ __forceinline__ double topn_fsum(const DevTopn &t, uint64_t v, const uint64_t *acc, uint32_t stride) {
    if (t.lo_acc < 0) return __longlong_as_double((long long)v);
    const uint64_t lo = acc[(uint64_t)t.lo_acc * stride];
    if (fx_hazard(v, lo, acc[(uint64_t)t.cnt_acc * stride]))
        __hip_atomic_fetch_or(t.flags, OVF_FX_PREC, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return fx_value(v, lo, t.fx_shift);
}

__device__ __forceinline__ void group_order(const DevTopn &t, uint64_t key, const uint64_t *acc,
                                            uint32_t stride, bool null_key, uint64_t &ord,
                                            uint64_t &tie) {
    uint64_t o;
    const uint64_t v = acc[(uint64_t)t.acc * stride];
    // every group of a TopN query passes here once: the shared-key-word check
    if (t.key_mode == KM_WIDE) (void)wide_kept(t.wide_acc, t.flags, acc, stride);
    switch (t.kind) {
    case TK_KEY: o = null_key ? 0 : key_order(t, key); break;
    case TK_ACC_I: o = v ^ 0x8000000000000000ULL; break;
    case TK_ACC_O: o = v; break;
    case TK_SUM_F:
        o = f64_ord((uint64_t)__double_as_longlong(topn_fsum(t, v, acc, stride)));
        break;
    case TK_AVG_I: {
        double d = (double)(int64_t)v / (double)acc[(uint64_t)t.cnt_acc * stride];
        o = f64_ord((uint64_t)__double_as_longlong(d));
        break;
    }
    default: {
        double d = topn_fsum(t, v, acc, stride) / (double)acc[(uint64_t)t.cnt_acc * stride];
        o = f64_ord((uint64_t)__double_as_longlong(d));
        break;
    }
    }
    ord = t.ascending ? ~o : o;
    tie = null_key ? ~0ULL : ~key_order(t, key);  // the NULL key sorts first
}

// (oa, ta) larger first, then the smaller producer slot. Empty entries always
// carry (0, 0, INVALID_SLOT), so the plain lexicographic order puts them after
// every real entry. Evaluated without branches: a short-circuit chain here
// compiles to divergent control flow whose mask bookkeeping dominates the
// block-wide selections.
__device__ __forceinline__ bool topn_better(uint64_t oa, uint64_t ta, uint32_t sa, uint64_t ob,
                                            uint64_t tb, uint32_t sb) {
    const bool o_gt = oa > ob, o_eq = oa == ob, t_gt = ta > tb, t_eq = ta == tb, s_lt = sa < sb;
    return o_gt | (o_eq & (t_gt | (t_eq & s_lt)));
}

template <int KMAX>
struct TopList {
    uint64_t o[KMAX], t[KMAX];
    uint32_t s[KMAX];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < KMAX; i++) { o[i] = 0; t[i] = 0; s[i] = INVALID_SLOT; }
    }
    // Replace the worst of the first K entries if the candidate beats it.
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

// Block-wide selection of the best K entries over the per-thread lists;
// winners' ids land in s_win[0..K) best-first (INVALID_SLOT when fewer).
template <int KMAX>
__device__ void block_select(TopList<KMAX> &L, uint32_t K, uint32_t *s_win) {
    __shared__ uint64_t s_o[16], s_t[16];
    __shared__ uint32_t s_s[16], s_owner[16];
    __shared__ uint32_t s_win_owner, s_valid;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
    // rounds: no more than the block's listed entries (often none: a producer
    // whose groups all fell below the TopN bound)
    if (threadIdx.x == 0) s_valid = 0;
    __syncthreads();
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < KMAX; i++) mine += (uint32_t)i < K && L.s[i] != INVALID_SLOT;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (lane == 0 && mine) atomicAdd(&s_valid, mine);
    __syncthreads();
    const uint32_t rounds = min(K, s_valid);
    for (uint32_t r = rounds + threadIdx.x; r < K; r += blockDim.x) s_win[r] = INVALID_SLOT;
    uint32_t used = 0;  // bit i: list entry i already emitted
    for (uint32_t r = 0; r < rounds; r++) {
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
                if (topn_better(s_o[w], s_t[w], s_s[w], wo, wt, ws)) {
                    wo = s_o[w]; wt = s_t[w]; ws = s_s[w]; wn = s_owner[w];
                }
            s_win[r] = ws;
            s_win_owner = ws == INVALID_SLOT ? 0xFFFFFFFFu : wn;
        }
        __syncthreads();
        if (s_win_owner == threadIdx.x && bi >= 0) used |= 1u << bi;
        __syncthreads();
    }
    __syncthreads();
}

// One round of wave-wide argmax under topn_better: every lane ends with the
// wave's best (o, t, s) and the lane that holds it.
__device__ __forceinline__ void wave_best(uint64_t &bo, uint64_t &bt, uint32_t &bs, uint32_t &owner) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t oo = __shfl_xor(bo, off), ot = __shfl_xor(bt, off);
        const uint32_t os = __shfl_xor(bs, off), ow = __shfl_xor(owner, off);
        if (topn_better(oo, ot, os, bo, bt, bs) || (!topn_better(bo, bt, bs, oo, ot, os) && ow < owner)) {
            bo = oo; bt = ot; bs = os; owner = ow;
        }
    }
}

// Exact top-K (K <= KMAX) of NI candidate entries per thread, best first,
// into s_win: each wave takes its K best by K rounds of wave argmax (no
// block barrier), then wave 0 picks the block's K best from the waves' lists.
// Entry j of a thread is (o[j], t[j], s[j]); s[j] = INVALID_SLOT is empty.
template <int KMAX, int NI>
__device__ void block_topk(const uint64_t (&o)[NI], const uint64_t (&t)[NI], const uint32_t (&s)[NI], uint32_t K,
                           uint32_t *s_win) {
    __shared__ uint64_t w_o[16 * KMAX], w_t[16 * KMAX];
    __shared__ uint32_t w_s[16 * KMAX];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
    // rounds: no more than the entries present (wave-wide counts by ballot)
    uint32_t nw = 0;
#pragma unroll
    for (int j = 0; j < NI; j++) nw += (uint32_t)__popcll(__ballot(s[j] != INVALID_SLOT));
    const uint32_t rounds1 = min(K, nw);
    if (lane == 0)
        for (uint32_t r = rounds1; r < K; r++) {
            w_o[wave * KMAX + r] = 0;
            w_t[wave * KMAX + r] = 0;
            w_s[wave * KMAX + r] = INVALID_SLOT;
        }
    uint32_t taken = 0;
    for (uint32_t r = 0; r < rounds1; r++) {
        uint64_t bo = 0, bt = 0;
        uint32_t bs = INVALID_SLOT;
        int bj = -1;
#pragma unroll
        for (int j = 0; j < NI; j++)
            if (!((taken >> j) & 1) && topn_better(o[j], t[j], s[j], bo, bt, bs)) {
                bo = o[j]; bt = t[j]; bs = s[j]; bj = j;
            }
        uint32_t owner = lane;
        wave_best(bo, bt, bs, owner);
        if (owner == lane && bj >= 0 && bs != INVALID_SLOT) taken |= 1u << bj;
        if (lane == 0) {
            w_o[wave * KMAX + r] = bo;
            w_t[wave * KMAX + r] = bt;
            w_s[wave * KMAX + r] = bs;
        }
    }
    __syncthreads();
    if (wave == 0) {
        constexpr int NE = (16 * KMAX + 63) / 64;
        uint64_t eo[NE], et[NE];
        uint32_t es[NE];
#pragma unroll
        for (int j = 0; j < NE; j++) {
            const uint32_t e = lane + 64 * j;
            const bool ok = e / KMAX < nwaves && e % KMAX < K;
            eo[j] = ok ? w_o[e] : 0;
            et[j] = ok ? w_t[e] : 0;
            es[j] = ok ? w_s[e] : INVALID_SLOT;
        }
        uint32_t ne = 0;
#pragma unroll
        for (int j = 0; j < NE; j++) ne += (uint32_t)__popcll(__ballot(es[j] != INVALID_SLOT));
        const uint32_t rounds2 = min(K, ne);
        for (uint32_t r = rounds2 + lane; r < K; r += 64) s_win[r] = INVALID_SLOT;
        uint32_t tk = 0;
        for (uint32_t r = 0; r < rounds2; r++) {
            uint64_t bo = 0, bt = 0;
            uint32_t bs = INVALID_SLOT;
            int bj = -1;
#pragma unroll
            for (int j = 0; j < NE; j++)
                if (!((tk >> j) & 1) && topn_better(eo[j], et[j], es[j], bo, bt, bs)) {
                    bo = eo[j]; bt = et[j]; bs = es[j]; bj = j;
                }
            uint32_t owner = lane;
            wave_best(bo, bt, bs, owner);
            if (owner == lane && bj >= 0 && bs != INVALID_SLOT) tk |= 1u << bj;
            if (lane == 0) s_win[r] = bs;
        }
    }
    __syncthreads();
}

// Appends this block's winners (producer-local ids in s_win) to the candidate
// buffer; fetch(id, key, acc_ptr, stride, null) describes a producer group.
// With a fixed slot (fixed_slot >= 0) the block owns candidates
// [fixed_slot * K, fixed_slot * K + K) and marks the ones it leaves unused as
// empty (cand_null = CAND_EMPTY); thousands of producers then need no shared
// counter, whose same-address atomics would serialise them.
constexpr uint8_t CAND_EMPTY = 2;
template <typename Fetch>
__device__ void topn_push(const DevTopn &t, DevPlan *plan, const uint32_t *s_win, Fetch fetch,
                          int64_t fixed_slot = -1) {
    __shared__ uint32_t s_nwin, s_base;
    if (threadIdx.x == 0) {
        uint32_t n = 0;
        for (uint32_t r = 0; r < t.k; r++) n += s_win[r] != INVALID_SLOT;
        s_nwin = n;
        if (fixed_slot >= 0) s_base = (uint32_t)fixed_slot * t.k;
        else s_base = n ? __hip_atomic_fetch_add(&plan->cand_count, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    }
    __syncthreads();
    if (fixed_slot >= 0)
        for (uint32_t r = s_nwin + threadIdx.x; r < t.k; r += blockDim.x)
            if (s_base + r < t.cand_cap) t.cand_null[s_base + r] = CAND_EMPTY;
    for (uint32_t r = threadIdx.x; r < s_nwin; r += blockDim.x) {
        uint32_t dst = s_base + r;
        if (dst >= t.cand_cap) {
            __hip_atomic_fetch_or(&plan->overflow, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        uint64_t key;
        const uint64_t *acc;
        uint32_t stride;
        bool null_key;
        fetch(s_win[r], key, acc, stride, null_key);
        uint64_t *out = t.cand_rec + (uint64_t)dst * t.rec_words;
        out[0] = key;
        for (uint32_t a = 0; a + 1 < t.rec_words; a++) out[1 + a] = acc[(uint64_t)a * stride];
        t.cand_null[dst] = null_key;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------
// Global table -> result (append) or TopN candidates
// ------------------------------------------------------------------------
__device__ __forceinline__ bool slot_occupied(const uint64_t *rec, uint64_t s, uint64_t cap) {
    return s < cap ? rec[0] != SENTINEL : rec[1] != 0;  // side records: COUNT(*) != 0
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

// Exclusive scan of the block counts in place, starting at the rows already
// in the result (plan->groups); plan->groups grows by the table's groups.
__global__ __launch_bounds__(1024) void k_scan_blocks(DevTable T, uint32_t *blk_counts,
                                                      uint64_t *blk_base) {
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
    const uint64_t start = T.plan->groups;
    uint64_t run = start + (threadIdx.x ? s_part[threadIdx.x - 1] : 0);
    for (uint32_t i = b0; i < b1; i++) {
        blk_base[i] = run;
        run += blk_counts[i];
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) T.plan->groups = start + s_part[blockDim.x - 1];
}

__global__ __launch_bounds__(256) void k_compact_write(DevTable T, const uint64_t *blk_base,
                                                       const DevOut *__restrict__ op) {
    const DevOut &o = *op;
    const uint64_t cap = T.plan->cap;
    const uint64_t nrec = cap + 2;
    const uint64_t nblk = (nrec + COMPACT_CHUNK - 1) / COMPACT_CHUNK;
    __shared__ uint32_t s_wave[4];
    __shared__ uint64_t s_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        if (threadIdx.x == 0) s_base = blk_base[blk];
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
            if (occ) write_group(o, rec[0], rec + 1, 1, s == cap + 1, s_base + before + rank);
            __syncthreads();
            if (threadIdx.x == 0) s_base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            __syncthreads();
        }
    }
}

// Tables of at most COMPACT_SMALL slots (the direct-mapped and small global
// tables) in one block of 1024 threads and one launch: thread t owns slots
// t + 1024 i and keeps their occupancy in a mask, thread 0 also the two side
// records; one reservation on plan->groups, then each thread writes its groups.
constexpr uint64_t COMPACT_SMALL = 4096;
__global__ __launch_bounds__(1024) void k_compact_small(DevTable T, const DevOut *__restrict__ op) {
    const DevOut &o = *op;
    const uint64_t cap = T.plan->cap;  // <= COMPACT_SMALL
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < COMPACT_SMALL / 1024; i++) {
        const uint64_t s = (uint64_t)i * 1024 + threadIdx.x;
        if (s < cap && T.rec[s * T.stride] != SENTINEL) m |= 1u << i;
    }
    if (threadIdx.x == 0)
        for (uint32_t j = 0; j < 2; j++)
            if (slot_occupied(T.rec + (cap + j) * T.stride, cap + j, cap)) m |= 1u << (16 + j);
    uint64_t row = block_reserve((uint32_t)__popc(m), &T.plan->groups);
    while (m) {
        const uint32_t i = (uint32_t)__ffs(m) - 1;
        m &= m - 1;
        const uint64_t s = i >= 16 ? cap + (i - 16) : (uint64_t)i * 1024 + threadIdx.x;
        const uint64_t *rec = T.rec + s * T.stride;
        write_group(o, rec[0], rec + 1, 1, s == cap + 1, row++);
    }
}

template <int KMAX>
__global__ __launch_bounds__(256) void k_topn_table(DevTable T, const DevTopn *__restrict__ tp) {
    const DevTopn &t = *tp;
    __shared__ uint32_t s_win[KMAX];
    const uint64_t cap = T.plan->cap, nrec = cap + 2;
    TopList<KMAX> L;
    L.init();
    const uint64_t per = (nrec + gridDim.x - 1) / gridDim.x;
    const uint64_t s0 = (uint64_t)blockIdx.x * per, s1 = min(nrec, s0 + per);
    const uint64_t lb = T.plan->topn_lb;  // groups below it cannot be in the top K
    for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const uint64_t *rec = T.rec + s * T.stride;
        if (!slot_occupied(rec, s, cap)) continue;
        uint64_t o, ti;
        group_order(t, rec[0], rec + 1, 1, s == cap + 1, o, ti);
        if (o >= lb) L.offer(o, ti, (uint32_t)s, t.k);
    }
    block_select<KMAX>(L, t.k, s_win);
    topn_push(t, T.plan, s_win, [&](uint32_t s, uint64_t &key, const uint64_t *&acc, uint32_t &stride, bool &nk) {
        const uint64_t *rec = T.rec + (uint64_t)s * T.stride;
        key = rec[0];
        acc = rec + 1;
        stride = 1;
        nk = s == cap + 1;
    });
}

// The final TopN over every producer's candidates, in two levels.
// Level 1: block b takes candidates [b * TOPN_L1, (b + 1) * TOPN_L1) and writes
// the indexes of its K best to lvl[b * KMAX ...] (INVALID_SLOT when fewer).
constexpr int TOPN_NI1 = 4;
constexpr uint32_t TOPN_L1 = 1024 * TOPN_NI1;
template <int KMAX>
__global__ __launch_bounds__(1024) void k_topn_level1(const DevTopn *__restrict__ tp, const DevPlan *__restrict__ plan,
                                                      uint32_t *lvl) {
    const DevTopn &t = *tp;
    __shared__ uint32_t s_win[KMAX];
    const uint32_t total = min(plan->cand_count, t.cand_cap);
    // no group below topn_lb (a lower bound of the K-th best value) can be in
    // the top K: such candidates count as empty
    const uint64_t lb = plan->topn_lb;
    uint64_t o[TOPN_NI1], ti[TOPN_NI1];
    uint32_t sl[TOPN_NI1];
#pragma unroll
    for (int j = 0; j < TOPN_NI1; j++) {
        const uint32_t i = blockIdx.x * TOPN_L1 + j * 1024 + threadIdx.x;
        o[j] = 0;
        ti[j] = 0;
        sl[j] = INVALID_SLOT;
        if (i < total && t.cand_null[i] != CAND_EMPTY) {
            const uint64_t *rec = t.cand_rec + (uint64_t)i * t.rec_words;
            group_order(t, rec[0], rec + 1, 1, t.cand_null[i] != 0, o[j], ti[j]);
            if (o[j] >= lb) sl[j] = i;
            else o[j] = ti[j] = 0;
        }
    }
    block_topk<KMAX, TOPN_NI1>(o, ti, sl, t.k, s_win);
    for (uint32_t r = threadIdx.x; r < KMAX; r += blockDim.x)
        lvl[blockIdx.x * KMAX + r] = r < t.k ? s_win[r] : INVALID_SLOT;
}

// Level 2: one block over the level-1 winners (nlvl blocks' lists); writes the
// result rows best first.
constexpr int TOPN_NI2 = 4;
template <int KMAX>
__global__ __launch_bounds__(1024) void k_topn_final(const DevTopn *__restrict__ tp, DevPlan *plan,
                                                     const DevOut *__restrict__ op, const uint32_t *lvl,
                                                     uint32_t nlvl) {
    const DevTopn &t = *tp;
    const DevOut &o = *op;
    __shared__ uint32_t s_win[KMAX];
    uint64_t ord[TOPN_NI2], tie[TOPN_NI2];
    uint32_t sl[TOPN_NI2];
#pragma unroll
    for (int j = 0; j < TOPN_NI2; j++) {
        const uint32_t e = j * 1024 + threadIdx.x;
        ord[j] = 0;
        tie[j] = 0;
        sl[j] = INVALID_SLOT;
        if (e < nlvl * KMAX && lvl[e] != INVALID_SLOT) {
            const uint32_t i = lvl[e];
            const uint64_t *rec = t.cand_rec + (uint64_t)i * t.rec_words;
            group_order(t, rec[0], rec + 1, 1, t.cand_null[i] != 0, ord[j], tie[j]);
            sl[j] = i;
        }
    }
    block_topk<KMAX, TOPN_NI2>(ord, tie, sl, t.k, s_win);
    for (uint32_t r = threadIdx.x; r < t.k; r += blockDim.x) {
        const uint32_t i = s_win[r];
        if (i == INVALID_SLOT) continue;
        const uint64_t *rec = t.cand_rec + (uint64_t)i * t.rec_words;
        write_group(o, rec[0], rec + 1, 1, t.cand_null[i] != 0, r);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t n = 0;
        for (uint32_t r = 0; r < t.k; r++) n += s_win[r] != INVALID_SLOT;
        plan->groups = n;
    }
}

}  // namespace gbagg
