// gbagg_radix.hpp -- the radix-partitioned strategy for high group counts.
//
//   k_sample            count the keys of a strided sample in a small table
//   k_heavy_hist/_threshold/_insert  keys frequent in the sample become the heavy-hitter
//                       table (a static, read-only LDS hash in the passes below)
//   k_rp_count          per workgroup and hash bucket: rows with a cold key;
//                       HyperLogLog sketch of all keys on the way
//   k_rp_scan_w/_p      exact partition offsets (partition p = top hash bits)
//   k_rp_scatter        fused filter + key hash; heavy keys aggregate in LDS
//                       atomics, cold rows are written to their partition at
//                       an exact offset (LDS cursor per partition)
//   k_rp_agg            one workgroup per partition: every group of the
//                       partition in an LDS hash table, then straight to the
//                       result (or its top-K to the TopN candidates)
//   k_heavy_fold/_emit  folds the per-workgroup heavy partial states
//
// Each group's final state is produced by exactly one workgroup, so the only
// device-scope atomics are one reservation per block and the rare rows that
// do not fit their LDS table (NULL keys, the SENTINEL key, LDS overflow),
// which merge into the small global fallback table.
#pragma once

#include "gbagg_kernels.hpp"

namespace gbagg {

constexpr int RP_BLOCK = 1024;   // count / scatter workgroup: 16 waves share one LDS table
constexpr int RA_BLOCK = 1024;   // per-partition aggregation workgroup
constexpr int RA_ROWS = 4;       // records per thread in flight

struct RadixParams {
    uint32_t W;             // workgroup slots over all batches
    uint32_t bucket_bits;   // count histogram: 1 << bucket_bits hash buckets
    uint32_t part_bits;     // partitions P = 1 << part_bits (<= bucket_bits)
    uint32_t *hist;         // [W][1 << bucket_bits]
    uint32_t *base;         // [W][P] rows of earlier workgroups in partition p
    uint32_t *ptotal;       // [P]
    uint64_t *pstart;       // [P + 1]
    uint64_t *recs;         // cold-row records, partition-major
    uint32_t rec_words;     // key + stored inputs (+ validity word)
    int32_t rec_pos[MAXU];  // record word of loaded slot u (>= 1), 0 = not stored
    uint32_t vmask_word;    // record word with the validity bits (0 = none)
    // heavy hitters
    uint64_t *hkeys;        // [hslots] open addressing, SENTINEL = empty
    uint32_t *hidx;         // [hslots] dense index of the key
    uint64_t *hdense;       // [hmax] key by dense index
    uint32_t hslots;
    uint32_t hmax;
    uint64_t *hpart;        // [W][nacc][hmax] per-workgroup partial states
    uint64_t *hfinal;       // [hmax][1 + nacc] folded states
    uint8_t *hll;           // [W][HLL_M]
    // sample
    uint64_t *skeys;        // [scap] keys, SENTINEL = empty
    uint32_t *scnt;         // [scap]
    uint32_t scap;
    uint32_t heavy_min;     // minimum sample count of a heavy key
};

// Everything a launch reads from device memory (kernel arguments would be
// spilled to scratch when indexed with a loop variable).
struct DevCtx {
    DevQuery q;
    DevOut o;
    DevTopn t;
    RadixParams rp;
};

__global__ void k_sample_init(const RadixParams *__restrict__ rpp) {
    const RadixParams &rp = *rpp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.scap; i += gridDim.x * blockDim.x) {
        rp.skeys[i] = SENTINEL;
        rp.scnt[i] = 0;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.hslots; i += gridDim.x * blockDim.x)
        rp.hkeys[i] = SENTINEL;
}

__global__ __launch_bounds__(256) void k_sample(const DevCtx *__restrict__ cx, const DevBatch *__restrict__ bp,
                                                uint32_t sample_n, uint64_t sample_stride, DevPlan *plan) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevBatch &b = *bp;
    const uint32_t mask = rp.scap - 1;
    uint32_t kept = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sample_n; i += gridDim.x * blockDim.x) {
        uint64_t r = (uint64_t)i * sample_stride;
        if (r >= b.nrows) continue;
        Row row;
        load_row(q, b, r, (1u << q.nkp) - 1, row);
        if (q.npred && !eval_preds(q, b, row)) continue;
        uint64_t w;
        if (make_key(q, b, row, w) != KEY_OK || w == SENTINEL) continue;
        kept++;
        uint32_t pos = (uint32_t)hash64(w) & mask;
        for (uint32_t probe = 0; probe < rp.scap; probe++) {
            uint64_t cur = __hip_atomic_load(&rp.skeys[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == SENTINEL) {
                uint64_t e = SENTINEL;
                __hip_atomic_compare_exchange_strong(&rp.skeys[pos], &e, w, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                cur = e == SENTINEL ? w : e;
            }
            if (cur == w) {
                atomicAdd(&rp.scnt[pos], 1u);
                break;
            }
            pos = (pos + 1) & mask;
        }
    }
    if (kept) atomicAdd((unsigned long long *)&plan->sample_rows, (unsigned long long)kept);
}

// Heavy-hitter selection over the sample table, in three small kernels:
//   k_heavy_hist       histogram of sample counts (counts above 63 share a bin)
//   k_heavy_threshold  smallest count >= heavy_min admitting at most hmax keys
//   k_heavy_insert     keys at or above it build the heavy table
__global__ __launch_bounds__(256) void k_heavy_hist(const RadixParams *__restrict__ rpp, DevPlan *plan,
                                                    uint32_t *ghist) {
    const RadixParams &rp = *rpp;
    __shared__ uint32_t hist[65];
    __shared__ unsigned long long s_distinct, s_f1;
    if (threadIdx.x < 65) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) { s_distinct = 0; s_f1 = 0; }
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.scap; i += gridDim.x * blockDim.x) {
        uint32_t c = rp.scnt[i];
        if (!c) continue;
        atomicAdd(&hist[min(c, 64u)], 1u);
        atomicAdd(&s_distinct, 1ULL);
        if (c == 1) atomicAdd(&s_f1, 1ULL);
    }
    __syncthreads();
    if (threadIdx.x < 65 && hist[threadIdx.x]) atomicAdd(&ghist[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_distinct) atomicAdd((unsigned long long *)&plan->sample_distinct, s_distinct);
        if (s_f1) atomicAdd((unsigned long long *)&plan->sample_f1, s_f1);
    }
}

__global__ void k_heavy_threshold(const RadixParams *__restrict__ rpp, DevPlan *plan, const uint32_t *ghist) {
    const RadixParams &rp = *rpp;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t t = 65, above = 0;
    for (uint32_t c = 64; c >= max(rp.heavy_min, 1u); c--) {
        if (above + ghist[c] > rp.hmax) break;
        above += ghist[c];
        t = c;
    }
    if (t == 65 && ghist[64] > 0) t = 64;  // more than hmax very frequent keys: take hmax of them
    plan->heavy_threshold = t;
    plan->heavy_count = 0;
}

__global__ __launch_bounds__(256) void k_heavy_insert(const RadixParams *__restrict__ rpp, DevPlan *plan) {
    const RadixParams &rp = *rpp;
    const uint32_t t = plan->heavy_threshold;
    const uint32_t hmask = rp.hslots - 1;
    if (t > 64) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.scap; i += gridDim.x * blockDim.x) {
        uint32_t c = rp.scnt[i];
        if (min(c, 64u) < t || c == 0) continue;
        uint32_t d = (uint32_t)atomicAdd((unsigned long long *)&plan->heavy_count, 1ULL);
        if (d >= rp.hmax) continue;
        uint64_t k = rp.skeys[i];
        atomicAdd((unsigned long long *)&plan->heavy_sample_rows, (unsigned long long)c);
        rp.hdense[d] = k;
        uint32_t pos = (uint32_t)hash64(k) & hmask;
        for (uint32_t probe = 0; probe < rp.hslots; probe++) {  // hslots >= 2 * hmax: finds a slot
            uint64_t e = SENTINEL;
            __hip_atomic_compare_exchange_strong(&rp.hkeys[pos], &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (e == SENTINEL) { rp.hidx[pos] = d; break; }
            pos = (pos + 1) & hmask;
        }
    }
}

// Static heavy table lookup in LDS: membership only.
__device__ __forceinline__ bool heavy_has(const uint64_t *hk, uint32_t hmask, uint64_t h, uint64_t k) {
    uint32_t pos = (uint32_t)h & hmask;
    for (;;) {
        uint64_t cur = hk[pos];
        if (cur == k) return true;
        if (cur == SENTINEL) return false;
        pos = (pos + 1) & hmask;
    }
}

// Static heavy table lookup in LDS: dense index or -1.
__device__ __forceinline__ int heavy_find(const uint64_t *hk, const uint32_t *hi, uint32_t hmask,
                                          uint64_t h, uint64_t k) {
    uint32_t pos = (uint32_t)h & hmask;
    for (;;) {
        uint64_t cur = hk[pos];
        if (cur == k) return (int)hi[pos];
        if (cur == SENTINEL) return -1;
        pos = (pos + 1) & hmask;
    }
}

template <class S>
__global__ __launch_bounds__(RP_BLOCK) void k_rp_count(const DevCtx *__restrict__ cx,
                                                       const DevBatch *__restrict__ bp, uint32_t wg_offset,
                                                       uint64_t rows_per_wg) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevBatch &b = *bp;
    const S sh(q);
    extern __shared__ uint64_t smem[];
    uint64_t *hk = smem;                                  // [hslots]
    uint32_t *cnt = (uint32_t *)(hk + rp.hslots);         // [1 << bucket_bits]
    uint32_t *reg = cnt + (1u << rp.bucket_bits);         // [HLL_M]
    const uint32_t nb = 1u << rp.bucket_bits;
    for (uint32_t i = threadIdx.x; i < rp.hslots; i += blockDim.x) hk[i] = rp.hkeys[i];
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) cnt[i] = 0;
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) reg[i] = 0;
    __syncthreads();
    const uint32_t hmask = rp.hslots - 1;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_wg;
    const uint64_t r1 = min(b.nrows, r0 + rows_per_wg);
    const uint32_t shift = 64 - rp.bucket_bits;
    for (uint64_t base = r0; base < r1; base += (uint64_t)RP_BLOCK * AGG_ROWS) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * RP_BLOCK + threadIdx.x;
            if (r < r1) load_row_s(sh, b, r, sh.nkp(), rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * RP_BLOCK + threadIdx.x;
            if (r >= r1) continue;
            if (sh.npred() && !eval_preds_s(sh, q, b, rows[j])) continue;
            uint64_t w;
            if (make_key_s(sh, q, b, rows[j], w) != KEY_OK || w == SENTINEL) continue;
            const uint64_t h = hash64(w);
            hll_add(reg, h);
            if (heavy_has(hk, hmask, h, w)) continue;
            __hip_atomic_fetch_add(&cnt[h >> shift], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    const uint32_t wg = wg_offset + blockIdx.x;
    uint32_t *out = rp.hist + (uint64_t)wg * nb;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) out[i] = cnt[i];
    uint8_t *hl = rp.hll + (uint64_t)wg * HLL_M;
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) hl[i] = (uint8_t)reg[i];
}

// Thread per partition: exclusive prefix over workgroups; partition totals.
__global__ void k_rp_scan_w(const RadixParams *__restrict__ rpp) {
    const RadixParams &rp = *rpp;
    const uint32_t P = 1u << rp.part_bits;
    const uint32_t shift = rp.bucket_bits - rp.part_bits;
    const uint32_t nb = 1u << rp.bucket_bits;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < rp.W; w++) {
            const uint32_t *h = rp.hist + (uint64_t)w * nb + ((uint64_t)p << shift);
            uint32_t c = 0;
            for (uint32_t j = 0; j < (1u << shift); j++) c += h[j];
            rp.base[(uint64_t)w * P + p] = run;
            run += c;
        }
        rp.ptotal[p] = run;
    }
}

// One block of 1024: exclusive scan of the partition totals.
__global__ __launch_bounds__(1024) void k_rp_scan_p(const RadixParams *__restrict__ rpp, DevPlan *plan) {
    const RadixParams &rp = *rpp;
    __shared__ uint64_t s_part[1024];
    const uint32_t P = 1u << rp.part_bits;
    const uint32_t per = (P + blockDim.x - 1) / blockDim.x;
    const uint32_t p0 = threadIdx.x * per, p1 = min(P, p0 + per);
    uint64_t local = 0;
    for (uint32_t p = p0; p < p1; p++) local += rp.ptotal[p];
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = threadIdx.x >= (uint32_t)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t p = p0; p < p1; p++) {
        rp.pstart[p] = run;
        run += rp.ptotal[p];
    }
    if (threadIdx.x == blockDim.x - 1) {
        rp.pstart[P] = s_part[blockDim.x - 1];
        plan->cold_rows = s_part[blockDim.x - 1];
    }
}

// Record word of slot u: compile-time for static shapes.
template <class S>
__device__ __forceinline__ int rec_pos_s(const RadixParams &rp, int u) {
    if constexpr (S::kStatic) return S::rec_pos(u);
    else return rp.rec_pos[u];
}

template <class S>
__device__ __forceinline__ int vmask_word_s(const RadixParams &rp) {
    if constexpr (S::kStatic) return S::any_nullable_input() ? 1 + S::rec_inputs() : 0;
    else return (int)rp.vmask_word;
}

template <class S>
__global__ __launch_bounds__(RP_BLOCK) void k_rp_scatter(const DevCtx *__restrict__ cx,
                                                         const DevBatch *__restrict__ bp, DevTable T,
                                                         uint32_t wg_offset, uint64_t rows_per_wg) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevBatch &b = *bp;
    const S sh(q);
    extern __shared__ uint64_t smem[];
    const uint32_t P = 1u << rp.part_bits;
    uint64_t *hk = smem;                                      // [hslots]
    uint64_t *hacc = hk + rp.hslots;                          // [nacc][hmax]
    uint32_t *hi = (uint32_t *)(hacc + (uint64_t)q.nacc * rp.hmax);  // [hslots]
    uint32_t *cur = hi + rp.hslots;                           // [P]
    const uint32_t wg = wg_offset + blockIdx.x;
    for (uint32_t i = threadIdx.x; i < rp.hslots; i += blockDim.x) {
        hk[i] = rp.hkeys[i];
        hi[i] = rp.hidx[i];
    }
    for (int a = 0; a < q.nacc; a++) {
        uint64_t id = acc_identity(q.acc_op[a]);
        for (uint32_t i = threadIdx.x; i < rp.hmax; i += blockDim.x) hacc[(uint64_t)a * rp.hmax + i] = id;
    }
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x)
        cur[p] = (uint32_t)rp.pstart[p] + rp.base[(uint64_t)wg * P + p];
    __syncthreads();
    const uint32_t hmask = rp.hslots - 1;
    const uint64_t cap = T.plan->cap;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_wg;
    const uint64_t r1 = min(b.nrows, r0 + rows_per_wg);
    const uint32_t shift = 64 - rp.part_bits;
    const uint32_t rw = rp.rec_words;
    const int vw = vmask_word_s<S>(rp);
    for (uint64_t base = r0; base < r1; base += (uint64_t)RP_BLOCK * AGG_ROWS) {
        Row rows[AGG_ROWS];
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * RP_BLOCK + threadIdx.x;
            if (r < r1) load_row_s(sh, b, r, sh.nu(), rows[j]);
        }
#pragma unroll
        for (int j = 0; j < AGG_ROWS; j++) {
            uint64_t r = base + (uint64_t)j * RP_BLOCK + threadIdx.x;
            if (r >= r1) continue;
            if (sh.npred() && !eval_preds_s(sh, q, b, rows[j])) continue;
            uint64_t w;
            int kr = make_key_s(sh, q, b, rows[j], w);
            if (kr == KEY_DROP) continue;
            if (side_row(q, T, cap, kr, w, rows[j])) continue;
            const uint64_t h = hash64(w);
            int d = heavy_find(hk, hi, hmask, h, w);
            if (d >= 0) {
                row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, rows[j], hacc + d, rp.hmax);
                continue;
            }
            const uint32_t p = P > 1 ? (uint32_t)(h >> shift) : 0;
            const uint32_t pos = __hip_atomic_fetch_add(&cur[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            uint64_t *rec = rp.recs + (uint64_t)pos * rw;
            rec[0] = w;
#pragma unroll
            for (int u = 0; u < MAXU; u++) {
                const int rpu = rec_pos_s<S>(rp, u);
                if (rpu) rec[rpu] = rows[j].v[u];
            }
            if (vw) rec[vw] = rows[j].vm;
        }
    }
    __syncthreads();
    uint64_t *out = rp.hpart + (uint64_t)wg * q.nacc * rp.hmax;
    for (uint64_t i = threadIdx.x; i < (uint64_t)q.nacc * rp.hmax; i += blockDim.x) out[i] = hacc[i];
}

// One workgroup per partition: all of the partition's groups in LDS.
template <class SH, int KMAX>
__global__ __launch_bounds__(RA_BLOCK) void k_rp_agg(const DevCtx *__restrict__ cx, DevTable T,
                                                     uint32_t S, uint32_t limit) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevOut &o = cx->o;
    const DevTopn &t = cx->t;
    const SH sh(q);
    const uint32_t rw = rp.rec_words;
    const int vw = vmask_word_s<SH>(rp);
    extern __shared__ uint64_t smem[];
    LdsTable lt = lds_table_setup(smem, q, S, limit);
    __syncthreads();
    const uint32_t p = blockIdx.x;
    const uint64_t r0 = rp.pstart[p], r1 = rp.pstart[p + 1];
    const uint64_t cap = T.plan->cap, mask = cap - 1;
    for (uint64_t base = r0; base < r1; base += (uint64_t)RA_BLOCK * RA_ROWS) {
        Row rows[RA_ROWS];
        uint64_t keys[RA_ROWS];
#pragma unroll
        for (int j = 0; j < RA_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * RA_BLOCK + threadIdx.x;
            if (r >= r1) continue;
            const uint64_t *rec = rp.recs + r * rw;
            keys[j] = rec[0];
            rows[j].vm = vw ? (uint32_t)rec[vw] : 0xFFFFFFFFu;
#pragma unroll
            for (int u = 0; u < MAXU; u++) {
                const int rpu = rec_pos_s<SH>(rp, u);
                rows[j].v[u] = rpu ? rec[rpu] : 0;
            }
        }
#pragma unroll
        for (int j = 0; j < RA_ROWS; j++) {
            const uint64_t r = base + (uint64_t)j * RA_BLOCK + threadIdx.x;
            if (r >= r1) continue;
            const uint64_t w = keys[j];
            const uint64_t h = hash64(w);
            int slot = lds_find(lt, h, w);
            if (slot >= 0) {
                row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, rows[j], lt.acc + slot, lt.S);
            } else {
                int64_t g = g_find(T, mask, h, w);
                if (g >= 0) row_update_s<__HIP_MEMORY_SCOPE_AGENT>(sh, rows[j], T.rec + (uint64_t)g * T.stride + 1, 1);
            }
        }
    }
    __syncthreads();
    if (!t.enabled) {
        uint32_t mine = 0;
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) mine += lt.key[i] != SENTINEL;
        uint64_t row = block_reserve(mine, &T.plan->groups);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x)
            if (lt.key[i] != SENTINEL) write_group(o, lt.key[i], lt.acc + i, S, false, row++);
        return;
    }
    __shared__ uint32_t s_win[KMAX];
    TopList<KMAX> L;
    L.init();
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        if (lt.key[i] == SENTINEL) continue;
        uint64_t ord, tie;
        group_order(t, lt.key[i], lt.acc + i, S, false, ord, tie);
        L.offer(ord, tie, i, t.k);
    }
    block_select<KMAX>(L, t.k, s_win);
    topn_push(t, T.plan, s_win, [&](uint32_t i, uint64_t &key, const uint64_t *&acc, uint32_t &stride, bool &nk) {
        key = lt.key[i];
        acc = lt.acc + i;
        stride = S;
        nk = false;
    });
}

// Folds the heavy partial states over workgroups: block = 64 heavy keys x 4
// workgroup phases, one accumulator per blockIdx.y. The fold order is fixed,
// so double sums are reproducible.
__device__ __forceinline__ uint64_t fold_acc(int op, uint64_t x, uint64_t y) {
    switch (op) {
    case ACC_SUM_F:
        return (uint64_t)__double_as_longlong(__longlong_as_double((long long)x) +
                                              __longlong_as_double((long long)y));
    case ACC_MIN_I: return (int64_t)y < (int64_t)x ? y : x;
    case ACC_MAX_I: return (int64_t)y > (int64_t)x ? y : x;
    case ACC_MIN_F: return y < x ? y : x;
    case ACC_MAX_F: return y > x ? y : x;
    default: return x + y;
    }
}

__global__ __launch_bounds__(256) void k_heavy_fold(const DevCtx *__restrict__ cx, DevTable T) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    __shared__ uint64_t s_part[4][64];
    const uint32_t hc = (uint32_t)min<uint64_t>(T.plan->heavy_count, rp.hmax);
    const uint32_t dl = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * 64 + dl;
    const int a = blockIdx.y;
    const int op = q.acc_op[a];
    uint64_t x = acc_identity(op);
    if (d < hc) {
        const uint64_t stride = (uint64_t)q.nacc * rp.hmax;
        const uint64_t *src = rp.hpart + (uint64_t)a * rp.hmax + d;
        uint32_t w = ph;
        for (; w + 12 < rp.W; w += 16) {
            uint64_t y0 = src[(uint64_t)w * stride], y1 = src[(uint64_t)(w + 4) * stride];
            uint64_t y2 = src[(uint64_t)(w + 8) * stride], y3 = src[(uint64_t)(w + 12) * stride];
            x = fold_acc(op, x, y0);
            x = fold_acc(op, x, y1);
            x = fold_acc(op, x, y2);
            x = fold_acc(op, x, y3);
        }
        for (; w < rp.W; w += 4) x = fold_acc(op, x, src[(uint64_t)w * stride]);
    }
    s_part[ph][dl] = x;
    __syncthreads();
    if (ph == 0 && d < hc) {
        uint64_t r = s_part[0][dl];
        for (int k = 1; k < 4; k++) r = fold_acc(op, r, s_part[k][dl]);
        uint64_t *dst = rp.hfinal + (uint64_t)d * (1 + q.nacc);
        dst[1 + a] = r;
        if (a == 0) dst[0] = rp.hdense[d];
    }
}

// Emits the folded heavy groups (full result or TopN candidates).
template <int KMAX>
__global__ __launch_bounds__(256) void k_heavy_emit(const DevCtx *__restrict__ cx, DevTable T) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevOut &o = cx->o;
    const DevTopn &t = cx->t;
    const uint32_t hc = (uint32_t)min<uint64_t>(T.plan->heavy_count, rp.hmax);
    const uint32_t stride = 1 + q.nacc;
    if (!t.enabled) {
        for (uint32_t d0 = blockIdx.x * blockDim.x; d0 < hc; d0 += gridDim.x * blockDim.x) {
            uint32_t d = d0 + threadIdx.x;
            const uint64_t *rec = rp.hfinal + (uint64_t)d * stride;
            bool have = d < hc && rec[1] != 0;
            uint64_t row = block_reserve(have ? 1u : 0u, &T.plan->groups);
            if (have) write_group(o, rec[0], rec + 1, 1, false, row);
        }
        return;
    }
    __shared__ uint32_t s_win[KMAX];
    TopList<KMAX> L;
    L.init();
    for (uint32_t d = threadIdx.x; d < hc; d += blockDim.x) {
        const uint64_t *rec = rp.hfinal + (uint64_t)d * stride;
        if (rec[1] == 0) continue;
        uint64_t ord, tie;
        group_order(t, rec[0], rec + 1, 1, false, ord, tie);
        L.offer(ord, tie, d, t.k);
    }
    block_select<KMAX>(L, t.k, s_win);
    topn_push(t, T.plan, s_win, [&](uint32_t d, uint64_t &key, const uint64_t *&acc, uint32_t &st, bool &nk) {
        key = rp.hfinal[(uint64_t)d * stride];
        acc = rp.hfinal + (uint64_t)d * stride + 1;
        st = 1;
        nk = false;
    });
}

}  // namespace gbagg
