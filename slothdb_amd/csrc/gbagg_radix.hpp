// gbagg_radix.hpp -- the radix-partitioned strategy for high group counts.
//
//   k_sample            count the keys of a strided sample in a small table
//   k_heavy_hist/_threshold/_insert  keys frequent in the sample become the heavy-hitter
//                       table (a static, read-only LDS hash in the passes below)
//   k_rp_count          per workgroup and hash bucket: rows with a cold key;
//                       HyperLogLog sketch of the cold keys on the way
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
constexpr uint32_t RA_SMAX = 8192;  // largest per-partition LDS table
// TopN partitions hold at most RA_TOPN_SMAX slots, so each thread ranks at most
// RA_TOPN_NI of them in registers and two workgroups fit a CU's registers
constexpr uint32_t RA_TOPN_SMAX = 4096;
constexpr int RA_TOPN_NI = (RA_TOPN_SMAX + RA_BLOCK - 1) / RA_BLOCK;
// Scatter workgroup roles: the first SC_PROD threads load and classify rows;
// the other SC_CONS threads only write cold records to HBM, taking them from
// an LDS ring. The vector-memory counter drains in issue order for loads and
// stores alike, so stores issued by the loading waves would hold up their
// later loads until the stores are acknowledged.
constexpr int SC_PROD = 768;
constexpr int SC_CONS = RP_BLOCK - SC_PROD;
constexpr uint32_t SC_SPIN_MAX = 1u << 24;  // bounded waits: an error, never a hang
constexpr uint32_t OVF_RING = 4u;           // DevPlan::overflow bit: ring wait bound hit

struct RadixParams {
    uint32_t W;             // count-pass workgroup slots over all batches (histogram rows)
    uint32_t W3;            // scatter-pass workgroup slots (heavy partial states)
    uint32_t g1;            // count workgroups per scatter workgroup
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
    uint64_t *hpart;        // [W3][nacc][hmax] per-scatter-workgroup partial states
    uint64_t *hfinal;       // [hmax][1 + nacc] folded states
    uint8_t *hll;           // [W][HLL_M]
    // sample
    uint64_t *skeys;        // [scap] keys, SENTINEL = empty
    uint32_t *scnt;         // [scap]
    uint32_t scap;
    uint32_t heavy_min;     // minimum sample count of a heavy key
    uint32_t ring;          // scatter LDS ring entries (power of two >= 256)
    uint32_t msplit_max;    // largest fanout whose scatter offsets come from the wave multisplit
};

// Everything a launch reads from device memory (kernel arguments would be
// spilled to scratch when indexed with a loop variable).
struct DevCtx {
    DevQuery q;
    DevOut o;
    DevTopn t;
    RadixParams rp;
};

// Clears the sample table, the heavy table and the count histogram of the
// heavy-key selection (ghist, 65 bins), and presets the sampled key range.
__global__ void k_sample_init(const RadixParams *__restrict__ rpp, DevPlan *plan, uint32_t *ghist) {
    const RadixParams &rp = *rpp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.scap; i += gridDim.x * blockDim.x) {
        rp.skeys[i] = SENTINEL;
        rp.scnt[i] = 0;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.hslots; i += gridDim.x * blockDim.x)
        rp.hkeys[i] = SENTINEL;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 65) ghist[threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            plan->key_min = INT64_MAX;
            plan->key_max = INT64_MIN;
            plan->rng_min = INT64_MAX;
            plan->rng_max = INT64_MIN;
            plan->heavy_count = 0;
        }
    }
}

// Sample keys are first counted in a small per-block LDS table, so a hot key
// costs one global atomic per block rather than one per sampled row.
__device__ __forceinline__ void sample_global_add(const RadixParams &rp, uint64_t w, uint32_t c) {
    const uint32_t mask = rp.scap - 1;
    uint32_t pos = (uint32_t)hash64(w) & mask;
    for (uint32_t probe = 0; probe < rp.scap; probe++) {
        uint64_t cur = __hip_atomic_load(&rp.skeys[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == SENTINEL) {
            uint64_t e = SENTINEL;
            __hip_atomic_compare_exchange_strong(&rp.skeys[pos], &e, w, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            cur = e == SENTINEL ? w : e;
        }
        if (cur == w) {
            atomicAdd(&rp.scnt[pos], c);
            return;
        }
        pos = (pos + 1) & mask;
    }
}

constexpr uint32_t SAMPLE_LDS = 1024;

// Sample rows i * sample_stride of the batches' concatenation.
__global__ __launch_bounds__(256) void k_sample(const DevCtx *__restrict__ cx, BatchRows br, uint32_t sample_n,
                                                uint64_t sample_stride, DevPlan *plan) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const uint64_t total = br.start[br.nb];
    __shared__ uint64_t lk[SAMPLE_LDS];
    __shared__ uint32_t lc[SAMPLE_LDS];
    __shared__ uint32_t s_fx[MAXU];
    for (uint32_t i = threadIdx.x; i < SAMPLE_LDS; i += blockDim.x) { lk[i] = SENTINEL; lc[i] = 0; }
    if (threadIdx.x < MAXU) s_fx[threadIdx.x] = 0;
    const uint32_t fxm = fx_slots(q);
    __syncthreads();
    uint32_t kept = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sample_n; i += gridDim.x * blockDim.x) {
        uint64_t r = (uint64_t)i * sample_stride;
        if (r >= total) continue;
        const uint32_t bi = batch_of(br, r);
        const DevBatch &b = br.b[bi];
        r -= br.start[bi];
        Row row;
        load_row(q, b, r, ((1u << q.nkp) - 1) | fxm, row);
        if (q.npred && !eval_preds(q, b, row)) continue;
        uint64_t w;
        if (make_key(q, b, row, w) != KEY_OK || w == SENTINEL) continue;
        kept++;
        for (int u = 0; u < MAXU; u++)  // largest exponents of the double SUM inputs
            if (((fxm & row.vm) >> u) & 1) {
                const uint32_t e = fx_exp_of(sel(row, u));
                if (e) __hip_atomic_fetch_max(&s_fx[u], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        uint32_t pos = (uint32_t)hash64(w) & (SAMPLE_LDS - 1);
        bool done = false;
        for (int probe = 0; probe < 16 && !done; probe++) {
            uint64_t cur = __hip_atomic_load(&lk[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == SENTINEL) {
                uint64_t e = SENTINEL;
                __hip_atomic_compare_exchange_strong(&lk[pos], &e, w, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                cur = e == SENTINEL ? w : e;
            }
            if (cur == w) {
                atomicAdd(&lc[pos], 1u);
                done = true;
            }
            pos = (pos + 1) & (SAMPLE_LDS - 1);
        }
        if (!done) sample_global_add(rp, w, 1);
    }
    // sampled rows: wave sums, then one device-scope add per block
    __shared__ uint32_t s_kept;
    if (threadIdx.x == 0) s_kept = 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_xor(kept, off);
    __shared__ uint32_t s_nocc;
    __shared__ uint16_t s_occ[SAMPLE_LDS];
    if (threadIdx.x == 0) s_nocc = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && kept) atomicAdd(&s_kept, kept);
    // occupied slots first gathered into a dense list (ballot ranks, one LDS
    // atomic per wave), so the device-scope adds run in as few rounds as the
    // block has distinct keys per thread, not SAMPLE_LDS / blockDim.x rounds
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t i0 = 0; i0 < SAMPLE_LDS; i0 += blockDim.x) {  // uniform
        const uint32_t i = i0 + threadIdx.x;
        const bool occ = i < SAMPLE_LDS && lc[i] != 0;
        const uint64_t m = __ballot(occ);
        uint32_t base = 0;
        if (lane == 0 && m)
            base = __hip_atomic_fetch_add(&s_nocc, (uint32_t)__popcll(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        base = __shfl(base, 0);
        if (occ) s_occ[base + __popcll(m & ((1ULL << lane) - 1))] = (uint16_t)i;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < s_nocc; j += blockDim.x) sample_global_add(rp, lk[s_occ[j]], lc[s_occ[j]]);
    __syncthreads();
    if (threadIdx.x == 0 && s_kept) atomicAdd((unsigned long long *)&plan->sample_rows, (unsigned long long)s_kept);
    if (threadIdx.x < MAXU && s_fx[threadIdx.x])
        __hip_atomic_fetch_max(&plan->fx_exp[threadIdx.x], s_fx[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Heavy-hitter selection over the sample table, in three small kernels:
//   k_heavy_hist       histogram of sample counts (counts above 63 share a bin)
//   k_heavy_threshold  smallest count >= heavy_min admitting at most hmax keys
//   k_heavy_insert     keys at or above it build the heavy table
__global__ __launch_bounds__(256) void k_heavy_hist(const RadixParams *__restrict__ rpp, DevPlan *plan,
                                                    uint32_t *ghist) {
    const RadixParams &rp = *rpp;
    __shared__ uint32_t hist[65];
    __shared__ unsigned long long s_distinct, s_f1, s_f2, s_f3, s_max, s_ent;
    __shared__ unsigned long long s_h16[16];
    __shared__ int64_t s_lo, s_hi;
    if (threadIdx.x < 65) hist[threadIdx.x] = 0;
    if (threadIdx.x < 16) s_h16[threadIdx.x] = 0;
    if (threadIdx.x == 0) { s_distinct = 0; s_f1 = 0; s_f2 = 0; s_f3 = 0; s_max = 0; s_ent = 0; s_lo = INT64_MAX; s_hi = INT64_MIN; }
    __syncthreads();
    int64_t lo = INT64_MAX, hi = INT64_MIN;  // sampled key word range (plan->key_min/max preset)
    unsigned long long dist = 0, f1 = 0, f2 = 0, f3 = 0, mx = 0, ent = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rp.scap; i += gridDim.x * blockDim.x) {
        uint32_t c = rp.scnt[i];
        if (!c) continue;
        dist++;
        // singletons (most keys) only count locally: bins 1 of both histograms
        // take the block's singleton count at the end
        if (c == 1) f1++;
        else {
            atomicAdd(&hist[min(c, 64u)], 1u);
            atomicAdd(&s_h16[min(15u, 31u - (uint32_t)__clz(c))], 1ULL);
        }
        f2 += c == 2;
        f3 += c == 3;
        mx = mx > c ? mx : c;
        // c * log2(c) in 2^-20 units: integer sums, so the features do not
        // depend on the order of the additions
        ent += (unsigned long long)((double)c * log2((double)c) * 1048576.0 + 0.5);
        lo = min(lo, (int64_t)rp.skeys[i]);
        hi = max(hi, (int64_t)rp.skeys[i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        dist += __shfl_xor(dist, off);
        f1 += __shfl_xor(f1, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (dist) atomicAdd(&s_distinct, dist);
        if (f1) {
            atomicAdd(&s_f1, f1);
            atomicAdd(&hist[1], (uint32_t)f1);
            atomicAdd(&s_h16[0], f1);
        }
    }
    if (f2) atomicAdd(&s_f2, f2);
    if (f3) atomicAdd(&s_f3, f3);
    if (mx) atomicMax(&s_max, mx);
    if (ent) atomicAdd(&s_ent, ent);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (int64_t)__shfl_xor(lo, off));
        hi = max(hi, (int64_t)__shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0 && lo <= hi) {
        __hip_atomic_fetch_min(&s_lo, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_max(&s_hi, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (threadIdx.x < 65 && hist[threadIdx.x]) atomicAdd(&ghist[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x < 16 && s_h16[threadIdx.x])
        atomicAdd((unsigned long long *)&plan->sample_hist[threadIdx.x], s_h16[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (s_distinct) atomicAdd((unsigned long long *)&plan->sample_distinct, s_distinct);
        if (s_f1) atomicAdd((unsigned long long *)&plan->sample_f1, s_f1);
        if (s_f2) atomicAdd((unsigned long long *)&plan->sample_f2, s_f2);
        if (s_f3) atomicAdd((unsigned long long *)&plan->sample_f3, s_f3);
        if (s_max) atomicMax((unsigned long long *)&plan->sample_max, s_max);
        if (s_ent) atomicAdd((unsigned long long *)&plan->sample_ent, s_ent);
        if (s_lo <= s_hi) {  // one device-scope update per block
            __hip_atomic_fetch_min(&plan->key_min, s_lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&plan->key_max, s_hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Smallest sample count >= heavy_min that admits at most hmax keys (65:
// none). Every k_heavy_insert block computes it from the 65-bin histogram.
__device__ __forceinline__ uint32_t heavy_threshold(const RadixParams &rp, const uint32_t *ghist) {
    uint32_t t = 65, above = 0;
    for (uint32_t c = 64; c >= max(rp.heavy_min, 1u); c--) {
        if (above + ghist[c] > rp.hmax) break;
        above += ghist[c];
        t = c;
    }
    if (t == 65 && ghist[64] > 0) t = 64;  // more than hmax very frequent keys: take hmax of them
    return t;
}

// Dense indexes come from one device-scope atomic per wave and round (ballot
// rank within the wave): same-address device atomics serialise.
// heavy_count starts at zero (k_sample_init). The heavy table serves only the
// radix strategy: for a sample of few distinct keys, which the default plan
// does not aggregate by radix, it is left empty unless radix is forced (a
// radix plan then runs without heavy keys, exactly, if the selector picks it).
__global__ __launch_bounds__(256) void k_heavy_insert(const RadixParams *__restrict__ rpp, DevPlan *plan,
                                                      const uint32_t *ghist, uint32_t force) {
    const RadixParams &rp = *rpp;
    // the default plan's rule for few distinct keys (also on the host): a count
    // pass launched before the plan is read runs only without it
    const uint64_t sd = plan->sample_distinct, sf1 = plan->sample_f1;
    const bool gate = !(sd <= 4096 && sf1 * 8 <= sd + 8);
    if (blockIdx.x == 0 && threadIdx.x == 0) plan->radix_gate = gate;
    __shared__ uint32_t s_t, s_hist[65];
    if (threadIdx.x < 65) s_hist[threadIdx.x] = ghist[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        s_t = heavy_threshold(rp, s_hist);
        if (blockIdx.x == 0) plan->heavy_threshold = s_t;
    }
    __syncthreads();
    const uint32_t t = s_t;
    const uint32_t hmask = rp.hslots - 1;
    if (t > 64 || !(gate || force)) return;
    const uint32_t lane = threadIdx.x & 63;
    // The block owns a contiguous range of sample slots. Pass 1 counts its
    // heavy keys, one device-scope atomic then reserves their dense indexes
    // (same-address device atomics serialise: one per block, not per wave),
    // pass 2 hands them out by wave ballot ranks.
    const uint32_t per = (rp.scap + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = min(rp.scap, blockIdx.x * per), r1 = min(rp.scap, r0 + per);
    __shared__ uint32_t s_cnt, s_base, s_run;
    __shared__ unsigned long long s_rows;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_run = 0;
        s_rows = 0;
    }
    __syncthreads();
    auto heavy_at = [&](uint32_t i, uint32_t &c) {
        c = i < r1 ? rp.scnt[i] : 0;
        return c != 0 && min(c, 64u) >= t;
    };
    for (uint32_t i0 = r0; i0 < r1; i0 += blockDim.x) {  // uniform
        uint32_t c;
        const uint64_t m = __ballot(heavy_at(i0 + threadIdx.x, c));
        if (lane == 0 && m) atomicAdd(&s_cnt, (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x == 0)
        s_base = s_cnt ? (uint32_t)atomicAdd((unsigned long long *)&plan->heavy_count, (unsigned long long)s_cnt) : 0;
    __syncthreads();
    unsigned long long rows = 0;
    for (uint32_t i0 = r0; i0 < r1; i0 += blockDim.x) {  // uniform
        const uint32_t i = i0 + threadIdx.x;
        uint32_t c;
        const bool heavy = heavy_at(i, c);
        const uint64_t m = __ballot(heavy);
        if (!m) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&s_run, (uint32_t)__popcll(m));
        base = __shfl(base, 0);
        const uint32_t d = s_base + base + (uint32_t)__popcll(m & ((1ULL << lane) - 1));
        const bool kept = heavy && d < rp.hmax;  // the table holds hmax keys
        if (!kept) continue;
        rows += c;
        uint64_t k = rp.skeys[i];
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
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rows += __shfl_xor(rows, off);
    if (lane == 0 && rows) atomicAdd(&s_rows, rows);
    __syncthreads();
    if (threadIdx.x == 0 && s_rows) atomicAdd((unsigned long long *)&plan->heavy_sample_rows, s_rows);
}

// Static heavy table lookup in LDS: membership only.
// Runs f(row) over rows [r0, r1) of a batch, R rows per thread per tile of
// BLOCK * R rows. Tile k+1's loads are issued before tile k's work, into the
// other of two register buffers (no copy between them), so a row's loads are
// consumed where the row is processed and never wait behind stores issued for
// earlier rows: the vector-memory counter is shared by loads and stores and
// drains in issue order. Row indexes are clamped to the last row, so every
// load is unconditional and in bounds.
template <int R, int BLOCK, class S, class F>
__device__ __forceinline__ void for_rows_pipelined(const S &sh, const DevBatch &b, uint64_t r0, uint64_t r1,
                                                   int nload, uint32_t tid, F &&f) {
    if (r0 >= r1) return;
    const uint64_t last = r1 - 1, step = (uint64_t)BLOCK * R;
    Row A[R], B[R];
    auto load = [&](uint64_t base, Row(&dst)[R]) {
#pragma unroll
        for (int j = 0; j < R; j++)
            load_row_s(sh, b, min(base + (uint64_t)j * BLOCK + tid, last), nload, dst[j]);
    };
    auto run = [&](uint64_t base, const Row(&src)[R]) {
#pragma unroll
        for (int j = 0; j < R; j++)
            if (base + (uint64_t)j * BLOCK + tid < r1) f(src[j]);
    };
    load(r0, A);
    for (uint64_t base = r0; base < r1; base += 2 * step) {
        load(base + step, B);
        run(base, A);
        load(base + 2 * step, A);
        run(base + step, B);
    }
}

// Linear probe of the heavy table in LDS with a single loop exit: the slot
// of k, or of the first empty slot. (Separate "found" and "empty" exits make
// the compiler keep two exit masks per iteration, several scalar
// instructions each.)
__device__ __forceinline__ uint32_t heavy_probe(const uint64_t *hk, uint32_t hmask, uint64_t h, uint64_t k,
                                                bool &found) {
    uint32_t pos = (uint32_t)h & hmask;
    uint64_t cur = hk[pos];
    while (cur != k && cur != SENTINEL) {
        pos = (pos + 1) & hmask;
        cur = hk[pos];
    }
    found = cur == k;
    return pos;
}

__device__ __forceinline__ bool heavy_has(const uint64_t *hk, uint32_t hmask, uint64_t h, uint64_t k) {
    bool found;
    heavy_probe(hk, hmask, h, k, found);
    return found;
}

// Static heavy table lookup in LDS: dense index or -1.
__device__ __forceinline__ int heavy_find(const uint64_t *hk, const uint32_t *hi, uint32_t hmask,
                                          uint64_t h, uint64_t k) {
    bool found;
    const uint32_t pos = heavy_probe(hk, hmask, h, k, found);
    return found ? (int)hi[pos] : -1;
}

// gate: null, or a word that must be non-zero for the pass to run (a count
// launched ahead of the plan decision, see DevPlan::radix_gate).
template <class S>
__global__ __launch_bounds__(RP_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_rp_count(const DevCtx *__restrict__ cx,
                                                                                                BatchRows br,
                                                                                                const uint32_t *gate) {
    if (gate && *gate == 0) return;
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
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
    const uint32_t shift = 64 - rp.bucket_bits;
    for_batch_rows(br, [&](const DevBatch &bb, uint64_t r0, uint64_t r1) {
        const DevBatch b = bb;  // column pointers stay in registers
        auto count_row = [&](const Row &row) {
            if (sh.npred() && !eval_preds_s(sh, q, b, row)) return;
            uint64_t w;
            if (make_key_s(sh, q, b, row, w) != KEY_OK || w == SENTINEL) return;
            const uint64_t h = hash64(w);
            if (heavy_has(hk, hmask, h, w)) return;
            hll_add(reg, h);  // the sketch counts the cold keys (the heavy ones are known)
            __hip_atomic_fetch_add(&cnt[h >> shift], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        for_rows_pipelined<S::kStatic ? AGG_ROWS : 1, RP_BLOCK>(sh, b, r0, r1, sh.nkp(), threadIdx.x, count_row);
    });
    __syncthreads();
    const uint32_t wg = blockIdx.x;
    uint32_t *out = rp.hist + (uint64_t)wg * nb;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) out[i] = cnt[i];
    uint8_t *hl = rp.hll + (uint64_t)wg * HLL_M;
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) hl[i] = (uint8_t)reg[i];
}

// Exclusive prefix over count workgroups per partition, and partition
// totals. Block = 64 partitions x 16 chunks of the workgroup range: chunk
// sums, then a prefix over the chunks in LDS, then each chunk writes its
// offsets.
constexpr int SCAN_PARTS = 64, SCAN_CHUNKS = 16;
__global__ __launch_bounds__(SCAN_PARTS * SCAN_CHUNKS) void k_rp_scan_w(const RadixParams *__restrict__ rpp) {
    const RadixParams &rp = *rpp;
    __shared__ uint32_t s_sum[SCAN_CHUNKS][SCAN_PARTS];
    const uint32_t P = 1u << rp.part_bits;
    const uint32_t shift = rp.bucket_bits - rp.part_bits;
    const uint32_t nb = 1u << rp.bucket_bits;
    const uint32_t pl = threadIdx.x % SCAN_PARTS, c = threadIdx.x / SCAN_PARTS;
    const uint32_t p = blockIdx.x * SCAN_PARTS + pl;
    const uint32_t per = (rp.W + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
    const uint32_t w0 = min(rp.W, c * per), w1 = min(rp.W, w0 + per);
    auto part_count = [&](uint32_t w) {
        const uint32_t *h = rp.hist + (uint64_t)w * nb + ((uint64_t)p << shift);
        if (shift == 0) return h[0];
        uint32_t x = 0;
        for (uint32_t j = 0; j < (1u << shift); j++) x += h[j];
        return x;
    };
    uint32_t sum = 0;
    if (p < P) {
#pragma unroll 8
        for (uint32_t w = w0; w < w1; w++) sum += part_count(w);
    }
    s_sum[c][pl] = sum;
    __syncthreads();
    if (p >= P) return;
    uint32_t run = 0;
    for (uint32_t k = 0; k < c; k++) run += s_sum[k][pl];
    for (uint32_t w = w0; w < w1; w++) {
        rp.base[(uint64_t)w * P + p] = run;
        run += part_count(w);
    }
    if (c == SCAN_CHUNKS - 1) rp.ptotal[p] = run;
}

// One block of 1024: exclusive scan of the partition totals (wave scans by
// shuffles, then the 16 wave totals).
__global__ __launch_bounds__(1024) void k_rp_scan_p(const RadixParams *__restrict__ rpp, DevPlan *plan) {
    const RadixParams &rp = *rpp;
    __shared__ uint64_t s_wave[16];
    const uint32_t P = 1u << rp.part_bits;
    const uint32_t per = (P + blockDim.x - 1) / blockDim.x;
    const uint32_t p0 = min(P, threadIdx.x * per), p1 = min(P, p0 + per);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // P <= 8192: at most 8 partitions per thread, loaded together
    uint32_t tot[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) tot[j] = p0 + j < p1 ? rp.ptotal[p0 + j] : 0;
    uint64_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) local += tot[j];
    uint64_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t before = 0, total = 0;
    for (uint32_t w = 0; w < 16; w++) {
        const uint64_t x = s_wave[w];
        before += w < wave ? x : 0;
        total += x;
    }
    uint64_t run = before + incl - local;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++)
        if (p0 + j < p1) {
            rp.pstart[p0 + j] = run;
            run += tot[j];
        }
    if (threadIdx.x == 0) {
        rp.pstart[P] = total;
        plan->cold_rows = total;
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

// LDS bytes of the scatter kernel.
__host__ __device__ inline size_t scatter_lds_bytes(uint32_t hslots, int nacc, uint32_t hmax, uint32_t P,
                                                    uint32_t ring, uint32_t rw) {
    return (size_t)hslots * 12 + ((size_t)nacc * hmax + 1) / 2 * 16 + (size_t)ring * (rw * 8 + 12) + (size_t)P * 4 + 8;
}

// Workgroup w's rows are those of count workgroups g1 * w .. g1 * w + g1 - 1
// (br.per = g1 * the count pass's rows per workgroup), which are contiguous
// in every partition.
template <class S>
__global__ __launch_bounds__(RP_BLOCK) void k_rp_scatter(const DevCtx *__restrict__ cx, BatchRows br, DevTable T) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const S sh(q);
    extern __shared__ uint64_t smem[];
    const uint32_t P = 1u << rp.part_bits;
    const uint32_t NR = rp.ring;
    const uint32_t rw = S::kStatic ? (uint32_t)S::rec_words() : rp.rec_words;
    uint64_t *hk = smem;                                      // [hslots]
    uint64_t *hacc = hk + rp.hslots;                          // [nacc][hmax]
    uint64_t *ring = hacc + (((uint64_t)q.nacc * rp.hmax + 1) & ~1ULL);  // [NR][rw] records in transit (16-B aligned)
    uint32_t *hi = (uint32_t *)(ring + (uint64_t)NR * rw);    // [hslots]
    uint32_t *cur = hi + rp.hslots;                           // [P]
    uint32_t *rpos = cur + P;                                 // [NR] destination record of the entry
    uint32_t *rpub = rpos + NR;                               // [NR] n + 1 once entry n is readable
    uint32_t *rfree = rpub + NR;                              // [NR] next entry number the slot may take
    uint32_t *rctl = rfree + NR;                              // [0] entries reserved, [1] producers done
    const uint32_t wg = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < rp.hslots; i += blockDim.x) {
        hk[i] = rp.hkeys[i];
        hi[i] = rp.hidx[i];
    }
    for (int a = 0; a < q.nacc; a++) {
        uint64_t id = acc_identity(q.acc_op[a]);
        for (uint32_t i = threadIdx.x; i < rp.hmax; i += blockDim.x) hacc[(uint64_t)a * rp.hmax + i] = id;
    }
    for (uint32_t i = threadIdx.x; i < NR; i += blockDim.x) {
        rpub[i] = 0;
        rfree[i] = i;
    }
    if (threadIdx.x < 2) rctl[threadIdx.x] = 0;
    const uint64_t cw = (uint64_t)blockIdx.x * rp.g1;
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x)
        cur[p] = (uint32_t)rp.pstart[p] + rp.base[cw * P + p];
    __syncthreads();
    const uint32_t hmask = rp.hslots - 1;
    const uint64_t cap = T.plan->cap;
    const uint32_t shift = 64 - rp.part_bits;
    const int vw = vmask_word_s<S>(rp);
    const uint32_t lane = threadIdx.x & 63;

    if (threadIdx.x < SC_PROD) {
        // ---- producers: filter, key, hash; heavy rows into LDS state, cold rows into the ring
        bool failed = false;
        for_batch_rows(br, [&](const DevBatch &bb, uint64_t r0, uint64_t r1) {
        const DevBatch b = bb;  // column pointers stay in registers
        auto scatter_row = [&](const Row &row) {
            if (sh.npred() && !eval_preds_s(sh, q, b, row)) return;
            uint64_t w;
            int kr = make_key_s(sh, q, b, row, w);
            if (kr == KEY_DROP) return;
            if (side_row(q, T, cap, kr, w, row)) return;
            const uint64_t h = hash64(w);
            int d = heavy_find(hk, hi, hmask, h, w);
            if (d >= 0) {
                row_update_s<__HIP_MEMORY_SCOPE_WORKGROUP>(sh, row, hacc + d, rp.hmax);
                return;
            }
            const uint32_t p = P > 1 ? (uint32_t)(h >> shift) : 0;
            // exact write offset within the workgroup's run of partition p: one
            // LDS atomic per row, or (msplit_max) the wave multisplit
            const uint32_t pos = P <= rp.msplit_max
                                     ? wave_multisplit<true>(cur, p)
                                     : __hip_atomic_fetch_add(&cur[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            // one reservation per wave: the first active lane takes the wave's entries
            const uint64_t act = __ballot(1);
            const uint32_t before = __popcll(act & ((1ULL << lane) - 1));
            uint32_t n0 = 0;
            if (before == 0)
                n0 = __hip_atomic_fetch_add(&rctl[0], (uint32_t)__popcll(act), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t n = __builtin_amdgcn_readfirstlane(n0) + before;
            const uint32_t i = n & (NR - 1);
            uint32_t spins = 0;
            while (__hip_atomic_load(&rfree[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != n) {
                if (++spins > SC_SPIN_MAX || failed) {
                    failed = true;
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            uint64_t *e = ring + (uint64_t)i * rw;
            if constexpr (S::kStatic && S::rec_words() % 2 == 0) {
                uint64_t tmp[S::rec_words()];
                tmp[0] = w;
#pragma unroll
                for (int u = 0; u < MAXU; u++)
                    if (S::rec_pos(u)) tmp[S::rec_pos(u)] = row.v[u];
                if (vw) tmp[vw] = row.vm;
#pragma unroll
                for (int k = 0; k < S::rec_words(); k += 2) *(u64x2 *)(e + k) = u64x2{tmp[k], tmp[k + 1]};
            } else {
                e[0] = w;
#pragma unroll
                for (int u = 0; u < MAXU; u++) {
                    const int rpu = rec_pos_s<S>(rp, u);
                    if (rpu) e[rpu] = row.v[u];
                }
                if (vw) e[vw] = row.vm;
            }
            rpos[i] = pos;
            __hip_atomic_store(&rpub[i], n + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        for_rows_pipelined<S::kStatic ? AGG_ROWS : 1, SC_PROD>(sh, b, r0, r1, sh.nu(), threadIdx.x, scatter_row);
        });
        if (failed) __hip_atomic_fetch_or(&T.plan->overflow, OVF_RING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lane == 0) __hip_atomic_fetch_add(&rctl[1], 64u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        // ---- consumers: entries c, c + SC_CONS, ... in order, to their partition offsets
        for (uint32_t n = threadIdx.x - SC_PROD;; n += SC_CONS) {
            const uint32_t i = n & (NR - 1);
            bool have = true;
            uint32_t spins = 0;
            while (__hip_atomic_load(&rpub[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != n + 1) {
                // producers done: every reserved entry is published, so n is past the end
                if (__hip_atomic_load(&rctl[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == (uint32_t)SC_PROD &&
                    n >= __hip_atomic_load(&rctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    have = false;
                    break;
                }
                if (++spins > SC_SPIN_MAX) {
                    __hip_atomic_fetch_or(&T.plan->overflow, OVF_RING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    have = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!have) break;
            const uint64_t *e = ring + (uint64_t)i * rw;
            uint64_t *rec = rp.recs + (uint64_t)rpos[i] * rw;
            if constexpr (S::kStatic && S::rec_words() % 2 == 0) {
#pragma unroll
                for (int k = 0; k < S::rec_words(); k += 2) gst((u64x2 *)(rec + k), *(const u64x2 *)(e + k));
            } else {
                for (uint32_t k = 0; k < rw; k++) gst(rec + k, e[k]);
            }
            __hip_atomic_store(&rfree[i], n + NR, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    uint64_t *out = rp.hpart + (uint64_t)wg * q.nacc * rp.hmax;
    for (uint64_t i = threadIdx.x; i < (uint64_t)q.nacc * rp.hmax; i += blockDim.x) out[i] = hacc[i];
}

// One workgroup per partition: all of the partition's groups in LDS.
template <class SH, int KMAX, bool TOPN>
__device__ __forceinline__ void rp_agg_body(const DevCtx *__restrict__ cx, DevTable T, uint32_t S,
                                            uint32_t limit) {
    const DevQuery &q = cx->q;
    const RadixParams &rp = cx->rp;
    const DevOut &o = cx->o;
    const DevTopn &t = cx->t;
    const SH sh(q);
    const uint32_t rw = rp.rec_words;
    const int vw = vmask_word_s<SH>(rp);
    extern __shared__ uint64_t smem[];
    LdsTable lt = lds_table_setup(smem, q, S, limit);
    __shared__ uint32_t s_spill;  // some row of this partition went to the fallback table
    if (threadIdx.x == 0) s_spill = 0;
    __syncthreads();
    const uint32_t p = blockIdx.x;
    const uint64_t r0 = rp.pstart[p], r1 = rp.pstart[p + 1];
    const uint64_t cap = T.plan->cap, mask = cap - 1;
    for (uint64_t base = r0; base < r1; base += (uint64_t)RA_BLOCK * RA_ROWS) {
        Row rows[RA_ROWS];
        uint64_t keys[RA_ROWS];
        // clamped to the partition's last record: every load is unconditional
#pragma unroll
        for (int j = 0; j < RA_ROWS; j++) {
            const uint64_t r = min(base + (uint64_t)j * RA_BLOCK + threadIdx.x, r1 - 1);
            const uint64_t *rec = rp.recs + r * rw;
            if constexpr (SH::kStatic && SH::rec_words() % 2 == 0) {
                uint64_t tmp[SH::rec_words()];
#pragma unroll
                for (int k = 0; k < SH::rec_words(); k += 2) {
                    const u64x2 v2 = gld((const u64x2 *)(rec + k));
                    tmp[k] = v2.x;
                    tmp[k + 1] = v2.y;
                }
                keys[j] = tmp[0];
                rows[j].vm = vw ? (uint32_t)tmp[vw] : 0xFFFFFFFFu;
#pragma unroll
                for (int u = 0; u < MAXU; u++) rows[j].v[u] = SH::rec_pos(u) ? tmp[SH::rec_pos(u)] : 0;
            } else {
                keys[j] = gld(rec);
                rows[j].vm = vw ? (uint32_t)gld(rec + vw) : 0xFFFFFFFFu;
#pragma unroll
                for (int u = 0; u < MAXU; u++) {
                    const int rpu = rec_pos_s<SH>(rp, u);
                    rows[j].v[u] = rpu ? gld(rec + rpu) : 0;
                }
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
                s_spill = 1;
                int64_t g = g_find(T, mask, h, w);
                if (g >= 0) row_update_s<__HIP_MEMORY_SCOPE_AGENT>(sh, rows[j], T.rec + (uint64_t)g * T.stride + 1, 1);
            }
        }
    }
    __syncthreads();
    // A key can reach the fallback table from one row while another row of
    // the same key, racing with the table filling up, claims an LDS slot. Such
    // a key must leave the operator once: fold its LDS state into its
    // fallback record, which the table emit produces, and drop the LDS copy.
    if (s_spill) {
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
            const uint64_t k = lt.key[i];
            if (k == SENTINEL) continue;
            const int64_t g = g_lookup(T, mask, hash64(k), k);
            if (g < 0) continue;
            g_merge(q, T, (uint64_t)g, lt.acc + i, S);
            lt.key[i] = SENTINEL;
        }
        __syncthreads();
    }
    if constexpr (!TOPN) {
        uint32_t mine = 0;
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) mine += lt.key[i] != SENTINEL;
        uint64_t row = block_reserve(mine, &T.plan->groups);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x)
            if (lt.key[i] != SENTINEL) write_group(o, lt.key[i], lt.acc + i, S, false, row++);
    } else {
        // The partition's candidates are its groups whose order value reaches
        // topn_lb, a lower bound of the query's K-th best value that other
        // partitions raise as they finish (any partition's K-th best is one).
        // Groups below it cannot be in the top K. When no more than K groups
        // pass, all of them are candidates and no selection runs.
        __shared__ uint32_t s_win[KMAX];
        __shared__ uint32_t s_nsurv, s_pos;
        if (threadIdx.x == 0) {
            s_nsurv = 0;
            s_pos = 0;
        }
        __syncthreads();
        const uint64_t lb = __hip_atomic_load(&T.plan->topn_lb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // at most RA_TOPN_NI slots per thread (S <= RA_TOPN_SMAX, RA_BLOCK threads)
        uint64_t o[RA_TOPN_NI], ti[RA_TOPN_NI];
        uint32_t sl[RA_TOPN_NI];
        uint32_t mine = 0;
#pragma unroll
        for (int j = 0; j < RA_TOPN_NI; j++) {
            const uint32_t i = threadIdx.x + j * RA_BLOCK;
            uint64_t oj = 0, tj = 0;
            sl[j] = INVALID_SLOT;
            if (i < S && lt.key[i] != SENTINEL) {
                group_order(t, lt.key[i], lt.acc + i, S, false, oj, tj);
                if (oj >= lb) {
                    sl[j] = i;
                    mine++;
                } else {
                    oj = 0;
                    tj = 0;
                }
            }
            o[j] = oj;
            ti[j] = tj;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
        if ((threadIdx.x & 63) == 0 && mine)
            __hip_atomic_fetch_add(&s_nsurv, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __syncthreads();
        const uint32_t nsurv = s_nsurv;
        if (nsurv <= t.k) {
#pragma unroll
            for (int j = 0; j < RA_TOPN_NI; j++)
                if (sl[j] != INVALID_SLOT)
                    s_win[__hip_atomic_fetch_add(&s_pos, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)] = sl[j];
            for (uint32_t r = nsurv + threadIdx.x; r < KMAX; r += blockDim.x) s_win[r] = INVALID_SLOT;
            __syncthreads();
        } else {
            block_topk<KMAX, RA_TOPN_NI>(o, ti, sl, t.k, s_win);
            if (threadIdx.x == 0) {
                const uint32_t kth = s_win[t.k - 1];
                uint64_t ok, tk;
                group_order(t, lt.key[kth], lt.acc + kth, S, false, ok, tk);
                __hip_atomic_fetch_max(&T.plan->topn_lb, ok, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        topn_push(
            t, T.plan, s_win,
            [&](uint32_t i, uint64_t &key, const uint64_t *&acc, uint32_t &stride, bool &nk) {
                key = lt.key[i];
                acc = lt.acc + i;
                stride = S;
                nk = false;
            },
            (int64_t)blockIdx.x);  // partition p owns candidates [p * K, p * K + K)
    }
}

// Full result: two workgroups per CU (LDS tables sized for it).
template <class SH>
__global__ __launch_bounds__(RA_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_rp_agg(
    const DevCtx *__restrict__ cx, DevTable T, uint32_t S, uint32_t limit) {
    rp_agg_body<SH, 1, false>(cx, T, S, limit);
}

// TopN: K candidates per partition; two workgroups per CU, like k_rp_agg.
template <class SH, int KMAX>
__global__ __launch_bounds__(RA_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_rp_agg_topn(const DevCtx *__restrict__ cx, DevTable T, uint32_t S,
                                                          uint32_t limit) {
    rp_agg_body<SH, KMAX, true>(cx, T, S, limit);
}

// Folds the heavy partial states over workgroups: block = 64 heavy keys x 4
// workgroup phases, one accumulator per blockIdx.y. Every accumulator combines
// with an associative, commutative integer operation (double sums are
// fixed-point words), so the folded states do not depend on the order in
// which rows reached the per-workgroup partials.
__device__ __forceinline__ uint64_t fold_acc(int op, uint64_t x, uint64_t y) { return acc_combine(op, x, y); }

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
        for (; w + 12 < rp.W3; w += 16) {
            uint64_t y0 = src[(uint64_t)w * stride], y1 = src[(uint64_t)(w + 4) * stride];
            uint64_t y2 = src[(uint64_t)(w + 8) * stride], y3 = src[(uint64_t)(w + 12) * stride];
            x = fold_acc(op, x, y0);
            x = fold_acc(op, x, y1);
            x = fold_acc(op, x, y2);
            x = fold_acc(op, x, y3);
        }
        for (; w < rp.W3; w += 4) x = fold_acc(op, x, src[(uint64_t)w * stride]);
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
    const uint64_t lb = T.plan->topn_lb;  // groups below it cannot be in the top K
    for (uint32_t d = threadIdx.x; d < hc; d += blockDim.x) {
        const uint64_t *rec = rp.hfinal + (uint64_t)d * stride;
        if (rec[1] == 0) continue;
        uint64_t ord, tie;
        group_order(t, rec[0], rec + 1, 1, false, ord, tie);
        if (ord >= lb) L.offer(ord, tie, d, t.k);
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
