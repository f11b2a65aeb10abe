// gbagg_batch.hpp -- many GROUP BY queries in one submission.
//
// A batch is a list of independent queries of the shape
//   SELECT k, agg(v)... FROM t_i GROUP BY k
// with one INTEGER/BIGINT key and up to four COUNT/SUM/MIN/MAX aggregates over
// INTEGER/BIGINT columns, no NULLs, no predicates, no TopN. The whole batch is
// six kernels captured once in a hipGraph and replayed:
//
//   k_b_hll       per row tile: HyperLogLog registers (p = 10) of its keys
//   k_b_estimate  one block per query: merged registers -> distinct estimate,
//                 which sizes the query's hash table and picks its mode (LDS
//                 first when the query's keys fit the LDS table or repeat
//                 within a tile, by the first tile's own estimate; else global);
//                 a query the learned selector planned on the host takes the
//                 selector's mode, and its predicted count may enlarge the table
//   k_b_layout    one block: table offsets (prefix sum of capacities)
//   k_b_init      fills every table with empty keys and identity states
//   k_b_agg       persistent work queue over (query, tile) items: LDS-first
//                 queries aggregate a tile in an LDS table and merge it into
//                 the query's global table; high-cardinality queries update the
//                 global table directly. Every wave leaves through the same
//                 item-counter check.
//   k_b_compact   per 256-slot chunk of a table: occupied records to the
//                 query's dense result rows (one reservation per block)
//
// Tables hold ceil(1.5 * estimate) + 64 slots rounded up to 256 (multiply-shift
// home slot, linear probing, bounded); a query whose table overflows is flagged
// and the host re-runs it on the single-query operator. So is a query with the
// key INT64_MIN, which equals the tables' empty-slot marker (the single-query
// operator keeps that key in a side record).
#pragma once

#include "gbagg_kernels.hpp"

namespace gbagg {

constexpr int BQ_MAXACC = 4;
constexpr uint32_t BQ_CHUNK = 256;           // table slots per compaction block (capacities are multiples)
constexpr uint32_t BQ_MAX_PROBE = 128;

enum BatchMode : uint32_t { BM_LDS = 0, BM_GLOBAL = 1 };

struct BatchQuery {
    const void *key;
    const void *val[BQ_MAXACC];
    uint64_t rows;
    uint32_t key64;       // key column is BIGINT
    uint32_t val64;       // bit a: value of accumulator a is BIGINT
    uint32_t nacc;
    uint32_t ops;         // AccOp of accumulator a in bits 4a..4a+3
    uint32_t item0;       // first work item of the query
    uint32_t nitems;      // work items (row tiles) of the query
    // filled on the device
    uint64_t est;         // estimated distinct keys
    uint64_t cap;         // table slots (multiple of BQ_CHUNK)
    uint64_t tbl_off;     // first slot of the query's table in the pool
    uint64_t groups;      // result rows
    uint32_t mode;        // BatchMode
    uint32_t overflow;    // table probe bound hit
    // the learned selector's plan of the query (host-set; sel = 0: none)
    uint64_t sel_est;     // predicted distinct keys
    uint32_t sel_mode;    // BatchMode from the predicted path
    uint32_t sel;
};

struct BatchTune {
    uint32_t tile_rows;   // rows per work item
    uint32_t lds_slots;   // LDS table slots of an LDS-first tile
    uint32_t block;       // threads per aggregation workgroup
    uint32_t wgs_per_cu;  // persistent workgroups per CU
};

struct BatchCtl {
    uint64_t total_slots;  // pool slots in use
    uint64_t pool_slots;   // pool slots allocated
    uint32_t next_item;    // work-queue head
    uint32_t n_items;
    uint32_t lds_slots;    // copy of BatchTune::lds_slots for the sketch's mode choice
    uint32_t cap_pct;      // table slots per estimated key, in percent (150)
    uint32_t mode_rule;    // 0: LDS-first iff est <= 4 lds_slots; 1: always; 2: also when keys repeat within a tile
    uint32_t pad;
};

struct BatchPool {
    uint64_t *keys;        // [pool_slots]
    uint64_t *acc;         // [BQ_MAXACC][pool_slots]
    int64_t *out_key;      // [pool_slots] result rows of query q at tbl_off ...
    int64_t *out_acc;      // [BQ_MAXACC][pool_slots]
};

__device__ __forceinline__ int bq_op(const BatchQuery &q, int a) { return (int)((q.ops >> (4 * a)) & 15u); }

__device__ __forceinline__ uint64_t bq_key(const BatchQuery &q, uint64_t r) {
    return q.key64 ? (uint64_t)gld((const int64_t *)q.key + r) : (uint64_t)(int64_t)gld((const int32_t *)q.key + r);
}
__device__ __forceinline__ uint64_t bq_val(const BatchQuery &q, int a, uint64_t r) {
    if (!q.val[a]) return 1;  // COUNT(*)
    return ((q.val64 >> a) & 1) ? (uint64_t)gld((const int64_t *)q.val[a] + r)
                                : (uint64_t)(int64_t)gld((const int32_t *)q.val[a] + r);
}

// ---- sketch: HyperLogLog per tile, then per query -------------------------
// Every key is read once: a sampled estimate misjudges near-uniform keys by
// several times, and an underestimate overflows the table.
constexpr int BQ_HLL_P = 10;
constexpr uint32_t BQ_HLL_M = 1u << BQ_HLL_P;

__global__ __launch_bounds__(256) void k_b_hll(const BatchQuery *__restrict__ bq, const BatchCtl *__restrict__ ctl,
                                               const uint2 *__restrict__ items, uint8_t *regs, BatchTune tune) {
    __shared__ uint32_t r[BQ_HLL_M];
    for (uint32_t item = blockIdx.x; item < ctl->n_items; item += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < BQ_HLL_M; i += blockDim.x) r[i] = 0;
        __syncthreads();
        const uint2 it = items[item];
        const BatchQuery &q = bq[it.x];
        const uint64_t r0 = (uint64_t)it.y * tune.tile_rows, r1 = min(q.rows, r0 + tune.tile_rows);
        for (uint64_t base = r0; base < r1; base += (uint64_t)blockDim.x * 4) {
            uint64_t kk[4];
#pragma unroll
            for (int j = 0; j < 4; j++) kk[j] = bq_key(q, min(base + (uint64_t)j * blockDim.x + threadIdx.x, r1 - 1));
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (base + (uint64_t)j * blockDim.x + threadIdx.x >= r1) continue;
                const uint64_t h = hash64(kk[j]);
                const uint64_t rest = h << BQ_HLL_P;
                const uint32_t rank = rest ? (uint32_t)__clzll((long long)rest) + 1 : (64 - BQ_HLL_P + 1);
                __hip_atomic_fetch_max(&r[h >> (64 - BQ_HLL_P)], rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < BQ_HLL_M; i += blockDim.x) regs[(uint64_t)item * BQ_HLL_M + i] = (uint8_t)r[i];
        __syncthreads();
    }
}

// One block per query: merge its tiles' registers, estimate (Flajolet et al.
// 2007, linear counting for small counts), size the table and pick the mode.
__global__ __launch_bounds__(256) void k_b_estimate(BatchQuery *__restrict__ bq, const BatchCtl *__restrict__ ctl,
                                                    const uint8_t *__restrict__ regs) {
    __shared__ double s_sum[256], s_sum1[256];
    __shared__ uint32_t s_zero[256], s_zero1[256];
    BatchQuery &q = bq[blockIdx.x];
    double sum = 0, sum1 = 0;
    uint32_t zeros = 0, zeros1 = 0;
    for (uint32_t i = threadIdx.x; i < BQ_HLL_M; i += blockDim.x) {
        uint32_t m = 0;
        for (uint32_t t = 0; t < q.nitems; t++) m = max(m, (uint32_t)regs[(uint64_t)(q.item0 + t) * BQ_HLL_M + i]);
        const uint32_t m1 = q.nitems ? regs[(uint64_t)q.item0 * BQ_HLL_M + i] : 0;  // the first tile alone
        sum += ldexp(1.0, -(int)m);
        zeros += m == 0;
        sum1 += ldexp(1.0, -(int)m1);
        zeros1 += m1 == 0;
    }
    s_sum[threadIdx.x] = sum;
    s_zero[threadIdx.x] = zeros;
    s_sum1[threadIdx.x] = sum1;
    s_zero1[threadIdx.x] = zeros1;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
            s_zero[threadIdx.x] += s_zero[threadIdx.x + off];
            s_sum1[threadIdx.x] += s_sum1[threadIdx.x + off];
            s_zero1[threadIdx.x] += s_zero1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = (double)BQ_HLL_M;
        auto hll = [&](double s, uint32_t z) {
            double e = 0.7213 / (1.0 + 1.079 / m) * m * m / s;
            if (e <= 2.5 * m && z > 0) e = m * log(m / (double)z);
            return e;
        };
        double est = hll(s_sum[0], s_zero[0]);
        const double tile_est = hll(s_sum1[0], s_zero1[0]);
        uint64_t e = (uint64_t)(est * 1.05) + 1;
        if (e > q.rows) e = q.rows;
        if (e < 1) e = 1;
        q.est = e;
        uint64_t cap = (e * ctl->cap_pct + 99) / 100 + 64;
        q.cap = (cap + BQ_CHUNK - 1) / BQ_CHUNK * BQ_CHUNK;
        // LDS-first pays when a tile's keys fit the LDS table, or repeat
        // within a tile (skewed keys: the frequent ones stay in LDS)
        const double tile_rows = (double)q.rows / (double)max(q.nitems, 1u);
        const bool fits = e <= (uint64_t)ctl->lds_slots * 4;
        const bool repeats = tile_rows >= 2.0 * tile_est;
        q.mode = (ctl->mode_rule == 1 || fits || (ctl->mode_rule == 2 && repeats)) ? BM_LDS : BM_GLOBAL;
        // a query the learned selector planned takes its mode; the table keeps
        // at least the HyperLogLog size
        if (q.sel && ctl->mode_rule != 1) {
            q.mode = q.sel_mode;
            if (q.sel_est > e) {
                cap = (min(q.sel_est, q.rows) * ctl->cap_pct + 99) / 100 + 64;
                q.cap = (cap + BQ_CHUNK - 1) / BQ_CHUNK * BQ_CHUNK;
            }
        }
        q.groups = 0;
        q.overflow = 0;
    }
}

// ---- layout: table offsets --------------------------------------------------
__global__ __launch_bounds__(1024) void k_b_layout(BatchQuery *__restrict__ bq, uint32_t nq, BatchCtl *ctl) {
    __shared__ uint64_t s_part[1024];
    const uint32_t per = (nq + blockDim.x - 1) / blockDim.x;
    const uint32_t q0 = threadIdx.x * per, q1 = min(nq, q0 + per);
    uint64_t local = 0;
    for (uint32_t q = q0; q < q1; q++) local += bq[q].cap;
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = threadIdx.x >= (uint32_t)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t q = q0; q < q1; q++) {
        bq[q].tbl_off = run;
        run += bq[q].cap;
    }
    if (threadIdx.x == blockDim.x - 1) {
        ctl->total_slots = s_part[blockDim.x - 1];
        ctl->next_item = 0;
    }
}

// Query owning pool slot s (tables are contiguous in query order).
__device__ __forceinline__ uint32_t bq_owner(const BatchQuery *bq, uint32_t nq, uint64_t s) {
    uint32_t lo = 0, hi = nq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (bq[mid].tbl_off <= s) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- init: empty keys, identity states --------------------------------------
__global__ __launch_bounds__(256) void k_b_init(const BatchQuery *__restrict__ bq, uint32_t nq,
                                                const BatchCtl *__restrict__ ctl, BatchPool pool) {
    const uint64_t total = min(ctl->total_slots, ctl->pool_slots);
    for (uint64_t c = blockIdx.x; c * BQ_CHUNK < total; c += gridDim.x) {
        const uint64_t s = c * BQ_CHUNK + threadIdx.x;  // chunks never straddle two tables
        const BatchQuery &q = bq[bq_owner(bq, nq, c * BQ_CHUNK)];
        pool.keys[s] = SENTINEL;
        for (int a = 0; a < BQ_MAXACC; a++)
            if (a < (int)q.nacc) pool.acc[(uint64_t)a * ctl->pool_slots + s] = acc_identity(bq_op(q, a));
    }
}

// ---- global table of one query ---------------------------------------------
__device__ __forceinline__ int64_t bq_find(const BatchQuery &q, const BatchPool &pool, uint64_t k) {
    const uint64_t h = hash64(k);
    uint64_t pos = ((uint64_t)(uint32_t)h * q.cap) >> 32;
#pragma nounroll
    for (uint32_t i = 0; i < BQ_MAX_PROBE; i++) {
        uint64_t *kp = pool.keys + q.tbl_off + pos;
        uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == SENTINEL) {
            uint64_t e = SENTINEL;
            __hip_atomic_compare_exchange_strong(kp, &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            cur = e == SENTINEL ? k : e;
        }
        if (cur == k) return (int64_t)(q.tbl_off + pos);
        pos = pos + 1 == q.cap ? 0 : pos + 1;
    }
    return -1;
}

template <int SCOPE>
__device__ __forceinline__ void bq_apply(uint64_t *p, int op, uint64_t x) {
    acc_apply<SCOPE>(p, op, x);
}

// ---- aggregation: persistent work queue over (query, tile) items -------------
__global__ void k_b_agg(BatchQuery *__restrict__ bq, BatchCtl *ctl, const uint2 *__restrict__ items,
                        BatchPool pool, BatchTune tune) {
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_item, s_fill;
    const uint32_t S = tune.lds_slots;
    uint64_t *lkey = smem;       // [S]
    uint64_t *lacc = smem + S;   // [BQ_MAXACC][S]
    for (;;) {
        if (threadIdx.x == 0) s_item = __hip_atomic_fetch_add(&ctl->next_item, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= ctl->n_items) break;  // every wave of the block sees the same item
        const uint2 it = items[item];
        BatchQuery &q = bq[it.x];
        const uint64_t r0 = (uint64_t)it.y * tune.tile_rows, r1 = min(q.rows, r0 + tune.tile_rows);
        const int nacc = (int)q.nacc;
        if (q.mode == BM_LDS) {
            for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
                lkey[i] = SENTINEL;
                for (int a = 0; a < nacc; a++) lacc[(uint64_t)a * S + i] = acc_identity(bq_op(q, a));
            }
            if (threadIdx.x == 0) s_fill = 0;
            __syncthreads();
            for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
                const uint64_t k = bq_key(q, r);
                if (k == SENTINEL) {  // INT64_MIN is the empty-slot marker: the query re-runs unbatched
                    q.overflow = 1;
                    continue;
                }
                // LDS slot of k: find or claim, bounded; else the global table
                uint32_t pos = (uint32_t)(((uint64_t)(uint32_t)hash64(k) * S) >> 32);
                int slot = -1;
                for (int p = 0; p < 16; p++) {
                    uint64_t cur = __hip_atomic_load(&lkey[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (cur == SENTINEL) {
                        if (__hip_atomic_load(&s_fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= S / 2) break;
                        uint64_t e = SENTINEL;
                        __hip_atomic_compare_exchange_strong(&lkey[pos], &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (e == SENTINEL) atomicAdd(&s_fill, 1u);
                        cur = e == SENTINEL ? k : e;
                    }
                    if (cur == k) {
                        slot = (int)pos;
                        break;
                    }
                    pos = pos + 1 == S ? 0 : pos + 1;
                }
                if (slot >= 0) {
                    for (int a = 0; a < nacc; a++)
                        bq_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(lacc + (uint64_t)a * S + slot, bq_op(q, a), bq_val(q, a, r));
                } else {
                    const int64_t g = bq_find(q, pool, k);
                    if (g < 0) {
                        q.overflow = 1;
                        continue;
                    }
                    for (int a = 0; a < nacc; a++)
                        bq_apply<__HIP_MEMORY_SCOPE_AGENT>(pool.acc + (uint64_t)a * ctl->pool_slots + g, bq_op(q, a),
                                                           bq_val(q, a, r));
                }
            }
            __syncthreads();
            // the tile's LDS states into the query's global table
            for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
                const uint64_t k = lkey[i];
                if (k == SENTINEL) continue;
                const int64_t g = bq_find(q, pool, k);
                if (g < 0) {
                    q.overflow = 1;
                    continue;
                }
                for (int a = 0; a < nacc; a++) {
                    const uint64_t x = lacc[(uint64_t)a * S + i];
                    if (acc_nontrivial(bq_op(q, a), x))
                        bq_apply<__HIP_MEMORY_SCOPE_AGENT>(pool.acc + (uint64_t)a * ctl->pool_slots + g, bq_op(q, a), x);
                }
            }
        } else {
            for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
                const uint64_t k = bq_key(q, r);
                if (k == SENTINEL) {
                    q.overflow = 1;
                    continue;
                }
                const int64_t g = bq_find(q, pool, k);
                if (g < 0) {
                    q.overflow = 1;
                    continue;
                }
                for (int a = 0; a < nacc; a++)
                    bq_apply<__HIP_MEMORY_SCOPE_AGENT>(pool.acc + (uint64_t)a * ctl->pool_slots + g, bq_op(q, a),
                                                       bq_val(q, a, r));
            }
        }
        __syncthreads();  // LDS table and s_item are reused by the next item
    }
}

// ---- compaction: table records to dense result rows -------------------------
__global__ __launch_bounds__(BQ_CHUNK) void k_b_compact(BatchQuery *__restrict__ bq, uint32_t nq,
                                                        const BatchCtl *__restrict__ ctl, BatchPool pool) {
    const uint64_t total = min(ctl->total_slots, ctl->pool_slots);
    for (uint64_t c = blockIdx.x; c * BQ_CHUNK < total; c += gridDim.x) {
        const uint64_t s = c * BQ_CHUNK + threadIdx.x;
        BatchQuery &q = bq[bq_owner(bq, nq, c * BQ_CHUNK)];
        const uint64_t k = pool.keys[s];
        const uint32_t mine = k != SENTINEL;
        const uint64_t row = block_reserve(mine, &q.groups);
        if (mine) {
            const uint64_t dst = q.tbl_off + row;
            pool.out_key[dst] = (int64_t)k;
            for (int a = 0; a < BQ_MAXACC; a++)
                if (a < (int)q.nacc)
                    pool.out_acc[(uint64_t)a * ctl->pool_slots + dst] = (int64_t)pool.acc[(uint64_t)a * ctl->pool_slots + s];
        }
    }
}

}  // namespace gbagg
