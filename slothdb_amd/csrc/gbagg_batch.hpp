// gbagg_batch.hpp -- many GROUP BY queries in one submission.
//
// A batch is a list of independent queries of the shape
//   SELECT k, agg(v)... FROM t_i GROUP BY k
// with one INTEGER/BIGINT key and up to four COUNT/SUM/MIN/MAX aggregates over
// INTEGER/BIGINT columns, no NULLs, no predicates, no TopN. The whole batch is
// eight kernels captured once in a hipGraph and replayed:
//
//   k_b_hll       per row tile: HyperLogLog registers (p = 10) of its keys and
//                 a histogram of the keys' top 10 hash bits
//   k_b_estimate  one block per query: merged registers -> distinct estimate,
//                 which sizes the query's hash table and picks its mode: LDS
//                 first when the query's keys fit the LDS table or repeat
//                 within a tile (by the first tile's own estimate), else radix
//                 partitions (else, beyond what 1024 partitions hold, the
//                 global table); a query the learned selector planned on the
//                 host takes the selector's mode, and its predicted count may
//                 enlarge the table
//   k_b_layout    one block: table offsets, record regions and partition work
//                 items (prefix sums)
//   k_b_rscan     one block per radix query: exact partition offsets per row
//                 tile from the histograms
//   k_b_init      fills every global table with empty keys and identity states
//   k_b_agg       persistent work queue over (query, tile) items: LDS-first
//                 queries aggregate a tile in an LDS table; a row that finds
//                 no room there, and at the tile's end every partial state of
//                 the table, becomes a record in the query's radix partitions
//                 (or, beyond what 1024 partitions hold, goes to the query's
//                 global table); radix queries write each row as a record
//                 (key, values) at its partition's exact offset; global-table
//                 queries update the table directly. Every wave leaves through
//                 the same item-counter check.
//   k_b_part      persistent work queue over the partitions of the queries
//                 that write records: all groups of one partition in an LDS
//                 table (raw rows and partial states combine alike), then
//                 straight to the query's dense result rows
//   k_b_compact   per 256-slot chunk of a global table: occupied records to
//                 the query's dense result rows (one reservation per block)
//
// Tables hold ceil(1.5 * estimate) + 64 slots rounded up to 256 (multiply-shift
// home slot, linear probing, bounded); a radix query's result rows use the
// same capacity. A query whose table, partition table or result rows overflow
// is flagged and the host re-runs it on the single-query operator. So is a
// query with the key INT64_MIN, which equals the tables' empty-slot marker
// (the single-query operator keeps that key in a side record).
//
// Per-row global atomics run at a fixed rate on this GPU, whatever the table
// size or the atomic's scope; a radix query instead moves each row twice
// through HBM as a record and does all of its hashing and atomics in LDS.
#pragma once

#include "gbagg_kernels.hpp"

namespace gbagg {

constexpr int BQ_MAXACC = 4;
constexpr uint32_t BQ_CHUNK = 256;           // table slots per compaction block (capacities are multiples)
constexpr uint32_t BQ_MAX_PROBE = 128;
constexpr uint32_t BQ_RBITS = 10;                 // radix: histogram of the top 10 hash bits per tile
constexpr uint32_t BQ_RBUCKETS = 1u << BQ_RBITS;  // (partitions = the top pbits <= 10 of them)
constexpr uint32_t BQ_PART_BLOCK = 512;           // k_b_part threads
constexpr uint32_t BQ_SEGS = 256;                 // k_b_part: tile segments mapped at a time (<= BQ_PART_BLOCK)
constexpr uint64_t BQ_PART_ROWS = 65536;          // radix: rows per partition the partition count aims at
constexpr uint32_t BQ_TILE_PROBES = 16;           // LDS-first tile table: probe bound (fill limit 1/2)

// BM_LDS: LDS-first tiles merged into a global table; BM_GLOBAL: global table
// per row; BM_RADIX: every row a record (key, value words); BM_LDS_RADIX:
// LDS-first tiles, rows without room and the tiles' partial states as records
// (key, one word per accumulator)
enum BatchMode : uint32_t { BM_LDS = 0, BM_GLOBAL = 1, BM_RADIX = 2, BM_LDS_RADIX = 3 };

__host__ __device__ __forceinline__ bool bq_records(uint32_t mode) { return mode == BM_RADIX || mode == BM_LDS_RADIX; }

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
    // radix mode (host-set record shape; the rest filled on the device)
    uint32_t rec_words;   // BM_RADIX: key + one word per accumulator with a value column;
                          // BM_LDS_RADIX: key + one word per accumulator (set on the device)
    uint32_t valmask;     // bit a: accumulator a reads a value column
    uint64_t rec_off;     // first word of the query's record region
    uint32_t pbits;       // partitions 1 << pbits
    uint32_t part0;       // first partition work item
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
    uint32_t mode_rule;    // 0: LDS-first iff est <= 4 lds_slots; 1: always; 2: also when keys repeat within a
                           // tile; 3: radix partitions for every query they can hold (tests)
    uint32_t radix;        // 1: queries not LDS-first take radix partitions when they fit
    uint32_t part_slots;   // k_b_part LDS table slots
    uint32_t n_parts;      // radix partition work items
    uint32_t next_part;    // their work-queue head
    uint64_t rec_words;    // record words in use
    uint64_t rec_cap;      // record words allocated
};

struct BatchPool {
    uint64_t *keys;        // [pool_slots]
    uint64_t *acc;         // [BQ_MAXACC][pool_slots]
    int64_t *out_key;      // [pool_slots] result rows of query q at tbl_off ...
    int64_t *out_acc;      // [BQ_MAXACC][pool_slots]
    uint32_t *bhist;       // [n_items][BQ_RBUCKETS] rows per top-hash bucket; then, for radix
                           // queries, each partition's first record of the tile
    uint32_t *pstart;      // [nq][BQ_RBUCKETS + 1] radix: partition starts in the query's records
    uint32_t *bfill;       // [n_items][BQ_RBUCKETS] BM_LDS_RADIX: records a tile wrote per partition
    uint64_t *recs;        // radix records, query regions at rec_off
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
                                               const uint2 *__restrict__ items, uint8_t *regs, uint32_t *bhist,
                                               BatchTune tune) {
    static_assert(BQ_RBITS == BQ_HLL_P, "register index and radix bucket are the same top hash bits");
    __shared__ uint32_t r[BQ_HLL_M];
    __shared__ uint32_t bh[BQ_RBUCKETS];
    for (uint32_t item = blockIdx.x; item < ctl->n_items; item += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < BQ_HLL_M; i += blockDim.x) {
            r[i] = 0;
            bh[i] = 0;
        }
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
                // registers only grow: a read first, the atomic only when it grows
                uint32_t *rp = &r[h >> (64 - BQ_HLL_P)];
                if (__hip_atomic_load(rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < rank)
                    __hip_atomic_fetch_max(rp, rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&bh[h >> (64 - BQ_RBITS)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < BQ_HLL_M; i += blockDim.x) {
            regs[(uint64_t)item * BQ_HLL_M + i] = (uint8_t)r[i];
            bhist[(uint64_t)item * BQ_RBUCKETS + i] = bh[i];
        }
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
        // radix partitions: each holds about 0.6 part_slots groups or fewer,
        // so the largest of them stays below the 7/8 fill limit, and at most
        // BQ_PART_ROWS rows (records), so one partition's workgroup does not
        // hold up the rest
        q.pbits = 0;
        if (ctl->radix || ctl->mode_rule == 3) {
            const uint64_t want = max<uint64_t>(e, q.sel ? min(q.sel_est, q.rows) : 0);
            const double per = 0.6 * (double)ctl->part_slots;
            uint32_t pb = 0;
            while (pb < BQ_RBITS && ((double)want > per * (double)(1u << pb) ||
                                     q.rows > (uint64_t)BQ_PART_ROWS << pb))
                pb++;
            if ((double)want <= per * (double)(1u << pb)) {
                if (ctl->mode_rule == 3 || q.mode == BM_GLOBAL) {
                    q.mode = BM_RADIX;
                } else {
                    q.mode = BM_LDS_RADIX;
                    q.rec_words = 1 + q.nacc;
                }
                q.pbits = pb;
            }
        }
        q.groups = 0;
        q.overflow = 0;
    }
}

// ---- layout: table offsets, record regions, partition work items ------------
// Exclusive prefix over the queries of val(q), written by put(q, offset);
// returns the total. Each thread takes `per` consecutive queries.
template <class V, class P>
__device__ uint64_t block_prefix(uint32_t nq, uint64_t *s_part, V val, P put) {
    const uint32_t per = (nq + blockDim.x - 1) / blockDim.x;
    const uint32_t q0 = threadIdx.x * per, q1 = min(nq, q0 + per);
    uint64_t local = 0;
    for (uint32_t q = q0; q < q1; q++) local += val(q);
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        uint64_t v = threadIdx.x >= (uint32_t)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t q = q0; q < q1; q++) {
        put(q, run);
        run += val(q);
    }
    const uint64_t total = s_part[blockDim.x - 1];
    __syncthreads();  // s_part is reused by the next call
    return total;
}

__global__ __launch_bounds__(1024) void k_b_layout(BatchQuery *__restrict__ bq, uint32_t nq, BatchCtl *ctl) {
    __shared__ uint64_t s_part[1024];
    const uint64_t slots = block_prefix(nq, s_part, [&](uint32_t q) { return bq[q].cap; },
                                        [&](uint32_t q, uint64_t o) { bq[q].tbl_off = o; });
    const uint64_t words = block_prefix(
        nq, s_part, [&](uint32_t q) { return bq_records(bq[q].mode) ? bq[q].rows * bq[q].rec_words : 0; },
        [&](uint32_t q, uint64_t o) { bq[q].rec_off = o; });
    const uint64_t parts = block_prefix(
        nq, s_part, [&](uint32_t q) { return bq_records(bq[q].mode) ? (uint64_t)1 << bq[q].pbits : 0; },
        [&](uint32_t q, uint64_t o) { bq[q].part0 = (uint32_t)o; });
    if (threadIdx.x == 0) {
        ctl->total_slots = slots;
        ctl->rec_words = words;
        ctl->n_parts = (uint32_t)parts;
        ctl->next_item = 0;
        ctl->next_part = 0;
    }
}

// ---- radix: exact partition offsets per tile ----------------------------------
// One block of BQ_RBUCKETS threads per query. Partition p of a query with
// pbits bits is buckets [p << (10 - pbits), (p + 1) << (10 - pbits)) of the
// tiles' histograms. The partitions are laid out one after another in the
// query's record region; within one, the tiles' rows follow in tile order.
// Each tile's histogram row is overwritten with the first record of each of
// its partitions (at the partition's first bucket). Thread b reads bucket b
// of every tile (coalesced); partition counts are LDS sums.
__global__ __launch_bounds__(BQ_RBUCKETS) void k_b_rscan(const BatchQuery *__restrict__ bq, BatchPool pool) {
    __shared__ uint32_t s_scan[BQ_RBUCKETS];
    __shared__ uint32_t s_tile[BQ_RBUCKETS];
    const BatchQuery &q = bq[blockIdx.x];
    if (!bq_records(q.mode)) return;
    const uint32_t P = 1u << q.pbits, sh = BQ_RBITS - q.pbits;
    const uint32_t b = threadIdx.x, pb = b >> sh;  // bucket b lies in partition pb
    uint32_t *const hist = pool.bhist + (uint64_t)q.item0 * BQ_RBUCKETS;
    // rows per bucket over all tiles, then per partition
    uint32_t tot = 0;
    for (uint32_t t = 0; t < q.nitems; t++) tot += hist[(uint64_t)t * BQ_RBUCKETS + b];
    s_scan[b] = 0;
    __syncthreads();
    __hip_atomic_fetch_add(&s_scan[pb], tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    const uint32_t ptot = b < P ? s_scan[b] : 0;
    __syncthreads();
    // exclusive scan of the partition totals (Hillis-Steele over the block)
    s_scan[b] = ptot;
    __syncthreads();
    for (uint32_t off = 1; off < BQ_RBUCKETS; off <<= 1) {
        const uint32_t v = b >= off ? s_scan[b - off] : 0;
        __syncthreads();
        s_scan[b] += v;
        __syncthreads();
    }
    uint32_t run = s_scan[b] - ptot;  // thread p < P: partition p's first record
    uint32_t *ps = pool.pstart + (uint64_t)blockIdx.x * (BQ_RBUCKETS + 1);
    if (b < P) ps[b] = run;
    if (b == P - 1) ps[P] = run + ptot;
    s_tile[b] = 0;
    __syncthreads();
    for (uint32_t t = 0; t < q.nitems; t++) {
        uint32_t *h = hist + (uint64_t)t * BQ_RBUCKETS;
        const uint32_t c = h[b];
        if (c) __hip_atomic_fetch_add(&s_tile[pb], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __syncthreads();  // every bucket of tile t is read before its row is overwritten
        if (b < P) {
            h[b << sh] = run;
            run += s_tile[b];
            s_tile[b] = 0;
        }
        __syncthreads();
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
        if (bq_records(q.mode)) continue;  // no table: its slots are the result rows of k_b_part
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
// Rows are taken BQ_U per thread at a time: every key and value load of the
// group is issued before the first row is used, so the loads overlap.
constexpr int BQ_U = 4;

struct BqRows {
    uint64_t k[BQ_U];
    uint64_t v[BQ_U][BQ_MAXACC];
};

// Loads rows base + j * blockDim.x + threadIdx.x (clamped to the tile's last
// row, so every load is in bounds) of query q.
__device__ __forceinline__ void bq_load(const BatchQuery &q, uint64_t base, uint64_t r1, BqRows &x) {
#pragma unroll
    for (int j = 0; j < BQ_U; j++) {
        const uint64_t r = min(base + (uint64_t)j * blockDim.x + threadIdx.x, r1 - 1);
        x.k[j] = bq_key(q, r);
#pragma unroll
        for (int a = 0; a < BQ_MAXACC; a++) x.v[j][a] = a < (int)q.nacc ? bq_val(q, a, r) : 0;
    }
}

// LDS slot of key k in the table lkey[S]: found or claimed while the fill is
// below `limit`; -1 when neither happens within `probes` slots.
__device__ __forceinline__ int bq_lds_slot(uint64_t *lkey, uint32_t S, uint32_t *fill, uint32_t limit,
                                           uint64_t k, uint32_t probes) {
    uint32_t pos = (uint32_t)(((uint64_t)(uint32_t)hash64(k) * S) >> 32);
    for (uint32_t p = 0; p < probes; p++) {
        uint64_t cur = __hip_atomic_load(&lkey[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == SENTINEL) {
            if (__hip_atomic_load(fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= limit) return -1;
            uint64_t e = SENTINEL;
            __hip_atomic_compare_exchange_strong(&lkey[pos], &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
            if (e == SENTINEL) __hip_atomic_fetch_add(fill, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            cur = e == SENTINEL ? k : e;
        }
        if (cur == k) return (int)pos;
        pos = pos + 1 == S ? 0 : pos + 1;
    }
    return -1;
}

// Partition cursor of a record: lanes of a wave in the same partition as the
// first active lane take consecutive offsets from one LDS atomic (ballot +
// popcount rank) when at least 8 of them share it; the others take one each.
__device__ __forceinline__ uint32_t bq_wave_offset(uint32_t *cur, bool active, uint32_t part) {
    const uint64_t act = __ballot(active);
    const int lane = threadIdx.x & 63;
    if (act) {
        const int first = __ffsll((unsigned long long)act) - 1;
        const uint32_t lead = __shfl(part, first);
        const uint64_t m = __ballot(active && part == lead);
        if (__popcll(m) >= 8) {
            uint32_t base = 0;
            if (lane == first)
                base = __hip_atomic_fetch_add(&cur[lead], (uint32_t)__popcll(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            base = __shfl(base, first);
            if ((m >> lane) & 1) return base + (uint32_t)__popcll(m & ((1ULL << lane) - 1));
        }
    }
    if (!active) return 0;
    return __hip_atomic_fetch_add(&cur[part], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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
        const uint64_t step = (uint64_t)blockDim.x * BQ_U;
        // a row into the query's global table
        // (value arrays are indexed with constants only, so they stay in registers)
        auto to_global = [&](uint64_t k, const uint64_t (&v)[BQ_MAXACC]) {
            const int64_t g = bq_find(q, pool, k);
            if (g < 0) {
                q.overflow = 1;
                return;
            }
#pragma unroll
            for (int a = 0; a < BQ_MAXACC; a++)
                if (a < nacc)
                    bq_apply<__HIP_MEMORY_SCOPE_AGENT>(pool.acc + (uint64_t)a * ctl->pool_slots + g, bq_op(q, a), v[a]);
        };
        if (q.mode == BM_LDS || q.mode == BM_LDS_RADIX) {
            // BM_LDS_RADIX: partition cursors after the table, from the tile's
            // first record of each partition
            const bool recs = q.mode == BM_LDS_RADIX;
            uint32_t *cur = (uint32_t *)(lacc + (uint64_t)nacc * S);  // [P]
            const uint32_t P = 1u << q.pbits, sh = BQ_RBITS - q.pbits;
            const uint32_t rw = q.rec_words;
            uint64_t *rq = pool.recs + q.rec_off;
            // a record (key, one word per accumulator) at its partition's cursor
            auto to_record = [&](uint64_t k, const uint64_t (&v)[BQ_MAXACC]) {
                const uint32_t part = q.pbits ? (uint32_t)(hash64(k) >> (64 - q.pbits)) : 0;
                const uint32_t pos = __hip_atomic_fetch_add(&cur[part], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                uint64_t *rec = rq + (uint64_t)pos * rw;
                gst(rec, k);
#pragma unroll
                for (int a = 0; a < BQ_MAXACC; a++)
                    if (a < nacc) gst(rec + 1 + a, v[a]);
            };
            for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
                lkey[i] = SENTINEL;
                for (int a = 0; a < nacc; a++) lacc[(uint64_t)a * S + i] = acc_identity(bq_op(q, a));
            }
            if (recs)
                for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
                    cur[i] = pool.bhist[(uint64_t)item * BQ_RBUCKETS + ((uint64_t)i << sh)];
            if (threadIdx.x == 0) s_fill = 0;
            __syncthreads();
            for (uint64_t base = r0; base < r1; base += step) {
                BqRows x;
                bq_load(q, base, r1, x);
#pragma unroll
                for (int j = 0; j < BQ_U; j++) {
                    const bool live = base + (uint64_t)j * blockDim.x + threadIdx.x < r1;
                    const uint64_t k = x.k[j];
                    if (!live) continue;
                    if (k == SENTINEL) {  // INT64_MIN is the empty-slot marker: the query re-runs unbatched
                        q.overflow = 1;
                        continue;
                    }
                    // LDS slot of k: find or claim, bounded; else a record or the global table
                    const int slot = bq_lds_slot(lkey, S, &s_fill, S / 2, k, BQ_TILE_PROBES);
                    if (slot >= 0) {
#pragma unroll
                        for (int a = 0; a < BQ_MAXACC; a++)
                            if (a < nacc)
                                bq_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(lacc + (uint64_t)a * S + slot, bq_op(q, a), x.v[j][a]);
                    } else if (recs) {
                        to_record(k, x.v[j]);
                    } else {
                        to_global(k, x.v[j]);
                    }
                }
            }
            __syncthreads();
            if (recs) {
                // the tile's partial states as records; a partition gets at
                // most as many records as the tile has rows in it (each state
                // stands for at least one row that did not become a record)
                for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
                    const uint64_t k = lkey[i];
                    if (k == SENTINEL) continue;
                    uint64_t v[BQ_MAXACC];
#pragma unroll
                    for (int a = 0; a < BQ_MAXACC; a++) v[a] = a < nacc ? lacc[(uint64_t)a * S + i] : 0;
                    to_record(k, v);
                }
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
                    pool.bfill[(uint64_t)item * BQ_RBUCKETS + i] =
                        cur[i] - pool.bhist[(uint64_t)item * BQ_RBUCKETS + ((uint64_t)i << sh)];
            }
            // the tile's LDS states into the query's global table
            for (uint32_t i = threadIdx.x; i < S && !recs; i += blockDim.x) {
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
        } else if (q.mode == BM_RADIX) {
            // each row as a record at its partition's next offset (the LDS
            // cursors start at the tile's first record of each partition)
            uint32_t *cur = (uint32_t *)smem;  // [P]
            const uint32_t P = 1u << q.pbits, sh = BQ_RBITS - q.pbits;
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
                cur[i] = pool.bhist[(uint64_t)item * BQ_RBUCKETS + ((uint64_t)i << sh)];
            __syncthreads();
            const uint32_t rw = q.rec_words;
            uint64_t *rq = pool.recs + q.rec_off;
            for (uint64_t base = r0; base < r1; base += step) {
                BqRows x;
                bq_load(q, base, r1, x);
#pragma unroll
                for (int j = 0; j < BQ_U; j++) {
                    const bool live = base + (uint64_t)j * blockDim.x + threadIdx.x < r1;
                    const uint64_t k = x.k[j];
                    if (live && k == SENTINEL) q.overflow = 1;  // the record is still written: offsets stay exact
                    const uint32_t part = q.pbits ? (uint32_t)(hash64(k) >> (64 - q.pbits)) : 0;
                    const uint32_t pos = bq_wave_offset(cur, live, part);
                    if (!live) continue;
                    uint64_t *rec = rq + (uint64_t)pos * rw;
                    gst(rec, k);
                    uint32_t w = 1;
#pragma unroll
                    for (int a = 0; a < BQ_MAXACC; a++)
                        if (a < nacc && ((q.valmask >> a) & 1)) gst(rec + w++, x.v[j][a]);
                }
            }
        } else {
            for (uint64_t base = r0; base < r1; base += step) {
                BqRows x;
                bq_load(q, base, r1, x);
#pragma unroll
                for (int j = 0; j < BQ_U; j++) {
                    if (base + (uint64_t)j * blockDim.x + threadIdx.x >= r1) continue;
                    if (x.k[j] == SENTINEL) {
                        q.overflow = 1;
                        continue;
                    }
                    to_global(x.k[j], x.v[j]);
                }
            }
        }
        __syncthreads();  // LDS table and s_item are reused by the next item
    }
}

// ---- radix partitions: one LDS table per partition ---------------------------
// Persistent workgroups take partition work items in order; item i belongs to
// the last query whose first item is <= i. A group that finds no LDS slot (the
// partition holds more groups than the fill limit) or no result row flags the
// query, which the host then re-runs unbatched.
__device__ __forceinline__ uint32_t bq_part_owner(const BatchQuery *bq, uint32_t nq, uint32_t item) {
    uint32_t lo = 0, hi = nq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (bq[mid].part0 <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(BQ_PART_BLOCK) void k_b_part(BatchQuery *__restrict__ bq, uint32_t nq, BatchCtl *ctl,
                                                          BatchPool pool) {
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_item, s_fill;
    __shared__ uint32_t s_seg[BQ_SEGS], s_pre[BQ_SEGS + 1];
    const uint32_t S = ctl->part_slots, limit = S / 8 * 7;
    uint64_t *lkey = smem;       // [S]
    uint64_t *lacc = smem + S;   // [BQ_MAXACC][S]
    for (;;) {
        if (threadIdx.x == 0) s_item = __hip_atomic_fetch_add(&ctl->next_part, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= ctl->n_parts) break;  // every wave of the block sees the same item
        const uint32_t qi = bq_part_owner(bq, nq, item);
        BatchQuery &q = bq[qi];
        const uint32_t p = item - q.part0;
        const uint32_t *ps = pool.pstart + (uint64_t)qi * (BQ_RBUCKETS + 1);
        const int nacc = (int)q.nacc;
        const uint32_t rw = q.rec_words;
        const bool lr = q.mode == BM_LDS_RADIX;  // records hold one word per accumulator
        const uint64_t *rq = pool.recs + q.rec_off;
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
            lkey[i] = SENTINEL;
            for (int a = 0; a < nacc; a++) lacc[(uint64_t)a * S + i] = acc_identity(bq_op(q, a));
        }
        if (threadIdx.x == 0) s_fill = 0;
        __syncthreads();
        // records rec_of(0 .. n - 1) of the query's region into the table
        auto run = [&](uint64_t n, auto rec_of) {
            for (uint64_t base = 0; base < n; base += (uint64_t)blockDim.x * BQ_U) {
                uint64_t kk[BQ_U], vv[BQ_U][BQ_MAXACC];
#pragma unroll
                for (int j = 0; j < BQ_U; j++) {  // clamped to the last record
                    const uint64_t *rec = rq + rec_of(min(base + (uint64_t)j * blockDim.x + threadIdx.x, n - 1)) * rw;
                    kk[j] = gld(rec);
#pragma unroll
                    for (int a = 0; a < BQ_MAXACC; a++) {
                        // accumulator a's word: 1 + a, or 1 + (value columns before a)
                        const uint32_t w = lr ? 1 + a : 1 + __popc(q.valmask & ((1u << a) - 1));
                        vv[j][a] = a < nacc && (lr || ((q.valmask >> a) & 1)) ? gld(rec + w) : 1;
                    }
                }
#pragma unroll
                for (int j = 0; j < BQ_U; j++) {
                    const bool live = base + (uint64_t)j * blockDim.x + threadIdx.x < n;
                    const uint64_t k = kk[j];
                    if (!live || k == SENTINEL) continue;  // SENTINEL: the query is flagged already
                    const int slot = bq_lds_slot(lkey, S, &s_fill, limit, k, BQ_MAX_PROBE);
                    if (slot < 0) {
                        q.overflow = 1;
                        continue;
                    }
#pragma unroll
                    for (int a = 0; a < BQ_MAXACC; a++)
                        if (a < nacc) bq_apply<__HIP_MEMORY_SCOPE_WORKGROUP>(lacc + (uint64_t)a * S + slot, bq_op(q, a), vv[j][a]);
                }
            }
        };
        if (!lr) {
            const uint64_t s0 = ps[p];
            run(ps[p + 1] - s0, [&](uint64_t i) { return s0 + i; });
        } else {
            // one segment per tile (its first record of the partition and its
            // count), BQ_SEGS tiles at a time: a prefix of their counts maps
            // each index of their concatenation to its segment, so the loads
            // of all of them are in flight together
            const uint32_t sh = BQ_RBITS - q.pbits;
            for (uint32_t t0 = 0; t0 < q.nitems; t0 += BQ_SEGS) {
                const uint32_t nt = min(BQ_SEGS, q.nitems - t0);
                if (threadIdx.x < nt) {
                    const uint64_t row = (uint64_t)(q.item0 + t0 + threadIdx.x) * BQ_RBUCKETS;
                    s_seg[threadIdx.x] = pool.bhist[row + ((uint64_t)p << sh)];
                    s_pre[threadIdx.x + 1] = pool.bfill[row + p];
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    s_pre[0] = 0;
                    for (uint32_t i = 1; i <= nt; i++) s_pre[i] += s_pre[i - 1];
                }
                __syncthreads();
                run(s_pre[nt], [&](uint64_t i) {
                    uint32_t lo = 0, hi = nt - 1;  // the last segment starting at or before i
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) / 2;
                        if (s_pre[mid] <= i) lo = mid;
                        else hi = mid - 1;
                    }
                    return (uint64_t)s_seg[lo] + (i - s_pre[lo]);
                });
                __syncthreads();  // s_seg / s_pre are refilled by the next group of tiles
            }
        }
        __syncthreads();
        // the partition's groups to the query's result rows
        uint32_t mine = 0;
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) mine += lkey[i] != SENTINEL;
        uint64_t row = block_reserve(mine, &q.groups);
        for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
            const uint64_t k = lkey[i];
            if (k == SENTINEL) continue;
            if (row >= q.cap) {  // more groups than the estimate allowed for
                q.overflow = 1;
                break;
            }
            const uint64_t dst = q.tbl_off + row++;
            pool.out_key[dst] = (int64_t)k;
            for (int a = 0; a < nacc; a++)
                pool.out_acc[(uint64_t)a * ctl->pool_slots + dst] = (int64_t)lacc[(uint64_t)a * S + i];
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
        if (bq_records(q.mode)) continue;  // the same for every thread of the block: k_b_part wrote its rows
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
