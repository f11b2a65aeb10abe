// gbagg_parquet.hpp -- Parquet row-group decode on the GPU: the step before
// the aggregate. It delivers the fused path's per-row-group columns
// (ParquetColumnData: i32 / i64 / f64 values, dictionary codes plus the
// dictionary, a validity byte per row; include/slothdb/storage/parquet.hpp:147-207)
// as HBM-resident gbagg_column_views, from a file image in host memory.
//
// Scope: flat schemas (REQUIRED or OPTIONAL top-level columns), uncompressed
// column chunks, INT32 / INT64 / DOUBLE / BYTE_ARRAY, data pages V1 and V2,
// PLAIN and PLAIN_DICTIONARY / RLE_DICTIONARY encodings. A BYTE_ARRAY column
// whose pages are all dictionary-encoded is delivered as dictionary codes;
// one with a PLAIN page (a writer's fallback once its dictionary grows too
// large) is delivered as its strings (offsets + bytes, GBAGG_STR_RAW), which
// the operator interns. The reference reader's page decoding is
// src/storage/parquet.cpp:1480-1660; the formats are the Parquet format
// specification's (Thrift compact protocol metadata, RLE / bit-packed hybrid
// levels and indices).
//
// Host: the footer and page headers, and the run headers of the definition
// levels (a few bytes per run; bit-packed runs are not read), collected over
// all columns of a decode call. Device: the column chunks' bytes (in place
// when the file image is resident in HBM), then
//   k_pq_streams  one workgroup per dictionary-encoded page: its index
//                 stream's run headers, then its indices to u32 (calls over
//                 many column chunks; the host reads them in small calls)
//   k_pq_hybrid   every run the host read, one thread per value: definition
//                 levels to validity bytes, dictionary indices to u32
//   k_pq_pages    every page of the call, one thread per non-NULL value:
//                 PLAIN values copied, dictionary indices gathered
//   (both in tiles of consecutive values whose runs or pages sit in LDS)
//   k_pq_count / k_pq_scan_top / k_pq_scatter
//                 OPTIONAL columns with NULLs: exclusive count of valid rows
//                 per row, then value[row] = dense[count] where valid
#pragma once

namespace gbagg {

// One RLE / bit-packed hybrid run of some column of a decode call. Runs are
// ordered by gstart, their first value in the call's concatenation of values.
struct PqRun {
    uint64_t src;     // bit-packed runs: device address of the first data byte
    uint64_t dst;     // device address of the run's first output element
    uint64_t gstart;
    uint32_t count;
    uint32_t value;   // RLE runs: the repeated value
    uint32_t bw;      // bit-packed runs: bit width (1..32); 0: an RLE run
    uint32_t dwidth;  // output element bytes: 1 (definition levels) or 4 (dictionary indices)
};

// Values are processed in tiles of 256 * items consecutive values per
// workgroup (items = 1..PQ_ITEMS, fewer for small calls so that they still
// spread over the whole chip). Two threads find the tile's first and last run
// (or page) by binary search in the call's table; when the tile spans at most
// PQ_LDS_RUNS of them they are staged in LDS and each value searches there,
// else in the table itself.
constexpr uint32_t PQ_ITEMS = 16;
constexpr uint32_t PQ_LDS_RUNS = 256;

// The last entry of t[lo..hi] whose gstart is <= i (t[lo].gstart <= i).
template <class R>
__device__ inline uint32_t pq_find(const R *t, uint32_t lo, uint32_t hi, uint64_t i) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (t[mid].gstart <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Little-endian loads at any byte address through the aligned words that
// cover it: the word below the address lies in the same 256-byte-aligned
// chunk buffer or file image, and the word above within the 16 readable
// bytes that follow every source range.
__device__ inline uint64_t pq_ld8(const uint8_t *p) {
    const uintptr_t a = (uintptr_t)p;
    const uint64_t *w = (const uint64_t *)(a & ~(uintptr_t)7);
    const uint32_t sh = (uint32_t)(a & 7) * 8;
    const uint64_t lo = w[0];
    return sh ? (lo >> sh) | (w[1] << (64 - sh)) : lo;
}

__device__ inline uint32_t pq_ld4(const uint8_t *p) {
    const uintptr_t a = (uintptr_t)p;
    const uint32_t *w = (const uint32_t *)(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3) * 8;
    const uint32_t lo = w[0];
    return sh ? (lo >> sh) | (w[1] << (32 - sh)) : lo;
}

// Runs the values of one tile [t0, t1) through f(i, entry): the tile's span
// of the table in LDS when it fits.
template <class R, class F>
__device__ inline void pq_tile(const R *__restrict__ t, uint32_t n, uint64_t t0, uint64_t t1, uint32_t items, R *s_t,
                               uint32_t *s_span, F &&f) {
    if (threadIdx.x == 0) s_span[0] = pq_find(t, 0, n - 1, t0);
    if (threadIdx.x == 64) s_span[1] = pq_find(t, 0, n - 1, t1 - 1);
    __syncthreads();
    const uint32_t lo = s_span[0], hi = s_span[1];
    const bool staged = hi - lo < PQ_LDS_RUNS;
    if (staged)
        for (uint32_t r = threadIdx.x; r <= hi - lo; r += blockDim.x) s_t[r] = t[lo + r];
    __syncthreads();
    uint32_t at = 0;  // the run of this thread's previous value: runs only move forward
#pragma unroll 4
    for (uint32_t k = 0; k < items; k++) {
        const uint64_t i = t0 + (uint64_t)k * 256 + threadIdx.x;
        if (i >= t1) break;
        if (staged) {
            at = pq_find(s_t, at, hi - lo, i);
            f(i, s_t[at]);
        } else {
            f(i, t[pq_find(t, lo, hi, i)]);
        }
    }
    __syncthreads();  // s_t and s_span are rewritten by the next tile
}

// Every run of a decode call, one thread per value: the repeated value or the
// bw bits at its bit offset (least-significant bit first; every source range
// carries 16 readable bytes past its end).
__global__ __launch_bounds__(256) void k_pq_hybrid(const PqRun *__restrict__ runs, uint32_t nruns, uint64_t total,
                                                   uint32_t items) {
    __shared__ PqRun s_t[PQ_LDS_RUNS];
    __shared__ uint32_t s_span[2];
    const uint64_t tile = 256ull * items;
    for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < total; t0 += (uint64_t)gridDim.x * tile) {
        const uint64_t t1 = t0 + tile < total ? t0 + tile : total;
        pq_tile(runs, nruns, t0, t1, items, s_t, s_span, [&](uint64_t i, const PqRun &r) {
            const uint64_t j = i - r.gstart;
            uint32_t v = r.value;
            if (r.bw) {
                const uint64_t bit = j * r.bw;
                const uint64_t w = pq_ld8((const uint8_t *)r.src + (bit >> 3));
                const uint64_t mask = r.bw == 32 ? 0xFFFFFFFFull : ((1ull << r.bw) - 1);
                v = (uint32_t)((w >> (bit & 7)) & mask);
            }
            if (r.dwidth == 1) ((uint8_t *)r.dst)[j] = (uint8_t)v;
            else ((uint32_t *)r.dst)[j] = v;
        });
    }
}

// The dictionary indices of one data page: an RLE / bit-packed hybrid stream
// at [src, end) (device addresses) holding `count` values of bw bits
// (1..32), written as u32 to dst. The run headers are read on the device.
struct PqStream {
    uint64_t src;
    uint64_t end;
    uint64_t dst;
    uint32_t count;
    uint32_t bw;
};

constexpr uint32_t PQ_SRUNS = 256;  // runs parsed per batch

// One workgroup per stream: one thread reads up to PQ_SRUNS run headers into
// LDS, then all threads write those runs' values; repeated until the stream
// is done. A header or run that leaves the stream sets bit 4 of *bad and ends
// the stream (the caller reports the error). Bit-packed data are read only
// inside the stream; each source range carries 16 readable bytes past its end.
__global__ __launch_bounds__(256) void k_pq_streams(const PqStream *__restrict__ st, uint32_t ns, uint32_t *bad) {
    __shared__ uint64_t s_src[PQ_SRUNS];      // bit-packed: address of the run's data; RLE: 0
    __shared__ uint32_t s_out[PQ_SRUNS + 1];  // first value of each run; s_out[nr]: the batch's end
    __shared__ uint32_t s_val[PQ_SRUNS];      // RLE: the repeated value
    __shared__ uint64_t s_p;
    __shared__ uint32_t s_nr, s_fail;
    for (uint32_t k = blockIdx.x; k < ns; k += gridDim.x) {
        const PqStream S = st[k];
        const uint64_t mask = S.bw >= 32 ? 0xFFFFFFFFull : ((1ull << S.bw) - 1);
        if (threadIdx.x == 0) s_p = S.src;
        uint32_t done = 0;
        while (done < S.count) {
            if (threadIdx.x == 0) {
                const uint8_t *p = (const uint8_t *)s_p, *end = (const uint8_t *)S.end;
                const uint32_t vbytes = (S.bw + 7) / 8;
                uint32_t nr = 0, d = done, fail = 0;
                while (nr < PQ_SRUNS && d < S.count) {
                    uint64_t h = 0;
                    uint32_t sh = 0;
                    bool whole = false;
                    while (p < end && sh < 35) {
                        const uint8_t b = *p++;
                        h |= (uint64_t)(b & 0x7F) << sh;
                        sh += 7;
                        if (!(b & 0x80)) {
                            whole = true;
                            break;
                        }
                    }
                    if (!whole) {
                        fail = 1;
                        break;
                    }
                    if (h & 1) {  // bit-packed: (h >> 1) groups of 8 values
                        const uint64_t bytes = (h >> 1) * S.bw;
                        if (bytes > (uint64_t)(end - p)) {
                            fail = 1;
                            break;
                        }
                        const uint64_t left = S.count - d;
                        const uint32_t cnt = (uint32_t)((h >> 1) * 8 < left ? (h >> 1) * 8 : left);
                        if (cnt) {
                            s_src[nr] = (uint64_t)p;
                            s_val[nr] = 0;
                            s_out[nr++] = d;
                            d += cnt;
                        }
                        p += bytes;
                    } else {  // RLE: (h >> 1) repeats of one value
                        if ((h >> 1) == 0 || (uint64_t)(end - p) < vbytes) {
                            fail = 1;
                            break;
                        }
                        uint32_t v = 0;
                        for (uint32_t b = 0; b < vbytes; b++) v |= (uint32_t)p[b] << (8 * b);
                        p += vbytes;
                        const uint64_t left = S.count - d;
                        s_src[nr] = 0;
                        s_val[nr] = v;
                        s_out[nr++] = d;
                        d += (uint32_t)((h >> 1) < left ? (h >> 1) : left);
                    }
                }
                s_out[nr] = d;
                s_nr = nr;
                s_p = (uint64_t)p;
                s_fail = fail;
            }
            __syncthreads();
            const uint32_t nr = s_nr, d1 = s_out[nr], fail = s_fail;
            uint32_t at = 0;  // the run of this thread's previous value: runs only move forward
            for (uint32_t i = done + threadIdx.x; i < d1; i += blockDim.x) {
                uint32_t hi = nr - 1;
                while (at < hi) {
                    const uint32_t mid = (at + hi + 1) / 2;
                    if (s_out[mid] <= i) at = mid;
                    else hi = mid - 1;
                }
                uint32_t v = s_val[at];
                if (s_src[at]) {
                    const uint64_t bit = (uint64_t)(i - s_out[at]) * S.bw;
                    v = (uint32_t)((pq_ld8((const uint8_t *)s_src[at] + (bit >> 3)) >> (bit & 7)) & mask);
                }
                ((uint32_t *)S.dst)[i] = v;
            }
            __syncthreads();  // the next batch rewrites the run table
            if (fail) {
                if (threadIdx.x == 0) __hip_atomic_fetch_or(bad, 4u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            done = d1;
        }
    }
}

// Dense (non-NULL) values of the columns of a decode call, by page: PLAIN
// pages are copied, dictionary pages gathered through their dictionary; an
// index outside the dictionary sets *bad.
struct PqSeg {
    uint64_t src;     // PLAIN: the page's values; dictionary: the dictionary's values
    uint64_t idx;     // dictionary: device address of the page's decoded indices
    uint64_t dst;     // device address of the page's first dense value
    uint64_t gstart;  // first value in the call's concatenation
    uint32_t count;
    uint32_t dict_n;  // dictionary entries (0: a PLAIN page)
    uint32_t width;   // value bytes: 4 or 8
    uint32_t pad;     // host planning: 1 while src is a remap entry index
};

__global__ __launch_bounds__(256) void k_pq_pages(const PqSeg *__restrict__ segs, uint32_t nseg, uint64_t total,
                                                  uint32_t items, uint32_t *bad) {
    __shared__ PqSeg s_t[PQ_LDS_RUNS];
    __shared__ uint32_t s_span[2];
    const uint64_t tile = 256ull * items;
    for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < total; t0 += (uint64_t)gridDim.x * tile) {
        const uint64_t t1 = t0 + tile < total ? t0 + tile : total;
        pq_tile(segs, nseg, t0, t1, items, s_t, s_span, [&](uint64_t i, const PqSeg &sg) {
            const uint64_t j = i - sg.gstart;
            uint64_t at = j;
            bool ok = true;
            if (sg.dict_n) {
                at = ((const uint32_t *)sg.idx)[j];
                ok = at < sg.dict_n;
                if (!ok) __hip_atomic_fetch_or(bad, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (sg.width == 4) ((uint32_t *)sg.dst)[j] = ok ? pq_ld4((const uint8_t *)sg.src + at * 4) : 0u;
            else ((uint64_t *)sg.dst)[j] = ok ? pq_ld8((const uint8_t *)sg.src + at * 8) : 0ull;
        });
    }
}

constexpr uint32_t PQ_SCAN_BLOCK = 1024;

// Valid rows per 1024-row block.
__global__ __launch_bounds__(PQ_SCAN_BLOCK) void k_pq_count(const uint8_t *__restrict__ valid, uint32_t n,
                                                           uint32_t *__restrict__ bsum) {
    __shared__ uint32_t s_w[PQ_SCAN_BLOCK / 64];
    const uint32_t i = blockIdx.x * PQ_SCAN_BLOCK + threadIdx.x;
    const uint32_t v = i < n ? valid[i] != 0 : 0;
    const uint32_t c = (uint32_t)__popcll(__ballot(v));
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < PQ_SCAN_BLOCK / 64; w++) t += s_w[w];
        bsum[blockIdx.x] = t;
    }
}

// One block: exclusive prefix of the block sums, in place (nb <= 1024 * 64).
__global__ __launch_bounds__(1024) void k_pq_scan_top(uint32_t *bsum, uint32_t nb) {
    __shared__ uint32_t s_part[1024];
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    uint32_t local = 0;
    for (uint32_t b = b0; b < b1; b++) local += bsum[b];
    s_part[threadIdx.x] = local;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t y = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += y;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t c = bsum[b];
        bsum[b] = run;
        run += c;
    }
}

// out[row] = dense[valid rows before row] where valid, 0 elsewhere.
template <class V>
__global__ __launch_bounds__(PQ_SCAN_BLOCK) void k_pq_scatter(const uint8_t *__restrict__ valid, uint32_t n,
                                                             const uint32_t *__restrict__ bsum, const V *__restrict__ dense,
                                                             uint32_t ndense, V *__restrict__ out) {
    __shared__ uint32_t s_w[PQ_SCAN_BLOCK / 64];
    const uint32_t i = blockIdx.x * PQ_SCAN_BLOCK + threadIdx.x;
    const bool v = i < n && valid[i] != 0;
    const uint64_t m = __ballot(v);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s_w[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = bsum[blockIdx.x];
    for (uint32_t w = 0; w < wv; w++) before += s_w[w];
    before += (uint32_t)__popcll(m & ((1ull << lane) - 1));
    if (i < n) out[i] = (v && before < ndense) ? dense[before] : (V)0;
}

// ---- BYTE_ARRAY columns with PLAIN pages: delivered as strings ------------
// (offsets + bytes, GBAGG_STR_RAW). Per value: its length and source address,
// from a PLAIN page (length-prefixed strings) or from the merged dictionary;
// then offsets by a scan of the lengths, and the bytes packed one after another.

// One PLAIN BYTE_ARRAY page of a decode call.
struct PqStrPage {
    uint64_t src, end;    // the page's values: 4-byte little-endian length, then the bytes
    uint64_t lens, srcs;  // the column's per-value lengths (u32) and source addresses (u64)
    uint32_t dense0, count;
};

// One thread per page follows its length prefixes (a sequential chain);
// a prefix or string past the page end sets bit 2 of *bad.
__global__ __launch_bounds__(64) void k_pq_str_plain(const PqStrPage *__restrict__ pages, uint32_t n, uint32_t *bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PqStrPage pg = pages[i];
    const uint8_t *p = (const uint8_t *)pg.src, *end = (const uint8_t *)pg.end;
    uint32_t *lens = (uint32_t *)pg.lens;
    uint64_t *srcs = (uint64_t *)pg.srcs;
    for (uint32_t j = 0; j < pg.count; j++) {
        bool ok = end - p >= 4;
        uint32_t len = 0;
        if (ok) {
            len = (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
            p += 4;
            ok = len <= (uint64_t)(end - p);
        }
        if (!ok) {
            __hip_atomic_fetch_or(bad, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (; j < pg.count; j++) {
                lens[pg.dense0 + j] = 0;
                srcs[pg.dense0 + j] = pg.src;
            }
            return;
        }
        lens[pg.dense0 + j] = len;
        srcs[pg.dense0 + j] = (uint64_t)p;
        p += len;
    }
}

// Values of dictionary-encoded pages of such a column: their strings in the
// merged dictionary (its device copy).
struct PqStrDict {
    uint64_t ids;            // the column's dense merged dictionary ids (u32)
    uint64_t doff, dbytes;   // the merged dictionary: offsets (u32) and bytes
    uint64_t lens, srcs;
    uint64_t gstart;         // first value in the call's concatenation of such values
    uint32_t dense0, count;
};

__global__ __launch_bounds__(256) void k_pq_str_dict(const PqStrDict *__restrict__ segs, uint32_t nseg,
                                                     uint64_t total) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (segs[mid].gstart <= i) lo = mid;
            else hi = mid - 1;
        }
        const PqStrDict sg = segs[lo];
        const uint64_t d = sg.dense0 + (i - sg.gstart);
        const uint32_t id = ((const uint32_t *)sg.ids)[d];
        const uint32_t *off = (const uint32_t *)sg.doff;
        ((uint32_t *)sg.lens)[d] = off[id + 1] - off[id];
        ((uint64_t *)sg.srcs)[d] = sg.dbytes + off[id];
    }
}

// Sum of the lengths per 1024-value block.
__global__ __launch_bounds__(PQ_SCAN_BLOCK) void k_pq_lsum(const uint32_t *__restrict__ len, uint32_t n,
                                                          uint32_t *__restrict__ bsum) {
    __shared__ uint32_t s_w[PQ_SCAN_BLOCK / 64];
    const uint32_t i = blockIdx.x * PQ_SCAN_BLOCK + threadIdx.x;
    uint32_t x = i < n ? len[i] : 0;
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < PQ_SCAN_BLOCK / 64; w++) t += s_w[w];
        bsum[blockIdx.x] = t;
    }
}

// off[i] = lengths before value i (bsum: exclusive block prefixes); off[n] = total.
__global__ __launch_bounds__(PQ_SCAN_BLOCK) void k_pq_lscan(const uint32_t *__restrict__ len, uint32_t n,
                                                           const uint32_t *__restrict__ bsum,
                                                           uint32_t *__restrict__ off) {
    __shared__ uint32_t s_w[PQ_SCAN_BLOCK / 64];
    const uint32_t i = blockIdx.x * PQ_SCAN_BLOCK + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t v = i < n ? len[i] : 0;
    uint32_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_w[wv] = x;
    __syncthreads();
    uint32_t before = bsum[blockIdx.x];
    for (uint32_t w = 0; w < wv; w++) before += s_w[w];
    if (i < n) {
        off[i] = before + x - v;
        if (i == n - 1) off[n] = before + x;
    }
}

__global__ __launch_bounds__(256) void k_pq_str_copy(const uint64_t *__restrict__ srcs, const uint32_t *__restrict__ len,
                                                     const uint32_t *__restrict__ off, uint32_t n, uint8_t *out) {
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < n; d += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t *s = (const uint8_t *)srcs[d];
        uint8_t *o = out + off[d];
        const uint32_t l = len[d];
        for (uint32_t k = 0; k < l; k++) o[k] = s[k];
    }
}

// OPTIONAL string columns with NULLs: row offsets from the dense offsets
// (a NULL row is an empty string); bsum holds the exclusive valid-row counts
// of the 1024-row blocks.
__global__ __launch_bounds__(PQ_SCAN_BLOCK) void k_pq_rowoff(const uint8_t *__restrict__ valid, uint32_t n,
                                                            const uint32_t *__restrict__ bsum,
                                                            const uint32_t *__restrict__ doff, uint32_t ndense,
                                                            uint32_t *__restrict__ roff) {
    __shared__ uint32_t s_w[PQ_SCAN_BLOCK / 64];
    const uint32_t i = blockIdx.x * PQ_SCAN_BLOCK + threadIdx.x;
    const bool v = i < n && valid[i] != 0;
    const uint64_t m = __ballot(v);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s_w[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = bsum[blockIdx.x];
    for (uint32_t w = 0; w < wv; w++) before += s_w[w];
    before += (uint32_t)__popcll(m & ((1ull << lane) - 1));
    if (i < n) {
        roff[i] = doff[min(before, ndense)];
        if (i == n - 1) roff[n] = doff[ndense];
    }
}

}  // namespace gbagg

namespace {

// ---- Thrift compact protocol (the Parquet footer and page headers) ---------
struct TReader {
    const uint8_t *p, *end;
    uint8_t byte() {
        if (p >= end) fail(GBAGG_INVALID, "parquet: truncated metadata");
        return *p++;
    }
    uint64_t varint() {
        uint64_t v = 0;
        for (int s = 0; s < 64; s += 7) {
            const uint8_t b = byte();
            v |= (uint64_t)(b & 0x7F) << s;
            if (!(b & 0x80)) return v;
        }
        fail(GBAGG_INVALID, "parquet: bad varint");
    }
    int64_t zz() {
        const uint64_t v = varint();
        return (int64_t)(v >> 1) ^ -(int64_t)(v & 1);
    }
    std::string binary() {
        const uint64_t n = varint();
        if (n > (uint64_t)(end - p)) fail(GBAGG_INVALID, "parquet: truncated binary field");
        std::string s((const char *)p, (size_t)n);
        p += n;
        return s;
    }
    void skip(int type, int depth = 0) {
        if (depth > 32) fail(GBAGG_INVALID, "parquet: metadata nested too deeply");
        switch (type) {
        case 1: case 2: return;                  // bool in the field header
        case 3: byte(); return;                  // i8
        case 4: case 5: case 6: varint(); return;  // i16 i32 i64
        case 7: for (int k = 0; k < 8; k++) byte(); return;
        case 8: binary(); return;
        case 9: case 10: {                        // list, set
            const uint8_t h = byte();
            uint64_t n = h >> 4;
            if (n == 15) n = varint();
            const int et = h & 15;
            for (uint64_t k = 0; k < n; k++) {
                if (et == 1 || et == 2) byte();
                else skip(et, depth + 1);
            }
            return;
        }
        case 11: {                                // map
            const uint64_t n = varint();
            if (!n) return;
            const uint8_t kv = byte();
            for (uint64_t k = 0; k < n; k++) {
                skip(kv >> 4, depth + 1);
                skip(kv & 15, depth + 1);
            }
            return;
        }
        case 12: struct_fields([&](int, int t) { skip(t, depth + 1); }); return;
        default: fail(GBAGG_INVALID, "parquet: bad thrift type");
        }
    }
    // Calls f(field id, type) for every field of a struct, through its stop byte.
    template <class F>
    void struct_fields(F &&f) {
        int last = 0;
        for (;;) {
            const uint8_t h = byte();
            if (h == 0) return;
            const int t = h & 15;
            const int d = h >> 4;
            const int id = d ? last + d : (int)zz();
            last = id;
            f(id, t);
        }
    }
    // List header: element count (element type returned in et).
    uint64_t list(int &et) {
        const uint8_t h = byte();
        uint64_t n = h >> 4;
        if (n == 15) n = varint();
        et = h & 15;
        return n;
    }
};

struct PqColumnMeta {
    int32_t type = -1, codec = 0;
    int64_t num_values = 0, total_size = 0, data_off = 0, dict_off = 0;
};
struct PqRowGroup {
    int64_t num_rows = 0;
    std::vector<PqColumnMeta> cols;
};
struct PqSchemaCol {
    std::string name;
    int32_t type = -1;
    int32_t repetition = 0;  // 0 REQUIRED, 1 OPTIONAL, 2 REPEATED
};

struct PqPage {
    int32_t type = -1, comp_size = 0, uncomp_size = 0;
    int32_t num_values = 0, encoding = 0, def_encoding = 3;
    int32_t v2_num_nulls = 0, v2_def_len = 0, v2_rep_len = 0;
    bool v2_compressed = true;
};

PqPage read_page_header(TReader &r) {
    PqPage pg;
    r.struct_fields([&](int id, int t) {
        if (id == 1 && t == 5) pg.type = (int32_t)r.zz();
        else if (id == 2 && t == 5) pg.uncomp_size = (int32_t)r.zz();
        else if (id == 3 && t == 5) pg.comp_size = (int32_t)r.zz();
        else if ((id == 5 || id == 7) && t == 12) {  // data page / dictionary page header
            r.struct_fields([&](int f, int ft) {
                if (f == 1 && ft == 5) pg.num_values = (int32_t)r.zz();
                else if (f == 2 && ft == 5) pg.encoding = (int32_t)r.zz();
                else if (id == 5 && f == 3 && ft == 5) pg.def_encoding = (int32_t)r.zz();
                else r.skip(ft);
            });
        } else if (id == 8 && t == 12) {  // data page header v2
            r.struct_fields([&](int f, int ft) {
                if (f == 1 && ft == 5) pg.num_values = (int32_t)r.zz();
                else if (f == 2 && ft == 5) pg.v2_num_nulls = (int32_t)r.zz();
                else if (f == 4 && ft == 5) pg.encoding = (int32_t)r.zz();
                else if (f == 5 && ft == 5) pg.v2_def_len = (int32_t)r.zz();
                else if (f == 6 && ft == 5) pg.v2_rep_len = (int32_t)r.zz();
                else if (f == 7 && (ft == 1 || ft == 2)) pg.v2_compressed = ft == 1;
                else r.skip(ft);
            });
        } else {
            r.skip(t);
        }
    });
    return pg;
}

// Run headers of an RLE / bit-packed hybrid stream of `n` values at
// [p, end) of the file, positions relative to `chunk0`. Returns the number of
// ones among the values when `count_ones` (definition levels, bw = 1).
// Run fields: src = byte offset from chunk0 (rebased by the caller), dst =
// the output element index out0 + k (the caller turns it into an address).
uint64_t hybrid_runs(const uint8_t *p, const uint8_t *end, const uint8_t *chunk0, uint32_t bw, uint32_t n,
                     uint32_t out0, std::vector<PqRun> &runs, bool count_ones) {
    if (bw > 32) fail(GBAGG_INVALID, "parquet: bit width above 32");
    TReader r{p, end};
    uint32_t done = 0;
    uint64_t ones = 0;
    const uint32_t vbytes = (bw + 7) / 8;
    while (done < n) {
        const uint64_t h = r.varint();
        if (h & 1) {  // bit-packed: (h >> 1) groups of 8 values
            const uint64_t groups = h >> 1;
            const uint64_t bytes = groups * bw;
            if (bytes > (uint64_t)(r.end - r.p)) fail(GBAGG_INVALID, "parquet: truncated bit-packed run");
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(groups * 8, n - done);
            if (cnt) runs.push_back(PqRun{bw ? (uint64_t)(r.p - chunk0) : 0, (uint64_t)(out0 + done), 0, cnt, 0, bw, 0});
            if (count_ones) {  // bw = 1: the first cnt bits, least-significant first, 64 at a time
                uint32_t k = 0;
                for (; k + 64 <= cnt; k += 64) {
                    uint64_t w;
                    std::memcpy(&w, r.p + k / 8, 8);
                    ones += (uint64_t)__builtin_popcountll(w);
                }
                for (; k + 8 <= cnt; k += 8) ones += (uint64_t)__builtin_popcount(r.p[k / 8]);
                if (k < cnt) ones += (uint64_t)__builtin_popcount(r.p[k / 8] & ((1u << (cnt - k)) - 1));
            }
            r.p += bytes;
            done += cnt;
        } else {  // RLE: (h >> 1) repeats of one value
            const uint64_t rep = h >> 1;
            uint32_t v = 0;
            for (uint32_t b = 0; b < vbytes; b++) v |= (uint32_t)r.byte() << (8 * b);
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(rep, n - done);
            if (cnt) runs.push_back(PqRun{0, (uint64_t)(out0 + done), 0, cnt, v, 0, 0});
            if (count_ones && v) ones += cnt;
            done += cnt;
            if (rep == 0) fail(GBAGG_INVALID, "parquet: empty RLE run");
        }
    }
    return ones;
}

struct PqCall;

}  // namespace

struct gbagg_parquet {
    const uint8_t *file = nullptr;
    uint64_t size = 0;
    std::vector<PqSchemaCol> cols;
    std::vector<PqRowGroup> rgs;
    int64_t num_rows = 0;
    hipStream_t stream = nullptr;
    // one decoded column: device values (and validity), its host dictionary;
    // kept until close so that an operator can consume several row groups
    struct Out {
        uint8_t *data = nullptr, *dense = nullptr, *valid = nullptr, *idx = nullptr, *bsum = nullptr;  // in the arena
        // BYTE_ARRAY delivered as strings: per-value lengths and sources,
        // offsets (dense, then per row), packed bytes, the merged dictionary
        uint8_t *slen = nullptr, *ssrc = nullptr, *soff = nullptr, *sbsum = nullptr, *sbytes = nullptr;
        uint8_t *roff = nullptr, *ddoff = nullptr, *ddbytes = nullptr;
        std::vector<uint32_t> remap;  // VARCHAR: host side of the remap tables' async copy
        std::vector<uint32_t> dict_off;
        std::vector<char> dict_bytes;
    };
    std::vector<std::unique_ptr<Out>> outs;
    DevBuf image;           // gbagg_parquet_upload: the whole file in HBM (+16 bytes)
    bool resident = false;
    DevBuf chunks, bad;     // per decode call: chunk bytes when the file is not resident
    // a launch's runs and pages go to the device in one copy from pinned
    // staging; two staging buffers alternate, each reused once the event
    // recorded after its copy has passed (the device tables themselves are
    // arena memory, never reused before reset)
    uint8_t *htab[2] = {nullptr, nullptr};
    size_t htab_bytes[2] = {0, 0};
    hipEvent_t htab_done[2] = {nullptr, nullptr};
    int htab_next = 0;
    uint32_t *hbad = nullptr;  // pinned host copy of the error bits, read after a wait
    // gbagg_parquet_decode_async: row groups planned into one pending call,
    // launched once flush_rows rows are pending and at gbagg_parquet_sync;
    // the error bits collect every launch since the last wait (armed)
    std::unique_ptr<PqCall> pending;
    uint64_t pending_rows = 0;
    uint64_t flush_rows = 4u << 20;
    bool armed = false;
    // decoded columns are bump-allocated from arenas of 256 MB or more (one
    // hipMalloc per arena, not per column), all kept until close
    // (gbagg_parquet_reset forgets them and reuses the arenas)
    std::vector<DevBuf> arenas;
    size_t arena = 0, arena_used = 0;  // current arena, bytes used in it
    uint8_t *alloc(size_t n) {
        n = (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
        while (arena < arenas.size() && arena_used + n > arenas[arena].bytes) {
            arena++;
            arena_used = 0;
        }
        if (arena == arenas.size()) {
            arenas.emplace_back();
            arenas.back().ensure(std::max<size_t>(n, (size_t)256 << 20));
            arena_used = 0;
        }
        uint8_t *p = arenas[arena].as<uint8_t>() + arena_used;
        arena_used += n;
        return p;
    }
    ~gbagg_parquet();
};

namespace {

void parquet_open(gbagg_parquet *pq, const uint8_t *file, uint64_t size) {
    if (!file || size < 12) fail(GBAGG_INVALID, "parquet: file too small");
    if (std::memcmp(file, "PAR1", 4) || std::memcmp(file + size - 4, "PAR1", 4))
        fail(GBAGG_INVALID, "parquet: missing PAR1 magic");
    uint32_t flen;
    std::memcpy(&flen, file + size - 8, 4);
    if ((uint64_t)flen + 12 > size) fail(GBAGG_INVALID, "parquet: bad footer length");
    pq->file = file;
    pq->size = size;
    TReader r{file + size - 8 - flen, file + size - 8};
    std::vector<PqSchemaCol> schema;
    r.struct_fields([&](int id, int t) {
        if (id == 2 && t == 9) {  // schema: list<SchemaElement>
            int et;
            const uint64_t n = r.list(et);
            for (uint64_t k = 0; k < n; k++) {
                PqSchemaCol c;
                int32_t children = 0;
                r.struct_fields([&](int f, int ft) {
                    if (f == 1 && ft == 5) c.type = (int32_t)r.zz();
                    else if (f == 3 && ft == 5) c.repetition = (int32_t)r.zz();
                    else if (f == 4 && ft == 8) c.name = r.binary();
                    else if (f == 5 && ft == 5) children = (int32_t)r.zz();
                    else r.skip(ft);
                });
                if (k == 0) continue;  // the root
                if (children) fail(GBAGG_INVALID, "parquet: nested schemas are not supported");
                schema.push_back(c);
            }
        } else if (id == 3 && t == 6) {
            pq->num_rows = r.zz();
        } else if (id == 4 && t == 9) {  // row_groups: list<RowGroup>
            int et;
            const uint64_t n = r.list(et);
            for (uint64_t k = 0; k < n; k++) {
                PqRowGroup g;
                r.struct_fields([&](int f, int ft) {
                    if (f == 1 && ft == 9) {  // columns: list<ColumnChunk>
                        int cet;
                        const uint64_t nc = r.list(cet);
                        for (uint64_t c = 0; c < nc; c++) {
                            PqColumnMeta m;
                            r.struct_fields([&](int cf, int cft) {
                                if (cf == 3 && cft == 12) {  // ColumnMetaData
                                    r.struct_fields([&](int mf, int mft) {
                                        if (mf == 1 && mft == 5) m.type = (int32_t)r.zz();
                                        else if (mf == 4 && mft == 5) m.codec = (int32_t)r.zz();
                                        else if (mf == 5 && mft == 6) m.num_values = r.zz();
                                        else if (mf == 7 && mft == 6) m.total_size = r.zz();
                                        else if (mf == 9 && mft == 6) m.data_off = r.zz();
                                        else if (mf == 11 && mft == 6) m.dict_off = r.zz();
                                        else r.skip(mft);
                                    });
                                } else {
                                    r.skip(cft);
                                }
                            });
                            g.cols.push_back(m);
                        }
                    } else if (f == 3 && ft == 6) {
                        g.num_rows = r.zz();
                    } else {
                        r.skip(ft);
                    }
                });
                pq->rgs.push_back(std::move(g));
            }
        } else {
            r.skip(t);
        }
    });
    pq->cols = std::move(schema);
    for (const auto &g : pq->rgs)
        if (g.cols.size() != pq->cols.size()) fail(GBAGG_INVALID, "parquet: row group column count differs from the schema");
}

gbagg_type parquet_type(int32_t physical) {
    switch (physical) {
    case 1: return GBAGG_TYPE_INTEGER;   // INT32
    case 2: return GBAGG_TYPE_BIGINT;    // INT64
    case 5: return GBAGG_TYPE_DOUBLE;    // DOUBLE
    case 6: return GBAGG_TYPE_VARCHAR;   // BYTE_ARRAY (dictionary codes)
    default: fail(GBAGG_INVALID, "parquet: physical type not supported (INT32, INT64, DOUBLE, BYTE_ARRAY)");
    }
}

// Chunk bytes [start, start + size) of column c of row group g.
void parquet_chunk_range(const gbagg_parquet *pq, uint32_t g, uint32_t c, int64_t &start, int64_t &size) {
    const PqColumnMeta &m = pq->rgs[g].cols[c];
    start = m.dict_off > 0 ? std::min(m.dict_off, m.data_off) : m.data_off;
    size = m.total_size;
    if (start < 4 || size < 0 || (uint64_t)(start + size) > pq->size - 8)
        fail(GBAGG_INVALID, "parquet: column chunk outside the file");
}

// The runs and pages of every column of one decode call, launched together.
struct PqCall {
    std::vector<PqRun> runs;
    std::vector<PqStream> streams;  // dictionary-index streams whose run headers the device reads
    // the device reads index run headers only in calls over many column chunks:
    // one stream is a chain of dependent reads (tens of microseconds), cheap
    // only when many run side by side; small calls read them on the host
    bool dev_streams = false;
    std::vector<PqSeg> segs;
    uint64_t nvals = 0, ndense = 0;
    struct Nulls {
        gbagg_parquet::Out *o;
        uint32_t nrows, dense, width;
    };
    std::vector<Nulls> nulls;  // OPTIONAL columns with a NULL in this chunk
    // BYTE_ARRAY columns delivered as strings
    std::vector<PqStrPage> spages;
    std::vector<PqStrDict> sdict;
    uint64_t nsdict = 0;
    struct Strs {
        gbagg_parquet::Out *o;
        uint32_t rows, dense;
        bool nulls;
    };
    std::vector<Strs> strs;
};

// Appends call b (planned on its own) to call a: b's value indexes follow a's.
void call_append(PqCall &a, PqCall &&b) {
    for (PqRun r : b.runs) {
        r.gstart += a.nvals;
        a.runs.push_back(r);
    }
    for (PqSeg g : b.segs) {
        g.gstart += a.ndense;
        a.segs.push_back(g);
    }
    for (PqStrDict d : b.sdict) {
        d.gstart += a.nsdict;
        a.sdict.push_back(d);
    }
    a.nvals += b.nvals;
    a.ndense += b.ndense;
    a.nsdict += b.nsdict;
    a.streams.insert(a.streams.end(), b.streams.begin(), b.streams.end());
    a.spages.insert(a.spages.end(), b.spages.begin(), b.spages.end());
    a.nulls.insert(a.nulls.end(), b.nulls.begin(), b.nulls.end());
    a.strs.insert(a.strs.end(), b.strs.begin(), b.strs.end());
}

}  // namespace

gbagg_parquet::~gbagg_parquet() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (int k = 0; k < 2; k++) {
        if (htab_done[k]) (void)hipEventDestroy(htab_done[k]);
        if (htab[k]) (void)hipHostFree(htab[k]);
    }
    if (hbad) (void)hipHostFree(hbad);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Host pass over column c of row groups [g0, g1), whose chunk bytes are on
// the device at dbufs[g - g0] (16 readable bytes past each): o's buffers (the
// row groups concatenated), the column's runs and pages appended to `call`,
// and the view. VARCHAR dictionaries merge into one dictionary of distinct
// strings; each row group's indices are gathered through its remap table.
void parquet_plan_column(gbagg_parquet *pq, uint32_t g0, uint32_t g1, uint32_t c, const uint8_t *const *dbufs,
                         gbagg_parquet::Out &o, gbagg_column_view &view, PqCall &call) {
    const PqSchemaCol &sc = pq->cols[c];
    if (sc.repetition == 2) fail(GBAGG_INVALID, "parquet: repeated columns are not supported");
    const gbagg_type type = parquet_type(sc.type);
    const uint32_t width = type == GBAGG_TYPE_INTEGER || type == GBAGG_TYPE_VARCHAR ? 4 : 8;
    const bool optional = sc.repetition == 1;
    const bool varchar = type == GBAGG_TYPE_VARCHAR;

    // this column's runs and pages go straight into the call's tables, with
    // dst (and idx) as row / dense indices until the buffers exist
    std::vector<PqRun> def_runs;            // levels: appended after the index runs once NULLs are known
    std::vector<PqRun> &idx_runs = call.runs;
    std::vector<PqSeg> &segs = call.segs;
    const size_t irun0 = call.runs.size(), seg0 = call.segs.size(), str0 = call.streams.size();
    std::vector<uint32_t> remap;            // VARCHAR: dictionary code -> merged id, row group after row group
    std::vector<PqStrPage> plain_pages;     // VARCHAR: PLAIN pages (the column is then delivered as strings)
    std::vector<PqStrDict> dict_pages;      // VARCHAR: dictionary-encoded pages, for that case
    uint64_t plain_bytes = 0, dict_vals = 0;
    std::unordered_map<std::string, uint32_t> ids;
    o.dict_off.assign(1, 0);
    o.dict_bytes.clear();
    uint64_t rows = 0, dense = 0;
    for (uint32_t g = g0; g < g1; g++) {
        const PqRowGroup &rg = pq->rgs[g];
        const PqColumnMeta &m = rg.cols[c];
        if (m.codec != 0) fail(GBAGG_INVALID, "parquet: compressed column chunks are not decoded on the GPU (codec " +
                                                  std::to_string(m.codec) + ")");
        if (m.type != sc.type) fail(GBAGG_INVALID, "parquet: column chunk type differs from the schema");
        int64_t start, csize;
        parquet_chunk_range(pq, g, c, start, csize);
        if (rg.num_rows < 0 || rows + (uint64_t)rg.num_rows > 0x7FFFFFFFull)
            fail(GBAGG_INVALID, "parquet: more than 2^31 rows in one decode");
        const uint32_t nrows = (uint32_t)rg.num_rows;
        const uint8_t *c0 = pq->file + start, *cend = c0 + csize;
        const uint8_t *dbuf = dbufs[g - g0];
        bool have_dict = false;
        uint64_t dict_src = 0;  // numeric: device address of the values; VARCHAR: first remap entry
        uint32_t dict_n = 0;
        uint32_t rgrows = 0;
        const uint8_t *p = c0;
        while (p < cend && rgrows < nrows) {
            TReader hr{p, cend};
            const PqPage pg = read_page_header(hr);
            const uint8_t *body = hr.p;
            if (pg.comp_size < 0 || pg.comp_size > cend - body) fail(GBAGG_INVALID, "parquet: page outside its chunk");
            const uint8_t *bend = body + pg.comp_size;
            p = bend;
            if (pg.type == 2) {  // dictionary page: PLAIN values
                if (pg.encoding != 0 && pg.encoding != 2) fail(GBAGG_INVALID, "parquet: dictionary page not PLAIN");
                if (pg.num_values < 0) fail(GBAGG_INVALID, "parquet: bad dictionary size");
                have_dict = true;
                dict_n = (uint32_t)pg.num_values;
                if (varchar) {
                    dict_src = remap.size();
                    const uint8_t *q = body;
                    for (uint32_t k = 0; k < dict_n; k++) {
                        if (bend - q < 4) fail(GBAGG_INVALID, "parquet: truncated dictionary");
                        uint32_t len;
                        std::memcpy(&len, q, 4);
                        q += 4;
                        if (len > (uint64_t)(bend - q)) fail(GBAGG_INVALID, "parquet: truncated dictionary string");
                        std::string str((const char *)q, len);
                        auto it = ids.find(str);
                        if (it == ids.end()) {
                            it = ids.emplace(std::move(str), (uint32_t)(o.dict_off.size() - 1)).first;
                            o.dict_bytes.insert(o.dict_bytes.end(), q, q + len);
                            o.dict_off.push_back((uint32_t)o.dict_bytes.size());
                        }
                        remap.push_back(it->second);
                        q += len;
                    }
                } else {
                    if ((uint64_t)dict_n * width > (uint64_t)pg.comp_size) fail(GBAGG_INVALID, "parquet: truncated dictionary");
                    dict_src = (uint64_t)(dbuf + (body - c0));
                }
                continue;
            }
            if (pg.type != 0 && pg.type != 3) continue;  // index pages
            if (pg.type == 3 && pg.v2_compressed && m.codec != 0) fail(GBAGG_INVALID, "parquet: compressed V2 page");
            if (pg.num_values < 0 || (uint32_t)pg.num_values > nrows - rgrows)
                fail(GBAGG_INVALID, "parquet: more values than the row group's rows");
            const uint32_t n = (uint32_t)pg.num_values;
            const uint8_t *vals = body;
            uint32_t nvalid = n;
            if (optional) {
                const uint8_t *lv, *lend;
                if (pg.type == 0) {  // V1: 4-byte length, then the hybrid levels
                    if (pg.def_encoding != 3) fail(GBAGG_INVALID, "parquet: definition levels not RLE");
                    if (bend - body < 4) fail(GBAGG_INVALID, "parquet: truncated levels");
                    uint32_t len;
                    std::memcpy(&len, body, 4);
                    lv = body + 4;
                    if (len > (uint64_t)(bend - lv)) fail(GBAGG_INVALID, "parquet: truncated levels");
                    lend = lv + len;
                } else {
                    if (pg.v2_rep_len) fail(GBAGG_INVALID, "parquet: repetition levels are not supported");
                    lv = body;
                    if (pg.v2_def_len < 0 || pg.v2_def_len > bend - body) fail(GBAGG_INVALID, "parquet: truncated levels");
                    lend = body + pg.v2_def_len;
                }
                const size_t r0 = def_runs.size();
                nvalid = (uint32_t)hybrid_runs(lv, lend, c0, 1, n, (uint32_t)(rows + rgrows), def_runs, true);
                for (size_t k = r0; k < def_runs.size(); k++)
                    if (def_runs[k].bw) def_runs[k].src = (uint64_t)(dbuf + def_runs[k].src);
                vals = lend;
            } else if (pg.type == 3) {
                if (pg.v2_def_len < 0 || pg.v2_rep_len < 0 || (int64_t)pg.v2_def_len + pg.v2_rep_len > bend - body)
                    fail(GBAGG_INVALID, "parquet: truncated levels");
                vals = body + pg.v2_def_len + pg.v2_rep_len;
            }
            if (pg.encoding == 0 && varchar) {  // PLAIN strings: walked on the device
                if (nvalid)
                    plain_pages.push_back(PqStrPage{(uint64_t)(dbuf + (vals - c0)), (uint64_t)(dbuf + (bend - c0)), 0, 0,
                                                    (uint32_t)dense, nvalid});
                plain_bytes += (uint64_t)(bend - vals);
            } else if (pg.encoding == 0) {  // PLAIN
                if ((uint64_t)nvalid * width > (uint64_t)(bend - vals)) fail(GBAGG_INVALID, "parquet: truncated values");
                if (nvalid) segs.push_back(PqSeg{(uint64_t)(dbuf + (vals - c0)), 0, dense, 0, nvalid, 0, width, 0});
            } else if (pg.encoding == 2 || pg.encoding == 8) {  // PLAIN_DICTIONARY / RLE_DICTIONARY
                if (!have_dict) fail(GBAGG_INVALID, "parquet: dictionary-encoded page before the dictionary");
                if (bend - vals < 1) fail(GBAGG_INVALID, "parquet: truncated indices");
                const uint32_t bw = vals[0];
                if (bw > 32) fail(GBAGG_INVALID, "parquet: index bit width above 32");
                if (nvalid) {
                    if (!dict_n) fail(GBAGG_INVALID, "parquet: dictionary-encoded values with an empty dictionary");
                    // bit width 0: every index is 0 (one RLE run); otherwise the
                    // device reads the stream's run headers (dst: a dense index for
                    // now), or the host does in small calls
                    if (bw == 0) {
                        idx_runs.push_back(PqRun{0, dense, 0, nvalid, 0, 0, 4});
                    } else if (call.dev_streams) {
                        call.streams.push_back(PqStream{(uint64_t)(dbuf + (vals + 1 - c0)),
                                                        (uint64_t)(dbuf + (bend - c0)), dense, nvalid, bw});
                    } else {
                        const size_t r0 = idx_runs.size();
                        hybrid_runs(vals + 1, bend, c0, bw, nvalid, (uint32_t)dense, idx_runs, false);
                        for (size_t k = r0; k < idx_runs.size(); k++)
                            if (idx_runs[k].bw) idx_runs[k].src = (uint64_t)(dbuf + idx_runs[k].src);
                    }
                    // VARCHAR: src is the row group's first remap entry, made an address below
                    segs.push_back(PqSeg{dict_src, dense, dense, 0, nvalid, dict_n, width, varchar ? 1u : 0u});
                    if (varchar) {
                        dict_pages.push_back(PqStrDict{0, 0, 0, 0, 0, 0, (uint32_t)dense, nvalid});
                        dict_vals += nvalid;
                    }
                }
            } else {
                fail(GBAGG_INVALID, "parquet: value encoding not supported (PLAIN, PLAIN_DICTIONARY, RLE_DICTIONARY)");
            }
            rgrows += n;
            dense += nvalid;
        }
        if (rgrows != nrows) fail(GBAGG_INVALID, "parquet: column chunk holds fewer values than the row group");
        rows += nrows;
    }

    // buffers (row groups concatenated), then device addresses in the tables
    o.dense = pq->alloc((size_t)dense * width);
    uint32_t *idx_p = call.runs.size() == irun0 && call.streams.size() == str0
                          ? nullptr
                          : (uint32_t *)(o.idx = pq->alloc((size_t)dense * 4));
    for (size_t k = str0; k < call.streams.size(); k++) call.streams[k].dst = (uint64_t)(idx_p + call.streams[k].dst);
    const uint32_t *remap_p = nullptr;
    if (!remap.empty()) {
        uint8_t *rp = pq->alloc(remap.size() * 4);
        o.remap = std::move(remap);
        HIPCHK(hipMemcpyAsync(rp, o.remap.data(), o.remap.size() * 4, hipMemcpyHostToDevice, pq->stream));
        remap_p = (const uint32_t *)rp;
    }
    for (size_t k = irun0; k < call.runs.size(); k++) {
        PqRun &r = call.runs[k];
        r.dst = (uint64_t)(idx_p + r.dst);
        r.dwidth = 4;
        r.gstart = call.nvals;
        call.nvals += r.count;
    }
    for (size_t k = seg0; k < call.segs.size(); k++) {
        PqSeg &sg = call.segs[k];
        const uint64_t d0 = sg.dst;
        if (sg.pad) sg.src = (uint64_t)(remap_p + sg.src);  // VARCHAR: through the remap table
        sg.pad = 0;
        sg.idx = sg.dict_n ? (uint64_t)(idx_p + sg.idx) : 0;
        sg.dst = (uint64_t)(o.dense + d0 * width);
        sg.gstart = call.ndense;
        call.ndense += sg.count;
    }
    view = gbagg_column_view{};
    view.type = type;
    if (varchar && !plain_pages.empty()) {
        // strings: lengths and sources per value, offsets, packed bytes
        uint32_t maxlen = 0;
        for (size_t k = 0; k + 1 < o.dict_off.size(); k++) maxlen = std::max(maxlen, o.dict_off[k + 1] - o.dict_off[k]);
        const uint64_t bound = plain_bytes + dict_vals * maxlen;
        if (bound > 0xFFFFFFFFull) fail(GBAGG_INVALID, "parquet: more than 4 GiB of strings in one decode");
        if (dense > (uint64_t)PQ_SCAN_BLOCK * 1024 * 64) fail(GBAGG_INVALID, "parquet: row groups too large");
        o.slen = pq->alloc((size_t)dense * 4);
        o.ssrc = pq->alloc((size_t)dense * 8);
        o.soff = pq->alloc((size_t)(dense + 1) * 4);
        o.sbsum = pq->alloc((size_t)((dense + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK) * 4);
        o.sbytes = pq->alloc((size_t)bound + 16);
        if (!dict_pages.empty()) {
            o.ddoff = pq->alloc(o.dict_off.size() * 4);
            o.ddbytes = pq->alloc(o.dict_bytes.size() + 16);
            HIPCHK(hipMemcpyAsync(o.ddoff, o.dict_off.data(), o.dict_off.size() * 4, hipMemcpyHostToDevice, pq->stream));
            if (!o.dict_bytes.empty())
                HIPCHK(hipMemcpyAsync(o.ddbytes, o.dict_bytes.data(), o.dict_bytes.size(), hipMemcpyHostToDevice,
                                      pq->stream));
        }
        for (PqStrPage pp : plain_pages) {
            pp.lens = (uint64_t)o.slen;
            pp.srcs = (uint64_t)o.ssrc;
            call.spages.push_back(pp);
        }
        for (PqStrDict ds : dict_pages) {
            ds.ids = (uint64_t)o.dense;
            ds.doff = (uint64_t)o.ddoff;
            ds.dbytes = (uint64_t)o.ddbytes;
            ds.lens = (uint64_t)o.slen;
            ds.srcs = (uint64_t)o.ssrc;
            ds.gstart = call.nsdict;
            call.nsdict += ds.count;
            call.sdict.push_back(ds);
        }
        const bool nulls = optional && dense < rows;
        if (nulls) {
            o.valid = pq->alloc(rows);
            o.bsum = pq->alloc((size_t)((rows + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK) * 4);
            o.roff = pq->alloc((size_t)(rows + 1) * 4);
            for (PqRun r : def_runs) {
                r.dst = (uint64_t)(o.valid + r.dst);
                r.dwidth = 1;
                r.gstart = call.nvals;
                call.nvals += r.count;
                call.runs.push_back(r);
            }
        }
        call.strs.push_back(PqCall::Strs{&o, (uint32_t)rows, (uint32_t)dense, nulls});
        view.data = nulls ? o.roff : o.soff;
        view.validity = nulls ? o.valid : nullptr;
        view.dict_bytes = (const char *)o.sbytes;
        view.flags = GBAGG_STR_RAW;
        return;
    }
    if (optional && dense < rows) {
        o.valid = pq->alloc(rows);
        o.data = pq->alloc((size_t)rows * width);
        o.bsum = pq->alloc((size_t)((rows + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK) * 4);
        for (PqRun r : def_runs) {
            r.dst = (uint64_t)(o.valid + r.dst);
            r.dwidth = 1;
            r.gstart = call.nvals;
            call.nvals += r.count;
            call.runs.push_back(r);
        }
        call.nulls.push_back(PqCall::Nulls{&o, (uint32_t)rows, (uint32_t)dense, width});
        view.data = o.data;
        view.validity = o.valid;
    } else {
        // REQUIRED, or OPTIONAL without a NULL in these chunks: the dense values are the rows
        view.data = o.dense;
        view.validity = nullptr;
    }
    if (varchar) {
        if (o.dict_bytes.empty()) o.dict_bytes.push_back(0);
        view.dict_offsets = o.dict_off.data();
        view.dict_bytes = o.dict_bytes.data();
        view.dict_size = (uint32_t)(o.dict_off.size() - 1);
        view.flags = 0;  // distinct strings, but one dictionary per decode call (not shared across batches)
    }
}

// The call's kernels: runs, then pages, then the NULL scatters; no wait.
void parquet_launch(gbagg_parquet *pq, PqCall &call) {
    hipStream_t s = pq->stream;
    const uint32_t cap = (uint32_t)device_info().cus * 8;
    // every table in one pinned staging buffer, one copy
    const size_t rb = call.runs.size() * sizeof(PqRun), sb = call.segs.size() * sizeof(PqSeg);
    const size_t pb = call.spages.size() * sizeof(PqStrPage), db = call.sdict.size() * sizeof(PqStrDict);
    const size_t qb = call.streams.size() * sizeof(PqStream);
    const size_t tb = rb + sb + pb + db + qb;
    uint8_t *dtab = nullptr;
    if (tb) {
        const int k = pq->htab_next;
        pq->htab_next ^= 1;
        if (pq->htab_done[k]) HIPCHK(hipEventSynchronize(pq->htab_done[k]));  // its last copy has been read
        else HIPCHK(hipEventCreateWithFlags(&pq->htab_done[k], hipEventDisableTiming));
        if (pq->htab_bytes[k] < tb) {
            if (pq->htab[k]) HIPCHK(hipHostFree(pq->htab[k]));
            pq->htab[k] = nullptr;
            pq->htab_bytes[k] = 0;
            HIPCHK(hipHostMalloc((void **)&pq->htab[k], tb, hipHostMallocDefault));
            pq->htab_bytes[k] = tb;
        }
        uint8_t *h = pq->htab[k];
        if (rb) std::memcpy(h, call.runs.data(), rb);
        if (sb) std::memcpy(h + rb, call.segs.data(), sb);
        if (pb) std::memcpy(h + rb + sb, call.spages.data(), pb);
        if (db) std::memcpy(h + rb + sb + pb, call.sdict.data(), db);
        if (qb) std::memcpy(h + rb + sb + pb + db, call.streams.data(), qb);
        dtab = pq->alloc(tb);
        HIPCHK(hipMemcpyAsync(dtab, h, tb, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(pq->htab_done[k], s));
    }
    // values per thread per tile: enough tiles for `cap` workgroups, at most PQ_ITEMS
    auto items_for = [&](uint64_t n) {
        return (uint32_t)std::min<uint64_t>(PQ_ITEMS, std::max<uint64_t>(1, n / ((uint64_t)cap * 256)));
    };
    if (!call.streams.empty())
        hipLaunchKernelGGL(k_pq_streams, dim3((uint32_t)std::min<size_t>(call.streams.size(), cap)), dim3(256), 0, s,
                           (const PqStream *)(dtab + rb + sb + pb + db),
                           (uint32_t)call.streams.size(), pq->bad.as<uint32_t>());
    if (!call.runs.empty()) {
        const uint32_t items = items_for(call.nvals);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((call.nvals + 256ull * items - 1) / (256ull * items), cap);
        hipLaunchKernelGGL(k_pq_hybrid, dim3(grid), dim3(256), 0, s, (const PqRun *)dtab,
                           (uint32_t)call.runs.size(), call.nvals, items);
    }
    if (!call.segs.empty()) {
        const uint32_t items = items_for(call.ndense);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((call.ndense + 256ull * items - 1) / (256ull * items), cap);
        hipLaunchKernelGGL(k_pq_pages, dim3(grid), dim3(256), 0, s, (const PqSeg *)(dtab + rb),
                           (uint32_t)call.segs.size(), call.ndense, items, pq->bad.as<uint32_t>());
    }
    if (!call.spages.empty())
        hipLaunchKernelGGL(k_pq_str_plain, dim3((uint32_t)((call.spages.size() + 63) / 64)), dim3(64), 0, s,
                           (const PqStrPage *)(dtab + rb + sb), (uint32_t)call.spages.size(),
                           pq->bad.as<uint32_t>());
    if (!call.sdict.empty()) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>((call.nsdict + 255) / 256, cap);
        hipLaunchKernelGGL(k_pq_str_dict, dim3(grid), dim3(256), 0, s,
                           (const PqStrDict *)(dtab + rb + sb + pb), (uint32_t)call.sdict.size(),
                           call.nsdict);
    }
    for (const auto &st : call.strs) {
        gbagg_parquet::Out &o = *st.o;
        if (st.dense) {
            const uint32_t nb = (st.dense + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK;
            hipLaunchKernelGGL(k_pq_lsum, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint32_t *)o.slen, st.dense,
                               (uint32_t *)o.sbsum);
            hipLaunchKernelGGL(k_pq_scan_top, dim3(1), dim3(1024), 0, s, (uint32_t *)o.sbsum, nb);
            hipLaunchKernelGGL(k_pq_lscan, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint32_t *)o.slen, st.dense,
                               (const uint32_t *)o.sbsum, (uint32_t *)o.soff);
            const uint32_t grid = (uint32_t)std::min<uint64_t>((st.dense + 255) / 256, cap);
            hipLaunchKernelGGL(k_pq_str_copy, dim3(grid), dim3(256), 0, s, (const uint64_t *)o.ssrc,
                               (const uint32_t *)o.slen, (const uint32_t *)o.soff, st.dense, o.sbytes);
        } else {
            HIPCHK(hipMemsetAsync(o.soff, 0, 4, s));
        }
        if (st.nulls) {
            const uint32_t nb = (st.rows + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK;
            if (nb > 1024u * 64) fail(GBAGG_INVALID, "parquet: row group too large");
            hipLaunchKernelGGL(k_pq_count, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid, st.rows,
                               (uint32_t *)o.bsum);
            hipLaunchKernelGGL(k_pq_scan_top, dim3(1), dim3(1024), 0, s, (uint32_t *)o.bsum, nb);
            hipLaunchKernelGGL(k_pq_rowoff, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid, st.rows,
                               (const uint32_t *)o.bsum, (const uint32_t *)o.soff, st.dense, (uint32_t *)o.roff);
        }
    }
    for (const auto &nl : call.nulls) {
        gbagg_parquet::Out &o = *nl.o;
        const uint32_t nb = (nl.nrows + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK;
        if (nb > 1024u * 64) fail(GBAGG_INVALID, "parquet: row group too large");
        hipLaunchKernelGGL(k_pq_count, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid, nl.nrows,
                           (uint32_t *)o.bsum);
        hipLaunchKernelGGL(k_pq_scan_top, dim3(1), dim3(1024), 0, s, (uint32_t *)o.bsum, nb);
        if (nl.width == 4)
            hipLaunchKernelGGL(k_pq_scatter<uint32_t>, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid,
                               nl.nrows, (const uint32_t *)o.bsum, (const uint32_t *)o.dense, nl.dense,
                               (uint32_t *)o.data);
        else
            hipLaunchKernelGGL(k_pq_scatter<uint64_t>, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid,
                               nl.nrows, (const uint32_t *)o.bsum, (const uint64_t *)o.dense, nl.dense,
                               (uint64_t *)o.data);
    }
    HIPCHK(hipGetLastError());
}

}  // namespace
