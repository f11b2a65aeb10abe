// gbagg_parquet.hpp -- Parquet row-group decode on the GPU: the step before
// the aggregate. It delivers the fused path's per-row-group columns
// (ParquetColumnData: i32 / i64 / f64 values, dictionary codes plus the
// dictionary, a validity byte per row; include/slothdb/storage/parquet.hpp:147-207)
// as HBM-resident gbagg_column_views, from a file image in host memory.
//
// Scope: flat schemas (REQUIRED or OPTIONAL top-level columns), uncompressed
// column chunks, INT32 / INT64 / DOUBLE / BYTE_ARRAY, data pages V1 and V2,
// PLAIN and PLAIN_DICTIONARY / RLE_DICTIONARY encodings (BYTE_ARRAY only
// dictionary-encoded). The reference reader's page decoding is
// src/storage/parquet.cpp:1480-1660; the formats are the Parquet format
// specification's (Thrift compact protocol metadata, RLE / bit-packed hybrid
// levels and indices).
//
// Host: the footer and page headers, and each RLE / bit-packed hybrid
// stream's run headers (a few bytes per run; bit-packed runs are not read).
// Device: every column chunk's bytes are uploaded once, then
//   k_pq_hybrid   one thread per value: its run (binary search), then the
//                 repeated value or the bw bits at its bit offset
//   k_pq_gather   dictionary pages: dense value i = dictionary[index i];
//                 PLAIN pages were copied into the dense values as they are
//   k_pq_count / k_pq_scan_top / k_pq_scatter
//                 OPTIONAL columns: exclusive count of valid rows per row,
//                 then value[row] = dense[count] where valid
#pragma once

namespace gbagg {

struct PqRun {
    uint64_t byte_off;  // packed runs: first data byte in the chunk buffer
    uint32_t out0;      // first output position
    uint32_t count;     // values in the run
    uint32_t value;     // RLE runs: the repeated value
    uint32_t bw;        // bit-packed runs: bit width (1..32); 0: an RLE run
};

// Decodes one hybrid stream set: out[i] for i < n (runs sorted by out0; each
// page may have its own bit width). T = uint8_t for levels, uint32_t for
// dictionary indices.
template <class T>
__global__ __launch_bounds__(256) void k_pq_hybrid(const uint8_t *__restrict__ buf, const PqRun *__restrict__ runs,
                                                   uint32_t nruns, T *__restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = nruns - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (runs[mid].out0 <= i) lo = mid;
        else hi = mid - 1;
    }
    const PqRun r = runs[lo];
    uint32_t v = r.value;
    if (r.bw) {
        const uint32_t bw = r.bw;
        // values are packed least-significant bit first; the buffer carries
        // 16 bytes of padding past its end, so the 8-byte window stays inside
        const uint64_t bit = (uint64_t)(i - r.out0) * bw;
        const uint8_t *p = buf + r.byte_off + (bit >> 3);
        uint64_t w = 0;
#pragma unroll
        for (int b = 0; b < 8; b++) w |= (uint64_t)p[b] << (8 * b);
        const uint64_t mask = bw == 32 ? 0xFFFFFFFFull : ((1ull << bw) - 1);
        v = (uint32_t)((w >> (bit & 7)) & mask);
    }
    out[i] = (T)v;
}

// Dense values of dictionary pages: dense[i] = dict[idx[i]] for i in the
// segments' ranges; an index outside the dictionary sets *bad.
struct PqSeg {
    uint32_t dense0;    // first dense value
    uint32_t count;
    uint64_t dict_off;  // dictionary values in the chunk buffer
    uint32_t dict_n;
    uint32_t pad;
};

template <class V>
__global__ __launch_bounds__(256) void k_pq_gather(const uint8_t *__restrict__ buf, const PqSeg *__restrict__ segs,
                                                   uint32_t nseg, const uint32_t *__restrict__ idx, V *__restrict__ dense,
                                                   uint32_t *bad) {
    const PqSeg s = segs[blockIdx.y];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < s.count; j += gridDim.x * blockDim.x) {
        const uint32_t k = idx[s.dense0 + j];
        V v = 0;
        if (k < s.dict_n) {
            __builtin_memcpy(&v, buf + s.dict_off + (uint64_t)k * sizeof(V), sizeof(V));
        } else {
            __hip_atomic_fetch_or(bad, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        dense[s.dense0 + j] = v;
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
            if (cnt) runs.push_back(bw ? PqRun{(uint64_t)(r.p - chunk0), out0 + done, cnt, 0, bw}
                                       : PqRun{0, out0 + done, cnt, 0, 0});
            if (count_ones)
                for (uint32_t k = 0; k < cnt; k++) ones += (r.p[k >> 3] >> (k & 7)) & 1;
            r.p += bytes;
            done += cnt;
        } else {  // RLE: (h >> 1) repeats of one value
            const uint64_t rep = h >> 1;
            uint32_t v = 0;
            for (uint32_t b = 0; b < vbytes; b++) v |= (uint32_t)r.byte() << (8 * b);
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(rep, n - done);
            if (cnt) runs.push_back(PqRun{0, out0 + done, cnt, v, 0});
            if (count_ones && v) ones += cnt;
            done += cnt;
            if (rep == 0) fail(GBAGG_INVALID, "parquet: empty RLE run");
        }
    }
    return ones;
}

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
        DevBuf data, dense, valid;
        std::vector<uint32_t> dict_off;
        std::vector<char> dict_bytes;
    };
    std::vector<std::unique_ptr<Out>> outs;
    // scratch reused by every decode
    DevBuf chunk, idx, bsum, runs, segs, bad;
    ~gbagg_parquet() {
        if (stream) (void)hipStreamDestroy(stream);
    }
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

// Decodes column c of row group g into o; fills the view.
void parquet_decode_column(gbagg_parquet *pq, uint32_t g, uint32_t c, gbagg_parquet::Out &o, gbagg_column_view &view) {
    gbagg_parquet &w = *pq;  // scratch buffers
    const PqRowGroup &rg = pq->rgs[g];
    const PqColumnMeta &m = rg.cols[c];
    const PqSchemaCol &sc = pq->cols[c];
    if (sc.repetition == 2) fail(GBAGG_INVALID, "parquet: repeated columns are not supported");
    if (m.codec != 0) fail(GBAGG_INVALID, "parquet: compressed column chunks are not decoded on the GPU (codec " +
                                              std::to_string(m.codec) + ")");
    const gbagg_type type = parquet_type(m.type);
    const uint32_t width = type == GBAGG_TYPE_INTEGER || type == GBAGG_TYPE_VARCHAR ? 4 : 8;
    const bool optional = sc.repetition == 1;
    const int64_t start = m.dict_off > 0 ? std::min(m.dict_off, m.data_off) : m.data_off;
    if (start < 4 || m.total_size < 0 || (uint64_t)(start + m.total_size) > pq->size - 8)
        fail(GBAGG_INVALID, "parquet: column chunk outside the file");
    if (rg.num_rows > 0x7FFFFFFF) fail(GBAGG_INVALID, "parquet: row group above 2^31 rows");
    const uint32_t nrows = (uint32_t)rg.num_rows;
    const uint8_t *c0 = pq->file + start, *cend = c0 + m.total_size;
    hipStream_t s = pq->stream;
    // the chunk's bytes, once, with padding for the 8-byte bit-window reads
    w.chunk.ensure((size_t)m.total_size + 16);
    HIPCHK(hipMemsetAsync(w.chunk.as<uint8_t>() + m.total_size, 0, 16, s));
    if (m.total_size) HIPCHK(hipMemcpyAsync(w.chunk.ptr, c0, (size_t)m.total_size, hipMemcpyHostToDevice, s));
    const uint8_t *dbuf = w.chunk.as<uint8_t>();

    std::vector<PqRun> def_runs, idx_runs;
    std::vector<PqSeg> segs;
    struct PlainCopy { uint64_t src; uint32_t dense0, count; };
    std::vector<PlainCopy> plains;
    bool have_dict = false;
    uint64_t dict_off = 0;
    uint32_t dict_n = 0;
    uint32_t rows = 0, dense = 0;
    o.dict_off.assign(1, 0);
    o.dict_bytes.clear();
    const uint8_t *p = c0;
    while (p < cend && rows < nrows) {
        TReader hr{p, cend};
        const PqPage pg = read_page_header(hr);
        const uint8_t *body = hr.p;
        if (pg.comp_size < 0 || pg.comp_size > cend - body) fail(GBAGG_INVALID, "parquet: page outside its chunk");
        const uint8_t *bend = body + pg.comp_size;
        p = bend;
        if (pg.type == 2) {  // dictionary page: PLAIN values
            if (pg.encoding != 0 && pg.encoding != 2) fail(GBAGG_INVALID, "parquet: dictionary page not PLAIN");
            have_dict = true;
            dict_n = (uint32_t)pg.num_values;
            dict_off = (uint64_t)(body - c0);
            if (type == GBAGG_TYPE_VARCHAR) {
                const uint8_t *q = body;
                for (uint32_t k = 0; k < dict_n; k++) {
                    if (bend - q < 4) fail(GBAGG_INVALID, "parquet: truncated dictionary");
                    uint32_t len;
                    std::memcpy(&len, q, 4);
                    q += 4;
                    if (len > (uint64_t)(bend - q)) fail(GBAGG_INVALID, "parquet: truncated dictionary string");
                    o.dict_bytes.insert(o.dict_bytes.end(), q, q + len);
                    o.dict_off.push_back((uint32_t)o.dict_bytes.size());
                    q += len;
                }
            } else if ((uint64_t)dict_n * width > (uint64_t)pg.comp_size) {
                fail(GBAGG_INVALID, "parquet: truncated dictionary");
            }
            continue;
        }
        if (pg.type != 0 && pg.type != 3) continue;  // index pages
        if (pg.type == 3 && pg.v2_compressed && m.codec != 0) fail(GBAGG_INVALID, "parquet: compressed V2 page");
        const uint32_t n = (uint32_t)pg.num_values;
        if (n > nrows - rows) fail(GBAGG_INVALID, "parquet: more values than the row group's rows");
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
                if (pg.v2_def_len > bend - body) fail(GBAGG_INVALID, "parquet: truncated levels");
                lend = body + pg.v2_def_len;
            }
            nvalid = (uint32_t)hybrid_runs(lv, lend, c0, 1, n, rows, def_runs, true);
            vals = lend;
        } else if (pg.type == 3) {
            vals = body + pg.v2_def_len + pg.v2_rep_len;
        }
        if (pg.encoding == 0) {  // PLAIN
            if (type == GBAGG_TYPE_VARCHAR) fail(GBAGG_INVALID, "parquet: BYTE_ARRAY pages must be dictionary-encoded");
            if ((uint64_t)nvalid * width > (uint64_t)(bend - vals)) fail(GBAGG_INVALID, "parquet: truncated values");
            plains.push_back(PlainCopy{(uint64_t)(vals - c0), dense, nvalid});
        } else if (pg.encoding == 2 || pg.encoding == 8) {  // PLAIN_DICTIONARY / RLE_DICTIONARY
            if (!have_dict) fail(GBAGG_INVALID, "parquet: dictionary-encoded page before the dictionary");
            if (bend - vals < 1) fail(GBAGG_INVALID, "parquet: truncated indices");
            const uint32_t bw = vals[0];
            if (bw > 32) fail(GBAGG_INVALID, "parquet: index bit width above 32");
            if (nvalid) {
                if (bw == 0) idx_runs.push_back(PqRun{0, dense, nvalid, 0, 0});
                else hybrid_runs(vals + 1, bend, c0, bw, nvalid, dense, idx_runs, false);
            }
            if (type != GBAGG_TYPE_VARCHAR && nvalid) segs.push_back(PqSeg{dense, nvalid, dict_off, dict_n, 0});
        } else {
            fail(GBAGG_INVALID, "parquet: value encoding not supported (PLAIN, PLAIN_DICTIONARY, RLE_DICTIONARY)");
        }
        rows += n;
        dense += nvalid;
    }
    if (rows != nrows) fail(GBAGG_INVALID, "parquet: column chunk holds fewer values than the row group");

    // dense values (non-NULL rows, in order), then the rows
    const size_t wbytes = (size_t)std::max<uint32_t>(dense, 1) * width;
    o.dense.ensure(wbytes);
    for (const auto &pc : plains)
        if (pc.count)
            HIPCHK(hipMemcpyAsync(o.dense.as<uint8_t>() + (size_t)pc.dense0 * width, dbuf + pc.src,
                                  (size_t)pc.count * width, hipMemcpyDeviceToDevice, s));
    auto launch_hybrid = [&](const std::vector<PqRun> &runs, void *out, uint32_t n, bool bytes) {
        if (runs.empty() || !n) return;
        w.runs.ensure(runs.size() * sizeof(PqRun));
        HIPCHK(hipMemcpyAsync(w.runs.ptr, runs.data(), runs.size() * sizeof(PqRun), hipMemcpyHostToDevice, s));
        const uint32_t grid = (n + 255) / 256;
        if (bytes)
            hipLaunchKernelGGL(k_pq_hybrid<uint8_t>, dim3(grid), dim3(256), 0, s, dbuf, w.runs.as<PqRun>(),
                               (uint32_t)runs.size(), (uint8_t *)out, n);
        else
            hipLaunchKernelGGL(k_pq_hybrid<uint32_t>, dim3(grid), dim3(256), 0, s, dbuf, w.runs.as<PqRun>(),
                               (uint32_t)runs.size(), (uint32_t *)out, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));  // the run table buffer is reused
    };
    // dictionary indices: straight into the dense array for BYTE_ARRAY codes
    if (!idx_runs.empty()) {
        uint32_t *idx_out;
        if (type == GBAGG_TYPE_VARCHAR) {
            idx_out = o.dense.as<uint32_t>();
        } else {
            w.idx.ensure((size_t)std::max<uint32_t>(dense, 1) * 4);
            idx_out = w.idx.as<uint32_t>();
        }
        launch_hybrid(idx_runs, idx_out, dense, false);
        if (type != GBAGG_TYPE_VARCHAR && !segs.empty()) {
            w.segs.ensure(segs.size() * sizeof(PqSeg));
            HIPCHK(hipMemcpyAsync(w.segs.ptr, segs.data(), segs.size() * sizeof(PqSeg), hipMemcpyHostToDevice, s));
            const dim3 grid(64, (uint32_t)segs.size());
            if (width == 4)
                hipLaunchKernelGGL(k_pq_gather<uint32_t>, grid, dim3(256), 0, s, dbuf, w.segs.as<PqSeg>(),
                                   (uint32_t)segs.size(), (const uint32_t *)idx_out, o.dense.as<uint32_t>(),
                                   w.bad.as<uint32_t>());
            else
                hipLaunchKernelGGL(k_pq_gather<uint64_t>, grid, dim3(256), 0, s, dbuf, w.segs.as<PqSeg>(),
                                   (uint32_t)segs.size(), (const uint32_t *)idx_out, o.dense.as<uint64_t>(),
                                   w.bad.as<uint32_t>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(s));  // the segment table buffer is reused
        }
    }
    view = gbagg_column_view{};
    view.type = type;
    if (optional) {
        o.valid.ensure(std::max<uint32_t>(nrows, 1));
        launch_hybrid(def_runs, o.valid.ptr, nrows, true);
        const uint32_t nb = (nrows + PQ_SCAN_BLOCK - 1) / PQ_SCAN_BLOCK;
        if (nb > 1024u * 64) fail(GBAGG_INVALID, "parquet: row group too large");
        w.bsum.ensure((size_t)std::max<uint32_t>(nb, 1) * 4);
        o.data.ensure((size_t)std::max<uint32_t>(nrows, 1) * width);
        if (nrows) {
            hipLaunchKernelGGL(k_pq_count, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s, (const uint8_t *)o.valid.ptr, nrows,
                               w.bsum.as<uint32_t>());
            hipLaunchKernelGGL(k_pq_scan_top, dim3(1), dim3(1024), 0, s, w.bsum.as<uint32_t>(), nb);
            if (width == 4)
                hipLaunchKernelGGL(k_pq_scatter<uint32_t>, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s,
                                   (const uint8_t *)o.valid.ptr, nrows, (const uint32_t *)w.bsum.ptr,
                                   (const uint32_t *)o.dense.ptr, dense, o.data.as<uint32_t>());
            else
                hipLaunchKernelGGL(k_pq_scatter<uint64_t>, dim3(nb), dim3(PQ_SCAN_BLOCK), 0, s,
                                   (const uint8_t *)o.valid.ptr, nrows, (const uint32_t *)w.bsum.ptr,
                                   (const uint64_t *)o.dense.ptr, dense, o.data.as<uint64_t>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(s));  // the block sums buffer is reused
        }
        view.data = o.data.ptr;
        view.validity = o.valid.as<uint8_t>();
    } else {
        view.data = o.dense.ptr;
        view.validity = nullptr;
    }
    HIPCHK(hipStreamSynchronize(s));  // the chunk buffer is reused by the next column
    if (type == GBAGG_TYPE_VARCHAR) {
        if (!have_dict) fail(GBAGG_INVALID, "parquet: BYTE_ARRAY column chunk without a dictionary");
        if (o.dict_bytes.empty()) o.dict_bytes.push_back(0);
        view.dict_offsets = o.dict_off.data();
        view.dict_bytes = o.dict_bytes.data();
        view.dict_size = (uint32_t)(o.dict_off.size() - 1);
        view.flags = 0;  // Parquet does not promise distinct dictionary entries
    }
}

}  // namespace
