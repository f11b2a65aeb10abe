// gbagg_strings.hpp -- VARCHAR columns that arrive as strings (offsets +
// bytes, one string per row) rather than as dictionary codes: the reference's
// string_t vectors (include/slothdb/common/types/string_type.hpp:19-96) and
// its non-dictionary VARCHAR group keys (RadixHashCountStr,
// src/execution/radix_count_agg.cpp:390-761, which groups on a 64-bit string
// hash). Here every distinct string of a column gets one dense code on the
// GPU, and the codes feed the operator's dictionary-code paths unchanged.
// Groups are exact: rows whose strings hash alike but differ in their bytes
// get different codes.
//
// A column's batches are interned together, as one concatenation of rows:
//   k_str_hash    per row: 64-bit hash of the bytes, offsets checked,
//                 HyperLogLog sketch per block (sizes the table)
//   then rounds of
//   k_str_probe   per unresolved row: linear probing from its hash's home
//                 slot; a slot with the row's hash and a code is compared
//                 byte for byte (equal: the row has its code; different:
//                 probe on); an empty slot is claimed for the hash; a slot
//                 with the hash and no code yet records the smallest row
//                 that reached it (atomicMin)
//   k_str_reps / k_str_scan / k_str_assign
//                 that row takes the next code (in row order), copies its
//                 string into the arena and publishes the code, with the
//                 string's first arena word, on the slot (a probe then loads
//                 the arena words and the code's length at the same time)
//   k_str_check   the other rows that stopped there compare their bytes with
//                 the code's string; rows with a different string probe on
//                 from the next slot in the next round
// until no row is unresolved (one round unless two distinct strings share a
// 64-bit hash). Codes are dense from 0, in order of the strings' first rows
// within each round.
//   k_str_match   VARCHAR predicates (=, <>, LIKE '%x%', NOT LIKE) once per
//                 distinct string, as the per-code match table the row
//                 kernels read
#pragma once

namespace gbagg {

constexpr uint32_t STR_NONE = 0xFFFFFFFFu;  // slot: no code yet; row: resolved or NULL
constexpr uint32_t STR_HOME = 0xFFFFFFFEu;  // row: not probed yet, start at the home slot
constexpr uint32_t STR_CAND = 0x80000000u;  // row's slot, flagged: the row lowered the slot's rep
constexpr uint32_t STR_REP = 0x40000000u;   // row's slot, flagged: the row is the slot's representative
constexpr uint64_t STR_FULL_TABLE = 1, STR_FULL_DICT = 2, STR_BAD_OFFSETS = 4;
// A row's key: 48 bits of its string's hash, and its length in the top 16
// bits (STR_LEN_LONG for 65535 bytes and more). Equal keys then mean equal
// lengths below STR_LEN_LONG, and a comparison needs the code's own length
// only for long strings.
constexpr uint64_t STR_KEY_HASH = (1ull << 48) - 1;
constexpr uint32_t STR_LEN_LONG = 0xFFFFu;
enum { SC_CODES = 0, SC_WORDS, SC_UNRESOLVED, SC_FLAGS, SC_ROWS, SC_BYTES, SC_N };

// Rows [row0, row0 + n) of the column's concatenation: one batch, given as
// strings (off: n + 1 ascending offsets into bytes), as dictionary codes
// (dcodes) whose strings are dictionary entries off/bytes (through remap), or
// as fixed-width rows of `fixed` bytes each (packed multi-column keys).
struct StrSeg {
    const uint32_t *off;
    const uint8_t *bytes;
    const uint8_t *valid;    // one byte per row, 0 = NULL; null: all valid
    const uint32_t *dcodes;  // null: a batch of strings
    const uint32_t *remap;   // batch code -> dictionary entry; null: identity
    uint64_t row0;
    uint64_t n;
    uint32_t dict_n;         // batch codes >= dict_n drop their row
    uint32_t fixed;          // bytes per row of fixed-width rows (0: strings or codes)
};

constexpr uint32_t STR_DROP = 0xFFFFFFFFu;  // output code of a dropped row (>= any dictionary size)

struct StrRow {
    const uint8_t *p;
    uint32_t len;
    bool valid;
};

// The string of row r of a segment (a valid row with an in-range code).
__device__ __forceinline__ StrRow str_at(const StrSeg &s, uint64_t r) {
    StrRow out;
    out.valid = true;
    if (s.fixed) {
        out.p = s.bytes + r * s.fixed;
        out.len = s.fixed;
        return out;
    }
    uint32_t o0, o1;
    if (s.dcodes) {
        const uint32_t c = gld(&s.dcodes[r]);
        const uint32_t e = s.remap ? gld(&s.remap[c]) : c;
        o0 = gld(&s.off[e]);
        o1 = gld(&s.off[e + 1]);
    } else {
        o0 = gld(&s.off[r]);
        o1 = gld(&s.off[r + 1]);
    }
    out.p = s.bytes + o0;
    out.len = o1 >= o0 ? o1 - o0 : 0;
    return out;
}

// One slot of the table: a probe reads one 16-byte record.
struct StrSlot {
    uint64_t key;  // stored hash, 0 = empty
    uint32_t val;  // code, STR_NONE = none yet
    uint32_t rep;  // no code yet: smallest row that stopped here; with a code: its first arena word
};

// One code: its string's first arena word and its length.
struct StrCode {
    uint32_t word;
    uint32_t len;
};

struct StrTab {
    StrSlot *slot;
    StrCode *code;
    uint32_t *arena;     // distinct strings, each from a word boundary, zero padded
    uint64_t *ctr;       // SC_* counters
    uint64_t mask;       // slots - 1
    uint64_t arena_words;
    uint64_t code_cap;
    uint64_t key_mask;   // hash bits kept in a row's key (all 48, unless a test narrows them)
    uint64_t seed;
};

__device__ __forceinline__ uint32_t str_seg(const StrSeg *segs, uint32_t nseg, uint64_t g) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (segs[mid].row0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// r (1..4) bytes from p, little endian, zero above them. Only the aligned
// words holding those bytes are read, so nothing outside the words of the
// string's own bytes is touched.
__device__ __forceinline__ uint32_t str_word(const uint8_t *p, uint32_t r) {
    const uintptr_t a = (uintptr_t)p & ~(uintptr_t)3;
    const uint32_t sh = (uint32_t)((uintptr_t)p & 3);
    const uint64_t lo = gld((const uint32_t *)a);
    const uint64_t hi = sh + r > 4 ? gld((const uint32_t *)(a + 4)) : 0;
    const uint32_t v = (uint32_t)(((hi << 32) | lo) >> (8 * sh));
    return r >= 4 ? v : (v & ((1u << (8 * r)) - 1));
}

// Strings of at most STR_SHORT bytes are read in one go, as 4 little-endian
// 64-bit words, zero above len (len >= 1). The aligned 8-byte words holding
// the bytes load together: indices past the last of them are clamped to it,
// so every load is unconditional and none waits on another. An aligned
// 8-byte word that holds a byte of the string lies in that byte's memory
// page; bytes outside the string are masked off.
constexpr uint32_t STR_SHORT = 32;

__device__ __forceinline__ void str_load5(const uint8_t *p, uint32_t len, uint64_t (&d)[5]) {
    const uint64_t *a = (const uint64_t *)((uintptr_t)p & ~(uintptr_t)7);
    const uint32_t last = ((uint32_t)((uintptr_t)p & 7) + len - 1) / 8;  // 0..4
#pragma unroll
    for (uint32_t i = 0; i < 5; i++) d[i] = gld(a + (i < last ? i : last));
}

__device__ __forceinline__ void str_extract4(const uint64_t (&d)[5], const uint8_t *p, uint32_t len,
                                             uint64_t (&w)[4]) {
    const uint32_t sh = 8 * (uint32_t)((uintptr_t)p & 7);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint64_t v = sh ? (d[j] >> sh) | (d[j + 1] << (64 - sh)) : d[j];
        const int rem = (int)len - 8 * j;
        w[j] = rem >= 8 ? v : rem <= 0 ? 0ull : v & ((1ull << (8 * rem)) - 1ull);
    }
}

__device__ __forceinline__ void str_words4(const uint8_t *p, uint32_t len, uint64_t (&w)[4]) {
    uint64_t d[5];
    str_load5(p, len, d);
    str_extract4(d, p, len, w);
}

__device__ __forceinline__ uint64_t str_fold(uint64_t a, uint64_t b) {
    return a * b ^ __umul64hi(a, b);
}

// 64-bit string hash: the length seeds the state, each 8-byte word (the last
// one zero padded) is folded in with a 64x64->128-bit multiply (low ^ high
// half), and SplitMix64's finalizer mixes the result.
__device__ __forceinline__ uint64_t str_hash(const uint8_t *p, uint32_t len, uint64_t seed) {
    uint64_t h = seed ^ ((uint64_t)len * 0x9E3779B97F4A7C15ull);
    if (len && len <= STR_SHORT) {
        uint64_t w[4];
        str_words4(p, len, w);
#pragma unroll
        for (uint32_t i = 0; i < 4; i++)
            if (8 * i < len) h = str_fold(h ^ w[i], 0xBF58476D1CE4E5B9ull) + 0x94D049BB133111EBull;
    } else {
        for (uint32_t i = 0; i < len; i += 8) {
            const uint32_t left = len - i;
            uint64_t w = str_word(p + i, left < 4 ? left : 4);
            if (left > 4) w |= (uint64_t)str_word(p + i + 4, left - 4 < 4 ? left - 4 : 4) << 32;
            h = str_fold(h ^ w, 0xBF58476D1CE4E5B9ull) + 0x94D049BB133111EBull;
        }
    }
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBull;
    return h ^ (h >> 31);
}

__device__ __forceinline__ uint64_t str_home(uint64_t key, uint64_t mask) {
    return ((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

// A row's string, as str_equal compares it: a short string's words are
// loaded once, before its probe sequence.
struct StrKey {
    const uint8_t *p;
    uint32_t len;
    uint64_t w[4];  // len 1..STR_SHORT: the string's words
};

__device__ __forceinline__ StrKey str_key(const uint8_t *p, uint32_t len) {
    StrKey k;
    k.p = p;
    k.len = len;
    if (len && len <= STR_SHORT) str_words4(p, len, k.w);
    return k;
}

// The row's string equals code c's string, whose arena copy starts at word w0
// (the slot holds it next to the code). The slot's key equals the row's, so
// the lengths are equal unless the row's is STR_LEN_LONG or more.
__device__ __forceinline__ bool str_equal(const StrTab &t, uint32_t c, uint32_t w0, const StrKey &k) {
    const uint32_t len = k.len;
    const uint32_t *a = t.arena + w0;
    if (len && len <= STR_SHORT) {
        uint64_t x[4];
        str_words4((const uint8_t *)a, len, x);
        return ((k.w[0] ^ x[0]) | (k.w[1] ^ x[1]) | (k.w[2] ^ x[2]) | (k.w[3] ^ x[3])) == 0;
    }
    if (len >= STR_LEN_LONG && gld(&t.code[c].len) != len) return false;
    for (uint32_t i = 0; i < len; i += 4) {
        const uint32_t left = len - i;
        if (str_word(k.p + i, left < 4 ? left : 4) != gld(&a[i / 4])) return false;
    }
    return true;
}

// Calls f(segment, row within it) for row g. A wave's rows are contiguous, so
// they are nearly always in one segment: then the segment is found and read
// once for the wave, with scalar loads, and its fields are uniform; otherwise
// each lane reads its own.
template <class F>
__device__ __forceinline__ void with_str_seg(const StrSeg *__restrict__ segs, uint32_t nseg, uint64_t g, F &&f) {
    const uint64_t g0 = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(g >> 32)) << 32 |
                        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)g);
    const uint32_t si = (uint32_t)__builtin_amdgcn_readfirstlane(str_seg(segs, nseg, g0));
    const StrSeg sw = segs[si];
    if (__all(g >= sw.row0 && g - sw.row0 < sw.n)) {
        f(sw, g - sw.row0);
    } else {
        const StrSeg sl = segs[str_seg(segs, nseg, g)];
        f(sl, g - sl.row0);
    }
}

__device__ __forceinline__ StrRow str_row(const StrSeg *segs, uint32_t nseg, uint64_t g) {
    StrRow out;
    with_str_seg(segs, nseg, g, [&](const StrSeg &s, uint64_t r) {
        out = str_at(s, r);
        out.valid = !s.valid || gld(&s.valid[r]) != 0;
    });
    return out;
}

__global__ __launch_bounds__(256) void k_str_init(StrTab t) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= t.mask;
         i += (uint64_t)gridDim.x * blockDim.x) {
        t.slot[i] = StrSlot{0, STR_NONE, STR_NONE};
    }
    if (blockIdx.x == 0 && threadIdx.x < SC_N) t.ctr[threadIdx.x] = 0;
}

// Every valid row back to its home slot (a restart with larger tables).
__global__ __launch_bounds__(256) void k_str_rehome(uint64_t total, const uint64_t *hash, uint32_t *pos) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x)
        pos[g] = hash[g] ? STR_HOME : STR_NONE;
}

// Hashes, offsets check, valid row and byte counts, HyperLogLog per block.
__global__ __launch_bounds__(256) void k_str_hash(const StrSeg *segs, uint32_t nseg, uint64_t total, StrTab t,
                                                  uint64_t *hash, uint32_t *pos, uint32_t *codes, uint8_t *hllp) {
    __shared__ uint32_t reg[HLL_M];
    __shared__ uint64_t bsum[2];
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) reg[i] = 0;
    if (threadIdx.x < 2) bsum[threadIdx.x] = 0;
    __syncthreads();
    uint64_t rows = 0, bytes = 0;
    bool bad = false;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) with_str_seg(segs, nseg, g, [&](const StrSeg &s, uint64_t r) {
        // the row's validity byte, code and offsets load before any of them
        // is used (a row without a validity array reads a stand-in byte)
        const uint8_t vb = gld(s.valid ? s.valid + r : (const uint8_t *)segs);
        const uint8_t *p;
        uint32_t len;
        bool drop = false, ok = true;
        if (s.fixed) {
            p = s.bytes + r * s.fixed;
            len = s.fixed;
        } else if (s.dcodes) {
            // a dictionary code outside its dictionary drops the row, as it
            // does for dictionary-coded keys
            const uint32_t c = gld(&s.dcodes[r]);
            drop = c >= s.dict_n;
            p = s.bytes;
            len = 0;
            if (!drop) {
                const uint32_t e = s.remap ? gld(&s.remap[c]) : c;
                const uint32_t o0 = gld(&s.off[e]), o1 = gld(&s.off[e + 1]);
                p = s.bytes + o0;
                len = o1 >= o0 ? o1 - o0 : 0;
            }
        } else {
            // every row's offsets ascend and stay within the batch's bytes
            // (offsets[n]); a row outside them is never read
            const uint32_t o0 = gld(&s.off[r]), o1 = gld(&s.off[r + 1]), end = gld(&s.off[s.n]);
            ok = o0 <= o1 && o1 <= end;
            if (!ok) bad = true;
            p = s.bytes + o0;
            len = o1 - o0;
        }
        const bool valid = !s.valid || vb != 0;
        drop = drop && valid;
        if (!valid || drop || !ok) {
            hash[g] = 0;
            pos[g] = STR_NONE;
            codes[g] = drop ? STR_DROP : 0;
            return;
        }
        const uint64_t full = str_hash(p, len, t.seed);
        const uint64_t k = (full & t.key_mask & STR_KEY_HASH) | (uint64_t)min(len, STR_LEN_LONG) << 48;
        hash[g] = k ? k : 1;
        pos[g] = STR_HOME;
        hll_add(reg, full);
        rows++;
        bytes += len;
    });
    if (bad) __hip_atomic_fetch_or(&t.ctr[SC_FLAGS], STR_BAD_OFFSETS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicAdd((unsigned long long *)&bsum[0], (unsigned long long)rows);
    atomicAdd((unsigned long long *)&bsum[1], (unsigned long long)bytes);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) hllp[(uint64_t)blockIdx.x * HLL_M + i] = (uint8_t)reg[i];
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&t.ctr[SC_ROWS], (unsigned long long)bsum[0]);
        atomicAdd((unsigned long long *)&t.ctr[SC_BYTES], (unsigned long long)bsum[1]);
    }
}

// Rows that stop at a slot without a code are listed for k_str_check, which
// then reads only them: workgroup b lists its rows in list[b * per_block ...)
// (per_block: at least the rows it visits) and their number in lcnt[b].
__global__ __launch_bounds__(256) void k_str_probe(const StrSeg *segs, uint32_t nseg, uint64_t total, StrTab t,
                                                   const uint64_t *hash, uint32_t *pos, uint32_t *codes,
                                                   uint32_t *list, uint32_t *lcnt, uint64_t per_block) {
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    uint32_t *mine = list + (uint64_t)blockIdx.x * per_block;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p0 = pos[g];
        if (p0 == STR_NONE) continue;
        const uint64_t h = hash[g];
        const StrRow row = str_row(segs, nseg, g);
        const StrKey key = str_key(row.p, row.len);
        uint64_t s = p0 == STR_HOME ? str_home(h, t.mask) : p0;
        uint32_t out = STR_NONE;
        for (uint64_t probes = 0;; probes++) {
            if (probes > t.mask) {
                __hip_atomic_fetch_or(&t.ctr[SC_FLAGS], STR_FULL_TABLE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            StrSlot *sl = &t.slot[s];
            // a key only ever changes from 0 to a row's key, so a cached read
            // is safe: a 0 is read again from the L2 before the
            // compare-and-swap (which returns the key in place if another row
            // was first). The hot strings' slots then stay in the L1.
            uint64_t k = gld(&sl->key);
            // code and arena word: set by an earlier kernel, or no code yet
            const uint64_t vr = gld((const uint64_t *)&sl->val);
            if (k == 0) k = __hip_atomic_load(&sl->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (k == 0) {
                uint64_t expect = 0;
                __hip_atomic_compare_exchange_strong(&sl->key, &expect, h, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                k = expect == 0 ? h : expect;
            }
            if (k == h) {
                const uint32_t c = (uint32_t)vr;
                if (c == STR_NONE) {
                    // no code yet: the smallest row here becomes the representative;
                    // only rows below the current one try (a hot string's rows
                    // would otherwise all hit one word with atomics)
                    const uint32_t cur = __hip_atomic_load(&sl->rep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bool cand = false;
                    if ((uint32_t)g < cur)
                        cand = (uint32_t)g < __hip_atomic_fetch_min(&sl->rep, (uint32_t)g, __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT);
                    out = (uint32_t)s | (cand ? STR_CAND : 0u);
                    break;
                }
                if (str_equal(t, c, (uint32_t)(vr >> 32), key)) {
                    codes[g] = c;
                    break;
                }
            }
            s = (s + 1) & t.mask;
        }
        pos[g] = out;
        // one reservation per wave in the workgroup's list
        const uint64_t m = __ballot(out != STR_NONE);
        if (m) {
            const uint32_t lane = __lane_id(), leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader)
                base = __hip_atomic_fetch_add(&s_n, (uint32_t)__popcll(m), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
            base = __shfl(base, (int)leader);
            if (out != STR_NONE) mine[base + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)g;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) lcnt[blockIdx.x] = s_n;
}

// k_str_reps flagged the row as its slot's representative.
__device__ __forceinline__ bool str_flagged_rep(uint32_t p, uint32_t &s) {
    s = p & ~STR_REP;
    return (p & (STR_CAND | STR_REP)) == STR_REP;
}

// Exclusive block-wide prefix sums of two values per thread (256 threads:
// four waves), and their block totals: shuffle scans within each wave, then
// the waves' totals through LDS.
__device__ __forceinline__ void str_block_scan(uint64_t a, uint64_t b, uint64_t &ea, uint64_t &eb, uint64_t &ta,
                                               uint64_t &tb) {
    __shared__ uint64_t sa[4], sb[4];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t xa = a, xb = b;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t ya = __shfl_up(xa, d), yb = __shfl_up(xb, d);
        if (lane >= d) {
            xa += ya;
            xb += yb;
        }
    }
    if (lane == 63) {
        sa[w] = xa;
        sb[w] = xb;
    }
    __syncthreads();
    uint64_t pa = 0, pb = 0;
    ta = 0;
    tb = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        if (i < w) {
            pa += sa[i];
            pb += sb[i];
        }
        ta += sa[i];
        tb += sb[i];
    }
    ea = pa + xa - a;
    eb = pb + xb - b;
    __syncthreads();  // sa, sb are written again by the next call
}

// Codes are numbered in row order of their representatives (the order of
// first appearance), by a scan instead of atomics: block b of the three
// kernels owns rows [b * chunk, (b + 1) * chunk).
//   k_str_reps    representatives and their arena words per block
//   k_str_scan    exclusive scan of the block counts from the codes and words
//                 already taken; the totals advance
//   k_str_assign  each representative takes its code, copies its string into
//                 the arena and publishes the code on its slot
__global__ __launch_bounds__(256) void k_str_reps(const StrSeg *segs, uint32_t nseg, uint64_t total, StrTab t,
                                                  uint32_t *pos, uint64_t chunk, uint64_t *blk) {
    const uint64_t r0 = (uint64_t)blockIdx.x * chunk, r1 = min(total, r0 + chunk);
    uint64_t cnt = 0, words = 0;
    // four rows per thread and step, their loads overlapping
    for (uint64_t g0 = r0 + threadIdx.x; g0 < r1; g0 += (uint64_t)blockDim.x * 4) {
        uint32_t p[4], rp[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t g = g0 + (uint64_t)j * blockDim.x;
            p[j] = g < r1 ? pos[g] : STR_NONE;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // candidates: rows that lowered their slot's rep word in k_str_probe
            const bool cand = p[j] != STR_NONE && p[j] != STR_HOME && (p[j] & STR_CAND);
            rp[j] = cand ? t.slot[p[j] & ~STR_CAND].rep : STR_NONE;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t g = g0 + (uint64_t)j * blockDim.x;
            if (rp[j] != (uint32_t)g) continue;
            // the row still holds its slot's rep word: the representative.
            // k_str_assign then knows it without the slots' rep words, which
            // it overwrites
            pos[g] = (p[j] & ~STR_CAND) | STR_REP;
            cnt++;
            words += (str_row(segs, nseg, g).len + 3) / 4;
        }
    }
    uint64_t ea, eb, ta, tb;
    str_block_scan(cnt, words, ea, eb, ta, tb);
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = ta;
        blk[2 * blockIdx.x + 1] = tb;
    }
}

// One block of 1024 threads, nblk <= 4096: blk pairs become their exclusive
// prefix sums plus the codes and words taken before this round.
__global__ __launch_bounds__(1024) void k_str_scan(uint64_t *blk, uint32_t nblk, uint64_t *ctr) {
    __shared__ uint64_t sa[1024], sb[1024];
    const uint32_t i = threadIdx.x;
    uint64_t a[4], b[4], ca = 0, cb = 0;
    for (int j = 0; j < 4; j++) {
        const uint32_t k = i * 4 + j;
        a[j] = k < nblk ? blk[2 * k] : 0;
        b[j] = k < nblk ? blk[2 * k + 1] : 0;
        ca += a[j];
        cb += b[j];
    }
    sa[i] = ca;
    sb[i] = cb;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint64_t xa = i >= d ? sa[i - d] : 0, xb = i >= d ? sb[i - d] : 0;
        __syncthreads();
        sa[i] += xa;
        sb[i] += xb;
        __syncthreads();
    }
    uint64_t pa = ctr[SC_CODES] + sa[i] - ca, pb = ctr[SC_WORDS] + sb[i] - cb;
    for (int j = 0; j < 4; j++) {
        const uint32_t k = i * 4 + j;
        if (k < nblk) {
            blk[2 * k] = pa;
            blk[2 * k + 1] = pb;
        }
        pa += a[j];
        pb += b[j];
    }
    __syncthreads();
    if (i == 1023) {
        ctr[SC_CODES] += sa[1023];
        ctr[SC_WORDS] += sb[1023];
    }
}

__global__ __launch_bounds__(256) void k_str_assign(const StrSeg *segs, uint32_t nseg, uint64_t total, StrTab t,
                                                    uint32_t *pos, uint32_t *codes, uint64_t chunk,
                                                    const uint64_t *blk) {
    const uint64_t r0 = (uint64_t)blockIdx.x * chunk, r1 = min(total, r0 + chunk);
    uint64_t bc = blk[2 * blockIdx.x], bw = blk[2 * blockIdx.x + 1];
    // each thread takes STR_AR consecutive rows of a tile: their loads overlap,
    // and one block scan serves STR_AR rows per thread
    constexpr int STR_AR = 4;
    for (uint64_t tile = r0; tile < r1; tile += (uint64_t)blockDim.x * STR_AR) {
        const uint64_t g0 = tile + (uint64_t)threadIdx.x * STR_AR;
        uint32_t sl[STR_AR];
        bool isrep[STR_AR];
#pragma unroll
        for (int j = 0; j < STR_AR; j++) isrep[j] = g0 + j < r1 && str_flagged_rep(pos[g0 + j], sl[j]);
        StrRow row[STR_AR];
        uint32_t wd[STR_AR];
        uint64_t cnt = 0, words = 0;
#pragma unroll
        for (int j = 0; j < STR_AR; j++) {
            wd[j] = 0;
            if (isrep[j]) {
                row[j] = str_row(segs, nseg, g0 + j);
                wd[j] = (row[j].len + 3) / 4;
                cnt++;
                words += wd[j];
            }
        }
        uint64_t ec, ew, tc, tw;
        str_block_scan(cnt, words, ec, ew, tc, tw);
        uint64_t c = bc + ec, w0 = bw + ew;
#pragma unroll
        for (int j = 0; j < STR_AR; j++) {
            if (!isrep[j]) continue;
            const uint64_t g = g0 + j;
            const StrRow &rw = row[j];
            if (c >= t.code_cap || w0 + wd[j] > t.arena_words) {
                __hip_atomic_fetch_or(&t.ctr[SC_FLAGS], STR_FULL_DICT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (rw.len && rw.len <= STR_SHORT) {
                    uint64_t w[4];
                    str_words4(rw.p, rw.len, w);
#pragma unroll
                    for (uint32_t i = 0; i < 8; i++)
                        if (4 * i < rw.len) gst(&t.arena[w0 + i], (uint32_t)(w[i / 2] >> (32 * (i & 1))));
                } else {
                    for (uint32_t i = 0; i < rw.len; i += 4) {
                        const uint32_t left = rw.len - i;
                        gst(&t.arena[w0 + i / 4], str_word(rw.p + i, left < 4 ? left : 4));
                    }
                }
                t.code[c] = StrCode{(uint32_t)w0, rw.len};
                // the code and its arena word; the representative is resolved
                gst((uint64_t *)&t.slot[sl[j]].val, c | w0 << 32);
                codes[g] = (uint32_t)c;
                pos[g] = STR_NONE;
            }
            c++;
            w0 += wd[j];
        }
        bc += tc;
        bw += tw;
    }
}

// The rows k_str_probe listed: nlist workgroup lists.
__global__ __launch_bounds__(256) void k_str_check(const StrSeg *segs, uint32_t nseg, StrTab t, uint32_t *pos,
                                                   uint32_t *codes, const uint32_t *list, const uint32_t *lcnt,
                                                   uint64_t per_block, uint32_t nlist) {
    uint64_t unresolved = 0;
    for (uint64_t i = 0, b = blockIdx.x;; i += blockDim.x) {
        while (b < nlist && i >= lcnt[b]) {
            b += gridDim.x;
            i = 0;
        }
        if (b >= nlist) break;
        if (i + threadIdx.x >= lcnt[b]) continue;
        const uint64_t g = list[b * per_block + i + threadIdx.x];
        const uint32_t p = pos[g];
        if (p == STR_NONE) continue;
        // representatives with a code were resolved by k_str_assign
        const uint32_t s = p & ~(STR_CAND | STR_REP);
        const uint64_t vr = gld((const uint64_t *)&t.slot[s].val);  // code and arena word
        const uint32_t c = (uint32_t)vr;
        if (c == STR_NONE) {  // the dictionary was full: the host restarts larger
            pos[g] = STR_NONE;
            continue;
        }
        const StrRow row = str_row(segs, nseg, g);
        if (str_equal(t, c, (uint32_t)(vr >> 32), str_key(row.p, row.len))) {
            codes[g] = c;
            pos[g] = STR_NONE;
        } else {
            pos[g] = (uint32_t)((s + 1) & t.mask);
            unresolved++;
        }
    }
    if (unresolved) atomicAdd((unsigned long long *)&t.ctr[SC_UNRESOLVED], (unsigned long long)unresolved);
}

// Result rows of a group column of interned strings: bytes of each row's
// string (0 for NULL rows), then the strings at the given offsets.
__global__ __launch_bounds__(256) void k_str_lengths(StrTab t, const uint32_t *ids, const uint8_t *valid,
                                                     uint64_t rows, uint32_t *len) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += (uint64_t)gridDim.x * blockDim.x)
        len[r] = valid[r] ? t.code[ids[r]].len : 0;
}

__global__ __launch_bounds__(256) void k_str_gather(StrTab t, const uint32_t *ids, const uint8_t *valid,
                                                    uint64_t rows, const uint64_t *off, uint8_t *out) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += (uint64_t)gridDim.x * blockDim.x) {
        if (!valid[r]) continue;
        const StrCode cd = t.code[ids[r]];
        const uint32_t len = cd.len;
        const uint32_t *a = t.arena + cd.word;
        uint8_t *o = out + off[r];
        for (uint32_t i = 0; i < len; i++) o[i] = (uint8_t)(a[i / 4] >> (8 * (i % 4)));
    }
}

// Codes relabelled in the byte order of their strings (rank[old] = new), so
// that MIN/MAX over codes is MIN/MAX over the strings; dropped rows keep
// STR_DROP.
__global__ __launch_bounds__(256) void k_str_relabel(uint32_t *codes, uint64_t total, const uint32_t *rank,
                                                     uint64_t ncodes) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = codes[g];
        if (c < ncodes) codes[g] = rank[c];
    }
}

// Byte i of code c's string.
__device__ __forceinline__ uint32_t str_byte(const uint32_t *a, uint32_t i) {
    return (a[i / 4] >> (8 * (i % 4))) & 0xFFu;
}

// Match table of one VARCHAR predicate over the column's codes: op 0 equal,
// 1 not equal, 2 contains (LIKE '%lit%'), 3 does not contain.
__global__ __launch_bounds__(256) void k_str_match(StrTab t, uint64_t ncodes, const uint8_t *lit, uint32_t litlen,
                                                   int op, uint8_t *match) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncodes;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const StrCode cd = t.code[c];
        const uint32_t len = cd.len;
        const uint32_t *a = t.arena + cd.word;
        bool r;
        if (op <= 1) {
            r = len == litlen;
            for (uint32_t i = 0; r && i < len; i++) r = str_byte(a, i) == lit[i];
            if (op == 1) r = !r;
        } else {
            r = litlen == 0;
            for (uint32_t s = 0; !r && s + litlen <= len; s++) {
                bool m = true;
                for (uint32_t i = 0; m && i < litlen; i++) m = str_byte(a, s + i) == lit[i];
                r = m;
            }
            if (op == 3) r = !r;
        }
        match[c] = r;
    }
}

// ---- Multi-column group keys -----------------------------------------
// Each row's group columns packed into `words` 32-bit words: one validity
// byte per column (0 = NULL) padded to a word, then each column's value at a
// word-aligned offset (NULL: zeros). DOUBLE values are canonical (-0.0 as
// 0.0, one NaN), so SQL-equal keys pack to equal bytes. The packed rows are
// interned like strings; their codes are the aggregate's one-column key.
enum { TK_I32 = 0, TK_I64 = 1, TK_F64 = 2, TK_CODE = 3 };

struct TupleComp {
    const void *data;      // values; TK_CODE: the interned codes of a VARCHAR column
    const uint8_t *valid;  // null: all valid
    uint32_t kind;
    uint32_t word;         // first word of the value in the packed row
};

// keep[r] = 0 drops the row: a code outside its dictionary, or a NULL key
// column when drop_null is set.
__global__ __launch_bounds__(256) void k_tuple_pack(const TupleComp *__restrict__ comps, uint32_t ncomp, uint64_t n,
                                                    uint32_t words, int drop_null, uint32_t *__restrict__ out,
                                                    uint8_t *__restrict__ keep) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t *o = out + r * words;
        uint32_t v0 = 0, v1 = 0;
        bool any_null = false, dropped = false;
        for (uint32_t j = 0; j < ncomp; j++) {
            const TupleComp c = comps[j];
            const bool v = !c.valid || c.valid[r] != 0;
            uint64_t x = 0;
            if (v) {
                if (c.kind == TK_I32) {
                    x = (uint32_t)((const int32_t *)c.data)[r];
                } else if (c.kind == TK_CODE) {
                    x = ((const uint32_t *)c.data)[r];
                    dropped |= x == STR_DROP;
                } else {
                    x = ((const uint64_t *)c.data)[r];
                    if (c.kind == TK_F64) {
                        if ((x << 1) == 0) x = 0;  // -0.0
                        if ((x & 0x7FF0000000000000ull) == 0x7FF0000000000000ull && (x & 0x000FFFFFFFFFFFFFull))
                            x = 0x7FF8000000000000ull;  // NaN
                    }
                }
            } else {
                any_null = true;
            }
            if (j < 4) v0 |= (v ? 1u : 0u) << (8 * j);
            else v1 |= (v ? 1u : 0u) << (8 * (j - 4));
            o[c.word] = (uint32_t)x;
            if (c.kind == TK_I64 || c.kind == TK_F64) o[c.word + 1] = (uint32_t)(x >> 32);
        }
        o[0] = v0;
        if (ncomp > 4) o[1] = v1;
        keep[r] = !(dropped || (drop_null && any_null));
    }
}

// The group columns of result rows, unpacked on the device from their tuple
// codes (for a second aggregation phase over this result).
struct TupleOut {
    uint32_t *data;  // 1 or 2 words per row
    uint8_t *valid;
    uint32_t word;   // first word of the value in the packed row
    uint32_t words;  // 1 or 2
};

__global__ __launch_bounds__(256) void k_tuple_unpack(const uint32_t *__restrict__ codes, uint64_t rows, StrTab t,
                                                      const TupleOut *__restrict__ outs, uint32_t ncomp) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t *tup = t.arena + t.code[codes[r]].word;
        for (uint32_t j = 0; j < ncomp; j++) {
            const TupleOut o = outs[j];
            o.valid[r] = (uint8_t)((tup[j / 4] >> (8 * (j % 4))) & 0xFF);
            o.data[r * o.words] = tup[o.word];
            if (o.words == 2) o.data[r * 2 + 1] = tup[o.word + 1];
        }
    }
}

}  // namespace gbagg
