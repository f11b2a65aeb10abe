// gbagg_noalloc.hpp -- GROUP BY int32 key with SUM and COUNT of int32 values
// (the shape of gbagg_hash_aggregate_i32) without any device allocation, for
// callers that cannot allocate: code under HIP graph capture, serving loops.
//
// The hash table lives in `scratch`, a file-scope __device__ array of fixed
// size that every GPU the library runs on holds until the process ends. Host
// code never passes the array's name to a kernel (in host code it is only a
// stand-in, not a device address): on its first call on a device the entry
// point resolves the array's address there with hipGetSymbolAddress, keeps it
// per device, and passes that pointer to the kernels.
//
// Every call on one GPU shares that GPU's array, and each call starts by
// clearing it, so calls that overlap on one GPU (two host threads, or two
// streams not ordered with each other) race and must not be made.
//
//   k_na_clear    fill kernel: empty keys, zero sums and counts, zero counter
//                 (no hipMemsetAsync, so the call can be captured in a graph)
//   k_na_agg      per-workgroup LDS cache of frequent keys, the rest into the
//                 scratch table with device-scope atomics; the cache is merged
//                 at the end
//   k_na_compact  occupied slots to the caller's output rows (one reservation
//                 per block); rows beyond `capacity` are counted, not written
//   k_na_total    the group count to the caller's device word
#pragma once

#include <mutex>

namespace gbagg {

constexpr uint32_t NA_SLOTS = 1u << 20;            // scratch table slots
constexpr uint64_t NA_MAX_ROWS = NA_SLOTS / 2;     // so the table is at most half full
constexpr uint32_t NA_LDS = 2048;                  // LDS cache slots per workgroup
constexpr uint32_t NA_BLOCK = 256;
constexpr size_t NA_WORDS = 3 * (size_t)NA_SLOTS + 32;  // keys, sums, counts, control

static __device__ __align__(256) uint64_t scratch[NA_WORDS];

struct NaTable {
    uint64_t *keys;    // [NA_SLOTS] sign-extended key, SENTINEL = empty (int32 keys never equal it)
    uint64_t *sums;    // [NA_SLOTS]
    uint64_t *counts;  // [NA_SLOTS]
    uint64_t *ctl;     // [0]: result rows reserved by k_na_compact
};

__host__ __device__ inline NaTable na_table(uint64_t *base) {
    return NaTable{base, base + NA_SLOTS, base + 2 * (size_t)NA_SLOTS, base + 3 * (size_t)NA_SLOTS};
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_clear(NaTable t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < NA_SLOTS; i += gridDim.x * blockDim.x) {
        t.keys[i] = SENTINEL;
        t.sums[i] = 0;
        t.counts[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 32) t.ctl[threadIdx.x] = 0;
}

// Slot of key k in the scratch table, claimed if absent. The table holds at
// most NA_MAX_ROWS keys in NA_SLOTS slots, so the probe always ends.
__device__ __forceinline__ uint32_t na_slot(const NaTable &t, uint64_t k) {
    uint32_t pos = (uint32_t)(((uint64_t)(uint32_t)hash64(k) * NA_SLOTS) >> 32);
    for (;;) {
        uint64_t cur = __hip_atomic_load(&t.keys[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == SENTINEL) {
            uint64_t e = SENTINEL;
            __hip_atomic_compare_exchange_strong(&t.keys[pos], &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            cur = e == SENTINEL ? k : e;
        }
        if (cur == k) return pos;
        pos = (pos + 1) & (NA_SLOTS - 1);
    }
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_agg(const int32_t *__restrict__ keys,
                                                     const int32_t *__restrict__ vals, uint64_t n, NaTable t) {
    __shared__ uint64_t lk[NA_LDS], ls[NA_LDS], lc[NA_LDS];
    __shared__ uint32_t fill;
    for (uint32_t i = threadIdx.x; i < NA_LDS; i += blockDim.x) {
        lk[i] = SENTINEL;
        ls[i] = 0;
        lc[i] = 0;
    }
    if (threadIdx.x == 0) fill = 0;
    __syncthreads();
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = (uint64_t)(int64_t)gld(keys + r);
        const uint64_t v = (uint64_t)(int64_t)gld(vals + r);
        // the LDS cache: find or claim within 16 probes while it is under 3/4 full
        uint32_t pos = (uint32_t)(((uint64_t)(uint32_t)hash64(k) * NA_LDS) >> 32);
        int slot = -1;
        for (int p = 0; p < 16; p++) {
            uint64_t cur = __hip_atomic_load(&lk[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == SENTINEL) {
                if (__hip_atomic_load(&fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= NA_LDS * 3 / 4) break;
                uint64_t e = SENTINEL;
                __hip_atomic_compare_exchange_strong(&lk[pos], &e, k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                if (e == SENTINEL)
                    __hip_atomic_fetch_add(&fill, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                cur = e == SENTINEL ? k : e;
            }
            if (cur == k) {
                slot = (int)pos;
                break;
            }
            pos = (pos + 1) & (NA_LDS - 1);
        }
        if (slot >= 0) {
            __hip_atomic_fetch_add(&ls[slot], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&lc[slot], 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            const uint32_t g = na_slot(t, k);
            __hip_atomic_fetch_add(&t.sums[g], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&t.counts[g], 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < NA_LDS; i += blockDim.x) {
        if (lk[i] == SENTINEL) continue;
        const uint32_t g = na_slot(t, lk[i]);
        __hip_atomic_fetch_add(&t.sums[g], ls[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&t.counts[g], lc[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_compact(NaTable t, int32_t *out_keys, int64_t *out_sums,
                                                         int64_t *out_counts, uint64_t capacity) {
    // every thread of the block takes part in each reservation
    for (uint32_t base = blockIdx.x * blockDim.x; base < NA_SLOTS; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t k = t.keys[i];
        const uint32_t mine = k != SENTINEL;
        const uint64_t row = block_reserve(mine, &t.ctl[0]);
        if (mine && row < capacity) {
            out_keys[row] = (int32_t)(int64_t)k;
            out_sums[row] = (int64_t)t.sums[i];
            out_counts[row] = (int64_t)t.counts[i];
        }
    }
}

__global__ void k_na_total(NaTable t, uint64_t *out_groups) { *out_groups = t.ctl[0]; }

constexpr int NA_MAX_DEVICES = 64;
uint64_t *g_na_scratch[NA_MAX_DEVICES] = {};  // scratch's address on each device
int g_na_cus[NA_MAX_DEVICES] = {};
std::mutex g_na_mutex;

}  // namespace gbagg

extern "C" {

gbagg_status gbagg_hash_aggregate_i32_noalloc(const int32_t *keys, const int32_t *values, uint64_t count,
                                              int32_t *out_keys, int64_t *out_sums, int64_t *out_counts,
                                              uint64_t capacity, uint64_t *out_groups, void *stream) {
    using namespace gbagg;
    auto err = [](gbagg_status st, const char *msg) {
        g_na_error = msg;
        return st;
    };
    if (count > NA_MAX_ROWS) return err(GBAGG_INVALID, "gbagg_hash_aggregate_i32_noalloc: at most 524288 rows");
    if (((!keys || !values) && count) || !out_groups || (capacity && (!out_keys || !out_sums || !out_counts)))
        return err(GBAGG_INVALID, "gbagg_hash_aggregate_i32_noalloc: null argument");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES)
        return err(GBAGG_ERROR, "gbagg_hash_aggregate_i32_noalloc: no usable current device");
    uint64_t *base;
    {
        std::lock_guard<std::mutex> lock(g_na_mutex);
        if (!g_na_scratch[dev]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
                return err(GBAGG_ERROR, "gbagg_hash_aggregate_i32_noalloc: hipGetSymbolAddress(scratch) failed");
            int cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
                return err(GBAGG_ERROR, "gbagg_hash_aggregate_i32_noalloc: no compute unit count");
            g_na_scratch[dev] = (uint64_t *)p;
            g_na_cus[dev] = cus;
        }
        base = g_na_scratch[dev];
    }
    const NaTable t = na_table(base);
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t cus = (uint32_t)g_na_cus[dev];
    hipLaunchKernelGGL(k_na_clear, dim3(cus * 4), dim3(NA_BLOCK), 0, s, t);
    if (count) {
        const uint64_t blocks = (count + NA_BLOCK * 4 - 1) / (NA_BLOCK * 4);
        const uint32_t g = (uint32_t)std::min<uint64_t>(blocks, (uint64_t)cus * 4);
        hipLaunchKernelGGL(k_na_agg, dim3(g), dim3(NA_BLOCK), 0, s, keys, values, count, t);
    }
    hipLaunchKernelGGL(k_na_compact, dim3(cus * 4), dim3(NA_BLOCK), 0, s, t, out_keys, out_sums, out_counts,
                       capacity);
    hipLaunchKernelGGL(k_na_total, dim3(1), dim3(1), 0, s, t, out_groups);
    if (hipGetLastError() != hipSuccess) return err(GBAGG_ERROR, "gbagg_hash_aggregate_i32_noalloc: launch failed");
    return GBAGG_OK;
}

}  // extern "C"
