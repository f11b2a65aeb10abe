// ref_shim.cpp -- extern "C" entry points over the reference's own aggregators,
// compiled together with the reference sources where they lie
// (src/execution/inline_row_agg.cpp, src/execution/radix_count_agg.cpp and the
// vendored include/third_party/unordered_dense.h) into oracle/_ref. TEST
// INFRASTRUCTURE ONLY: used to pin the oracle (tests/test_oracle.py) and to
// write golden vectors (tests/golden/make_golden.py).

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "slothdb/execution/inline_row_agg.hpp"
#include "slothdb/execution/radix_count_agg.hpp"
#include "third_party/unordered_dense.h"

using namespace slothdb;

extern "C" {

uint64_t ref_hash_u64(uint64_t x) { return ankerl::unordered_dense::hash<uint64_t>{}(x); }
uint64_t ref_hash_i64(int64_t x) { return ankerl::unordered_dense::hash<int64_t>{}(x); }
uint64_t ref_hash_i32(int32_t x) { return ankerl::unordered_dense::hash<int32_t>{}(x); }

// RadixCountAgg driven the way physical_planner.cpp:7519-7586 drives it:
// per-thread ScatterRow, one MergeRadix per radix, EmitTopK(k) (k <= 0: all).
int ref_radix_count(const int64_t *keys, uint64_t n, int threads, int k, int64_t *out_keys,
                    int64_t *out_counts, uint64_t cap, uint64_t *out_n) {
    if (threads < 1) threads = 1;
    RadixCountAgg agg(threads);
    agg.ReserveExpectedRows((int64_t)n);
    std::vector<std::thread> ts;
    uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        uint64_t b = std::min(n, per * t), e = std::min(n, per * (t + 1));
        ts.emplace_back([&, t, b, e] {
            for (uint64_t r = b; r < e; r++) agg.ScatterRow(t, keys[r]);
        });
    }
    for (auto &t : ts) t.join();
    ts.clear();
    for (int r = 0; r < RadixCountAgg::N_RADIX; r++)
        ts.emplace_back([&, r] { agg.MergeRadix(r); });
    for (auto &t : ts) t.join();
    auto res = agg.EmitTopK(k);
    if (res.size() > cap) return 1;
    for (size_t i = 0; i < res.size(); i++) {
        out_keys[i] = res[i].key;
        out_counts[i] = res[i].count;
    }
    *out_n = res.size();
    return 0;
}

}  // extern "C"

// InlineRowAgg<N> (inline_row_agg.cpp:22-266) over packed uint64 keys with N
// int64 aggregate inputs laid out column-major vals[a * n + r] and validity
// valid[a * n + r]. Output rows: key, count_star, sum[N], cnt[N].
template <int N>
static int run_inline(const uint64_t *keys, const int64_t *vals, const uint8_t *valid, uint64_t n,
                      int threads, int k, uint64_t *out_key, int64_t *out_cs, int64_t *out_sum,
                      int64_t *out_cnt, uint64_t cap, uint64_t *out_n) {
    InlineRowAgg<N> agg(threads);
    std::vector<std::thread> ts;
    uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        uint64_t b = std::min(n, per * t), e = std::min(n, per * (t + 1));
        ts.emplace_back([&, t, b, e] {
            int64_t v[N];
            uint8_t ok[N];
            for (uint64_t r = b; r < e; r++) {
                for (int a = 0; a < N; a++) {
                    v[a] = vals[(uint64_t)a * n + r];
                    ok[a] = valid ? valid[(uint64_t)a * n + r] : 1;
                }
                agg.Update(t, keys[r], v, ok);
            }
        });
    }
    for (auto &t : ts) t.join();
    ts.clear();
    for (int s = 0; s < InlineRowAgg<N>::N_RADIX; s++) ts.emplace_back([&, s] { agg.MergeShard(s); });
    for (auto &t : ts) t.join();
    auto res = agg.EmitTopK(k);
    if (res.size() > cap) return 1;
    for (size_t i = 0; i < res.size(); i++) {
        out_key[i] = res[i].key;
        out_cs[i] = res[i].count_star;
        for (int a = 0; a < N; a++) {
            out_sum[i * N + a] = res[i].sum[a];
            out_cnt[i * N + a] = res[i].cnt[a];
        }
    }
    *out_n = res.size();
    return 0;
}

extern "C" int ref_inline_row_agg(int num_aggs, const uint64_t *keys, const int64_t *vals,
                       const uint8_t *valid, uint64_t n, int threads, int k, uint64_t *out_key,
                       int64_t *out_cs, int64_t *out_sum, int64_t *out_cnt, uint64_t cap,
                       uint64_t *out_n) {
    if (threads < 1) threads = 1;
    switch (num_aggs) {
    case 1: return run_inline<1>(keys, vals, valid, n, threads, k, out_key, out_cs, out_sum, out_cnt, cap, out_n);
    case 2: return run_inline<2>(keys, vals, valid, n, threads, k, out_key, out_cs, out_sum, out_cnt, cap, out_n);
    case 3: return run_inline<3>(keys, vals, valid, n, threads, k, out_key, out_cs, out_sum, out_cnt, cap, out_n);
    case 4: return run_inline<4>(keys, vals, valid, n, threads, k, out_key, out_cs, out_sum, out_cnt, cap, out_n);
    default: return 2;
    }
}
