// Seeded, counter-based synthetic data for the GROUP BY workloads.
//
// Every value is a pure function of (seed, stream, row), so any row range can
// be generated independently and in parallel, on any machine, and the CPU
// oracle and the GPU operator always see identical inputs.
//
//   rnd(seed, stream, i) = sm64(stream_key(seed, stream) + i)
//   sm64 = SplitMix64 output function (Steele, Lea, Flood 2014; constants
//          from Vigna's reference splitmix64.c), a bijection on uint64.
//
// Distributions:
//   uniform integer [lo, hi]   lo + mulhi64(rnd, hi - lo + 1)   (Lemire 2019)
//   uniform double [0, 1)      (rnd >> 11) * 2^-53
//   bounded Zipf(s, N)         Vose alias table over w_r = (r+1)^-s, r < N;
//                              column = mulhi64(rndA, N), coin = u01(rndB)
//   Normal(mu, sigma)          Box-Muller cosine branch, u1 in (0,1]
//   sparse BIGINT key ids      id = sm64(rank ^ key_salt)  (bijective, so
//                              distinct ranks give distinct ids)
//
// Built as a host-only shared library; it is test and benchmark input
// plumbing and is never called by the aggregate operator itself.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

inline uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint32_t stream) {
    return sm64(seed * 0x9E3779B97F4A7C15ULL ^ ((uint64_t)stream + 1) * 0xD1B54A32D192ED03ULL);
}

inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

template <typename Fn>
void parallel_rows(uint64_t n, int threads, Fn &&fn) {
    if (threads <= 1 || n < 65536) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> ts;
    uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        uint64_t b = per * t, e = std::min(n, b + per);
        if (b >= e) break;
        ts.emplace_back([&, b, e] { fn(b, e); });
    }
    for (auto &t : ts) t.join();
}

// Vose's alias method (Vose 1991, "A linear algorithm for generating random
// numbers with a given distribution"). Built single-threaded so the table is
// bit-identical on every run.
struct Alias {
    double s;
    uint64_t n;
    std::vector<double> prob;
    std::vector<uint32_t> alias;

    Alias(double s_, uint64_t n_) : s(s_), n(n_), prob(n_), alias(n_) {
        std::vector<double> w(n);
        double total = 0.0;
        for (uint64_t r = 0; r < n; r++) {
            w[r] = std::pow((double)(r + 1), -s);
            total += w[r];
        }
        std::vector<uint32_t> small, large;
        small.reserve(n);
        large.reserve(n);
        for (uint64_t r = 0; r < n; r++) {
            w[r] = w[r] * (double)n / total;
            if (w[r] < 1.0) small.push_back((uint32_t)r);
            else large.push_back((uint32_t)r);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t l = small.back(); small.pop_back();
            uint32_t g = large.back(); large.pop_back();
            prob[l] = w[l];
            alias[l] = g;
            w[g] = (w[g] + w[l]) - 1.0;
            if (w[g] < 1.0) small.push_back(g);
            else large.push_back(g);
        }
        for (uint32_t g : large) { prob[g] = 1.0; alias[g] = g; }
        for (uint32_t l : small) { prob[l] = 1.0; alias[l] = l; }
    }

    inline uint64_t sample(uint64_t ra, uint64_t rb) const {
        uint64_t i = mulhi64(ra, n);
        return (u01(rb) < prob[i]) ? i : (uint64_t)alias[i];
    }
};

std::mutex g_alias_mu;
std::vector<std::unique_ptr<Alias>> g_alias_cache;

const Alias *get_alias(double s, uint64_t n) {
    std::lock_guard<std::mutex> lk(g_alias_mu);
    for (auto &a : g_alias_cache)
        if (a->s == s && a->n == n) return a.get();
    if (g_alias_cache.size() >= 8) g_alias_cache.erase(g_alias_cache.begin());
    g_alias_cache.emplace_back(new Alias(s, n));
    return g_alias_cache.back().get();
}

// Stream ids: one per column role, so columns are independent.
enum : uint32_t {
    S_KEY_A = 1, S_KEY_B = 2, S_V1 = 3, S_V2 = 4, S_V3A = 5, S_V3B = 6,
    S_C3_A = 11, S_C3_BA = 12, S_C3_BB = 13, S_C3_C = 14,
    S_DICT_LEN = 21, S_DICT_CHR = 22,
    S_SALT = 31,
};

inline int64_t uniform_i64(uint64_t r, int64_t lo, int64_t hi) {
    uint64_t range = (uint64_t)(hi - lo) + 1;
    return lo + (int64_t)mulhi64(r, range);
}

}  // namespace

extern "C" {

uint64_t sdg_rand(uint64_t seed, uint32_t stream, uint64_t i) {
    return sm64(stream_key(seed, stream) + i);
}

uint64_t sdg_sm64(uint64_t x) { return sm64(x); }

// Sparse BIGINT id for a Zipf rank. Bijective in `rank` for a fixed seed.
int64_t sdg_key_id(uint64_t seed, uint64_t rank) {
    return (int64_t)sm64(rank ^ stream_key(seed, S_SALT));
}

// Bounded Zipf(s, n) ranks for rows [row0, row0+count) of `stream_pair`.
int sdg_zipf_ranks(uint64_t seed, uint32_t stream_a, uint32_t stream_b, double s,
                   uint64_t n, uint64_t row0, uint64_t count, uint64_t *out, int threads) {
    if (n == 0 || n > 0xFFFFFFFFULL) return 1;
    const Alias *al = get_alias(s, n);
    uint64_t ka = stream_key(seed, stream_a), kb = stream_key(seed, stream_b);
    parallel_rows(count, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            out[i] = al->sample(sm64(ka + row0 + i), sm64(kb + row0 + i));
    });
    return 0;
}

// Exact expected number of distinct ranks in `rows` draws from Zipf(s, n):
// sum_r 1 - (1 - p_r)^rows.
double sdg_zipf_expected_distinct(double s, uint64_t n, double rows) {
    double total = 0.0;
    for (uint64_t r = 0; r < n; r++) total += std::pow((double)(r + 1), -s);
    double e = 0.0;
    for (uint64_t r = 0; r < n; r++) {
        double p = std::pow((double)(r + 1), -s) / total;
        e += -std::expm1(rows * std::log1p(-p));
    }
    return e;
}

// Plain aggregate, CPU-runnable workload: k int32 uniform [0, nkeys),
// v int64 uniform [vlo, vhi].
int sdg_gen_plain(uint64_t seed, uint64_t row0, uint64_t n, int64_t nkeys, int64_t vlo,
                  int64_t vhi, int32_t *k, int64_t *v, int threads) {
    uint64_t kk = stream_key(seed, S_KEY_A), kv = stream_key(seed, S_V1);
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            k[i] = (int32_t)uniform_i64(sm64(kk + row0 + i), 0, nkeys - 1);
            v[i] = uniform_i64(sm64(kv + row0 + i), vlo, vhi);
        }
    });
    return 0;
}

// High-cardinality workload: k int64 = sparse id of a Zipf(s, nkeys) rank,
// v1 int64 uniform [-1e9, 1e9], v2 int32 uniform over all int32,
// v3 double Normal(100, 25). Any output pointer may be null (column skipped).
int sdg_gen_highcard(uint64_t seed, uint64_t row0, uint64_t n, uint64_t nkeys, double s,
                     int64_t *k, int64_t *v1, int32_t *v2, double *v3, int threads) {
    if (nkeys == 0 || nkeys > 0xFFFFFFFFULL) return 1;
    const Alias *al = k ? get_alias(s, nkeys) : nullptr;
    uint64_t ka = stream_key(seed, S_KEY_A), kb = stream_key(seed, S_KEY_B);
    uint64_t salt = stream_key(seed, S_SALT);
    uint64_t k1 = stream_key(seed, S_V1), k2 = stream_key(seed, S_V2);
    uint64_t k3a = stream_key(seed, S_V3A), k3b = stream_key(seed, S_V3B);
    const double two_pi = 6.283185307179586;
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint64_t r = row0 + i;
            if (k) {
                uint64_t rank = al->sample(sm64(ka + r), sm64(kb + r));
                k[i] = (int64_t)sm64(rank ^ salt);
            }
            if (v1) v1[i] = uniform_i64(sm64(k1 + r), -1000000000LL, 1000000000LL);
            if (v2) v2[i] = (int32_t)(uint32_t)(sm64(k2 + r) >> 32);
            if (v3) {
                double a = ((double)(sm64(k3a + r) >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
                double c = u01(sm64(k3b + r));
                v3[i] = 100.0 + 25.0 * (std::sqrt(-2.0 * std::log(a)) * std::cos(two_pi * c));
            }
        }
    });
    return 0;
}

// Dictionary of `dict_size` distinct strings, lengths uniform [4, 32],
// alphabet [a-z0-9]. offsets has dict_size + 1 entries; bytes must hold
// dict_size * 32 bytes. Duplicates are redrawn with the next attempt index,
// so the dictionary is deterministic. Returns total bytes used, or -1.
int64_t sdg_gen_dict(uint64_t seed, uint32_t dict_size, uint32_t *offsets, char *bytes) {
    static const char alphabet[] = "abcdefghijklmnopqrstuvwxyz0123456789";
    uint64_t kl = stream_key(seed, S_DICT_LEN), kc = stream_key(seed, S_DICT_CHR);
    std::unordered_set<std::string> seen;
    seen.reserve((size_t)dict_size * 2);
    uint64_t pos = 0;
    std::string s;
    for (uint32_t j = 0; j < dict_size; j++) {
        for (uint64_t attempt = 0;; attempt++) {
            uint64_t id = ((uint64_t)j << 8) | attempt;
            if (attempt > 255) return -1;
            uint32_t len = 4 + (uint32_t)mulhi64(sm64(kl + id), 29);
            s.resize(len);
            for (uint32_t c = 0; c < len; c++)
                s[c] = alphabet[mulhi64(sm64(kc + id * 32 + c), 36)];
            if (seen.insert(s).second) break;
        }
        offsets[j] = (uint32_t)pos;
        std::memcpy(bytes + pos, s.data(), s.size());
        pos += s.size();
    }
    offsets[dict_size] = (uint32_t)pos;
    return (int64_t)pos;
}

// Composite key + filter workload: a int32 uniform [0, a_card), b = Zipf(s, dict_size)
// rank used as the dictionary code, c int32 uniform [0, c_card).
int sdg_gen_composite(uint64_t seed, uint64_t row0, uint64_t n, int64_t a_card, uint32_t dict_size,
                      double s, int64_t c_card, int32_t *a, uint32_t *b, int32_t *c, int threads) {
    const Alias *al = get_alias(s, dict_size);
    uint64_t kA = stream_key(seed, S_C3_A), kBa = stream_key(seed, S_C3_BA);
    uint64_t kBb = stream_key(seed, S_C3_BB), kC = stream_key(seed, S_C3_C);
    parallel_rows(n, threads, [&](uint64_t bb, uint64_t e) {
        for (uint64_t i = bb; i < e; i++) {
            uint64_t r = row0 + i;
            a[i] = (int32_t)uniform_i64(sm64(kA + r), 0, a_card - 1);
            b[i] = (uint32_t)al->sample(sm64(kBa + r), sm64(kBb + r));
            c[i] = (int32_t)uniform_i64(sm64(kC + r), 0, c_card - 1);
        }
    });
    return 0;
}

// Mixed-query workload: one query's int32 keys (Zipf(s, distinct) rank through a
// 32-bit bijective mix) and int64 values uniform [vlo, vhi].
int sdg_gen_query_i32(uint64_t seed, uint64_t n, uint64_t distinct, double s, int64_t vlo,
                      int64_t vhi, int32_t *k, int64_t *v, int threads) {
    if (distinct == 0 || distinct > 0xFFFFFFFFULL) return 1;
    const Alias *al = get_alias(s, distinct);
    uint64_t ka = stream_key(seed, S_KEY_A), kb = stream_key(seed, S_KEY_B);
    uint64_t kv = stream_key(seed, S_V1);
    uint32_t salt = (uint32_t)stream_key(seed, S_SALT);
    parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t rank = (uint32_t)al->sample(sm64(ka + i), sm64(kb + i));
            // 32-bit bijection: xor salt, then an odd multiply and xorshift.
            uint32_t x = rank ^ salt;
            x *= 0x9E3779B1u;
            x ^= x >> 16;
            k[i] = (int32_t)x;
            if (v) v[i] = uniform_i64(sm64(kv + i), vlo, vhi);
        }
    });
    return 0;
}

// A dictionary-coded string column as its strings (offsets + bytes): row i is
// dictionary entry codes[i]. out_offsets has n + 1 entries; out_bytes holds
// the returned byte count (call with out_bytes = nullptr to size it).
int64_t sdg_strings_from_codes(const uint32_t *dict_off, const char *dict_bytes, const uint32_t *codes,
                               uint64_t n, uint32_t *out_offsets, char *out_bytes, int threads) {
    uint64_t pos = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (out_offsets) out_offsets[i] = (uint32_t)pos;
        pos += dict_off[codes[i] + 1] - dict_off[codes[i]];
        if (pos > 0xFFFFFFFFull) return -1;
    }
    if (out_offsets) out_offsets[n] = (uint32_t)pos;
    if (out_bytes && out_offsets)
        parallel_rows(n, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++)
                std::memcpy(out_bytes + out_offsets[i], dict_bytes + dict_off[codes[i]],
                            dict_off[codes[i] + 1] - dict_off[codes[i]]);
        });
    return (int64_t)pos;
}

}  // extern "C"
