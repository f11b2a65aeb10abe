// Host-side checks of the exact double SUM arithmetic shared with the device
// (slothdb_amd/csrc/gbagg_device.hpp): fx_scale, the hi/lo split, the LO word
// combination with flags, fx_value's single rounding and fx_hazard. Built with
// hipcc (the header is HIP), run on the CPU.
#include "../../slothdb_amd/csrc/gbagg_device.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace gbagg;

static int failures = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
            failures++;                                \
        }                                              \
    } while (0)

static uint64_t bits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }

// What a row adds to the two words (host mirror of fx_part without the
// overflow flag).
static void parts(double v, int s, uint64_t &hi, uint64_t &lo) {
    bool ok, inexact;
    const int64_t x = fx_scale(bits(v), s, ok, inexact);
    hi = (uint64_t)(x >> FX_LO_BITS);
    lo = ((uint64_t)x & ((1ULL << FX_LO_BITS) - 1)) | (inexact ? FX_INEXACT : 0);
}

int main() {
    std::mt19937_64 rng(12345);
    // 1. fx_scale is round-half-even of v * 2^s, flags what it rounds, and
    //    rejects |v| >= 2^(63 - s)
    for (int i = 0; i < 200000; i++) {
        const int s = (int)(rng() % 200) - 60;
        double v = std::ldexp((double)(rng() >> 11) / 9007199254740992.0, (int)(rng() % 140) - 70);
        if (rng() & 1) v = -v;
        bool ok, inexact;
        const int64_t x = fx_scale(bits(v), s, ok, inexact);
        const double scaled = std::ldexp(v, s);  // exact: no overflow or underflow in this range
        const bool fits = std::fabs(scaled) < 9223372036854775808.0;
        CHECK(ok == fits, "bound v=%a s=%d", v, s);
        if (!ok) continue;
        const double r = std::nearbyint(scaled);  // round half to even
        CHECK(x == (int64_t)r, "scale v=%a s=%d x=%lld want %.0f", v, s, (long long)x, r);
        CHECK(inexact == (r != scaled), "inexact v=%a s=%d", v, s);
    }
    // subnormals and zeros
    {
        bool ok, inexact;
        CHECK(fx_scale(bits(5e-324), FX_SHIFT_MAX, ok, inexact) == 1024 && ok && !inexact, "min subnormal");
        CHECK(fx_scale(bits(-0.0), 10, ok, inexact) == 0 && ok && !inexact, "-0.0");
    }
    // 2. sums: the words of many rows, combined in any order, give the
    //    correctly rounded exact sum of the rounded inputs
    for (int t = 0; t < 2000; t++) {
        const int n = 1 + (int)(rng() % 3000);
        const int emax = (int)(rng() % 80) - 40;
        const int s = 62 - emax;  // |v| < 2^(emax + 1) fits
        std::vector<double> v(n);
        for (auto &x : v) {
            x = std::ldexp((double)(rng() >> 11) / 9007199254740992.0, emax - (int)(rng() % 70));
            if (rng() & 1) x = -x;
        }
        if (t % 3 == 0)  // cancellation: big values that cancel, plus small ones
            for (int i = 0; i + 1 < n; i += 2) v[i + 1] = -v[i] + (rng() % 7 == 0 ? std::ldexp(1.0, emax - 50) : 0.0);
        uint64_t hi = 0, lo = 0, hi2 = 0, lo2 = 0;
        __int128 exact = 0;
        bool any_inexact = false;
        for (int i = 0; i < n; i++) {
            uint64_t h, l;
            parts(v[i], s, h, l);
            hi += h;
            lo = fx_lo_combine(lo, l);
            bool ok, inexact;
            exact += fx_scale(bits(v[i]), s, ok, inexact);
            any_inexact |= inexact;
        }
        for (int i = n - 1; i >= 0; i--) {  // reverse order: the same words
            uint64_t h, l;
            parts(v[i], s, h, l);
            hi2 += h;
            lo2 = fx_lo_combine(l, lo2);
        }
        CHECK(hi == hi2 && lo == lo2, "order dependence n=%d", n);
        CHECK(((lo & FX_INEXACT) != 0) == any_inexact, "inexact flag");
        const double want = std::ldexp((double)exact, -s);  // one rounding (__floattidf)
        const double got = fx_value(hi, lo, s);
        CHECK(got == want, "value n=%d s=%d got %a want %a", n, s, got, want);
    }
    // 3. specials in the low word
    {
        uint64_t lo = 0;
        lo = fx_lo_combine(lo, FX_POS_INF);
        CHECK(fx_value(0, lo, 0) == INFINITY, "+inf");
        lo = fx_lo_combine(lo, 5);
        CHECK(fx_value(0, lo, 0) == INFINITY, "+inf with finite");
        lo = fx_lo_combine(lo, FX_NEG_INF);
        CHECK(std::isnan(fx_value(0, lo, 0)), "+inf and -inf");
        CHECK(std::isnan(fx_value(0, FX_SPECIAL, 0)), "nan");
        CHECK(fx_value(7, FX_NEG_INF, 3) == -INFINITY, "-inf");
    }
    // 4. the reference-test cancellation: 1e16, 1, -1e16, 1 (x1000) sums to 2000 exactly
    {
        uint64_t hi = 0, lo = 0;
        const int s = 1085 - (int)fx_exp_of(bits(1e16));
        for (int i = 0; i < 1000; i++)
            for (double x : {1e16, 1.0, -1e16, 1.0}) {
                uint64_t h, l;
                parts(x, s, h, l);
                hi += h;
                lo = fx_lo_combine(lo, l);
            }
        CHECK(fx_value(hi, lo, s) == 2000.0, "cancellation %a", fx_value(hi, lo, s));
        CHECK(!fx_hazard(hi, lo, 4000), "no hazard when exact");
    }
    // 5. hazard: rounded inputs whose sum is small against their count
    {
        uint64_t hi = 0, lo = 0;
        const int s = 10;
        for (int i = 0; i < 100; i++) {
            uint64_t h, l;
            parts(1e-5 * (i + 1), s, h, l);
            hi += h;
            lo = fx_lo_combine(lo, l);
        }
        CHECK(fx_hazard(hi, lo, 100), "tiny rounded inputs");
        uint64_t h2 = 0, l2 = 0;
        for (int i = 0; i < 100; i++) {
            uint64_t h, l;
            parts(1e6 + 1e-9 * i, 30, h, l);
            h2 += h;
            l2 = fx_lo_combine(l2, l);
        }
        CHECK((l2 & FX_INEXACT) && !fx_hazard(h2, l2, 100), "large sum of rounded inputs");
    }
    // 6. MIN/MAX images: NaN of either sign loses to every number, numbers keep order
    {
        const double nn = std::nan(""), neg_nan = -std::nan("");
        CHECK(f64_ord_mm(bits(nn), true) > f64_ord_mm(bits(INFINITY), true), "min nan");
        CHECK(f64_ord_mm(bits(neg_nan), false) < f64_ord_mm(bits(-INFINITY), false), "max nan");
        CHECK(f64_ord_mm(bits(nn), true) < acc_identity(ACC_MIN_F), "min nan vs identity");
        CHECK(f64_ord_mm(bits(nn), false) > acc_identity(ACC_MAX_F), "max nan vs identity");
        CHECK(std::isnan(__builtin_bit_cast(double, f64_unord_mm(ORD_NAN_MIN))), "nan back");
        CHECK(f64_unord_mm(f64_ord_mm(bits(-2.5), true)) == bits(-2.5), "round trip");
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
