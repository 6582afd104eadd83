// Host check of the face kernels' quotient by 6 (fvg::quot6 in fenice_amd/csrc/fvg_device.hpp, compiled here for the
// host from the same source): every argument in its range must give the bits of the IEEE division x / 6.0, and
// range_fits_quot6 must bound the WENO5-Z numerators built from values it accepts. Prints one line per class of
// arguments with the number tried and the number that differ; exits nonzero on any difference.
// usage: quot6_check <samples per class> <seed>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../fenice_amd/csrc/fvg_device.hpp"

namespace {

uint64_t g_state;
uint64_t next_u64() {   // splitmix64
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double bits_to(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
uint64_t to_bits(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
}
// a double of either sign with a random significand and a binary exponent uniform in [lo, hi]
double random_in(int lo, int hi) {
    const uint64_t r = next_u64();
    const int e = lo + (int)(next_u64() % (uint64_t)(hi - lo + 1));
    const uint64_t sig = r & ((1ull << 52) - 1);
    const uint64_t b = ((uint64_t)(e + 1023) << 52) | sig | ((r >> 63) << 63);
    return bits_to(b);
}

long long g_bad = 0;
bool check(double x) {
    const volatile double six = 6.0;   // the host's IEEE division, not folded into a multiplication
    const double want = x / six, got = fvg::quot6(x);
    if (to_bits(want) != to_bits(got)) {
        if (g_bad < 5) std::printf("  differs: x = %a: x / 6 = %a, quot6 = %a\n", x, want, got);
        ++g_bad;
        return false;
    }
    return true;
}

}   // namespace

int main(int argc, char** argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 1000000;
    g_state = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    long long tried, bad;
    auto report = [&](const char* what) {
        std::printf("%-58s tried %lld differ %lld\n", what, tried, g_bad - bad);
    };

    // 1. random significands over the whole range of quot6
    tried = 0, bad = g_bad;
    for (long long s = 0; s < n; ++s, ++tried) check(random_in(-960, 904));
    report("random significands, 2^-960 .. 2^905");

    // 2. multiples of 6 (exact quotients) and their neighbours within 4 ulp
    tried = 0, bad = g_bad;
    for (long long s = 0; s < n / 9; ++s) {
        const double q = random_in(-955, 900), x = q * 6.0;   // 6 q is exact unless q's significand needs > 50 bits
        double y = x;
        for (int k = 0; k < 4; ++k) y = std::nextafter(y, 0.0);
        for (int k = 0; k < 9; ++k, ++tried, y = std::nextafter(y, 2.0 * x)) check(y);
    }
    report("multiples of 6 and neighbours within 4 ulp");

    // 3. arguments whose quotient lies next to a power of two (x near 6 * 2^e, both sides)
    tried = 0, bad = g_bad;
    for (int e = -955; e <= 900; ++e) {
        const double x = std::ldexp(6.0, e);
        double lo = x, hi = x;
        for (int k = 0; k < 64; ++k, tried += 4) {
            lo = std::nextafter(lo, 0.0);
            hi = std::nextafter(hi, 2.0 * x);
            check(lo);
            check(hi);
            check(-lo);
            check(-hi);
        }
    }
    report("quotients next to powers of two");

    // 4. the ends of the range and zeros
    tried = 0, bad = g_bad;
    const double ends[] = {0.0, -0.0, 0x1p-960, -0x1p-960, 0x1p905, -0x1p905, 0x1.fffffffffffffp904, 0x1.0000000000001p-960,
                           1.0, -1.0, 3.0, 6.0, 9.0, 1.0 / 3.0, 2.0 / 3.0};
    for (double x : ends) {
        check(x);
        ++tried;
    }
    report("zeros, range ends, small integers");

    // 5. WENO5-Z candidate numerators from stencil values that range_fits_quot6 accepts (including exact zeros and
    //    values at the ends of its range); each numerator must be in quot6's range and give the division's bits
    tried = 0, bad = g_bad;
    long long out_of_range = 0;
    for (long long s = 0; s < n / 3; ++s) {
        double c[5];
        for (int q = 0; q < 5; ++q) {
            const uint64_t r = next_u64() % 16;
            if (r == 0) c[q] = 0.0;
            else if (r == 1) c[q] = random_in(-900, -890);
            else if (r == 2) c[q] = random_in(890, 899);
            else if (r < 6) c[q] = random_in(-4, 4);   // clustered magnitudes: cancellation in the sums
            else c[q] = random_in(-60, 60);
            if (!fvg::range_fits_quot6(c[q])) return 2;
        }
        const uint64_t shape = next_u64() % 8;
        if (shape == 0) c[0] = c[1] = c[2] = c[3] = c[4] = 0.0;            // zero numerators
        else if (shape == 1) c[1] = c[2] = c[3] = c[4] = c[0];               // numerators 6 c: exact quotients
        const double num[3] = {2.0 * c[0] - 7.0 * c[1] + 11.0 * c[2], -c[1] + 5.0 * c[2] + 2.0 * c[3],
                               2.0 * c[2] + 5.0 * c[3] - c[4]};
        for (double x : num) {
            const double m = std::fabs(x);
            if (!(x == 0.0 || (m >= 0x1p-960 && m <= 0x1p905))) ++out_of_range;
            check(x);
            ++tried;
        }
    }
    report("WENO5-Z numerators of values in range_fits_quot6");
    std::printf("numerators outside quot6's range: %lld\n", out_of_range);

    // 6. range_fits_quot6 itself: its bounds, non-finite values and NaN
    const bool fits_ok = fvg::range_fits_quot6(0.0) && fvg::range_fits_quot6(-0.0) && fvg::range_fits_quot6(0x1p-900) &&
                         fvg::range_fits_quot6(-0x1p900) && !fvg::range_fits_quot6(0x1.fffffffffffffp-901) &&
                         !fvg::range_fits_quot6(0x1.0000000000001p900) && !fvg::range_fits_quot6(INFINITY) &&
                         !fvg::range_fits_quot6(-INFINITY) && !fvg::range_fits_quot6(NAN) &&
                         !fvg::range_fits_quot6(4.9e-324);
    std::printf("range_fits_quot6 bounds: %s\n", fits_ok ? "ok" : "WRONG");

    // 7. outside the range the three-operation form is not the division (why callers keep the IEEE division there)
    const double sub = 9.0 * 0x1p-1074;
    const volatile double six = 6.0;
    std::printf("outside the range: x = 9 * 2^-1074: x / 6 = %a, quot6 = %a\n", sub / six, fvg::quot6(sub));

    std::printf("total differing: %lld\n", g_bad);
    return (g_bad == 0 && out_of_range == 0 && fits_ok) ? 0 : 1;
}
