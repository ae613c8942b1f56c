// Host-side checks of two pieces of the device arithmetic, compiled from the
// library's own headers (tests/test_libm_host.py builds and runs this):
//
//   libm_check tanh N   vh::glibc_tanh against the host C library's tanh on
//                       N arguments: uniform on [0, 25), log-uniform on
//                       [2^-60, 2^6), negative, and every double within
//                       2^16 ulps of each branch threshold of tanh and expm1;
//   libm_check exp N    vh::fdlibm_exp against the host exp on N arguments in
//                       [-745, 710]: at most one ulp apart (prints the
//                       largest distance in ulps);
//   libm_check div N    vh::DivFast (reciprocal + Markstein correction, with
//                       the division operator outside its proven range)
//                       against the division operator on N pairs with
//                       exponents spread over the whole double range.
//
// Prints "checked <n> mismatches <m>" and the first mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../cpu-vh_amd/csrc/vh_device.hpp"
#include "../cpu-vh_amd/csrc/vh_libm.hpp"

static uint64_t state = 0x2545F4914F6CDD1Dull;
static uint64_t next_u64() {  // SplitMix64
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unit() { return std::ldexp((double)(next_u64() >> 11), -53); }
static uint64_t bits(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}
static double from_bits(uint64_t b) {
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

static long bad = 0, checked = 0;
static void cmp(const char *what, double x, double y, double want, double got) {
    ++checked;
    const bool same = bits(want) == bits(got) || (std::isnan(want) && std::isnan(got));
    if (!same) {
        if (bad < 8) std::printf("%s x=%a y=%a want=%a got=%a\n", what, x, y, want, got);
        ++bad;
    }
}

static void check_tanh(long n) {
    for (long i = 0; i < n; ++i) {
        double x;
        switch (i & 3) {
            case 0: x = 25.0 * unit(); break;
            case 1: x = std::exp2(-60.0 + 66.0 * unit()); break;
            case 2: x = -std::exp2(-60.0 + 66.0 * unit()); break;
            default: x = 1.5 * unit(); break;
        }
        cmp("tanh", x, 0, std::tanh(x), vh::glibc_tanh(x));
    }
    // branch thresholds of tanh (2^-55, 1, 22) and of expm1 at x = 2|t|
    // (ln2/2, 3 ln2/2, 2^-54, 56 ln2)
    const double th[] = {0x1p-55, 1.0, 22.0, 0.5 * 0.6931471805599453, 0.5 * 1.0397207708399179, 0x1p-55,
                         0.5 * 38.816242111356935};
    for (double t : th)
        for (int64_t d = -(1 << 16); d <= (1 << 16); ++d) {
            const double x = from_bits(bits(t) + d);
            cmp("tanh", x, 0, std::tanh(x), vh::glibc_tanh(x));
            cmp("tanh", -x, 0, std::tanh(-x), vh::glibc_tanh(-x));
        }
    const double special[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 0x1p-1074, 0x1p-1022, 1e-300, 700.0, 1e300};
    for (double x : special) cmp("tanh", x, 0, std::tanh(x), vh::glibc_tanh(x));
}

static long max_ulps = 0;
static void check_exp(long n) {
    for (long i = 0; i < n; ++i) {
        double x;
        switch (i & 3) {
            case 0: x = -745.0 + 1455.0 * unit(); break;  // the whole finite range
            case 1: x = -30.0 + 40.0 * unit(); break;      // what zeta/s feeds it
            case 2: x = std::ldexp(unit() - 0.5, -20); break;
            default: x = -3.0 + 6.0 * unit(); break;
        }
        const double want = std::exp(x), got = vh::fdlibm_exp(x);
        ++checked;
        const long d = (long)(bits(want) > bits(got) ? bits(want) - bits(got) : bits(got) - bits(want));
        if (d > max_ulps) max_ulps = d;
        if (d > 1) {
            if (bad < 8) std::printf("exp x=%a want=%a got=%a (%ld ulps)\n", x, want, got, d);
            ++bad;
        }
    }
    std::printf("largest distance %ld ulps\n", max_ulps);
}

static double spread(int emin, int emax) {
    const uint64_t r = next_u64();
    const double m = 1.0 + unit();
    const int e = emin + (int)(r % (uint64_t)(emax - emin + 1));
    const double v = std::ldexp(m, e);
    return (r >> 63) ? -v : v;
}

static void check_div(long n) {
    for (long i = 0; i < n; ++i) {
        double a, b;
        switch (i & 3) {
            case 0: a = spread(-1074, 1023); b = spread(-1074, 1023); break;  // everything
            case 1: a = spread(-60, 60); b = spread(-60, 60); break;          // the values the step sees
            case 2: a = spread(-900, -800); b = spread(-100, 100); break;     // near the low edge of the range
            default: a = spread(900, 1023); b = spread(-100, 100); break;     // near the high edge
        }
        if ((i & 255) == 7) a = 0.0;
        const vh::Recip R = vh::make_recip(b);
        cmp("div", a, b, a / b, vh::DivFast()(a, R));
    }
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const long n = std::atol(argv[2]);
    if (!std::strcmp(argv[1], "tanh")) check_tanh(n);
    else if (!std::strcmp(argv[1], "div")) check_div(n);
    else if (!std::strcmp(argv[1], "exp")) check_exp(n);
    else return 2;
    std::printf("checked %ld mismatches %ld\n", checked, bad);
    return bad ? 1 : 0;
}
