// Hyperbolic tangent that rounds exactly like the host C library's tanh, and
// an exponential whose every bit is fixed by its source (fdlibm_exp, below).
//
// The regulator of the viscous step (REG_SCHEME_1,
// rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:775,780) calls the C
// library's double tanh.  The GNU C library's tanh (2.35, the generic
// double-precision code used on x86-64) follows the fdlibm method
// (Sun Microsystems, freely distributable; K.C. Ng's e_tanh/s_expm1):
//
//   |x| < 2^-55        tanh x = x (1 + x)
//   2^-55 <= |x| < 1   tanh x = -t / (t + 2),    t = expm1(-2|x|)
//   1 <= |x| < 22      tanh x = 1 - 2 / (t + 2), t = expm1(2|x|)
//   |x| >= 22          tanh x = 1
//
// and expm1 reduces x = k ln2 + r with ln2 split into a 32-bit-exact high
// part and a low part, evaluates a rational approximation of r (exp r - 1)
// built from the even polynomial R1(r^2/2) of degree 5 in r^2/2, and scales
// by 2^k.  The library evaluates the polynomial as
// R1 + h2 R2 + h4 R3 (h2 = hxs^2, h4 = h2^2); the code below does the same,
// in the same operand order, with no fused multiply-adds (the file must be
// compiled with -ffp-contract=off), so every intermediate rounds as the
// library's does.  The coefficients are the published fdlibm ones.
//
// The same source runs on the host (tests compare it with the host library
// over 10^8 arguments) and on the device (regulator).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vh {

__host__ __device__ inline uint32_t hi_word(double x) {
    return (uint32_t)((uint64_t)__builtin_bit_cast(uint64_t, x) >> 32);
}
__host__ __device__ inline uint32_t lo_word(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
__host__ __device__ inline double with_hi_word(double x, uint32_t hi) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (b & 0xffffffffull));
}

// exp(x) - 1 for finite x (the tanh argument range needs |x| < 44).
__host__ __device__ inline double glibc_expm1(double x) {
    const double one = 1.0, huge = 1.0e+300, tiny = 1.0e-300;
    const double o_threshold = 7.09782712893383973096e+02;
    const double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42 fee00000
    const double ln2_lo = 1.90821492927058770002e-10;  // 0x3dea39ef 35793c76
    const double invln2 = 1.44269504088896338700e+00;  // 0x3ff71547 652b82fe
    const double Q1 = -3.33333333333331316428e-02;     // 0xbfa11111 111110f4
    const double Q2 = 1.58730158725481460165e-03;      // 0x3f5a01a0 19fe5585
    const double Q3 = -7.93650757867487942473e-05;     // 0xbf14ce19 9eaadbb7
    const double Q4 = 4.00821782732936239552e-06;      // 0x3ed0cfca 86e65239
    const double Q5 = -2.01099218183624371326e-07;     // 0xbe8afdb7 6e09c32d

    uint32_t hx = hi_word(x);
    const uint32_t xsb = hx & 0x80000000u;
    hx &= 0x7fffffffu;
    if (hx >= 0x4043687Au) {  // |x| >= 56 ln2
        if (hx >= 0x40862E42u) {
            if (hx >= 0x7ff00000u) {
                if (((hx & 0xfffffu) | lo_word(x)) != 0) return x + x;
                return xsb == 0 ? x : -1.0;
            }
            if (x > o_threshold) return huge * huge;
        }
        if (xsb != 0) return tiny - one;
    }
    double hi, lo, c = 0.0, t;
    int k;
    if (hx > 0x3fd62e42u) {      // |x| > ln2 / 2
        if (hx < 0x3FF0A2B2u) {  // and |x| < 3 ln2 / 2
            if (xsb == 0) {
                hi = x - ln2_hi;
                lo = ln2_lo;
                k = 1;
            } else {
                hi = x + ln2_hi;
                lo = -ln2_lo;
                k = -1;
            }
        } else {
            k = (int)(invln2 * x + (xsb == 0 ? 0.5 : -0.5));
            t = (double)k;
            hi = x - t * ln2_hi;  // exact
            lo = t * ln2_lo;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (hx < 0x3c900000u) {  // |x| < 2^-54
        t = huge + x;
        return x - (t - (huge + x));
    } else {
        k = 0;
    }
    const double hfx = 0.5 * x;
    const double hxs = x * hfx;
    const double R1 = one + hxs * Q1, h2 = hxs * hxs;
    const double R2 = Q2 + hxs * Q3, h4 = h2 * h2;
    const double R3 = Q4 + hxs * Q5;
    const double r1 = R1 + h2 * R2 + h4 * R3;
    t = 3.0 - r1 * hfx;
    double e = hxs * ((r1 - t) / (6.0 - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e -= hxs;
    if (k == -1) return 0.5 * (x - e) - 0.5;
    if (k == 1) {
        if (x < -0.25) return -2.0 * (e - (x + 0.5));
        return one + 2.0 * (x - e);
    }
    double y;
    if (k <= -2 || k > 56) {
        y = one - (e - x);
        y = with_hi_word(y, hi_word(y) + ((uint32_t)k << 20));
        return y - one;
    }
    if (k < 20) {
        t = with_hi_word(one, 0x3ff00000u - (0x200000u >> k));  // 1 - 2^-k
        y = t - (e - x);
        y = with_hi_word(y, hi_word(y) + ((uint32_t)k << 20));
    } else {
        t = with_hi_word(one, (uint32_t)(0x3ff - k) << 20);  // 2^-k
        y = x - (e + t);
        y += one;
        y = with_hi_word(y, hi_word(y) + ((uint32_t)k << 20));
    }
    return y;
}

__host__ __device__ inline double glibc_tanh(double x) {
    const double one = 1.0, two = 2.0, tiny = 1.0e-300;
    const uint32_t jx = hi_word(x);
    const uint32_t ix = jx & 0x7fffffffu;
    const bool pos = (jx & 0x80000000u) == 0;
    if (ix >= 0x7ff00000u) return pos ? one / x + one : one / x - one;  // inf, NaN
    double z;
    if (ix < 0x40360000u) {  // |x| < 22
        if ((ix | lo_word(x)) == 0) return x;
        if (ix < 0x3c800000u) return x * (one + x);  // |x| < 2^-55
        if (ix >= 0x3ff00000u) {                     // |x| >= 1
            const double t = glibc_expm1(two * __builtin_fabs(x));
            z = one - two / (t + two);
        } else {
            const double t = glibc_expm1(-two * __builtin_fabs(x));
            z = -t / (t + two);
        }
    } else {
        z = one - tiny;
    }
    return pos ? z : -z;
}

// exp(x) by the fdlibm method (Sun's e_exp.c): x = k ln2 + r with ln2
// split into a 32-bit-exact high part and a low part, |r| <= ln2/2, then
// exp(r) = 1 + r + r c/(2 - c) with c = r - r^2 (P1 + r^2 (P2 + ... P5 r^2))
// from the published minimax coefficients, scaled by 2^k.  Within an ulp
// of exp.  It serves the bulk-viscosity extension (zeta/s, reached only with
// the table's own sound speed, which the reference never uses), where the
// device and the tests must agree bit for bit; the host C library's exp is
// a different algorithm.
__host__ __device__ inline double fdlibm_exp(double x) {
    const double one = 1.0, huge = 1.0e+300, twom1000 = 9.33263618503218878990e-302;
    const double o_threshold = 7.09782712893383973096e+02, u_threshold = -7.45133219101941108420e+02;
    const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10;
    const double invln2 = 1.44269504088896338700e+00;
    const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
                 P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
    uint32_t hx = hi_word(x);
    const int xsb = (int)((hx >> 31) & 1);
    hx &= 0x7fffffffu;
    if (hx >= 0x40862E42u) {  // |x| >= 709.78...
        if (hx >= 0x7ff00000u) {
            if (((hx & 0xfffffu) | lo_word(x)) != 0) return x + x;  // NaN
            return xsb == 0 ? x : 0.0;                             // exp(+-inf) = {inf, 0}
        }
        if (x > o_threshold) return huge * huge;
        if (x < u_threshold) return twom1000 * twom1000;
    }
    double hi = 0.0, lo = 0.0;
    int k = 0;
    if (hx > 0x3fd62e42u) {      // |x| > ln2 / 2
        if (hx < 0x3FF0A2B2u) {  // and |x| < 3 ln2 / 2
            hi = xsb == 0 ? x - ln2HI : x + ln2HI;
            lo = xsb == 0 ? ln2LO : -ln2LO;
            k = 1 - xsb - xsb;
        } else {
            k = (int)(invln2 * x + (xsb == 0 ? 0.5 : -0.5));
            const double t = (double)k;
            hi = x - t * ln2HI;  // exact
            lo = t * ln2LO;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {  // |x| < 2^-28
        return one + x;
    }
    const double t = x * x;
    const double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
    if (k == 0) return one - ((x * c) / (c - 2.0) - x);
    const double y = one - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) return with_hi_word(y, hi_word(y) + ((uint32_t)k << 20));
    return with_hi_word(y, hi_word(y) + ((uint32_t)(k + 1000) << 20)) * twom1000;
}

}  // namespace vh
