// Number formatting of the driver's text output: printf's "%.<p>f" (p <= 8)
// and "%g" characters, as the reference's fprintf and ostream write them.
//
// Fast path: the value scaled by an exact power of ten (10^0 .. 10^22 are
// exact doubles) in one multiplication or division, which is within half an
// ulp of the exact scaled value; its rounding to an integer is then the
// correctly rounded one unless the scaled value lies within a few ulps of a
// half-integer.  Those cases, and values out of the fast path's range, go
// through std::to_chars, which rounds exactly (ties to even, as printf).
#pragma once
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace cpuvh {

inline const double *exact_pow10() {
    static const double P[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    return P;
}

// Rounds y >= 0 (y < 2^53, within y*2^-53 of the exact value it stands for)
// to the nearest integer; false when the exact value may be a tie or on the
// other side of one.
inline bool round_scaled(double y, uint64_t *q) {
    const double f = (double)(uint64_t)y;  // floor: 0 <= y < 2^53
    const double r = y - f;                 // exact for y < 2^53
    const double tol = y * 0x1p-51 + 0x1p-60;
    if (std::fabs(r - 0.5) <= tol) return false;
    *q = (uint64_t)f + (r > 0.5 ? 1u : 0u);
    return true;
}

// The decimal digits of v, zero-padded to at least width, written backwards
// so that they end at end; returns their start.
inline char *digits_back(char *end, uint64_t v, int width) {
    char *p = end;
    do {
        *--p = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (end - p < width) *--p = '0';
    return p;
}

inline void put_fixed_slow(std::string &out, double v, int prec) {
    char tmp[400];
    const auto r = std::to_chars(tmp, tmp + sizeof tmp, v, std::chars_format::fixed, prec);
    if (r.ec == std::errc()) {
        out.append(tmp, r.ptr);
    } else {  // beyond 1e390: cannot happen for a finite double at these precisions
        const int len = std::snprintf(tmp, sizeof tmp, "%.*f", prec, v);
        out.append(tmp, (size_t)(len < (int)sizeof tmp - 1 ? len : (int)sizeof tmp - 1));
    }
}

// printf("%.<prec>f", v), 0 <= prec <= 8
inline void put_fixed(std::string &out, double v, int prec) {
    const double a = std::fabs(v);
    uint64_t q;
    if (prec < 0 || prec > 8 || !(a < 9.0e6) || !round_scaled(a * exact_pow10()[prec], &q)) {
        put_fixed_slow(out, v, prec);
        return;
    }
    char dig[24], buf[32];
    const char *d = digits_back(dig + sizeof dig, q, prec + 1);
    const int nint = (int)(dig + sizeof dig - d) - prec;
    char *o = buf;
    if (std::signbit(v)) *o++ = '-';
    for (int i = 0; i < nint; ++i) *o++ = d[i];
    if (prec > 0) {
        *o++ = '.';
        for (int i = 0; i < prec; ++i) *o++ = d[nint + i];
    }
    out.append(buf, (size_t)(o - buf));
}

inline void put_g_slow(std::string &out, double v) {
    char tmp[32];
    const auto r = std::to_chars(tmp, tmp + sizeof tmp, v, std::chars_format::general, 6);
    out.append(tmp, r.ptr);
}

// printf("%g", v): six significant digits, fixed notation for decimal
// exponents -4 .. 5 and exponent notation otherwise, trailing zeros and a
// bare decimal point removed
inline void put_g(std::string &out, double v) {
    const double a = std::fabs(v);
    if (a == 0) {
        out += std::signbit(v) ? "-0" : "0";
        return;
    }
    if (!(a >= 1e-16 && a < 1e27)) {  // out of the exact-power range, inf, nan
        put_g_slow(out, v);
        return;
    }
    const double *P = exact_pow10();
    uint64_t bits;
    std::memcpy(&bits, &a, sizeof bits);
    const int e2 = (int)((bits >> 52) & 0x7ff) - 1023;  // a is normal here: a >= 1e-16
    int X = (int)std::floor(e2 * 0.30102999566398120);  // decimal exponent, maybe one low
    uint64_t q = 0;
    for (int pass = 0;; ++pass) {
        const int k = 5 - X;  // a * 10^k in [1e5, 1e6) when X is right
        if (k > 22 || k < -22 || pass > 2) {
            put_g_slow(out, v);
            return;
        }
        const double y = k >= 0 ? a * P[k] : a / P[-k];
        if (!round_scaled(y, &q)) {
            put_g_slow(out, v);
            return;
        }
        if (q < 100000) {
            --X;
            continue;
        }
        if (q > 1000000) {
            ++X;
            continue;
        }
        if (q == 1000000) {  // rounded up into the next decade
            q = 100000;
            ++X;
        }
        break;
    }
    char d[6];
    for (int i = 5; i >= 0; --i) {
        d[i] = (char)('0' + q % 10);
        q /= 10;
    }
    int nd = 6;
    while (nd > 1 && d[nd - 1] == '0') --nd;
    char buf[32], *o = buf;
    if (std::signbit(v)) *o++ = '-';
    if (X >= -4 && X < 6) {
        if (X >= 0) {
            for (int i = 0; i <= X; ++i) *o++ = d[i];
            if (nd > X + 1) {
                *o++ = '.';
                for (int i = X + 1; i < nd; ++i) *o++ = d[i];
            }
        } else {
            *o++ = '0';
            *o++ = '.';
            for (int i = 0; i < -X - 1; ++i) *o++ = '0';
            for (int i = 0; i < nd; ++i) *o++ = d[i];
        }
    } else {
        *o++ = d[0];
        if (nd > 1) {
            *o++ = '.';
            for (int i = 1; i < nd; ++i) *o++ = d[i];
        }
        *o++ = 'e';
        *o++ = X < 0 ? '-' : '+';
        char ex[8];
        const char *e = digits_back(ex + sizeof ex, (uint64_t)(X < 0 ? -X : X), 2);
        while (e < ex + sizeof ex) *o++ = *e++;
    }
    out.append(buf, (size_t)(o - buf));
}

}  // namespace cpuvh
