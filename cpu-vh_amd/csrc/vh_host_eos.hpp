// Host-side equation of state of the product (drivers need e_F = e(T_F) and
// e_reg = e(T_reg) exactly as the reference's host code computes them) and the
// builder of the glibc powf(x, 0.25f) exception list used by the device.
#pragma once
#include <cstdint>
#include <vector>

namespace vh {

struct HostEos {
    int kind = 0;  // 0 conformal, 1 table
    double e_min = 0, e_spacing = 0, e_max = 0;
    std::vector<double> p_tb, T_tb;

    // rhic/src/EquationOfState.cpp:292-327 on an in-memory copy of the file
    void load(const double *raw);
    double interp(double e, const std::vector<double> &tb) const;  // :329-360
    double pressure(double e) const;                                // :44-107
    double cs2(double e) const;                                     // :109-160
    double temperature(double e) const;                             // :162-199
    double energy_density(double T) const;                          // :201-290
    // True when the reference's table c_s^2 (a one-segment difference quotient
    // clamped to [0.01, 1/3]) is exactly 1/3 for every e: every pressure
    // increment is positive by far more than rounding and below 3*de.
    bool cs2_is_one_third() const;
};

struct QuarterRootExceptions {
    uint32_t lo = 0, hi = 0;  // x-bit range covered: [2^-32, 2^16)
    int shift = 15;
    std::vector<uint32_t> key, val;  // sorted x bits, glibc powf(x, 0.25f) bits
    std::vector<uint32_t> bucket;    // offsets per (x - lo) >> shift, size nb+1
};

// Built once per process from this host's glibc (the library the reference's
// powf resolves to); a few hundred thousand entries.
const QuarterRootExceptions &quarter_root_exceptions();

}  // namespace vh
