// Device-side physics of the viscous-hydro step (gfx950, FP64).
//
// Arithmetic is written to round exactly like cpu-vh's compiled C++: every
// expression keeps the reference's operand order, the file is compiled with
// -ffp-contract=off (no fused multiply-add unless written as fma()), and the
// reference's single-precision islands (sqrtf/fabsf/powf on double operands)
// are reproduced bit for bit:
//   * sqrtf(x)   == (float)sqrt((double)x): the double square root of a float
//                   rounded back to float is the correctly rounded float root;
//   * powf(x,.25f) (glibc) == the correctly rounded fourth root, except on a
//                   list of inputs where glibc's result differs; that list is
//                   computed at start-up from the host's own glibc and looked
//                   up here (eos_fourth_root).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace vh {

constexpr double HBARC = 0.197326938;   // rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:886
constexpr double EOS_FACTOR = 13.8997;  // rhic/include/EquationOfState.h:19
constexpr double THETA = 1.1;           // rhic/include/FluxLimiter.h:13
constexpr int EOS_LEN = 100000;         // rhic/src/EquationOfState.cpp:295
constexpr int MAX_ITERS_NR = 10000;     // rhic/src/EnergyMomentumTensor.cpp:18-20
constexpr int MAX_ITERS_BI = 1000000;
constexpr double ACC = 1.0e-3;
constexpr double delta_pipi = 1.33333;  // rhic/src/SourceTerms.cpp:50-53
constexpr double tau_pipi = 1.42857;
constexpr double delta_PiPi = 0.666667;
constexpr double lambda_piPi = 1.2;

enum : int { NAN_NONE = 0, NAN_ENERGY = 1, NAN_ROOT = 2, NAN_REG_SHEAR = 3, NAN_REG_BULK = 4 };

// Division by a divisor used several times: one correctly rounded reciprocal
// r = RN(1/b) and, per numerator, q0 = RN(a*r), rem = a - b*q0 (exact with
// FMA), q = RN(q0 + rem*r).  By Markstein's theorem (IBM J. Res. Dev. 34,
// 1990; Muller et al., Handbook of Floating-Point Arithmetic, ch. 5) q is the
// correctly rounded a/b, i.e. the division operator's value, when nothing
// underflows or overflows.  Sufficient here: |b| in [2^-100, 2^100] (so r is
// normal) and |q0| in [2^-860, 2^1000] (then rem is a multiple of 2^-1064,
// exactly representable, and q is normal); a == +-0 gives a*r, the
// correctly signed zero.
//
// DivFast uses the reciprocal form and, for the lanes whose quotient leaves
// that range (tiny or huge values, NaN), the division operator; DivExact
// always uses the division operator.  Either way every result equals the
// division operator's.
struct Recip {
    double b, r;
    bool ok;
};
__host__ __device__ __forceinline__ Recip make_recip(double b) {
    Recip R;
    R.b = b;
    R.r = 1.0 / b;
    const double ab = __builtin_fabs(b);
    R.ok = ab >= 0x1p-100 && ab <= 0x1p100;
    return R;
}
struct DivFast {
    __host__ __device__ __forceinline__ double operator()(double a, const Recip &R) {
        const double q0 = a * R.r;
        const double rem = __builtin_fma(-R.b, q0, a);
        double q = __builtin_fma(rem, R.r, q0);
        const double aq = __builtin_fabs(q0);
        const bool zero = a == 0.0;
        if (zero) q = q0;
        if (__builtin_expect(!(R.ok && ((aq >= 0x1p-860 && aq <= 0x1p1000) || zero)), 0)) q = a / R.b;
        return q;
    }
};
struct DivExact {
    __device__ __forceinline__ double operator()(double a, const Recip &R) { return a / R.b; }
};

// Quotient policies of the root solve and the primitive recovery.  QuotExact
// is the division operator (the FP64 path).  QuotFast (mixed form only) is a
// hardware reciprocal refined by two Newton steps and one residual
// correction: within an ulp of the quotient, not correctly rounded.  With
// DivRecip the table lookups multiply by the precomputed reciprocal spacing.
struct QuotExact {
    static __device__ __forceinline__ double q(double a, double b) { return a / b; }
    // the reference's sqrtf of a double (a float island): correctly rounded
    static __device__ __forceinline__ float sqrt_island(float x) { return (float)__builtin_sqrt((double)x); }
};
struct QuotFast {
    // hardware reciprocal and one Newton step: relative error about 2^-46
    static __device__ __forceinline__ double rcp1(double b) {
        const double r = __builtin_amdgcn_rcp(b);
        return __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    }
    static __device__ __forceinline__ double rcp(double b) {
        double r = __builtin_amdgcn_rcp(b);
        r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
        return __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    }
    static __device__ __forceinline__ double q(double a, double b) {
        double r = __builtin_amdgcn_rcp(b);
        r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
        r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
        const double q0 = a * r;
        return __builtin_fma(__builtin_fma(-b, q0, a), r, q0);
    }
    // the hardware's single-precision square root (within an ulp)
    static __device__ __forceinline__ float sqrt_island(float x) { return __builtin_amdgcn_sqrtf(x); }
};
struct DivRecip {
    __device__ __forceinline__ double operator()(double a, const Recip &R) { return a * R.r; }
};

struct EosDev {
    int kind;        // 0 conformal, 1 table
    int cs2_const;   // table: host verified that the reference's c_s^2 is 1/3 for every e
    int cs2_mode;    // table: 0 the reference's delta_e/delta_p, 1 the sound speed delta_p/delta_e
    double e_min, e_spacing, e_max;
    double r_spacing;  // RN(1/e_spacing)
    const double *p_tb;
    const double *T_tb;
    // FP32 segments of the pressure table for the mixed form's face solves:
    // entry k = {p_k, d_k, d_k / de, (d_{k+1} - d_k) / de}, d_k = p_{k+1} - p_k
    // formed in FP64 (the last segment's slope continues past the table)
    const float4 *p4;
    // glibc powf(x, 0.25f) exceptions: sorted (x bits, result bits) pairs,
    // bucketed by (xbits - exc_lo) >> exc_shift.
    const uint2 *exc;
    const uint32_t *exc_bucket;
    uint32_t exc_lo, exc_hi;
    int exc_shift;
};

__device__ __forceinline__ float f32_sqrt(float x) { return (float)__builtin_sqrt((double)x); }
__device__ __forceinline__ float f32_fabs(float x) { return __builtin_fabsf(x); }

// glibc powf(x, 0.25f) for the conformal temperature (EquationOfState.cpp:196).
__device__ __forceinline__ float eos_fourth_root(const EosDev &E, float x) {
    float r = (float)__builtin_sqrt(__builtin_sqrt((double)x));
    uint32_t xb = __float_as_uint(x);
    if (xb >= E.exc_lo && xb < E.exc_hi) {
        uint32_t b = (xb - E.exc_lo) >> E.exc_shift;
        uint32_t lo = E.exc_bucket[b], hi = E.exc_bucket[b + 1];
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            uint32_t k = E.exc[mid].x;
            if (k < xb) lo = mid + 1;
            else hi = mid;
        }
        if (lo < E.exc_bucket[b + 1] && E.exc[lo].x == xb) r = __uint_as_float(E.exc[lo].y);
    }
    return r;
}

// rhic/src/EquationOfState.cpp:329-360
template <class DV>
__device__ __forceinline__ double eos_interp(const EosDev &E, double e, const double *__restrict__ tb, DV &div) {
    const double e0 = E.e_min, de = E.e_spacing;
    const Recip Rde{de, E.r_spacing, true};
    int idx = (int)div(e - e0, Rde);
    idx = idx < EOS_LEN - 2 ? idx : EOS_LEN - 2;
    idx = idx > 0 ? idx : 0;
    const double frac = div(e - (idx * de + e0), Rde);
    double t1 = tb[idx], t2 = tb[idx + 1];
    return t1 * (1. - frac) + t2 * frac;
}

template <bool TABLE, class DV>
__device__ __forceinline__ double eos_pressure(const EosDev &E, double e, DV &div) {
    if constexpr (TABLE) {
        double f = eos_interp(E, e, E.p_tb, div);
        return 1.e-15 < f ? f : 1.e-15;  // std::max(1.e-15, f)
    } else {
        return div(e, Recip{3., 1. / 3., true});
    }
}

// Mixed form: the same linear interpolation from one scaled coordinate
// x = (e - e_min) / de (idx = trunc(x), frac = x - idx), within a few ulps of
// the reference's two divisions.
__device__ __forceinline__ double eos_pressure_fast(const EosDev &E, double e) {
    const double x = (e - E.e_min) * E.r_spacing;
    int idx = (int)x;
    idx = idx < EOS_LEN - 2 ? idx : EOS_LEN - 2;
    idx = idx > 0 ? idx : 0;
    const double frac = x - (double)idx;
    const double t1 = E.p_tb[idx], t2 = E.p_tb[idx + 1];
    const double f = __builtin_fma(frac, t2 - t1, t1);
    return 1.e-15 < f ? f : 1.e-15;
}

template <bool TABLE, class DV>
__device__ __forceinline__ double eos_cs2(const EosDev &E, double e, DV &div) {
    if constexpr (TABLE) {
        if (E.cs2_const) return 1. / 3.;
        double v_min = 0.01, v_max = 1. / 3.;
        double delta_e = E.e_spacing;
        double p = eos_pressure<true>(E, e, div);
        double p_r = eos_pressure<true>(E, e + delta_e, div);
        double delta_p = p_r - p;
        double v = E.cs2_mode ? delta_p / delta_e : delta_e / delta_p;
        v = v_max < v ? v_max : v;  // std::min(v_max, v)
        v = v_min < v ? v : v_min;  // std::max(v_min, .)
        return v;
    } else {
        return 1. / 3.;
    }
}

template <bool TABLE, class DV>
__device__ __forceinline__ double eos_temperature(const EosDev &E, double e, DV &div) {
    if constexpr (TABLE) {
        double f = eos_interp(E, e, E.T_tb, div);
        return 1.e-15 < f ? f : 1.e-15;
    } else {
        return (double)eos_fourth_root(E, (float)(e / EOS_FACTOR));
    }
}

// rhic/src/FluxLimiter.cpp:14-32: sign(0) = +1, minmod = (sx+sy)*min(|x|,|y|)/2
__device__ __forceinline__ int isign(double x) { return x < 0 ? -1 : 1; }
__device__ __forceinline__ double minmod(double x, double y) {
    return (isign(x) + isign(y)) * __builtin_fmin(__builtin_fabs(x), __builtin_fabs(y)) / 2.0;
}
__device__ __forceinline__ double approx_deriv(double x, double y, double z) {
    double l = THETA * (y - x);
    double c = (z - x) / 2.0;
    double r = THETA * (z - y);
    return minmod(l, minmod(c, r));
}

// Newton-Raphson + bisection root solve of the table EoS,
// rhic/src/EnergyMomentumTensor.cpp:26-111, with the division operator
// throughout.  Callers instantiate it inside loops so each kernel carries
// few copies.
struct RootResult {
    double e;
    int nan;
};
// One Newton step of energyDensityFromConservedVariables
// (rhic/src/EnergyMomentumTensor.cpp:26-47) from e0; bad is set when the
// step is NaN (the reference's "f/fp is nan").
template <class QD, class DV>
__device__ __forceinline__ double nr_step(const EosDev &E, double e0, double M0, double M, double Pi, DV &dx,
                                          bool &bad) {
    if constexpr (std::is_same_v<QD, QuotFast>) {
        // mixed form: the same iteration with two reciprocals in place of
        // five quotients.  R = 1/(e0 p) gives 1/e0 and 1/p; with
        // X = c (B + D H - c cst2 D M0 / e0) and cst2 e0 H = p H the step
        // f / fp = f (1 - X / (p H))^-1 is f p H / (p H - X).
        const double p = eos_pressure_fast(E, e0);
        const double cs2 = eos_cs2<true>(E, e0, dx);
        const double R = QuotFast::rcp1(e0 * p);
        const double ie0 = p * R, ip = e0 * R;
        const double cst2 = p * ie0;
        const double A = __builtin_fma(-M0, cst2, M0) + Pi;
        const double B = M0 * (M0 + Pi) - M;
        const double H = (double)QuotFast::sqrt_island(f32_fabs((float)__builtin_fma(A, A, 4 * cst2 * B)));
        const double D = (A - H) * (0.5 * e0 * ip);
        const double f = e0 + D;
        const double c = cs2 - cst2;
        const double X = c * (B + D * H - c * cst2 * D * M0 * ie0);
        const double pH = p * H;
        const double step = f * pH * QuotFast::rcp1(pH - X);
        bad = __builtin_isnan(step);
        return e0 - step;
    }
    double p = eos_pressure<true>(E, e0, dx);
    double cs2 = eos_cs2<true>(E, e0, dx);
    double cst2 = QD::q(p, e0);
    double A = M0 * (1.0 - cst2) + Pi;
    double B = M0 * (M0 + Pi) - M;
    double H = (double)QD::sqrt_island(f32_fabs((float)(A * A + 4 * cst2 * B)));
    double D = QD::q(A - H, 2 * cst2);
    double f = e0 + D;
    double fp = 1.0 - QD::q((cs2 - cst2) * (B + D * H - QD::q((cs2 - cst2) * cst2 * D * M0, e0)), cst2 * e0 * H);
    const double step = QD::q(f, fp);
    bad = __builtin_isnan(step);
    return e0 - step;
}

// the reference's stop test, evaluated in float (:44); the mixed form
// evaluates it in double
template <class QD>
__device__ __forceinline__ bool nr_converged(double e, double e0) {
    if constexpr (std::is_same_v<QD, QuotFast>) return __builtin_fabs(e - e0) <= ACC * __builtin_fabs(e);
    return (double)f32_fabs((float)(e - e0)) <= ACC * (double)f32_fabs((float)e);
}

// bisection fallback after MAX_ITERS_NR Newton steps (:48-111)
template <class QD, class DV>
__device__ __forceinline__ double table_bisect(const EosDev &E, double ePrev, double M0, double M, double Pi) {
    DV dx;
    double e_lo = 0.0, e_mid = ePrev, e_hi = 2.0 * (ePrev + 1.0);
    for (int j = 0; j < MAX_ITERS_BI; ++j) {
        double p_lo = eos_pressure<true>(E, e_lo, dx), p_mid = eos_pressure<true>(E, e_mid, dx),
               p_hi = eos_pressure<true>(E, e_hi, dx);
        double cst2_lo = QD::q(p_lo, e_lo), cst2_mid = QD::q(p_mid, e_mid), cst2_hi = QD::q(p_hi, e_hi);
        double A_lo = M0 * (1.0 - cst2_lo) + Pi;
        double A_mid = M0 * (1.0 - cst2_mid) + Pi;
        double A_hi = M0 * (1.0 - cst2_hi) + Pi;
        double B = M0 * (M0 + Pi) - M;
        double H_lo = (double)QD::sqrt_island(f32_fabs((float)(A_lo * A_lo + 4 * cst2_lo * B)));
        double H_mid = (double)QD::sqrt_island(f32_fabs((float)(A_mid * A_mid + 4 * cst2_mid * B)));
        double H_hi = (double)QD::sqrt_island(f32_fabs((float)(A_hi * A_hi + 4 * cst2_hi * B)));
        double D_lo = QD::q(A_lo - H_lo, 2 * cst2_lo);
        double D_mid = QD::q(A_mid - H_mid, 2 * cst2_mid);
        double D_hi = QD::q(A_hi - H_hi, 2 * cst2_hi);
        double f_lo = e_lo + D_lo, f_mid = e_mid + D_mid, f_hi = e_hi + D_hi;
        if ((e_hi - e_lo) < (ACC * e_hi) / 2.0) return e_mid;
        if (f_lo * f_mid < 0.0) {
            e_hi = e_mid;
            e_mid = (e_lo + e_hi) / 2.0;
        } else if (f_mid * f_hi < 0.0) {
            e_lo = e_mid;
            e_mid = (e_lo + e_hi) / 2.0;
        } else {
            // no sign change (or NaN): the bracket never moves again, so the
            // reference's remaining iterations all repeat this one and its
            // loop ends with ePrev -- return that now
            return ePrev;
        }
    }
    return ePrev;
}

// Table-EoS root solve (rhic/src/EnergyMomentumTensor.cpp:22-111): Newton
// from ePrev, then bisection.
template <class QD = QuotExact, class DV = DivExact>
__device__ __forceinline__ RootResult table_root(const EosDev &E, double ePrev, double M0, double M, double Pi) {
    DV dx;
    double e0 = ePrev;
    for (int j = 0; j < MAX_ITERS_NR; ++j) {
        bool bad;
        const double e = nr_step<QD, DV>(E, e0, M0, M, Pi, dx, bad);
        if (bad) return RootResult{ePrev, NAN_ROOT};
        if (nr_converged<QD>(e, e0)) return RootResult{e, NAN_NONE};
        e0 = e;
    }
    return RootResult{table_bisect<QD, DV>(E, ePrev, M0, M, Pi), NAN_NONE};
}

// The segment of E.p4 holding e (clamped to the table's ends, where its
// line extends) and e's fraction within it.
__device__ __forceinline__ float4 eos_segment_f32(const EosDev &E, float e, float emin, float rsp, float &frac) {
    const float x = (e - emin) * rsp;
    int idx = (int)x;
    idx = idx < EOS_LEN - 2 ? idx : EOS_LEN - 2;
    idx = idx > 0 ? idx : 0;
    frac = x - (float)idx;
    return E.p4[idx];
}

// The mixed form's Newton iteration in FP32: the reference's iteration
// (EnergyMomentumTensor.cpp:26-47, the formulas of nr_step's QuotFast form)
// in single precision from the FP64 moments, with B = M0 (M0 + Pi) - M
// formed in FP64 (it cancels for fast flow), hardware reciprocals and square
// root, and the reference's stop test |e - e0| <= 1e-3 |e| evaluated in FP32,
// as the reference evaluates it (fabsf).  One FP32 segment per step gives p
// and, unless CONST_CS2 (the table's c_s^2 is 1/3 for every e), c_s^2(e0) as
// eos_cs2's one-segment quotient: the pressure one table spacing up at the
// same fraction, clamped alike.  The bisection fallback is the FP64 one.
template <bool CONST_CS2>
__device__ __forceinline__ RootResult table_root_f32(const EosDev &E, double ePrev, double M0d, double Md, double Pid) {
    const float M0 = (float)M0d, Pi = (float)Pid;
    const float B = (float)(M0d * (M0d + Pid) - Md);
    const float emin = (float)E.e_min, rsp = (float)E.r_spacing;
    float e0 = (float)ePrev;
    int how = 0;  // 1: converged, 2: NaN step, 0: iterations spent
    for (int j = 0; j < MAX_ITERS_NR; ++j) {
        float frac;
        const float4 T = eos_segment_f32(E, e0, emin, rsp, frac);
        float p = __builtin_fmaf(frac, T.y, T.x);
        p = 1.e-15f < p ? p : 1.e-15f;
        float cs2 = 1.0f / 3.0f;
        if constexpr (!CONST_CS2) {
            const float dpde = __builtin_fmaf(frac, T.w, T.z);
            float v = E.cs2_mode ? dpde : __builtin_amdgcn_rcpf(dpde);
            v = (1.0f / 3.0f) < v ? (1.0f / 3.0f) : v;
            v = 0.01f < v ? v : 0.01f;
            cs2 = v;
        }
        const float R = __builtin_amdgcn_rcpf(e0 * p);
        const float ie0 = p * R, ip = e0 * R;
        const float cst2 = p * ie0;
        const float A = __builtin_fmaf(-M0, cst2, M0) + Pi;
        const float H = __builtin_amdgcn_sqrtf(__builtin_fabsf(__builtin_fmaf(A, A, 4 * cst2 * B)));
        const float D = (A - H) * (0.5f * e0 * ip);
        const float f = e0 + D;
        const float c = cs2 - cst2;
        const float X = c * (B + D * H - c * cst2 * D * M0 * ie0);
        const float pH = p * H;
        const float step = f * pH * __builtin_amdgcn_rcpf(pH - X);
        if (__builtin_isnan(step)) {
            how = 2;
            break;
        }
        const float e = e0 - step;
        const bool done = __builtin_fabsf(e - e0) <= 1.0e-3f * __builtin_fabsf(e);
        e0 = e;
        if (done) {
            how = 1;
            break;
        }
    }
    if (how == 1) return RootResult{(double)e0, NAN_NONE};
    if (how == 2) return RootResult{ePrev, NAN_ROOT};
    return RootResult{table_bisect<QuotFast, DivRecip>(E, ePrev, M0d, Md, Pid), NAN_NONE};
}

// energyDensityFromConservedVariables, rhic/src/EnergyMomentumTensor.cpp:22-115
template <bool TABLE, class QD = QuotExact, class DV = DivExact>
__device__ __forceinline__ double energy_from_conserved(const EosDev &E, double ePrev, double M0, double M, double Pi,
                                                        int &nan) {
    if constexpr (!TABLE) {
        (void)E; (void)ePrev; (void)Pi; (void)nan;
        return (double)f32_fabs((float)((double)QD::sqrt_island(f32_fabs((float)(4 * M0 * M0 - 3 * M))) - M0));
    } else {
        const RootResult r = table_root<QD, DV>(E, ePrev, M0, M, Pi);
        if (r.nan) nan = r.nan;
        return r.e;
    }
}

struct Prim {
    double e, p, ut, ux, uy, un;
};

// getInferredVariables, rhic/src/EnergyMomentumTensor.cpp:117-233, in two
// parts: the energy density from the conserved variables (the closed form
// below ePrev = 0.1, else the root solve), then pressure, floors and
// velocity.  ttt..pitn are q[0..7]; Pi = q[14] when bulk is evolved.
struct Moments {
    double M0, M1, M2, M3, M;
};

__device__ __forceinline__ Moments moments(double t, double ttt, double ttx, double tty, double ttn, double pitt,
                                           double pitx, double pity, double pitn) {
    Moments m;
    m.M0 = ttt - pitt;
    m.M1 = ttx - pitx;
    m.M2 = tty - pity;
    m.M3 = ttn - pitn;
    m.M = m.M1 * m.M1 + m.M2 * m.M2 + t * t * m.M3 * m.M3;
    return m;
}

template <bool TABLE, class DV, class QD>
__device__ __forceinline__ Prim recover_finish(const EosDev &E, const Moments &m, double e, double Pi, int &nan,
                                               DV &div) {
    Prim r;
    r.e = e;
    if (__builtin_isnan(r.e)) nan = NAN_ENERGY;
    constexpr bool FAST = std::is_same_v<QD, QuotFast>;
    if constexpr (FAST && TABLE) r.p = eos_pressure_fast(E, r.e);
    else r.p = eos_pressure<TABLE>(E, r.e, div);
    if (r.e < 1.e-7) {
        r.e = 1.e-7;
        r.p = 1.e-7;
    }
    double P = r.p + Pi;
    double Ei;
    if constexpr (FAST) Ei = QuotFast::rcp1(r.e + P);
    else Ei = QD::q(1.0, r.e + P);
    r.ut = (double)QD::sqrt_island(f32_fabs((float)((m.M0 + P) * Ei)));
    double E2;
    if constexpr (FAST) E2 = Ei * QuotFast::rcp1(r.ut);
    else E2 = QD::q(Ei, r.ut);
    r.ux = m.M1 * E2;
    r.uy = m.M2 * E2;
    r.un = m.M3 * E2;
    if (r.ut > 1.0e3) {
        r.ut = 1.0;
        r.ux = 0.0;
        r.uy = 0.0;
        r.un = 0.0;
    }
    return r;
}

// getInferredVariables from the moments of the conserved variables.
template <bool TABLE, class DV, class QD = QuotExact>
__device__ __forceinline__ Prim recover_moments(const EosDev &E, const Moments &m, double Pi, double ePrev, int &nan,
                                                DV &div) {
    double e;
    if (ePrev <= 0.1) {
        if constexpr (std::is_same_v<QD, QuotFast>) e = m.M0 - m.M * QuotFast::rcp1(m.M0);
        else e = m.M0 - QD::q(m.M, m.M0);
    } else  // the FP64 path keeps the division operator inside the root solve's table lookups
        e = energy_from_conserved<TABLE, QD, std::conditional_t<std::is_same_v<QD, QuotFast>, DivRecip, DivExact>>(
            E, ePrev, m.M0, m.M, Pi, nan);
    return recover_finish<TABLE, DV, QD>(E, m, e, Pi, nan, div);
}

template <bool TABLE, class DV, class QD = QuotExact>
__device__ __forceinline__ Prim recover(const EosDev &E, double t, double ttt, double ttx, double tty, double ttn,
                                        double pitt, double pitx, double pity, double pitn, double Pi, double ePrev,
                                        int &nan, DV &div) {
    const Moments m = moments(t, ttt, ttx, tty, ttn, pitt, pitx, pity, pitn);
    return recover_moments<TABLE, DV, QD>(E, m, Pi, ePrev, nan, div);
}

// Four root solves from one start value, stepped together: the table-EoS
// Newton iteration and bisection of table_root for each state, unchanged,
// with the four iterations interleaved so that their table lookups and
// divisions overlap (independent chains).  Each state takes exactly the
// steps table_root takes for it; rn[k] is its NaN site.
template <class DV>
__device__ __forceinline__ void table_root4(const EosDev &E, double ePrev, const Moments (&m)[4], const double (&Pi)[4],
                                            double (&e)[4], int (&rn)[4]) {
    DV dx;
    double e0[4];
    bool done[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        e0[k] = ePrev;
        done[k] = false;
        rn[k] = NAN_NONE;
        e[k] = ePrev;
    }
    for (int j = 0; j < MAX_ITERS_NR; ++j) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!done[k]) {
                bool bad;
                const double en = nr_step<QuotExact, DV>(E, e0[k], m[k].M0, m[k].M, Pi[k], dx, bad);
                if (bad) {
                    e[k] = ePrev;
                    rn[k] = NAN_ROOT;
                    done[k] = true;
                } else if (nr_converged<QuotExact>(en, e0[k])) {
                    e[k] = en;
                    done[k] = true;
                } else {
                    e0[k] = en;
                }
            }
            any = any || !done[k];
        }
        if (!any) break;
    }
#pragma unroll 1
    for (int k = 0; k < 4; ++k)
        if (!done[k]) e[k] = table_bisect<QuotExact, DV>(E, ePrev, m[k].M0, m[k].M, Pi[k]);
}

// getInferredVariables for four states sharing ePrev (the four half-site
// states of one cell's two faces along one direction): recover() of each,
// with the four root solves interleaved.  NaN sites are recorded in the
// order the four recover() calls would record them.
template <bool TABLE, class DV>
__device__ __forceinline__ void recover4(const EosDev &E, double t, const double (&x)[4][9], double ePrev, int &nan,
                                         DV &div, Prim (&out)[4]) {
    Moments m[4];
    double Pi[4], e[4];
    int rn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        m[k] = moments(t, x[k][0], x[k][1], x[k][2], x[k][3], x[k][4], x[k][5], x[k][6], x[k][7]);
        Pi[k] = x[k][8];
        rn[k] = NAN_NONE;
    }
    if (ePrev <= 0.1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) e[k] = m[k].M0 - QuotExact::q(m[k].M, m[k].M0);
    } else if constexpr (TABLE) {
        table_root4<DivExact>(E, ePrev, m, Pi, e, rn);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int dummy = NAN_NONE;
            e[k] = energy_from_conserved<false>(E, ePrev, m[k].M0, m[k].M, Pi[k], dummy);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (rn[k]) nan = rn[k];
        out[k] = recover_finish<TABLE, DV, QuotExact>(E, m[k], e[k], Pi[k], nan, div);
    }
}

}  // namespace vh
