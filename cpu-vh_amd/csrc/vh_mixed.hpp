// Mixed-precision stage kernel of the viscous-hydro step (gfx950).
//
// Same equations as the FP64 path (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:866-924
// and the files it calls), re-associated for speed:
//   * storage is FP32: q (high parts), u, previous u, e, p; the energy-momentum
//     components T^{tau mu} also keep an FP32 low part, so their stored value
//     is the FP64 accumulation to ~2^-48 (compensated storage);
//   * limited slopes, half-site values, face fluxes and the source terms run
//     in FP32 with fused multiply-adds and reciprocal multiplies;
//   * the cell's conserved-to-primitive root solve, the EoS table lookups and
//     the accumulation acc = q + dt*S + r_x + r_y + r_eta (and the stage-2
//     convex combination) run in FP64, the face states' Newton iterations in
//     FP32 (from FP64 moments); the regulator's factors are FP32;
//   * each face flux is evaluated once and shared by the two cells it
//     separates (the reference evaluates it from both sides, with each
//     cell's own e as the root solve's starting value; here each side of a
//     face starts from the e of the cell on that side).
// A block of 256 threads owns a TX x TY x TZ tile of one event.  All its
// threads first evaluate the tile's faces (the x, y and eta faces in one
// list, (TX+1)*TY*TZ + ... of them) into LDS, then each thread finishes one
// cell: source terms, the combination with its faces' fluxes, the root
// solve, the regulator and the stores.  Every wave does both kinds of work,
// so no wave idles at the barrier and no single wave carries the
// per-cell tail.
#pragma once
#include "vh_stage.hpp"

namespace vh {
namespace mx {

constexpr float delta_pipi = 1.33333f;  // rhic/src/SourceTerms.cpp:50-53
constexpr float tau_pipi = 1.42857f;
constexpr float delta_PiPi = 0.666667f;
constexpr float lambda_piPi = 1.2f;
constexpr float THETA_F = 1.1f;         // rhic/include/FluxLimiter.h:13

// Device pointers of one stage (pooled state of a batch).  The fields of a
// group lie cap cells apart in the pool, so a group is one base pointer and
// field n of it is at(base, n): a few kernel arguments instead of one pointer
// per field, which kept scalar registers from spilling.
struct MxFields {
    const float *q;    // stage input, high parts (15 fields)
    const float *ql;   // stage input, low parts of T^{tau mu} (4)
    const float *qn;   // q^n (stage 2)
    const float *qnl;
    float *qo;
    float *qol;
    const float *u;    // stage input velocity (4)
    const float *up;   // velocity of the time derivative
    const float *e, *p;  // stage input e, p
    float *uo;
    float *eo, *po;
    unsigned long cap;  // cells between consecutive fields of a group
    int stage;            // 0: first stage, 1: second stage
    int off;              // step offset inside a captured graph
    int face_mode;        // see faces_mx
    int *err;
    // The empty asm makes the stride opaque at each use, so field addresses
    // are formed (scalar instructions) where they are used; hoisted, all of
    // them at once overflow the scalar registers into spills.
    template <class T>
    __device__ __forceinline__ T *at(T *base, int n) const {
        unsigned long c = cap;
        asm volatile("" : "+s"(c));
        return base + c * (unsigned long)n;
    }
};

// rhic/src/FluxLimiter.cpp:14-32 in FP32: minmod(a, minmod(b, c)) of
// a = theta (y - x), b = (z - x) / 2, c = theta (z - y) is min(|a|,|b|,|c|)
// with the common sign when all three share one, else zero (up to the sign
// of a zero result).  a and c share a sign exactly when y - x and z - y do,
// and then z - x has it too (a float difference of distinct values is never
// zero with subnormals kept), so the test is a c > 0 (products below the
// smallest subnormal, differences under 1e-22, give zero).
// S scales the result (minmod is positively homogeneous): the half-site step
// takes half the slope with the halves folded into the three candidates.
template <int S2 = 0>
__device__ __forceinline__ float deriv_f(float x, float y, float z) {
    constexpr float s = S2 ? 0.5f : 1.0f;
    const float a = (s * THETA_F) * (y - x), b = (z - x) * (s * 0.5f), c = (s * THETA_F) * (z - y);
    const float m = __builtin_fminf(__builtin_fabsf(a), __builtin_fminf(__builtin_fabsf(b), __builtin_fabsf(c)));
    return a * c > 0.f ? __builtin_copysignf(m, a) : 0.f;
}

__device__ __forceinline__ int clampi(int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); }

// FP32 reciprocal of the hardware (1 ulp), in place of a division sequence
// that scales operands for subnormal results
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Field access by a 32-bit cell index: a scalar base plus a 32-bit byte
// offset, which the hardware adds itself (global loads with a scalar base),
// instead of per-lane 64-bit address arithmetic.  The pool holds fewer than
// 2^30 cells (checked when a mixed batch is created).
__device__ __forceinline__ float ldf(const float *__restrict__ base, unsigned idx) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (idx << 2));
}
__device__ __forceinline__ void stf(float *__restrict__ base, unsigned idx, float v) {
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (idx << 2)) = v;
}

// Boost-invariant 2+1D events (nz = 1 with u^eta, pi^{tau eta}, pi^{x eta}
// and pi^{y eta} zero initially): T^{tau eta} (q[3]), pi^{tau eta} (q[7]),
// pi^{x eta} (q[10]) and pi^{y eta} (q[12]) are zero and stay exactly zero
// under the update (every term of their fluxes and sources carries one of
// them, u^eta or an eta derivative, all zero), so the BI kernels neither
// load, evolve nor store them.
constexpr bool dead_field(int n) { return n == 3 || n == 7 || n == 10 || n == 12; }
template <bool BI>
constexpr bool field_live(int n) { return !BI || !dead_field(n); }
// LDS slot of a live field
template <bool BI>
constexpr int field_slot(int n) { return BI ? n - (n > 3) - (n > 7) - (n > 10) - (n > 12) : n; }

template <bool TABLE>
__device__ __forceinline__ float temperature_f(const EosDev &E, float e) {
    if constexpr (TABLE) {
        DivRecip dx;
        return (float)eos_temperature<true>(E, (double)e, dx);
    } else {
        return __builtin_sqrtf(__builtin_sqrtf(e * (float)(1.0 / EOS_FACTOR)));
    }
}

// zeta_over_s in FP32 (the hardware's exponential)
__device__ __forceinline__ float zeta_over_s_f(const EvDesc &G, float T) {
    const bool ext = G.zeta_Tpeak_GeV > 0.0;
    const float x = ext ? T * (float)(HBARC / G.zeta_Tpeak_GeV) : T * (1.0f / 1.01355f);
    float z;
    if (x > 1.05f) z = 0.9f * __expf(-(x - 1) * 40.0f) + 0.25f * __expf(-(x - 1) * (1.0f / 0.13f)) + 0.001f;
    else if (x < 0.995f) z = 0.9f * __expf((x - 1) * 400.0f) + 0.22f * __expf((x - 1) * (1.0f / 0.022f)) + 0.03f;
    else z = -13.77f * x * x + 27.55f * x - 13.45f;
    return ext ? (float)(G.zeta_peak / 0.33) * z : z;
}

// tau_Pi^-1 = 15 a^2 T / (zeta/s), held at 1/dt at most with beta_Pi scaled
// alike (the Navier-Stokes value -zeta theta is kept; explicit RK2 is unstable
// for dt/tau_Pi > 2)
__device__ __forceinline__ void bulk_relaxation(const EvDesc &G, float a2, float T, float rdt, float &tauPiInv,
                                                float &beta_Pi) {
    tauPiInv = 15 * a2 * T * frcp(zeta_over_s_f(G, T));
    if (tauPiInv > rdt) {
        beta_Pi *= rdt * frcp(tauPiInv);
        tauPiInv = rdt;
    }
}

// Source terms of one cell (loadSourceTerms2 + setPimunuSourceTerms,
// rhic/src/SourceTerms.cpp:55-313,421-612) in FP32, and the pi-gradient
// terms of loadSourceTermsX/Y/Z (:316-419) for each direction: Sd[d][mu].
template <int NC, bool TABLE, bool DIM3>
__device__ __forceinline__ void source_mx(const MxFields &F, const EvDesc &G, const EosDev &E, unsigned s, int i, int j,
                                          int k, float t, float ecell, const float (&q)[NC], float (&S)[NC],
                                          float (&Sd)[3][4]) {
#pragma clang fp contract(fast)
    const float rdx = frcp((float)G.dx), rdy = frcp((float)G.dy), rdz = frcp((float)G.dz);
    const float rt = frcp(t), rdt = frcp((float)G.dt);
    const float r3 = 1.0f / 3.0f, r5 = 1.0f / 5.0f;
    const float ttt = q[0], ttx = q[1], tty = q[2], ttn = q[3];
    const float pitt = q[4], pitx = q[5], pity = q[6], pitn = q[7], pixx = q[8], pixy = q[9], pixn = q[10],
                piyy = q[11], piyn = q[12], pinn = q[13];
    const float Pi = NC == 15 ? q[NC - 1] : 0.0f;
    const float *__restrict__ utv = F.at(F.u, 0);
    const float *__restrict__ uxv = F.at(F.u, 1);
    const float *__restrict__ uyv = F.at(F.u, 2);
    const float *__restrict__ unv = F.at(F.u, 3);
    const float *__restrict__ pv = F.p;
    const float p = ldf(pv, s);
    const float ut = ldf(utv, s), ux = ldf(uxv, s), uy = ldf(uyv, s), un = ldf(unv, s);
    const float utp = ldf(F.at(F.up, 0), s), uxp = ldf(F.at(F.up, 1), s), uyp = ldf(F.at(F.up, 2), s), unp = ldf(F.at(F.up, 3), s);
    const unsigned strY = G.nx, strZ = (unsigned)G.nx * G.ny;
    const unsigned sl = i > 0 ? s - 1 : s, sr = i < G.nx - 1 ? s + 1 : s;
    const unsigned sb = j > 0 ? s - strY : s, stp = j < G.ny - 1 ? s + strY : s;
    const unsigned sk = k > 0 ? s - strZ : s, sf = k < G.nz - 1 ? s + strZ : s;
    const float dxut = deriv_f(ldf(utv, sl), ut, ldf(utv, sr)) * rdx;
    const float dxux = deriv_f(ldf(uxv, sl), ux, ldf(uxv, sr)) * rdx;
    const float dxuy = deriv_f(ldf(uyv, sl), uy, ldf(uyv, sr)) * rdx;
    const float dxun = deriv_f(ldf(unv, sl), un, ldf(unv, sr)) * rdx;
    const float dyut = deriv_f(ldf(utv, sb), ut, ldf(utv, stp)) * rdy;
    const float dyux = deriv_f(ldf(uxv, sb), ux, ldf(uxv, stp)) * rdy;
    const float dyuy = deriv_f(ldf(uyv, sb), uy, ldf(uyv, stp)) * rdy;
    const float dyun = deriv_f(ldf(unv, sb), un, ldf(unv, stp)) * rdy;
    float dnut = 0.f, dnux = 0.f, dnuy = 0.f, dnun = 0.f, dnp = 0.f;
    if constexpr (DIM3) {
        dnut = deriv_f(ldf(utv, sk), ut, ldf(utv, sf)) * rdz;
        dnux = deriv_f(ldf(uxv, sk), ux, ldf(uxv, sf)) * rdz;
        dnuy = deriv_f(ldf(uyv, sk), uy, ldf(uyv, sf)) * rdz;
        dnun = deriv_f(ldf(unv, sk), un, ldf(unv, sf)) * rdz;
        dnp = deriv_f(ldf(pv, sk), p, ldf(pv, sf)) * rdz;
    }
    const float dxp = deriv_f(ldf(pv, sl), p, ldf(pv, sr)) * rdx;
    const float dyp = deriv_f(ldf(pv, sb), p, ldf(pv, stp)) * rdy;

    const float pP = p + Pi;
    const float tnn = (ecell + pP) * un * un + pP * rt * rt + pinn;
    const float rut = frcp(ut);
    const float vx = ux * rut, vy = uy * rut, vn = un * rut;
    const float dxvx = (dxux - vx * dxut) * rut;
    const float dyvy = (dyuy - vy * dyut) * rut;
    const float dnvn = (dnun - vn * dnut) * rut;
    const float dkvk = dxvx + dyvy + dnvn;
    S[0] = -(ttt * rt + t * tnn) + dkvk * (pitt - p - Pi) - vx * dxp - vy * dyp - vn * dnp;
    S[1] = -ttx * rt - dxp + dkvk * pitx;
    S[2] = -tty * rt - dyp + dkvk * pity;
    S[3] = -(3 * ttn) * rt - dnp * (rt * rt) + dkvk * pitn;

    // pi-gradient terms of each direction: d_a(pi^{tau mu}) v_a - d_a(pi^{a mu})
    {
        const unsigned nbm[3] = {sl, sb, sk}, nbp[3] = {sr, stp, sf};
        const float fac[3] = {0.5f * rdx, 0.5f * rdy, 0.5f * rdz};
        const float vd[3] = {vx, vy, vn};
#pragma unroll
        for (int d = 0; d < (DIM3 ? 3 : 2); ++d) {
            auto dq = [&](int n) { return (ldf(F.at(F.q, n), nbp[d]) - ldf(F.at(F.q, n), nbm[d])) * fac[d]; };
#pragma unroll
            for (int mu = 0; mu < 4; ++mu) Sd[d][mu] = dq(4 + mu) * vd[d] - dq(pi_a_mu(d, mu));
            if constexpr (NC == 15) {
                const float dPi = dq(14);
                Sd[d][0] -= vd[d] * dPi;
                Sd[d][d + 1] -= d == 2 ? dPi * (rt * rt) : dPi;
            }
        }
        if constexpr (!DIM3) Sd[2][0] = Sd[2][1] = Sd[2][2] = Sd[2][3] = 0.f;
    }

    const float e = ecell;
    const float T = temperature_f<TABLE>(E, e);
    const float d_etabar = (float)etabar_of_T(G, (double)T);
    const float taupiInv = T * r5 * frcp(d_etabar);
    const float beta_pi = (e + p) * r5;
    DivRecip dx64;
    const float cs2 = (float)eos_cs2<TABLE>(E, (double)e, dx64);
    const float a = 1.0f / 3.0f - cs2;
    const float a2 = a * a;
    float beta_Pi = 15 * a2 * (e + p);
    const float lambda_Pipi = 8 * a * r5;
    float tauPiInv = 0.0f;
    if constexpr (NC == 15) {
        // bulkViscosityToEntropyDensity (SourceTerms.cpp:40-48), or its
        // rescaled form with the event's peak (zeta_over_s); relaxation no
        // faster than 1/dt, as in the FP64 step
        if (a2 != 0.0f) bulk_relaxation(G, a2, T, rdt, tauPiInv, beta_Pi);
    }
    const float ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;
    const float rt2 = rt * rt, rt3 = rt2 * rt, r2t = 0.5f * rt;
    const float dtut = (ut - utp) * rdt;
    const float dtux = (ux - uxp) * rdt;
    const float dtuy = (uy - uyp) * rdt;
    const float dtun = (un - unp) * rdt;
    const float Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
    const float DuxUpper = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const float Dux = -DuxUpper;
    const float DuyUpper = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const float Duy = -DuyUpper;
    const float DunUpper = ut * dtun + ux * dxun + uy * dyun + un * dnun + (2 * ut * un) * rt;
    const float Dun = -t2 * DunUpper;
    const float dut = Dut - t * un * un;
    const float dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const float duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const float dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
    const float theta = (ut) * rt + dtut + dxux + dyuy + dnun;
    const float stt = -t * ut * un2 + (dtut - ut * dut) + ((ut2 - 1) * theta) * r3;
    const float stx = -(t * un2 * ux) / 2.0f + (dtux - dxut) / 2.0f - (ux * dut + ut * dux) / 2.0f + (ut * ux * theta) * r3;
    const float sty = -(t * un2 * uy) / 2.0f + (dtuy - dyut) / 2.0f - (uy * dut + ut * duy) / 2.0f + (ut * uy * theta) * r3;
    const float stn = (-un * (2.0f * ut2 + t2 * un2)) * r2t + (dtun - (dnut) * rt2) / 2.0f -
                       (un * dut + ut * dun) / 2.0f + (ut * un * theta) * r3;
    const float sxx = -(dxux + ux * dux) + ((1.0f + ux * ux) * theta) * r3;
    const float sxy = -(dxuy + dyux) / 2.0f - (uy * dux + ux * duy) / 2.0f + (ux * uy * theta) * r3;
    const float sxn = (-ut * ux * un) * rt - (dxun + (dnux) * rt2) / 2.0f - (un * dux + ux * dun) / 2.0f + (ux * un * theta) * r3;
    const float syy = -(dyuy + uy * duy) + ((1.0f + uy * uy) * theta) * r3;
    const float syn = (-ut * uy * un) * rt - (dyun + (dnuy) * rt2) / 2.0f - (un * duy + uy * dun) / 2.0f + (uy * un * theta) * r3;
    const float snn = (-ut * (1.0f + 2.0f * t2 * un2)) * rt3 - (dnun) * rt2 - un * dun + ((rt2 + un2) * theta) * r3;
    const float wtx = (-dtux - dxut + ut * dux - ux * dut - t * un2 * ux) / 2.0f;
    const float wty = (-dtuy - dyut + ut * duy - uy * dut - t * un2 * uy) / 2.0f;
    const float wtn = (-t2 * dtun - 2.0f * t * un - dnut - ut * Dun - t2 * un * dut - t3 * un * un2) / 2.0f;
    const float wxy = (dxuy - dyux + ux * duy - uy * dux) / 2.0f;
    const float wxn = (t2 * dxun - dnux - ux * Dun - t2 * un * dux) / 2.0f;
    const float wyn = (t2 * dyun - dnuy - uy * Dun - t2 * un * duy) / 2.0f;
    const float wxt = wtx, wyt = wty, wnt = (wtn) * rt2, wyx = -wxy, wnx = (-wxn) * rt2, wny = (-wyn) * rt2;
    const float I1tt = 2.0f * ut * (pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun);
    const float I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy +
                        (pitn * ux + pixn * ut) * Dun;
    const float I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy +
                        (pitn * uy + piyn * ut) * Dun;
    const float I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux + (pity * un + piyn * ut) * Duy +
                        (pitn * un + pinn * ut) * Dun;
    const float I1xx = 2.0f * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
    const float I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy +
                        (pixn * uy + piyn * ux) * Dun;
    const float I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux + (pixy * un + piyn * ux) * Duy +
                        (pixn * un + pinn * ux) * Dun;
    const float I1yy = 2.0f * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
    const float I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux + (piyy * un + piyn * uy) * Duy +
                        (piyn * un + pinn * uy) * Dun;
    const float I1nn = 2.0f * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun);
    const float I2tt = theta * pitt, I2tx = theta * pitx, I2ty = theta * pity, I2tn = theta * pitn,
                 I2xx = theta * pixx, I2xy = theta * pixy, I2xn = theta * pixn, I2yy = theta * piyy,
                 I2yn = theta * piyn, I2nn = theta * pinn;
    const float I3tt = 2.0f * (pitx * wtx + pity * wty + pitn * wtn);
    const float I3tx = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
    const float I3ty = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
    const float I3tn = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
    const float I3xx = 2.0f * (pitx * wxt + pixy * wxy + pixn * wxn);
    const float I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
    const float I3xn = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
    const float I3yy = 2.0f * (pity * wyt + pixy * wyx + piyn * wyn);
    const float I3yn = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
    const float I3nn = 2.0f * (pitn * wnt + pixn * wnx + piyn * wny);
    const float ux2 = ux * ux, uy2 = uy * uy;
    const float ps = pitt * stt - 2.0f * pitx * stx - 2.0f * pity * sty + pixx * sxx + 2.0f * pixy * sxy + piyy * syy -
                      2.0f * pitn * stn * t2 + 2.0f * pixn * sxn * t2 + 2.0f * piyn * syn * t2 + pinn * snn * t2 * t2;
    const float I4tt = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - ((1.0f - ut2) * ps) * r3;
    const float I4tx = (pitt * stx + pitx * stt) / 2.0f - (pitx * sxx + pixx * stx) / 2.0f -
                        (pity * sxy + pixy * sty) / 2.0f - t2 * (pitn * sxn + pixn * stn) / 2.0f + ((ut * ux) * ps) * r3;
    const float I4ty = (pitt * sty + pity * stt) / 2.0f - (pitx * sxy + pixy * stx) / 2.0f -
                        (pity * syy + piyy * sty) / 2.0f - t2 * (pitn * syn + piyn * stn) / 2.0f + ((ut * uy) * ps) * r3;
    const float I4tn = (pitt * stn + pitn * stt) / 2.0f - (pitx * sxn + pixn * stx) / 2.0f -
                        (pity * syn + piyn * sty) / 2.0f - t2 * (pitn * snn + pinn * stn) / 2.0f + ((ut * un) * ps) * r3;
    const float I4xx = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + ((1.0f + ux2) * ps) * r3;
    const float I4xy = (pitx * sty + pity * stx) / 2.0f - (pixx * sxy + pixy * sxx) / 2.0f -
                        (pixy * syy + piyy * sxy) / 2.0f - t2 * (pixn * syn + piyn * sxn) / 2.0f + ((ux * uy) * ps) * r3;
    const float I4xn = (pitx * stn + pitn * stx) / 2.0f - (pixx * sxn + pixn * sxx) / 2.0f -
                        (pixy * syn + piyn * sxy) / 2.0f - t2 * (pixn * snn + pinn * sxn) / 2.0f + ((ux * un) * ps) * r3;
    const float I4yy = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + ((1.0f + uy2) * ps) * r3;
    const float I4yn = (pity * stn + pitn * sty) / 2.0f - (pixy * sxn + pixn * sxy) / 2.0f -
                        (piyy * syn + piyn * syy) / 2.0f - t2 * (piyn * snn + pinn * syn) / 2.0f + ((uy * un) * ps) * r3;
    const float I4nn = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + ((rt2 + un2) * ps) * r3;
    const float Itt = I1tt + delta_pipi * I2tt - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
    const float Itx = I1tx + delta_pipi * I2tx - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
    const float Ity = I1ty + delta_pipi * I2ty - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
    const float Itn = I1tn + delta_pipi * I2tn - I3tn + tau_pipi * I4tn - lambda_piPi * Pi * stn;
    const float Ixx = I1xx + delta_pipi * I2xx - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
    const float Ixy = I1xy + delta_pipi * I2xy - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
    const float Ixn = I1xn + delta_pipi * I2xn - I3xn + tau_pipi * I4xn - lambda_piPi * Pi * sxn;
    const float Iyy = I1yy + delta_pipi * I2yy - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
    const float Iyn = I1yn + delta_pipi * I2yn - I3yn + tau_pipi * I4yn - lambda_piPi * Pi * syn;
    const float Inn = I1nn + delta_pipi * I2nn - I3nn + tau_pipi * I4nn - lambda_piPi * Pi * snn;
    S[4] = (2.0f * beta_pi * stt - pitt * taupiInv - Itt - 2.0f * un * t * pitn) * rut + pitt * dkvk;
    S[5] = (2.0f * beta_pi * stx - pitx * taupiInv - Itx - un * t * pixn) * rut + pitx * dkvk;
    S[6] = (2.0f * beta_pi * sty - pity * taupiInv - Ity - un * t * piyn) * rut + pity * dkvk;
    S[7] = (2.0f * beta_pi * stn - pitn * taupiInv - Itn - un * t * pinn - (ut * pitn + un * pitt) * rt) * rut +
           pitn * dkvk;
    S[8] = (2.0f * beta_pi * sxx - pixx * taupiInv - Ixx) * rut + pixx * dkvk;
    S[9] = (2.0f * beta_pi * sxy - pixy * taupiInv - Ixy) * rut + pixy * dkvk;
    S[10] = (2.0f * beta_pi * sxn - pixn * taupiInv - Ixn - (ut * pixn + un * pitx) * rt) * rut + pixn * dkvk;
    S[11] = (2.0f * beta_pi * syy - piyy * taupiInv - Iyy) * rut + piyy * dkvk;
    S[12] = (2.0f * beta_pi * syn - piyn * taupiInv - Iyn - (ut * piyn + un * pity) * rt) * rut + piyn * dkvk;
    S[13] = (2.0f * beta_pi * snn - pinn * taupiInv - Inn - (2 * (ut * pinn + un * pitn)) * rt) * rut + pinn * dkvk;
    if constexpr (NC == 15) {
        const float dPi = -beta_Pi * theta - Pi * tauPiInv - delta_PiPi * Pi * theta + lambda_Pipi * ps;
        S[14] = (dPi) * rut + Pi * dkvk;
    } else {
        (void)beta_Pi; (void)lambda_Pipi; (void)tauPiInv;
    }

}

// source_mx for a boost-invariant 2+1D cell: the same expressions with
// u^eta, pi^{tau eta}, pi^{x eta}, pi^{y eta} and every eta derivative zero
// (S[3], S[7], S[10], S[12] and Sd[.][3] are zero and not formed).
template <int NC, bool TABLE>
__device__ __forceinline__ void source_mx_bi(const MxFields &F, const EvDesc &G, const EosDev &E, unsigned s, int i,
                                             int j, float t, float ecell, const float (&q)[NC], float (&S)[NC],
                                             float (&Sd)[3][4]) {
#pragma clang fp contract(fast)
    const float rdx = frcp((float)G.dx), rdy = frcp((float)G.dy);
    const float rt = frcp(t), rdt = frcp((float)G.dt);
    const float r3 = 1.0f / 3.0f, r5 = 1.0f / 5.0f;
    const float ttt = q[0], ttx = q[1], tty = q[2];
    const float pitt = q[4], pitx = q[5], pity = q[6], pixx = q[8], pixy = q[9], piyy = q[11], pinn = q[13];
    const float Pi = NC == 15 ? q[NC - 1] : 0.0f;
    const float *__restrict__ utv = F.at(F.u, 0);
    const float *__restrict__ uxv = F.at(F.u, 1);
    const float *__restrict__ uyv = F.at(F.u, 2);
    const float *__restrict__ pv = F.p;
    const float p = ldf(pv, s);
    const float ut = ldf(utv, s), ux = ldf(uxv, s), uy = ldf(uyv, s);
    const float utp = ldf(F.at(F.up, 0), s), uxp = ldf(F.at(F.up, 1), s), uyp = ldf(F.at(F.up, 2), s);
    const unsigned strY = G.nx;
    const unsigned sl = i > 0 ? s - 1 : s, sr = i < G.nx - 1 ? s + 1 : s;
    const unsigned sb = j > 0 ? s - strY : s, stp = j < G.ny - 1 ? s + strY : s;
    const float dxut = deriv_f(ldf(utv, sl), ut, ldf(utv, sr)) * rdx;
    const float dxux = deriv_f(ldf(uxv, sl), ux, ldf(uxv, sr)) * rdx;
    const float dxuy = deriv_f(ldf(uyv, sl), uy, ldf(uyv, sr)) * rdx;
    const float dyut = deriv_f(ldf(utv, sb), ut, ldf(utv, stp)) * rdy;
    const float dyux = deriv_f(ldf(uxv, sb), ux, ldf(uxv, stp)) * rdy;
    const float dyuy = deriv_f(ldf(uyv, sb), uy, ldf(uyv, stp)) * rdy;
    const float dxp = deriv_f(ldf(pv, sl), p, ldf(pv, sr)) * rdx;
    const float dyp = deriv_f(ldf(pv, sb), p, ldf(pv, stp)) * rdy;

    const float pP = p + Pi;
    const float tnn = pP * rt * rt + pinn;
    const float rut = frcp(ut);
    const float vx = ux * rut, vy = uy * rut;
    const float dxvx = (dxux - vx * dxut) * rut;
    const float dyvy = (dyuy - vy * dyut) * rut;
    const float dkvk = dxvx + dyvy;
    S[0] = -(ttt * rt + t * tnn) + dkvk * (pitt - p - Pi) - vx * dxp - vy * dyp;
    S[1] = -ttx * rt - dxp + dkvk * pitx;
    S[2] = -tty * rt - dyp + dkvk * pity;
    S[3] = 0.0f;

    {
        const unsigned nbm[2] = {sl, sb}, nbp[2] = {sr, stp};
        const float fac[2] = {0.5f * rdx, 0.5f * rdy};
        const float vd[2] = {vx, vy};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            auto dq = [&](int n) { return (ldf(F.at(F.q, n), nbp[d]) - ldf(F.at(F.q, n), nbm[d])) * fac[d]; };
#pragma unroll
            for (int mu = 0; mu < 3; ++mu) Sd[d][mu] = dq(4 + mu) * vd[d] - dq(pi_a_mu(d, mu));
            Sd[d][3] = 0.0f;
            if constexpr (NC == 15) {
                const float dPi = dq(14);
                Sd[d][0] -= vd[d] * dPi;
                Sd[d][d + 1] -= dPi;
            }
        }
        Sd[2][0] = Sd[2][1] = Sd[2][2] = Sd[2][3] = 0.f;
    }

    const float e = ecell;
    const float T = temperature_f<TABLE>(E, e);
    const float d_etabar = (float)etabar_of_T(G, (double)T);
    const float taupiInv = T * r5 * frcp(d_etabar);
    const float beta_pi = (e + p) * r5;
    DivRecip dx64;
    const float cs2 = (float)eos_cs2<TABLE>(E, (double)e, dx64);
    const float a = 1.0f / 3.0f - cs2;
    const float a2 = a * a;
    float beta_Pi = 15 * a2 * (e + p);
    const float lambda_Pipi = 8 * a * r5;
    float tauPiInv = 0.0f;
    if constexpr (NC == 15) {
        if (a2 != 0.0f) bulk_relaxation(G, a2, T, rdt, tauPiInv, beta_Pi);
    }
    const float ut2 = ut * ut, t2 = t * t, t3 = t * t2;
    const float rt2 = rt * rt, rt3 = rt2 * rt;
    const float dtut = (ut - utp) * rdt;
    const float dtux = (ux - uxp) * rdt;
    const float dtuy = (uy - uyp) * rdt;
    const float Dut = ut * dtut + ux * dxut + uy * dyut;
    const float dux = ut * dtux + ux * dxux + uy * dyux;
    const float duy = ut * dtuy + ux * dxuy + uy * dyuy;
    const float Dux = -dux, Duy = -duy;
    const float dut = Dut;
    const float theta = (ut) * rt + dtut + dxux + dyuy;
    const float stt = (dtut - ut * dut) + ((ut2 - 1) * theta) * r3;
    const float stx = (dtux - dxut) / 2.0f - (ux * dut + ut * dux) / 2.0f + (ut * ux * theta) * r3;
    const float sty = (dtuy - dyut) / 2.0f - (uy * dut + ut * duy) / 2.0f + (ut * uy * theta) * r3;
    const float sxx = -(dxux + ux * dux) + ((1.0f + ux * ux) * theta) * r3;
    const float sxy = -(dxuy + dyux) / 2.0f - (uy * dux + ux * duy) / 2.0f + (ux * uy * theta) * r3;
    const float syy = -(dyuy + uy * duy) + ((1.0f + uy * uy) * theta) * r3;
    const float snn = (-ut) * rt3 + (rt2 * theta) * r3;
    const float wtx = (-dtux - dxut + ut * dux - ux * dut) / 2.0f;
    const float wty = (-dtuy - dyut + ut * duy - uy * dut) / 2.0f;
    const float wxy = (dxuy - dyux + ux * duy - uy * dux) / 2.0f;
    const float wxt = wtx, wyt = wty, wyx = -wxy;
    const float I1tt = 2.0f * ut * (pitt * Dut + pitx * Dux + pity * Duy);
    const float I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy;
    const float I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy;
    const float I1xx = 2.0f * ux * (pitx * Dut + pixx * Dux + pixy * Duy);
    const float I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy;
    const float I1yy = 2.0f * uy * (pity * Dut + pixy * Dux + piyy * Duy);
    const float I3tt = 2.0f * (pitx * wtx + pity * wty);
    const float I3tx = pitt * wxt + pity * wxy + pixx * wtx + pixy * wty;
    const float I3ty = pitt * wyt + pitx * wyx + pixy * wtx + piyy * wty;
    const float I3xx = 2.0f * (pitx * wxt + pixy * wxy);
    const float I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy;
    const float I3yy = 2.0f * (pity * wyt + pixy * wyx);
    const float ps = pitt * stt - 2.0f * pitx * stx - 2.0f * pity * sty + pixx * sxx + 2.0f * pixy * sxy + piyy * syy +
                     pinn * snn * t2 * t2;
    const float I4tt = (pitt * stt - pitx * stx - pity * sty) - ((1.0f - ut2) * ps) * r3;
    const float I4tx = (pitt * stx + pitx * stt) / 2.0f - (pitx * sxx + pixx * stx) / 2.0f -
                       (pity * sxy + pixy * sty) / 2.0f + ((ut * ux) * ps) * r3;
    const float I4ty = (pitt * sty + pity * stt) / 2.0f - (pitx * sxy + pixy * stx) / 2.0f -
                       (pity * syy + piyy * sty) / 2.0f + ((ut * uy) * ps) * r3;
    const float I4xx = (pitx * stx - pixx * sxx - pixy * sxy) + ((1.0f + ux * ux) * ps) * r3;
    const float I4xy = (pitx * sty + pity * stx) / 2.0f - (pixx * sxy + pixy * sxx) / 2.0f -
                       (pixy * syy + piyy * sxy) / 2.0f + ((ux * uy) * ps) * r3;
    const float I4yy = (pity * sty - pixy * sxy - piyy * syy) + ((1.0f + uy * uy) * ps) * r3;
    const float I4nn = (-t2 * pinn * snn) + (rt2 * ps) * r3;
    const float Itt = I1tt + delta_pipi * (theta * pitt) - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
    const float Itx = I1tx + delta_pipi * (theta * pitx) - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
    const float Ity = I1ty + delta_pipi * (theta * pity) - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
    const float Ixx = I1xx + delta_pipi * (theta * pixx) - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
    const float Ixy = I1xy + delta_pipi * (theta * pixy) - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
    const float Iyy = I1yy + delta_pipi * (theta * piyy) - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
    const float Inn = delta_pipi * (theta * pinn) + tau_pipi * I4nn - lambda_piPi * Pi * snn;
    S[4] = (2.0f * beta_pi * stt - pitt * taupiInv - Itt) * rut + pitt * dkvk;
    S[5] = (2.0f * beta_pi * stx - pitx * taupiInv - Itx) * rut + pitx * dkvk;
    S[6] = (2.0f * beta_pi * sty - pity * taupiInv - Ity) * rut + pity * dkvk;
    S[7] = 0.0f;
    S[8] = (2.0f * beta_pi * sxx - pixx * taupiInv - Ixx) * rut + pixx * dkvk;
    S[9] = (2.0f * beta_pi * sxy - pixy * taupiInv - Ixy) * rut + pixy * dkvk;
    S[10] = 0.0f;
    S[11] = (2.0f * beta_pi * syy - piyy * taupiInv - Iyy) * rut + piyy * dkvk;
    S[12] = 0.0f;
    S[13] = (2.0f * beta_pi * snn - pinn * taupiInv - Inn - (2 * (ut * pinn)) * rt) * rut + pinn * dkvk;
    if constexpr (NC == 15) {
        const float dPi = -beta_Pi * theta - Pi * tauPiInv - delta_PiPi * Pi * theta + lambda_Pipi * ps;
        S[14] = (dPi) * rut + Pi * dkvk;
    } else {
        (void)beta_Pi; (void)lambda_Pipi; (void)tauPiInv;
    }
}

// The face speed u^dir / u^tau of one half-site state (localPropagationSpeed
// and the flux functions' u^i/u^tau, LocalPropagationSpeed.cpp:14-22,
// FluxFunctions.cpp:15-25) from its recovered e, without forming u: with
// u^tau^2 = (M0 + P)/(e + P) and u^i = M_i/((e + P) u^tau) the speed is
// M_i/(M0 + P).  The recovery's rest-frame cap (u^tau > 1e3 gives u = (1,0,0,0),
// EnergyMomentumTensor.cpp:223-232) and its floors are kept.
template <bool TABLE>
__device__ __forceinline__ float face_speed_mx(const EosDev &E, const Moments &m, double Pi, double ePrev, int dir,
                                               int &nan) {
    double e;
    if (ePrev <= 0.1) e = m.M0 - m.M * QuotFast::rcp1(m.M0);
    else if (TABLE) {  // the Newton iteration in FP32
        const RootResult r = E.cs2_const ? table_root_f32<true>(E, ePrev, m.M0, m.M, Pi)
                                         : table_root_f32<false>(E, ePrev, m.M0, m.M, Pi);
        if (r.nan) nan = r.nan;
        e = r.e;
    } else e = energy_from_conserved<TABLE, QuotFast, DivRecip>(E, ePrev, m.M0, m.M, Pi, nan);
    if (__builtin_isnan(e)) nan = NAN_ENERGY;
    // the speed is FP32, and so is what sets it
    float ef = (float)e, pf;
    if constexpr (TABLE) {
        float fr;
        const float4 T = eos_segment_f32(E, ef, (float)E.e_min, (float)E.r_spacing, fr);
        pf = __builtin_fmaf(fr, T.y, T.x);
        pf = 1.e-15f < pf ? pf : 1.e-15f;
    } else {
        DivRecip div;
        pf = (float)eos_pressure<false>(E, e, div);
    }
    if (e < 1.e-7) {
        ef = 1.e-7f;
        pf = 1.e-7f;
    }
    const float P = pf + (float)Pi;
    const float A = (float)m.M0 + P;
    if (__builtin_fabsf(A) > 1.0e6f * __builtin_fabsf(ef + P)) return 0.0f;
    const float Mi = (float)(dir == 0 ? m.M1 : (dir == 1 ? m.M2 : m.M3));
    return Mi * frcp(A);
}

// One face flux between cells a = b-1 and b along dir, b = (i, j, k).  The
// limited slopes of the two cells come from the 4-point line (clamped at the
// lattice edges, which equals the reference's ghost cells), the half-site
// values are qL = q_a + D_a/2 and qR = q_b - D_b/2
// (HalfSiteExtrapolation.cpp:12-26), each side gets an FP64 root solve, and
// H = (v_R qR + v_L qL - a (qR - qL)) / 2 with a = max(|v_L|, |v_R|)
// (SemiDiscreteKurganovTadmorScheme.cpp:14-62).
//
// The reference evaluates a face once per cell: cell a (forward face) and
// cell b (backward face) each start both root solves from their own e
// (EnergyMomentumTensor.cpp:188-191: e <= 0.1 selects the closed-form
// branch, otherwise it is the Newton start).  HA is cell a's view, HB cell
// b's.  face_mode 0: one evaluation, each side started from the e of the
// cell on that side (HA = HB, only HA stored); 1: both views, the
// reference's semantics (cell b's view is evaluated only where it can
// differ); 2: cell a's view, plus cell b's view only where the two cells fall
// on different sides of the e <= 0.1 branch.
template <int NC, bool TABLE, bool TWO, int NF, bool BI, int NL>
__device__ __forceinline__ void face_mx(const MxFields &F, const EvDesc &G, const EosDev &E, int dir, int i, int j,
                                        int k, double t, float (&HA)[NL][NF], float (&HB)[TWO ? NL : 1][TWO ? NF : 1],
                                        int f) {
#pragma clang fp contract(fast)
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int ndir = dir == 0 ? nx : (dir == 1 ? ny : nz);
    const unsigned stride = dir == 0 ? 1u : (dir == 1 ? (unsigned)nx : (unsigned)nx * ny);
    const int cb = dir == 0 ? i : (dir == 1 ? j : k);
    const unsigned s0 = (unsigned)G.off + (unsigned)((dir == 0 ? 0 : i) +
                                                     nx * ((dir == 1 ? 0 : j) + ny * (dir == 2 ? 0 : k)));
    const unsigned p0 = s0 + clampi(cb - 2, ndir) * stride, p1 = s0 + clampi(cb - 1, ndir) * stride;
    const unsigned p2 = s0 + clampi(cb, ndir) * stride, p3 = s0 + clampi(cb + 1, ndir) * stride;
    float qL[NC], qR[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        if (!field_live<BI>(n)) {
            qL[n] = qR[n] = 0.0f;
            continue;
        }
        const float *__restrict__ fld = F.at(F.q, n);
        const float x0 = ldf(fld, p0), x1 = ldf(fld, p1), x2 = ldf(fld, p2), x3 = ldf(fld, p3);
        qL[n] = x1 + deriv_f<1>(x0, x1, x2);
        qR[n] = x2 - deriv_f<1>(x1, x2, x3);
    }
    const double ea = ldf(F.e, p1), eb = ldf(F.e, p2);
    const int mode = F.face_mode;
    // cell b's view differs from cell a's only where the two start values
    // differ and at least one of them leads to the Newton solve (at e <= 0.1
    // both states take the closed form whatever the start)
    const bool lowA = ea <= 0.1, lowB = eb <= 0.1;
    const bool viewB = TWO && (mode == 1 ? (ea != eb && !(lowA && lowB)) : (mode == 2 && lowA != lowB));
    // v[0], v[1]: the left and right states' speeds in cell a's view (both
    // solves started from ea); v[2], v[3] in cell b's view (from eb).  The
    // moments of a state serve both of its views.
    float v[4];
    int nan = NAN_NONE;
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {  // 0: left state, 1: right state
        double x[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            const int n = m < 8 ? m : NC - 1;
            x[m] = (m == 8 && NC == 14) ? 0.0 : (double)(side == 0 ? qL[n] : qR[n]);
        }
        const Moments mo = moments(t, x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7]);
#pragma unroll 1
        for (int view = 0; view < (viewB ? 2 : 1); ++view) {
            const double guess = view == 1 ? eb : ((!TWO || mode == 0) && side == 1 ? eb : ea);
            const float vs = face_speed_mx<TABLE>(E, mo, x[8], guess, dir, nan);
            const int kk = side + 2 * view;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q == kk) v[q] = vs;
        }
    }
    if (nan != NAN_NONE) report_nan(F.err, nan, p2);
    // H = (v_R qR + v_L qL - a (qR - qL)) / 2, with the halves taken into
    // the speeds (exact: a power of two)
    const float aA = __builtin_fmaxf(__builtin_fabsf(v[0]), __builtin_fabsf(v[1]));
    const float hRA = 0.5f * v[1], hLA = 0.5f * v[0], haA = 0.5f * aA;
#pragma unroll
    for (int n = 0; n < NC; ++n)
        if (field_live<BI>(n)) HA[field_slot<BI>(n)][f] = hRA * qR[n] + hLA * qL[n] - haA * (qR[n] - qL[n]);
    if constexpr (TWO) {
        if (!viewB) {
            v[2] = v[0];
            v[3] = v[1];
        }
        const float aB = __builtin_fmaxf(__builtin_fabsf(v[2]), __builtin_fabsf(v[3]));
        const float hRB = 0.5f * v[3], hLB = 0.5f * v[2], haB = 0.5f * aB;
#pragma unroll
        for (int n = 0; n < NC; ++n)
            if (field_live<BI>(n)) HB[field_slot<BI>(n)][f] = hRB * qR[n] + hLB * qL[n] - haB * (qR[n] - qL[n]);
    }
}

// REG_SCHEME_1 (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:689-860) in
// FP32 for the mixed form: the same factors (rho_max = 1, xi0 = 0.1, applied
// only at e <= e_reg) scale pi (pv[0..9]) and Pi (pv[10]).
template <int NC>
__device__ __forceinline__ void regulate_mx(const EvDesc &G, double t64, float (&pv)[NC - 4], const Prim &w,
                                            int &nan) {
#pragma clang fp contract(fast)
    if (w.e > G.e_reg) return;
    const float t = (float)t64, t2 = t * t;
    const float pitt = pv[0], pitx = pv[1], pity = pv[2], pitn = pv[3], pixx = pv[4], pixy = pv[5], pixn = pv[6],
                piyy = pv[7], piyn = pv[8], pinn = pv[9];
    const float ut = (float)w.ut, ux = (float)w.ux, uy = (float)w.uy, un = (float)w.un;
    const float e = (float)w.e, p = (float)w.p;
    const float pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx + 2 * pixy * pixy +
                       piyy * piyy - 2 * pitn * pitn * t2 + 2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 +
                       pinn * pinn * t2 * t2;
    const float spipi = __builtin_sqrtf(__builtin_fabsf(pipi));
    const float pimumu = pitt - pixx - piyy - pinn * t2;
    const float piu0 = -(pitn * t2 * un) + pitt * ut - pitx * ux - pity * uy;
    const float piu1 = -(pixn * t2 * un) + pitx * ut - pixx * ux - pixy * uy;
    const float piu2 = -(piyn * t2 * un) + pity * ut - pixy * ux - piyy * uy;
    const float piu3 = -(pinn * t2 * un) + pitn * ut - pixn * ux - piyn * uy;
    const float a1 = spipi * __builtin_amdgcn_rsqf(e * e + 3 * p * p);
    const float inv = frcp(0.1f * spipi);  // pi = 0: 0 * inf gives NaN, which fmax passes over, as 0 / 0 in FP64
    const float a2 = pimumu * inv, a3 = piu0 * inv, a4 = piu1 * inv, a5 = piu2 * inv, a6 = piu3 * inv;
    const float rho = __builtin_fmaxf(__builtin_fmaxf(a1, a2), __builtin_fmaxf(__builtin_fmaxf(a3, a4),
                                                                                 __builtin_fmaxf(a5, a6)));
    float facShear = __builtin_fabsf(rho) < 1.e-7f ? 1.0f : tanhf(rho) * frcp(rho);
#pragma unroll
    for (int n = 0; n < 10; ++n) pv[n] *= facShear;
    if (__builtin_isnan(facShear)) nan = NAN_REG_SHEAR;
    if constexpr (NC == 15) {
        const float Pi = pv[10];
        const float rb = __builtin_fabsf(Pi) * frcp(p);
        const float facBulk = __builtin_fabsf(rb) < 1.e-7f ? 1.0f : tanhf(rb) * frcp(rb);
        pv[10] *= facBulk;
        if (__builtin_isnan(facBulk)) nan = NAN_REG_BULK;
    }
}

constexpr int MX_THREADS = 256;

template <int TX, int TY, int TZ, bool DIM3>
struct Tile {
    static constexpr int NCELL = TX * TY * TZ;
    static constexpr int FX = (TX + 1) * TY * TZ, FY = TX * (TY + 1) * TZ, FZ = DIM3 ? TX * TY * (TZ + 1) : 0;
    // faces of the largest direction: one LDS face buffer serves every direction in turn
    static constexpr int FMAX = FX > FY ? (FX > FZ ? FX : FZ) : (FY > FZ ? FY : FZ);
    static_assert(NCELL <= MX_THREADS, "one cell per thread");
    static_assert(DIM3 || TZ == 1, "2+1D tiles are one cell thick");
};

// One RK stage for every event of a batch: block b is tile (b - blk_first[b])
// of event blk_ev[b], one cell per thread.  Direction by direction (x, y,
// eta) the block evaluates every face of the tile in that direction into one
// LDS face buffer and each cell adds its (H_before - H_after)/dx; then each
// cell forms q + dt (S + divergence + S_x + S_y + S_eta) with the FP64
// accumulation, recovers (e, p, u), regulates and stores.  TWO: both views
// of a face are stored (face modes 1, 2).
template <int NC, bool TABLE, bool DIM3, int TX, int TY, int TZ, bool TWO, bool BI>
__device__ __forceinline__ void stage_mixed_body(const MxFields &F, const EosDev &E, const EvDesc *__restrict__ ev,
                                                 const int *__restrict__ blk_ev, const int *__restrict__ blk_first) {
#pragma clang fp contract(fast)
    static_assert(!(BI && DIM3), "boost invariance is a 2+1D property");
    using T = Tile<TX, TY, TZ, DIM3>;
    constexpr int NL = BI ? NC - 4 : NC;  // fields evolved
    __shared__ float HA[NL][T::FMAX];
    __shared__ float HB[TWO ? NL : 1][TWO ? T::FMAX : 1];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int eid = __builtin_amdgcn_readfirstlane(blk_ev[blk]);
    const int tile = blk - __builtin_amdgcn_readfirstlane(blk_first[blk]);
    const EvDesc &G = ev[eid];
    if (!G.active) return;  // a removed event whose blocks are still in the table (block-uniform)
    const int ntx = (G.nx + TX - 1) / TX, nty = (G.ny + TY - 1) / TY, ntz = (G.nz + TZ - 1) / TZ;
    if (tile < 0 || tile >= ntx * nty * ntz) return;  // block-uniform: no barrier is skipped by part of a block
    const int i0 = (tile % ntx) * TX, j0 = ((tile / ntx) % nty) * TY, k0 = (tile / (ntx * nty)) * TZ;
    const StepTimes st = G.tab[G.base + F.off];
    const double t = F.stage ? st.t2 : st.t1;
    const double t_new = st.t2;

    const int c = threadIdx.x;
    const int lx = c % TX, ly = (c / TX) % TY, lz = c / (TX * TY);
    const int i = i0 + lx, j = j0 + ly, k = k0 + lz;
    const bool live = c < T::NCELL && i < G.nx && j < G.ny && k < G.nz;
    // the flux divergence of every direction, summed in FP32 in the
    // reference's order x, y, eta; the only per-cell state live across the
    // face passes
    float fdiv[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) fdiv[n] = 0.0f;
#pragma unroll 1
    for (int d = 0; d < (DIM3 ? 3 : 2); ++d) {
        // faces of direction d: (fx, fy, fz) lies between cells b - e_d and
        // b = (i0 + fx, j0 + fy, k0 + fz)
        const int ex = TX + (d == 0), ey = TY + (d == 1), ez = TZ + (d == 2);
        for (int f = threadIdx.x; f < ex * ey * ez; f += MX_THREADS) {
            const int fx = f % ex, fy = (f / ex) % ey, fz = f / (ex * ey);
            const int fi = i0 + fx, fj = j0 + fy, fk = k0 + fz;
            const bool need = (d == 0 ? fi - 1 < G.nx : fi < G.nx) && (d == 1 ? fj - 1 < G.ny : fj < G.ny) &&
                              (d == 2 ? fk - 1 < G.nz : fk < G.nz);
            if (need) face_mx<NC, TABLE, TWO, T::FMAX, BI, NL>(F, G, E, d, fi, fj, fk, t, HA, HB, f);
        }
        __syncthreads();
        if (live) {
            const float rd = 1.0f / (float)(d == 0 ? G.dx : (d == 1 ? G.dy : G.dz));
            const int fb = lx + ex * (ly + ey * lz);  // the face before the cell; the one after is fb + fstep
            const int fstep = d == 0 ? 1 : (d == 1 ? ex : ex * ey);
#pragma unroll
            for (int n = 0; n < NC; ++n) {
                if (!field_live<BI>(n)) continue;
                // the face before the cell is seen from its b side, the face after from its a side
                const int sl = field_slot<BI>(n);
                const float hb = TWO ? HB[TWO ? sl : 0][TWO ? fb : 0] : HA[sl][fb];
                fdiv[n] += (hb - HA[sl][fb + fstep]) * rd;
            }
        }
        if (d + 1 < (DIM3 ? 3 : 2)) __syncthreads();  // the next direction reuses the face buffer
    }
    if (!live) return;
    const unsigned s = (unsigned)G.off + (unsigned)(i + G.nx * (j + G.ny * k));
    // acc = q + dt (S + sum_d (divergence_d + S_d)), formed in FP64; T^{tau mu}
    // stays FP64 (it is stored as FP32 plus an FP32 low part), pi and Pi are
    // kept as the FP32 values they are stored as
    double acc[4];
    float pv[NC - 4];
    {
        float q[NC], S[NC], Sd[3][4];
#pragma unroll
        for (int n = 0; n < NC; ++n) q[n] = field_live<BI>(n) ? ldf(F.at(F.q, n), s) : 0.0f;
        if constexpr (BI) source_mx_bi<NC, TABLE>(F, G, E, s, i, j, (float)t, ldf(F.e, s), q, S, Sd);
        else source_mx<NC, TABLE, DIM3>(F, G, E, s, i, j, k, (float)t, ldf(F.e, s), q, S, Sd);
        const float dt = (float)G.dt;
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            float r = fdiv[n];
            if (!field_live<BI>(n)) {
                if (n < 4) acc[n] = 0.0;
                else pv[n - 4] = 0.0f;
            } else if (n < 4) {
                r += DIM3 ? (Sd[0][n] + Sd[1][n]) + Sd[2][n] : Sd[0][n] + Sd[1][n];
                const double qn = (double)q[n] + (double)ldf(F.at(F.ql, n), s);
                acc[n] = (qn + (double)(dt * S[n])) + (double)(dt * r);
            } else {
                pv[n - 4] = (float)(((double)q[n] + (double)(dt * S[n])) + (double)(dt * r));
            }
        }
    }
    if (F.stage == 1) {
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            if (!field_live<BI>(n)) continue;
            if (n < 4) acc[n] = (acc[n] + ((double)ldf(F.at(F.qn, n), s) + (double)ldf(F.at(F.qnl, n), s))) * 0.5;
            else pv[n - 4] = (pv[n - 4] + ldf(F.at(F.qn, n), s)) * 0.5f;
        }
    }
    int nan = NAN_NONE;
    DivRecip div;
    const Prim w = recover<TABLE, DivRecip, QuotFast>(E, t_new, acc[0], acc[1], acc[2], acc[3], (double)pv[0],
                                                      (double)pv[1], (double)pv[2], (double)pv[3],
                                                      NC == 15 ? (double)pv[NC - 5] : 0.0, (double)ldf(F.e, s), nan,
                                                      div);
    regulate_mx<NC>(G, t_new, pv, w, nan);
    // a boost-invariant event's eta components are not stored: only BI
    // kernels read its cells, and downloads give those fields as zero
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        if (!field_live<BI>(n)) continue;
        const float hi = (float)acc[n];
        stf(F.at(F.qo, n), s, hi);
        stf(F.at(F.qol, n), s, (float)(acc[n] - (double)hi));
    }
#pragma unroll
    for (int n = 4; n < NC; ++n)
        if (field_live<BI>(n)) stf(F.at(F.qo, n), s, pv[n - 4]);
    stf(F.eo, s, (float)w.e);
    stf(F.po, s, (float)w.p);
    stf(F.at(F.uo, 0), s, (float)w.ut);
    stf(F.at(F.uo, 1), s, (float)w.ux);
    stf(F.at(F.uo, 2), s, (float)w.uy);
    if constexpr (!BI) stf(F.at(F.uo, 3), s, (float)w.un);
    if (nan != NAN_NONE) report_nan(F.err, nan, s);
}

// 2+1D: 128 VGPRs (four waves per SIMD) with a 12-byte spill measured
// faster than 3 waves without; 3+1D keeps the register budget the compiler
// picks (it would spill ~110 bytes at four waves).
template <int NC, bool TABLE, bool DIM3, int TX, int TY, int TZ, bool TWO, bool BI>
__global__ __launch_bounds__(MX_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void stage_mixed_kernel(
    const MxFields F, const EosDev E, const EvDesc *__restrict__ ev, const int *__restrict__ blk_ev,
    const int *__restrict__ blk_first) {
    stage_mixed_body<NC, TABLE, DIM3, TX, TY, TZ, TWO, BI>(F, E, ev, blk_ev, blk_first);
}
template <int NC, bool TABLE, bool DIM3, int TX, int TY, int TZ, bool TWO>
__global__ __launch_bounds__(MX_THREADS) void stage_mixed_kernel_3d(const MxFields F, const EosDev E,
                                                                    const EvDesc *__restrict__ ev,
                                                                    const int *__restrict__ blk_ev,
                                                                    const int *__restrict__ blk_first) {
    stage_mixed_body<NC, TABLE, DIM3, TX, TY, TZ, TWO, false>(F, E, ev, blk_ev, blk_first);
}

}  // namespace mx
}  // namespace vh
