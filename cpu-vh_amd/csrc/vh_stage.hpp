// One fused RK stage of the viscous-hydro step on gfx950.
//
// The reference runs seven OpenMP passes per stage over the grid
// (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:895-906 and :912-923:
// source, X, Y, Z fluxes, [convex combination], primitive recovery,
// regulation, ghost refresh).  Here one thread owns one interior cell and does
// all of it in registers:
//     acc = q + dt*S                                   (eulerStepKernelSource)
//     acc += ((-H_{+1/2} + H_{-1/2})/d + S_d)*dt       (eulerStepKernelX/Y/Z, that order)
//     acc = (acc + q^n)/2                              (stage 2 only)
//     (e,p,u) = recover(acc, ePrev = e[s]); regulate pi in acc; store.
// Ghost cells are replaced by clamp-to-edge neighbour indices (every stencil
// is axis-aligned, so the reference's boundary copies are the clamped values).
// Outputs go to buffers distinct from the inputs, so neighbours read the
// stage's input state while the new state is written.
#pragma once
#include "vh_device.hpp"

namespace vh {

// Times of one step (rungeKutta2's t before and after t += dt, :900) and
// pow(t, 2) as the reference's glibc computes it (SourceTerms.cpp:415,587).
struct StepTimes {
    double t1, t1sq, t2, t2sq;
};

struct StageParams {
    const double *q[15];   // stage input conserved variables
    const double *u[4];    // stage input velocity
    const double *up[4];   // velocity used for the time derivative
    const double *e;       // stage input e (also ePrev of every recovery)
    const double *p;       // stage input p
    const double *qn[15];  // q^n for the convex combination (stage 2)
    double *qo[15];
    double *uo[4];
    double *eo;
    double *po;
    double *regfac;        // optional
    const StepTimes *tab;  // per-step times (graph replays), or null: use st
    const int *base;       // step index of graph node offset 0
    int off;
    int stage;             // 0: first stage, 1: second stage
    StepTimes st;
    double dt, dx, dy, dz, e_reg;
    double etas_min, etas_Tk_GeV, etas_slope_lo, etas_slope_hi;
    int nx, ny, nz;
    int *err;              // err[0] site, err[1..2] cell (64-bit)
};

__device__ __forceinline__ void report_nan(int *err, int site, long s) {
    if (atomicCAS(err, 0, site) == 0) {
        err[1] = (int)(s & 0xffffffff);
        err[2] = (int)(s >> 32);
    }
}

__device__ __forceinline__ double etabar_of_T(const StageParams &P, double T) {
    if (P.etas_slope_lo == 0.0 && P.etas_slope_hi == 0.0) return P.etas_min;
    double TGeV = T * HBARC;
    if (TGeV > P.etas_Tk_GeV) return P.etas_min + P.etas_slope_hi * (TGeV - P.etas_Tk_GeV);
    return P.etas_min + P.etas_slope_lo * (P.etas_Tk_GeV - TGeV);
}

// KT central-upwind flux at the face between points b and c of the line
// (a, b, c, d): rhic/src/SemiDiscreteKurganovTadmorScheme.cpp:14-62 with the
// half-site extrapolations of rhic/src/HalfSiteExtrapolation.cpp:12-26
// (forward face of cell i: a,b,c,d = i-1,i,i+1,i+2; backward: i-2,i-1,i,i+1).
template <int NC, int DIR, bool TABLE>
__device__ __forceinline__ void face_flux(const StageParams &P, const EosDev &E, long a, long b, long c, long d,
                                          double ePrev, double t, double (&H)[NC], int &nan) {
    double qR[NC], qL[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        const double *__restrict__ f = P.q[n];
        double xa = f[a], xb = f[b], xc = f[c], xd = f[d];
        qR[n] = xc - approx_deriv(xb, xc, xd) / 2;
        qL[n] = xb + approx_deriv(xa, xb, xc) / 2;
    }
    const double PiR = NC == 15 ? qR[NC - 1] : 0.0, PiL = NC == 15 ? qL[NC - 1] : 0.0;
    Prim R = recover<TABLE>(E, t, qR[0], qR[1], qR[2], qR[3], qR[4], qR[5], qR[6], qR[7], PiR, ePrev, nan);
    Prim L = recover<TABLE>(E, t, qL[0], qL[1], qL[2], qL[3], qL[4], qL[5], qL[6], qL[7], PiL, ePrev, nan);
    const double vR = DIR == 0 ? R.ux : (DIR == 1 ? R.uy : R.un);
    const double vL = DIR == 0 ? L.ux : (DIR == 1 ? L.uy : L.un);
    // localPropagationSpeed: fmax(rho(left), rho(right)), SpectralRadius.cpp:14-24
    const double amax = __builtin_fmax(__builtin_fabs(vL / L.ut), __builtin_fabs(vR / R.ut));
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        double FqR = vR * qR[n] / R.ut;  // FluxFunctions.cpp:15-25
        double FqL = vL * qL[n] / L.ut;
        double res = FqR + FqL - amax * (qR[n] - qL[n]);
        res /= 2;
        H[n] = res;
    }
}

// One direction's contribution (eulerStepKernelX/Y/Z, :300-634, with
// loadSourceTermsX/Y/Z, rhic/src/SourceTerms.cpp:316-419), added into acc.
template <int NC, int DIR, bool TABLE>
__device__ __forceinline__ void dir_contrib(const StageParams &P, const EosDev &E, long s, int coord, int ncoord,
                                            long stride, double ePrev, double t, double t2pow, double ut_c,
                                            double v_c, double (&acc)[NC], int &nan) {
    auto nb = [&](int o) -> long {
        int c = coord + o;
        c = c < 0 ? 0 : (c >= ncoord ? ncoord - 1 : c);
        return s + (long)(c - coord) * stride;
    };
    const long sm2 = nb(-2), sm1 = nb(-1), sp1 = nb(1), sp2 = nb(2);
    const double d = DIR == 0 ? P.dx : (DIR == 1 ? P.dy : P.dz);
    double r[NC];
    {
        double H[NC];
        face_flux<NC, DIR, TABLE>(P, E, sm1, s, sp1, sp2, ePrev, t, H, nan);
#pragma unroll
        for (int n = 0; n < NC; ++n) r[n] = -H[n];
    }
    {
        double H[NC];
        face_flux<NC, DIR, TABLE>(P, E, sm2, sm1, s, sp1, ePrev, t, H, nan);
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            r[n] += H[n];
            r[n] /= d;
        }
    }
    const double fac = 1 / d / 2;
    auto dq = [&](int n) { return (P.q[n][sp1] - P.q[n][sm1]) * fac; };
    const double v = v_c / ut_c;
    double S0, S1, S2, S3;
    if constexpr (DIR == 0) {
        double dxpitt = dq(4), dxpitx = dq(5), dxpity = dq(6), dxpitn = dq(7), dxpixx = dq(8), dxpixy = dq(9),
               dxpixn = dq(10);
        if constexpr (NC == 14) {
            S0 = dxpitt * v - dxpitx;
            S1 = dxpitx * v - dxpixx;
        } else {
            double dxPi = dq(14);
            S0 = dxpitt * v - dxpitx - v * dxPi;
            S1 = dxpitx * v - dxpixx - dxPi;
        }
        S2 = dxpity * v - dxpixy;
        S3 = dxpitn * v - dxpixn;
    } else if constexpr (DIR == 1) {
        double dypitt = dq(4), dypitx = dq(5), dypity = dq(6), dypitn = dq(7), dypixy = dq(9), dypiyy = dq(11),
               dypiyn = dq(12);
        if constexpr (NC == 14) {
            S0 = dypitt * v - dypity;
            S2 = dypity * v - dypiyy;
        } else {
            double dyPi = dq(14);
            S0 = dypitt * v - dypity - v * dyPi;
            S2 = dypity * v - dypiyy - dyPi;
        }
        S1 = dypitx * v - dypixy;
        S3 = dypitn * v - dypiyn;
    } else {
        double dnpitt = dq(4), dnpitx = dq(5), dnpity = dq(6), dnpitn = dq(7), dnpixn = dq(10), dnpiyn = dq(12),
               dnpinn = dq(13);
        if constexpr (NC == 14) {
            S0 = dnpitt * v - dnpitn;
            S3 = dnpitn * v - dnpinn;
        } else {
            double dnPi = dq(14);
            S0 = dnpitt * v - dnpitn - v * dnPi;
            S3 = dnpitn * v - dnpinn - dnPi / t2pow;
        }
        S1 = dnpitx * v - dnpixn;
        S2 = dnpity * v - dnpiyn;
    }
    r[0] += S0; r[0] *= P.dt;
    r[1] += S1; r[1] *= P.dt;
    r[2] += S2; r[2] *= P.dt;
    r[3] += S3; r[3] *= P.dt;
#pragma unroll
    for (int n = 4; n < NC; ++n) r[n] *= P.dt;
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] += r[n];
}

// loadSourceTerms2 + setPimunuSourceTerms (rhic/src/SourceTerms.cpp:421-612,
// :55-313), then acc = q + dt*S (eulerStepKernelSource :233-296).
template <int NC, bool TABLE>
__device__ __forceinline__ void source_init(const StageParams &P, const EosDev &E, long s, int i, int j, int k,
                                            double t, double t2pow, double ecell, double (&acc)[NC]) {
    const double ttt = acc[0], ttx = acc[1], tty = acc[2], ttn = acc[3];
    const double pitt = acc[4], pitx = acc[5], pity = acc[6], pitn = acc[7], pixx = acc[8], pixy = acc[9],
                 pixn = acc[10], piyy = acc[11], piyn = acc[12], pinn = acc[13];
    const double Pi = NC == 15 ? acc[NC - 1] : 0.0;
    const double *__restrict__ utv = P.u[0];
    const double *__restrict__ uxv = P.u[1];
    const double *__restrict__ uyv = P.u[2];
    const double *__restrict__ unv = P.u[3];
    const double *__restrict__ pv = P.p;
    const double p = pv[s];
    const double ut = utv[s], ux = uxv[s], uy = uyv[s], un = unv[s];
    const double utp = P.up[0][s], uxp = P.up[1][s], uyp = P.up[2][s], unp = P.up[3][s];
    const long strY = P.nx, strZ = (long)P.nx * P.ny;
    const long sl = i > 0 ? s - 1 : s, sr = i < P.nx - 1 ? s + 1 : s;
    const long sb = j > 0 ? s - strY : s, stp = j < P.ny - 1 ? s + strY : s;
    const long sk = k > 0 ? s - strZ : s, sf = k < P.nz - 1 ? s + strZ : s;
    const double dxut = approx_deriv(utv[sl], ut, utv[sr]) / P.dx;
    const double dxux = approx_deriv(uxv[sl], ux, uxv[sr]) / P.dx;
    const double dxuy = approx_deriv(uyv[sl], uy, uyv[sr]) / P.dx;
    const double dxun = approx_deriv(unv[sl], un, unv[sr]) / P.dx;
    const double dyut = approx_deriv(utv[sb], ut, utv[stp]) / P.dy;
    const double dyux = approx_deriv(uxv[sb], ux, uxv[stp]) / P.dy;
    const double dyuy = approx_deriv(uyv[sb], uy, uyv[stp]) / P.dy;
    const double dyun = approx_deriv(unv[sb], un, unv[stp]) / P.dy;
    const double dnut = approx_deriv(utv[sk], ut, utv[sf]) / P.dz;
    const double dnux = approx_deriv(uxv[sk], ux, uxv[sf]) / P.dz;
    const double dnuy = approx_deriv(uyv[sk], uy, uyv[sf]) / P.dz;
    const double dnun = approx_deriv(unv[sk], un, unv[sf]) / P.dz;
    const double dxp = approx_deriv(pv[sl], p, pv[sr]) / P.dx;
    const double dyp = approx_deriv(pv[sb], p, pv[stp]) / P.dy;
    const double dnp = approx_deriv(pv[sk], p, pv[sf]) / P.dz;

    double S[NC];
    const double pP = p + Pi;
    const double tnn = (ecell + pP) * un * un + pP / t / t + pinn;  // Tnn(e, p+Pi, un, pinn, t)
    const double vx = ux / ut, vy = uy / ut, vn = un / ut;
    const double dxvx = (dxux - vx * dxut) / ut;
    const double dyvy = (dyuy - vy * dyut) / ut;
    const double dnvn = (dnun - vn * dnut) / ut;
    const double dkvk = dxvx + dyvy + dnvn;
    const double src0 = 0.0;  // dynamical sources are zero outside IC type 12
    S[0] = src0 - (ttt / t + t * tnn) + dkvk * (pitt - p - Pi) - vx * dxp - vy * dyp - vn * dnp;
    S[1] = src0 - ttx / t - dxp + dkvk * pitx;
    S[2] = src0 - tty / t - dyp + dkvk * pity;
    S[3] = src0 - 3 * ttn / t - dnp / t2pow + dkvk * pitn;

    const double e = ecell;
    const double T = eos_temperature<TABLE>(E, e);
    const double d_etabar = etabar_of_T(P, T);
    const double taupiInv = T / 5 / d_etabar;
    const double beta_pi = (e + p) / 5;
    const double cs2 = eos_cs2<TABLE>(E, e);
    const double a = 1.0 / 3.0 - cs2;
    const double a2 = a * a;
    const double beta_Pi = 15 * a2 * (e + p);
    const double lambda_Pipi = 8 * a / 5;
    double tauPiInv = 0.0;
    if constexpr (NC == 15) {
        // bulkViscosityToEntropyDensity (SourceTerms.cpp:40-48); with a == 0 the
        // quotient is +0 for any positive finite zeta/s, so exp is skipped.
        if (a2 != 0.0) {
            double x = T / 1.01355, zetabar;
            if (x > 1.05) zetabar = 0.9 * exp(-(x - 1) / 0.025) + 0.25 * exp(-(x - 1) / 0.13) + 0.001;
            else if (x < 0.995) zetabar = 0.9 * exp((x - 1) / 0.0025) + 0.22 * exp((x - 1) / 0.022) + 0.03;
            else zetabar = -13.77 * x * x + 27.55 * x - 13.45;
            tauPiInv = 15 * a2 * T / zetabar;
        }
    }
    const double ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;
    const double d_dt = P.dt;
    const double dtut = (ut - utp) / d_dt;
    const double dtux = (ux - uxp) / d_dt;
    const double dtuy = (uy - uyp) / d_dt;
    const double dtun = (un - unp) / d_dt;
    const double Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
    const double DuxUpper = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const double Dux = -DuxUpper;
    const double DuyUpper = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const double Duy = -DuyUpper;
    const double DunUpper = ut * dtun + ux * dxun + uy * dyun + un * dnun + 2 * ut * un / t;
    const double Dun = -t2 * DunUpper;
    const double dut = Dut - t * un * un;
    const double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
    const double theta = ut / t + dtut + dxux + dyuy + dnun;
    const double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1) * theta / 3.0;
    const double stx = -(t * un2 * ux) / 2.0 + (dtux - dxut) / 2.0 - (ux * dut + ut * dux) / 2.0 + ut * ux * theta / 3.0;
    const double sty = -(t * un2 * uy) / 2.0 + (dtuy - dyut) / 2.0 - (uy * dut + ut * duy) / 2.0 + ut * uy * theta / 3.0;
    const double stn = -un * (2.0 * ut2 + t2 * un2) / (2.0 * t) + (dtun - dnut / t2) / 2.0 -
                       (un * dut + ut * dun) / 2.0 + ut * un * theta / 3.0;
    const double sxx = -(dxux + ux * dux) + (1.0 + ux * ux) * theta / 3.0;
    const double sxy = -(dxuy + dyux) / 2.0 - (uy * dux + ux * duy) / 2.0 + ux * uy * theta / 3.0;
    const double sxn = -ut * ux * un / t - (dxun + dnux / t2) / 2.0 - (un * dux + ux * dun) / 2.0 + ux * un * theta / 3.0;
    const double syy = -(dyuy + uy * duy) + (1.0 + uy * uy) * theta / 3.0;
    const double syn = -ut * uy * un / t - (dyun + dnuy / t2) / 2.0 - (un * duy + uy * dun) / 2.0 + uy * un * theta / 3.0;
    const double snn = -ut * (1.0 + 2.0 * t2 * un2) / t3 - dnun / t2 - un * dun + (1.0 / t2 + un2) * theta / 3.0;
    const double wtx = (-dtux - dxut + ut * dux - ux * dut - t * un2 * ux) / 2.0;
    const double wty = (-dtuy - dyut + ut * duy - uy * dut - t * un2 * uy) / 2.0;
    const double wtn = (-t2 * dtun - 2.0 * t * un - dnut - ut * Dun - t2 * un * dut - t3 * un * un2) / 2.0;
    const double wxy = (dxuy - dyux + ux * duy - uy * dux) / 2.0;
    const double wxn = (t2 * dxun - dnux - ux * Dun - t2 * un * dux) / 2.0;
    const double wyn = (t2 * dyun - dnuy - uy * Dun - t2 * un * duy) / 2.0;
    const double wxt = wtx, wyt = wty, wnt = wtn / t2, wyx = -wxy, wnx = -wxn / t2, wny = -wyn / t2;
    const double I1tt = 2.0 * ut * (pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun);
    const double I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy +
                        (pitn * ux + pixn * ut) * Dun;
    const double I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy +
                        (pitn * uy + piyn * ut) * Dun;
    const double I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux + (pity * un + piyn * ut) * Duy +
                        (pitn * un + pinn * ut) * Dun;
    const double I1xx = 2.0 * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
    const double I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy +
                        (pixn * uy + piyn * ux) * Dun;
    const double I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux + (pixy * un + piyn * ux) * Duy +
                        (pixn * un + pinn * ux) * Dun;
    const double I1yy = 2.0 * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
    const double I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux + (piyy * un + piyn * uy) * Duy +
                        (piyn * un + pinn * uy) * Dun;
    const double I1nn = 2.0 * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun);
    const double I2tt = theta * pitt, I2tx = theta * pitx, I2ty = theta * pity, I2tn = theta * pitn,
                 I2xx = theta * pixx, I2xy = theta * pixy, I2xn = theta * pixn, I2yy = theta * piyy,
                 I2yn = theta * piyn, I2nn = theta * pinn;
    const double I3tt = 2.0 * (pitx * wtx + pity * wty + pitn * wtn);
    const double I3tx = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
    const double I3ty = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
    const double I3tn = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
    const double I3xx = 2.0 * (pitx * wxt + pixy * wxy + pixn * wxn);
    const double I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
    const double I3xn = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
    const double I3yy = 2.0 * (pity * wyt + pixy * wyx + piyn * wyn);
    const double I3yn = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
    const double I3nn = 2.0 * (pitn * wnt + pixn * wnx + piyn * wny);
    const double ux2 = ux * ux, uy2 = uy * uy;
    const double ps = pitt * stt - 2.0 * pitx * stx - 2.0 * pity * sty + pixx * sxx + 2.0 * pixy * sxy + piyy * syy -
                      2.0 * pitn * stn * t2 + 2.0 * pixn * sxn * t2 + 2.0 * piyn * syn * t2 + pinn * snn * t2 * t2;
    const double I4tt = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - (1.0 - ut2) * ps / 3.0;
    const double I4tx = (pitt * stx + pitx * stt) / 2.0 - (pitx * sxx + pixx * stx) / 2.0 -
                        (pity * sxy + pixy * sty) / 2.0 - t2 * (pitn * sxn + pixn * stn) / 2.0 + (ut * ux) * ps / 3.0;
    const double I4ty = (pitt * sty + pity * stt) / 2.0 - (pitx * sxy + pixy * stx) / 2.0 -
                        (pity * syy + piyy * sty) / 2.0 - t2 * (pitn * syn + piyn * stn) / 2.0 + (ut * uy) * ps / 3.0;
    const double I4tn = (pitt * stn + pitn * stt) / 2.0 - (pitx * sxn + pixn * stx) / 2.0 -
                        (pity * syn + piyn * sty) / 2.0 - t2 * (pitn * snn + pinn * stn) / 2.0 + (ut * un) * ps / 3.0;
    const double I4xx = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + (1.0 + ux2) * ps / 3.0;
    const double I4xy = (pitx * sty + pity * stx) / 2.0 - (pixx * sxy + pixy * sxx) / 2.0 -
                        (pixy * syy + piyy * sxy) / 2.0 - t2 * (pixn * syn + piyn * sxn) / 2.0 + (ux * uy) * ps / 3.0;
    const double I4xn = (pitx * stn + pitn * stx) / 2.0 - (pixx * sxn + pixn * sxx) / 2.0 -
                        (pixy * syn + piyn * sxy) / 2.0 - t2 * (pixn * snn + pinn * sxn) / 2.0 + (ux * un) * ps / 3.0;
    const double I4yy = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + (1.0 + uy2) * ps / 3.0;
    const double I4yn = (pity * stn + pitn * sty) / 2.0 - (pixy * sxn + pixn * sxy) / 2.0 -
                        (piyy * syn + piyn * syy) / 2.0 - t2 * (piyn * snn + pinn * syn) / 2.0 + (uy * un) * ps / 3.0;
    const double I4nn = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + (1.0 / t2 + un2) * ps / 3.0;
    const double Itt = I1tt + delta_pipi * I2tt - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
    const double Itx = I1tx + delta_pipi * I2tx - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
    const double Ity = I1ty + delta_pipi * I2ty - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
    const double Itn = I1tn + delta_pipi * I2tn - I3tn + tau_pipi * I4tn - lambda_piPi * Pi * stn;
    const double Ixx = I1xx + delta_pipi * I2xx - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
    const double Ixy = I1xy + delta_pipi * I2xy - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
    const double Ixn = I1xn + delta_pipi * I2xn - I3xn + tau_pipi * I4xn - lambda_piPi * Pi * sxn;
    const double Iyy = I1yy + delta_pipi * I2yy - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
    const double Iyn = I1yn + delta_pipi * I2yn - I3yn + tau_pipi * I4yn - lambda_piPi * Pi * syn;
    const double Inn = I1nn + delta_pipi * I2nn - I3nn + tau_pipi * I4nn - lambda_piPi * Pi * snn;
    S[4] = (2.0 * beta_pi * stt - pitt * taupiInv - Itt - 2.0 * un * t * pitn) / ut + pitt * dkvk;
    S[5] = (2.0 * beta_pi * stx - pitx * taupiInv - Itx - un * t * pixn) / ut + pitx * dkvk;
    S[6] = (2.0 * beta_pi * sty - pity * taupiInv - Ity - un * t * piyn) / ut + pity * dkvk;
    S[7] = (2.0 * beta_pi * stn - pitn * taupiInv - Itn - un * t * pinn - (ut * pitn + un * pitt) / t) / ut +
           pitn * dkvk;
    S[8] = (2.0 * beta_pi * sxx - pixx * taupiInv - Ixx) / ut + pixx * dkvk;
    S[9] = (2.0 * beta_pi * sxy - pixy * taupiInv - Ixy) / ut + pixy * dkvk;
    S[10] = (2.0 * beta_pi * sxn - pixn * taupiInv - Ixn - (ut * pixn + un * pitx) / t) / ut + pixn * dkvk;
    S[11] = (2.0 * beta_pi * syy - piyy * taupiInv - Iyy) / ut + piyy * dkvk;
    S[12] = (2.0 * beta_pi * syn - piyn * taupiInv - Iyn - (ut * piyn + un * pity) / t) / ut + piyn * dkvk;
    S[13] = (2.0 * beta_pi * snn - pinn * taupiInv - Inn - 2 * (ut * pinn + un * pitn) / t) / ut + pinn * dkvk;
    if constexpr (NC == 15) {
        const double dPi = -beta_Pi * theta - Pi * tauPiInv - delta_PiPi * Pi * theta + lambda_Pipi * ps;
        S[14] = dPi / ut + Pi * dkvk;
    } else {
        (void)beta_Pi; (void)lambda_Pipi; (void)tauPiInv;
    }
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] = acc[n] + P.dt * S[n];
}

// REG_SCHEME_1, rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:689-860, one cell.
template <int NC>
__device__ __forceinline__ double regulate(const StageParams &P, double t, double (&acc)[NC], const Prim &w,
                                           int &nan) {
    const double pitt = acc[4], pitx = acc[5], pity = acc[6], pitn = acc[7], pixx = acc[8], pixy = acc[9],
                 pixn = acc[10], piyy = acc[11], piyn = acc[12], pinn = acc[13];
    const double Pi = NC == 15 ? acc[NC - 1] : 0.0;
    const double ut = w.ut, ux = w.ux, uy = w.uy, un = w.un;
    const double t2 = t * t;
    double facShear = 1.0, facBulk = 1.0;
    if (!(w.e > P.e_reg)) {  // the factors are overridden to 1 above e_reg (:835)
        const double xi0 = 0.1, rhomax = 1.0;
        const double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx + 2 * pixy * pixy +
                            piyy * piyy - 2 * pitn * pitn * t2 + 2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 +
                            pinn * pinn * t2 * t2;
        const double spipi = __builtin_sqrt(__builtin_fabs(pipi));
        const double pimumu = pitt - pixx - piyy - pinn * t * t;
        const double piu0 = -(pitn * t2 * un) + pitt * ut - pitx * ux - pity * uy;
        const double piu1 = -(pixn * t2 * un) + pitx * ut - pixx * ux - pixy * uy;
        const double piu2 = -(piyn * t2 * un) + pity * ut - pixy * ux - piyy * uy;
        const double piu3 = -(pinn * t2 * un) + pitn * ut - pixn * ux - piyn * uy;
        const double a1 = spipi / rhomax / (double)f32_sqrt((float)(w.e * w.e + 3 * w.p * w.p));
        const double a2 = pimumu / xi0 / rhomax / spipi;
        const double a3 = piu0 / xi0 / rhomax / spipi;
        const double a4 = piu1 / xi0 / rhomax / spipi;
        const double a5 = piu2 / xi0 / rhomax / spipi;
        const double a6 = piu3 / xi0 / rhomax / spipi;
        const double a12 = __builtin_fmax(a1, a2), a34 = __builtin_fmax(a3, a4), a56 = __builtin_fmax(a5, a6);
        const double rho = __builtin_fmax(a12, __builtin_fmax(a34, a56));
        facShear = tanh(rho) / rho;
        if (__builtin_fabs(rho) < 1.e-7) facShear = 1.0;
        if constexpr (NC == 15) {
            const double rhoBulk = __builtin_fabs(Pi) / w.p;
            facBulk = tanh(rhoBulk) / rhoBulk;
            if (__builtin_fabs(rhoBulk) < 1.e-7) facBulk = 1.0;
        }
    }
#pragma unroll
    for (int n = 4; n < 14; ++n) acc[n] *= facShear;
    if (__builtin_isnan(facShear)) nan = NAN_REG_SHEAR;
    if constexpr (NC == 15) {
        acc[NC - 1] *= facBulk;
        if (__builtin_isnan(facBulk)) nan = NAN_REG_BULK;
    }
    return facShear;
}

template <int NC, bool TABLE, bool DIM3>
__global__ __launch_bounds__(256) void stage_kernel(const StageParams P, const EosDev E) {
    const long N = (long)P.nx * P.ny * P.nz;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int i = (int)(s % P.nx);
    const long rest = s / P.nx;
    const int j = (int)(rest % P.ny);
    const int k = (int)(rest / P.ny);
    const StepTimes st = P.tab ? P.tab[*P.base + P.off] : P.st;
    const double t = P.stage ? st.t2 : st.t1;
    const double t2pow = P.stage ? st.t2sq : st.t1sq;
    const double t_new = st.t2;
    int nan = NAN_NONE;

    const double ecell = P.e[s];
    double acc[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] = P.q[n][s];
    source_init<NC, TABLE>(P, E, s, i, j, k, t, t2pow, ecell, acc);

    const double ut_c = P.u[0][s];
    dir_contrib<NC, 0, TABLE>(P, E, s, i, P.nx, 1, ecell, t, t2pow, ut_c, P.u[1][s], acc, nan);
    dir_contrib<NC, 1, TABLE>(P, E, s, j, P.ny, P.nx, ecell, t, t2pow, ut_c, P.u[2][s], acc, nan);
    if constexpr (DIM3) {
        dir_contrib<NC, 2, TABLE>(P, E, s, k, P.nz, (long)P.nx * P.ny, ecell, t, t2pow, ut_c, P.u[3][s], acc, nan);
    } else {
        // With one rapidity cell every point of the eta line is the cell itself:
        // both face fluxes coincide and the contribution is exactly +0.
#pragma unroll
        for (int n = 0; n < NC; ++n) acc[n] += 0.0;
    }
    if (P.stage == 1) {
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            acc[n] += P.qn[n][s];
            acc[n] /= 2;
        }
    }
    const Prim w = recover<TABLE>(E, t_new, acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], acc[6], acc[7],
                                  NC == 15 ? acc[NC - 1] : 0.0, ecell, nan);
    const double fs = regulate<NC>(P, t_new, acc, w, nan);
#pragma unroll
    for (int n = 0; n < NC; ++n) P.qo[n][s] = acc[n];
    P.eo[s] = w.e;
    P.po[s] = w.p;
    P.uo[0][s] = w.ut;
    P.uo[1][s] = w.ux;
    P.uo[2][s] = w.uy;
    P.uo[3][s] = w.un;
    if (P.regfac) P.regfac[s] = fs;
    if (nan != NAN_NONE) report_nan(P.err, nan, s);
}

}  // namespace vh
