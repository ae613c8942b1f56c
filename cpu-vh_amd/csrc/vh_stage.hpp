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
#include "vh_libm.hpp"

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
    double *regfac;        // optional: regFacShear of the last stage
    double *regfacb;       // optional: regFacBulk of the last stage (bulk builds)
    const StepTimes *tab;  // per-step times (graph replays), or null: use st
    const int *base;       // step index of graph node offset 0
    int off;
    int stage;             // 0: first stage, 1: second stage
    StepTimes st;
    double dt, dx, dy, dz, e_reg;
    double etas_min, etas_Tk_GeV, etas_slope_lo, etas_slope_hi;
    double zeta_Tpeak_GeV, zeta_peak;
    int nx, ny, nz;
    int *err;              // err[0] site, err[1..2] cell (64-bit)
};

// One event's lattice, transport parameters and clock.  A single-event
// launch builds it from StageParams; in a batch, event k's copy sits in a
// device array and its cells occupy [off, off + nx*ny*nz) of every pooled
// field.  Step n of the event runs at tab[n] (host-computed times).
struct EvDesc {
    int nx, ny, nz, active;
    long off;
    double dx, dy, dz, dt, e_reg;
    double etas_min, etas_Tk_GeV, etas_slope_lo, etas_slope_hi;
    double zeta_Tpeak_GeV, zeta_peak;
    const StepTimes *tab;
    int base;  // step index of graph node offset 0
    int pad;
};

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  The renumbering hands each XCD runs of
// consecutive blocks, the runs dealt round-robin: within a run a block's
// stencil neighbours (the blocks just before and after it) share its L2, and
// the interleaved runs spread a launch whose blocks cost unequally (the
// fireball's root solves against the closed form outside it) evenly over the
// XCDs, where one contiguous eighth per XCD leaves the XCD with the costliest
// eighth last.  Runs are XCD_RUN blocks long, or an eighth of a smaller
// launch.  A bijection on [0, nblocks): the blocks past the last whole round
// of runs keep their order.
constexpr int NUM_XCD = 8, XCD_RUN = 64;
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int run = nblocks / NUM_XCD < XCD_RUN ? nblocks / NUM_XCD : XCD_RUN;
    if (run == 0) return b;
    const int round = NUM_XCD * run, full = nblocks / round * round;
    if (b >= full) return b;
    const int k = b % NUM_XCD, j = b / NUM_XCD;  // the j-th block XCD k receives
    return (j / run * NUM_XCD + k) * run + j % run;
}

__device__ __forceinline__ void report_nan(int *err, int site, long s) {
    if (atomicCAS(err, 0, site) == 0) {
        err[1] = (int)(s & 0xffffffff);
        err[2] = (int)(s >> 32);
    }
}

template <class GEO>
__device__ __forceinline__ double etabar_of_T(const GEO &G, double T) {
    if (G.etas_slope_lo == 0.0 && G.etas_slope_hi == 0.0) return G.etas_min;
    double TGeV = T * HBARC;
    if (TGeV > G.etas_Tk_GeV) return G.etas_min + G.etas_slope_hi * (TGeV - G.etas_Tk_GeV);
    return G.etas_min + G.etas_slope_lo * (G.etas_Tk_GeV - TGeV);
}

// zeta/s at T (fm^-1): the reference's bulkViscosityToEntropyDensity
// (rhic/src/SourceTerms.cpp:40-48) of x = T/1.01355, or, with
// zeta_Tpeak_GeV > 0, the same shape of x = T/T_peak scaled so that its peak
// (0.33 at x = 1) becomes zeta_peak.  The exponentials are fdlibm_exp (this
// is reached only where c_s^2 != 1/3, which the reference's EoS never gives).
template <class GEO>
__device__ __forceinline__ double zeta_over_s(const GEO &G, double T) {
    const bool ext = G.zeta_Tpeak_GeV > 0.0;
    const double x = ext ? T * HBARC / G.zeta_Tpeak_GeV : T / 1.01355;
    double z;
    if (x > 1.05) z = 0.9 * fdlibm_exp(-(x - 1) / 0.025) + 0.25 * fdlibm_exp(-(x - 1) / 0.13) + 0.001;
    else if (x < 0.995) z = 0.9 * fdlibm_exp((x - 1) / 0.0025) + 0.22 * fdlibm_exp((x - 1) / 0.022) + 0.03;
    else z = -13.77 * x * x + 27.55 * x - 13.45;
    return ext ? G.zeta_peak / 0.33 * z : z;
}

// Primitive state on one side of a face, with the reciprocal of u^tau used by
// every flux of that side.
struct Side {
    double v;  // u^x, u^y or u^eta (direction of the face normal)
    Recip ut;
};

// One face's KT speed: localPropagationSpeed = fmax(rho(left), rho(right))
// (LocalPropagationSpeed.cpp:14-22, SpectralRadius.cpp:14-24).
template <class DV>
__device__ __forceinline__ double face_speed(const Side &R, const Side &L, DV &div) {
    return __builtin_fmax(__builtin_fabs(div(L.v, L.ut)), __builtin_fabs(div(R.v, R.ut)));
}

// Flux of one variable at one face: (F(qR) + F(qL) - a (qR - qL)) / 2 with
// F(q) = v q / u^tau (SemiDiscreteKurganovTadmorScheme.cpp:53-61,
// FluxFunctions.cpp:15-25).
template <class DV>
__device__ __forceinline__ double face_var_flux(const Side &R, const Side &L, double a, double qR, double qL, DV &div) {
    const double FqR = div(R.v * qR, R.ut);
    const double FqL = div(L.v * qL, L.ut);
    double res = FqR + FqL - a * (qR - qL);
    res /= 2;
    return res;
}

template <int DIR>
__device__ __forceinline__ Side make_side(const Prim &w) {
    return Side{DIR == 0 ? w.ux : (DIR == 1 ? w.uy : w.un), make_recip(w.ut)};
}

// The three limited slopes of one variable on the 5-point line
// (qmm, qm, q, qp, qpp): D(qmm,qm,q), D(qm,q,qp), D(q,qp,qpp).  The four
// half-site values of the two faces follow from them exactly as
// HalfSiteExtrapolation.cpp:12-26 forms them (the middle slope is the one both
// the forward-left and the backward-right extrapolations evaluate).
struct Line {
    double qm, q, qp, Dm, Dc, Dp;
};
__device__ __forceinline__ Line make_line(const double *__restrict__ f, long smm, long sm, long s, long sp,
                                          long spp) {
    const double xmm = f[smm], xm = f[sm], x = f[s], xp = f[sp], xpp = f[spp];
    return Line{xm, x, xp, approx_deriv(xmm, xm, x), approx_deriv(xm, x, xp), approx_deriv(x, xp, xpp)};
}
__device__ __forceinline__ double right_fwd(const Line &l) { return l.qp - l.Dp / 2; }
__device__ __forceinline__ double left_fwd(const Line &l) { return l.q + l.Dc / 2; }
__device__ __forceinline__ double right_bwd(const Line &l) { return l.q - l.Dc / 2; }
__device__ __forceinline__ double left_bwd(const Line &l) { return l.qm + l.Dm / 2; }

// Components of pi^{tau mu} and pi^{a mu} (a = x, y, eta) in the conserved
// ordering, for the pi-gradient terms of loadSourceTermsX/Y/Z
// (SourceTerms.cpp:316-419): S[mu] = d_a(pi^{tau mu}) v_a - d_a(pi^{a mu}).
__device__ __forceinline__ int pi_a_mu(int dir, int mu) {
    // x: pitx pixx pixy pixn | y: pity pixy piyy piyn | eta: pitn pixn piyn pinn
    const int tab = dir == 0 ? 0x0A090805 : (dir == 1 ? 0x0C0B0906 : 0x0D0C0A07);
    return (tab >> (8 * mu)) & 0xff;
}

// One direction's contribution (eulerStepKernelX/Y/Z, :300-634, with
// loadSourceTermsX/Y/Z, rhic/src/SourceTerms.cpp:316-419), handed to sink
// variable by variable: r = ((-H_{+1/2} + H_{-1/2}) / d [+ S_d]) * dt.  dir
// is wave-uniform, so the three directions share one body.  Every face flux
// uses the owning cell's e as the root solve's previous value, as the
// reference does; the four half-site root solves run in a loop so the solver
// is instantiated once.
template <int NC, bool TABLE, class DV, class SINK, class GEO>
__device__ __forceinline__ void dir_contrib(const StageParams &P, const GEO &G, const EosDev &E, long s, int dir, int coord,
                                            int ncoord, long stride, double ePrev, double t, const Recip &rt2pow,
                                            const Recip &rut, double v_c, SINK &&sink, int &nan, DV &div) {
    auto nb = [&](int o) -> long {
        int c = coord + o;
        c = c < 0 ? 0 : (c >= ncoord ? ncoord - 1 : c);
        return s + (long)(c - coord) * stride;
    };
    const long sm2 = nb(-2), sm1 = nb(-1), sp1 = nb(1), sp2 = nb(2);
    const double d = dir == 0 ? G.dx : (dir == 1 ? G.dy : G.dz);
    const Recip rd = make_recip(d);

    // Half-site values of the recovery inputs (q[0..7], Pi) on both faces.
    double qRf[9], qLf[9], qRb[9], qLb[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        if (m == 8 && NC == 14) { qRf[m] = qLf[m] = qRb[m] = qLb[m] = 0.0; continue; }
        const int n = m < 8 ? m : NC - 1;
        const Line l = make_line(P.q[n], sm2, sm1, s, sp1, sp2);
        qRf[m] = right_fwd(l); qLf[m] = left_fwd(l);
        qRb[m] = right_bwd(l); qLb[m] = left_bwd(l);
    }
    // Four root solves: forward-right, forward-left, backward-right,
    // backward-left, all from the cell's e, stepped together.
    Side sd[4];
    {
        double x[4][9];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            x[0][m] = qRf[m];
            x[1][m] = qLf[m];
            x[2][m] = qRb[m];
            x[3][m] = qLb[m];
        }
        Prim w[4];
        recover4<TABLE>(E, t, x, ePrev, nan, div, w);
#pragma unroll
        for (int q = 0; q < 4; ++q) sd[q] = Side{dir == 0 ? w[q].ux : (dir == 1 ? w[q].uy : w[q].un), make_recip(w[q].ut)};
    }
    const Side Rf = sd[0], Lf = sd[1], Rb = sd[2], Lb = sd[3];
    const double af = face_speed(Rf, Lf, div), ab = face_speed(Rb, Lb, div);

    // pi-gradient terms of T^{tau mu} (central differences, fac = 1/d/2)
    const double fac = 1 / d / 2;
    auto dq = [&](int n) { return (P.q[n][sp1] - P.q[n][sm1]) * fac; };
    const double v = div(v_c, rut);
    double S[4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) S[mu] = dq(4 + mu) * v - dq(pi_a_mu(dir, mu));
    if constexpr (NC == 15) {
        const double dPi = dq(14);
        S[0] = S[0] - v * dPi;
        // S[a] -= dPi, with dPi / pow(t,2) along eta (SourceTerms.cpp:345,380,415)
        const double dPa = dir == 2 ? div(dPi, rt2pow) : dPi;
#pragma unroll
        for (int mu = 1; mu < 4; ++mu)
            if (mu == dir + 1) S[mu] = S[mu] - dPa;
    }

    // Per-variable fluxes of both faces, then r.
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        const Line l = make_line(P.q[n], sm2, sm1, s, sp1, sp2);
        const double Hf = face_var_flux(Rf, Lf, af, right_fwd(l), left_fwd(l), div);
        const double Hb = face_var_flux(Rb, Lb, ab, right_bwd(l), left_bwd(l), div);
        double r = -Hf;
        r += Hb;
        r = div(r, rd);
        if (n < 4) r += S[n];
        r *= G.dt;
        sink(n, r);
        __builtin_amdgcn_sched_barrier(0);  // keep one variable's loads live at a time
    }
}

// Divisions by the cell's u^tau, by t, t^2, t^3, 2t, pow(t,2), the lattice
// spacings, dt, 3 and 5 go through correctly rounded reciprocals (divr),
// which return the division operator's value.  Each variable's result
// q + dt*S goes to sink(n, value) as soon as its S is known, with q read
// again from memory there, so the fifteen inputs are not held in registers
// across the whole evaluation.
template <int NC, bool TABLE, class DV, class GEO, class SINK>
__device__ __forceinline__ void source_init(const StageParams &P, const GEO &G, const EosDev &E, long s, int i, int j, int k,
                                            double t, const Recip &rt2pow, const Recip &rut, double ecell,
                                            SINK &&sink, DV &div) {
    const Recip rdx = make_recip(G.dx), rdy = make_recip(G.dy), rdz = make_recip(G.dz);
    const Recip rt = make_recip(t), rdt = make_recip(G.dt);
    const Recip r3{3.0, 1.0 / 3.0, true}, r5{5.0, 1.0 / 5.0, true};
    auto put = [&](int n, double Sn) { sink(n, P.q[n][s] + G.dt * Sn); };
    const double ttt = P.q[0][s], ttx = P.q[1][s], tty = P.q[2][s], ttn = P.q[3][s];
    const double pitt = P.q[4][s], pitx = P.q[5][s], pity = P.q[6][s], pitn = P.q[7][s], pixx = P.q[8][s],
                 pixy = P.q[9][s], pixn = P.q[10][s], piyy = P.q[11][s], piyn = P.q[12][s], pinn = P.q[13][s];
    const double Pi = NC == 15 ? P.q[NC - 1][s] : 0.0;
    const double *__restrict__ utv = P.u[0];
    const double *__restrict__ uxv = P.u[1];
    const double *__restrict__ uyv = P.u[2];
    const double *__restrict__ unv = P.u[3];
    const double *__restrict__ pv = P.p;
    const double p = pv[s];
    const double ut = utv[s], ux = uxv[s], uy = uyv[s], un = unv[s];
    const double utp = P.up[0][s], uxp = P.up[1][s], uyp = P.up[2][s], unp = P.up[3][s];
    const long strY = G.nx, strZ = (long)G.nx * G.ny;
    const long sl = i > 0 ? s - 1 : s, sr = i < G.nx - 1 ? s + 1 : s;
    const long sb = j > 0 ? s - strY : s, stp = j < G.ny - 1 ? s + strY : s;
    const long sk = k > 0 ? s - strZ : s, sf = k < G.nz - 1 ? s + strZ : s;
    const double dxut = div(approx_deriv(utv[sl], ut, utv[sr]), rdx);
    const double dxux = div(approx_deriv(uxv[sl], ux, uxv[sr]), rdx);
    const double dxuy = div(approx_deriv(uyv[sl], uy, uyv[sr]), rdx);
    const double dxun = div(approx_deriv(unv[sl], un, unv[sr]), rdx);
    const double dyut = div(approx_deriv(utv[sb], ut, utv[stp]), rdy);
    const double dyux = div(approx_deriv(uxv[sb], ux, uxv[stp]), rdy);
    const double dyuy = div(approx_deriv(uyv[sb], uy, uyv[stp]), rdy);
    const double dyun = div(approx_deriv(unv[sb], un, unv[stp]), rdy);
    const double dnut = div(approx_deriv(utv[sk], ut, utv[sf]), rdz);
    const double dnux = div(approx_deriv(uxv[sk], ux, uxv[sf]), rdz);
    const double dnuy = div(approx_deriv(uyv[sk], uy, uyv[sf]), rdz);
    const double dnun = div(approx_deriv(unv[sk], un, unv[sf]), rdz);
    const double dxp = div(approx_deriv(pv[sl], p, pv[sr]), rdx);
    const double dyp = div(approx_deriv(pv[sb], p, pv[stp]), rdy);
    const double dnp = div(approx_deriv(pv[sk], p, pv[sf]), rdz);

    double S[NC];
    const double pP = p + Pi;
    const double tnn = (ecell + pP) * un * un + div(div(pP, rt), rt) + pinn;  // Tnn(e, p+Pi, un, pinn, t)
    const double vx = div(ux, rut), vy = div(uy, rut), vn = div(un, rut);
    const double dxvx = div(dxux - vx * dxut, rut);
    const double dyvy = div(dyuy - vy * dyut, rut);
    const double dnvn = div(dnun - vn * dnut, rut);
    const double dkvk = dxvx + dyvy + dnvn;
    const double src0 = 0.0;  // dynamical sources are zero outside IC type 12
    S[0] = src0 - (div(ttt, rt) + t * tnn) + dkvk * (pitt - p - Pi) - vx * dxp - vy * dyp - vn * dnp;
    S[1] = src0 - div(ttx, rt) - dxp + dkvk * pitx;
    S[2] = src0 - div(tty, rt) - dyp + dkvk * pity;
    S[3] = src0 - div(3 * ttn, rt) - div(dnp, rt2pow) + dkvk * pitn;
#pragma unroll
    for (int n = 0; n < 4; ++n) put(n, S[n]);
    __builtin_amdgcn_sched_barrier(0);

    const double e = ecell;
    const double T = eos_temperature<TABLE>(E, e, div);
    const double d_etabar = etabar_of_T(G, T);
    const double taupiInv = div(T, r5) / d_etabar;
    const double beta_pi = div(e + p, r5);
    const double cs2 = eos_cs2<TABLE>(E, e, div);
    const double a = 1.0 / 3.0 - cs2;
    const double a2 = a * a;
    double beta_Pi = 15 * a2 * (e + p);
    const double lambda_Pipi = div(8 * a, r5);
    double tauPiInv = 0.0;
    if constexpr (NC == 15) {
        // bulkViscosityToEntropyDensity (SourceTerms.cpp:40-48); with a == 0 the
        // quotient is +0 for any positive finite zeta/s, so it is skipped.
        if (a2 != 0.0) {
            tauPiInv = 15 * a2 * T / zeta_over_s(G, T);
            // Bulk relaxation no faster than 1/dt (tau_Pi >= dt), with beta_Pi
            // scaled alike so that the Navier-Stokes value -zeta theta stays:
            // the explicit RK2 step is unstable for dt/tau_Pi > 2, which a
            // small zeta/s reaches.  Never active where c_s^2 = 1/3 (a = 0).
            const double cap = div(1.0, rdt);
            if (tauPiInv > cap) {
                beta_Pi = beta_Pi * (cap / tauPiInv);
                tauPiInv = cap;
            }
        }
    }
    const double ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;
    const Recip rt2 = make_recip(t2), rt3 = make_recip(t3), r2t = make_recip(2.0 * t);
        const double dtut = div(ut - utp, rdt);
    const double dtux = div(ux - uxp, rdt);
    const double dtuy = div(uy - uyp, rdt);
    const double dtun = div(un - unp, rdt);
    const double Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
    const double DuxUpper = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const double Dux = -DuxUpper;
    const double DuyUpper = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const double Duy = -DuyUpper;
    const double DunUpper = ut * dtun + ux * dxun + uy * dyun + un * dnun + div(2 * ut * un, rt);
    const double Dun = -t2 * DunUpper;
    const double dut = Dut - t * un * un;
    const double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
    const double theta = div(ut, rt) + dtut + dxux + dyuy + dnun;
    const double stt = -t * ut * un2 + (dtut - ut * dut) + div((ut2 - 1) * theta, r3);
    const double stx = -(t * un2 * ux) / 2.0 + (dtux - dxut) / 2.0 - (ux * dut + ut * dux) / 2.0 + div(ut * ux * theta, r3);
    const double sty = -(t * un2 * uy) / 2.0 + (dtuy - dyut) / 2.0 - (uy * dut + ut * duy) / 2.0 + div(ut * uy * theta, r3);
    const double stn = div(-un * (2.0 * ut2 + t2 * un2), r2t) + (dtun - div(dnut, rt2)) / 2.0 -
                       (un * dut + ut * dun) / 2.0 + div(ut * un * theta, r3);
    const double sxx = -(dxux + ux * dux) + div((1.0 + ux * ux) * theta, r3);
    const double sxy = -(dxuy + dyux) / 2.0 - (uy * dux + ux * duy) / 2.0 + div(ux * uy * theta, r3);
    const double sxn = div(-ut * ux * un, rt) - (dxun + div(dnux, rt2)) / 2.0 - (un * dux + ux * dun) / 2.0 + div(ux * un * theta, r3);
    const double syy = -(dyuy + uy * duy) + div((1.0 + uy * uy) * theta, r3);
    const double syn = div(-ut * uy * un, rt) - (dyun + div(dnuy, rt2)) / 2.0 - (un * duy + uy * dun) / 2.0 + div(uy * un * theta, r3);
    const double snn = div(-ut * (1.0 + 2.0 * t2 * un2), rt3) - div(dnun, rt2) - un * dun + div((rt2.r + un2) * theta, r3);
    const double wtx = (-dtux - dxut + ut * dux - ux * dut - t * un2 * ux) / 2.0;
    const double wty = (-dtuy - dyut + ut * duy - uy * dut - t * un2 * uy) / 2.0;
    const double wtn = (-t2 * dtun - 2.0 * t * un - dnut - ut * Dun - t2 * un * dut - t3 * un * un2) / 2.0;
    const double wxy = (dxuy - dyux + ux * duy - uy * dux) / 2.0;
    const double wxn = (t2 * dxun - dnux - ux * Dun - t2 * un * dux) / 2.0;
    const double wyn = (t2 * dyun - dnuy - uy * Dun - t2 * un * duy) / 2.0;
    const double wxt = wtx, wyt = wty, wnt = div(wtn, rt2), wyx = -wxy, wnx = div(-wxn, rt2), wny = div(-wyn, rt2);
    const double ux2 = ux * ux, uy2 = uy * uy;
    const double ps = pitt * stt - 2.0 * pitx * stx - 2.0 * pity * sty + pixx * sxx + 2.0 * pixy * sxy + piyy * syy -
                      2.0 * pitn * stn * t2 + 2.0 * pixn * sxn * t2 + 2.0 * piyn * syn * t2 + pinn * snn * t2 * t2;
    // One component of pi^{mu nu} at a time: its I1..I4 terms, I and S, with
    // a scheduling barrier after each so the ten components' terms are not all
    // live at once (the same operations in the same order as written out in
    // the reference, only computed component by component).
    {
        const double I1tt = 2.0 * ut * (pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun);
        const double I2tt = theta * pitt;
        const double I3tt = 2.0 * (pitx * wtx + pity * wty + pitn * wtn);
        const double I4tt = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - div((1.0 - ut2) * ps, r3);
        const double Itt = I1tt + delta_pipi * I2tt - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
        S[4] = div(2.0 * beta_pi * stt - pitt * taupiInv - Itt - 2.0 * un * t * pitn, rut) + pitt * dkvk;
        put(4, S[4]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy +
                            (pitn * ux + pixn * ut) * Dun;
        const double I2tx = theta * pitx;
        const double I3tx = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
        const double I4tx = (pitt * stx + pitx * stt) / 2.0 - (pitx * sxx + pixx * stx) / 2.0 -
                            (pity * sxy + pixy * sty) / 2.0 - t2 * (pitn * sxn + pixn * stn) / 2.0 + div((ut * ux) * ps, r3);
        const double Itx = I1tx + delta_pipi * I2tx - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
        S[5] = div(2.0 * beta_pi * stx - pitx * taupiInv - Itx - un * t * pixn, rut) + pitx * dkvk;
        put(5, S[5]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy +
                            (pitn * uy + piyn * ut) * Dun;
        const double I2ty = theta * pity;
        const double I3ty = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
        const double I4ty = (pitt * sty + pity * stt) / 2.0 - (pitx * sxy + pixy * stx) / 2.0 -
                            (pity * syy + piyy * sty) / 2.0 - t2 * (pitn * syn + piyn * stn) / 2.0 + div((ut * uy) * ps, r3);
        const double Ity = I1ty + delta_pipi * I2ty - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
        S[6] = div(2.0 * beta_pi * sty - pity * taupiInv - Ity - un * t * piyn, rut) + pity * dkvk;
        put(6, S[6]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux + (pity * un + piyn * ut) * Duy +
                            (pitn * un + pinn * ut) * Dun;
        const double I2tn = theta * pitn;
        const double I3tn = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
        const double I4tn = (pitt * stn + pitn * stt) / 2.0 - (pitx * sxn + pixn * stx) / 2.0 -
                            (pity * syn + piyn * sty) / 2.0 - t2 * (pitn * snn + pinn * stn) / 2.0 + div((ut * un) * ps, r3);
        const double Itn = I1tn + delta_pipi * I2tn - I3tn + tau_pipi * I4tn - lambda_piPi * Pi * stn;
        S[7] = div(2.0 * beta_pi * stn - pitn * taupiInv - Itn - un * t * pinn - div(ut * pitn + un * pitt, rt), rut) +
               pitn * dkvk;
        put(7, S[7]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1xx = 2.0 * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
        const double I2xx = theta * pixx;
        const double I3xx = 2.0 * (pitx * wxt + pixy * wxy + pixn * wxn);
        const double I4xx = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + div((1.0 + ux2) * ps, r3);
        const double Ixx = I1xx + delta_pipi * I2xx - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
        S[8] = div(2.0 * beta_pi * sxx - pixx * taupiInv - Ixx, rut) + pixx * dkvk;
        put(8, S[8]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy +
                            (pixn * uy + piyn * ux) * Dun;
        const double I2xy = theta * pixy;
        const double I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
        const double I4xy = (pitx * sty + pity * stx) / 2.0 - (pixx * sxy + pixy * sxx) / 2.0 -
                            (pixy * syy + piyy * sxy) / 2.0 - t2 * (pixn * syn + piyn * sxn) / 2.0 + div((ux * uy) * ps, r3);
        const double Ixy = I1xy + delta_pipi * I2xy - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
        S[9] = div(2.0 * beta_pi * sxy - pixy * taupiInv - Ixy, rut) + pixy * dkvk;
        put(9, S[9]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux + (pixy * un + piyn * ux) * Duy +
                            (pixn * un + pinn * ux) * Dun;
        const double I2xn = theta * pixn;
        const double I3xn = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
        const double I4xn = (pitx * stn + pitn * stx) / 2.0 - (pixx * sxn + pixn * sxx) / 2.0 -
                            (pixy * syn + piyn * sxy) / 2.0 - t2 * (pixn * snn + pinn * sxn) / 2.0 + div((ux * un) * ps, r3);
        const double Ixn = I1xn + delta_pipi * I2xn - I3xn + tau_pipi * I4xn - lambda_piPi * Pi * sxn;
        S[10] = div(2.0 * beta_pi * sxn - pixn * taupiInv - Ixn - div(ut * pixn + un * pitx, rt), rut) + pixn * dkvk;
        put(10, S[10]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1yy = 2.0 * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
        const double I2yy = theta * piyy;
        const double I3yy = 2.0 * (pity * wyt + pixy * wyx + piyn * wyn);
        const double I4yy = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + div((1.0 + uy2) * ps, r3);
        const double Iyy = I1yy + delta_pipi * I2yy - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
        S[11] = div(2.0 * beta_pi * syy - piyy * taupiInv - Iyy, rut) + piyy * dkvk;
        put(11, S[11]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux + (piyy * un + piyn * uy) * Duy +
                            (piyn * un + pinn * uy) * Dun;
        const double I2yn = theta * piyn;
        const double I3yn = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
        const double I4yn = (pity * stn + pitn * sty) / 2.0 - (pixy * sxn + pixn * sxy) / 2.0 -
                            (piyy * syn + piyn * syy) / 2.0 - t2 * (piyn * snn + pinn * syn) / 2.0 + div((uy * un) * ps, r3);
        const double Iyn = I1yn + delta_pipi * I2yn - I3yn + tau_pipi * I4yn - lambda_piPi * Pi * syn;
        S[12] = div(2.0 * beta_pi * syn - piyn * taupiInv - Iyn - div(ut * piyn + un * pity, rt), rut) + piyn * dkvk;
        put(12, S[12]);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        const double I1nn = 2.0 * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun);
        const double I2nn = theta * pinn;
        const double I3nn = 2.0 * (pitn * wnt + pixn * wnx + piyn * wny);
        const double I4nn = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + div((rt2.r + un2) * ps, r3);
        const double Inn = I1nn + delta_pipi * I2nn - I3nn + tau_pipi * I4nn - lambda_piPi * Pi * snn;
        S[13] = div(2.0 * beta_pi * snn - pinn * taupiInv - Inn - div(2 * (ut * pinn + un * pitn), rt), rut) + pinn * dkvk;
        put(13, S[13]);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (NC == 15) {
        const double dPi = -beta_Pi * theta - Pi * tauPiInv - delta_PiPi * Pi * theta + lambda_Pipi * ps;
        S[14] = div(dPi, rut) + Pi * dkvk;
        put(14, S[14]);
    } else {
        (void)beta_Pi; (void)lambda_Pipi; (void)tauPiInv;
    }
}

// REG_SCHEME_1, rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:689-860, one cell.
template <int NC, class GEO>
__device__ __forceinline__ double regulate(const GEO &G, double t, double (&acc)[NC], const Prim &w,
                                           int &nan, double &fb_out) {
    const double pitt = acc[4], pitx = acc[5], pity = acc[6], pitn = acc[7], pixx = acc[8], pixy = acc[9],
                 pixn = acc[10], piyy = acc[11], piyn = acc[12], pinn = acc[13];
    const double Pi = NC == 15 ? acc[NC - 1] : 0.0;
    const double ut = w.ut, ux = w.ux, uy = w.uy, un = w.un;
    const double t2 = t * t;
    double facShear = 1.0, facBulk = 1.0;
    if (!(w.e > G.e_reg)) {  // the factors are overridden to 1 above e_reg (:835)
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
        facShear = glibc_tanh(rho) / rho;
        if (__builtin_fabs(rho) < 1.e-7) facShear = 1.0;
        if constexpr (NC == 15) {
            const double rhoBulk = __builtin_fabs(Pi) / w.p;
            facBulk = glibc_tanh(rhoBulk) / rhoBulk;
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
    fb_out = facBulk;
    return facShear;
}

template <int NC>
struct CellOut {
    double acc[NC];
    Prim w;
    double fs, fb;
};

template <int NC>
__device__ __forceinline__ void store_cell(const StageParams &P, long s, const CellOut<NC> &o, int nan) {
#pragma unroll
    for (int n = 0; n < NC; ++n) P.qo[n][s] = o.acc[n];
    P.eo[s] = o.w.e;
    P.po[s] = o.w.p;
    P.uo[0][s] = o.w.ut;
    P.uo[1][s] = o.w.ux;
    P.uo[2][s] = o.w.uy;
    P.uo[3][s] = o.w.un;
    if (P.regfac) P.regfac[s] = o.fs;
    if (NC == 15 && P.regfacb) P.regfacb[s] = o.fb;
    if (nan != NAN_NONE) report_nan(P.err, nan, s);
}

// Wave-specialised stage kernel.  A block owns 64 consecutive cells and has
// one wave per role: wave 0 evaluates acc = q + dt*S (the source pass), waves
// 1..3 the X, Y, eta contributions r_d of the same 64 cells.  The partial
// results meet in LDS and wave 0 sums them in the reference's order
// ((acc + r_x) + r_y) + r_eta, applies the convex combination of stage 2,
// recovers (e, p, u), regulates and stores.
constexpr int CELLS_PER_BLOCK = 64;

// Where an event's cells start in the field arrays, and the times of the
// current step: a single-event launch reads them from its parameters, a batch
// from the event's descriptor.
__device__ __forceinline__ long cell_base(const StageParams &) { return 0; }
__device__ __forceinline__ long cell_base(const EvDesc &G) { return G.off; }
__device__ __forceinline__ StepTimes step_times(const StageParams &P, const StageParams &) {
    return P.tab ? P.tab[*P.base + P.off] : P.st;
}
__device__ __forceinline__ StepTimes step_times(const StageParams &P, const EvDesc &G) {
    return G.tab[G.base + P.off];
}

template <int NC, bool TABLE, bool DIM3, class GEO>
__device__ __forceinline__ void stage_body(const StageParams &P, const EosDev &E, const GEO &G, long blk) {
    constexpr int ROLES = DIM3 ? 4 : 3;
    __shared__ double part[ROLES][NC][CELLS_PER_BLOCK];
    __shared__ int flags[ROLES][CELLS_PER_BLOCK];  // NaN site of each role
    const int role = __builtin_amdgcn_readfirstlane(threadIdx.x / CELLS_PER_BLOCK);
    const int lane = threadIdx.x % CELLS_PER_BLOCK;
    const long N = (long)G.nx * G.ny * G.nz;
    const long sl = blk * CELLS_PER_BLOCK + lane;
    const bool valid = sl >= 0 && sl < N && role < ROLES;
    const long sc = valid ? sl : 0;
    const long s = cell_base(G) + sc;  // index into the (pooled) field arrays
    const int i = (int)(sc % G.nx);
    const long rest = sc / G.nx;
    const int j = (int)(rest % G.ny);
    const int k = (int)(rest / G.ny);
    const StepTimes st = step_times(P, G);
    const double t = P.stage ? st.t2 : st.t1;
    const double t2pow = P.stage ? st.t2sq : st.t1sq;
    const double t_new = st.t2;
    if (valid) {
        DivFast div;
        int nan = NAN_NONE;
        const double ecell = P.e[s];
        const Recip rut = make_recip(P.u[0][s]);
        const Recip rt2pow = make_recip(t2pow);
        if (role == 0) {
            source_init<NC, TABLE>(P, G, E, s, i, j, k, t, rt2pow, rut, ecell,
                                   [&](int n, double a) { part[0][n][lane] = a; }, div);
        } else {
            auto put = [&](int n, double r) { part[role][n][lane] = r; };
            const int dir = role - 1;
            const int coord = dir == 0 ? i : (dir == 1 ? j : k);
            const int ncoord = dir == 0 ? G.nx : (dir == 1 ? G.ny : G.nz);
            const long stride = dir == 0 ? 1 : (dir == 1 ? (long)G.nx : (long)G.nx * G.ny);
            dir_contrib<NC, TABLE>(P, G, E, s, dir, coord, ncoord, stride, ecell, t, rt2pow, rut, P.u[1 + dir][s], put,
                                   nan, div);
        }
        flags[role][lane] = nan;
    }
    __syncthreads();
    if (role != 0 || !valid) return;
    // NaN site in the reference's detection order: X, Y, eta fluxes, then the cell
    int nan = NAN_NONE;
    for (int r = 1; r < ROLES && nan == NAN_NONE; ++r) nan = flags[r][lane];
    CellOut<NC> o;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        double a = part[0][n][lane];
        a += part[1][n][lane];
        a += part[2][n][lane];
        if constexpr (DIM3) a += part[3][n][lane];
        else a += 0.0;  // one rapidity cell: the eta contribution is exactly +0
        o.acc[n] = a;
    }
    if (P.stage == 1) {
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            o.acc[n] += P.qn[n][s];
            o.acc[n] /= 2;
        }
    }
    DivFast div;
    int nan_cell = NAN_NONE;
    o.w = recover<TABLE>(E, t_new, o.acc[0], o.acc[1], o.acc[2], o.acc[3], o.acc[4], o.acc[5], o.acc[6], o.acc[7],
                         NC == 15 ? o.acc[NC - 1] : 0.0, P.e[s], nan_cell, div);
    o.fs = regulate<NC>(G, t_new, o.acc, o.w, nan_cell, o.fb);
    store_cell<NC>(P, s, o, nan != NAN_NONE ? nan : nan_cell);
}

template <int NC, bool TABLE, bool DIM3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void stage_kernel(const StageParams P,
                                                                                             const EosDev E) {
    stage_body<NC, TABLE, DIM3>(P, E, P, xcd_block(blockIdx.x, gridDim.x));
}

// Many events in one launch: block b belongs to event blk_ev[b] and is that
// event's (b - blk_first[b])-th block of 64 cells.  The per-event fields of P
// are ignored; its field pointers address the pooled state.
template <int NC, bool TABLE, bool DIM3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void stage_kernel_batch(
    const StageParams P, const EosDev E, const EvDesc *__restrict__ ev, const int *__restrict__ blk_ev,
    const int *__restrict__ blk_first) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int e = __builtin_amdgcn_readfirstlane(blk_ev[b]);
    const int b0 = __builtin_amdgcn_readfirstlane(blk_first[b]);
    const EvDesc &G = ev[e];
    if (!G.active) return;  // a removed event whose blocks are still in the table (block-uniform)
    stage_body<NC, TABLE, DIM3>(P, E, G, (long)b - b0);
}

}  // namespace vh
