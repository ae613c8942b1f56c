// Host-side initial conditions of the product (C ABI in include/cpuvh_amd.h).
//
// * Seeded Monte-Carlo Glauber energy density: Woods-Saxon nucleons, wounded
//   nucleons from a black-disk sigma_NN cut, one Gaussian per participant, and
//   the reference's longitudinal profile (a plateau of full width
//   rapidityMean with Gaussian tails of variance rapidityVariance,
//   rhic/src/InitialConditions.cpp:1046-1064) plus the reference's 1e-3
//   floor (:1134).  The reference's own generator uses glibc rand() with a
//   fixed seed and hard-coded nuclear parameters
//   (rhic/src/MonteCarloGlauberModel.cpp:18-111); this one takes every
//   parameter and a 64-bit seed so events are reproducible and independent.
// * Type-11 file reader/writer: input/Tmunu.dat, 20 columns per line read
//   through float (rhic/src/InitialConditions.cpp:41-105).
// * The reference's own types 2 and 4 (optical and Monte-Carlo Glauber), its
//   analytic test states (types 3, 5-9) and its other file formats (types 1
//   and 10).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cpuvh_amd.h"

namespace {

// SplitMix64 (Steele, Lea & Flood 2014): a 64-bit counter passed through a
// fixed bijective mixer.
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

void sample_nucleus(SplitMix64 &rng, int A, double R, double a, double *x, double *y) {
    const double rmax = R + 12 * a;
    double M = 0;  // max of r^2 rho(r) on [0, rmax]
    for (int i = 0; i <= 4000; ++i) {
        const double r = rmax * i / 4000.0;
        M = std::fmax(M, r * r / (1 + std::exp((r - R) / a)));
    }
    M *= 1.01;
    int m = 0;
    while (m < A) {
        const double r = rmax * rng.uniform();
        const double u = rng.uniform();
        if (u * M < r * r / (1 + std::exp((r - R) / a))) {
            const double ct = 2 * rng.uniform() - 1, phi = 2 * M_PI * rng.uniform();
            const double st = std::sqrt(1 - ct * ct);
            x[m] = r * st * std::cos(phi);
            y[m] = r * st * std::sin(phi);
            ++m;
        }
    }
}

}  // namespace

extern "C" vh_status vh_ic_mc_glauber(const vh_glauber_params *gp, int nx, int ny, int nz, double dx, double dy,
                                      double dz, double *e_out, vh_glauber_info *info) {
    if (!gp || !e_out || nx < 1 || ny < 1 || nz < 1 || gp->A < 1 || !(gp->sigma_nn_fm2 > 0) || !(gp->width > 0))
        return VH_ERR_ARG;
    SplitMix64 rng{gp->seed};
    const int A = gp->A;
    double b = gp->b;
    if (b < 0) b = gp->b_max * rng.uniform();
    std::vector<double> x1(A), y1(A), x2(A), y2(A);
    sample_nucleus(rng, A, gp->R, gp->a, x1.data(), y1.data());
    sample_nucleus(rng, A, gp->R, gp->a, x2.data(), y2.data());
    for (int i = 0; i < A; ++i) {
        x1[i] -= b / 2;
        x2[i] += b / 2;
    }
    const double d2 = gp->sigma_nn_fm2 / M_PI;
    std::vector<int> w1(A, 0), w2(A, 0);
    int ncoll = 0;
    for (int i = 0; i < A; ++i)
        for (int j = 0; j < A; ++j) {
            const double ddx = x1[i] - x2[j], ddy = y1[i] - y2[j];
            if (ddx * ddx + ddy * ddy < d2) {
                ++w1[i];
                ++w2[j];
                ++ncoll;
            }
        }
    std::vector<double> px, py;
    for (int i = 0; i < A; ++i) {
        if (w1[i]) { px.push_back(x1[i]); py.push_back(y1[i]); }
        if (w2[i]) { px.push_back(x2[i]); py.push_back(y2[i]); }
    }
    // transverse profile: normalised Gaussians of width w around each participant
    std::vector<double> eT((size_t)nx * ny, 0.0);
    const double w = gp->width, norm = 1.0 / (2 * M_PI * w * w), cut = 6 * w;
    for (size_t m = 0; m < px.size(); ++m) {
        for (int j = 0; j < ny; ++j) {
            const double y = (j - (ny - 1) / 2.0) * dy;
            if (std::fabs(y - py[m]) > cut) continue;
            for (int i = 0; i < nx; ++i) {
                const double x = (i - (nx - 1) / 2.0) * dx;
                if (std::fabs(x - px[m]) > cut) continue;
                const double r2 = (x - px[m]) * (x - px[m]) + (y - py[m]) * (y - py[m]);
                eT[(size_t)i + (size_t)nx * j] += norm * std::exp(-r2 / (2 * w * w));
            }
        }
    }
    // longitudinal profile, InitialConditions.cpp:1046-1064 (integer (nz-1)/2 as there)
    std::vector<double> eL(nz);
    for (int k = 0; k < nz; ++k) {
        const double eta = (k - (nz - 1) / 2) * dz;
        const double es = std::fabs(eta) - gp->eta_flat / 2;
        const double arg = -es * es / gp->eta_variance / 2 * (es < 0 ? 0.0 : 1.0);
        eL[k] = std::exp(arg);
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                e_out[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] =
                    gp->e0 * eT[(size_t)i + (size_t)nx * j] * eL[k] + gp->e_floor;
    if (info) {
        info->b = b;
        info->n_part = (int)px.size();
        info->n_coll = ncoll;
    }
    return VH_OK;
}

// The reference's own type-4 generator, reproduced value for value: the C
// library's rand() seeded with the reference's fixed seed, Woods-Saxon
// nucleons with its hard-coded radius, diffuseness and acceptance bound,
// wounded nucleons from the black-disk cut sqrt(0.1 sigma_NN[mb] / pi),
// participants listed in the reference's interleaved order (projectile i,
// then target i), unnormalised Gaussians of width 0.96 fm summed in that
// order, times e0 and the longitudinal profile, plus the 1e-3 floor
// (rhic/src/MonteCarloGlauberModel.cpp:18-111, rhic/src/GlauberModel.cpp:19-24,
// rhic/src/InitialConditions.cpp:1046-1064, :1107-1140).  Products are
// written as the reference's g++ -O3 build evaluates them (pow(d, 2) is d*d).
namespace {

double classic_uniform() { return (double)std::rand() / (double)RAND_MAX; }

void classic_nucleus(int A, double *x, double *y) {
    const double n0 = 0.17, thick = 0.54;  // GlauberModel.cpp:20-22
    const double radius = 1.12 * std::pow((double)A, 1. / 3.) - 0.86 * std::pow((double)A, -1. / 3.);
    const double bound = 1.01 * 4.5310551374155095, rmax = 20;  // MonteCarloGlauberModel.cpp:19-20
    int m = 0;
    while (m < A) {
        const double r = rmax * (double)std::rand() / (double)RAND_MAX;
        const double u = classic_uniform();
        if (u < r * r * (n0 / (1 + std::exp((r - radius) / thick))) / bound) {
            const double v = 2 * (classic_uniform() - 0.5);
            const double f = 2 * M_PI * (double)std::rand() / (double)RAND_MAX;
            x[m] = r * std::sqrt(1 - v * v) * std::cos(f);
            y[m] = r * std::sqrt(1 - v * v) * std::sin(f);
            ++m;
        }
    }
}

}  // namespace

extern "C" vh_status vh_ic_mc_glauber_classic(int A, double b, double sigma_nn_mb, double e0, double eta_flat,
                                              double eta_variance, int nx, int ny, int nz, double dx, double dy,
                                              double dz, double *e_out, int *n_wounded) {
    if (!e_out || nx < 1 || ny < 1 || nz < 1 || A < 1) return VH_ERR_ARG;
    std::srand(1328398221u);  // MonteCarloGlauberModel.cpp:92
    std::vector<double> x1(A), y1(A), x2(A), y2(A);
    classic_nucleus(A, x1.data(), y1.data());
    classic_nucleus(A, x2.data(), y2.data());
    for (int i = 0; i < A; ++i) {
        x1[i] -= b / 2;
        x2[i] += b / 2;
    }
    const double dn = std::sqrt(0.1 * sigma_nn_mb / M_PI);
    std::vector<int> hit1(A, 0), hit2(A, 0);
    for (int i = 0; i < A; ++i)
        for (int j = 0; j < A; ++j) {
            const double ddx = x1[i] - x2[j], ddy = y1[i] - y2[j];
            double d = ddx * ddx;
            d += ddy * ddy;
            if (std::sqrt(d) < dn) {
                ++hit1[i];
                ++hit2[j];
            }
        }
    std::vector<double> px, py;
    for (int i = 0; i < A; ++i) {
        if (hit1[i] > 0) { px.push_back(x1[i]); py.push_back(y1[i]); }
        if (hit2[i] > 0) { px.push_back(x2[i]); py.push_back(y2[i]); }
    }
    const double sig = 0.96;  // MonteCarloGlauberModel.cpp:90
    std::vector<double> eT((size_t)nx * ny, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (size_t n = 0; n < px.size(); ++n) {
                const double x = (i - ((double)nx - 1.) / 2.) * dx - px[n];
                const double y = (j - ((double)ny - 1.) / 2.) * dy - py[n];
                eT[(size_t)i + (size_t)nx * j] += std::exp(-x * x / 2 / sig / sig - y * y / 2 / sig / sig);
            }
    std::vector<double> eL(nz);
    for (int k = 0; k < nz; ++k) {
        const double eta = (k - (nz - 1) / 2) * dz;
        const double es = std::fabs(eta) - eta_flat / 2;
        eL[k] = std::exp(-es * es / eta_variance / 2 * (es < 0 ? 0.0 : 1.0));
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double eTr = e0 * eT[(size_t)i + (size_t)nx * j];
                e_out[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = (eTr * eL[k]) + 1.e-3;
            }
    if (n_wounded) *n_wounded = (int)px.size();
    return VH_OK;
}

// The reference's IC type 2, the optical Glauber profile
// (rhic/src/InitialConditions.cpp:1069-1105, rhic/src/GlauberModel.cpp:19-108):
// nuclear thickness T(x, y) = integral of the Woods-Saxon density along z
// (n0 = 0.17 fm^-3, R = 1.12 A^(1/3) - 0.86 A^(-1/3), d = 0.54 fm), a mixture
// alpha * binary collisions + (1 - alpha) * wounded nucleons, each
// normalised to 1 at the origin of a central collision, with the reference's
// own expressions (the nucleus at +b/2 and the one at -b/2 use the same
// thickness function; the wounded-nucleon term is written as in
// GlauberModel.cpp:50-59), times e0 and the longitudinal profile, plus the
// 1e-3 floor.  The reference integrates with GSL's QAGS to 1e-8 relative;
// here composite 8-point Gauss-Legendre over z in [0, R + 40 d] (error below
// 1e-13), so profiles agree to the reference's integration tolerance.
namespace {

double ws_density(double r, double A) {
    const double n0 = 0.17, d = 0.54;
    const double R = 1.12 * std::pow(A, 1. / 3.) - 0.86 * std::pow(A, -1. / 3.);
    return n0 / (1 + std::exp((r - R) / d));
}

double thickness(double x, double y, double A) {
    static const double gx[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                 -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                                 0.7966664774136267,  0.9602898564975363};
    static const double gw[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                                 0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
    const double R = 1.12 * std::pow(A, 1. / 3.) - 0.86 * std::pow(A, -1. / 3.);
    const double zmax = R + 40 * 0.54;
    const int panels = 256;
    const double h = zmax / panels, r2 = x * x + y * y;
    double sum = 0;
    for (int k = 0; k < panels; ++k) {
        const double mid = (k + 0.5) * h;
        double ps = 0;
        for (int g = 0; g < 8; ++g) {
            const double z = mid + 0.5 * h * gx[g];
            ps += gw[g] * ws_density(std::sqrt(r2 + z * z), A);
        }
        sum += 0.5 * h * ps;
    }
    return 2 * sum;  // both halves of the z axis
}

double wounded(double TAminus, double TAplus, double TBminus, double TBplus, double A, double B, double snn) {
    (void)TAminus; (void)TBplus;
    const double fm_to_mb = 0.1;
    double part = 1 - snn * TBminus * fm_to_mb / B;
    double res = TAplus * (1 - std::pow(1 - part, B));
    part = 1 - snn * TAplus * fm_to_mb / A;
    res += TBminus * (1 - std::pow(1 - part, A));
    return res;
}

}  // namespace

extern "C" vh_status vh_ic_optical_glauber(int A, double b, double sigma_nn_mb, double alpha, double e0,
                                           double eta_flat, double eta_variance, int nx, int ny, int nz, double dx,
                                           double dy, double dz, double *e_out) {
    if (!e_out || nx < 1 || ny < 1 || nz < 1 || A < 1) return VH_ERR_ARG;
    const double Ad = A;
    const double T0 = thickness(0, 0, Ad);
    const double nbc_norm = 1. / (sigma_nn_mb * T0 * T0);
    const double wn_norm = 1. / wounded(T0, T0, T0, T0, Ad, Ad, sigma_nn_mb);
    std::vector<double> eT((size_t)nx * ny);
    // transverse points are independent: columns i spread over host threads
    auto columns = [&](int first, int stride) {
        for (int i = first; i < nx; i += stride) {
            const double x = (i - (nx - 1) / 2.) * dx;
            for (int j = 0; j < ny; ++j) {
                const double y = (j - (ny - 1) / 2.) * dy;
                const double TAminus = thickness(x - b / 2, y, Ad), TAplus = thickness(x + b / 2, y, Ad);
                double ed = nbc_norm * (sigma_nn_mb * TAplus * TAminus);
                ed *= alpha;
                ed += (1 - alpha) * wn_norm * wounded(TAminus, TAplus, TAminus, TAplus, Ad, Ad, sigma_nn_mb);
                eT[(size_t)i + (size_t)j * nx] = ed;
            }
        }
    };
    const int nthreads = std::max(1, std::min({16, nx, (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(columns, t, nthreads);
    columns(0, nthreads);
    for (std::thread &th : pool) th.join();
    for (int k = 0; k < nz; ++k) {
        const double eta = (k - (nz - 1) / 2) * dz;
        const double es = std::fabs(eta) - eta_flat / 2;
        const double eL = std::exp(-es * es / eta_variance / 2 * (es < 0 ? 0.0 : 1.0));
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double eTr = e0 * eT[(size_t)i + (size_t)j * nx];
                e_out[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = (eTr * eL) + 1.e-3;
            }
    }
    return VH_OK;
}

// rhic/src/InitialConditions.cpp:41-105: values pass through float, lines in
// i-outer, k-inner order; columns x y z e p ut ux uy un pitt..pinn Pi.
extern "C" vh_status vh_ic_read_tmunu(const char *path, int nx, int ny, int nz, double *e, double *p, double *ut,
                                      double *ux, double *uy, double *un, double *pi10, double *Pi) {
    if (!path || !e || !p || !ut || !ux || !uy || !un || !pi10) return VH_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return VH_ERR_ARG;
    const size_t N = (size_t)nx * ny * nz;
    float v[20];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (std::fscanf(f, "%f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f\n", &v[0], &v[1],
                                &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12],
                                &v[13], &v[14], &v[15], &v[16], &v[17], &v[18], &v[19]) != 20) {
                    std::fclose(f);
                    return VH_ERR_ARG;
                }
                const size_t s = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                e[s] = v[3]; p[s] = v[4]; ut[s] = v[5]; ux[s] = v[6]; uy[s] = v[7]; un[s] = v[8];
                for (int c = 0; c < 10; ++c) pi10[c * N + s] = v[9 + c];
                if (Pi) Pi[s] = v[19];
            }
    std::fclose(f);
    return VH_OK;
}

extern "C" vh_status vh_ic_write_tmunu(const char *path, int nx, int ny, int nz, double dx, double dy, double dz,
                                       const double *e, const double *p, const double *ut, const double *ux,
                                       const double *uy, const double *un, const double *pi10, const double *Pi) {
    if (!path || !e || !p || !ut || !ux || !uy || !un || !pi10) return VH_ERR_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return VH_ERR_ARG;
    const size_t N = (size_t)nx * ny * nz;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t s = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                std::fprintf(f, "%.9g %.9g %.9g", (i - (nx - 1) / 2.0) * dx, (j - (ny - 1) / 2.0) * dy,
                             (k - (nz - 1) / 2.0) * dz);
                const double vals[7] = {e[s], p[s], ut[s], ux[s], uy[s], un[s], 0.0};
                for (int c = 0; c < 6; ++c) std::fprintf(f, " %.9g", (double)(float)vals[c]);
                for (int c = 0; c < 10; ++c) std::fprintf(f, " %.9g", (double)(float)pi10[c * N + s]);
                std::fprintf(f, " %.9g\n", Pi ? (double)(float)Pi[s] : 0.0);
            }
    std::fclose(f);
    return VH_OK;
}

// The reference's analytic initial states (rhic/src/InitialConditions.cpp):
// p = e/3 whatever the EoS, pi^{mu nu} and Pi zero, u^mu as written there.
//   3: ideal Gubser flow at tau = 1, q = 1 (:1146-1190)
//   5: Sod shock tube in x (:1259-1298)     9: Sod shock tube across y = x (:1300-1338)
//   6: implosion in a box (:1340-1378)      7: Rayleigh-Taylor instability (:1380-1429)
//   8: Gaussian pulse with a transverse shear flow (:1431-1478)
// At r = 0 the Gubser flow is defined as u^mu = (1, 0, 0, 0); the reference
// evaluates sinh(phi) x / r there and gets NaN on grids with a centre cell.
extern "C" vh_status vh_ic_test_problem(int type, double e0, int nx, int ny, int nz, double dx, double dy,
                                        double *e, double *p, double *ut, double *ux, double *uy, double *un) {
    if (!e || !p || !ut || !ux || !uy || !un || nx < 1 || ny < 1 || nz < 1) return VH_ERR_ARG;
    if (type != 3 && (type < 5 || type > 9)) return VH_ERR_ARG;
    const double Lx = ((nx - 1) / 2.) * dx, Ly = ((ny - 1) / 2.) * dy;
    for (int j = 0; j < ny; ++j) {
        const double y = (j - (ny - 1) / 2.) * dy;
        for (int i = 0; i < nx; ++i) {
            const double x = (i - (nx - 1) / 2.) * dx;
            double ed = 0, vx = 0, vy = 0;
            switch (type) {
            case 3: {
                const double T = 1.9048812623618392 /
                                 std::pow(1 + std::pow(1 - std::pow(x, 2) - std::pow(y, 2), 2) +
                                              2 * (1 + std::pow(x, 2) + std::pow(y, 2)),
                                          0.3333333333333333);
                const double r = std::sqrt(x * x + y * y);
                const double phi = std::atanh(2.0 * r / (1.0 + 1.0 + x * x + y * y));
                ed = e0 * std::pow(T, 4);
                if (r > 0) {
                    vx = std::sinh(phi) * x / r;
                    vy = std::sinh(phi) * y / r;
                }
                break;
            }
            case 5: ed = x > 0 ? 0.00778147 : 0.124503; break;
            case 9: ed = y > x ? 0.00778147 : 0.124503; break;
            case 6: ed = std::sqrt(x * x + y * y) <= 0.15 ? 0.124503 : 0.00778147; break;
            case 7: {
                const double gravity = 0.1, rhoTop = 2.0, rhoBot = 1.0, pert = 0.01;
                const double yloc = 0.5 + pert * std::cos(M_PI * x);
                ed = y > yloc ? rhoTop * gravity * (1 - y) : rhoTop * gravity * (1 - yloc) + rhoBot * gravity * (yloc - y);
                break;
            }
            case 8: {
                const double xc = 0.5, yc = 0.5, beta = 50.0;
                const double pr = 1.0 + 1e-1 * std::exp(-beta * ((x - xc) * (x - xc) + (y - yc) * (y - yc)));
                ed = pr / (1.4 - 1.);
                vy = (1. + std::cos(2 * M_PI * x / Lx)) * (1 + std::cos(2 * M_PI * y / Ly));
                break;
            }
            }
            const double vt = std::sqrt(1 + vx * vx + vy * vy);
            for (int k = 0; k < nz; ++k) {
                const size_t s = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                e[s] = ed;
                p[s] = ed / 3.0;
                ux[s] = vx;
                uy[s] = vy;
                un[s] = 0.0;
                ut[s] = vt;
            }
        }
    }
    return VH_OK;
}

namespace {

// One input/<name>.dat of IC type 10: "x y z value" lines read through float,
// i outer, k inner (rhic/src/InitialConditions.cpp:123-146).
bool read_field_file(const char *dir, const char *name, int nx, int ny, int nz, double *out) {
    std::string path = std::string(dir) + "/input/" + name + ".dat";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    float x, y, z, v;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (std::fscanf(f, "%f %f %f %f\n", &x, &y, &z, &v) != 4) {
                    std::fclose(f);
                    return false;
                }
                out[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = v;
            }
    std::fclose(f);
    return true;
}

}  // namespace

// IC type 10 (rhic/src/InitialConditions.cpp:108-466): input/e.dat, ux, uy,
// un (nz > 1) and pitt..pinn (the eta components only for nz > 1; zero
// otherwise), each "x y z value" per line.  u^tau = sqrt(1 + ux^2 + uy^2 +
// t0^2 un^2): the reference stores the square (:226), which is not a unit
// four-velocity.  The pressure and Pi depend on the EoS and are the
// caller's.  VH_ERR_ARG names no file; the caller checks which exist.
extern "C" vh_status vh_ic_read_fields(const char *dir, int nx, int ny, int nz, double t0, double *e, double *ut,
                                       double *ux, double *uy, double *un, double *pi10) {
    if (!dir || !e || !ut || !ux || !uy || !un || !pi10 || nx < 1 || ny < 1 || nz < 1) return VH_ERR_ARG;
    const size_t N = (size_t)nx * ny * nz;
    std::fill(un, un + N, 0.0);
    std::fill(pi10, pi10 + 10 * N, 0.0);
    if (!read_field_file(dir, "e", nx, ny, nz, e) || !read_field_file(dir, "ux", nx, ny, nz, ux) ||
        !read_field_file(dir, "uy", nx, ny, nz, uy) || (nz > 1 && !read_field_file(dir, "un", nx, ny, nz, un)))
        return VH_ERR_ARG;
    for (size_t s = 0; s < N; ++s) ut[s] = std::sqrt(1.0 + ux[s] * ux[s] + uy[s] * uy[s] + t0 * t0 * un[s] * un[s]);
    static const char *pin[10] = {"pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"};
    for (int c = 0; c < 10; ++c) {
        const bool eta = c == 3 || c == 6 || c == 8;
        if (eta && nz == 1) continue;
        if (!read_field_file(dir, pin[c], nx, ny, nz, pi10 + c * N)) return VH_ERR_ARG;
    }
    return VH_OK;
}

// IC type 1 (rhic/src/InitialConditions.cpp:1196-1257): the Israel-Stewart
// Gubser profile from a file of one line per transverse cell, i outer, j
// inner, columns x y e ux uy pixx piyy pixy pitt pitx pity pinn read as
// double; the same values on every rapidity slice, p = e/3, un and the eta
// components of pi zero.
extern "C" vh_status vh_ic_read_gubser(const char *path, int nx, int ny, int nz, double *e, double *p, double *ut,
                                       double *ux, double *uy, double *un, double *pi10) {
    if (!path || !e || !p || !ut || !ux || !uy || !un || !pi10 || nx < 1 || ny < 1 || nz < 1) return VH_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return VH_ERR_ARG;
    const size_t N = (size_t)nx * ny * nz;
    double v[12];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", &v[0], &v[1], &v[2], &v[3], &v[4],
                            &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11]) != 12) {
                std::fclose(f);
                return VH_ERR_ARG;
            }
            const double pv[10] = {v[8], v[9], v[10], 0.0, v[5], v[7], 0.0, v[6], 0.0, v[11]};
            for (int k = 0; k < nz; ++k) {
                const size_t s = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                e[s] = v[2];
                p[s] = v[2] / 3.0;
                ux[s] = v[3];
                uy[s] = v[4];
                un[s] = 0.0;
                ut[s] = std::sqrt(1.0 + v[3] * v[3] + v[4] * v[4]);
                for (int c = 0; c < 10; ++c) pi10[c * N + s] = pv[c];
            }
        }
    std::fclose(f);
    return VH_OK;
}
