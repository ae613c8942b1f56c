// cpuvh_run — the command-line hydro driver on top of the C ABI.
//
// Usage (the reference's options, rhic/src/CommandLineArguments.cpp:17-23):
//   cpuvh_run --hydro --config <dir> --output <dir>       (short: -h -c -o; -t is accepted)
//   cpuvh_run --config <dir> --print-config              (the parsed parameters, one per line)
//
// Reads the reference's rhic-conf files with the reference's reader semantics
// (rhic/src/LatticeParameters.cpp:68-157, rhic/src/InitialConditionParameters.cpp:75-150,
// rhic/src/HydroParameters.cpp:52-100): values are taken by position, one per
// line, keys ignored, whitespace removed, the text after '=' converted with
// atoi/atof; a missing file leaves its parameters at zero.  Options the
// reference fixes at compile time (bulk viscosity, the EoS) come from an
// optional, keyed <config>/gpu.properties, so the rhic-conf files stay
// unchanged:
//   bulk=0|1                 evolve Pi (the reference's -DPI build)           [0]
//   eos=conformal|table      the EoS (CONFORMAL_EOS / HOTQCDHRG_EOS)          [conformal]
//   eosFile=<path>           table file, 4 x 100000 doubles                   [EoS/hrg_hotqcd_eos_binary.dat]
//   device=<n>               HIP device                                       [0]
//   snapshotEvery=<n>        steps between snapshots (reference: 100)          [100]
//
// The time loop is the reference's run() (rhic/src/HydroPlugin.cpp:74-321),
// in its order: for n = 1 .. nt+1 (nt+1 steps), snapshot when (n-1) % 100 == 0,
// then on every n % coarse == 0 the freezeout cube flags (the previous
// slice starts at zero) and the count of cells above e_F — the run stops once
// that count has been zero three times (the count of zeros is never reset) —
// then one RK2 step at t, then t = t0 + n dt.  Consecutive steps with no host
// work between them are replayed from a HIP graph (vh_steps).
//
// Initial conditions (rhic/src/InitialConditions.cpp:1508-1581): types 0
// (constant energy density, :953-981), 2 (the optical Glauber, :1069-1105),
// 4 (the reference's MC-Glauber, bit for bit), with u = (1,0,0,0) and pi = 0
// or its Navier-Stokes value (:492-646); the analytic test states 3 and 5-9
// (vh_ic_test_problem); the files of types 1
// (<root>/rhic/rhic-trunk/src/test/resources/gubser/viscous/gubserIC.dat),
// 10 (input/e.dat, ux.dat, ...) and 11 (input/Tmunu.dat), under the working
// directory; 13 and 14 from the vectors of the HYDRO class
// (include/cpuvh_hydro.hpp).  Type 12 (dynamical sources) exits as the
// reference does for an undefined type.
//
// Output: the snapshot files of rhic/src/HydroPlugin.cpp:40-72 in the format
// of rhic/src/FileIO.cpp:14-59 (<name>_<t %.3f>.dat, "x y z value" lines, x
// fastest), and surface.dat as freezeout/FreezeOut.cpp:200-361 writes it: for
// every cube (2+1D) or hypercube (3+1D) between the previous and the current
// coarse step that the surface cuts, in the reference's cube order (x, then
// y, then eta), each surface element's position, dsigma_mu = t * normal, and
// u^x u^y u^eta e T P pi^xx pi^xy pi^xeta pi^yy pi^yeta Pi linearly
// interpolated at its centroid — 20 columns, default stream formatting.  The
// GPU flags the cut cubes (vh_fo_flag) and builds their elements and the
// variables at the centroids (vh_fo_surface); the host orders and formats.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <getopt.h>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/cpuvh_amd.h"
#include "../../include/cpuvh_hydro.hpp"
#include "vh_format.hpp"

namespace {

constexpr double HBARC = 0.197326938;  // rhic/src/HydroPlugin.cpp:112

// printf's "%.Nf" and "%g" (6 significant digits, an ostream's default output
// of a double) appended to a string, exactly and without printf's cost
using cpuvh::put_fixed;
using cpuvh::put_g;

// One positional value per line: whitespace removed, the text after the
// first '=' (the whole line when there is none).  When the file has run out,
// std::getline leaves the previous line in place, so the previous value
// repeats — as in the reference's readers.
class PositionalReader {
  public:
    explicit PositionalReader(const std::string &path) : in_(path) {}
    bool ok() const { return in_.is_open(); }
    std::string next() {
        std::getline(in_, line_);
        line_.erase(std::remove_if(line_.begin(), line_.end(), [](unsigned char c) { return std::isspace(c); }),
                    line_.end());
        const size_t eq = line_.find('=');
        line_ = line_.substr(eq == std::string::npos ? 0 : eq + 1);
        return line_;
    }
    int next_int() { return std::atoi(next().c_str()); }
    double next_double() { return std::atof(next().c_str()); }

  private:
    std::ifstream in_;
    std::string line_;
};

struct Lattice {  // rhic/include/LatticeParameters.h:17-34
    int nx = 0, ny = 0, nz = 0, nt = 0;
    double dx = 0, dy = 0, dz = 0, dt = 0;
    int coarse = 0;
};
struct InitCond {  // rhic/include/InitialConditionParameters.h
    int type = 0, n_source_files = 0, A = 0;
    double e0 = 0, sigma_nn_mb = 0, b = 0, f_binary = 0, eta_variance = 0, eta_flat = 0;
    int source_type = 0;
};
struct HydroPar {  // rhic/include/HydroParameters.h:15-23
    double t0 = 0, etabar = 0, T_F_GeV = 0;
    int pi_navier_stokes = 0;
    double T_reg_GeV = 0;
    int do_freezeout = 0;
};
struct Extra {
    int bulk = 0;
    int eos_table = 0;
    std::string eos_file = "EoS/hrg_hotqcd_eos_binary.dat";
    int device = 0;
    int snapshot_every = 100;
};

void read_config(const std::string &dir, Lattice &L, InitCond &I, HydroPar &H, Extra &X) {
    {
        PositionalReader r(dir + "/lattice.properties");
        if (r.ok()) {
            L.nx = r.next_int(); L.ny = r.next_int(); L.nz = r.next_int(); L.nt = r.next_int();
            L.dx = r.next_double(); L.dy = r.next_double(); L.dz = r.next_double(); L.dt = r.next_double();
            L.coarse = r.next_int();
        } else {
            std::fprintf(stderr, "No configuration file found for lattice parameters\n");
        }
    }
    {
        PositionalReader r(dir + "/ic.properties");
        if (r.ok()) {
            I.type = r.next_int(); I.n_source_files = r.next_int(); I.A = r.next_int();
            I.e0 = r.next_double(); I.sigma_nn_mb = r.next_double(); I.b = r.next_double();
            I.f_binary = r.next_double(); I.eta_variance = r.next_double(); I.eta_flat = r.next_double();
            I.source_type = r.next_int();
        } else {
            std::fprintf(stderr, "No configuration file found for initial condition parameters\n");
        }
    }
    {
        PositionalReader r(dir + "/hydro.properties");
        if (r.ok()) {
            H.t0 = r.next_double(); H.etabar = r.next_double(); H.T_F_GeV = r.next_double();
            H.pi_navier_stokes = r.next_int(); H.T_reg_GeV = r.next_double(); H.do_freezeout = r.next_int();
        } else {
            std::fprintf(stderr, "No configuration file found for hydro parameters\n");
        }
    }
    std::ifstream in(dir + "/gpu.properties");
    std::string line;
    while (std::getline(in, line)) {
        line.erase(std::remove_if(line.begin(), line.end(), [](unsigned char c) { return std::isspace(c); }),
                   line.end());
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = line.substr(0, eq), val = line.substr(eq + 1);
        if (key == "bulk") X.bulk = std::atoi(val.c_str());
        else if (key == "eos") X.eos_table = (val == "table" || val == "1");
        else if (key == "eosFile") X.eos_file = val;
        else if (key == "device") X.device = std::atoi(val.c_str());
        else if (key == "snapshotEvery") X.snapshot_every = std::atoi(val.c_str());
        else std::fprintf(stderr, "gpu.properties: unknown key '%s' ignored\n", key.c_str());
    }
}

[[noreturn]] void die(const char *what, vh_status st) {
    std::printf("%s failed: %s\n", what, vh_status_string(st));
    std::exit(-1);
}

void check(const char *what, vh_status st) {
    if (st != VH_OK) die(what, st);
}

// The reference stops the process on the first NaN with the message of the
// site that saw it (EnergyMomentumTensor.cpp:42,197-203;
// FullyDiscreteKurganovTadmorScheme.cpp:849,854).
void check_sync(vh_handle *h, vh_status st) {
    if (st == VH_ERR_NAN) {
        int site = 0;
        int64_t cell = -1;
        vh_nan_info(h, &site, &cell);
        static const char *msg[] = {"NaN", "Found e nan inside getInferredVariables", "f/fp is nan",
                                    "Found facShear nan", "Found facBulk nan"};
        std::printf("%s \n", msg[site >= 0 && site <= 4 ? site : 0]);
        std::exit(-1);
    }
    check("step", st);
}

// pi^{mu nu} (and Pi) at their Navier-Stokes values for a fluid at rest in
// Bjorken expansion, rhic/src/InitialConditions.cpp:530-596.
void navier_stokes_pi(vh_handle *h, const HydroPar &H, int bulk, size_t N, const std::vector<double> &e,
                      const std::vector<double> &p, std::vector<double> &pi, std::vector<double> &Pi) {
    const double t = H.t0, etabar = H.etabar;
    for (size_t s = 0; s < N; ++s) {
        double T = vh_eos_temperature(h, e[s]);
        if (T == 0) T = 1.e-3;
        const double pinn = -4.0 / (3.0 * t * t * t) * etabar * (e[s] + p[s]) / T;
        for (int c = 0; c < 10; ++c) pi[c * N + s] = 0;
        pi[4 * N + s] = -t * t * pinn / 2;  // pixx
        pi[7 * N + s] = -t * t * pinn / 2;  // piyy
        pi[9 * N + s] = pinn;
        if (bulk) {
            const double x = T / 1.01355;
            double zetabar = -13.77 * x * x + 27.55 * x - 13.45;
            if (x > 1.05) zetabar = 0.9 * std::exp(-(x - 1) / 0.025) + 0.25 * std::exp(-(x - 1) / 0.13) + 0.001;
            else if (x < 0.995) zetabar = 0.9 * std::exp((x - 1) / 0.0025) + 0.22 * std::exp((x - 1) / 0.022) + 0.03;
            Pi[s] = -zetabar * (e[s] + p[s]) / T / t;
        }
    }
}

// Initial state on the host, interior-only, x fastest.
void initial_state(vh_handle *h, const Lattice &L, const InitCond &I, const HydroPar &H, int bulk,
                   const std::string &root, const cpuvh::InitialTmunu *vec) {
    const size_t N = (size_t)L.nx * L.ny * L.nz;
    std::vector<double> e(N), p(N), ut(N, 1.0), ux(N, 0.0), uy(N, 0.0), un(N, 0.0), pi(10 * N, 0.0), Pi(N, 0.0);
    std::printf("Setting initial conditions: ");
    switch (I.type) {
    case 0: {
        std::printf("constant energy density.\n");
        const double ed = vh_eos_energy_density(h, 3.05);  // InitialConditions.cpp:958-959
        for (size_t s = 0; s < N; ++s) { e[s] = ed; p[s] = vh_eos_pressure(h, ed); }
        break;
    }
    case 2: {
        std::printf("Continous optical Glauber.\n");
        check("vh_ic_optical_glauber",
              vh_ic_optical_glauber(I.A, I.b, I.sigma_nn_mb, I.f_binary, I.e0, I.eta_flat, I.eta_variance, L.nx, L.ny,
                                    L.nz, L.dx, L.dy, L.dz, e.data()));
        for (size_t s = 0; s < N; ++s) p[s] = vh_eos_pressure(h, e[s]);
        break;
    }
    case 4: {
        std::printf("Monte carlo Glauber.\n");
        int nw = 0;
        check("vh_ic_mc_glauber_classic",
              vh_ic_mc_glauber_classic(I.A, I.b, I.sigma_nn_mb, I.e0, I.eta_flat, I.eta_variance, L.nx, L.ny, L.nz,
                                       L.dx, L.dy, L.dz, e.data(), &nw));
        std::printf("==> Found %d wounded nucleons.\n", nw);
        for (size_t s = 0; s < N; ++s) p[s] = vh_eos_pressure(h, e[s]);
        break;
    }
    case 1: {
        std::printf("Isreal-Stewart hydrodynamic Gubser flow test.\n");
        const std::string path = root + "/rhic/rhic-trunk/src/test/resources/gubser/viscous/gubserIC.dat";
        if (vh_ic_read_gubser(path.c_str(), L.nx, L.ny, L.nz, e.data(), p.data(), ut.data(), ux.data(), uy.data(),
                              un.data(), pi.data()) != VH_OK) {
            std::printf("Could not read %s\n", path.c_str());
            std::exit(-1);
        }
        break;
    }
    case 3: case 5: case 6: case 7: case 8: case 9: {
        static const char *name[10] = {"", "", "", "Ideal hydrodynamic Gubser flow test.",
                                       "", "Relativistic Sod shock-tube test.", "Implosion in a box test.",
                                       "Rayleigh-Taylor instability test.", "Implosion in a box test.",
                                       "Relativistic 2d Sod shock-tube test."};
        std::printf("%s\n", name[I.type]);
        check("vh_ic_test_problem", vh_ic_test_problem(I.type, I.e0, L.nx, L.ny, L.nz, L.dx, L.dy, e.data(), p.data(),
                                                        ut.data(), ux.data(), uy.data(), un.data()));
        break;
    }
    case 10: {
        std::printf("Reading initial T ^mu nu from input/e.dat , input/p.dat , etc... \n");
        // The reference reports a missing file and then fails in fclose
        // (InitialConditions.cpp:125-146); here the run stops at the report.
        static const char *files[14] = {"e", "ux", "uy", "un", "pitt", "pitx", "pity", "pitn",
                                        "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"};
        for (const char *f : files) {
            const bool eta = !std::strcmp(f, "un") || !std::strcmp(f, "pitn") || !std::strcmp(f, "pixn") ||
                             !std::strcmp(f, "piyn");
            if (eta && L.nz == 1) continue;
            const std::string path = root + "/input/" + f + ".dat";
            if (FILE *fp = std::fopen(path.c_str(), "r")) {
                std::fclose(fp);
            } else {
                std::printf("Couldn't open %s.dat!\n", f);
                std::exit(-1);
            }
        }
        if (vh_ic_read_fields(root.c_str(), L.nx, L.ny, L.nz, H.t0, e.data(), ut.data(), ux.data(), uy.data(),
                              un.data(), pi.data()) != VH_OK) {
            std::printf("Could not read the input/*.dat files of %s\n", root.c_str());
            std::exit(-1);
        }
        for (size_t s = 0; s < N; ++s) {
            p[s] = vh_eos_pressure(h, e[s]);
            if (bulk) Pi[s] = e[s] / 3. - p[s];  // InitialConditions.cpp:455-464
        }
        break;
    }
    case 11: {
        std::printf("Reading initial T ^mu nu from /input/Tmunu.dat \n");
        const std::string path = root + "/input/Tmunu.dat";
        if (vh_ic_read_tmunu(path.c_str(), L.nx, L.ny, L.nz, e.data(), p.data(), ut.data(), ux.data(), uy.data(),
                             un.data(), pi.data(), Pi.data()) != VH_OK) {
            std::printf("Could not read %s\n", path.c_str());
            std::exit(-1);
        }
        if (!bulk) std::fill(Pi.begin(), Pi.end(), 0.0);
        break;
    }
    case 13: case 14: {
        // HYDRO's vectors (rhic/src/InitialConditions.cpp:1593-1711): e in
        // GeV/fm^3 over hbar c plus the 1e-3 floor, p(e); type 13 at rest with
        // pi = Pi = 0, type 14 with u (components below 1e-7 taken as 0) and
        // pi, Pi over hbar c
        std::printf(I.type == 13 ? "Initialize from initial energy density vector ...\n"
                                 : "Initialize from preequilibrium vectors ...\n");
        auto need = [&](const std::vector<double> &v, const char *name) {
            if (v.size() != N) {
                std::printf("HYDRO vector %s has %zu values, the lattice %zu\n", name, v.size(), N);
                std::exit(-1);
            }
        };
        if (!vec) {
            std::printf("Initial condition type %d takes its state from the HYDRO wrapper's vectors\n", I.type);
            std::exit(-1);
        }
        need(vec->e, "initial_energy_density");
        for (size_t s = 0; s < N; ++s) {
            e[s] = (vec->e[s] / HBARC) + 1.0e-3;
            p[s] = vh_eos_pressure(h, e[s]);
        }
        if (I.type == 14) {
            const std::vector<double> *comp[4] = {&vec->ux, &vec->uy, &vec->un, nullptr};
            const char *cname[3] = {"initial_ux", "initial_uy", "initial_un"};
            for (int c = 0; c < 3; ++c) need(*comp[c], cname[c]);
            const std::vector<double> *piv[10] = {&vec->pitt, &vec->pitx, &vec->pity, &vec->pitn, &vec->pixx,
                                                  &vec->pixy, &vec->pixn, &vec->piyy, &vec->piyn, &vec->pinn};
            for (int c = 0; c < 10; ++c) need(*piv[c], "initial_pi");
            if (bulk) need(vec->Pi, "initial_Pi");
            for (size_t s = 0; s < N; ++s) {
                auto cut = [](double x) { return std::fabs(x) < 1.0e-7 ? 0.0 : x; };
                ux[s] = cut(vec->ux[s]);
                uy[s] = cut(vec->uy[s]);
                un[s] = cut(vec->un[s]);
                ut[s] = vec->reference_ut ? 1.0 + ux[s] * ux[s] + uy[s] * uy[s] + un[s] * un[s]
                                          : std::sqrt(1.0 + ux[s] * ux[s] + uy[s] * uy[s] + H.t0 * H.t0 * un[s] * un[s]);
                for (int c = 0; c < 10; ++c) pi[c * N + s] = (*piv[c])[s] / HBARC;
                if (bulk) Pi[s] = vec->Pi[s] / HBARC;
            }
        }
        break;
    }
    default:
        std::printf("Initial condition type %d is not available in this build. Exiting ...\n", I.type);
        std::exit(-1);
    }
    if (I.type == 0 || I.type == 2 || I.type == 4) {  // setFluidVelocityInitialCondition + setPimunuInitialCondition
        if (H.pi_navier_stokes == 1) {
            std::printf("Initialize \\pi^\\mu\\nu to its asymptotic Navier-Stokes value.\n");
            if (bulk) std::printf("Initialize \\Pi to its asymptotic Navier-Stokes value.\n");
            navier_stokes_pi(h, H, bulk, N, e, p, pi, Pi);
        } else {
            std::printf("Initialize \\pi^\\mu\\nu to zero.\n");
        }
    }
    check("vh_upload_ic", vh_upload_ic(h, e.data(), p.data(), ut.data(), ux.data(), uy.data(), un.data(), pi.data(),
                                       bulk ? Pi.data() : nullptr));
}

void write_snapshot(const std::vector<double> &v, double t, const std::string &dir, const char *name,
                    const Lattice &L) {
    char fname[4096];
    std::snprintf(fname, sizeof fname, "%s/%s_%.3f.dat", dir.c_str(), name, t);
    FILE *fp = std::fopen(fname, "w");
    if (!fp) {
        std::printf("cannot write %s\n", fname);
        std::exit(-1);
    }
    // the lines (k, then j, then i, x fastest) are formatted in parallel in
    // blocks of rows and written in order
    const long rows = (long)L.ny * L.nz;
    constexpr long ROWS_PER_BLOCK = 64;
    const long nblocks = (rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
    std::vector<std::string> text(nblocks);
    // each coordinate's "%.3f\t" text, formatted once
    auto coords = [](int n, double d) {
        std::vector<std::string> c(n);
        for (int i = 0; i < n; ++i) {
            put_fixed(c[i], (i - (n - 1) / 2.) * d, 3);
            c[i] += '\t';
        }
        return c;
    };
    const std::vector<std::string> xs = coords(L.nx, L.dx), ys = coords(L.ny, L.dy), zs = coords(L.nz, L.dz);
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < nblocks; ++b) {
        std::string &out = text[b];
        for (long r = b * ROWS_PER_BLOCK; r < std::min(rows, (b + 1) * ROWS_PER_BLOCK); ++r) {
            const int j = (int)(r % L.ny), k = (int)(r / L.ny);
            const std::string yz = ys[j] + zs[k];
            const size_t s0 = (size_t)r * L.nx;
            for (int i = 0; i < L.nx; ++i) {  // "%.3f\t%.3f\t%.3f\t%.8f\n"
                out += xs[i];
                out += yz;
                put_fixed(out, v[s0 + i], 8);
                out += '\n';
            }
        }
    }
    for (const std::string &t : text) std::fwrite(t.data(), 1, t.size(), fp);
    std::fclose(fp);
}

// outputDynamicalQuantities, rhic/src/HydroPlugin.cpp:40-72, with the
// thermal vorticity the reference computes after every step
// (THERMAL_VORTICITY is on in its default build, DynamicalVariables.h:15):
// wxy, beta_x and beta_y of the current state.  Before the first step the
// regulation factors and the vorticity are the reference's calloc'd zeros.
void snapshot(vh_handle *h, double t, const std::string &dir, const Lattice &L, int bulk, bool before_first_step,
              std::vector<double> &buf) {
    struct F { const char *name; int id; };
    const F shear[] = {{"e", 15}, {"ux", 18}, {"uy", 19}, {"ut", 17}, {"pixx", 8}, {"pixy", 9}, {"pinn", 13},
                       {"regFacShear", 25}};
    const F bulkf[] = {{"Pi", 14}, {"regFacBulk", 26}};
    const F vort[] = {{"wxy", 30}, {"beta_x", 34}, {"beta_y", 35}};
    auto one = [&](const F &f) {
        if (before_first_step && (f.id == 25 || f.id == 26 || f.id >= 27)) std::fill(buf.begin(), buf.end(), 0.0);
        else check_sync(h, vh_download(h, f.id, buf.data()));
        write_snapshot(buf, t, dir, f.name, L);
    };
    for (const F &f : shear) one(f);
    if (bulk)
        for (const F &f : bulkf) one(f);
    if (!before_first_step) check_sync(h, vh_thermal_vorticity(h));
    for (const F &f : vort) one(f);
}

// Sorts a[0..n) by .first (all keys in [0, limit)): least-significant digit
// first, 11 bits per stable counting pass; tmp is scratch of the same size.
template <class T>
void radix_sort_by_key(std::vector<T> &a, std::vector<T> &tmp, int64_t limit) {
    const size_t n = a.size();
    tmp.resize(n);
    constexpr int BITS = 11, R = 1 << BITS;
    std::vector<size_t> count(R);
    for (int shift = 0; shift < 63 && (limit - 1) >> shift; shift += BITS) {
        std::fill(count.begin(), count.end(), 0);
        for (size_t i = 0; i < n; ++i) ++count[(a[i].first >> shift) & (R - 1)];
        size_t sum = 0;
        for (int d = 0; d < R; ++d) {
            const size_t c = count[d];
            count[d] = sum;
            sum += c;
        }
        for (size_t i = 0; i < n; ++i) tmp[count[(a[i].first >> shift) & (R - 1)]++] = a[i];
        a.swap(tmp);
    }
}

// Per-coarse-step work space of the surface writer, kept across calls.
struct SurfaceWork {
    std::vector<std::pair<int64_t, int64_t>> order, scratch;  // (reference-order key, index into ids)
    std::vector<int> ne;
    std::vector<int64_t> off;
    std::vector<double> el, vars;
    std::vector<std::string> text;
};

// callFOFinder2p1D / callFOFinder3p1D (FreezeOut.cpp:200-361) for the nf cut
// cubes ids (x-fastest cube index, increasing, as vh_fo_flag returns them):
// the elements and the ten variables at their centroids come from the
// device (vh_fo_surface, the Cornelius construction and
// interpolateVariable3D/4D, cube by cube in the ids' order); the cubes are
// visited in the reference's order (x outer, eta inner), one line per
// element, the text built on all host threads and written afterwards in that
// order.  Adds the seconds spent sorting, on the device's elements and
// variables, and formatting to phase[0], phase[1] and phase[3].
void write_surface(std::ofstream &out, const std::vector<int64_t> &ids, int64_t nf, int dim, const Lattice &L,
                   const double *spacing, double cell_tau, double t, vh_handle *h, SurfaceWork &W, double *phase) {
    using clk = std::chrono::steady_clock;
    auto lap = [](clk::time_point &a) {
        const clk::time_point b = clk::now();
        const double s = std::chrono::duration<double>(b - a).count();
        a = b;
        return s;
    };
    clk::time_point t_lap = clk::now();
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int64_t cx = nx - 1, cy = ny - 1;
    W.order.resize(nf);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < nf; ++m) {  // key: (ix, iy, iz) lexicographic
        const int64_t c = ids[m], ix = c % cx, iy = (c / cx) % cy, iz = c / (cx * cy);
        W.order[m] = {(ix * cy + iy) * (int64_t)std::max(nz - 1, 1) + iz, m};
    }
    radix_sort_by_key(W.order, W.scratch, cx * cy * std::max(nz - 1, 1));
    phase[0] += lap(t_lap);
    const int per = 2 * dim;
    W.ne.resize(std::max<int64_t>(nf, 1));
    int64_t total = 0;
    const int64_t room = (int64_t)(W.vars.size() / 10);
    vh_status st = vh_fo_surface(h, spacing, W.ne.data(), W.el.data(), W.vars.data(), room, &total);
    if (st == VH_ERR_CAPACITY) {  // grow, and build again
        W.el.resize((size_t)total * per + total / 2 * per);
        W.vars.resize((size_t)total * 10 + total / 2 * 10);
        st = vh_fo_surface(h, spacing, W.ne.data(), W.el.data(), W.vars.data(), (int64_t)(W.vars.size() / 10),
                           &total);
    }
    if (st != VH_OK) {
        std::printf("freezeout surface: %s\n", vh_status_string(st));
        std::exit(-1);
    }
    W.off.resize(nf + 1);
    W.off[0] = 0;
    for (int64_t m = 0; m < nf; ++m) W.off[m + 1] = W.off[m] + W.ne[m];
    phase[1] += lap(t_lap);
    W.text.resize(nf);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < nf; ++r) {
        std::string &os = W.text[r];
        os.clear();
        const int64_t m = W.order[r].second, c = ids[m];
        const int ix = (int)(c % cx), iy = (int)((c / cx) % cy), iz = dim == 4 ? (int)(c / (cx * cy)) : 0;
        const double cell_x = (double)ix * L.dx - (((double)(nx - 1)) / 2.0 * L.dx);
        const double cell_y = (double)iy * L.dy - (((double)(ny - 1)) / 2.0 * L.dy);
        const double cell_z = dim == 4 ? (double)iz * L.dz - (((double)(nz - 1)) / 2.0 * L.dz) : 0.0;
        for (int e = 0; e < W.ne[m]; ++e) {
            const double *cen = W.el.data() + (size_t)(W.off[m] + e) * per, *nor = cen + dim;
            const double *var = W.vars.data() + (size_t)(W.off[m] + e) * 10;
            const double tau = cen[0] + cell_tau, x = cen[1] + cell_x, y = cen[2] + cell_y;
            const double eta = dim == 4 ? cen[3] + cell_z : 0.0;
            const double ds0 = t * nor[0], ds1 = t * nor[1], ds2 = t * nor[2], ds3 = dim == 4 ? t * nor[3] : 0.0;
            const double eps = var[3];
            const double T = vh_eos_temperature(h, eps), P = vh_eos_pressure(h, eps);
            // an ostream's default double output is printf's %g (6 significant digits)
            const double col[20] = {tau,    x,      y,      eta,    ds0,    ds1, ds2, ds3,    var[0], var[1],
                                    var[2], eps,    T,      P,      var[4], var[5], var[6], var[7], var[8], var[9]};
            for (int q = 0; q < 20; ++q) {
                put_g(os, col[q]);
                os += q < 19 ? ' ' : '\n';
            }
        }
    }
    phase[3] += lap(t_lap);
    for (int64_t r = 0; r < nf; ++r) out << W.text[r];
    out.flush();
}

void run(const Lattice &L, const InitCond &I, const HydroPar &H, const Extra &X, const std::string &root,
         const std::string &out, const cpuvh::InitialTmunu *vec) {
    const auto run_start = std::chrono::steady_clock::now();
    std::vector<double> table;
    if (X.eos_table) {
        table.resize(4 * 100000);
        FILE *f = std::fopen(X.eos_file.c_str(), "rb");
        if (!f || std::fread(table.data(), sizeof(double), table.size(), f) != table.size()) {
            std::printf("Could not read the EoS table %s\n", X.eos_file.c_str());
            std::exit(-1);
        }
        std::fclose(f);
    }
    vh_config cfg{};
    cfg.nx = L.nx; cfg.ny = L.ny; cfg.nz = L.nz;
    cfg.dx = L.dx; cfg.dy = L.dy; cfg.dz = L.dz; cfg.dt = L.dt; cfg.t0 = H.t0;
    cfg.etabar = H.etabar;
    cfg.T_reg_GeV = H.T_reg_GeV;
    cfg.bulk = X.bulk;
    cfg.eos_kind = X.eos_table ? VH_EOS_TABLE : VH_EOS_CONFORMAL;
    cfg.eos_table = X.eos_table ? table.data() : nullptr;
    cfg.etas_Tk_GeV = 0.154;
    cfg.device = X.device;
    vh_handle *h = nullptr;
    check("vh_create", vh_create(&cfg, &h));

    const int nt = L.nt, nx = L.nx, ny = L.ny, nz = L.nz, coarse = L.coarse;
    const double t0 = H.t0, dt = L.dt;
    const double TF = H.T_F_GeV / HBARC;
    const double eF = vh_eos_energy_density(h, TF);
    std::printf("Grid size = %d x %d x %d\n", nx, ny, nz);
    std::printf("spatial resolution = (%.3f, %.3f, %.3f)\n", L.dx, L.dy, L.dz);
    std::printf("freezeout temperature = %.3f [fm^-1] (eF = %.3f [fm^-4])\n", TF, eF);
    std::printf("Coarse graining FO surface by factor %d in proper time\n", coarse);
    std::printf("Regulating viscous currents using : Regulation Scheme 1\n");
    if (coarse < 1) {
        std::printf("coarseProperTimeFOFactor must be positive\n");
        std::exit(-1);
    }

    int dim = 0;
    if (nx > 1 && ny > 1 && nz > 1) dim = 4;
    else if (nx > 1 && ny > 1 && nz < 2) dim = 3;
    else std::printf("simulation is not in 3+1D or 2+1D; freezeout finder will not work!\n");
    std::ofstream surface(out + "/surface.dat");
    if (!surface) {
        std::printf("cannot write %s/surface.dat\n", out.c_str());
        std::exit(-1);
    }

    initial_state(h, L, I, H, X.bulk, root, vec);
    check("vh_init_conserved", vh_init_conserved(h, t0));
    const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - run_start).count();

    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> buf(N);
    check("vh_download", vh_download(h, 25, buf.data()));  // record regFacShear from the first step on
    if (X.bulk) check("vh_download", vh_download(h, 26, buf.data()));
    const size_t ncubes = (size_t)std::max(nx - 1, 1) * std::max(ny - 1, 1) * std::max(nz - 1, 1);
    // the flagged cubes and their corner energies; the device keeps the
    // previous and current coarse slices of e and the interpolated variables
    // (the previous ones start at zero, HydroPlugin.cpp:159-170)
    std::vector<int64_t> ids;
    SurfaceWork work;
    if (H.do_freezeout && dim) ids.resize(ncubes);
    const double spacing[4] = {dt * coarse, L.dx, L.dy, L.dz};  // HydroPlugin.cpp:130-145
    const int sc = nx / 2, sj = ny / 2, sk = nz / 2;  // the reference's centre cell (HydroPlugin.cpp:211-214)
    const size_t sctr = (size_t)sc + (size_t)nx * (sj + (size_t)ny * sk);

    double t = t0, total_ms = 0;
    // freezeout: flags; sort, elements, variables, text; output
    double fo_flag_s = 0, fo_phase[4] = {0, 0, 0, 0}, fo_write_s = 0;
    double snap_ms = 0;
    int nsteps = 0, zero_counts = 0, pend_first = 0, pend_count = 0;
    using clk = std::chrono::steady_clock;
    // queued steps run as one graph replay; their wall time counts as step time
    auto flush = [&]() {
        if (!pend_count) return;
        const auto a = clk::now();
        check_sync(h, vh_steps(h, pend_first, pend_count));
        check_sync(h, vh_sync(h));
        total_ms += std::chrono::duration<double, std::milli>(clk::now() - a).count();
        pend_count = 0;
    };
    const int wf = X.snapshot_every > 0 ? X.snapshot_every : 100;
    for (int n = 1; n <= nt + 1; ++n) {
        const bool snap = (n - 1) % wf == 0, fo = n % coarse == 0;
        if (snap || fo) flush();
        if (snap) {
            double ec = 0;
            check_sync(h, vh_download(h, 15, buf.data()));
            ec = buf[sctr];
            check_sync(h, vh_download(h, 16, buf.data()));
            const double pc = buf[sctr];
            std::printf("n = %d:%d (t = %.3f),\t (e, p) = (%.3f, %.3f) [fm^-4],\t (T = %.3f [GeV]),\t", n - 1, nt, t,
                        ec, pc, vh_eos_temperature(h, ec) * HBARC);
            const auto a = clk::now();
            snapshot(h, t, out, L, X.bulk, n == 1, buf);
            snap_ms += std::chrono::duration<double, std::milli>(clk::now() - a).count();
        }
        if (fo) {
            if (H.do_freezeout && dim) {
                // swapAndSetHydroVariables (FreezeOut.cpp:57-97) on the device, then the cut cubes
                const auto a = clk::now();
                int64_t nf = 0;
                check_sync(h, vh_fo_flag(h, eF, ids.data(), nullptr, (int64_t)ids.size(), &nf));
                const auto b = clk::now();
                const double before = fo_phase[0] + fo_phase[1] + fo_phase[2] + fo_phase[3];
                write_surface(surface, ids, nf, dim, L, spacing, t0 + (double)n * dt, t, h, work, fo_phase);
                const double build = fo_phase[0] + fo_phase[1] + fo_phase[2] + fo_phase[3] - before;
                fo_flag_s += std::chrono::duration<double>(b - a).count();
                fo_write_s += std::chrono::duration<double>(clk::now() - b).count() - build;
            }
            int64_t above = 0;
            check_sync(h, vh_count_above(h, eF, &above));
            if (above == 0) zero_counts += 1;
            if (zero_counts >= 3) {
                std::printf("\nAll cells have dropped below freezeout energy density\n");
                break;
            }
        }
        if (!pend_count) pend_first = n;
        ++pend_count;
        if (snap) {
            const double before = total_ms;
            flush();
            std::printf("(Elapsed time: %.3f ms)\n", total_ms - before);
        }
        ++nsteps;
        t = t0 + n * dt;
    }
    flush();
    std::printf("Average time/step: %.3f ms\n", nsteps ? total_ms / nsteps : 0.0);
    std::printf("Setup: %.3f s (EoS table, device, initial state)\n", setup_s);
    std::printf("Snapshots: %.3f s\n", snap_ms * 1e-3);
    if (H.do_freezeout && dim)
        std::printf("Freezeout: %.3f s flagging, %.3f s ordering, %.3f s elements and variables (device), "
                    "%.3f s text, %.3f s writing\n",
                    fo_flag_s, fo_phase[0], fo_phase[1], fo_phase[3], fo_write_s);
    std::printf("Steps taken: %d (t = %.3f)\n", nsteps, t);
    surface.close();
    check("vh_destroy", vh_destroy(h));
}

}  // namespace

namespace cpuvh {

// HydroWrapper.cpp:106-136 (and the reference's main): the options, the
// rhic-conf readers and the run.
int run_main(int argc, char **argv, const InitialTmunu *vec) {
    bool run_hydro = false, run_test = false, print_config = false;
    const char *config = nullptr, *output = nullptr;
    static const option opts[] = {{"test", optional_argument, nullptr, 't'},
                                  {"hydro", optional_argument, nullptr, 'h'},
                                  {"output", required_argument, nullptr, 'o'},
                                  {"config", required_argument, nullptr, 'c'},
                                  {"print-config", no_argument, nullptr, 'p'},
                                  {nullptr, 0, nullptr, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "t::h::o:c:", opts, nullptr)) != -1) {
        if (c == 't') run_test = true;
        else if (c == 'h') run_hydro = true;
        else if (c == 'o') output = optarg;
        else if (c == 'c') config = optarg;
        else if (c == 'p') print_config = true;
        else return 64;
    }
    char cwd[4096];
    const std::string root = getcwd(cwd, sizeof cwd) ? cwd : ".";
    std::printf("configDirectory = %s\n", config ? config : "(null)");
    std::printf("outputDirectory = %s\n", output ? output : "(null)");
    std::printf("runHydro = %s\n", run_hydro ? "True" : "False");
    std::printf("runTest = %s\n", run_test ? "True" : "False");
    std::printf("%s\n", vh_build_info());
    Lattice L;
    InitCond I;
    HydroPar H;
    Extra X;
    read_config(config ? config : ".", L, I, H, X);
    if (print_config) {
        std::printf("lattice %d %d %d %d %.17g %.17g %.17g %.17g %d\n", L.nx, L.ny, L.nz, L.nt, L.dx, L.dy, L.dz, L.dt,
                    L.coarse);
        std::printf("ic %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %d\n", I.type, I.n_source_files, I.A, I.e0,
                    I.sigma_nn_mb, I.b, I.f_binary, I.eta_variance, I.eta_flat, I.source_type);
        std::printf("hydro %.17g %.17g %.17g %d %.17g %d\n", H.t0, H.etabar, H.T_F_GeV, H.pi_navier_stokes, H.T_reg_GeV,
                    H.do_freezeout);
        std::printf("gpu %d %d %s %d %d\n", X.bulk, X.eos_table, X.eos_file.c_str(), X.device, X.snapshot_every);
    }
    const auto a = std::chrono::steady_clock::now();
    if (run_hydro) {
        run(L, I, H, X, root, output ? output : ".", vec);
        std::printf("Done hydro.\n");
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    std::printf("Hydro took %f seconds\n", sec);
    return 0;
}

// HydroWrapper.cpp:64-104: config "rhic-conf", output "cpu-vh_output" under
// the working directory.
int run_no_cli(const InitialTmunu *vec) {
    const char *config = "rhic-conf", *output = "cpu-vh_output";
    char cwd[4096];
    const std::string root = getcwd(cwd, sizeof cwd) ? cwd : ".";
    std::printf("configDirectory = %s\n", config);
    std::printf("outputDirectory = %s\n", output);
    Lattice L;
    InitCond I;
    HydroPar H;
    Extra X;
    read_config(config, L, I, H, X);
    const auto a = std::chrono::steady_clock::now();
    run(L, I, H, X, root, output, vec);
    std::printf("Done hydro.\n");
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    std::printf("Hydro took %f seconds\n", sec);
    return 0;
}

}  // namespace cpuvh
