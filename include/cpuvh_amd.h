/*
 * cpuvh_amd.h — C ABI of the MI355X (gfx950) viscous-hydro step.
 *
 * Drop-in boundary for cpu-vh's rungeKutta2 hot path.  The reference operates
 * on global SoA arrays with two ghost layers per side
 * (rhic/include/DynamicalVariables.h:47-113, rhic/src/DynamicalVariables.cpp:35-167);
 * here the state lives in HBM behind an opaque handle and crosses the boundary
 * as interior-only, x-fastest host arrays (index i + nx*(j + ny*k), the
 * reference's columnMajorLinearIndex order without ghosts).  Units as in the
 * reference: e, p in fm^-4, u^mu contravariant in (tau, x, y, eta_s).
 *
 * Field ids (vh_download / vh_upload_field):
 *   0..13  ttt ttx tty ttn pitt pitx pity pitn pixx pixy pixn piyy piyn pinn
 *   14     Pi (bulk builds only)
 *   15 e   16 p   17..20 ut ux uy un   21..24 previous-step ut ux uy un
 *   25     regulation factor of the last stage (regFacShear)
 *   26     bulk regulation factor of the last stage (regFacBulk; bulk builds only)
 *   27..32 thermal vorticity wtx wty wtn wxy wxn wyn, 33..36 thermal velocity
 *          beta_t beta_x beta_y beta_n (from the last vh_thermal_vorticity)
 * Fields 25 and 26 are recorded only once requested: the first vh_download of
 * either returns ones and turns its recording on from the next step on
 * (the reference writes both in regulateDissipativeCurrents,
 * FullyDiscreteKurganovTadmorScheme.cpp:848,853).
 *
 * All entry points return a vh_status; none aborts the process.
 */
#ifndef CPUVH_AMD_H
#define CPUVH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VH_OK = 0,
    VH_ERR_ARG = -1,        /* invalid argument or configuration */
    VH_ERR_HIP = -2,        /* HIP runtime error */
    VH_ERR_NOMEM = -3,      /* device allocation failed */
    VH_ERR_NAN = -4,        /* NaN in recovery or regulation (reference: printf + exit(-1)) */
    VH_ERR_STATE = -5,      /* call out of order (e.g. step before init) */
    VH_ERR_CAPACITY = -6    /* output buffer too small; required size returned */
} vh_status;

/* Where the NaN of VH_ERR_NAN came from (reference message in brackets). */
typedef enum {
    VH_NAN_NONE = 0,
    VH_NAN_ENERGY = 1,      /* "Found e nan inside getInferredVariables" EnergyMomentumTensor.cpp:197-203 */
    VH_NAN_ROOT = 2,        /* "f/fp is nan" EnergyMomentumTensor.cpp:42 */
    VH_NAN_REG_SHEAR = 3,   /* "Found facShear nan" FullyDiscreteKurganovTadmorScheme.cpp:849 */
    VH_NAN_REG_BULK = 4     /* "Found facBulk nan" FullyDiscreteKurganovTadmorScheme.cpp:854 */
} vh_nan_site;

typedef enum { VH_EOS_CONFORMAL = 0, VH_EOS_TABLE = 1 } vh_eos_kind;

/* Replaces the reference's LatticeParameters + HydroParameters structs
 * (rhic/include/LatticeParameters.h:17-34, rhic/include/HydroParameters.h:15-23)
 * plus the compile-time switches PI (bulk) and *_EOS
 * (rhic/include/DynamicalVariables.h:13-15, rhic/include/EquationOfState.h:13-15). */
typedef struct {
    int nx, ny, nz;            /* numLatticePointsX/Y/Rapidity (interior points) */
    double dx, dy, dz;         /* latticeSpacingX/Y/Rapidity */
    double dt;                 /* latticeSpacingProperTime */
    double t0;                 /* initialProperTimePoint */
    double etabar;             /* shearViscosityToEntropyDensity (eta/s, or its minimum) */
    double T_reg_GeV;          /* regulationTemperatureGeV */
    int bulk;                  /* 1: evolve Pi (reference built with -DPI) */
    int eos_kind;              /* vh_eos_kind */
    const double *eos_table;   /* VH_EOS_TABLE: 4*100000 little-endian doubles (e, P, s, T in GeV units),
                                  the layout of EoS/hrg_hotqcd_eos_binary.dat; copied */
    /* eta/s(T) extension (no reference counterpart: the reference's eta/s is constant,
       rhic/src/SourceTerms.cpp:69).  eta/s = etabar + slope_hi*(T-Tk) above Tk and
       etabar + slope_lo*(Tk-T) below, T in GeV.  Both slopes 0 reproduces the reference. */
    double etas_Tk_GeV, etas_slope_lo, etas_slope_hi;
    int device;                /* HIP device ordinal */
    /* Bulk transport extensions (no reference counterpart; the zero values give
       the reference).  cs2_mode 0: the reference's speedOfSoundSquared, the
       one-table-segment quotient delta_e/delta_p clamped to [0.01, 1/3]
       (EquationOfState.cpp:141-153; 1/3 for the bundled table, which makes every
       bulk coefficient (1/3 - c_s^2)^2 vanish, SourceTerms.cpp:75-82); 1: the
       table's sound speed delta_p/delta_e over the same segment, same clamp.
       zeta_Tpeak_GeV > 0: zeta/s(T) = zeta_peak/0.33 * f(T/zeta_Tpeak_GeV), f the
       reference's shape (SourceTerms.cpp:40-48, peak 0.33 at its argument 1),
       so zeta/s peaks at zeta_Tpeak_GeV with the value zeta_peak; 0: the
       reference's zeta/s, f(T/1.01355 fm^-1).  zeta_peak 0 (with zeta_Tpeak_GeV
       > 0) is zeta/s = 0.  Where the bulk terms are reached (cs2_mode 1) the
       exponentials of f are vh's own fdlibm-method exp, the same on the device
       and in tests, and the bulk relaxation rate 15 (1/3 - c_s^2)^2 T / (zeta/s)
       is held at 1/dt at most, with beta_Pi scaled alike (the Navier-Stokes
       value -zeta theta is kept): the explicit RK2 step is unstable beyond
       dt/tau_Pi = 2, which a small zeta/s reaches.  None of this changes a
       configuration with cs2_mode 0, where (1/3 - c_s^2) = 0. */
    int cs2_mode;
    double zeta_Tpeak_GeV, zeta_peak;
} vh_config;

typedef struct vh_handle vh_handle;

/* Allocates the device state (allocateHostMemory, DynamicalVariables.cpp:43-167). */
vh_status vh_create(const vh_config *cfg, vh_handle **out);
vh_status vh_destroy(vh_handle *h);

/* Initial state, the type-11 reader's assignments (rhic/src/InitialConditions.cpp:41-105):
 * e, p, u^mu, previous u^mu := u^mu, pi^{mu nu} (10 arrays back to back, pitt..pinn),
 * Pi.  Host arrays of nx*ny*nz doubles.  Every array but e and p may be NULL:
 * u^tau is then 1 everywhere, the others 0 (a fluid at rest without
 * dissipative currents, as the Glauber and constant initial states start),
 * and nothing is copied for it. */
vh_status vh_upload_ic(vh_handle *h, const double *e, const double *p, const double *ut, const double *ux,
                       const double *uy, const double *un, const double *pi10, const double *Pi);

/* setConservedVariables (rhic/src/DynamicalVariables.cpp:169-218) followed by the
 * ghost refresh of HydroPlugin.cpp:204 (implicit here: clamp-to-edge indexing). */
vh_status vh_init_conserved(vh_handle *h, double t);

/* rungeKutta2(t, dt, q, Q, ...) + setCurrentConservedVariables()
 * (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:866-924, rhic/src/HydroPlugin.cpp:292,300).
 * Asynchronous on the handle's stream; a NaN is reported by the next synchronising call. */
vh_status vh_step(vh_handle *h, double t);

/* count steps n = n_first .. n_first+count-1 of the driver loop, each at
 * t = t0 + (n-1)*dt as HydroPlugin.cpp:226,307 advances it, replayed from a
 * captured HIP graph. */
vh_status vh_steps(vh_handle *h, int n_first, int count);

/* Waits for the stream and returns VH_ERR_NAN if any stage hit a NaN. */
vh_status vh_sync(vh_handle *h);
vh_status vh_nan_info(vh_handle *h, int *site, int64_t *cell);

/* calculateThermalVorticity (rhic/src/Vorticity.cpp:6-171), which the
 * reference runs after every step (rhic/src/HydroPlugin.cpp:302-305) and
 * writes (wxy, beta_x, beta_y) into its snapshots (:63-70): the thermal
 * vorticity w_{mu nu} = -(d_mu beta_nu - d_nu beta_mu)/2 of beta_mu = u_mu/T
 * from the current state (u, the previous step's u, e), bit for bit the
 * reference's, into the handle's fields 27..32 (wtx wty wtn wxy wxn wyn) and
 * 33..36 (beta_t beta_x beta_y beta_n) for vh_download.  Computing it on
 * demand gives what the reference holds after the same steps.  Enqueued on
 * the handle's stream.  Fields 27..36 are VH_ERR_ARG before the first call. */
vh_status vh_thermal_vorticity(vh_handle *h);

/* Host copy of one field (ids above), nx*ny*nz doubles.  Synchronising. */
vh_status vh_download(vh_handle *h, int field, double *out);
vh_status vh_upload_field(vh_handle *h, int field, const double *in);

/* checkForCellsAboveTc (freezeout/FreezeOut.cpp:364-381): number of interior
 * cells with e > eF. */
vh_status vh_count_above(vh_handle *h, double eF, int64_t *count);

/* Freezeout cell flagging for one coarse step: the reference keeps the
 * previous and current coarse e slices (swapAndSetHydroVariables,
 * freezeout/FreezeOut.cpp:57-97; the previous slice starts as zeros,
 * HydroPlugin.cpp:159) and Cornelius builds elements only in cubes with
 * 0 < #(corners >= eF) < 2^dim (cornelius.cpp:2201-2212, :2257-2272) over
 * ix<nx-1, iy<ny-1 (, iz<nz-1) (FreezeOut.cpp:207-212, :288-291).
 * This call shifts the slices (previous := last current, current := e) and
 * writes the flagged cube ids (ix + (nx-1)*(iy + (ny-1)*iz)) in increasing
 * order into ids (capacity cap), the count into *n_flagged, and, when corners
 * is not NULL, the 2^dim corner energies of each flagged cube (previous slice
 * first, then current; within a slice x fastest).  VH_ERR_CAPACITY if cap is
 * too small (*n_flagged still set). */
vh_status vh_fo_flag(vh_handle *h, double eF, int64_t *ids, double *corners, int64_t cap, int64_t *n_flagged);

/* The ten variables surface.dat interpolates (u^x u^y u^eta e pi^xx pi^xy
 * pi^xeta pi^yy pi^yeta Pi): the reference keeps their previous and current
 * coarse slices beside e's (swapAndSetHydroVariables, FreezeOut.cpp:57-97),
 * and vh_fo_flag shifts them with e's (Pi stays zero without bulk).  At n
 * points inside cubes: cubes[m] a cube id numbered as vh_fo_flag numbers
 * them, frac[4m .. 4m+3] the point's offset from the cube's first corner in
 * units of (coarse time step, dx, dy, deta), the fourth unused in 2+1D;
 * vars[10m + v] = interpolateVariable3D/4D (FreezeOut.cpp:182-198, called at
 * :248-262 and :327-343), the terms of the multilinear interpolation summed
 * in the reference's order.  VH_ERR_ARG before the first vh_fo_flag call or
 * for a cube id out of range.  Synchronising. */
vh_status vh_fo_interp(vh_handle *h, const int64_t *cubes, const double *frac, int64_t n, double *vars);

/* The surface elements of the cubes the last vh_fo_flag call flagged, built
 * on the device: replaces the reference's per-cube host loop
 * (callFOFinder2p1D / callFOFinder3p1D, freezeout/FreezeOut.cpp:200-361, with
 * Cornelius::find_surface_3d/4d, freezeout/cornelius-c++-1.3/cornelius.cpp:
 * 2192-2287) and the interpolation at the centroids (:248-262, :327-343).
 * Each element is the one vh_surface_cube (2+1D) / vh_surface_hypercube
 * (3+1D) builds from the cube's corner energies at that call's level, bit for
 * bit.  widths: the cell's widths (coarse dtau, dx, dy[, deta]).  For the
 * flagged cubes in vh_fo_flag's order, n_el[m] receives cube m's element
 * count; elements (2*dim doubles per element: centroid relative to the
 * cube's first corner, then the normal, dim = 3 in 2+1D and 4 in 3+1D) and
 * vars (the ten variables of vh_fo_interp at the centroid, 10 doubles per
 * element) follow cube by cube in that order.  cap: capacity in elements;
 * *n_elements receives the total.  VH_ERR_CAPACITY when cap is too small
 * (n_el and *n_elements are set: call again with room); VH_ERR_STATE for a
 * cube whose surface cannot be closed; VH_ERR_ARG before the first vh_fo_flag
 * call.  Synchronising. */
vh_status vh_fo_surface(vh_handle *h, const double *widths, int *n_el, double *elements, double *vars, int64_t cap,
                        int64_t *n_elements);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): vh_fo_flag's
 * crossing test and ordered compaction on caller-owned DEVICE memory,
 * allocating nothing.  eprev, ecur: the previous and current coarse slices,
 * nx*ny*nz doubles each (x fastest); ids: at most cap cube ids, numbered and
 * ordered as vh_fo_flag's (ids beyond cap are not written); *n_flagged (a
 * device int64) receives the full count.  Enqueued on `stream` (a
 * hipStream_t, NULL for the null stream) of the current device, with no
 * synchronisation, so the call can be captured into a graph; the results are
 * there once the stream reaches them.  The kernels' per-block counts live in
 * one fixed device array of the library per GPU (512 KiB, kept until the
 * process ends): at most VH_FO_NOALLOC_MAX_CUBES cubes ((nx-1)(ny-1)(nz-1),
 * nz-1 read as 1 in 2+1D), else VH_ERR_ARG; calls that overlap on one GPU race
 * for that array. */
#define VH_FO_NOALLOC_MAX_CUBES 16777216L
vh_status vh_fo_flag_noalloc(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                             int64_t *ids, int64_t cap, int64_t *n_flagged, void *stream);

/* Host-side EoS of the handle (EquationOfState.cpp:44-290): used by drivers for
 * e_F = e(T_F) and by tests. */
double vh_eos_energy_density(const vh_handle *h, double T);
double vh_eos_pressure(const vh_handle *h, double e);
double vh_eos_temperature(const vh_handle *h, double e);

/* Average duration (ms) of the last timed vh_steps call's stage kernels,
 * from HIP events on the handle's stream; 0 if none. */
double vh_last_kernel_ms(const vh_handle *h);

/* getInferredVariables (rhic/include/EnergyMomentumTensor.h:13-16,
 * rhic/src/EnergyMomentumTensor.cpp:117-233) for n states at time t, on the
 * device, with the handle's EoS and bulk setting and the arithmetic the stage
 * kernels apply to every cell: q[k*n + i] is conserved variable k of state i
 * (k < 14, or 15 with bulk: ttt ttx tty ttn pitt ... pinn [Pi]), ePrev[i] its
 * previous energy density; out[j*n + i] = e, p, u^tau, u^x, u^y, u^eta
 * (j = 0..5).  Where the reference prints and exits on a NaN
 * (EnergyMomentumTensor.cpp:42,197-203) nan_site[i] (optional) records the
 * site -- 1 e, 2 the root solve's f/fp -- and the call returns VH_ERR_NAN
 * after writing every state.  Synchronising. */
vh_status vh_inferred_variables(vh_handle *h, double t, const double *q, const double *ePrev, int64_t n, double *out,
                                int32_t *nan_site);

/* ---- initial conditions (host side; no device needed) ------------------ */

/* Seeded Monte-Carlo Glauber energy density (the reference's IC type 4,
 * rhic/src/InitialConditions.cpp:1120-1140 with its generator in
 * rhic/src/MonteCarloGlauberModel.cpp:18-111, here with every nuclear
 * parameter and a 64-bit seed exposed).  e = e0 * T(x,y) * L(eta) + e_floor,
 * T a sum of normalised Gaussians (width) over wounded nucleons, L the
 * reference's plateau + Gaussian tails (InitialConditions.cpp:1046-1064). */
typedef struct {
    int A;                 /* nucleons per nucleus */
    double R, a;           /* Woods-Saxon radius and diffuseness [fm] */
    double sigma_nn_fm2;   /* inelastic NN cross section [fm^2] */
    double b;              /* impact parameter [fm]; < 0: uniform in [0, b_max] from the seed */
    double b_max;
    double width;          /* participant Gaussian width [fm] */
    double eta_flat;       /* rapidityMean: full width of the plateau */
    double eta_variance;   /* rapidityVariance of the tails */
    double e0;             /* normalisation [fm^-2] */
    double e_floor;        /* additive floor [fm^-4] (reference: 1e-3) */
    uint64_t seed;
} vh_glauber_params;

typedef struct {
    double b;
    int n_part, n_coll;
} vh_glauber_info;

vh_status vh_ic_mc_glauber(const vh_glauber_params *gp, int nx, int ny, int nz, double dx, double dy, double dz,
                           double *e_out, vh_glauber_info *info);

/* ---- freezeout surface elements (host side; no device needed) ----------
 * The constant-value surface inside one lattice cell, built as the
 * reference's surface finder Cornelius 1.3 builds it
 * (freezeout/cornelius-c++-1.3/cornelius.cpp: Cornelius::find_surface_2d
 * :2127-2151, find_surface_3d :2161-2246, find_surface_4d :2248-2287).
 * corners: 4, 8 or 16 values, the corner at (i d0, j d1, k d2, l d3) at index
 * ((i*2 + j)*2 + k)*2 + l (as many indices as dimensions); E0 the level; d
 * the cell's edge lengths.  Each element is its centroid (relative to corner
 * 0) followed by its normal vector, which points to the side below E0 and has
 * the element's length, area or volume as its norm: 4, 6 or 8 doubles per
 * element.  VH_ERR_CAPACITY if more than cap elements (*n_elements still
 * set). */
vh_status vh_surface_square(const double *corners, double E0, const double *d, double *elements, int cap,
                            int *n_elements);
vh_status vh_surface_cube(const double *corners, double E0, const double *d, double *elements, int cap,
                          int *n_elements);
vh_status vh_surface_hypercube(const double *corners, double E0, const double *d, double *elements, int cap,
                               int *n_elements);

/* The reference's own IC type 4 (rhic/src/InitialConditions.cpp:1107-1140,
 * rhic/src/MonteCarloGlauberModel.cpp:18-111, rhic/src/GlauberModel.cpp:19-24):
 * the C library's rand() seeded with its fixed seed 1328398221, so the same
 * ic.properties give the same profile as the reference on this C library.
 * A = numberOfNucleonsPerNuclei, b = impactParameter [fm], sigma_nn_mb =
 * scatteringCrossSectionNN, e0 = initialEnergyDensity, eta_flat/eta_variance =
 * rapidityMean/rapidityVariance; e includes the reference's 1e-3 floor.
 * Reseeds the process-wide rand() state, as the reference does. */
vh_status vh_ic_mc_glauber_classic(int A, double b, double sigma_nn_mb, double e0, double eta_flat,
                                   double eta_variance, int nx, int ny, int nz, double dx, double dy, double dz,
                                   double *e_out, int *n_wounded);

/* The reference's IC type 2, the optical Glauber profile
 * (rhic/src/InitialConditions.cpp:1069-1105, rhic/src/GlauberModel.cpp:19-108):
 * alpha (fractionOfBinaryCollisions) * binary collisions + (1 - alpha) *
 * wounded nucleons from the Woods-Saxon thickness, normalised to 1 at the
 * centre of a central collision, times e0 and the longitudinal profile, plus
 * the 1e-3 floor.  The thickness integral is accurate to 1e-13 (the reference
 * asks GSL for 1e-8). */
vh_status vh_ic_optical_glauber(int A, double b, double sigma_nn_mb, double alpha, double e0, double eta_flat,
                                double eta_variance, int nx, int ny, int nz, double dx, double dy, double dz,
                                double *e_out);

/* The reference's analytic initial states, type = its initialConditionType
 * (rhic/src/InitialConditions.cpp): 3 ideal Gubser flow (:1146-1190; u = 0
 * at r = 0, where the reference divides 0 by 0), 5 Sod shock tube (:1259),
 * 6 implosion in a box (:1340), 7 Rayleigh-Taylor (:1380), 8 Gaussian pulse
 * (:1431), 9 diagonal Sod shock tube (:1300).  p = e/3 as there; pi and Pi
 * are zero.  e0 is used by type 3 only. */
vh_status vh_ic_test_problem(int type, double e0, int nx, int ny, int nz, double dx, double dy, double *e, double *p,
                             double *ut, double *ux, double *uy, double *un);

/* IC type 10 reader (rhic/src/InitialConditions.cpp:108-466): dir/input/
 * e.dat, ux.dat, uy.dat, un.dat and pitt.dat..pinn.dat ("x y z value" lines
 * through float, i outer, k inner; the eta files only when nz > 1).  u^tau is
 * normalised (the reference stores 1 + u_perp^2 + t0^2 un^2 without the
 * square root, :226).  The caller sets p = p(e) and, with bulk, Pi = e/3 - p,
 * as :200 and :455-464 do. */
vh_status vh_ic_read_fields(const char *dir, int nx, int ny, int nz, double t0, double *e, double *ut, double *ux,
                            double *uy, double *un, double *pi10);

/* IC type 1 reader (rhic/src/InitialConditions.cpp:1196-1257): one line per
 * transverse cell, i outer, "x y e ux uy pixx piyy pixy pitt pitx pity pinn",
 * repeated over rapidity; p = e/3. */
vh_status vh_ic_read_gubser(const char *path, int nx, int ny, int nz, double *e, double *p, double *ut, double *ux,
                            double *uy, double *un, double *pi10);

/* input/Tmunu.dat (IC type 11) reader, values through float as the reference
 * reads them (rhic/src/InitialConditions.cpp:41-105); Pi may be NULL.  The
 * writer produces a file that reader (and the reference) accept. */
vh_status vh_ic_read_tmunu(const char *path, int nx, int ny, int nz, double *e, double *p, double *ut, double *ux,
                           double *uy, double *un, double *pi10, double *Pi);
vh_status vh_ic_write_tmunu(const char *path, int nx, int ny, int nz, double dx, double dy, double dz,
                            const double *e, const double *p, const double *ut, const double *ux, const double *uy,
                            const double *un, const double *pi10, const double *Pi);

/* Builds (without launching) the HIP graph vh_steps replays for this count,
 * so a timed vh_steps call measures replay only.  No counterpart in the
 * reference (which has no graphs). */
vh_status vh_prepare_steps(vh_handle *h, int count);

/* ---- batches of events ---------------------------------------------------
 * No counterpart in the reference, which runs one event per process
 * (rhic/src/HydroPlugin.cpp:74-321).  A batch holds many independent events
 * in one pooled device state; each event keeps its own lattice, spacings,
 * clock, eta/s and T_reg, and vh_batch_steps advances every resident event by
 * the same number of steps (two launches per RK stage: all 2+1D events, then
 * all 3+1D events).  Results of an event are bit-identical to running it
 * alone through vh_handle.  bulk and the EoS are per batch.  One host thread
 * at a time per batch: its calls share staging buffers and streams. */
/* Arithmetic of a batch.  VH_PREC_FP64: the drop-in path (each event
 * bit-identical to vh_handle).  VH_PREC_MIXED: FP32 storage (with an FP32
 * low part kept for T^{tau mu}) and FP32 slopes, face fluxes, source terms
 * and regulator factors; FP64 root solve, EoS table lookups and
 * conserved-variable accumulation; each face flux evaluated once for the two
 * cells it separates.  No counterpart in the reference (FP64 only). */
typedef enum { VH_PREC_FP64 = 0, VH_PREC_MIXED = 1 } vh_precision;

typedef struct {
    int device;
    int64_t capacity_cells;    /* cells of all resident events together (61 doubles, or 73 floats, each) */
    int max_events;
    int max_steps;             /* steps an event may take before it must be removed */
    int bulk;
    int eos_kind;
    const double *eos_table;   /* as vh_config.eos_table */
    int precision;             /* vh_precision */
    int cs2_mode;              /* as vh_config.cs2_mode */
    int fo_vars;               /* 1: vh_batch_fo_flag also keeps each event's previous and current slices of
                                  the ten surface variables (20 doubles a cell), for vh_batch_fo_surface */
} vh_batch_config;

/* One event's parameters: vh_config's per-event fields. */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, dt, t0;
    double etabar, T_reg_GeV;
    double etas_Tk_GeV, etas_slope_lo, etas_slope_hi;
    double zeta_Tpeak_GeV, zeta_peak;  /* as vh_config's */
} vh_event_params;

typedef struct vh_batch vh_batch;

vh_status vh_batch_create(const vh_batch_config *cfg, vh_batch **out);
vh_status vh_batch_destroy(vh_batch *b);
/* Admits one event (vh_upload_ic + vh_init_conserved at t0 for it, with
 * vh_upload_ic's NULL arrays): first-fit placement of its nx*ny*nz cells;
 * VH_ERR_CAPACITY when no free range or event slot is large enough.
 * Returns once the arrays are copied (the caller may reuse them at once);
 * the upload and the initial-state kernel are queued on the batch's stream
 * without waiting, so every later call sees the event's state, and a
 * device fault in them is reported by the next call that waits. */
vh_status vh_batch_add(vh_batch *b, const vh_event_params *prm, const double *e, const double *p, const double *ut,
                       const double *ux, const double *uy, const double *un, const double *pi10, const double *Pi,
                       int *id_out);
/* Releases an event's cells and slot.  Synchronising. */
vh_status vh_batch_remove(vh_batch *b, int id);
/* Every resident event takes count steps (event steps n..n+count-1 with n its
 * own next step), replayed from a captured HIP graph.  Asynchronous. */
vh_status vh_batch_steps(vh_batch *b, int count);
vh_status vh_batch_sync(vh_batch *b);
/* NaN report of the batch: site (vh_nan_site), event id and cell within it. */
vh_status vh_batch_nan_info(vh_batch *b, int *site, int *event, int64_t *cell);
vh_status vh_batch_clear_error(vh_batch *b);
/* counts[id] = cells of event id with e > eF (-1 for free slots); max_events entries. */
vh_status vh_batch_counts_above(vh_batch *b, double eF, int64_t *counts);
/* The reference's run loop for every resident event at once, stop rule on
 * the device (rhic/src/HydroPlugin.cpp:226-308, the check and its 3-strike
 * stop at :266-281): nchecks times, coarse steps then the count of cells
 * above eF; an event whose count is zero gets a strike (strikes are never
 * reset, counting from its vh_batch_add) and at the third it stops being
 * stepped, at that check, without a host round trip.  One captured graph per
 * (coarse, nchecks, eF, layout).  On return stopped[id] is 1 for an event
 * that has stopped (now or in an earlier call), 0 for one still running, -1
 * for a free id; steps_done[id] its steps (either array may be NULL).  A
 * stopped event stays resident, and is no longer stepped, until
 * vh_batch_remove.  VH_ERR_NAN as vh_batch_sync (the outputs are still
 * filled); VH_ERR_CAPACITY if a running event would pass max_steps.  No
 * counterpart in the reference, which checks one event on the host. */
vh_status vh_batch_run(vh_batch *b, double eF, int coarse, int nchecks, int32_t *stopped, int32_t *steps_done);
/* As vh_download / vh_fo_flag for one event. */
vh_status vh_batch_download(vh_batch *b, int id, int field, double *out);
vh_status vh_batch_fo_flag(vh_batch *b, int id, double eF, int64_t *ids, double *corners, int64_t cap,
                           int64_t *n_flagged);
/* vh_batch_fo_flag for several events in one pass (ids[0..n), distinct;
 * no corners): every event's flag kernels queued, then one download of the
 * counts and one of the ids, so two synchronisations in all.  cube_ids[k]
 * receives event k's cut cubes' ids when they fit its cap[k];
 * n_flagged[k] their count.  VH_ERR_CAPACITY when an event's cubes exceed
 * its cap (every count is set, every event's slices have moved and its
 * surface can be built; only its ids are not copied).  No counterpart in
 * the reference (one event, one check). */
vh_status vh_batch_fo_flags(vh_batch *b, int n, const int *ids, double eF, int64_t *const *cube_ids,
                            const int64_t *cap, int64_t *n_flagged);
/* As vh_fo_surface for one event of a batch created with fo_vars = 1, on the
 * cubes its last vh_batch_fo_flag call flagged (mixed-precision events'
 * fields widened to FP64 first).  It reads only the copies that flag call
 * made and runs on a stream of its own, so steps enqueued after the flag
 * (vh_batch_steps) proceed while it builds; the event's next
 * vh_batch_fo_flag must come after it returns. */
vh_status vh_batch_fo_surface(vh_batch *b, int id, const double *widths, int *n_el, double *elements, double *vars,
                              int64_t cap, int64_t *n_elements);
/* Surfaces of several events of such a batch in one pass: as
 * vh_batch_fo_surface for each of ids[0..n) (distinct), with one element
 * launch over all their cut cubes and two synchronisations in all, on the
 * same stream of its own.  Event k: widths[4k..4k+3] its cell widths (the
 * fourth ignored for nz = 1), n_el[k] room for its flag count, elements[k]
 * and vars[k] room for cap[k] elements, n_elements[k] its element count.
 * VH_ERR_CAPACITY when an event's elements exceed its cap: every count is
 * set and the events with room have their elements; a second call for the
 * others, with larger arrays, builds theirs.  No counterpart in the
 * reference (one event, one surface call). */
vh_status vh_batch_fo_surfaces(vh_batch *b, int n, const int *ids, const double *widths, int *const *n_el,
                               double *const *elements, double *const *vars, const int64_t *cap,
                               int64_t *n_elements);
vh_status vh_batch_event_info(vh_batch *b, int id, int *active, int *steps_done, int64_t *cells);
vh_status vh_batch_free_cells(vh_batch *b, int64_t *free_cells, int64_t *largest_range);
/* Mixed form: times a stage launch of each tile variant in turn, reps rounds,
 * on the resident events (the state is not advanced) and keeps the 2+1D and
 * 3+1D tiles by median time, a later variant replacing the one kept only
 * when more than 2 % faster (variant 0 is the fastest over whole runs of 2+1D
 * Glauber batches).  FP64 batches: no variants, *tile = -1. */
vh_status vh_batch_autotune(vh_batch *b, int reps, int *tile2d, int *tile3d);
vh_status vh_batch_set_tiles(vh_batch *b, int tile2d, int tile3d);
/* Mixed form: how face fluxes are shared.  The reference evaluates each face
 * twice, once per neighbouring cell, each time starting both root solves
 * from that cell's e (EnergyMomentumTensor.cpp:188-191 switches formula at
 * e <= 0.1; otherwise e is the Newton start).  1 (default): both
 * evaluations, as the reference (the second only where it can differ);
 * 0: one evaluation per face, each side started from its own cell's e
 * (faster, but it departs from the reference where the two starts differ);
 * 2: one evaluation, from the left cell's e, plus the right cell's
 * evaluation where the two cells sit on different sides of e = 0.1. */
vh_status vh_batch_set_face_mode(vh_batch *b, int mode);
/* Duration (ms) of the last vh_batch_steps graph, from HIP events; 0 if none. */
double vh_batch_last_ms(const vh_batch *b);
double vh_batch_eos_energy_density(const vh_batch *b, double T);

/* ---- serving -------------------------------------------------------------
 * No counterpart in the reference.  An open-loop request loop on one device:
 * requests (seeded MC-Glauber events, a share of them 3+1D) arrive as a
 * Poisson process in wall-clock time; each is admitted into one vh_batch as
 * soon as its state fits the HBM budget (first fit, later requests may pass
 * one that does not fit yet), stays resident while every resident request
 * advances by `coarse` steps per round, and leaves when no cell is above
 * e(T_F) at three consecutive checks (HydroPlugin.cpp:266-281).  Initial
 * states are generated before the clock starts (the requests' payload). */
typedef struct {
    double rate_per_s;         /* mean arrival rate */
    int n_requests;
    uint64_t seed;
    double frac_3d;            /* share of 3+1D requests */
    int side2_min, side2_max;  /* 2+1D: nx = ny uniform in [min, max], nz = 1 */
    int side3_min, side3_max;  /* 3+1D: nx = ny uniform in [min, max] */
    int nz3_min, nz3_max;      /* 3+1D: nz uniform in [min, max] */
    double dx, dt, t0;         /* dy = deta = dx */
    double etabar, etas_Tk_GeV, etas_slope_lo, etas_slope_hi, T_reg_GeV;
    double T_F_GeV;
    int coarse;                /* steps per round (the freezeout check interval) */
    double e0, e_floor;        /* MC-Glauber normalisation and floor (Au+Au, b uniform in [0, 10] fm) */
    int64_t hbm_budget_bytes;  /* device bytes for resident state */
    int precision;             /* vh_precision */
    int bulk;
    int max_steps;             /* per request */
    int max_resident;          /* event slots */
    int collect_surface;       /* 1: flag freezeout cubes of every resident request at every check;
                                  2: also build their surface elements on the device (vh_batch_fo_surface)
                                  and hand each request's to on_surface; the next check's steps then run
                                  while the surfaces of requests that stay are built, and a request
                                  admitted meanwhile joins the steps after them */
    int device;
    int digest;                /* 1: a completed request's final state is downloaded and hashed (digest) */
    /* with collect_surface 2 (may be NULL): called for each resident request
     * at each check with that check's n surface elements, 20 doubles each in
     * surface.dat's columns (FreezeOut.cpp:243-270: tau x y eta, dsigma_mu
     * = t * normal, u^x u^y u^eta e T P pi^xx pi^xy pi^xeta pi^yy pi^yeta
     * Pi), cubes in the reference's order.  t is the request's state time
     * t0 + steps * dt and an element's tau is measured from t + dt, as the
     * reference's are.  The rows are valid during the call only. */
    void (*on_surface)(void *user, int request, int64_t n, const double *rows);
    void *user;
} vh_serve_config;

typedef struct {
    double arrival_s, admit_s, done_s;  /* seconds from the start of the run */
    int nx, ny, nz;
    int steps;                          /* steps taken */
    int status;                         /* VH_OK, VH_ERR_NAN (blew up), VH_ERR_CAPACITY (never fits) */
    uint64_t glauber_seed;
    int64_t fo_cubes;                   /* freezeout cubes flagged, summed over checks */
    double e_sum;                       /* sum of e over the final state (identity checks) */
    uint64_t digest;                    /* with cfg.digest: hash of the final state's bits, 0 otherwise:
                                         * h = 14695981039346656037; for each field f of vh_download's
                                         * numbering 0..nc-1, 15, 16, 17..20: h = h * 1099511628211 +
                                         * sum_i bits(x_i) * (2i + 1), all mod 2^64 */
    int64_t fo_elements;                /* with collect_surface 2: surface elements, summed over checks */
} vh_serve_request;

typedef struct {
    int completed, failed;
    double wall_s, events_per_s, zone_upd_per_s;
    double p50_s, p99_s, mean_s;        /* latency: done - arrival, completed requests */
    double occupancy_mean, occupancy_peak;  /* resident state bytes / budget */
} vh_serve_stats;

/* eos_table: NULL for the conformal EoS; requests[] gets n_requests entries. */
vh_status vh_serve_run(const vh_serve_config *cfg, const double *eos_table, vh_serve_stats *stats,
                       vh_serve_request *requests);

const char *vh_status_string(vh_status s);
const char *vh_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* CPUVH_AMD_H */
