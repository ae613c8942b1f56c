// Test infrastructure only: a C-ABI driver around the reference's own hot-path
// functions, compiled from the reference sources where they lie (see
// oracle/ref.mk).  Nothing in the product links or loads this file; tests,
// the smoke check and bench.py's CPU baseline use the resulting
// oracle/_ref/libcpuvh_ref_*.so as the checker / the timed reference.
//
// The driver reproduces the state set-up of the type-11 initial condition
// reader (rhic/src/InitialConditions.cpp:41-105: e, p, u, up = u, pi, Pi from
// interior arrays), then calls the reference's setConservedVariables
// (rhic/src/DynamicalVariables.cpp:169-218), setGhostCells (:220-227),
// rungeKutta2 (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:866-924),
// setCurrentConservedVariables (rhic/src/DynamicalVariables.cpp:386-388),
// calculateThermalVorticity (rhic/src/Vorticity.cpp:6-171, called after each
// step at rhic/src/HydroPlugin.cpp:302-305), checkForCellsAboveTc and the
// freezeout finder (freezeout/FreezeOut.cpp).
//
// Field numbering (shared with the oracle and the product's C-ABI):
//   0..13 ttt ttx tty ttn pitt pitx pity pitn pixx pixy pixn piyy piyn pinn,
//   14 Pi, 15 e, 16 p, 17..20 ut ux uy un, 21..24 up (previous-step u),
//   25 regFacShear, 27..32 the thermal vorticity wtx wty wtn wxy wxn wyn,
//   33..36 the thermal velocity beta_t beta_x beta_y beta_n.
// Interior arrays are x-fastest: idx = i + nx*(j + ny*k).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>
#include <unistd.h>

#include "DynamicalVariables.h"
#include "LatticeParameters.h"
#include "HydroParameters.h"
#include "EquationOfState.h"
#include "FullyDiscreteKurganovTadmorScheme.h"
#include "EnergyMomentumTensor.h"
#include "FreezeOut.h"
#include "Vorticity.h"

// freezeout/FreezeOut.cpp and freezeout/Memory.cpp (compiled in ref_fo_unit.cpp)
double ****calloc4dArray(double ****array, int dim1, int dim2, int dim3, int dim4);
double *****calloc5dArray(double *****array, int dim1, int dim2, int dim3, int dim4, int dim5);
void free4dArray(double ****array, int dim1, int dim2, int dim3);
void free5dArray(double *****array, int dim1, int dim2, int dim3, int dim4);
double ***calloc3dArray(double ***array, int dim1, int dim2, int dim3);
void swapAndSetHydroVariables(double ****, double *****, CONSERVED_VARIABLES *const __restrict__,
                              PRECISION *const __restrict__, FLUID_VELOCITY *const __restrict__, int, int, int);
void callFOFinder3p1D(int dim, int nx, int ny, int nz, int n, double t0, double dt, double t, double dx,
                      double dy, double dz, double *lattice_spacing, double freezeoutEnergyDensity,
                      double ****hyperCube4D, double ***hyperCube3D, double ****energy_density_evoution,
                      double *****hydrodynamic_evoution, std::ofstream &freezeoutSurfaceFile,
                      std::vector<FO_Element> &fo_surf, EOS eqnOfState);
void callFOFinder2p1D(int dim, int nx, int ny, int nz, int n, double t0, double dt, double t, double dx,
                      double dy, double dz, double *lattice_spacing, double freezeoutEnergyDensity,
                      double ****hyperCube4D, double ***hyperCube3D, double ****energy_density_evoution,
                      double *****hydrodynamic_evoution, std::ofstream &freezeoutSurfaceFile,
                      std::vector<FO_Element> &fo_surf, EOS eqnOfState);
int checkForCellsAboveTc(int nx, int ny, int nz, double freezeoutEnergyDensity, PRECISION *e);

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
LatticeParameters g_lat;
HydroParameters g_hydro;
int g_nx, g_ny, g_nz, g_ncx, g_ncy, g_ncz;
bool g_alloc = false;

inline int sidx(int i, int j, int k) { return (i + 2) + g_ncx * ((j + 2) + g_ncy * (k + 2)); }

PRECISION *field_ptr(int f) {
    switch (f) {
    case 0: return q->ttt;  case 1: return q->ttx;  case 2: return q->tty;  case 3: return q->ttn;
    case 4: return q->pitt; case 5: return q->pitx; case 6: return q->pity; case 7: return q->pitn;
    case 8: return q->pixx; case 9: return q->pixy; case 10: return q->pixn; case 11: return q->piyy;
    case 12: return q->piyn; case 13: return q->pinn;
#ifdef PI
    case 14: return q->Pi;
#endif
    case 15: return e; case 16: return p;
    case 17: return u->ut; case 18: return u->ux; case 19: return u->uy; case 20: return u->un;
    case 21: return up->ut; case 22: return up->ux; case 23: return up->uy; case 24: return up->un;
    case 25: return regFacShear;
    case 27: return wmunu->wtx; case 28: return wmunu->wty; case 29: return wmunu->wtn;
    case 30: return wmunu->wxy; case 31: return wmunu->wxn; case 32: return wmunu->wyn;
    case 33: return beta_mu->beta_t; case 34: return beta_mu->beta_x; case 35: return beta_mu->beta_y;
    case 36: return beta_mu->beta_n;
    default: return nullptr;
    }
}

// freezeout state
double ****g_eevo = nullptr;
double *****g_hevo = nullptr;
double ****g_cube4 = nullptr;
double ***g_cube3 = nullptr;
double g_lspace[4];
int g_fo_dim = 0;
std::ofstream g_surf;
std::vector<FO_Element> g_fo_vec;
}  // namespace

REF_API int ref_has_bulk(void) {
#ifdef PI
    return 1;
#else
    return 0;
#endif
}

REF_API int ref_is_conformal(void) {
#ifdef CONFORMAL_EOS
    return 1;
#else
    return 0;
#endif
}

// eos_dir: directory holding EoS/hrg_hotqcd_eos_binary.dat (lattice builds only);
// the reference reads that path relative to the working directory
// (rhic/src/EquationOfState.cpp:299).
REF_API int ref_setup(int nx, int ny, int nz, int nt, double dx, double dy, double dz, double dt, int foFactor,
                      double t0, double etabar, double TfGeV, double TregGeV, const char *eos_dir) {
    if (g_alloc) return -1;  // one lattice per process image (the reference keeps globals)
    g_nx = nx; g_ny = ny; g_nz = nz;
    g_ncx = nx + N_GHOST_CELLS; g_ncy = ny + N_GHOST_CELLS; g_ncz = nz + N_GHOST_CELLS;
    g_lat.numLatticePointsX = nx;
    g_lat.numLatticePointsY = ny;
    g_lat.numLatticePointsRapidity = nz;
    g_lat.numComputationalLatticePointsX = g_ncx;
    g_lat.numComputationalLatticePointsY = g_ncy;
    g_lat.numComputationalLatticePointsRapidity = g_ncz;
    g_lat.numProperTimePoints = nt;
    g_lat.coarseProperTimeFOFactor = foFactor;
    g_lat.latticeSpacingX = dx;
    g_lat.latticeSpacingY = dy;
    g_lat.latticeSpacingRapidity = dz;
    g_lat.latticeSpacingProperTime = dt;
    g_lat.fluxLimiterTheta = 1.1;
    g_hydro.initialProperTimePoint = t0;
    g_hydro.shearViscosityToEntropyDensity = etabar;
    g_hydro.freezeoutTemperatureGeV = TfGeV;
    g_hydro.initializePimunuNavierStokes = 0;
    g_hydro.regulationTemperatureGeV = TregGeV;
    g_hydro.doFreezeOut = 1;
#ifndef CONFORMAL_EOS
    {
        char cwd[4096];
        if (!getcwd(cwd, sizeof cwd)) return -2;
        if (!eos_dir || chdir(eos_dir) != 0) return -3;
        eqnOfState.loadEOSHotQCDHRG();
        if (chdir(cwd) != 0) return -4;
    }
#else
    (void)eos_dir;
#endif
    allocateHostMemory(g_ncx * g_ncy * g_ncz);
    g_alloc = true;
    return 0;
}

// Interior arrays; pi10 = 10 arrays back to back (pitt..pinn), Pi may be null.
// Same assignments as the type-11 reader: up = u (InitialConditions.cpp:83-86).
REF_API void ref_set_ic(const double *e_in, const double *p_in, const double *ut, const double *ux,
                        const double *uy, const double *un, const double *pi10, const double *Pi_in) {
    const long N = (long)g_nx * g_ny * g_nz;
    PRECISION *pis[10] = {q->pitt, q->pitx, q->pity, q->pitn, q->pixx, q->pixy, q->pixn, q->piyy, q->piyn, q->pinn};
    for (int k = 0; k < g_nz; ++k)
        for (int j = 0; j < g_ny; ++j)
            for (int i = 0; i < g_nx; ++i) {
                long a = i + (long)g_nx * (j + (long)g_ny * k);
                int s = sidx(i, j, k);
                e[s] = e_in[a];
                p[s] = p_in[a];
                u->ut[s] = ut[a]; u->ux[s] = ux[a]; u->uy[s] = uy[a]; u->un[s] = un[a];
                up->ut[s] = ut[a]; up->ux[s] = ux[a]; up->uy[s] = uy[a]; up->un[s] = un[a];
                for (int c = 0; c < 10; ++c) pis[c][s] = pi10[c * N + a];
#ifdef PI
                q->Pi[s] = Pi_in ? Pi_in[a] : 0.0;
#else
                (void)Pi_in;
#endif
            }
}

// HydroPlugin.cpp:202-204
REF_API void ref_init_conserved(double t) {
    setConservedVariables(t, &g_lat);
    setGhostCells(q, e, p, u, &g_lat);
}

// HydroPlugin.cpp:292,300
REF_API void ref_step(double t) {
    rungeKutta2(t, g_lat.latticeSpacingProperTime, q, Q, &g_lat, &g_hydro);
    setCurrentConservedVariables();
}

// HydroPlugin.cpp:302-305 (the fields land at 27..36 of ref_get)
REF_API void ref_vorticity(double t) {
    calculateThermalVorticity(t, g_lat.latticeSpacingProperTime, q, Q, &g_lat, &g_hydro);
}

REF_API int ref_get(int field, double *out) {
    PRECISION *src = field_ptr(field);
    if (!src) return -1;
    for (int k = 0; k < g_nz; ++k)
        for (int j = 0; j < g_ny; ++j)
            for (int i = 0; i < g_nx; ++i) out[i + (long)g_nx * (j + (long)g_ny * k)] = src[sidx(i, j, k)];
    return 0;
}

REF_API int ref_count_above(double eF) { return checkForCellsAboveTc(g_nx, g_ny, g_nz, eF, e); }

REF_API double ref_eos_energy_density(double T) { return eqnOfState.equilibriumEnergyDensity(T); }
REF_API double ref_eos_pressure(double ed) { return eqnOfState.equilibriumPressure(ed); }
REF_API double ref_eos_temperature(double ed) { return eqnOfState.effectiveTemperature(ed); }
REF_API double ref_eos_cs2(double ed) { return eqnOfState.speedOfSoundSquared(ed); }

// Primitive recovery of one cell (rhic/src/EnergyMomentumTensor.cpp:117-233).
REF_API void ref_get_inferred(double t, const double *qcell, double ePrev, double *out6) {
    getInferredVariables(t, qcell, ePrev, &out6[0], &out6[1], &out6[2], &out6[3], &out6[4], &out6[5]);
}

// Freezeout: the per-coarse-step sequence of HydroPlugin.cpp:136-182,243-263.
REF_API int ref_fo_begin(const char *surface_path) {
    if (g_nx > 1 && g_ny > 1 && g_nz > 1) g_fo_dim = 4;
    else if (g_nx > 1 && g_ny > 1 && g_nz < 2) g_fo_dim = 3;
    else return -1;
    g_lspace[0] = g_lat.latticeSpacingProperTime * g_lat.coarseProperTimeFOFactor;
    g_lspace[1] = g_lat.latticeSpacingX;
    g_lspace[2] = g_lat.latticeSpacingY;
    g_lspace[3] = g_lat.latticeSpacingRapidity;
    g_eevo = calloc4dArray(g_eevo, 2, g_nx, g_ny, g_nz);
    g_hevo = calloc5dArray(g_hevo, 16, 2, g_nx, g_ny, g_nz);
    g_cube4 = calloc4dArray(g_cube4, 2, 2, 2, 2);
    g_cube3 = calloc3dArray(g_cube3, 2, 2, 2);
    g_surf.open(surface_path);
    return g_surf.is_open() ? 0 : -2;
}

REF_API void ref_fo_call(int n, double t, double eF) {
    swapAndSetHydroVariables(g_eevo, g_hevo, q, e, u, g_nx, g_ny, g_nz);
    const double t0 = g_hydro.initialProperTimePoint, dt = g_lat.latticeSpacingProperTime;
    if (g_fo_dim == 4)
        callFOFinder3p1D(4, g_nx, g_ny, g_nz, n, t0, dt, t, g_lat.latticeSpacingX, g_lat.latticeSpacingY,
                         g_lat.latticeSpacingRapidity, g_lspace, eF, g_cube4, g_cube3, g_eevo, g_hevo, g_surf,
                         g_fo_vec, eqnOfState);
    else
        callFOFinder2p1D(3, g_nx, g_ny, g_nz, n, t0, dt, t, g_lat.latticeSpacingX, g_lat.latticeSpacingY,
                         g_lat.latticeSpacingRapidity, g_lspace, eF, g_cube4, g_cube3, g_eevo, g_hevo, g_surf,
                         g_fo_vec, eqnOfState);
}

// The previous coarse slice of e as held by the freezeout buffers (x-fastest).
REF_API void ref_fo_slice(int which, double *out) {
    for (int k = 0; k < g_nz; ++k)
        for (int j = 0; j < g_ny; ++j)
            for (int i = 0; i < g_nx; ++i) out[i + (long)g_nx * (j + (long)g_ny * k)] = g_eevo[which][i][j][k];
}

REF_API void ref_fo_end(void) {
    g_surf.close();
    free4dArray(g_eevo, 2, g_nx, g_ny);
    free5dArray(g_hevo, 16, 2, g_nx, g_ny);
    g_eevo = nullptr;
    g_hevo = nullptr;
}
