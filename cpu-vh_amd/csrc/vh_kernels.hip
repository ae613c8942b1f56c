// gfx950 kernels of the viscous-hydro step and their launchers.
#include "vh_kernels.hpp"

namespace vh {

// setConservedVariables, rhic/src/DynamicalVariables.cpp:169-218, with
// Ttt/Ttx/Tty/Ttn of rhic/src/EnergyMomentumTensor.cpp:292-306 (p -> p + Pi).
__global__ __launch_bounds__(256) void init_conserved_kernel(double *__restrict__ ttt, double *__restrict__ ttx,
                                                             double *__restrict__ tty, double *__restrict__ ttn,
                                                             const double *__restrict__ pitt,
                                                             const double *__restrict__ pitx,
                                                             const double *__restrict__ pity,
                                                             const double *__restrict__ pitn,
                                                             const double *__restrict__ Pi,
                                                             const double *__restrict__ e,
                                                             const double *__restrict__ p,
                                                             const double *__restrict__ ut,
                                                             const double *__restrict__ ux,
                                                             const double *__restrict__ uy,
                                                             const double *__restrict__ un, long N) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const double Pi_s = Pi ? Pi[s] : 0.;
    const double e_s = e[s], pp = p[s] + Pi_s, ut_s = ut[s];
    ttt[s] = (e_s + pp) * ut_s * ut_s - pp + pitt[s];
    ttx[s] = (e_s + pp) * ut_s * ux[s] + pitx[s];
    tty[s] = (e_s + pp) * ut_s * uy[s] + pity[s];
    ttn[s] = (e_s + pp) * ut_s * un[s] + pitn[s];
}

// checkForCellsAboveTc (freezeout/FreezeOut.cpp:364-381) as a wavefront
// reduction: one 64-bit atomic per wave.
__global__ __launch_bounds__(256) void count_above_kernel(const double *__restrict__ e, long N, double eF,
                                                          unsigned long long *__restrict__ out) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool above = s < N && e[s] > eF;
    const unsigned long long mask = __ballot(above);
    if ((threadIdx.x & 63) == 0 && mask) atomicAdd(out, (unsigned long long)__popcll(mask));
}

// Freezeout cube flags (the crossing test of cornelius.cpp:2201-2212 /
// :2257-2272 on the cubes FreezeOut.cpp:207-212 / :288-291 visits), with
// wave-aggregated append of the flagged cube ids.
__global__ __launch_bounds__(256) void fo_flag_kernel(const double *__restrict__ eprev,
                                                      const double *__restrict__ ecur, int nx, int ny, int nz,
                                                      double eF, long long *__restrict__ ids, long long cap,
                                                      unsigned long long *__restrict__ count) {
    const bool d4 = nz > 1;
    const long mx = nx - 1, my = ny - 1, mz = d4 ? nz - 1 : 1;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool flag = false;
    if (c < mx * my * mz) {
        const long ix = c % mx, iy = (c / mx) % my, iz = c / (mx * my);
        const long base = ix + (long)nx * (iy + (long)ny * iz);
        const long sx = 1, sy = nx, sz = (long)nx * ny;
        int above = 0, tot = 0;
        for (int dz = 0; dz <= (d4 ? 1 : 0); ++dz)
            for (int dy = 0; dy <= 1; ++dy)
                for (int dx = 0; dx <= 1; ++dx) {
                    const long s = base + dx * sx + dy * sy + dz * sz;
                    above += (eprev[s] >= eF) + (ecur[s] >= eF);
                    tot += 2;
                }
        flag = above != 0 && above != tot;
    }
    const unsigned long long mask = __ballot(flag);
    const int lane = threadIdx.x & 63;
    unsigned long long basepos = 0;
    if (lane == 0 && mask) basepos = atomicAdd(count, (unsigned long long)__popcll(mask));
    basepos = __shfl(basepos, 0);
    if (flag) {
        const unsigned long long below = mask & ((1ull << lane) - 1ull);
        const unsigned long long pos = basepos + __popcll(below);
        if ((long long)pos < cap) ids[pos] = c;
    }
}

// Corner energies of the flagged cubes for the host surface builder.
__global__ __launch_bounds__(256) void fo_corners_kernel(const double *__restrict__ eprev,
                                                         const double *__restrict__ ecur, int nx, int ny, int nz,
                                                         const long long *__restrict__ ids, long long n,
                                                         double *__restrict__ corners) {
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const bool d4 = nz > 1;
    const long mx = nx - 1, my = ny - 1;
    const long c = ids[m];
    const long ix = c % mx, iy = (c / mx) % my, iz = c / (mx * my);
    const long base = ix + (long)nx * (iy + (long)ny * iz);
    const int per = d4 ? 16 : 8, half = per / 2;
    int q = 0;
    for (int dz = 0; dz <= (d4 ? 1 : 0); ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx, ++q) {
                const long s = base + dx + (long)nx * (dy + (long)ny * dz);
                corners[m * per + q] = eprev[s];
                corners[m * per + half + q] = ecur[s];
            }
}

template <int NC, bool TABLE, bool DIM3>
static hipError_t launch_stage_t(const StageParams &P, const EosDev &E, hipStream_t st) {
    const long N = (long)P.nx * P.ny * P.nz;
    const unsigned grid = (unsigned)((N + CELLS_PER_BLOCK - 1) / CELLS_PER_BLOCK);
    const unsigned threads = (DIM3 ? 4 : 3) * CELLS_PER_BLOCK;
    hipLaunchKernelGGL((stage_kernel<NC, TABLE, DIM3>), dim3(grid), dim3(threads), 0, st, P, E);
    return hipGetLastError();
}

hipError_t launch_stage(const StageParams &P, const EosDev &E, int nc, bool dim3d, hipStream_t st) {
    const bool tab = E.kind == 1;
    if (nc == 14) {
        if (tab) return dim3d ? launch_stage_t<14, true, true>(P, E, st) : launch_stage_t<14, true, false>(P, E, st);
        return dim3d ? launch_stage_t<14, false, true>(P, E, st) : launch_stage_t<14, false, false>(P, E, st);
    }
    if (tab) return dim3d ? launch_stage_t<15, true, true>(P, E, st) : launch_stage_t<15, true, false>(P, E, st);
    return dim3d ? launch_stage_t<15, false, true>(P, E, st) : launch_stage_t<15, false, false>(P, E, st);
}

hipError_t launch_init_conserved(double *const q[15], const double *Pi, const double *e, const double *p,
                                 const double *const u[4], long N, hipStream_t st) {
    const unsigned grid = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(init_conserved_kernel, dim3(grid), dim3(256), 0, st, q[0], q[1], q[2], q[3], q[4], q[5], q[6],
                       q[7], Pi, e, p, u[0], u[1], u[2], u[3], N);
    return hipGetLastError();
}

hipError_t launch_count_above(const double *e, long N, double eF, unsigned long long *out, hipStream_t st) {
    const unsigned grid = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(count_above_kernel, dim3(grid), dim3(256), 0, st, e, N, eF, out);
    return hipGetLastError();
}

hipError_t launch_fo_flag(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                          long long *ids, long long cap, unsigned long long *count, hipStream_t st) {
    const long M = (long)(nx - 1) * (ny - 1) * (nz > 1 ? nz - 1 : 1);
    if (M <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((M + 255) / 256);
    hipLaunchKernelGGL(fo_flag_kernel, dim3(grid), dim3(256), 0, st, eprev, ecur, nx, ny, nz, eF, ids, cap, count);
    return hipGetLastError();
}

hipError_t launch_fo_corners(const double *eprev, const double *ecur, int nx, int ny, int nz, const long long *ids,
                             long long n, double *corners, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(fo_corners_kernel, dim3(grid), dim3(256), 0, st, eprev, ecur, nx, ny, nz, ids, n, corners);
    return hipGetLastError();
}

}  // namespace vh
