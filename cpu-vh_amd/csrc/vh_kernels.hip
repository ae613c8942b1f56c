// gfx950 kernels of the viscous-hydro step and their launchers.
#include <algorithm>
#include <vector>

#include "vh_kernels.hpp"
#include "vh_fo_interp.hpp"

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

// One batched event's admission in one pass: the uploaded primitives (a NULL
// array is zero, u^tau one) into the event's cells of e, p, u and the
// previous u, pi and Pi into q, and T^{tau mu} formed as init_conserved_kernel
// forms them; the first thread also opens the event's stop-rule slots.
__global__ __launch_bounds__(256) void admit_fp64_kernel(const AdmitSrc S, double *__restrict__ q,
                                                         double *__restrict__ e0, double *__restrict__ p0,
                                                         double *__restrict__ uc, double *__restrict__ up, long cap,
                                                         long N, long off, int nc, int *__restrict__ strikes,
                                                         int *__restrict__ stop) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) {
        *strikes = 0;
        *stop = -1;
    }
    if (c >= N) return;
    const long s = off + c;
    const double e_s = S.e[c], p_s = S.p[c];
    double u[4];
    for (int k = 0; k < 4; ++k) u[k] = S.u[k] ? S.u[k][c] : (k == 0 ? 1.0 : 0.0);
    double pi[10];
    for (int n = 0; n < 10; ++n) pi[n] = S.pi10 ? S.pi10[n * N + c] : 0.0;
    const double Pi_s = (nc == 15 && S.Pi) ? S.Pi[c] : 0.0;
    e0[s] = e_s;
    p0[s] = p_s;
    for (int k = 0; k < 4; ++k) {
        uc[k * cap + s] = u[k];
        up[k * cap + s] = u[k];
    }
    for (int n = 0; n < 10; ++n) q[(4 + n) * cap + s] = pi[n];
    if (nc == 15) q[14 * cap + s] = Pi_s;
    const double pp = p_s + Pi_s;
    q[s] = (e_s + pp) * u[0] * u[0] - pp + pi[0];
    q[cap + s] = (e_s + pp) * u[0] * u[1] + pi[1];
    q[2 * cap + s] = (e_s + pp) * u[0] * u[2] + pi[2];
    q[3 * cap + s] = (e_s + pp) * u[0] * u[3] + pi[3];
}

hipError_t launch_admit_fp64(const AdmitSrc &S, double *q, double *e0, double *p0, double *uc, double *up, long cap,
                             long N, long off, int nc, int *strikes, int *stop, hipStream_t st) {
    if (N <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(admit_fp64_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, S, q, e0, p0, uc, up,
                       cap, N, off, nc, strikes, stop);
    return hipGetLastError();
}

// checkForCellsAboveTc (freezeout/FreezeOut.cpp:364-381) as a wavefront
// reduction: one 64-bit atomic per wave.
// A bounded grid strides over the cells; each wave sums its ballots, the
// block its four waves in LDS, and one atomic per block adds the total (an
// atomic per wave on the one counter serialises and sets the time).
constexpr int COUNT_BLOCKS = 1024;
__global__ __launch_bounds__(256) void count_above_kernel(const double *__restrict__ e, long N, double eF,
                                                          unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[4];
    unsigned long long n = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long base = (long)blockIdx.x * blockDim.x; base < N; base += stride) {  // the same trip count block-wide
        const long s = base + threadIdx.x;
        const bool above = s < N && e[s] > eF;
        n += (unsigned long long)__popcll(__ballot(above));
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(out, t);
    }
}

// Freezeout cube flags: the crossing test of cornelius.cpp:2201-2212 /
// :2257-2272 on the cubes FreezeOut.cpp:207-212 / :288-291 visits (cube c
// below (nx-1)(ny-1)(nz-1 or 1), x fastest).
__device__ __forceinline__ bool fo_cut(const double *__restrict__ eprev, const double *__restrict__ ecur, int nx,
                                       int ny, int nz, double eF, long c) {
    const bool d4 = nz > 1;
    const long mx = nx - 1, my = ny - 1;
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
    return above != 0 && above != tot;
}

// The flagged ids in increasing order without atomics (an append with a
// returning atomic per wave serialises on the one counter): pass 1 counts
// the cut cubes of each block of 256, pass 2 turns the counts into offsets,
// pass 3 recomputes the flags and writes each id at its block's offset plus
// its rank in the block.
__global__ __launch_bounds__(256) void fo_flag_count_kernel(const double *__restrict__ eprev,
                                                            const double *__restrict__ ecur, int nx, int ny, int nz,
                                                            double eF, long M, unsigned long long *__restrict__ blk) {
    __shared__ unsigned int part[4];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const bool flag = c < M && fo_cut(eprev, ecur, nx, ny, nz, eF, c);
    const unsigned long long mask = __ballot(flag);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = (unsigned int)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = (unsigned long long)part[0] + part[1] + part[2] + part[3];
}

// exclusive prefix sum of blk[0 .. nb) in place and the total into *count;
// one block of 1024 threads walks the counts in chunks of 1024
__global__ __launch_bounds__(1024) void fo_flag_scan_kernel(unsigned long long *__restrict__ blk, long nb,
                                                           unsigned long long *__restrict__ count) {
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long carry_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (long base = 0; base < nb; base += 1024) {
        const long i = base + threadIdx.x;
        const unsigned long long v = i < nb ? blk[i] : 0ull;
        unsigned long long x = v;  // inclusive scan within the wave
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned long long y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned long long woff = 0;
        for (int k = 0; k < w; ++k) woff += wsum[k];
        const unsigned long long carry = carry_s;
        if (i < nb) blk[i] = carry + woff + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + woff + x;  // the chunk's total added
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = carry_s;
}

__global__ __launch_bounds__(256) void fo_flag_write_kernel(const double *__restrict__ eprev,
                                                            const double *__restrict__ ecur, int nx, int ny, int nz,
                                                            double eF, long M,
                                                            const unsigned long long *__restrict__ blk,
                                                            long long *__restrict__ ids, long long cap) {
    __shared__ unsigned int part[4];
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const bool flag = c < M && fo_cut(eprev, ecur, nx, ny, nz, eF, c);
    const unsigned long long mask = __ballot(flag);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = (unsigned int)__popcll(mask);
    __syncthreads();
    unsigned long long off = blk[blockIdx.x];
    for (int k = 0; k < w; ++k) off += part[k];
    if (flag) {
        const unsigned long long pos = off + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
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

// interpolateVariable3D/4D (FreezeOut.cpp:182-198) of the ten freezeout
// variables at points inside cubes: one thread per (point, variable), so the
// output vars[10 m + v] is written contiguously.  cubes[] were checked to lie
// in [0, (nx-1)(ny-1)(nz-1 or 1)) by the caller.
__global__ __launch_bounds__(256) void fo_interp_kernel(FoSlices S, int nx, int ny, int nz,
                                                        const long long *__restrict__ cubes,
                                                        const double *__restrict__ frac, long long n,
                                                        double *__restrict__ vars) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * FO_NVARS) return;
    const long long m = g / FO_NVARS;
    const int v = (int)(g - m * FO_NVARS);
    const long mx = nx - 1, my = ny - 1;
    const long c = cubes[m];
    const long ix = c % mx, iy = (c / mx) % my, iz = c / (mx * my);
    const long sy = nx, sz = (long)nx * ny;
    const long base = ix + sy * iy + sz * iz;
    const double *f = frac + 4 * m;
    const double *p0 = S.v[0][v], *p1 = S.v[1][v];
    if (nz > 1) {
        double a[2][2][2][2];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y)
                for (int z = 0; z < 2; ++z) {
                    const long s = base + x + sy * y + sz * z;
                    a[0][x][y][z] = p0[s];
                    a[1][x][y][z] = p1[s];
                }
        vars[g] = fo_lin4(f[0], f[1], f[2], f[3], a);
    } else {
        double a[2][2][2];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) {
                const long s = base + x + sy * y;
                a[0][x][y] = p0[s];
                a[1][x][y] = p1[s];
            }
        vars[g] = fo_lin3(f[0], f[1], f[2], a);
    }
}

// calculateThermalVorticity (rhic/src/Vorticity.cpp:6-171), called after
// each step (HydroPlugin.cpp:302-305): beta_mu = u_mu / T with
// u_mu = (u^t, -u^x, -u^y, -u^eta) and T the centre cell's (the neighbours'
// beta are divided by it as well, as the reference does), the time
// derivative backwards over dt from the previous step's u, the spatial ones
// central with the neighbours clamped at the edges (the reference's ghost
// copies), w_{mu nu} = -(d_mu beta_nu - d_nu beta_mu) / 2; the operations in
// the reference's order, so the results are its bit for bit.
template <bool TABLE>
__global__ __launch_bounds__(256) void vorticity_kernel(VortArgs A, EosDev E) {
    const long N = (long)A.nx * A.ny * A.nz;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int i = (int)(s % A.nx), j = (int)((s / A.nx) % A.ny), k = (int)(s / ((long)A.nx * A.ny));
    const long sy = A.nx, sz = (long)A.nx * A.ny;
    const long nb[6] = {s + ((i > 0) ? -1 : 0),          s + ((i < A.nx - 1) ? 1 : 0),
                        s + ((j > 0) ? -sy : 0),         s + ((j < A.ny - 1) ? sy : 0),
                        s + ((k > 0) ? -sz : 0),         s + ((k < A.nz - 1) ? sz : 0)};
    DivExact dv;
    const double T = eos_temperature<TABLE>(E, A.e[s], dv);
    double b[4], bp[4], bn[6][4];
    b[0] = A.u[0][s] / T;
    bp[0] = A.up[0][s] / T;
    for (int c = 1; c < 4; ++c) {
        b[c] = (A.u[c][s] * -1.0) / T;
        bp[c] = (A.up[c][s] * -1.0) / T;
    }
    for (int m = 0; m < 6; ++m) {
        bn[m][0] = A.u[0][nb[m]] / T;
        for (int c = 1; c < 4; ++c) bn[m][c] = (A.u[c][nb[m]] * -1.0) / T;
    }
    double dtb[4], dxb[4], dyb[4], dnb[4];
    for (int c = 0; c < 4; ++c) {
        dtb[c] = (b[c] - bp[c]) / A.dt;
        dxb[c] = (bn[1][c] - bn[0][c]) / (2.0 * A.dx);
        dyb[c] = (bn[3][c] - bn[2][c]) / (2.0 * A.dy);
        dnb[c] = (bn[5][c] - bn[4][c]) / (2.0 * A.dz);
    }
    double *o = A.out;
    o[0 * N + s] = (-1.0 / 2.0) * (dtb[1] - dxb[0]);
    o[1 * N + s] = (-1.0 / 2.0) * (dtb[2] - dyb[0]);
    o[2 * N + s] = (-1.0 / 2.0) * (dtb[3] - dnb[0]);
    o[3 * N + s] = (-1.0 / 2.0) * (dxb[2] - dyb[1]);
    o[4 * N + s] = (-1.0 / 2.0) * (dxb[3] - dnb[1]);
    o[5 * N + s] = (-1.0 / 2.0) * (dyb[3] - dnb[2]);
    for (int c = 0; c < 4; ++c) o[(6 + c) * N + s] = b[c];
}

hipError_t launch_vorticity(const VortArgs &A, const EosDev &E, hipStream_t st) {
    const long N = (long)A.nx * A.ny * A.nz;
    const unsigned grid = (unsigned)((N + 255) / 256);
    if (E.kind == 1) hipLaunchKernelGGL(vorticity_kernel<true>, dim3(grid), dim3(256), 0, st, A, E);
    else hipLaunchKernelGGL(vorticity_kernel<false>, dim3(grid), dim3(256), 0, st, A, E);
    return hipGetLastError();
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

template <int NC, bool TABLE, bool DIM3>
static hipError_t launch_stage_batch_t(const StageParams &P, const EosDev &E, const EvDesc *ev, const int *blk_ev,
                                       const int *blk_first, int nblocks, hipStream_t st) {
    const unsigned threads = (DIM3 ? 4 : 3) * CELLS_PER_BLOCK;
    hipLaunchKernelGGL((stage_kernel_batch<NC, TABLE, DIM3>), dim3(nblocks), dim3(threads), 0, st, P, E, ev, blk_ev,
                       blk_first);
    return hipGetLastError();
}

hipError_t launch_stage_batch(const StageParams &P, const EosDev &E, int nc, bool dim3d, const EvDesc *ev,
                              const int *blk_ev, const int *blk_first, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    const bool tab = E.kind == 1;
#define VH_SB(NC_, T_, D_) launch_stage_batch_t<NC_, T_, D_>(P, E, ev, blk_ev, blk_first, nblocks, st)
    if (nc == 14) {
        if (tab) return dim3d ? VH_SB(14, true, true) : VH_SB(14, true, false);
        return dim3d ? VH_SB(14, false, true) : VH_SB(14, false, false);
    }
    if (tab) return dim3d ? VH_SB(15, true, true) : VH_SB(15, true, false);
    return dim3d ? VH_SB(15, false, true) : VH_SB(15, false, false);
#undef VH_SB
}

// Per-event count of cells with e > eF over a table of 64-cell chunks: chunk b
// covers cells [start[b], start[b] + 64) of event ev_of[b].
__global__ __launch_bounds__(64) void count_above_batch_kernel(const double *__restrict__ e, double eF,
                                                               const EvDesc *__restrict__ ev,
                                                               const int *__restrict__ ev_of,
                                                               const long *__restrict__ start,
                                                               unsigned long long *__restrict__ counts) {
    const int id = ev_of[blockIdx.x];
    const EvDesc &D = ev[id];
    if (!D.active) return;
    const long N = (long)D.nx * D.ny * D.nz;
    const long sl = start[blockIdx.x] + threadIdx.x;
    const bool above = sl < N && e[D.off + sl] > eF;
    const unsigned long long mask = __ballot(above);
    if (threadIdx.x == 0 && mask) atomicAdd(&counts[id], (unsigned long long)__popcll(mask));
}

hipError_t launch_count_above_batch(const double *e, double eF, const EvDesc *ev, const int *ev_of,
                                    const long *start, int nchunks, unsigned long long *counts, hipStream_t st) {
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_above_batch_kernel, dim3(nchunks), dim3(64), 0, st, e, eF, ev, ev_of, start, counts);
    return hipGetLastError();
}

// Zero of the per-event counts inside a captured graph (a fill kernel: no
// memset node in a graph).
__global__ __launch_bounds__(64) void zero_counts_kernel(unsigned long long *__restrict__ counts, int n) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k < n) counts[k] = 0;
}

hipError_t launch_zero_counts(unsigned long long *counts, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_counts_kernel, dim3((n + 63) / 64), dim3(64), 0, st, counts, n);
    return hipGetLastError();
}

// The reference's stop rule on the device (rhic/src/HydroPlugin.cpp:266-281):
// an active event whose count of cells above e_F is zero gets a strike (the
// strikes are never reset); at the third it stops (inactive: its blocks
// return at once) and the check is recorded.
__global__ __launch_bounds__(64) void strike_kernel(EvDesc *__restrict__ ev, const unsigned long long *__restrict__ counts,
                                                    int *__restrict__ strikes, int *__restrict__ stop_check, int n,
                                                    int check) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= n || !ev[k].active || counts[k] != 0) return;
    if (++strikes[k] >= 3) {
        ev[k].active = 0;
        stop_check[k] = check;
    }
}

hipError_t launch_strikes(EvDesc *ev, const unsigned long long *counts, int *strikes, int *stop_check, int n,
                          int check, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(strike_kernel, dim3((n + 63) / 64), dim3(64), 0, st, ev, counts, strikes, stop_check, n, check);
    return hipGetLastError();
}

// Every event's clock moves on by count steps (the last node of a batch graph).
__global__ __launch_bounds__(64) void advance_clocks_kernel(EvDesc *__restrict__ ev, int n, int count) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k < n) ev[k].base += count;
}

// getInferredVariables (rhic/src/EnergyMomentumTensor.cpp:117-233) on
// independent states, one thread each, through the same recover() the stage
// kernel applies to its cell.
template <bool TABLE>
__global__ __launch_bounds__(256) void inferred_kernel(const EosDev E, int nc, double t, const double *__restrict__ q,
                                                       const double *__restrict__ ePrev, long n,
                                                       double *__restrict__ out, int *__restrict__ nan_out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = q[k * n + i];
    const double Pi = nc == 15 ? q[14 * n + i] : 0.0;
    DivFast div;
    int nan = NAN_NONE;
    const Prim w = recover<TABLE>(E, t, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], Pi, ePrev[i], nan, div);
    out[i] = w.e;
    out[n + i] = w.p;
    out[2 * n + i] = w.ut;
    out[3 * n + i] = w.ux;
    out[4 * n + i] = w.uy;
    out[5 * n + i] = w.un;
    nan_out[i] = nan;
}

hipError_t launch_inferred(const EosDev &E, int nc, double t, const double *q, const double *ePrev, long n,
                           double *out, int *nan, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (E.kind == 1)
        hipLaunchKernelGGL(inferred_kernel<true>, dim3(grid), dim3(256), 0, st, E, nc, t, q, ePrev, n, out, nan);
    else
        hipLaunchKernelGGL(inferred_kernel<false>, dim3(grid), dim3(256), 0, st, E, nc, t, q, ePrev, n, out, nan);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_kernel(double *__restrict__ out, long n, double v) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = v;
}

hipError_t launch_fill(double *out, long n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n, v);
    return hipGetLastError();
}

// The single-event step clock moves on by count steps (last node of a graph).
__global__ void advance_base_kernel(int *__restrict__ base, int count) {
    if (threadIdx.x == 0) *base += count;
}

hipError_t launch_advance_base(int *base, int count, hipStream_t st) {
    hipLaunchKernelGGL(advance_base_kernel, dim3(1), dim3(64), 0, st, base, count);
    return hipGetLastError();
}

hipError_t launch_advance_clocks(EvDesc *ev, int n, int count, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(advance_clocks_kernel, dim3((n + 63) / 64), dim3(64), 0, st, ev, n, count);
    return hipGetLastError();
}

hipError_t launch_init_conserved(double *const q[15], const double *Pi, const double *e, const double *p,
                                 const double *const u[4], long N, hipStream_t st) {
    const unsigned grid = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(init_conserved_kernel, dim3(grid), dim3(256), 0, st, q[0], q[1], q[2], q[3], q[4], q[5], q[6],
                       q[7], Pi, e, p, u[0], u[1], u[2], u[3], N);
    return hipGetLastError();
}

hipError_t launch_count_above(const double *e, long N, double eF, unsigned long long *out, hipStream_t st) {
    if (N <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<long>((N + 255) / 256, COUNT_BLOCKS);
    hipLaunchKernelGGL(count_above_kernel, dim3(grid), dim3(256), 0, st, e, N, eF, out);
    return hipGetLastError();
}

hipError_t launch_fo_flag(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                          long long *ids, long long cap, unsigned long long *count, unsigned long long *blk,
                          hipStream_t st) {
    const long M = (long)(nx - 1) * (ny - 1) * (nz > 1 ? nz - 1 : 1);
    if (M <= 0) return hipSuccess;
    const long nb = fo_flag_blocks(M);
    hipLaunchKernelGGL(fo_flag_count_kernel, dim3((unsigned)nb), dim3(256), 0, st, eprev, ecur, nx, ny, nz, eF, M,
                       blk);
    hipLaunchKernelGGL(fo_flag_scan_kernel, dim3(1), dim3(1024), 0, st, blk, nb, count);
    hipLaunchKernelGGL(fo_flag_write_kernel, dim3((unsigned)nb), dim3(256), 0, st, eprev, ecur, nx, ny, nz, eF, M,
                       (const unsigned long long *)blk, ids, cap);
    return hipGetLastError();
}

hipError_t launch_fo_corners(const double *eprev, const double *ecur, int nx, int ny, int nz, const long long *ids,
                             long long n, double *corners, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(fo_corners_kernel, dim3(grid), dim3(256), 0, st, eprev, ecur, nx, ny, nz, ids, n, corners);
    return hipGetLastError();
}

hipError_t launch_fo_interp(const FoSlices &S, int nx, int ny, int nz, const long long *cubes, const double *frac,
                            long long n, double *vars, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n * FO_NVARS + 255) / 256);
    hipLaunchKernelGGL(fo_interp_kernel, dim3(grid), dim3(256), 0, st, S, nx, ny, nz, cubes, frac, n, vars);
    return hipGetLastError();
}

hipError_t DeviceEos::init(const HostEos &h, int kind, int cs2_mode) {
    E = EosDev{};
    E.kind = kind;
    hipError_t r;
    if (kind == 1) {
        E.cs2_mode = cs2_mode;
        E.cs2_const = cs2_mode == 0 && h.cs2_is_one_third() ? 1 : 0;
        E.e_min = h.e_min;
        E.e_spacing = h.e_spacing;
        E.r_spacing = 1.0 / h.e_spacing;
        E.e_max = h.e_max;
        const size_t tb = sizeof(double) * h.p_tb.size();
        if ((r = hipMalloc(&ptb, tb)) != hipSuccess || (r = hipMalloc(&Ttb, tb)) != hipSuccess) return r;
        if ((r = hipMemcpy(ptb, h.p_tb.data(), tb, hipMemcpyHostToDevice)) != hipSuccess ||
            (r = hipMemcpy(Ttb, h.T_tb.data(), tb, hipMemcpyHostToDevice)) != hipSuccess)
            return r;
        E.p_tb = ptb;
        E.T_tb = Ttb;
        const size_t n = h.p_tb.size();
        std::vector<float4> seg(n);
        for (size_t k = 0; k < n; ++k) {
            const size_t a = std::min(k, n - 2);  // past the last segment its slope continues
            const double d = h.p_tb[a + 1] - h.p_tb[a];
            const size_t b = std::min(k + 1, n - 2);
            const double d1 = h.p_tb[b + 1] - h.p_tb[b];
            seg[k] = make_float4((float)h.p_tb[k], (float)d, (float)(d * E.r_spacing),
                                 (float)((d1 - d) * E.r_spacing));
        }
        if ((r = hipMalloc(&p4, sizeof(float4) * n)) != hipSuccess ||
            (r = hipMemcpy(p4, seg.data(), sizeof(float4) * n, hipMemcpyHostToDevice)) != hipSuccess)
            return r;
        E.p4 = p4;
    } else {
        const QuarterRootExceptions &X = quarter_root_exceptions();
        std::vector<uint2> pairs(std::max<size_t>(1, X.key.size()));
        for (size_t i = 0; i < X.key.size(); ++i) pairs[i] = make_uint2(X.key[i], X.val[i]);
        if ((r = hipMalloc(&exc, sizeof(uint2) * pairs.size())) != hipSuccess ||
            (r = hipMalloc(&bucket, sizeof(uint32_t) * X.bucket.size())) != hipSuccess)
            return r;
        if ((r = hipMemcpy(exc, pairs.data(), sizeof(uint2) * pairs.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (r = hipMemcpy(bucket, X.bucket.data(), sizeof(uint32_t) * X.bucket.size(), hipMemcpyHostToDevice)) !=
                hipSuccess)
            return r;
        E.exc = exc;
        E.exc_bucket = bucket;
        E.exc_lo = X.lo;
        E.exc_hi = X.hi;
        E.exc_shift = X.shift;
    }
    return hipSuccess;
}

void DeviceEos::release() {
    (void)hipFree(ptb);
    (void)hipFree(Ttb);
    (void)hipFree(p4);
    (void)hipFree(exc);
    (void)hipFree(bucket);
    ptb = Ttb = nullptr;
    p4 = nullptr;
    exc = nullptr;
    bucket = nullptr;
}

}  // namespace vh
