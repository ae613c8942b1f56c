// Launchers of the mixed-precision kernels (tile variants for the autotuner).
#include "vh_kernels.hpp"
#include "vh_mixed.hpp"

namespace vh {

namespace mx {

// setConservedVariables (rhic/src/DynamicalVariables.cpp:169-218) in FP64 from
// FP64 primitives, stored as the mixed form's FP32 fields.
__global__ __launch_bounds__(256) void init_mixed_kernel(const MxFields F, const AdmitSrc S, float *__restrict__ up,
                                                         long N, long off, int nc, int *__restrict__ strikes,
                                                         int *__restrict__ stop) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) {
        *strikes = 0;
        *stop = -1;
    }
    if (c >= N) return;
    const long s = off + c;
    double u[4];
    for (int k = 0; k < 4; ++k) u[k] = S.u[k] ? S.u[k][c] : (k == 0 ? 1.0 : 0.0);
    double pi[10];
    for (int n = 0; n < 10; ++n) pi[n] = S.pi10 ? S.pi10[n * N + c] : 0.0;
    const double Pb = (nc == 15 && S.Pi) ? S.Pi[c] : 0.0;
    const double e = S.e[c], p = S.p[c];
    const double P = p + Pb;
    const double w = e + P;
    double T4[4];
    T4[0] = w * u[0] * u[0] - P + pi[0];
    T4[1] = w * u[0] * u[1] + pi[1];
    T4[2] = w * u[0] * u[2] + pi[2];
    T4[3] = w * u[0] * u[3] + pi[3];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const float hi = (float)T4[n];
        F.at(F.qo, n)[s] = hi;
        F.at(F.qol, n)[s] = (float)(T4[n] - (double)hi);
    }
    for (int n = 0; n < 10; ++n) F.at(F.qo, 4 + n)[s] = (float)pi[n];
    if (nc == 15) F.at(F.qo, 14)[s] = (float)Pb;
    F.eo[s] = (float)e;
    F.po[s] = (float)p;
    for (int k = 0; k < 4; ++k) {
        F.at(F.uo, k)[s] = (float)u[k];
        F.at(up, k)[s] = (float)u[k];
    }
}

}  // namespace mx

namespace {

template <int NC, bool TABLE, bool DIM3, int TX, int TY, int TZ, bool BI = false>
hipError_t launch_mx(const mx::MxFields &F, const EosDev &E, const EvDesc *ev, const int *blk_ev,
                     const int *blk_first, int nblocks, hipStream_t st) {
    if constexpr (DIM3) {
        if (F.face_mode == 0)
            hipLaunchKernelGGL((mx::stage_mixed_kernel_3d<NC, TABLE, DIM3, TX, TY, TZ, false>), dim3(nblocks),
                               dim3(mx::MX_THREADS), 0, st, F, E, ev, blk_ev, blk_first);
        else
            hipLaunchKernelGGL((mx::stage_mixed_kernel_3d<NC, TABLE, DIM3, TX, TY, TZ, true>), dim3(nblocks),
                               dim3(mx::MX_THREADS), 0, st, F, E, ev, blk_ev, blk_first);
    } else {
        if (F.face_mode == 0)
            hipLaunchKernelGGL((mx::stage_mixed_kernel<NC, TABLE, DIM3, TX, TY, TZ, false, BI>), dim3(nblocks),
                               dim3(mx::MX_THREADS), 0, st, F, E, ev, blk_ev, blk_first);
        else
            hipLaunchKernelGGL((mx::stage_mixed_kernel<NC, TABLE, DIM3, TX, TY, TZ, true, BI>), dim3(nblocks),
                               dim3(mx::MX_THREADS), 0, st, F, E, ev, blk_ev, blk_first);
    }
    return hipGetLastError();
}

// Tile variants: the face list of a tile fills whole passes of 256 threads
// closely (16x15: 511 faces, 240 cells; 6x6x6: 756 faces, 216 cells).
// bi: every event of the launch is a boost-invariant 2+1D event (the
// kernels that neither load nor evolve the eta components).
template <int NC, bool TABLE>
hipError_t launch_mx_tile(const mx::MxFields &F, const EosDev &E, bool dim3d, int tile, bool bi, const EvDesc *ev,
                          const int *blk_ev, const int *blk_first, int nblocks, hipStream_t st) {
    if (!dim3d && bi) {
        switch (tile) {
        case 0: return launch_mx<NC, TABLE, false, 16, 15, 1, true>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 1: return launch_mx<NC, TABLE, false, 32, 7, 1, true>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 2: return launch_mx<NC, TABLE, false, 64, 3, 1, true>(F, E, ev, blk_ev, blk_first, nblocks, st);
        }
    } else if (!dim3d) {
        switch (tile) {
        case 0: return launch_mx<NC, TABLE, false, 16, 15, 1>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 1: return launch_mx<NC, TABLE, false, 32, 7, 1>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 2: return launch_mx<NC, TABLE, false, 64, 3, 1>(F, E, ev, blk_ev, blk_first, nblocks, st);
        }
    } else {
        switch (tile) {
        case 0: return launch_mx<NC, TABLE, true, 6, 6, 6>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 1: return launch_mx<NC, TABLE, true, 8, 8, 3>(F, E, ev, blk_ev, blk_first, nblocks, st);
        case 2: return launch_mx<NC, TABLE, true, 8, 4, 6>(F, E, ev, blk_ev, blk_first, nblocks, st);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace

const int MIXED_TILES_2D[MIXED_NTILE][3] = {{16, 15, 1}, {32, 7, 1}, {64, 3, 1}};
const int MIXED_TILES_3D[MIXED_NTILE][3] = {{6, 6, 6}, {8, 8, 3}, {8, 4, 6}};

hipError_t launch_stage_mixed(const mx::MxFields &F, const EosDev &E, int nc, bool dim3d, int tile, bool bi,
                              const EvDesc *ev, const int *blk_ev, const int *blk_first, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    if (tile < 0 || tile >= MIXED_NTILE || (bi && dim3d)) return hipErrorInvalidValue;
    const bool tab = E.kind == 1;
    if (nc == 14)
        return tab ? launch_mx_tile<14, true>(F, E, dim3d, tile, bi, ev, blk_ev, blk_first, nblocks, st)
                   : launch_mx_tile<14, false>(F, E, dim3d, tile, bi, ev, blk_ev, blk_first, nblocks, st);
    return tab ? launch_mx_tile<15, true>(F, E, dim3d, tile, bi, ev, blk_ev, blk_first, nblocks, st)
               : launch_mx_tile<15, false>(F, E, dim3d, tile, bi, ev, blk_ev, blk_first, nblocks, st);
}

hipError_t launch_init_mixed(const mx::MxFields &F, const AdmitSrc &S, float *up, long N, long off, int nc,
                             int *strikes, int *stop, hipStream_t st) {
    if (N <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mx::init_mixed_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, F, S, up, N, off,
                       nc, strikes, stop);
    return hipGetLastError();
}

// Per-event count of cells with e > eF (FP32 state, compared in FP64).  Each
// wave takes one 64-cell entry of the block table, the block four; one
// thread adds the waves' counts per run of equal events, so one atomic per
// event run replaces one per 64 cells.
__global__ __launch_bounds__(256) void count_above_batch_f_kernel(const float *__restrict__ e, double eF,
                                                                  const EvDesc *__restrict__ ev,
                                                                  const int *__restrict__ ev_of,
                                                                  const long *__restrict__ start, int nblocks,
                                                                  unsigned long long *__restrict__ counts) {
    __shared__ int wid[4];
    __shared__ unsigned long long wcnt[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long k = (long)blockIdx.x * 4 + w;  // this wave's table entry
    int id = -1;
    bool above = false;
    if (k < nblocks) {
        const int eid = ev_of[k];
        const EvDesc &D = ev[eid];
        if (D.active) {
            id = eid;
            const long N = (long)D.nx * D.ny * D.nz;
            const long sl = start[k] + lane;
            above = sl < N && (double)e[D.off + sl] > eF;
        }
    }
    const unsigned long long mask = __ballot(above);
    if (lane == 0) {
        wid[w] = id;
        wcnt[w] = (unsigned long long)__popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cur = -1;
        unsigned long long sum = 0;
        for (int q = 0; q < 4; ++q) {
            if (wid[q] != cur) {
                if (cur >= 0 && sum) atomicAdd(&counts[cur], sum);
                cur = wid[q];
                sum = 0;
            }
            sum += wcnt[q];
        }
        if (cur >= 0 && sum) atomicAdd(&counts[cur], sum);
    }
}

hipError_t launch_count_above_batch_f(const float *e, double eF, const EvDesc *ev, const int *ev_of,
                                      const long *start, int nblocks, unsigned long long *counts, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_above_batch_f_kernel, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, st, e, eF, ev,
                       ev_of, start, nblocks, counts);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void widen_kernel(const float *__restrict__ hi, const float *__restrict__ lo,
                                                    double *__restrict__ out, long n) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    out[c] = lo ? (double)hi[c] + (double)lo[c] : (double)hi[c];
}

hipError_t launch_widen(const float *hi, const float *lo, double *out, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hi, lo, out, n);
    return hipGetLastError();
}

// Several widenings in one launch: blockIdx.y picks the array; a NULL hi
// fills that output with zeros.
__global__ __launch_bounds__(256) void widen_set_kernel(const WidenSet S, long n) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (c >= n) return;
    const float *hi = S.hi[k], *lo = S.lo[k];
    S.out[k][c] = !hi ? 0.0 : (lo ? (double)hi[c] + (double)lo[c] : (double)hi[c]);
}

hipError_t launch_widen_set(const WidenSet &S, long n, hipStream_t st) {
    if (n <= 0 || S.count <= 0) return hipSuccess;
    if (S.count > WidenSet::MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(widen_set_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)S.count), dim3(256), 0, st, S, n);
    return hipGetLastError();
}

}  // namespace vh
