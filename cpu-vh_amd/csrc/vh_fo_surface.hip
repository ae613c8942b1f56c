// Freezeout surface elements on the device: the elements of every cut cube
// of a coarse step (callFOFinder2p1D / callFOFinder3p1D,
// freezeout/FreezeOut.cpp:200-361, with the Cornelius construction of
// vh_surface.hpp, the same source the host C ABI runs) and the ten
// variables at their centroids (interpolateVariable3D/4D, :182-198).
//
// One thread per cut cube.  One pass builds each cube's elements and the
// interpolated variables into a slot of SLOT_ELEMS elements per cube and
// records the count; a scan of the counts gives every cube its offset, and a
// gather moves the slots' elements to those offsets — so the elements come
// out compact and in the flagged cubes' order.  A cube with more elements than
// a slot holds (none in the runs measured) sends the whole call to a second
// pass that builds every cube again straight into its offset.  The
// hypercube construction keeps the eight cubes' segments in private memory
// (about 12 KB a thread), hence blocks of 64 threads.
#include <algorithm>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "vh_fo_interp.hpp"
#include "vh_kernels.hpp"
#include "vh_surface.hpp"

namespace vh {

namespace {

constexpr int ELEM_BLOCK = 64;
constexpr int SLOT_ELEMS = 4;

// A flagged cube's corner energies (fo_corners_kernel's layout: previous
// slice then current, x fastest within a slice) in the element builder's
// order [tau][x][y][eta].
__device__ inline void builder_corners(const double *fc, bool d4, double *c) {
    const int half = d4 ? 8 : 4, nzc = d4 ? 2 : 1;
    for (int it = 0; it < 2; ++it)
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y)
                for (int z = 0; z < nzc; ++z) c[((it * 2 + x) * 2 + y) * nzc + z] = fc[it * half + (z * 2 + y) * 2 + x];
}

// SLOTS: cube m's elements go to slot m (SLOT_ELEMS elements), its count to
// counts[m] (a count above the slot's size raises *overflow and writes
// nothing), a failure to *err; otherwise to offsets[m] .. offsets[m + 1].
template <bool SLOTS>
__global__ __launch_bounds__(ELEM_BLOCK) void fo_elem_write_kernel(
    const double *__restrict__ corners, const long long *__restrict__ cubes, long long n, bool d4, double eF, FoCell g,
    const int *__restrict__ offsets, FoSlices S, int nx, int ny, double *__restrict__ elements,
    double *__restrict__ vars, int *__restrict__ counts, int *__restrict__ err, int *__restrict__ overflow) {
    const long long m = (long long)blockIdx.x * ELEM_BLOCK + threadIdx.x;
    if (m >= n) return;
    double c[16];
    builder_corners(corners + m * (d4 ? 16 : 8), d4, c);
    const long long off = SLOTS ? m * SLOT_ELEMS : offsets[m];
    const int cap = SLOTS ? SLOT_ELEMS : offsets[m + 1] - offsets[m];
    if (!SLOTS && cap == 0) return;
    const int per = d4 ? 8 : 6, dim = d4 ? 4 : 3;
    double *el = elements + off * per;
    int ne = 0;
    const int r = d4 ? surf::hypercube_elements(c, eF, g.d, el, cap, &ne) : surf::cube_elements(c, eF, g.d, el, cap, &ne);
    if constexpr (SLOTS) {
        counts[m] = ne;
        if (r == surf::SURF_STATE) err[0] = 1;  // any failing cube will do: one plain store
        if (r != surf::SURF_OK) {
            if (r == surf::SURF_CAPACITY) overflow[0] = 1;
            return;
        }
    }
    // the ten variables at each centroid: the cell fraction cen/width, then
    // the multilinear interpolation of the two coarse slices
    const long mx = nx - 1, my = ny - 1;
    const long cid = cubes[m];
    const long ix = cid % mx, iy = (cid / mx) % my, iz = cid / (mx * my);
    const long sy = nx, sz = (long)nx * ny;
    const long base = ix + sy * iy + sz * iz;
    for (int e = 0; e < ne; ++e) {
        const double *cen = el + e * per;
        double f[4] = {0, 0, 0, 0};
        for (int k = 0; k < dim; ++k) f[k] = cen[k] / g.d[k];
        double *out = vars + (off + e) * FO_NVARS;
        for (int v = 0; v < FO_NVARS; ++v) {
            const double *p0 = S.v[0][v], *p1 = S.v[1][v];
            if (d4) {
                double a[2][2][2][2];
                for (int x = 0; x < 2; ++x)
                    for (int y = 0; y < 2; ++y)
                        for (int z = 0; z < 2; ++z) {
                            const long s = base + x + sy * y + sz * z;
                            a[0][x][y][z] = p0[s];
                            a[1][x][y][z] = p1[s];
                        }
                out[v] = fo_lin4(f[0], f[1], f[2], f[3], a);
            } else {
                double a[2][2][2];
                for (int x = 0; x < 2; ++x)
                    for (int y = 0; y < 2; ++y) {
                        const long s = base + x + sy * y;
                        a[0][x][y] = p0[s];
                        a[1][x][y] = p1[s];
                    }
                out[v] = fo_lin3(f[0], f[1], f[2], a);
            }
        }
    }
}

// Moves cube m's counts[m] elements and variables from its slot to offsets[m].
__global__ __launch_bounds__(256) void fo_elem_gather_kernel(const double *__restrict__ sel,
                                                              const double *__restrict__ svar,
                                                              const int *__restrict__ counts,
                                                              const int *__restrict__ offsets, long long n, int per,
                                                              double *__restrict__ elements, double *__restrict__ vars) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long m = g / SLOT_ELEMS;
    const int e = (int)(g - m * SLOT_ELEMS);
    if (m >= n || e >= counts[m]) return;
    const long long src = m * SLOT_ELEMS + e, dst = (long long)offsets[m] + e;
    for (int k = 0; k < per; ++k) elements[dst * per + k] = sel[src * per + k];
    for (int v = 0; v < FO_NVARS; ++v) vars[dst * FO_NVARS + v] = svar[src * FO_NVARS + v];
}

}  // namespace

// Device memory grown on demand (the stream is drained before a buffer is
// replaced).
static hipError_t grow(void **buf, size_t *have, size_t need, hipStream_t st) {
    if (need <= *have) return hipSuccess;
    hipError_t r = hipStreamSynchronize(st);
    if (r != hipSuccess) return r;
    const size_t want = std::max(need, 2 * *have);
    (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    if ((r = hipMalloc(buf, want)) != hipSuccess) return r;
    *have = want;
    return hipSuccess;
}

hipError_t fo_download(FoSurfaceWork &W, hipStream_t st, int n, void *const *dst, const void *const *src,
                       const size_t *bytes) {
    size_t total = 0;
    for (int k = 0; k < n; ++k) total += (bytes[k] + 255) & ~(size_t)255;
    hipError_t r;
    if (total > W.host_bytes) {
        if ((r = hipStreamSynchronize(st)) != hipSuccess) return r;
        const size_t want = std::max(total, 2 * W.host_bytes);
        (void)hipHostFree(W.host);
        W.host = nullptr;
        W.host_bytes = 0;
        if ((r = hipHostMalloc((void **)&W.host, want)) != hipSuccess) return r;
        W.host_bytes = want;
    }
    size_t at = 0;
    for (int k = 0; k < n; ++k) {
        if (bytes[k] && (r = hipMemcpyAsync(W.host + at, src[k], bytes[k], hipMemcpyDeviceToHost, st)) != hipSuccess)
            return r;
        at += (bytes[k] + 255) & ~(size_t)255;
    }
    if ((r = hipStreamSynchronize(st)) != hipSuccess) return r;
    at = 0;
    for (int k = 0; k < n; ++k) {
        if (bytes[k]) std::memcpy(dst[k], W.host + at, bytes[k]);
        at += (bytes[k] + 255) & ~(size_t)255;
    }
    return hipSuccess;
}

hipError_t fo_surface_run(FoSurfaceWork &W, hipStream_t st, const double *eprev, const double *ecur, int nx, int ny,
                          int nz, const long long *ids, long long n, double eF, const FoCell &g, const FoSlices &S,
                          int *n_el, double *elements, double *vars, long long cap, long long *total, int *bad) {
    *total = 0;
    *bad = 0;
    if (n <= 0) return hipSuccess;
    const bool d4 = nz > 1;
    const int dim = d4 ? 4 : 3, per = d4 ? 16 : 8, epd = 2 * dim;
    // staging: corners | counts | offsets (n + 1) | error and overflow words |
    // slot elements | slot variables | scan scratch
    size_t temp_bytes = 0;
    hipError_t r = fo_elem_scan(nullptr, temp_bytes, nullptr, nullptr, n, st);
    if (r != hipSuccess) return r;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t cb = al(sizeof(double) * per * n), nb = al(sizeof(int) * n), ob = al(sizeof(int) * (n + 1)),
                 eb = 256, seb = al(sizeof(double) * n * SLOT_ELEMS * epd),
                 svb = al(sizeof(double) * n * SLOT_ELEMS * FO_NVARS);
    if ((r = grow(&W.buf, &W.buf_bytes, cb + nb + ob + eb + seb + svb + temp_bytes, st)) != hipSuccess) return r;
    char *base = (char *)W.buf;
    double *d_corn = (double *)base;
    int *d_cnt = (int *)(base + cb), *d_off = (int *)(base + cb + nb), *d_bad = (int *)(base + cb + nb + ob);
    int *d_ovf = d_bad + 1;
    double *d_sel = (double *)(base + cb + nb + ob + eb), *d_svar = (double *)(base + cb + nb + ob + eb + seb);
    void *d_tmp = base + cb + nb + ob + eb + seb + svb;
    int h_total = 0, flags[2] = {0, 0};
#define FOCHK(call)                               \
    do {                                          \
        if ((r = (call)) != hipSuccess) return r; \
    } while (0)
    const unsigned grid = (unsigned)((n + ELEM_BLOCK - 1) / ELEM_BLOCK);
    FOCHK(launch_fo_corners(eprev, ecur, nx, ny, nz, ids, n, d_corn, st));
    FOCHK(hipMemsetAsync(d_off, 0, sizeof(int), st));
    FOCHK(hipMemsetAsync(d_bad, 0, 2 * sizeof(int), st));
    hipLaunchKernelGGL(fo_elem_write_kernel<true>, dim3(grid), dim3(ELEM_BLOCK), 0, st, d_corn, ids, n, d4, eF, g,
                       nullptr, S, nx, ny, d_sel, d_svar, d_cnt, d_bad, d_ovf);
    FOCHK(hipGetLastError());
    FOCHK(fo_elem_scan(d_tmp, temp_bytes, d_cnt, d_off + 1, n, st));
    {
        void *const dst[3] = {n_el, &h_total, flags};
        const void *const src[3] = {d_cnt, d_off + n, d_bad};
        const size_t bytes[3] = {sizeof(int) * n, sizeof(int), 2 * sizeof(int)};
        FOCHK(fo_download(W, st, 3, dst, src, bytes));
    }
    *total = h_total;
    *bad = flags[0];
    if (*bad || h_total == 0 || h_total > cap) return hipSuccess;
    const size_t need = sizeof(double) * (size_t)h_total * (epd + FO_NVARS);
    FOCHK(grow((void **)&W.el, &W.el_bytes, need, st));
    double *d_el = W.el, *d_vars = W.el + (size_t)h_total * epd;
    if (flags[1]) {  // a cube outgrew its slot: build every cube again at its offset
        hipLaunchKernelGGL(fo_elem_write_kernel<false>, dim3(grid), dim3(ELEM_BLOCK), 0, st, d_corn, ids, n, d4, eF,
                           g, d_off, S, nx, ny, d_el, d_vars, nullptr, nullptr, nullptr);
    } else {
        const unsigned ggrid = (unsigned)((n * SLOT_ELEMS + 255) / 256);
        hipLaunchKernelGGL(fo_elem_gather_kernel, dim3(ggrid), dim3(256), 0, st, d_sel, d_svar, d_cnt, d_off, n, epd,
                           d_el, d_vars);
    }
    FOCHK(hipGetLastError());
    {
        void *const dst[2] = {elements, vars};
        const void *const src[2] = {d_el, d_vars};
        const size_t bytes[2] = {sizeof(double) * (size_t)h_total * epd, sizeof(double) * (size_t)h_total * FO_NVARS};
        FOCHK(fo_download(W, st, 2, dst, src, bytes));
    }
#undef FOCHK
    return hipSuccess;
}

void fo_surface_free(FoSurfaceWork &W) {
    (void)hipFree(W.buf);
    (void)hipFree(W.el);
    (void)hipHostFree(W.host);
    W = FoSurfaceWork{};
}

hipError_t fo_elem_scan(void *temp, size_t &temp_bytes, const int *counts, int *offsets_plus1, long long n,
                        hipStream_t st) {
    return hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, counts, offsets_plus1, (int)n, st);
}

}  // namespace vh
