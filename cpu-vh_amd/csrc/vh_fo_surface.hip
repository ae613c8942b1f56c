// Freezeout surface elements on the device: the elements of every cut cube
// of a coarse step (callFOFinder2p1D / callFOFinder3p1D,
// freezeout/FreezeOut.cpp:200-361, with the Cornelius construction of
// vh_surface.hpp, the same source the host C ABI runs) and the ten
// variables at their centroids (interpolateVariable3D/4D, :182-198).
//
// One thread per cut cube, and one pass over the cut cubes of several events
// (a batch's resident events at one check) at once.  One launch builds each
// cube's elements and the interpolated variables into a slot of SLOT_ELEMS
// elements per cube and records the count; a scan of the counts gives every
// cube its offset, and a gather moves the slots' elements to those offsets —
// so each event's elements come out compact and in its flagged cubes' order.
// A cube with more elements than a slot holds (none in the runs measured)
// sends the whole pass to a second launch that builds every cube again
// straight into its offset.  The hypercube construction keeps the eight
// cubes' segments in private memory (about 12 KB a thread), hence blocks of
// 64 threads.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

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

// One event of a pass as the kernels see it: its cubes' ids, its first cube
// in the pass, where its compacted elements start (doubles, set after the
// scan), lattice, level, widths and variable slices.
struct FoDevItem {
    const long long *ids;
    long long cube0, base;
    int nx, ny, d4;
    double eF;
    FoCell g;
    FoSlices S;
};

// The item holding pass cube m: the last with c0[k] <= m (c0 increasing).
__device__ inline int item_of(const long long *__restrict__ c0, int n_items, long long m) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c0[mid] <= m) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// SLOTS: pass cube m's elements go to slot m (SLOT_ELEMS elements of up to 8
// doubles), its count to counts[m] (a count above the slot's size raises
// *overflow and writes nothing), a failure to *err; otherwise its item's
// compacted elements from offsets[m] on.  Corners: 16 doubles a pass cube
// from the item's first, 8 or 16 a cube within it.
template <bool SLOTS>
__global__ __launch_bounds__(ELEM_BLOCK) void fo_elem_write_kernel(
    const FoDevItem *__restrict__ items, const long long *__restrict__ c0, int n_items, long long ntot,
    const double *__restrict__ corners, const int *__restrict__ offsets, double *__restrict__ elements,
    double *__restrict__ vars, int *__restrict__ counts, int *__restrict__ err, int *__restrict__ overflow) {
    const long long m = (long long)blockIdx.x * ELEM_BLOCK + threadIdx.x;
    if (m >= ntot) return;
    const FoDevItem &I = items[item_of(c0, n_items, m)];
    const bool d4 = I.d4 != 0;
    const long long j = m - I.cube0;  // the cube within its item
    const FoCell g = I.g;
    double c[16];
    builder_corners(corners + I.cube0 * 16 + j * (d4 ? 16 : 8), d4, c);
    const int per = d4 ? 8 : 6, dim = d4 ? 4 : 3;
    long long voff;
    int cap;
    double *el;
    if constexpr (SLOTS) {
        voff = m * SLOT_ELEMS;
        cap = SLOT_ELEMS;
        el = elements + m * (SLOT_ELEMS * 8);
    } else {
        voff = offsets[m];
        cap = offsets[m + 1] - offsets[m];
        if (cap == 0) return;
        el = elements + I.base + (voff - offsets[I.cube0]) * per;
    }
    int ne = 0;
    const int r = d4 ? surf::hypercube_elements(c, I.eF, g.d, el, cap, &ne) : surf::cube_elements(c, I.eF, g.d, el, cap, &ne);
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
    const long mx = I.nx - 1, my = I.ny - 1;
    const long cid = I.ids[j];
    const long ix = cid % mx, iy = (cid / mx) % my, iz = cid / (mx * my);
    const long sy = I.nx, sz = (long)I.nx * I.ny;
    const long base = ix + sy * iy + sz * iz;
    for (int e = 0; e < ne; ++e) {
        const double *cen = el + e * per;
        double f[4] = {0, 0, 0, 0};
        for (int k = 0; k < dim; ++k) f[k] = cen[k] / g.d[k];
        double *out = vars + (voff + e) * FO_NVARS;
        for (int v = 0; v < FO_NVARS; ++v) {
            const double *p0 = I.S.v[0][v], *p1 = I.S.v[1][v];
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

// Moves pass cube m's counts[m] elements and variables from its slot to its
// item's compacted elements and the pass's variables at offsets[m].
__global__ __launch_bounds__(256) void fo_elem_gather_kernel(const FoDevItem *__restrict__ items,
                                                              const long long *__restrict__ c0, int n_items,
                                                              const double *__restrict__ sel,
                                                              const double *__restrict__ svar,
                                                              const int *__restrict__ counts,
                                                              const int *__restrict__ offsets, long long ntot,
                                                              double *__restrict__ elements, double *__restrict__ vars) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long m = g / SLOT_ELEMS;
    const int e = (int)(g - m * SLOT_ELEMS);
    if (m >= ntot || e >= counts[m]) return;
    const FoDevItem &I = items[item_of(c0, n_items, m)];
    const int per = I.d4 ? 8 : 6;
    const long long src = m * SLOT_ELEMS + e, dst = (long long)offsets[m] + e;
    const double *from = sel + m * (SLOT_ELEMS * 8) + e * per;
    double *to = elements + I.base + (dst - offsets[I.cube0]) * per;
    for (int k = 0; k < per; ++k) to[k] = from[k];
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

hipError_t fo_surface_many(FoSurfaceWork &W, hipStream_t st, FoItem *items, int n_items, int *bad) {
    *bad = 0;
    long long ntot = 0;
    int live = 0;
    for (int k = 0; k < n_items; ++k) {
        FoItem &I = items[k];
        *I.total = 0;
        I.cube0 = ntot;
        if (I.n > 0) {
            ntot += I.n;
            ++live;
        }
    }
    if (ntot == 0) return hipSuccess;
    if (ntot > INT_MAX / (2 * SLOT_ELEMS)) return hipErrorInvalidValue;  // int counts and offsets
    // staging: corners (16 doubles a cube) | counts | offsets (ntot + 1) |
    // error and overflow words | item table and first cubes | slot elements
    // | slot variables | scan scratch
    size_t temp_bytes = 0;
    hipError_t r = fo_elem_scan(nullptr, temp_bytes, nullptr, nullptr, ntot, st);
    if (r != hipSuccess) return r;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t table = sizeof(FoDevItem) * live + sizeof(long long) * live;
    const size_t cb = al(sizeof(double) * 16 * ntot), nb = al(sizeof(int) * ntot), ob = al(sizeof(int) * (ntot + 1)),
                 eb = 256, tb = al(table), seb = al(sizeof(double) * ntot * SLOT_ELEMS * 8),
                 svb = al(sizeof(double) * ntot * SLOT_ELEMS * FO_NVARS);
    if ((r = grow(&W.buf, &W.buf_bytes, cb + nb + ob + eb + tb + seb + svb + temp_bytes, st)) != hipSuccess) return r;
    if (table > W.up_bytes) {
        if ((r = hipStreamSynchronize(st)) != hipSuccess) return r;
        const size_t want = std::max(table, 2 * W.up_bytes);
        (void)hipHostFree(W.up);
        W.up = nullptr;
        W.up_bytes = 0;
        if ((r = hipHostMalloc((void **)&W.up, want)) != hipSuccess) return r;
        W.up_bytes = want;
    }
    char *base = (char *)W.buf;
    double *d_corn = (double *)base;
    int *d_cnt = (int *)(base + cb), *d_off = (int *)(base + cb + nb), *d_bad = (int *)(base + cb + nb + ob);
    int *d_ovf = d_bad + 1;
    char *d_table = base + cb + nb + ob + eb;
    const FoDevItem *d_items = (const FoDevItem *)d_table;
    const long long *d_c0 = (const long long *)(d_table + sizeof(FoDevItem) * live);
    double *d_sel = (double *)(d_table + tb), *d_svar = (double *)(d_table + tb + seb);
    void *d_tmp = d_table + tb + seb + svb;
    // the live items in pass order (the host copy stays untouched while its
    // upload may be in flight: until the next synchronisation)
    FoDevItem *h_items = (FoDevItem *)W.up;
    long long *h_c0 = (long long *)(W.up + sizeof(FoDevItem) * live);
    std::vector<int> of(live);
    for (int k = 0, q = 0; k < n_items; ++k) {
        const FoItem &I = items[k];
        if (I.n <= 0) continue;
        h_items[q] = FoDevItem{I.ids, I.cube0, 0, I.nx, I.ny, I.nz > 1 ? 1 : 0, I.eF, I.g, I.S};
        h_c0[q] = I.cube0;
        of[q++] = k;
    }
#define FOCHK(call)                               \
    do {                                          \
        if ((r = (call)) != hipSuccess) return r; \
    } while (0)
    FOCHK(hipMemcpyAsync(d_table, W.up, table, hipMemcpyHostToDevice, st));
    for (int q = 0; q < live; ++q) {
        const FoItem &I = items[of[q]];
        FOCHK(launch_fo_corners(I.eprev, I.ecur, I.nx, I.ny, I.nz, I.ids, I.n, d_corn + I.cube0 * 16, st));
    }
    FOCHK(hipMemsetAsync(d_off, 0, sizeof(int), st));
    FOCHK(hipMemsetAsync(d_bad, 0, 2 * sizeof(int), st));
    const unsigned grid = (unsigned)((ntot + ELEM_BLOCK - 1) / ELEM_BLOCK);
    hipLaunchKernelGGL(fo_elem_write_kernel<true>, dim3(grid), dim3(ELEM_BLOCK), 0, st, d_items, d_c0, live, ntot,
                       d_corn, nullptr, d_sel, d_svar, d_cnt, d_bad, d_ovf);
    FOCHK(hipGetLastError());
    FOCHK(fo_elem_scan(d_tmp, temp_bytes, d_cnt, d_off + 1, ntot, st));
    int flags[2] = {0, 0};
    {
        std::vector<void *> dst(live + 1);
        std::vector<const void *> src(live + 1);
        std::vector<size_t> bytes(live + 1);
        for (int q = 0; q < live; ++q) {
            const FoItem &I = items[of[q]];
            dst[q] = I.n_el;
            src[q] = d_cnt + I.cube0;
            bytes[q] = sizeof(int) * I.n;
        }
        dst[live] = flags;
        src[live] = d_bad;
        bytes[live] = 2 * sizeof(int);
        FOCHK(fo_download(W, st, live + 1, dst.data(), src.data(), bytes.data()));
    }
    *bad = flags[0];
    // every item's element count and where its elements go: [item elements,
    // 6 or 8 doubles each, in pass order | the pass's variables]
    long long all = 0, dbl = 0;
    std::vector<long long> e0(live), at(live);
    for (int q = 0; q < live; ++q) {
        const FoItem &I = items[of[q]];
        long long t = 0;
        for (long long m = 0; m < I.n; ++m) t += I.n_el[m];
        *I.total = t;
        e0[q] = all;
        at[q] = dbl;
        h_items[q].base = dbl;
        all += t;
        dbl += t * (I.nz > 1 ? 8 : 6);
    }
    if (*bad || all == 0) return hipSuccess;
    bool any = false;
    for (int q = 0; q < live; ++q) any |= *items[of[q]].total <= items[of[q]].cap;
    if (!any) return hipSuccess;
    FOCHK(grow((void **)&W.el, &W.el_bytes, sizeof(double) * (size_t)(dbl + all * FO_NVARS), st));
    double *d_el = W.el, *d_vars = W.el + dbl;
    FOCHK(hipMemcpyAsync(d_table, W.up, table, hipMemcpyHostToDevice, st));  // with the bases
    if (flags[1]) {  // a cube outgrew its slot: build every cube again at its offset
        hipLaunchKernelGGL(fo_elem_write_kernel<false>, dim3(grid), dim3(ELEM_BLOCK), 0, st, d_items, d_c0, live, ntot,
                           d_corn, d_off, d_el, d_vars, nullptr, nullptr, nullptr);
    } else {
        const unsigned ggrid = (unsigned)((ntot * SLOT_ELEMS + 255) / 256);
        hipLaunchKernelGGL(fo_elem_gather_kernel, dim3(ggrid), dim3(256), 0, st, d_items, d_c0, live, d_sel, d_svar,
                           d_cnt, d_off, ntot, d_el, d_vars);
    }
    FOCHK(hipGetLastError());
    {
        std::vector<void *> dst;
        std::vector<const void *> src;
        std::vector<size_t> bytes;
        for (int q = 0; q < live; ++q) {
            const FoItem &I = items[of[q]];
            const long long t = *I.total;
            if (t == 0 || t > I.cap) continue;
            dst.push_back(I.elements);
            src.push_back(d_el + at[q]);
            bytes.push_back(sizeof(double) * (size_t)t * (I.nz > 1 ? 8 : 6));
            dst.push_back(I.vars);
            src.push_back(d_vars + e0[q] * FO_NVARS);
            bytes.push_back(sizeof(double) * (size_t)t * FO_NVARS);
        }
        FOCHK(fo_download(W, st, (int)dst.size(), dst.data(), src.data(), bytes.data()));
    }
#undef FOCHK
    return hipSuccess;
}

hipError_t fo_surface_run(FoSurfaceWork &W, hipStream_t st, const double *eprev, const double *ecur, int nx, int ny,
                          int nz, const long long *ids, long long n, double eF, const FoCell &g, const FoSlices &S,
                          int *n_el, double *elements, double *vars, long long cap, long long *total, int *bad) {
    FoItem I;
    I.eprev = eprev;
    I.ecur = ecur;
    I.ids = ids;
    I.n = n;
    I.nx = nx;
    I.ny = ny;
    I.nz = nz;
    I.eF = eF;
    I.g = g;
    I.S = S;
    I.n_el = n_el;
    I.elements = elements;
    I.vars = vars;
    I.cap = cap;
    I.total = total;
    return fo_surface_many(W, st, &I, 1, bad);
}

void fo_surface_free(FoSurfaceWork &W) {
    (void)hipFree(W.buf);
    (void)hipFree(W.el);
    (void)hipHostFree(W.host);
    (void)hipHostFree(W.up);
    W = FoSurfaceWork{};
}

hipError_t fo_elem_scan(void *temp, size_t &temp_bytes, const int *counts, int *offsets_plus1, long long n,
                        hipStream_t st) {
    return hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, counts, offsets_plus1, (int)n, st);
}

}  // namespace vh
