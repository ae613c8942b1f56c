// C ABI of the MI355X viscous-hydro step (include/cpuvh_amd.h).
//
// Device state per handle (FP64 SoA, interior cells only, x fastest):
//   q[3][nc]  conserved sets rotating through the roles q^n / stage / next
//   u[3][4]   velocity sets rotating through u / previous u / stage u
//   e[2], p[2] current and stage-1 values
// One stream per handle; vh_steps replays a captured HIP graph of
// [stage 1, stage 2] x count with per-step times read from a device table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cpuvh_amd.h"
#include "vh_host_eos.hpp"
#include "vh_kernels.hpp"

using namespace vh;

struct vh_handle {
    vh_config cfg{};
    int nc = 14;
    long N = 0;
    bool dim3d = true;
    int dev = 0;
    hipStream_t stream = nullptr;
    double *pool = nullptr;
    double *q[3][15] = {};
    double *u[3][4] = {};
    double *e[2] = {}, *p[2] = {};
    int iq_cur = 0, iq_stage = 1, iq_next = 2;
    int iu_cur = 0, iu_prev = 1, iu_stage = 2;
    bool conserved_ready = false;
    double *regfac = nullptr;   // regFacShear output, allocated on first request
    double *regfacb = nullptr;  // regFacBulk output (bulk builds), likewise
    int *d_err = nullptr;
    unsigned long long *d_count = nullptr;
    // freezeout slices
    double *fo_prev = nullptr, *fo_cur = nullptr;
    long long *fo_ids = nullptr;
    // the ten interpolated variables' slices, [2][10][N]; slot fo_vslot is current
    double *fo_vars = nullptr;
    int fo_vslot = 0;
    // staging for corners and interpolation points, grown on demand
    double *fo_buf = nullptr;
    size_t fo_buf_bytes = 0;
    // the last vh_fo_flag call's count and level (-1: none yet), and the
    // device surface elements and their variables, grown on demand
    long long fo_n = -1;
    double fo_eF = 0;
    FoSurfaceWork fo_surf;
    // thermal vorticity and velocity, 10 x N, from the last vh_thermal_vorticity
    double *vort = nullptr;
    // equation of state
    HostEos heos;
    EosDev eos{};
    DeviceEos deos;
    double e_reg = 0;
    // per-step times for graph replays
    StepTimes *d_times = nullptr;
    int times_cap = 0;
    int *d_base = nullptr;
    int *h_base = nullptr;  // pinned
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms = 0;
};

namespace {

thread_local std::string g_last_error;

}  // namespace

void vh::set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

vh_status hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return VH_ERR_HIP;
}

#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return hip_fail(_e, #call);   \
    } while (0)

void fill_times(const vh_config &c, StepTimes *out, int n_lo, int n_hi) {
    // HydroPlugin.cpp:226,307: step n runs at t = t0 + (n-1)*dt; rungeKutta2
    // advances t += dt between the stages (:900).
    for (int n = n_lo; n < n_hi; ++n) {
        const double t = (n == 1) ? c.t0 : c.t0 + (n - 1) * c.dt;
        const double t2 = t + c.dt;
        out[n - n_lo] = StepTimes{t, pow(t, 2), t2, pow(t2, 2)};
    }
}

vh_status ensure_times(vh_handle *h, int n_max) {
    if (n_max < h->times_cap) return VH_OK;
    int cap = std::max(4096, h->times_cap);
    while (cap <= n_max) cap *= 2;
    std::vector<StepTimes> tab(cap);
    fill_times(h->cfg, tab.data() + 1, 1, cap);
    tab[0] = StepTimes{0, 0, 0, 0};
    StepTimes *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(StepTimes) * cap));
    HIPCHK(hipMemcpy(d, tab.data(), sizeof(StepTimes) * cap, hipMemcpyHostToDevice));
    if (h->d_times) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(h->d_times));
    }
    for (auto &kv : h->graphs) hipGraphExecDestroy(kv.second);
    h->graphs.clear();
    h->d_times = d;
    h->times_cap = cap;
    return VH_OK;
}

StageParams stage_params(const vh_handle *h, int stage, int iq_in, int iu_in, int iu_prev_used, int ie_in,
                         int iq_out, int iu_out, int ie_out) {
    StageParams P{};
    for (int n = 0; n < 15; ++n) {
        P.q[n] = h->q[iq_in][n];
        P.qn[n] = h->q[h->iq_cur][n];
        P.qo[n] = h->q[iq_out][n];
    }
    for (int c = 0; c < 4; ++c) {
        P.u[c] = h->u[iu_in][c];
        P.up[c] = h->u[iu_prev_used][c];
        P.uo[c] = h->u[iu_out][c];
    }
    P.e = h->e[ie_in];
    P.p = h->p[ie_in];
    P.eo = h->e[ie_out];
    P.po = h->p[ie_out];
    P.regfac = h->regfac;
    P.regfacb = h->regfacb;
    P.stage = stage;
    P.dt = h->cfg.dt;
    P.dx = h->cfg.dx;
    P.dy = h->cfg.dy;
    P.dz = h->cfg.dz;
    P.e_reg = h->e_reg;
    P.etas_min = h->cfg.etabar;
    P.etas_Tk_GeV = h->cfg.etas_Tk_GeV;
    P.etas_slope_lo = h->cfg.etas_slope_lo;
    P.etas_slope_hi = h->cfg.etas_slope_hi;
    P.zeta_Tpeak_GeV = h->cfg.zeta_Tpeak_GeV;
    P.zeta_peak = h->cfg.zeta_peak;
    P.nx = h->cfg.nx;
    P.ny = h->cfg.ny;
    P.nz = h->cfg.nz;
    P.err = h->d_err;
    return P;
}

// Enqueues one step (two stage kernels) and rotates the buffer roles.
// With tab != null the kernels read their times from tab[*base + off].
vh_status enqueue_step(vh_handle *h, const StepTimes &st, const StepTimes *tab, int off) {
    // stage 1: q, u (prev up) -> qS, uS, e[1], p[1]
    StageParams P1 = stage_params(h, 0, h->iq_cur, h->iu_cur, h->iu_prev, 0, h->iq_stage, h->iu_stage, 1);
    // stage 2: qS, uS (prev u), q^n -> Q, u into the old "previous" set, e[0], p[0]
    StageParams P2 = stage_params(h, 1, h->iq_stage, h->iu_stage, h->iu_cur, 1, h->iq_next, h->iu_prev, 0);
    P1.st = P2.st = st;
    P1.tab = P2.tab = tab;
    P1.base = P2.base = h->d_base;
    P1.off = P2.off = off;
    HIPCHK(launch_stage(P1, h->eos, h->nc, h->dim3d, h->stream));
    HIPCHK(launch_stage(P2, h->eos, h->nc, h->dim3d, h->stream));
    // setCurrentConservedVariables (q <-> Q) and swapFluidVelocity(&up, &u)
    std::swap(h->iq_cur, h->iq_next);
    std::swap(h->iu_cur, h->iu_prev);
    return VH_OK;
}

vh_status check_err(vh_handle *h) {
    int err[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(err, h->d_err, sizeof(int) * 3, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return err[0] ? VH_ERR_NAN : VH_OK;
}

}  // namespace

extern "C" {

const char *vh_status_string(vh_status s) {
    switch (s) {
    case VH_OK: return "ok";
    case VH_ERR_ARG: return "invalid argument";
    case VH_ERR_HIP: return g_last_error.empty() ? "HIP error" : g_last_error.c_str();
    case VH_ERR_NOMEM: return "device allocation failed";
    case VH_ERR_NAN: return "NaN in primitive recovery or regulation";
    case VH_ERR_STATE: return "call out of order";
    case VH_ERR_CAPACITY: return "output capacity too small";
    }
    return "unknown status";
}

const char *vh_build_info(void) {
    return "cpuvh_amd: gfx950 HIP, FP64, -ffp-contract=off, fused stage kernels";
}

vh_status vh_create(const vh_config *cfg, vh_handle **out) {
    if (!cfg || !out) return VH_ERR_ARG;
    *out = nullptr;
    const vh_config &c = *cfg;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1 || !(c.dx > 0) || !(c.dy > 0) || !(c.dz > 0) || !(c.dt > 0) ||
        !(c.t0 > 0) || !(c.etabar > 0) || (c.eos_kind != VH_EOS_CONFORMAL && c.eos_kind != VH_EOS_TABLE) ||
        (c.eos_kind == VH_EOS_TABLE && !c.eos_table) || (long)c.nx * c.ny * c.nz >= (1L << 31) ||
        (c.cs2_mode != 0 && c.cs2_mode != 1) || !(c.zeta_Tpeak_GeV >= 0.0) || !(c.zeta_peak >= 0.0))
        return VH_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_last_error = "no HIP device";
        return VH_ERR_HIP;
    }
    if (c.device < 0 || c.device >= ndev) return VH_ERR_ARG;
    vh_handle *h = new vh_handle();
    h->cfg = c;
    h->cfg.eos_table = nullptr;
    h->dev = c.device;
    h->nc = c.bulk ? 15 : 14;
    h->N = (long)c.nx * c.ny * c.nz;
    h->dim3d = c.nz > 1;
    auto fail = [&](vh_status s) {
        vh_destroy(h);
        return s;
    };
    if (hipSetDevice(h->dev) != hipSuccess) return fail(hip_fail(hipErrorInvalidDevice, "hipSetDevice"));
    if (c.eos_kind == VH_EOS_TABLE) h->heos.load(c.eos_table);
    // rungeKutta2 :884-888: e_reg = e(T_reg / hbarc), constant per run
    h->e_reg = h->heos.energy_density(c.T_reg_GeV / 0.197326938);

    const long N = h->N;
    const size_t nfield = 3 * 15 + 3 * 4 + 4;
    if (hipMalloc(&h->pool, sizeof(double) * nfield * N) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipMemset(h->pool, 0, sizeof(double) * nfield * N) != hipSuccess) return fail(VH_ERR_HIP);
    double *ptr = h->pool;
    for (int s = 0; s < 3; ++s)
        for (int n = 0; n < 15; ++n, ptr += N) h->q[s][n] = ptr;
    for (int s = 0; s < 3; ++s)
        for (int k = 0; k < 4; ++k, ptr += N) h->u[s][k] = ptr;
    for (int s = 0; s < 2; ++s, ptr += N) h->e[s] = ptr;
    for (int s = 0; s < 2; ++s, ptr += N) h->p[s] = ptr;
    if (hipMalloc(&h->d_err, sizeof(int) * 4) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipMemset(h->d_err, 0, sizeof(int) * 4) != hipSuccess) return fail(VH_ERR_HIP);
    if (hipMalloc(&h->d_count, sizeof(unsigned long long)) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipMalloc(&h->d_base, sizeof(int)) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipHostMalloc(&h->h_base, sizeof(int), hipHostMallocDefault) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(VH_ERR_HIP);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return fail(VH_ERR_HIP);

    if (hipError_t r = h->deos.init(h->heos, c.eos_kind, c.cs2_mode); r != hipSuccess)
        return fail(r == hipErrorOutOfMemory ? VH_ERR_NOMEM : hip_fail(r, "equation of state upload"));
    h->eos = h->deos.E;
    vh_status s = ensure_times(h, 1);
    if (s != VH_OK) return fail(s);
    *out = h;
    return VH_OK;
}

vh_status vh_destroy(vh_handle *h) {
    if (!h) return VH_ERR_ARG;
    hipSetDevice(h->dev);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (auto &kv : h->graphs) hipGraphExecDestroy(kv.second);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    hipFree(h->pool);
    hipFree(h->regfac);
    hipFree(h->regfacb);
    hipFree(h->d_err);
    hipFree(h->d_count);
    hipFree(h->d_base);
    if (h->h_base) hipHostFree(h->h_base);
    hipFree(h->fo_prev);
    hipFree(h->fo_cur);
    hipFree(h->fo_ids);
    hipFree(h->fo_vars);
    hipFree(h->fo_buf);
    fo_surface_free(h->fo_surf);
    hipFree(h->vort);
    h->deos.release();
    hipFree(h->d_times);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return VH_OK;
}

vh_status vh_upload_ic(vh_handle *h, const double *e, const double *p, const double *ut, const double *ux,
                       const double *uy, const double *un, const double *pi10, const double *Pi) {
    if (!h || !e || !p) return VH_ERR_ARG;
    HIPCHK(hipSetDevice(h->dev));
    const size_t b = sizeof(double) * h->N;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->e[0], e, b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->p[0], p, b, hipMemcpyHostToDevice));
    // a NULL array is zero (u^tau: one)
    const double *uu[4] = {ut, ux, uy, un};
    for (int c = 0; c < 4; ++c) {
        if (uu[c]) {
            HIPCHK(hipMemcpy(h->u[h->iu_cur][c], uu[c], b, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(h->u[h->iu_prev][c], uu[c], b, hipMemcpyHostToDevice));  // up = u
        } else {
            HIPCHK(launch_fill(h->u[h->iu_cur][c], h->N, c == 0 ? 1.0 : 0.0, h->stream));
            HIPCHK(launch_fill(h->u[h->iu_prev][c], h->N, c == 0 ? 1.0 : 0.0, h->stream));
        }
    }
    for (int n = 0; n < 10; ++n) {
        if (pi10) HIPCHK(hipMemcpy(h->q[h->iq_cur][4 + n], pi10 + (size_t)n * h->N, b, hipMemcpyHostToDevice));
        else HIPCHK(launch_fill(h->q[h->iq_cur][4 + n], h->N, 0.0, h->stream));
    }
    if (h->nc == 15) {
        if (Pi) HIPCHK(hipMemcpy(h->q[h->iq_cur][14], Pi, b, hipMemcpyHostToDevice));
        else HIPCHK(launch_fill(h->q[h->iq_cur][14], h->N, 0.0, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->conserved_ready = false;
    return VH_OK;
}

vh_status vh_init_conserved(vh_handle *h, double t) {
    if (!h) return VH_ERR_ARG;
    (void)t;
    HIPCHK(hipSetDevice(h->dev));
    double *q[15];
    for (int n = 0; n < 15; ++n) q[n] = h->q[h->iq_cur][n];
    const double *u[4];
    for (int c = 0; c < 4; ++c) u[c] = h->u[h->iu_cur][c];
    HIPCHK(launch_init_conserved(q, h->nc == 15 ? h->q[h->iq_cur][14] : nullptr, h->e[0], h->p[0], u, h->N,
                                 h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->conserved_ready = true;
    return VH_OK;
}

vh_status vh_step(vh_handle *h, double t) {
    if (!h) return VH_ERR_ARG;
    if (!h->conserved_ready) return VH_ERR_STATE;
    HIPCHK(hipSetDevice(h->dev));
    const double t2 = t + h->cfg.dt;
    const StepTimes st{t, pow(t, 2), t2, pow(t2, 2)};
    return enqueue_step(h, st, nullptr, 0);
}

// Graph of [stage 1, stage 2] x count for the current buffer-role state,
// count <= MAX_GRAPH_STEPS, ending with a node that moves the step clock on.
static vh_status get_graph(vh_handle *h, int count, hipGraphExec_t *out) {
    const int parity = h->iq_cur * 3 + h->iu_cur;  // buffer-role state the graph was captured in
    const auto key = std::make_pair(count, parity);
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
        const int iq0 = h->iq_cur, iqn0 = h->iq_next, iu0 = h->iu_cur, iup0 = h->iu_prev;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        vh_status cs = VH_OK;
        for (int k = 0; k < count && cs == VH_OK; ++k) cs = enqueue_step(h, StepTimes{}, h->d_times, k);
        if (cs == VH_OK) {  // the next graph replayed continues at step *d_base + count
            const hipError_t ae = launch_advance_base(h->d_base, count, h->stream);
            if (ae != hipSuccess) cs = hip_fail(ae, "advance step clock");
        }
        hipError_t ce = hipStreamEndCapture(h->stream, &g);
        // capture only records the launches: restore the roles, replay applies them
        h->iq_cur = iq0; h->iq_next = iqn0; h->iu_cur = iu0; h->iu_prev = iup0;
        if (cs != VH_OK) return cs;
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate");
        it = h->graphs.emplace(key, ge).first;
    }
    *out = it->second;
    return VH_OK;
}

// The graphs that replay `count` steps from the current buffer roles, in
// order (captured where missing; the stream must be idle).
static vh_status graphs_for(vh_handle *h, int count, std::vector<hipGraphExec_t> *out) {
    const int iq0 = h->iq_cur, iqn0 = h->iq_next, iu0 = h->iu_cur, iup0 = h->iu_prev;
    vh_status s = VH_OK;
    for (int done = 0; done < count && s == VH_OK;) {
        const int c = std::min(count - done, MAX_GRAPH_STEPS);
        hipGraphExec_t ge = nullptr;
        s = get_graph(h, c, &ge);
        if (out) out->push_back(ge);
        for (int k = 0; k < c; ++k) {
            std::swap(h->iq_cur, h->iq_next);
            std::swap(h->iu_cur, h->iu_prev);
        }
        done += c;
    }
    h->iq_cur = iq0; h->iq_next = iqn0; h->iu_cur = iu0; h->iu_prev = iup0;
    return s;
}

vh_status vh_prepare_steps(vh_handle *h, int count) {
    if (!h || count < 1) return VH_ERR_ARG;
    if (!h->conserved_ready) return VH_ERR_STATE;
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    return graphs_for(h, count, nullptr);
}

vh_status vh_steps(vh_handle *h, int n_first, int count) {
    if (!h || n_first < 1 || count < 0) return VH_ERR_ARG;
    if (!h->conserved_ready) return VH_ERR_STATE;
    if (count == 0) return VH_OK;
    HIPCHK(hipSetDevice(h->dev));
    vh_status s = ensure_times(h, n_first + count);
    if (s != VH_OK) return s;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<hipGraphExec_t> ges;
    s = graphs_for(h, count, &ges);
    if (s != VH_OK) return s;
    *h->h_base = n_first;
    HIPCHK(hipMemcpyAsync(h->d_base, h->h_base, sizeof(int), hipMemcpyHostToDevice, h->stream));
    // graphs of at most MAX_GRAPH_STEPS steps, back to back: each one's last
    // node moves *d_base on, so the next continues the clock
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (hipGraphExec_t ge : ges) HIPCHK(hipGraphLaunch(ge, h->stream));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    for (int k = 0; k < count; ++k) {
        std::swap(h->iq_cur, h->iq_next);
        std::swap(h->iu_cur, h->iu_prev);
    }
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_kernel_ms = ms / (2.0 * count);
    return VH_OK;
}

double vh_last_kernel_ms(const vh_handle *h) { return h ? h->last_kernel_ms : 0.0; }

vh_status vh_inferred_variables(vh_handle *h, double t, const double *q, const double *ePrev, int64_t n, double *out,
                                int32_t *nan_site) {
    if (!h || !q || !ePrev || !out || n < 0) return VH_ERR_ARG;
    if (n == 0) return VH_OK;
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t nq = (size_t)h->nc * n;
    double *dq = nullptr;
    HIPCHK(hipMalloc(&dq, sizeof(double) * (nq + 7 * (size_t)n) + sizeof(int) * n));
    double *de = dq + nq, *dout = de + n;
    int *dnan = reinterpret_cast<int *>(dout + 6 * (size_t)n);
    hipError_t r = hipMemcpy(dq, q, sizeof(double) * nq, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(de, ePrev, sizeof(double) * n, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = launch_inferred(h->eos, h->nc, t, dq, de, n, dout, dnan, h->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(h->stream);
    if (r == hipSuccess) r = hipMemcpy(out, dout, sizeof(double) * 6 * n, hipMemcpyDeviceToHost);
    std::vector<int> nan(n);
    if (r == hipSuccess) r = hipMemcpy(nan.data(), dnan, sizeof(int) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dq);
    if (r != hipSuccess) return hip_fail(r, "vh_inferred_variables");
    bool any = false;
    for (int64_t i = 0; i < n; ++i) {
        if (nan_site) nan_site[i] = nan[i];
        any = any || nan[i] != 0;
    }
    return any ? VH_ERR_NAN : VH_OK;
}

vh_status vh_sync(vh_handle *h) {
    if (!h) return VH_ERR_ARG;
    HIPCHK(hipSetDevice(h->dev));
    return check_err(h);
}

vh_status vh_nan_info(vh_handle *h, int *site, int64_t *cell) {
    if (!h) return VH_ERR_ARG;
    int err[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(err, h->d_err, sizeof(int) * 3, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (site) *site = err[0];
    if (cell) *cell = (int64_t)(uint32_t)err[1] | ((int64_t)err[2] << 32);
    return VH_OK;
}

static double *field_ptr(vh_handle *h, int field) {
    if (field >= 0 && field < h->nc) return h->q[h->iq_cur][field];
    if (field == 15) return h->e[0];
    if (field == 16) return h->p[0];
    if (field >= 17 && field <= 20) return h->u[h->iu_cur][field - 17];
    if (field >= 21 && field <= 24) return h->u[h->iu_prev][field - 21];
    if (field == 25) return h->regfac;
    if (field == 26) return h->nc == 15 ? h->regfacb : nullptr;
    if (field >= 27 && field <= 36) return h->vort ? h->vort + (field - 27) * h->N : nullptr;
    return nullptr;
}

vh_status vh_download(vh_handle *h, int field, double *out) {
    if (!h || !out) return VH_ERR_ARG;
    HIPCHK(hipSetDevice(h->dev));
    if (field == 26 && h->nc != 15) return VH_ERR_ARG;
    double **rf = field == 25 ? &h->regfac : field == 26 ? &h->regfacb : nullptr;
    if (rf && !*rf) {
        // enable regulation-factor output from the next step on
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMalloc(rf, sizeof(double) * h->N));
        for (long s = 0; s < h->N; ++s) out[s] = 1.0;
        std::vector<double> ones(h->N, 1.0);
        HIPCHK(hipMemcpy(*rf, ones.data(), sizeof(double) * h->N, hipMemcpyHostToDevice));
        for (auto &kv : h->graphs) hipGraphExecDestroy(kv.second);
        h->graphs.clear();
        return VH_OK;
    }
    double *src = field_ptr(h, field);
    if (!src) return VH_ERR_ARG;
    HIPCHK(hipMemcpyAsync(out, src, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->stream));
    return check_err(h);
}

vh_status vh_thermal_vorticity(vh_handle *h) {
    if (!h) return VH_ERR_ARG;
    const vh_config &c = h->cfg;
    HIPCHK(hipSetDevice(h->dev));
    if (!h->vort) HIPCHK(hipMalloc(&h->vort, sizeof(double) * 10 * h->N));
    VortArgs A{};
    for (int m = 0; m < 4; ++m) {
        A.u[m] = h->u[h->iu_cur][m];
        A.up[m] = h->u[h->iu_prev][m];
    }
    A.e = h->e[0];
    A.out = h->vort;
    A.nx = c.nx;
    A.ny = c.ny;
    A.nz = c.nz;
    A.dt = c.dt;
    A.dx = c.dx;
    A.dy = c.dy;
    A.dz = c.dz;
    HIPCHK(launch_vorticity(A, h->eos, h->stream));
    return VH_OK;
}

vh_status vh_upload_field(vh_handle *h, int field, const double *in) {
    if (!h || !in) return VH_ERR_ARG;
    HIPCHK(hipSetDevice(h->dev));
    double *dst = field_ptr(h, field);
    if (!dst || field == 25 || field == 26) return VH_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(dst, in, sizeof(double) * h->N, hipMemcpyHostToDevice));
    return VH_OK;
}

vh_status vh_count_above(vh_handle *h, double eF, int64_t *count) {
    if (!h || !count) return VH_ERR_ARG;
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipMemsetAsync(h->d_count, 0, sizeof(unsigned long long), h->stream));
    HIPCHK(launch_count_above(h->e[0], h->N, eF, h->d_count, h->stream));
    unsigned long long c = 0;
    HIPCHK(hipMemcpyAsync(&c, h->d_count, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    vh_status s = check_err(h);
    *count = (int64_t)c;
    return s;
}

// surface.dat's interpolated variables as field ids, in the reference's order
// (FreezeOut.cpp:57-97): u^x u^y u^eta e pi^xx pi^xy pi^xeta pi^yy pi^yeta Pi
static constexpr int FO_FIELD[FO_NVARS] = {18, 19, 20, 15, 8, 9, 10, 11, 12, 14};

// device staging of at least bytes (the stream is idle when it grows)
static hipError_t fo_staging(vh_handle *h, size_t bytes) {
    if (bytes <= h->fo_buf_bytes) return hipSuccess;
    hipError_t r = hipStreamSynchronize(h->stream);
    if (r != hipSuccess) return r;
    const size_t want = std::max(bytes, h->fo_buf_bytes * 2);
    hipFree(h->fo_buf);
    h->fo_buf = nullptr;
    h->fo_buf_bytes = 0;
    if ((r = hipMalloc(&h->fo_buf, want)) != hipSuccess) return r;
    h->fo_buf_bytes = want;
    return hipSuccess;
}

vh_status vh_fo_flag(vh_handle *h, double eF, int64_t *ids, double *corners, int64_t cap, int64_t *n_flagged) {
    if (!h || !n_flagged || (cap > 0 && !ids)) return VH_ERR_ARG;
    const vh_config &c = h->cfg;
    if (c.nx < 2 || c.ny < 2) return VH_ERR_ARG;  // the reference's finder needs 2+1D or 3+1D
    HIPCHK(hipSetDevice(h->dev));
    const size_t b = sizeof(double) * h->N;
    const long M = (long)(c.nx - 1) * (c.ny - 1) * (c.nz > 1 ? c.nz - 1 : 1);
    if (!h->fo_prev) {
        HIPCHK(hipMalloc(&h->fo_prev, b));
        HIPCHK(hipMalloc(&h->fo_cur, b));
        // the ids, then the flag kernels' per-block scratch
        HIPCHK(hipMalloc(&h->fo_ids, sizeof(long long) * (std::max(1L, M) + fo_flag_blocks(M))));
        HIPCHK(hipMemsetAsync(h->fo_prev, 0, b, h->stream));  // calloc'd slices, HydroPlugin.cpp:159
        HIPCHK(hipMemsetAsync(h->fo_cur, 0, b, h->stream));
    }
    std::swap(h->fo_prev, h->fo_cur);
    HIPCHK(hipMemcpyAsync(h->fo_cur, h->e[0], b, hipMemcpyDeviceToDevice, h->stream));
    // the interpolated variables move with e (Pi stays zero without bulk)
    if (!h->fo_vars) {
        HIPCHK(hipMalloc(&h->fo_vars, 2 * FO_NVARS * b));
        HIPCHK(hipMemsetAsync(h->fo_vars, 0, 2 * FO_NVARS * b, h->stream));
    }
    h->fo_vslot ^= 1;
    for (int v = 0; v < FO_NVARS; ++v) {
        const double *src = field_ptr(h, FO_FIELD[v]);
        if (src)
            HIPCHK(hipMemcpyAsync(h->fo_vars + (h->fo_vslot * FO_NVARS + v) * h->N, src, b, hipMemcpyDeviceToDevice,
                                  h->stream));
    }
    HIPCHK(hipMemsetAsync(h->d_count, 0, sizeof(unsigned long long), h->stream));
    HIPCHK(launch_fo_flag(h->fo_prev, h->fo_cur, c.nx, c.ny, c.nz, eF, h->fo_ids, M, h->d_count,
                          (unsigned long long *)(h->fo_ids + std::max(1L, M)), h->stream));
    unsigned long long n = 0;
    HIPCHK(hipMemcpyAsync(&n, h->d_count, sizeof(n), hipMemcpyDeviceToHost, h->stream));
    vh_status s = check_err(h);
    if (s != VH_OK) return s;
    *n_flagged = (int64_t)n;
    h->fo_n = (long long)n;
    h->fo_eF = eF;
    if ((int64_t)n > cap) return VH_ERR_CAPACITY;
    if (n == 0) return VH_OK;
    // the ids come out in increasing order
    static_assert(sizeof(long long) == sizeof(int64_t), "cube ids are 64-bit");
    {
        void *const dst[1] = {ids};
        const void *const src[1] = {h->fo_ids};
        const size_t nb[1] = {sizeof(long long) * n};
        HIPCHK(fo_download(h->fo_surf, h->stream, 1, dst, src, nb));
    }
    if (corners) {
        const int per = c.nz > 1 ? 16 : 8;
        HIPCHK(fo_staging(h, sizeof(double) * per * n));
        HIPCHK(launch_fo_corners(h->fo_prev, h->fo_cur, c.nx, c.ny, c.nz, h->fo_ids, (long long)n, h->fo_buf,
                                 h->stream));
        HIPCHK(hipMemcpyAsync(corners, h->fo_buf, sizeof(double) * per * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return VH_OK;
}

vh_status vh_fo_interp(vh_handle *h, const int64_t *cubes, const double *frac, int64_t n, double *vars) {
    if (!h || n < 0 || (n > 0 && (!cubes || !frac || !vars))) return VH_ERR_ARG;
    const vh_config &c = h->cfg;
    if (c.nx < 2 || c.ny < 2) return VH_ERR_ARG;
    if (!h->fo_vars) {
        vh::set_last_error("vh_fo_interp: no vh_fo_flag call yet");
        return VH_ERR_ARG;
    }
    const int64_t M = (int64_t)(c.nx - 1) * (c.ny - 1) * (c.nz > 1 ? c.nz - 1 : 1);
    for (int64_t m = 0; m < n; ++m)
        if (cubes[m] < 0 || cubes[m] >= M) {
            vh::set_last_error("vh_fo_interp: cube id out of range");
            return VH_ERR_ARG;
        }
    if (n == 0) return VH_OK;
    HIPCHK(hipSetDevice(h->dev));
    // staging: ids, then 4 fractions and 10 variables per point
    const size_t ib = sizeof(long long) * n, fb = sizeof(double) * 4 * n, vb = sizeof(double) * FO_NVARS * n;
    HIPCHK(fo_staging(h, ib + fb + vb));
    char *base = (char *)h->fo_buf;
    long long *d_cubes = (long long *)base;
    double *d_frac = (double *)(base + ib), *d_vars = (double *)(base + ib + fb);
    static_assert(sizeof(long long) == sizeof(int64_t), "cube ids are 64-bit");
    HIPCHK(hipMemcpyAsync(d_cubes, cubes, ib, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(d_frac, frac, fb, hipMemcpyHostToDevice, h->stream));
    FoSlices S;
    for (int v = 0; v < FO_NVARS; ++v) {
        S.v[1][v] = h->fo_vars + (h->fo_vslot * FO_NVARS + v) * h->N;
        S.v[0][v] = h->fo_vars + ((h->fo_vslot ^ 1) * FO_NVARS + v) * h->N;
    }
    HIPCHK(launch_fo_interp(S, c.nx, c.ny, c.nz, d_cubes, d_frac, (long long)n, d_vars, h->stream));
    HIPCHK(hipMemcpyAsync(vars, d_vars, vb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_err(h);
}

vh_status vh_fo_surface(vh_handle *h, const double *widths, int *n_el, double *elements, double *vars, int64_t cap,
                        int64_t *n_elements) {
    if (!h || !widths || !n_elements || cap < 0 || (cap > 0 && (!elements || !vars))) return VH_ERR_ARG;
    const vh_config &c = h->cfg;
    if (h->fo_n < 0) {
        vh::set_last_error("vh_fo_surface: no vh_fo_flag call yet");
        return VH_ERR_ARG;
    }
    const long long n = h->fo_n;
    *n_elements = 0;
    if (n == 0) return VH_OK;
    if (!n_el) return VH_ERR_ARG;
    const int dim = c.nz > 1 ? 4 : 3;
    FoCell g{};
    for (int k = 0; k < dim; ++k) {
        if (!(widths[k] > 0)) return VH_ERR_ARG;
        g.d[k] = widths[k];
    }
    HIPCHK(hipSetDevice(h->dev));
    FoSlices S;
    for (int v = 0; v < FO_NVARS; ++v) {
        S.v[1][v] = h->fo_vars + (h->fo_vslot * FO_NVARS + v) * h->N;
        S.v[0][v] = h->fo_vars + ((h->fo_vslot ^ 1) * FO_NVARS + v) * h->N;
    }
    long long total = 0;
    int bad = 0;
    HIPCHK(fo_surface_run(h->fo_surf, h->stream, h->fo_prev, h->fo_cur, c.nx, c.ny, c.nz, h->fo_ids, n, h->fo_eF, g,
                          S, n_el, elements, vars, cap, &total, &bad));
    vh_status s = check_err(h);
    if (s != VH_OK) return s;
    if (bad) {
        vh::set_last_error("vh_fo_surface: a cut cube whose surface cannot be closed");
        return VH_ERR_STATE;
    }
    *n_elements = total;
    return total > cap ? VH_ERR_CAPACITY : VH_OK;
}

double vh_eos_energy_density(const vh_handle *h, double T) { return h ? h->heos.energy_density(T) : NAN; }
double vh_eos_pressure(const vh_handle *h, double e) { return h ? h->heos.pressure(e) : NAN; }
double vh_eos_temperature(const vh_handle *h, double e) { return h ? h->heos.temperature(e) : NAN; }

}  // extern "C"
