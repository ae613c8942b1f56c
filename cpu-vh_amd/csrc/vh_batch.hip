// Batches of events (include/cpuvh_amd.h, vh_batch_*): many independent
// lattices advance together, one launch per RK stage for all 2+1D events and
// one for all 3+1D events.  Each event keeps its own lattice, transport
// parameters and clock; the reference runs one event per process
// (rhic/src/HydroPlugin.cpp:74-321).
//
// Device state: one pool of 61 fields x capacity cells (the vh_handle layout;
// FP64, or FP32 plus the low parts of T^{tau mu} in the mixed form), each event
// owning a contiguous range [off, off + N) of every field; ranges come from a
// first-fit allocator, so admission is by HBM footprint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cpuvh_amd.h"
#include "vh_host_eos.hpp"
#include "vh_kernels.hpp"
#include "vh_mixed.hpp"

using namespace vh;

namespace {

vh_status bfail(hipError_t e, const char *what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return VH_ERR_HIP;
}

#define BCHK(call)                                        \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) return bfail(_e, #call);    \
    } while (0)

struct Event {
    bool used = false;
    vh_event_params prm{};
    long off = 0, N = 0;
    int steps_done = 0;
    double e_reg = 0;
    double *fo_prev = nullptr, *fo_cur = nullptr;
    long long *fo_ids = nullptr;
    // with cfg.fo_vars: the ten surface variables' coarse slices, [2][10][N]
    // (slot fo_vslot current); the last vh_batch_fo_flag's count and level
    double *fo_vars = nullptr;
    int fo_vslot = 0;
    long long fo_n = -1;
    double fo_eF = 0;
    bool bi = false;  // boost-invariant 2+1D: nz = 1, u^eta = pi^{tau eta} = pi^{x eta} = pi^{y eta} = 0
    bool stopped = false;  // met the stop rule in vh_batch_run; no longer stepped
    long stop_at = 0;      // the batch's step count when it stopped
};

}  // namespace

struct vh_batch {
    vh_batch_config cfg{};
    int nc = 14, dev = 0;
    hipStream_t stream = nullptr;
    // vh_batch_fo_surface's stream: it works from the flag's copies, so a
    // caller may build surfaces while steps run on `stream`
    hipStream_t side = nullptr;
    long cap = 0;
    double *pool = nullptr;
    double *q[3][15] = {};
    double *u[3][4] = {};
    double *e[2] = {}, *p[2] = {};
    int iq_cur = 0, iq_stage = 1, iq_next = 2;
    int iu_cur = 0, iu_prev = 1, iu_stage = 2;
    HostEos heos;
    DeviceEos deos;
    int *d_err = nullptr;
    std::vector<Event> ev;
    FoSurfaceWork fo_surf;  // vh_batch_fo_surface's device buffers
    std::map<long, long> free_seg;  // offset -> length
    EvDesc *d_ev = nullptr;
    std::vector<EvDesc> h_ev;
    StepTimes *d_times = nullptr;  // max_events x (max_steps + 1)
    int *d_blk_ev = nullptr, *d_blk_first = nullptr;
    // block table: [2+1D blocks | 3+1D blocks]; in a mixed batch the 2+1D
    // part starts with the nblk2bi blocks of boost-invariant events, which
    // have a launch of their own
    int nblk2 = 0, nblk3 = 0, nblk2bi = 0;
    // blocks of removed events still in the table (their kernels return at
    // once); the table is rebuilt when they exceed a tenth of it or an
    // event is added, so removing an event keeps the captured graphs
    long dead_blocks = 0;
    // 64-cell chunks of every event (counting kernels of the mixed form)
    int *d_cnt_ev = nullptr;
    long *d_cnt_start = nullptr;
    int ncnt = 0;
    // mixed form: FP32 fields, tile variants of the 2+1D and 3+1D launches
    bool mixed = false;
    float *fpool = nullptr;
    // staging of admitted events' FP64 initial states: a device buffer and two
    // pinned host slots used in turn (grown on demand); a slot's event marks
    // the end of the copies that read it
    double *stage_buf = nullptr;
    size_t stage_bytes = 0;
    struct HostSlot {
        char *h = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
    } hslot[2];
    int hslot_next = 0;
    float *qf[3][15] = {}, *qfl[3][4] = {}, *uf[3][4] = {}, *ef[2] = {}, *pf[2] = {};
    int tile2 = 0, tile3 = 0;
    int face_mode = 1;  // the reference's face semantics (vh_batch_set_face_mode)
    bool layout_dirty = true;
    int layout_version = 0;
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
    unsigned long long *d_counts = nullptr;
    // device stop rule of vh_batch_run: strikes per event and the check at
    // which an event stopped (-1: running)
    int *d_strikes = nullptr, *d_stop = nullptr;
    // steps taken by the batch: the state buffers swap roles every step, so
    // a stopped event's latest state sits in the buffers that were current
    // when it stopped, current again after an even number of steps
    long step_count = 0;
    std::map<std::tuple<int, int, int, int, int, uint64_t>, hipGraphExec_t> run_graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;  // ev0/ev1 bracket a graph launch not yet read
    double last_ms = 0;
};

namespace {

void drop_graphs(vh_batch *b) {
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    for (auto &kv : b->run_graphs) (void)hipGraphExecDestroy(kv.second);
    b->run_graphs.clear();
}

long alloc_cells(vh_batch *b, long n) {
    for (auto it = b->free_seg.begin(); it != b->free_seg.end(); ++it) {
        if (it->second >= n) {
            const long off = it->first, len = it->second;
            b->free_seg.erase(it);
            if (len > n) b->free_seg[off + n] = len - n;
            return off;
        }
    }
    return -1;
}

void free_cells(vh_batch *b, long off, long n) {
    auto it = b->free_seg.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != b->free_seg.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_seg.erase(nx);
    }
    if (it != b->free_seg.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_seg.erase(it);
        }
    }
}

// Blocks of one event's stage launch: 64-cell chunks (FP64) or tiles (mixed).
long event_blocks(const vh_batch *b, const Event &E) {
    if (!b->mixed) return (E.N + CELLS_PER_BLOCK - 1) / CELLS_PER_BLOCK;
    const int *T = E.prm.nz > 1 ? MIXED_TILES_3D[b->tile3] : MIXED_TILES_2D[b->tile2];
    return (long)((E.prm.nx + T[0] - 1) / T[0]) * ((E.prm.ny + T[1] - 1) / T[1]) * ((E.prm.nz + T[2] - 1) / T[2]);
}

vh_status rebuild_layout(vh_batch *b) {
    if (!b->layout_dirty) return VH_OK;
    std::vector<int> bev, bfirst, cev;
    std::vector<long> cstart;
    int n2 = 0, n2bi = 0;
    // pass 0: boost-invariant 2+1D events of a mixed batch; 1: other 2+1D; 2: 3+1D
    for (int pass = 0; pass < 3; ++pass)
        for (int id = 0; id < (int)b->ev.size(); ++id) {
            const Event &E = b->ev[id];
            if (!E.used) continue;
            const int cls = E.prm.nz > 1 ? 2 : (b->mixed && E.bi ? 0 : 1);
            if (cls != pass) continue;
            const long nb = event_blocks(b, E);
            // each class goes to a launch of its own, indexing its part of
            // the table from 0: first is relative to that part.
            const int first = (int)bev.size() - (pass == 0 ? 0 : (pass == 1 ? n2bi : n2));
            for (long k = 0; k < nb; ++k) {
                bev.push_back(id);
                bfirst.push_back(first);
            }
            if (pass == 0) n2bi += (int)nb;
            if (pass <= 1) n2 += (int)nb;
            for (long c = 0; c < E.N; c += 64) {
                cev.push_back(id);
                cstart.push_back(c);
            }
        }
    b->nblk2 = n2;
    b->nblk2bi = n2bi;
    b->nblk3 = (int)bev.size() - n2;
    // every block must name a tile inside its event: 0 <= k - first < blocks
    for (size_t k = 0; k < bev.size(); ++k) {
        const long seg = k < (size_t)n2bi ? 0 : (k < (size_t)n2 ? n2bi : n2);
        const long local = (long)k - seg - bfirst[k];
        if (local < 0 || local >= event_blocks(b, b->ev[bev[k]])) {
            set_last_error("batch block table inconsistent");
            return VH_ERR_STATE;
        }
    }
    b->ncnt = (int)cev.size();
    // captured graphs and queued launches read the old tables: let them
    // finish before the tables are freed and the graphs destroyed
    BCHK(hipStreamSynchronize(b->stream));
    (void)hipFree(b->d_blk_ev);
    (void)hipFree(b->d_blk_first);
    (void)hipFree(b->d_cnt_ev);
    (void)hipFree(b->d_cnt_start);
    b->d_blk_ev = b->d_blk_first = b->d_cnt_ev = nullptr;
    b->d_cnt_start = nullptr;
    if (!bev.empty()) {
        BCHK(hipMalloc(&b->d_blk_ev, sizeof(int) * bev.size()));
        BCHK(hipMalloc(&b->d_blk_first, sizeof(int) * bev.size()));
        BCHK(hipMemcpy(b->d_blk_ev, bev.data(), sizeof(int) * bev.size(), hipMemcpyHostToDevice));
        BCHK(hipMemcpy(b->d_blk_first, bfirst.data(), sizeof(int) * bev.size(), hipMemcpyHostToDevice));
        BCHK(hipMalloc(&b->d_cnt_ev, sizeof(int) * cev.size()));
        BCHK(hipMalloc(&b->d_cnt_start, sizeof(long) * cstart.size()));
        BCHK(hipMemcpy(b->d_cnt_ev, cev.data(), sizeof(int) * cev.size(), hipMemcpyHostToDevice));
        BCHK(hipMemcpy(b->d_cnt_start, cstart.data(), sizeof(long) * cstart.size(), hipMemcpyHostToDevice));
    }
    drop_graphs(b);
    b->dead_blocks = 0;
    b->layout_dirty = false;
    ++b->layout_version;
    return VH_OK;
}

StageParams batch_params(const vh_batch *b, int stage, int iq_in, int iu_in, int iu_prev_used, int ie_in,
                         int iq_out, int iu_out, int ie_out) {
    StageParams P{};
    for (int n = 0; n < 15; ++n) {
        P.q[n] = b->q[iq_in][n];
        P.qn[n] = b->q[b->iq_cur][n];
        P.qo[n] = b->q[iq_out][n];
    }
    for (int c = 0; c < 4; ++c) {
        P.u[c] = b->u[iu_in][c];
        P.up[c] = b->u[iu_prev_used][c];
        P.uo[c] = b->u[iu_out][c];
    }
    P.e = b->e[ie_in];
    P.p = b->p[ie_in];
    P.eo = b->e[ie_out];
    P.po = b->p[ie_out];
    P.stage = stage;
    P.err = b->d_err;
    return P;
}

mx::MxFields mixed_fields(const vh_batch *b, int stage, int iq_in, int iu_in, int iu_prev_used, int ie_in,
                          int iq_out, int iu_out, int ie_out) {
    mx::MxFields F{};
    F.q = b->qf[iq_in][0];
    F.qn = b->qf[b->iq_cur][0];
    F.qo = b->qf[iq_out][0];
    F.ql = b->qfl[iq_in][0];
    F.qnl = b->qfl[b->iq_cur][0];
    F.qol = b->qfl[iq_out][0];
    F.u = b->uf[iu_in][0];
    F.up = b->uf[iu_prev_used][0];
    F.uo = b->uf[iu_out][0];
    F.cap = (unsigned long)b->cap;
    F.e = b->ef[ie_in];
    F.p = b->pf[ie_in];
    F.eo = b->ef[ie_out];
    F.po = b->pf[ie_out];
    F.stage = stage;
    F.face_mode = b->face_mode;
    F.err = b->d_err;
    return F;
}

// The launches of one stage (both RK stages: stage = 0, 1) for all events.
vh_status enqueue_stage(vh_batch *b, int stage, int off) {
    const int *be2 = b->d_blk_ev, *bf2 = b->d_blk_first;
    const int *be3 = b->d_blk_ev ? b->d_blk_ev + b->nblk2 : nullptr;
    const int *bf3 = b->d_blk_first ? b->d_blk_first + b->nblk2 : nullptr;
    if (b->mixed) {
        mx::MxFields F = stage == 0 ? mixed_fields(b, 0, b->iq_cur, b->iu_cur, b->iu_prev, 0, b->iq_stage, b->iu_stage, 1)
                                    : mixed_fields(b, 1, b->iq_stage, b->iu_stage, b->iu_cur, 1, b->iq_next, b->iu_prev, 0);
        F.off = off;
        const int nbi = b->nblk2bi;
        BCHK(launch_stage_mixed(F, b->deos.E, b->nc, false, b->tile2, true, b->d_ev, be2, bf2, nbi, b->stream));
        BCHK(launch_stage_mixed(F, b->deos.E, b->nc, false, b->tile2, false, b->d_ev, be2 ? be2 + nbi : nullptr,
                                bf2 ? bf2 + nbi : nullptr, b->nblk2 - nbi, b->stream));
        BCHK(launch_stage_mixed(F, b->deos.E, b->nc, true, b->tile3, false, b->d_ev, be3, bf3, b->nblk3, b->stream));
        return VH_OK;
    }
    StageParams P = stage == 0 ? batch_params(b, 0, b->iq_cur, b->iu_cur, b->iu_prev, 0, b->iq_stage, b->iu_stage, 1)
                               : batch_params(b, 1, b->iq_stage, b->iu_stage, b->iu_cur, 1, b->iq_next, b->iu_prev, 0);
    P.off = off;
    P.tab = b->d_times;
    BCHK(launch_stage_batch(P, b->deos.E, b->nc, false, b->d_ev, be2, bf2, b->nblk2, b->stream));
    BCHK(launch_stage_batch(P, b->deos.E, b->nc, true, b->d_ev, be3, bf3, b->nblk3, b->stream));
    return VH_OK;
}

vh_status enqueue_batch_step(vh_batch *b, int off) {
    vh_status s = enqueue_stage(b, 0, off);
    if (s == VH_OK) s = enqueue_stage(b, 1, off);
    if (s != VH_OK) return s;
    std::swap(b->iq_cur, b->iq_next);
    std::swap(b->iu_cur, b->iu_prev);
    return VH_OK;
}

// flip: an odd number of steps since the event stopped (its state is in the
// buffers that are now the previous ones)
// vh_download's field numbers of the ten surface variables (u^x u^y u^eta e
// pi^xx pi^xy pi^xeta pi^yy pi^yeta Pi)
constexpr int FO_SURF_FIELD[FO_NVARS] = {18, 19, 20, 15, 8, 9, 10, 11, 12, 14};

double *field_of(vh_batch *b, int field, bool flip) {
    const int iq = flip ? b->iq_next : b->iq_cur, iu = flip ? b->iu_prev : b->iu_cur;
    const int iup = flip ? b->iu_cur : b->iu_prev;
    if (field >= 0 && field < b->nc) return b->q[iq][field];
    if (field == 15) return b->e[0];
    if (field == 16) return b->p[0];
    if (field >= 17 && field <= 20) return b->u[iu][field - 17];
    if (field >= 21 && field <= 24) return b->u[iup][field - 21];
    return nullptr;
}

bool field_of_mixed(vh_batch *b, int field, bool flip, const float **hi, const float **lo) {
    const int iq = flip ? b->iq_next : b->iq_cur, iu = flip ? b->iu_prev : b->iu_cur;
    const int iup = flip ? b->iu_cur : b->iu_prev;
    *lo = nullptr;
    if (field >= 0 && field < b->nc) {
        *hi = b->qf[iq][field];
        if (field < 4) *lo = b->qfl[iq][field];
        return true;
    }
    if (field == 15) *hi = b->ef[0];
    else if (field == 16) *hi = b->pf[0];
    else if (field >= 17 && field <= 20) *hi = b->uf[iu][field - 17];
    else if (field >= 21 && field <= 24) *hi = b->uf[iup][field - 21];
    else return false;
    return true;
}

}  // namespace

extern "C" {

vh_status vh_batch_create(const vh_batch_config *cfg, vh_batch **out) {
    if (!cfg || !out) return VH_ERR_ARG;
    *out = nullptr;
    const vh_batch_config &c = *cfg;
    if (c.capacity_cells < 1 || c.max_events < 1 || c.max_steps < 1 ||
        (c.eos_kind != VH_EOS_CONFORMAL && c.eos_kind != VH_EOS_TABLE) || (c.eos_kind == VH_EOS_TABLE && !c.eos_table))
        return VH_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_last_error("no HIP device");
        return VH_ERR_HIP;
    }
    if (c.device < 0 || c.device >= ndev) return VH_ERR_ARG;
    vh_batch *b = new vh_batch();
    auto fail = [&](vh_status s) {
        vh_batch_destroy(b);
        return s;
    };
    b->cfg = c;
    b->cfg.eos_table = nullptr;
    b->dev = c.device;
    b->nc = c.bulk ? 15 : 14;
    b->cap = (long)c.capacity_cells;
    if (hipSetDevice(b->dev) != hipSuccess) return fail(VH_ERR_HIP);
    if (c.eos_kind == VH_EOS_TABLE) b->heos.load(c.eos_table);
    if (c.cs2_mode != 0 && c.cs2_mode != 1) return fail(VH_ERR_ARG);
    if (b->deos.init(b->heos, c.eos_kind, c.cs2_mode) != 0) return fail(VH_ERR_NOMEM);
    if (c.precision != VH_PREC_FP64 && c.precision != VH_PREC_MIXED) return fail(VH_ERR_ARG);
    b->mixed = c.precision == VH_PREC_MIXED;
    if (b->mixed && b->cap >= (1L << 30)) return fail(VH_ERR_ARG);  // 32-bit byte offsets per field
    const size_t nfield = 3 * 15 + 3 * 4 + 4;
    if (!b->mixed) {
        if (hipMalloc(&b->pool, sizeof(double) * nfield * b->cap) != hipSuccess) return fail(VH_ERR_NOMEM);
        double *ptr = b->pool;
        for (int s = 0; s < 3; ++s)
            for (int n = 0; n < 15; ++n, ptr += b->cap) b->q[s][n] = ptr;
        for (int s = 0; s < 3; ++s)
            for (int k = 0; k < 4; ++k, ptr += b->cap) b->u[s][k] = ptr;
        for (int s = 0; s < 2; ++s, ptr += b->cap) b->e[s] = ptr;
        for (int s = 0; s < 2; ++s, ptr += b->cap) b->p[s] = ptr;
    } else {
        if (hipMalloc(&b->fpool, sizeof(float) * (nfield + 3 * 4) * b->cap) != hipSuccess) return fail(VH_ERR_NOMEM);
        float *ptr = b->fpool;
        for (int s = 0; s < 3; ++s)
            for (int n = 0; n < 15; ++n, ptr += b->cap) b->qf[s][n] = ptr;
        for (int s = 0; s < 3; ++s)
            for (int n = 0; n < 4; ++n, ptr += b->cap) b->qfl[s][n] = ptr;
        for (int s = 0; s < 3; ++s)
            for (int k = 0; k < 4; ++k, ptr += b->cap) b->uf[s][k] = ptr;
        for (int s = 0; s < 2; ++s, ptr += b->cap) b->ef[s] = ptr;
        for (int s = 0; s < 2; ++s, ptr += b->cap) b->pf[s] = ptr;
    }
    if (hipMalloc(&b->d_err, sizeof(int) * 4) != hipSuccess || hipMemset(b->d_err, 0, sizeof(int) * 4) != hipSuccess)
        return fail(VH_ERR_NOMEM);
    if (hipMalloc(&b->d_ev, sizeof(EvDesc) * c.max_events) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipMalloc(&b->d_counts, sizeof(unsigned long long) * c.max_events) != hipSuccess) return fail(VH_ERR_NOMEM);
    if (hipMalloc(&b->d_strikes, sizeof(int) * c.max_events) != hipSuccess ||
        hipMalloc(&b->d_stop, sizeof(int) * c.max_events) != hipSuccess)
        return fail(VH_ERR_NOMEM);
    if (hipMalloc(&b->d_times, sizeof(StepTimes) * (size_t)c.max_events * (c.max_steps + 1)) != hipSuccess)
        return fail(VH_ERR_NOMEM);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess)
        return fail(VH_ERR_HIP);
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return fail(VH_ERR_HIP);
    for (auto &H : b->hslot)
        if (hipEventCreateWithFlags(&H.done, hipEventDisableTiming) != hipSuccess) return fail(VH_ERR_HIP);
    b->ev.resize(c.max_events);
    b->h_ev.resize(c.max_events);
    std::memset(b->h_ev.data(), 0, sizeof(EvDesc) * c.max_events);
    if (hipMemcpy(b->d_ev, b->h_ev.data(), sizeof(EvDesc) * c.max_events, hipMemcpyHostToDevice) != hipSuccess)
        return fail(VH_ERR_HIP);
    b->free_seg[0] = b->cap;
    *out = b;
    return VH_OK;
}

vh_status vh_batch_destroy(vh_batch *b) {
    if (!b) return VH_ERR_ARG;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    drop_graphs(b);
    for (auto &E : b->ev) {
        (void)hipFree(E.fo_prev);
        (void)hipFree(E.fo_cur);
        (void)hipFree(E.fo_ids);
        (void)hipFree(E.fo_vars);
    }
    fo_surface_free(b->fo_surf);
    (void)hipFree(b->pool);
    (void)hipFree(b->fpool);
    (void)hipFree(b->stage_buf);
    for (auto &H : b->hslot) {
        (void)hipHostFree(H.h);
        if (H.done) (void)hipEventDestroy(H.done);
    }
    (void)hipFree(b->d_cnt_ev);
    (void)hipFree(b->d_cnt_start);
    (void)hipFree(b->d_err);
    (void)hipFree(b->d_ev);
    (void)hipFree(b->d_counts);
    (void)hipFree(b->d_strikes);
    (void)hipFree(b->d_stop);
    (void)hipFree(b->d_times);
    (void)hipFree(b->d_blk_ev);
    (void)hipFree(b->d_blk_first);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    b->deos.release();
    delete b;
    return VH_OK;
}

vh_status vh_batch_add(vh_batch *b, const vh_event_params *prm, const double *e, const double *p, const double *ut,
                       const double *ux, const double *uy, const double *un, const double *pi10, const double *Pi,
                       int *id_out) {
    if (!b || !prm || !e || !p || !id_out) return VH_ERR_ARG;
    const vh_event_params &c = *prm;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1 || !(c.dx > 0) || !(c.dy > 0) || !(c.dz > 0) || !(c.dt > 0) ||
        !(c.t0 > 0) || !(c.etabar > 0) || !(c.zeta_Tpeak_GeV >= 0.0) || !(c.zeta_peak >= 0.0))
        return VH_ERR_ARG;
    int id = -1;
    for (int k = 0; k < (int)b->ev.size(); ++k)
        if (!b->ev[k].used) {
            id = k;
            break;
        }
    if (id < 0) return VH_ERR_CAPACITY;
    const long N = (long)c.nx * c.ny * c.nz;
    BCHK(hipSetDevice(b->dev));
    const long off = alloc_cells(b, N);
    if (off < 0) return VH_ERR_CAPACITY;
    Event &E = b->ev[id];
    E = Event();
    E.used = true;
    E.prm = c;
    E.off = off;
    E.N = N;
    E.steps_done = 0;
    E.e_reg = b->heos.energy_density(c.T_reg_GeV / 0.197326938);
    E.bi = c.nz == 1;
    for (long s = 0; s < N && E.bi && un; ++s) E.bi = un[s] == 0.0;
    for (long s = 0; s < N && E.bi && pi10; ++s)
        E.bi = pi10[3 * N + s] == 0.0 && pi10[6 * N + s] == 0.0 && pi10[8 * N + s] == 0.0;
    EvDesc &D = b->h_ev[id];
    D = EvDesc{};
    D.nx = c.nx; D.ny = c.ny; D.nz = c.nz; D.active = 1;
    D.off = off;
    D.dx = c.dx; D.dy = c.dy; D.dz = c.dz; D.dt = c.dt;
    D.e_reg = E.e_reg;
    D.etas_min = c.etabar;
    D.etas_Tk_GeV = c.etas_Tk_GeV;
    D.etas_slope_lo = c.etas_slope_lo;
    D.etas_slope_hi = c.etas_slope_hi;
    D.zeta_Tpeak_GeV = c.zeta_Tpeak_GeV;
    D.zeta_peak = c.zeta_peak;
    D.tab = b->d_times + (size_t)id * (b->cfg.max_steps + 1);
    D.base = 1;
    auto undo = [&](hipError_t r, const char *what) {
        free_cells(b, off, N);
        E = Event();
        D = EvDesc{};
        return bfail(r, what);
    };
    // One pinned slot holds the arrays given (packed: e, p, u^tau, u^x, u^y,
    // u^eta, pi (10 N), Pi), the event's clock and its descriptor; three
    // asynchronous copies take it to the device and one kernel writes the
    // event's cells.  The call returns without waiting: the slot is reused
    // two admissions later, once its copies have ended.
    const double *src[8] = {e, p, ut, ux, uy, un, pi10, Pi};
    const long len[8] = {1, 1, 1, 1, 1, 1, 10, 1};
    long nd = 0;
    for (int k = 0; k < 8; ++k)
        if (src[k]) nd += len[k];
    const size_t arr_bytes = sizeof(double) * (size_t)(nd * N);
    const size_t tab_n = (size_t)b->cfg.max_steps + 1;
    const size_t tab_at = (arr_bytes + 255) & ~(size_t)255, desc_at = tab_at + sizeof(StepTimes) * tab_n;
    const size_t slot_bytes = desc_at + sizeof(EvDesc);
    vh_batch::HostSlot &H = b->hslot[b->hslot_next];
    b->hslot_next ^= 1;
    hipError_t r = hipEventSynchronize(H.done);
    if (r != hipSuccess) return undo(r, "admission staging");
    if (H.bytes < slot_bytes) {
        (void)hipHostFree(H.h);
        H.h = nullptr;
        H.bytes = 0;
        if ((r = hipHostMalloc((void **)&H.h, slot_bytes)) != hipSuccess) return undo(r, "admission staging");
        H.bytes = slot_bytes;
    }
    if (b->stage_bytes < arr_bytes) {  // the device buffer: drained before it is replaced
        if ((r = hipStreamSynchronize(b->stream)) != hipSuccess) return undo(r, "admission staging");
        (void)hipFree(b->stage_buf);
        b->stage_buf = nullptr;
        b->stage_bytes = 0;
        if ((r = hipMalloc(&b->stage_buf, arr_bytes)) != hipSuccess) return undo(r, "admission staging");
        b->stage_bytes = arr_bytes;
    }
    AdmitSrc A{};
    const double **dst[8] = {&A.e, &A.p, &A.u[0], &A.u[1], &A.u[2], &A.u[3], &A.pi10, &A.Pi};
    long at = 0;
    for (int k = 0; k < 8; ++k) {
        if (!src[k]) continue;
        std::memcpy((double *)H.h + at * N, src[k], sizeof(double) * (size_t)(len[k] * N));
        *dst[k] = b->stage_buf + at * N;
        at += len[k];
    }
    // the event's clock (HydroPlugin.cpp:226,307): pow(t, 2) as the reference
    // writes it (both compilers make it t * t, which is exact)
    StepTimes *tab = (StepTimes *)(H.h + tab_at);
    tab[0] = StepTimes{0, 0, 0, 0};
    for (size_t n = 1; n < tab_n; ++n) {
        const double t = (n == 1) ? c.t0 : c.t0 + (double)(n - 1) * c.dt;
        const double t2 = t + c.dt;
        tab[n] = StepTimes{t, pow(t, 2), t2, pow(t2, 2)};
    }
    std::memcpy(H.h + desc_at, &D, sizeof(EvDesc));
    r = hipMemcpyAsync(b->stage_buf, H.h, arr_bytes, hipMemcpyHostToDevice, b->stream);
    if (r == hipSuccess)
        r = hipMemcpyAsync(b->d_times + (size_t)id * tab_n, tab, sizeof(StepTimes) * tab_n, hipMemcpyHostToDevice,
                           b->stream);
    if (r == hipSuccess)
        r = hipMemcpyAsync(b->d_ev + id, H.h + desc_at, sizeof(EvDesc), hipMemcpyHostToDevice, b->stream);
    if (r == hipSuccess) r = hipEventRecord(H.done, b->stream);
    if (r == hipSuccess) {
        if (!b->mixed) {
            r = launch_admit_fp64(A, b->q[b->iq_cur][0], b->e[0], b->p[0], b->u[b->iu_cur][0], b->u[b->iu_prev][0],
                                  b->cap, N, off, b->nc, b->d_strikes + id, b->d_stop + id, b->stream);
        } else {
            mx::MxFields F{};
            F.qo = b->qf[b->iq_cur][0];
            F.qol = b->qfl[b->iq_cur][0];
            F.uo = b->uf[b->iu_cur][0];
            F.cap = (unsigned long)b->cap;
            F.eo = b->ef[0];
            F.po = b->pf[0];
            r = launch_init_mixed(F, A, b->uf[b->iu_prev][0], N, off, b->nc, b->d_strikes + id, b->d_stop + id,
                                  b->stream);
        }
    }
    if (r != hipSuccess) {  // nothing of the event may stay in flight once its cells are free again
        (void)hipStreamSynchronize(b->stream);
        return undo(r, "initial state");
    }
    b->layout_dirty = true;
    *id_out = id;
    return VH_OK;
}

vh_status vh_batch_remove(vh_batch *b, int id) {
    if (!b || id < 0 || id >= (int)b->ev.size() || !b->ev[id].used) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    BCHK(hipStreamSynchronize(b->stream));
    Event &E = b->ev[id];
    if (!b->layout_dirty) b->dead_blocks += event_blocks(b, E);
    free_cells(b, E.off, E.N);
    (void)hipFree(E.fo_prev);
    (void)hipFree(E.fo_cur);
    (void)hipFree(E.fo_ids);
    (void)hipFree(E.fo_vars);
    E = Event();
    b->h_ev[id].active = 0;
    BCHK(hipMemcpy(b->d_ev + id, &b->h_ev[id], sizeof(EvDesc), hipMemcpyHostToDevice));
    if (10 * b->dead_blocks > (long)b->nblk2 + b->nblk3) b->layout_dirty = true;
    return VH_OK;
}

namespace {

// Graph of `count` batched steps (count <= MAX_GRAPH_STEPS) from the current
// buffer roles, ending with the node that moves every event's clock on.
vh_status steps_graph(vh_batch *b, int count, hipGraphExec_t *out) {
    const auto key = std::make_tuple(count, b->iq_cur * 3 + b->iu_cur, b->layout_version);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        const int iq0 = b->iq_cur, iqn0 = b->iq_next, iu0 = b->iu_cur, iup0 = b->iu_prev;
        hipGraph_t g = nullptr;
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        vh_status cs = VH_OK;
        for (int k = 0; k < count && cs == VH_OK; ++k) cs = enqueue_batch_step(b, k);
        if (cs == VH_OK) {
            hipError_t ae = launch_advance_clocks(b->d_ev, (int)b->ev.size(), count, b->stream);
            if (ae != hipSuccess) cs = bfail(ae, "advance clocks");
        }
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        b->iq_cur = iq0; b->iq_next = iqn0; b->iu_cur = iu0; b->iu_prev = iup0;
        if (cs != VH_OK) return cs;
        if (ce != hipSuccess) return bfail(ce, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return bfail(ie, "hipGraphInstantiate");
        it = b->graphs.emplace(key, ge).first;
    }
    *out = it->second;
    return VH_OK;
}

// Graph of `nchecks` rounds of (coarse steps, per-event count above eF,
// strike) numbered from check_off, ending with the clock node
// (coarse * nchecks <= MAX_GRAPH_STEPS unless one round is longer).
vh_status run_graph(vh_batch *b, double eF, int coarse, int nchecks, int check_off, hipGraphExec_t *out) {
    const int nev = (int)b->ev.size();
    // the graph's count nodes carry eF: its bits are part of the key
    uint64_t eF_bits = 0;
    std::memcpy(&eF_bits, &eF, sizeof eF_bits);
    const auto key = std::make_tuple(coarse, nchecks, check_off, b->iq_cur * 3 + b->iu_cur, b->layout_version, eF_bits);
    auto it = b->run_graphs.find(key);
    if (it == b->run_graphs.end()) {
        const int iq0 = b->iq_cur, iqn0 = b->iq_next, iu0 = b->iu_cur, iup0 = b->iu_prev;
        hipGraph_t g = nullptr;
        BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        vh_status cs = VH_OK;
        for (int c = 0; c < nchecks && cs == VH_OK; ++c) {
            for (int k = 0; k < coarse && cs == VH_OK; ++k) cs = enqueue_batch_step(b, c * coarse + k);
            hipError_t ke = launch_zero_counts(b->d_counts, nev, b->stream);
            if (ke == hipSuccess)
                ke = b->mixed ? launch_count_above_batch_f(b->ef[0], eF, b->d_ev, b->d_cnt_ev, b->d_cnt_start,
                                                           b->ncnt, b->d_counts, b->stream)
                              : launch_count_above_batch(b->e[0], eF, b->d_ev, b->d_cnt_ev, b->d_cnt_start,
                                                         b->ncnt, b->d_counts, b->stream);
            if (ke == hipSuccess)
                ke = launch_strikes(b->d_ev, b->d_counts, b->d_strikes, b->d_stop, nev, check_off + c, b->stream);
            if (cs == VH_OK && ke != hipSuccess) cs = bfail(ke, "stop rule");
        }
        if (cs == VH_OK) {
            hipError_t ae = launch_advance_clocks(b->d_ev, nev, coarse * nchecks, b->stream);
            if (ae != hipSuccess) cs = bfail(ae, "advance clocks");
        }
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        b->iq_cur = iq0; b->iq_next = iqn0; b->iu_cur = iu0; b->iu_prev = iup0;
        if (cs != VH_OK) return cs;
        if (ce != hipSuccess) return bfail(ce, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return bfail(ie, "hipGraphInstantiate");
        it = b->run_graphs.emplace(key, ge).first;
    }
    *out = it->second;
    return VH_OK;
}

void rotate_roles(vh_batch *b, int steps) {
    for (int k = 0; k < steps; ++k) {
        std::swap(b->iq_cur, b->iq_next);
        std::swap(b->iu_cur, b->iu_prev);
    }
}

}  // namespace

vh_status vh_batch_steps(vh_batch *b, int count) {
    if (!b || count < 0) return VH_ERR_ARG;
    if (count == 0) return VH_OK;
    BCHK(hipSetDevice(b->dev));
    for (const Event &E : b->ev)
        if (E.used && !E.stopped && E.steps_done + count > b->cfg.max_steps) return VH_ERR_CAPACITY;
    vh_status s = rebuild_layout(b);
    if (s != VH_OK) return s;
    if (b->nblk2 + b->nblk3 == 0) return VH_OK;
    // capture first (on an idle stream), then replay the chunks back to back
    BCHK(hipStreamSynchronize(b->stream));
    std::vector<hipGraphExec_t> ges;
    const int iq0 = b->iq_cur, iqn0 = b->iq_next, iu0 = b->iu_cur, iup0 = b->iu_prev;
    for (int done = 0; done < count && s == VH_OK;) {
        const int c = std::min(count - done, MAX_GRAPH_STEPS);
        hipGraphExec_t ge = nullptr;
        s = steps_graph(b, c, &ge);
        ges.push_back(ge);
        rotate_roles(b, c);
        done += c;
    }
    b->iq_cur = iq0; b->iq_next = iqn0; b->iu_cur = iu0; b->iu_prev = iup0;
    if (s != VH_OK) return s;
    BCHK(hipEventRecord(b->ev0, b->stream));
    for (hipGraphExec_t ge : ges) BCHK(hipGraphLaunch(ge, b->stream));
    BCHK(hipEventRecord(b->ev1, b->stream));
    b->timed = true;
    rotate_roles(b, count);
    for (Event &E : b->ev)
        if (E.used && !E.stopped) E.steps_done += count;
    b->step_count += count;
    return VH_OK;
}

vh_status vh_batch_run(vh_batch *b, double eF, int coarse, int nchecks, int32_t *stopped, int32_t *steps_done) {
    if (!b || coarse < 1 || nchecks < 1) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    const int total = coarse * nchecks;
    for (const Event &E : b->ev)
        if (E.used && !E.stopped && E.steps_done + total > b->cfg.max_steps) return VH_ERR_CAPACITY;
    vh_status s = rebuild_layout(b);
    if (s != VH_OK) return s;
    const int nev = (int)b->ev.size();
    if (b->nblk2 + b->nblk3 > 0) {
        // whole checks per graph, at most MAX_GRAPH_STEPS steps (one check if
        // a check is longer); captured first on an idle stream
        const int per = std::max(1, MAX_GRAPH_STEPS / coarse);
        BCHK(hipStreamSynchronize(b->stream));
        std::vector<hipGraphExec_t> ges;
        const int iq0 = b->iq_cur, iqn0 = b->iq_next, iu0 = b->iu_cur, iup0 = b->iu_prev;
        for (int done = 0; done < nchecks && s == VH_OK;) {
            const int c = std::min(nchecks - done, per);
            hipGraphExec_t ge = nullptr;
            s = run_graph(b, eF, coarse, c, done, &ge);
            ges.push_back(ge);
            rotate_roles(b, c * coarse);
            done += c;
        }
        b->iq_cur = iq0; b->iq_next = iqn0; b->iu_cur = iu0; b->iu_prev = iup0;
        if (s != VH_OK) return s;
        BCHK(hipEventRecord(b->ev0, b->stream));
        for (hipGraphExec_t ge : ges) BCHK(hipGraphLaunch(ge, b->stream));
        BCHK(hipEventRecord(b->ev1, b->stream));
        b->timed = true;
        rotate_roles(b, total);
    }
    const long step_count0 = b->step_count;
    b->step_count += total;
    std::vector<int> stop(nev, -1);
    BCHK(hipMemcpyAsync(stop.data(), b->d_stop, sizeof(int) * nev, hipMemcpyDeviceToHost, b->stream));
    const vh_status st = vh_batch_sync(b);
    for (int id = 0; id < nev; ++id) {
        Event &E = b->ev[id];
        if (E.used && !E.stopped) {
            if (stop[id] >= 0) {
                E.stopped = true;
                E.steps_done += coarse * (stop[id] + 1);
                E.stop_at = step_count0 + coarse * (stop[id] + 1);
            } else {
                E.steps_done += total;
            }
        }
        if (stopped) stopped[id] = E.used ? (E.stopped ? 1 : 0) : -1;
        if (steps_done) steps_done[id] = E.used ? E.steps_done : -1;
    }
    return st;
}

vh_status vh_batch_sync(vh_batch *b) {
    if (!b) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    int err[4] = {0, 0, 0, 0};
    BCHK(hipMemcpyAsync(err, b->d_err, sizeof(int) * 3, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    if (b->timed) {
        float ms = 0;
        BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->last_ms = ms;
        b->timed = false;
    }
    return err[0] ? VH_ERR_NAN : VH_OK;
}

double vh_batch_last_ms(const vh_batch *b) { return b ? b->last_ms : 0.0; }

vh_status vh_batch_counts_above(vh_batch *b, double eF, int64_t *counts) {
    if (!b || !counts) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    vh_status s = rebuild_layout(b);
    if (s != VH_OK) return s;
    BCHK(hipMemsetAsync(b->d_counts, 0, sizeof(unsigned long long) * b->ev.size(), b->stream));
    if (b->mixed)
        BCHK(launch_count_above_batch_f(b->ef[0], eF, b->d_ev, b->d_cnt_ev, b->d_cnt_start, b->ncnt, b->d_counts,
                                        b->stream));
    else
        BCHK(launch_count_above_batch(b->e[0], eF, b->d_ev, b->d_cnt_ev, b->d_cnt_start, b->ncnt, b->d_counts,
                                      b->stream));
    std::vector<unsigned long long> c(b->ev.size());
    BCHK(hipMemcpyAsync(c.data(), b->d_counts, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost,
                        b->stream));
    s = vh_batch_sync(b);
    for (size_t id = 0; id < c.size(); ++id) counts[id] = b->ev[id].used ? (int64_t)c[id] : -1;
    return s;
}

vh_status vh_batch_download(vh_batch *b, int id, int field, double *out) {
    if (!b || !out || id < 0 || id >= (int)b->ev.size() || !b->ev[id].used) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    const Event &E = b->ev[id];
    const bool flip = E.stopped && ((b->step_count - E.stop_at) & 1);
    if (!b->mixed) {
        double *src = field_of(b, field, flip);
        if (!src) return VH_ERR_ARG;
        BCHK(hipMemcpyAsync(out, src + E.off, sizeof(double) * E.N, hipMemcpyDeviceToHost, b->stream));
        return vh_batch_sync(b);
    }
    const float *hi = nullptr, *lo = nullptr;
    if (!field_of_mixed(b, field, flip, &hi, &lo)) return VH_ERR_ARG;
    // the eta components of a boost-invariant event are zero and not stored
    if (E.bi && (field == 3 || field == 7 || field == 10 || field == 12 || field == 20 || field == 24)) {
        std::fill(out, out + E.N, 0.0);
        return vh_batch_sync(b);
    }
    double *tmp = nullptr;
    BCHK(hipMalloc(&tmp, sizeof(double) * E.N));
    hipError_t r = launch_widen(hi + E.off, lo ? lo + E.off : nullptr, tmp, E.N, b->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(out, tmp, sizeof(double) * E.N, hipMemcpyDeviceToHost, b->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (r != hipSuccess) return bfail(r, "vh_batch_download");
    return vh_batch_sync(b);
}

vh_status vh_batch_event_info(vh_batch *b, int id, int *active, int *steps_done, int64_t *cells) {
    if (!b || id < 0 || id >= (int)b->ev.size()) return VH_ERR_ARG;
    const Event &E = b->ev[id];
    if (active) *active = E.used ? 1 : 0;
    if (steps_done) *steps_done = E.steps_done;
    if (cells) *cells = E.N;
    return VH_OK;
}

vh_status vh_batch_free_cells(vh_batch *b, int64_t *free_cells, int64_t *largest) {
    if (!b) return VH_ERR_ARG;
    long tot = 0, big = 0;
    for (auto &kv : b->free_seg) {
        tot += kv.second;
        big = std::max(big, kv.second);
    }
    if (free_cells) *free_cells = tot;
    if (largest) *largest = big;
    return VH_OK;
}

// The device work of one event's flag call, queued on the batch's stream:
// the slices move on (e and, with fo_vars, the ten surface variables) and the
// cut cubes' ids go to the event's id buffer in increasing order, their count
// to *count (device).
static vh_status fo_flag_enqueue(vh_batch *b, int id, double eF, unsigned long long *count) {
    Event &E = b->ev[id];
    const vh_event_params &c = E.prm;
    if (c.nx < 2 || c.ny < 2) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    const size_t bytes = sizeof(double) * E.N;
    const long M = (long)(c.nx - 1) * (c.ny - 1) * (c.nz > 1 ? c.nz - 1 : 1);
    if (!E.fo_prev) {
        BCHK(hipMalloc(&E.fo_prev, bytes));
        BCHK(hipMalloc(&E.fo_cur, bytes));
        // the ids, then the flag kernels' per-block scratch
        BCHK(hipMalloc(&E.fo_ids, sizeof(long long) * (std::max(1L, M) + fo_flag_blocks(M))));
        BCHK(hipMemsetAsync(E.fo_prev, 0, bytes, b->stream));
        BCHK(hipMemsetAsync(E.fo_cur, 0, bytes, b->stream));
    }
    std::swap(E.fo_prev, E.fo_cur);
    if (b->mixed) BCHK(launch_widen(b->ef[0] + E.off, nullptr, E.fo_cur, E.N, b->stream));
    else BCHK(hipMemcpyAsync(E.fo_cur, b->e[0] + E.off, bytes, hipMemcpyDeviceToDevice, b->stream));
    if (b->cfg.fo_vars) {
        // the ten surface variables move with e (the eta components of a
        // boost-invariant event and Pi without bulk are zero)
        if (!E.fo_vars) {
            BCHK(hipMalloc(&E.fo_vars, 2 * FO_NVARS * bytes));
            BCHK(launch_fill(E.fo_vars, 2 * FO_NVARS * E.N, 0.0, b->stream));
        }
        E.fo_vslot ^= 1;
        const bool flip = E.stopped && ((b->step_count - E.stop_at) & 1);
        WidenSet W{};  // the mixed form's slices: one launch
        for (int v = 0; v < FO_NVARS; ++v) {
            double *dst = E.fo_vars + (E.fo_vslot * FO_NVARS + v) * E.N;
            const int f = FO_SURF_FIELD[v];
            const bool zero = (f >= b->nc && f < 15) || (b->mixed && E.bi && (f == 10 || f == 12 || f == 20));
            if (b->mixed) {
                const float *hi = nullptr, *lo = nullptr;
                if (!zero && !field_of_mixed(b, f, flip, &hi, &lo)) return VH_ERR_STATE;
                W.hi[W.count] = zero ? nullptr : hi + E.off;
                W.lo[W.count] = zero || !lo ? nullptr : lo + E.off;
                W.out[W.count++] = dst;
            } else if (zero) {
                BCHK(launch_fill(dst, E.N, 0.0, b->stream));
            } else {
                BCHK(hipMemcpyAsync(dst, field_of(b, f, flip) + E.off, bytes, hipMemcpyDeviceToDevice, b->stream));
            }
        }
        if (b->mixed) BCHK(launch_widen_set(W, E.N, b->stream));
    }
    BCHK(launch_fo_flag(E.fo_prev, E.fo_cur, c.nx, c.ny, c.nz, eF, E.fo_ids, M, count,
                        (unsigned long long *)(E.fo_ids + std::max(1L, M)), b->stream));
    return VH_OK;
}

vh_status vh_batch_fo_flag(vh_batch *b, int id, double eF, int64_t *ids, double *corners, int64_t cap,
                           int64_t *n_flagged) {
    if (!b || !n_flagged || id < 0 || id >= (int)b->ev.size() || !b->ev[id].used || (cap > 0 && !ids))
        return VH_ERR_ARG;
    Event &E = b->ev[id];
    const vh_event_params &c = E.prm;
    if (c.nx < 2 || c.ny < 2) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    BCHK(hipMemsetAsync(b->d_counts, 0, sizeof(unsigned long long), b->stream));
    vh_status fs = fo_flag_enqueue(b, id, eF, b->d_counts);
    if (fs != VH_OK) return fs;
    unsigned long long n = 0;
    BCHK(hipMemcpyAsync(&n, b->d_counts, sizeof(n), hipMemcpyDeviceToHost, b->stream));
    vh_status s = vh_batch_sync(b);
    if (s != VH_OK) return s;
    *n_flagged = (int64_t)n;
    E.fo_n = (long long)n;
    E.fo_eF = eF;
    if ((int64_t)n > cap) return VH_ERR_CAPACITY;
    if (n == 0) return VH_OK;
    // the ids come out in increasing order
    {
        void *const dst[1] = {ids};
        const void *const src[1] = {E.fo_ids};
        const size_t nb[1] = {sizeof(long long) * n};
        BCHK(fo_download(b->fo_surf, b->stream, 1, dst, src, nb));
    }
    if (corners) {
        const int per = c.nz > 1 ? 16 : 8;
        double *dc = nullptr;
        BCHK(hipMalloc(&dc, sizeof(double) * per * n));
        hipError_t le = launch_fo_corners(E.fo_prev, E.fo_cur, c.nx, c.ny, c.nz, E.fo_ids, (long long)n, dc, b->stream);
        if (le == hipSuccess) le = hipMemcpyAsync(corners, dc, sizeof(double) * per * n, hipMemcpyDeviceToHost, b->stream);
        if (le == hipSuccess) le = hipStreamSynchronize(b->stream);
        (void)hipFree(dc);
        if (le != hipSuccess) return bfail(le, "fo corners");
    }
    return VH_OK;
}

vh_status vh_batch_fo_flags(vh_batch *b, int n, const int *ids, double eF, int64_t *const *cube_ids,
                            const int64_t *cap, int64_t *n_flagged) {
    if (!b || n < 0 || n > (int)b->ev.size() || (n > 0 && (!ids || !cube_ids || !cap || !n_flagged)))
        return VH_ERR_ARG;
    for (int k = 0; k < n; ++k) {
        const int id = ids[k];
        if (id < 0 || id >= (int)b->ev.size() || !b->ev[id].used || cap[k] < 0 || (cap[k] > 0 && !cube_ids[k]))
            return VH_ERR_ARG;
        const vh_event_params &c = b->ev[id].prm;
        if (c.nx < 2 || c.ny < 2) return VH_ERR_ARG;
        for (int j = 0; j < k; ++j)
            if (ids[j] == id) return VH_ERR_ARG;  // an event's slices move once per pass
    }
    if (n == 0) return VH_OK;
    BCHK(hipSetDevice(b->dev));
    BCHK(hipMemsetAsync(b->d_counts, 0, sizeof(unsigned long long) * n, b->stream));
    for (int k = 0; k < n; ++k) {
        vh_status fs = fo_flag_enqueue(b, ids[k], eF, b->d_counts + k);
        if (fs != VH_OK) return fs;
    }
    std::vector<unsigned long long> cnt(n);
    {
        void *const dst[1] = {cnt.data()};
        const void *const src[1] = {b->d_counts};
        const size_t nb[1] = {sizeof(unsigned long long) * n};
        BCHK(fo_download(b->fo_surf, b->stream, 1, dst, src, nb));
    }
    vh_status s = vh_batch_sync(b);  // a NaN or other device error reported by the kernels
    if (s != VH_OK) return s;
    std::vector<void *> dst;
    std::vector<const void *> src;
    std::vector<size_t> nb;
    bool room = true;
    for (int k = 0; k < n; ++k) {
        Event &E = b->ev[ids[k]];
        n_flagged[k] = (int64_t)cnt[k];
        E.fo_n = (long long)cnt[k];
        E.fo_eF = eF;
        if ((int64_t)cnt[k] > cap[k]) {
            room = false;
            continue;
        }
        if (cnt[k] == 0) continue;
        dst.push_back(cube_ids[k]);
        src.push_back(E.fo_ids);
        nb.push_back(sizeof(long long) * cnt[k]);
    }
    // the ids come out in increasing order
    if (!dst.empty()) BCHK(fo_download(b->fo_surf, b->stream, (int)dst.size(), dst.data(), src.data(), nb.data()));
    return room ? VH_OK : VH_ERR_CAPACITY;
}

vh_status vh_batch_fo_surface(vh_batch *b, int id, const double *widths, int *n_el, double *elements, double *vars,
                              int64_t cap, int64_t *n_elements) {
    if (!b || !widths || !n_elements || cap < 0 || (cap > 0 && (!elements || !vars)) || id < 0 ||
        id >= (int)b->ev.size() || !b->ev[id].used)
        return VH_ERR_ARG;
    Event &E = b->ev[id];
    const vh_event_params &c = E.prm;
    if (!b->cfg.fo_vars || E.fo_n < 0) {
        vh::set_last_error("vh_batch_fo_surface: needs fo_vars in the batch's configuration and a vh_batch_fo_flag "
                           "call for the event");
        return VH_ERR_ARG;
    }
    *n_elements = 0;
    if (E.fo_n == 0) return VH_OK;
    if (!n_el) return VH_ERR_ARG;
    const int dim = c.nz > 1 ? 4 : 3;
    FoCell g{};
    for (int k = 0; k < dim; ++k) {
        if (!(widths[k] > 0)) return VH_ERR_ARG;
        g.d[k] = widths[k];
    }
    BCHK(hipSetDevice(b->dev));
    FoSlices S;
    for (int v = 0; v < FO_NVARS; ++v) {
        S.v[1][v] = E.fo_vars + (E.fo_vslot * FO_NVARS + v) * E.N;
        S.v[0][v] = E.fo_vars + ((E.fo_vslot ^ 1) * FO_NVARS + v) * E.N;
    }
    long long total = 0;
    int bad = 0;
    // fo_surface_run returns with the side stream drained; the flag call that
    // made its inputs drained `stream`
    BCHK(fo_surface_run(b->fo_surf, b->side, E.fo_prev, E.fo_cur, c.nx, c.ny, c.nz, E.fo_ids, E.fo_n, E.fo_eF, g, S,
                        n_el, elements, vars, cap, &total, &bad));
    if (bad) {
        vh::set_last_error("vh_batch_fo_surface: a cut cube whose surface cannot be closed");
        return VH_ERR_STATE;
    }
    *n_elements = total;
    return total > cap ? VH_ERR_CAPACITY : VH_OK;
}

vh_status vh_batch_fo_surfaces(vh_batch *b, int n, const int *ids, const double *widths, int *const *n_el,
                               double *const *elements, double *const *vars, const int64_t *cap,
                               int64_t *n_elements) {
    if (!b || n < 0 || (n > 0 && (!ids || !widths || !n_el || !elements || !vars || !cap || !n_elements)))
        return VH_ERR_ARG;
    if (!b->cfg.fo_vars) {
        vh::set_last_error("vh_batch_fo_surfaces: needs fo_vars in the batch's configuration");
        return VH_ERR_ARG;
    }
    std::vector<FoItem> items(n);
    std::vector<long long> total(n, 0);
    for (int k = 0; k < n; ++k) {
        const int id = ids[k];
        if (id < 0 || id >= (int)b->ev.size() || !b->ev[id].used || cap[k] < 0 ||
            (cap[k] > 0 && (!elements[k] || !vars[k])))
            return VH_ERR_ARG;
        for (int j = 0; j < k; ++j)
            if (ids[j] == id) return VH_ERR_ARG;  // one pass builds an event's surface once
        const Event &E = b->ev[id];
        const vh_event_params &c = E.prm;
        if (E.fo_n < 0) {
            vh::set_last_error("vh_batch_fo_surfaces: an event without a vh_batch_fo_flag call");
            return VH_ERR_ARG;
        }
        FoItem &I = items[k];
        const int dim = c.nz > 1 ? 4 : 3;
        for (int j = 0; j < dim; ++j) {
            if (!(widths[4 * k + j] > 0)) return VH_ERR_ARG;
            I.g.d[j] = widths[4 * k + j];
        }
        if (E.fo_n > 0 && !n_el[k]) return VH_ERR_ARG;
        I.eprev = E.fo_prev;
        I.ecur = E.fo_cur;
        I.ids = E.fo_ids;
        I.n = E.fo_n;
        I.nx = c.nx;
        I.ny = c.ny;
        I.nz = c.nz;
        I.eF = E.fo_eF;
        for (int v = 0; v < FO_NVARS; ++v) {
            I.S.v[1][v] = E.fo_vars + (E.fo_vslot * FO_NVARS + v) * E.N;
            I.S.v[0][v] = E.fo_vars + ((E.fo_vslot ^ 1) * FO_NVARS + v) * E.N;
        }
        I.n_el = n_el[k];
        I.elements = elements[k];
        I.vars = vars[k];
        I.cap = cap[k];
        I.total = &total[k];
    }
    BCHK(hipSetDevice(b->dev));
    int bad = 0;
    // returns with the side stream drained; the flag calls that made its
    // inputs drained `stream`
    BCHK(fo_surface_many(b->fo_surf, b->side, items.data(), n, &bad));
    if (bad) {
        vh::set_last_error("vh_batch_fo_surfaces: a cut cube whose surface cannot be closed");
        return VH_ERR_STATE;
    }
    bool room = true;
    for (int k = 0; k < n; ++k) {
        n_elements[k] = total[k];
        room &= total[k] <= cap[k];
    }
    return room ? VH_OK : VH_ERR_CAPACITY;
}

double vh_batch_eos_energy_density(const vh_batch *b, double T) { return b ? b->heos.energy_density(T) : NAN; }

vh_status vh_batch_autotune(vh_batch *b, int reps, int *tile2d, int *tile3d) {
    if (!b || reps < 1) return VH_ERR_ARG;
    if (tile2d) *tile2d = -1;
    if (tile3d) *tile3d = -1;
    if (!b->mixed) return VH_OK;
    BCHK(hipSetDevice(b->dev));
    BCHK(hipStreamSynchronize(b->stream));
    // Stage-1 launches write only the stage buffers, which the next step
    // overwrites, so timing them leaves the events' state untouched.  The
    // variants are timed in turn, reps rounds, and each keeps its median;
    // they are listed fastest-first for whole runs of the batching workload
    // (16x15 before 32x7 before 64x3), and a later one is taken only where it
    // beats the earlier choice by more than MARGIN, since on the state just
    // admitted the first two are within a few per cent of each other.
    constexpr float MARGIN = 0.02f;
    const int start[2] = {b->tile2, b->tile3};
    int best[2] = {b->tile2, b->tile3};
    for (int dim = 0; dim < 2; ++dim) {
        std::vector<float> t[MIXED_NTILE];
        bool any = true;
        for (int r = -1; r < reps && any; ++r)  // round -1 warms each variant up
            for (int v = 0; v < MIXED_NTILE; ++v) {
                if (dim == 0) b->tile2 = v;
                else b->tile3 = v;
                b->layout_dirty = true;
                vh_status s = rebuild_layout(b);
                if (s != VH_OK) return s;
                if ((dim == 0 ? b->nblk2 : b->nblk3) == 0) {
                    any = false;
                    break;
                }
                mx::MxFields F = mixed_fields(b, 0, b->iq_cur, b->iu_cur, b->iu_prev, 0, b->iq_stage, b->iu_stage, 1);
                const int *be = b->d_blk_ev + (dim ? b->nblk2 : 0), *bf = b->d_blk_first + (dim ? b->nblk2 : 0);
                const int nb = dim ? b->nblk3 : b->nblk2;
                // the 2+1D part is two launches (boost-invariant events first)
                const int nbi = dim ? 0 : b->nblk2bi;
                BCHK(hipEventRecord(b->ev0, b->stream));
                BCHK(launch_stage_mixed(F, b->deos.E, b->nc, dim == 1, v, true, b->d_ev, be, bf, nbi, b->stream));
                BCHK(launch_stage_mixed(F, b->deos.E, b->nc, dim == 1, v, false, b->d_ev, be + nbi, bf + nbi,
                                        nb - nbi, b->stream));
                BCHK(hipEventRecord(b->ev1, b->stream));
                BCHK(hipStreamSynchronize(b->stream));
                float ms = 0;
                BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                b->timed = false;
                if (r >= 0) t[v].push_back(ms);
            }
        if (!any) {
            best[dim] = start[dim];
        } else {
            float med[MIXED_NTILE];
            for (int v = 0; v < MIXED_NTILE; ++v) {
                std::sort(t[v].begin(), t[v].end());
                med[v] = t[v][t[v].size() / 2];
            }
            int pick = 0;
            for (int v = 1; v < MIXED_NTILE; ++v)
                if (med[v] < (1.0f - MARGIN) * med[pick]) pick = v;
            best[dim] = pick;
        }
        if (dim == 0) b->tile2 = best[0];
        else b->tile3 = best[1];
    }
    b->layout_dirty = true;
    vh_status s = rebuild_layout(b);
    if (tile2d) *tile2d = b->tile2;
    if (tile3d) *tile3d = b->tile3;
    return s;
}

vh_status vh_batch_set_tiles(vh_batch *b, int tile2d, int tile3d) {
    if (!b || tile2d < 0 || tile2d >= MIXED_NTILE || tile3d < 0 || tile3d >= MIXED_NTILE) return VH_ERR_ARG;
    if (!b->mixed) return VH_OK;
    b->tile2 = tile2d;
    b->tile3 = tile3d;
    b->layout_dirty = true;
    return VH_OK;
}

vh_status vh_batch_set_face_mode(vh_batch *b, int mode) {
    if (!b || mode < 0 || mode > 2) return VH_ERR_ARG;
    if (b->face_mode != mode) {
        BCHK(hipSetDevice(b->dev));
        BCHK(hipStreamSynchronize(b->stream));  // no graph is destroyed while it runs
        b->face_mode = mode;
        drop_graphs(b);  // the mode is baked into captured launches
    }
    return VH_OK;
}

vh_status vh_batch_nan_info(vh_batch *b, int *site, int *event, int64_t *cell) {
    if (!b) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    int err[4] = {0, 0, 0, 0};
    BCHK(hipMemcpyAsync(err, b->d_err, sizeof(int) * 3, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    const long s = (long)(uint32_t)err[1] | ((long)err[2] << 32);
    int ev = -1;
    long c = -1;
    if (err[0])
        for (int id = 0; id < (int)b->ev.size(); ++id)
            if (b->ev[id].used && s >= b->ev[id].off && s < b->ev[id].off + b->ev[id].N) {
                ev = id;
                c = s - b->ev[id].off;
            }
    if (site) *site = err[0];
    if (event) *event = ev;
    if (cell) *cell = c;
    return VH_OK;
}

vh_status vh_batch_clear_error(vh_batch *b) {
    if (!b) return VH_ERR_ARG;
    BCHK(hipSetDevice(b->dev));
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemset(b->d_err, 0, sizeof(int) * 4));
    return VH_OK;
}

}  // extern "C"
