// Launch interface between the C-ABI layer and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include "vh_stage.hpp"
#include "vh_host_eos.hpp"

namespace vh {

// Message vh_status_string returns for VH_ERR_HIP (per host thread).
void set_last_error(const std::string &msg);

// Device copy of an equation of state: the table columns (table EoS) or the
// powf(x, 0.25f) exception list (conformal EoS), and the EosDev view of them.
struct DeviceEos {
    EosDev E{};
    double *ptb = nullptr, *Ttb = nullptr;
    float4 *p4 = nullptr;
    uint2 *exc = nullptr;
    uint32_t *bucket = nullptr;
    hipError_t init(const HostEos &h, int kind, int cs2_mode = 0);  // on the current device
    void release();
};

hipError_t launch_stage(const StageParams &P, const EosDev &E, int nc, bool dim3d, hipStream_t st);
hipError_t launch_stage_batch(const StageParams &P, const EosDev &E, int nc, bool dim3d, const EvDesc *ev,
                              const int *blk_ev, const int *blk_first, int nblocks, hipStream_t st);
hipError_t launch_count_above_batch(const double *e, double eF, const EvDesc *ev, const int *ev_of,
                                    const long *start, int nchunks, unsigned long long *counts, hipStream_t st);
hipError_t launch_advance_clocks(EvDesc *ev, int n, int count, hipStream_t st);
hipError_t launch_advance_base(int *base, int count, hipStream_t st);

// Steps per captured HIP graph.  Longer runs replay several graphs back to
// back (each one's last node moves the step clock on), so no single launch
// puts more than about a hundred kernel dispatches in flight.
constexpr int MAX_GRAPH_STEPS = 40;
hipError_t launch_zero_counts(unsigned long long *counts, int n, hipStream_t st);
// getInferredVariables on n states (SoA: q[k*n + i], k < nc; out[j*n + i],
// j = e, p, u^tau, u^x, u^y, u^eta; nan[i] the NaN site, 0 if none)
hipError_t launch_inferred(const EosDev &E, int nc, double t, const double *q, const double *ePrev, long n,
                           double *out, int *nan, hipStream_t st);
hipError_t launch_strikes(EvDesc *ev, const unsigned long long *counts, int *strikes, int *stop_check, int n,
                          int check, hipStream_t st);
hipError_t launch_init_conserved(double *const q[15], const double *Pi, const double *e, const double *p,
                                 const double *const u[4], long N, hipStream_t st);
// One event's initial state as uploaded to the device; a NULL array is absent
// (u^tau one, every other field zero).
struct AdmitSrc {
    const double *e, *p, *u[4], *pi10, *Pi;
};
// A batched FP64 event's admission: e, p, u and the previous u, pi and Pi
// (q[(4 + n) * cap + off ..]) and T^{tau mu} (q[n * cap + off ..]); strikes
// and stop are the event's stop-rule slots (0 and -1).
hipError_t launch_admit_fp64(const AdmitSrc &S, double *q, double *e0, double *p0, double *uc, double *up, long cap,
                             long N, long off, int nc, int *strikes, int *stop, hipStream_t st);
hipError_t launch_count_above(const double *e, long N, double eF, unsigned long long *out, hipStream_t st);
// The cut cubes' ids in increasing order into ids (capacity cap), their
// number into *count; blk is scratch of fo_flag_blocks(cubes) entries.
inline long fo_flag_blocks(long cubes) { return (cubes + 255) / 256; }
hipError_t launch_fo_flag(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                          long long *ids, long long cap, unsigned long long *count, unsigned long long *blk,
                          hipStream_t st);
hipError_t launch_fo_corners(const double *eprev, const double *ecur, int nx, int ny, int nz, const long long *ids,
                             long long n, double *corners, hipStream_t st);
// The ten variables the freezeout surface interpolates (u^x u^y u^eta e pi^xx
// pi^xy pi^xeta pi^yy pi^yeta Pi, FreezeOut.cpp:57-97): previous (v[0]) and
// current (v[1]) coarse slices, nx*ny*nz doubles each.
constexpr int FO_NVARS = 10;
struct FoSlices {
    const double *v[2][FO_NVARS];
};
hipError_t launch_fo_interp(const FoSlices &S, int nx, int ny, int nz, const long long *cubes, const double *frac,
                            long long n, double *vars, hipStream_t st);
// Surface elements of the flagged cubes (vh_fo_surface.hip): d = the cell's
// widths (coarse dtau, dx, dy, deta); fo_elem_scan: inclusive scan of the
// element counts into offsets[1 .. n] (hipcub; temp NULL queries temp_bytes).
struct FoCell {
    double d[4];
};
hipError_t fo_elem_scan(void *temp, size_t &temp_bytes, const int *counts, int *offsets_plus1, long long n,
                        hipStream_t st);
// The whole surface of one coarse step for the n flagged cubes ids (device):
// corners, the elements and variables into per-cube slots, the counts' scan,
// then, when there is room (total <= cap) and no cube failed (*bad), the
// elements and variables compacted and copied to the host arrays n_el (n
// counts), elements and vars.  Synchronising.  W's buffers grow on demand.
struct FoSurfaceWork {
    void *buf = nullptr;
    size_t buf_bytes = 0;
    double *el = nullptr;
    size_t el_bytes = 0;
    // pinned host staging of the downloads (a copy into pageable memory is
    // staged by the runtime in synchronous chunks)
    char *host = nullptr;
    size_t host_bytes = 0;
    // pinned staging of the pass's item table (uploaded; reused only after
    // the pass has synchronised)
    char *up = nullptr;
    size_t up_bytes = 0;
};
// Device-to-host copies of n (src, bytes) pieces through W's pinned staging:
// one synchronisation of st, then each piece copied to its destination.
hipError_t fo_download(FoSurfaceWork &W, hipStream_t st, int n, void *const *dst, const void *const *src,
                       const size_t *bytes);
hipError_t fo_surface_run(FoSurfaceWork &W, hipStream_t st, const double *eprev, const double *ecur, int nx, int ny,
                          int nz, const long long *ids, long long n, double eF, const FoCell &g, const FoSlices &S,
                          int *n_el, double *elements, double *vars, long long cap, long long *total, int *bad);
// One event's part of a surface pass over several events (fo_surface_many):
// its coarse slices of e, its n flagged cubes ids (device), lattice, level,
// cell widths and variable slices; the host arrays its results go to
// (n_el: n counts; elements and vars: room for cap elements) and its element
// count (*total).  cube0 is set by the pass (the event's first cube in it).
struct FoItem {
    const double *eprev = nullptr, *ecur = nullptr;
    const long long *ids = nullptr;
    long long n = 0;
    int nx = 0, ny = 0, nz = 1;
    double eF = 0;
    FoCell g{};
    FoSlices S{};
    long long cube0 = 0;
    int *n_el = nullptr;
    double *elements = nullptr, *vars = nullptr;
    long long cap = 0;
    long long *total = nullptr;
};
// The surfaces of several events in one pass: one element launch over all
// their cubes, one scan, two synchronisations.  Every item's *total and n_el
// are set; the elements and variables of each item with room (total <= cap)
// are copied, unless a cube failed (*bad: nothing copied).  Items with n = 0
// are skipped (total 0).
hipError_t fo_surface_many(FoSurfaceWork &W, hipStream_t st, FoItem *items, int n_items, int *bad);
void fo_surface_free(FoSurfaceWork &W);

// Thermal vorticity (calculateThermalVorticity): u and up the step's and the
// previous step's velocity, e the step's energy density; out: 10 arrays of
// nx*ny*nz, wtx wty wtn wxy wxn wyn then beta_t beta_x beta_y beta_n.
struct VortArgs {
    const double *u[4], *up[4], *e;
    double *out;
    int nx, ny, nz;
    double dt, dx, dy, dz;
};
hipError_t launch_vorticity(const VortArgs &A, const EosDev &E, hipStream_t st);

// Mixed-precision form (FP32 storage; FP64 root solve, EoS and accumulation).
namespace mx {
struct MxFields;
}
constexpr int MIXED_NTILE = 3;
extern const int MIXED_TILES_2D[MIXED_NTILE][3];  // (TX, TY, TZ) of each 2+1D tile variant
extern const int MIXED_TILES_3D[MIXED_NTILE][3];
// bi: the launch holds only boost-invariant 2+1D events (nz = 1, u^eta and
// the pi^{. eta} components zero), whose eta components stay zero.
hipError_t launch_stage_mixed(const mx::MxFields &F, const EosDev &E, int nc, bool dim3d, int tile, bool bi,
                              const EvDesc *ev, const int *blk_ev, const int *blk_first, int nblocks, hipStream_t st);
// A mixed batch's admission: F.qo/F.qol/F.uo/F.eo/F.po of the event's cells
// from the FP64 upload, the previous u (up, field stride F.cap) equal to u,
// and the stop-rule slots as launch_admit_fp64 opens them.
hipError_t launch_init_mixed(const mx::MxFields &F, const AdmitSrc &S, float *up, long N, long off, int nc,
                             int *strikes, int *stop, hipStream_t st);
hipError_t launch_count_above_batch_f(const float *e, double eF, const EvDesc *ev, const int *ev_of,
                                      const long *start, int nblocks, unsigned long long *counts, hipStream_t st);
hipError_t launch_widen(const float *hi, const float *lo, double *out, long n, hipStream_t st);
// count arrays widened (or, with hi NULL, zeroed) in one launch
struct WidenSet {
    static constexpr int MAX = 10;
    const float *hi[MAX], *lo[MAX];
    double *out[MAX];
    int count;
};
hipError_t launch_widen_set(const WidenSet &S, long n, hipStream_t st);
// out[0 .. n) = v (a fill kernel: usable where a memset node is not)
hipError_t launch_fill(double *out, long n, double v, hipStream_t st);

}  // namespace vh
