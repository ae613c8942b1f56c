// Launch interface between the C-ABI layer and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "vh_stage.hpp"

namespace vh {

hipError_t launch_stage(const StageParams &P, const EosDev &E, int nc, bool dim3d, hipStream_t st);
hipError_t launch_init_conserved(double *const q[15], const double *Pi, const double *e, const double *p,
                                 const double *const u[4], long N, hipStream_t st);
hipError_t launch_count_above(const double *e, long N, double eF, unsigned long long *out, hipStream_t st);
hipError_t launch_fo_flag(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                          long long *ids, long long cap, unsigned long long *count, hipStream_t st);
hipError_t launch_fo_corners(const double *eprev, const double *ecur, int nx, int ny, int nz, const long long *ids,
                             long long n, double *corners, hipStream_t st);

}  // namespace vh
