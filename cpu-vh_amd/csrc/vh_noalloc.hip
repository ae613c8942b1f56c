// vh_fo_flag_noalloc: the freezeout crossing test and ordered compaction of
// vh_fo_flag on caller-owned device slices, with the kernels' per-block
// counts in a fixed device array of the library instead of memory allocated
// per call, so that callers which cannot allocate (code under HIP graph
// capture, tight serving loops) can flag cubes.  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cpuvh_amd.h"
#include "vh_kernels.hpp"

namespace {

constexpr long MAX_BLOCKS = VH_FO_NOALLOC_MAX_CUBES / 256;
constexpr int MAX_DEVICES = 64;

// per-block counts of the flag kernels (fo_flag_blocks(M) entries); each GPU
// holds its own copy until the process ends
static __device__ __align__(256) unsigned long long scratch[MAX_BLOCKS];

std::mutex g_mu;
unsigned long long *g_scratch[MAX_DEVICES] = {};  // device address of scratch, per device

}  // namespace

extern "C" vh_status vh_fo_flag_noalloc(const double *eprev, const double *ecur, int nx, int ny, int nz, double eF,
                                        int64_t *ids, int64_t cap, int64_t *n_flagged, void *stream) {
    if (!eprev || !ecur || !n_flagged || (cap > 0 && !ids) || cap < 0 || nx < 2 || ny < 2 || nz < 1)
        return VH_ERR_ARG;
    const long M = (long)(nx - 1) * (ny - 1) * (nz > 1 ? nz - 1 : 1);
    if (M > VH_FO_NOALLOC_MAX_CUBES) {
        vh::set_last_error("vh_fo_flag_noalloc: " + std::to_string(M) + " cubes exceed the limit of " +
                           std::to_string((long)VH_FO_NOALLOC_MAX_CUBES));
        return VH_ERR_ARG;
    }
    int dev = 0;
    hipError_t r = hipGetDevice(&dev);
    if (r != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        vh::set_last_error(std::string("vh_fo_flag_noalloc: device: ") + hipGetErrorString(r));
        return VH_ERR_HIP;
    }
    unsigned long long *blk = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {  // once per device, with that device current
            void *p = nullptr;
            r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (r != hipSuccess || !p) {
                vh::set_last_error(std::string("vh_fo_flag_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(r));
                return VH_ERR_HIP;
            }
            g_scratch[dev] = static_cast<unsigned long long *>(p);
        }
        blk = g_scratch[dev];
    }
    // every block writes its own count and the scan writes the total, so
    // nothing is cleared first
    r = vh::launch_fo_flag(eprev, ecur, nx, ny, nz, eF, reinterpret_cast<long long *>(ids), (long long)cap,
                           reinterpret_cast<unsigned long long *>(n_flagged), blk,
                           static_cast<hipStream_t>(stream));
    if (r != hipSuccess) {
        vh::set_last_error(std::string("vh_fo_flag_noalloc: ") + hipGetErrorString(r));
        return VH_ERR_HIP;
    }
    return VH_OK;
}
