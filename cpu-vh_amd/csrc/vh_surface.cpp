// Freezeout surface elements of one lattice cell (C ABI in include/cpuvh_amd.h)
// on the host: the construction is vh_surface.hpp's, which the device's
// element kernels share.
#include "vh_surface.hpp"

#include "../../include/cpuvh_amd.h"

namespace {
vh_status to_status(int r) {
    return r == vh::surf::SURF_OK ? VH_OK : (r == vh::surf::SURF_STATE ? VH_ERR_STATE : VH_ERR_CAPACITY);
}
}  // namespace

extern "C" vh_status vh_surface_square(const double *corners, double E0, const double *d, double *elements,
                                       int cap, int *n_elements) {
    if (!corners || !d || !n_elements) return VH_ERR_ARG;
    return to_status(vh::surf::square_elements(corners, E0, d, elements, cap, n_elements));
}

extern "C" vh_status vh_surface_cube(const double *corners, double E0, const double *d, double *elements, int cap,
                                     int *n_elements) {
    if (!corners || !d || !n_elements) return VH_ERR_ARG;
    // without an output buffer a cut cube is a capacity error, as before
    const int r = vh::surf::cube_elements(corners, E0, d, elements, elements ? cap : -1, n_elements);
    return to_status(r);
}

extern "C" vh_status vh_surface_hypercube(const double *corners, double E0, const double *d, double *elements,
                                          int cap, int *n_elements) {
    if (!corners || !d || !n_elements) return VH_ERR_ARG;
    const int r = vh::surf::hypercube_elements(corners, E0, d, elements, elements ? cap : -1, n_elements);
    return to_status(r);
}
