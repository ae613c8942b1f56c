// Multilinear interpolation of the freezeout variables inside a cell, shared
// by the interpolation kernel and the surface-element kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace vh {

// Multilinear interpolation inside a cell (FreezeOut.cpp:13-56), the terms
// summed in the reference's order; a[t][x][y][z] corner values.
__device__ inline double fo_lin3(double x0, double x1, double x2, const double (&a)[2][2][2]) {
    return ((1 - x0) * (1 - x1) * (1 - x2) * a[0][0][0]) + ((x0) * (1 - x1) * (1 - x2) * a[1][0][0]) +
           ((1 - x0) * (x1) * (1 - x2) * a[0][1][0]) + ((1 - x0) * (1 - x1) * (x2) * a[0][0][1]) +
           ((x0) * (x1) * (1 - x2) * a[1][1][0]) + ((x0) * (1 - x1) * (x2) * a[1][0][1]) +
           ((1 - x0) * (x1) * (x2) * a[0][1][1]) + ((x0) * (x1) * (x2) * a[1][1][1]);
}
__device__ inline double fo_lin4(double x0, double x1, double x2, double x3, const double (&a)[2][2][2][2]) {
    return ((1 - x0) * (1 - x1) * (1 - x2) * (1 - x3) * a[0][0][0][0]) +
           ((x0) * (1 - x1) * (1 - x2) * (1 - x3) * a[1][0][0][0]) +
           ((1 - x0) * (x1) * (1 - x2) * (1 - x3) * a[0][1][0][0]) +
           ((1 - x0) * (1 - x1) * (x2) * (1 - x3) * a[0][0][1][0]) +
           ((1 - x0) * (1 - x1) * (1 - x2) * (x3) * a[0][0][0][1]) + ((x0) * (x1) * (1 - x2) * (1 - x3) * a[1][1][0][0]) +
           ((x0) * (1 - x1) * (x2) * (1 - x3) * a[1][0][1][0]) + ((x0) * (1 - x1) * (1 - x2) * (x3) * a[1][0][0][1]) +
           ((1 - x0) * (x1) * (x2) * (1 - x3) * a[0][1][1][0]) + ((1 - x0) * (x1) * (1 - x2) * (x3) * a[0][1][0][1]) +
           ((1 - x0) * (1 - x1) * (x2) * (x3) * a[0][0][1][1]) + ((x0) * (x1) * (x2) * (1 - x3) * a[1][1][1][0]) +
           ((x0) * (x1) * (1 - x2) * (x3) * a[1][1][0][1]) + ((x0) * (1 - x1) * (x2) * (x3) * a[1][0][1][1]) +
           ((1 - x0) * (x1) * (x2) * (x3) * a[0][1][1][1]) + ((x0) * (x1) * (x2) * (x3) * a[1][1][1][1]);
}

}  // namespace vh
