// Numeric helpers shared by every LARE3D kernel: the reference's min/max
// semantics, the limited-slope core of the remap and the Cartesian metric
// from the 1D deltas.
#pragma once

#include "lare_dev.h"

// Every kernel is launched with 256 threads per block; the bound lets the
// compiler use up to 256 VGPRs per lane before spilling.
#define KERNEL __global__ __launch_bounds__(256)

// Reference min/max semantics (include/pp/utils.h:277-335).
__device__ __forceinline__ double pmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double pmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double pmin_list3(double a, double b, double c)
{
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}
__device__ __forceinline__ double pmin_list4(double a, double b, double c, double d)
{
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    if (d < m) m = d;
    return m;
}

static constexpr double kSixth = 1.0 / 6.0;  // constants.h:46

// The limited-slope core shared by the fifteen remap flux kernels
// (xremap.cpp:304-316 and twins).
__device__ __forceinline__ double vl_di(double sign_v, double dfi, double dfu, double phi, double dxci, double dxcu, double dxbu)
{
    double Da = (2.0 - phi) * fabs(dfi) / dxci + (1.0 + phi) * fabs(dfu) / dxcu;
    Da = Da * kSixth;
    double ss = 0.5 * ((dfi >= 0.0 ? 1.0 : -1.0) + (dfu >= 0.0 ? 1.0 : -1.0));
    return sign_v * ss * pmin_list3(fabs(Da) * dxbu, fabs(dfi), fabs(dfu));
}


// Metric of LARE3D::grid (setup.cpp:337-390) for Cartesian geometry, from
// the 1D deltas: the same products, in the same order, as the stored arrays.
#define DXAB(iy, iz) (DYB(iy) * DZB(iz))
#define DYAB(ix, iz) (DXB(ix) * DZB(iz))
#define DZAB(ix, iy) (DXB(ix) * DYB(iy))
#define CV(ix, iy, iz) (DXB(ix) * (DYB(iy) * DZB(iz)))
#define DXAC(iy, iz) (DYB(iy) * DZC(iz))
#define DYAC(ix, iz) (DXB(ix) * DZC(iz))
