// Device helpers shared by the LARE3D kernels: field names, the
// reference's min/max semantics and the limited-slope core of the remap.
#pragma once

#include "lare_dev.h"

#define RHO A.f[0]
#define EE A.f[1]
#define EI A.f[2]
#define VX A.f[3]
#define VY A.f[4]
#define VZ A.f[5]
#define BX A.f[6]
#define BY A.f[7]
#define BZ A.f[8]
#define VX1 A.f[9]
#define VY1 A.f[10]
#define VZ1 A.f[11]
#define DM A.f[12]

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
