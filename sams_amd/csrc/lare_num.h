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

// Quotients by a shared divisor, equal bit for bit to IEEE a / d (and to
// ddiv).  QDiv keeps d and y = 1/d rounded to nearest (computed once), so a
// quotient costs m = a * y and q = fma(-fma(d, m, -a), y, m): ddiv's last
// steps without its scaling, v_div_fmas and v_div_fixup (fma(d, m, -a) is
// exactly -r).  The scaling steps return their operand unchanged and
// v_div_fmas is a plain fma unless an operand or the quotient is near the
// exponent limits, and v_div_fixup then only sets the sign; so for |a| and d
// in [2^-300, 2^301) the two agree.  For d > 0 a zero a gives the signed zero
// a / d gives (-fma(d, m, -a) is -0 there, so the sum keeps m's sign).  Any
// other a, and every a when d is negative or out of that range, marks the
// quotient slow: the caller recomputes it with ddiv.
struct QDiv {
    double d, y;
    bool ok;
};
// |v| in [2^-300, 2^301): the biased exponent field in 723..1323.  Compared
// as doubles (a NaN compares false, so it is out of the range like an
// infinity; zeros and subnormals are below it).
__device__ __forceinline__ bool qdiv_in_range(double v)
{
    const double a = __builtin_fabs(v);
    return a >= 0x1p-300 && a < 0x1p301;
}
__device__ __forceinline__ QDiv qdiv_of(double d)
{
    const bool ok = d > 0.0 && qdiv_in_range(d);
    return QDiv{d, rcp_nearest(d), ok};
}
// The quotient, with `slow` set where it may differ from a / d.
__device__ __forceinline__ double qdiv_f(double a, const QDiv &R, bool &slow)
{
    const double m = a * R.y;
    const double q = fma(-fma(R.d, m, -a), R.y, m);
    slow = slow || !R.ok || !(qdiv_in_range(a) || a == 0.0);
    return q;
}
__device__ __forceinline__ double qdiv(double a, const QDiv &R)
{
    bool slow = false;
    double q = qdiv_f(a, R, slow);
    if (__builtin_expect(slow, 0)) q = ddiv(a, R.d);
    return q;
}

// The limited-slope core shared by the fifteen remap flux kernels
// (xremap.cpp:304-316 and twins).
__device__ __forceinline__ double vl_di(double sign_v, double dfi, double dfu, double phi, double dxci, double dxcu, double dxbu)
{
    double Da = ddiv((2.0 - phi) * fabs(dfi), dxci) + ddiv((1.0 + phi) * fabs(dfu), dxcu);
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
