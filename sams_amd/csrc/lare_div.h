// Correctly rounded FP64 division for the LARE3D kernels.
//
// hipcc lowers a / d (FP64, gfx950) to: v_div_scale of d and of a,
// y = v_rcp_f64(d') refined by two Newton steps (e = fma(-d', y, 1),
// y = fma(y, e, y)), m = a' * y, r = fma(-d', m, a'), q = v_div_fmas(r, y,
// m), then v_div_fixup.  After two steps y can still be one unit in the last
// place away from 1/d' rounded to nearest: 1 - 5 * 2^-53 gets y = 1 + 2^-51
// where the nearest is 1 + 3 * 2^-52, and the divisors 1 - k * 2^-53 with
// small odd k (whose reciprocals lie next to a midpoint) fail the same way
// in every binade.  The last step then rounds quotients that lie close to a
// midpoint the wrong way: 1.3999999999999992 / (1 - 5 * 2^-53) comes out as
// 1.4 where IEEE division (x86 divsd) gives 1.4000000000000001.
//
// ddiv keeps the scaling, v_div_fmas and v_div_fixup of that sequence and
// adds a third Newton step.  From a y within one unit of 1/d', that step
// yields 1/d' rounded to nearest for every d' whose significand is not all
// ones; for an all-ones significand 2^e (2 - 2^-52) the step ties to
// 2^(-e-1), while the nearest is 2^(-e-1) (1 + 2^-52), one more in the last
// bit, which rcp_nearest sets.  With y = RN(1/d'), m = RN(a' y) is within one
// unit of a'/d', r = a' - d' m is exact and RN(m + r y) is the correctly
// rounded quotient.  Both facts are Markstein's (P. Markstein, IBM J. Res.
// Dev. 34(1), 1990; restated in Muller et al., Handbook of Floating-Point
// Arithmetic); the parity tests check ddiv against x86 division on the
// divisors above and on quotients next to midpoints.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

// 1/d rounded to nearest, for d normal and away from the exponent limits
// (|d| in [2^-1020, 2^1020]).
__device__ __forceinline__ double rcp_nearest(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    // all-ones significand: |mantissa| (frexp, in [1/2, 1)) is 1 - 2^-53.
    // There y is 2^(-e-1) (the tie) or already 2^(-e-1) (1 + 2^-52) (when
    // the second step had reached it): both have the high word of 2^(-e-1)
    // and a low word of 0 or 1, and the nearest has the low word 1.
    const bool ones = __builtin_fabs(__builtin_amdgcn_frexp_mant(d)) == 0x1.fffffffffffffp-1;
    return __hiloint2double(__double2hiint(y), ones ? 1 : __double2loint(y));
}

// a / d, correctly rounded (IEEE 754 round to nearest even), for all inputs
// whose quotient is a normal number, zero, infinite or NaN.
__device__ __forceinline__ double ddiv(double a, double d)
{
    bool dflag, nflag;
    const double ds = __builtin_amdgcn_div_scale(a, d, false, &dflag);
    const double ns = __builtin_amdgcn_div_scale(a, d, true, &nflag);
    const double y = rcp_nearest(ds);
    const double m = ns * y;
    const double r = fma(-ds, m, ns);
    return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(r, y, m, nflag), d, a);
}

// Integer and single-precision operands are not meant for ddiv.
template <class A, class B,
          std::enable_if_t<std::is_integral<A>::value || std::is_integral<B>::value ||
                               std::is_same<A, float>::value || std::is_same<B, float>::value,
                           int> = 0>
double ddiv(A, B) = delete;
