// Device-side layout shared by the LARE3D kernels.
//
// Every 3D array of a grid lives in one padded box covering indices
// -2..n+2 in each dimension (the widest reference extent, the fully
// staggered one of setup.cpp:51-61), z fastest.  The z pitch is rounded up
// to a multiple of 8 doubles so that rows start on 64-byte boundaries.
// Because all arrays share one box, one linear index addresses the same
// (ix,iy,iz) in every array; arrays whose reference extent is smaller simply
// never touch the extra cells.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lare {

struct Geo {
    int nx, ny, nz;
    int sy;    // pitch of y (doubles)
    int sx;    // pitch of x (doubles)
    int base;  // linear index of (0,0,0)
};

// Inclusive index range in reference indexing.
struct R3 {
    int x0, x1, y0, y1, z0, z1;
};

// Device scalars of one grid: everything a step reads or produces that
// must not round-trip through the host.
struct Scal {
    double dt;          // current dt (previous step's dt until set_dt runs)
    double dtr;         // dt recorded by set_dt (lagran.cpp:442)
    double dt_sub;      // dt of a resistive substep
    unsigned long long dtmin_bits;  // running CFL minimum (non-negative doubles order as integers)
    int64_t nstep;      // completed steps
};

// Fields and work arrays of one grid (device pointers, shared layout).
struct Arr {
    // registered variables, in enum lare_field order
    double *f[13];
    // Lagrangian-phase arrays (setup.cpp:170-212) and temporaries
    double *p_visc, *eta, *cv, *cv1, *bx1, *by1, *bz1;
    double *alpha1, *alpha2, *alpha3, *visc_heat, *pressure, *p_e, *p_i;
    double *rho_v, *cv_v, *fx, *fy, *fz, *fxv, *fyv, *fzv;
    double *flux_x, *flux_y, *flux_z, *curlb, *cs, *cs_v;
    // metric (setup.cpp:337-390)
    double *dxab, *dyab, *dzab, *dxac, *dyac, *dzac;
    // remap work arrays (remap.cpp:30-35 and the per-sweep flux)
    double *rho1, *cv2, *cvc1, *db1, *rrho_v, *rho_v1, *rflux, *cv2v, *wsh, *dmv;
    // fluxes of the fused sweep (two CT or energy fluxes, three momentum)
    double *fl[5];
    // 1D axis deltas, index i stored at [i + 2]
    double *dxb, *dyb, *dzb, *dxc, *dyc, *dzc;
    Scal *s;
};

struct Dev {
    Geo g;
    Arr a;
};

}  // namespace lare

// Linear index of (i,j,k) in the shared box.
#define IX(i, j, k) (G.base + (i) * G.sx + (j) * G.sy + (k))
// 1D deltas
#define DXB(i) (A.dxb[(i) + 2])
#define DYB(i) (A.dyb[(i) + 2])
#define DZB(i) (A.dzb[(i) + 2])
#define DXC(i) (A.dxc[(i) + 2])
#define DYC(i) (A.dyc[(i) + 2])
#define DZC(i) (A.dzc[(i) + 2])
