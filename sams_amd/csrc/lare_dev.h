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

#include "lare_div.h"

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
    // metric factors of LARE3D::grid (setup.cpp:276-335), computed on the
    // host: hy, hyc at [ix + 2]; hz, hzc, hz1, hz2 at [(ix + 2) (ny + 5) + iy + 2];
    // the x positions xb and cot(yb) of the geometric forces (lagran.cpp:711-721)
    // at [i + 2] (1 and unused in Cartesian geometry)
    double *hy, *hyc, *hz, *hzc, *hz1, *hz2, *xbv, *cotv;
    Scal *s;
};

struct Dev {
    Geo g;
    Arr a;
};

// ------------------------------------------------------------------------
// Grid-group engine (the product kernel set): every launch covers a group of
// grids, a single grid being a group of one.  Arrays are typed by the
// storage type T of the grid (double, or float for FP32 storage); loads
// widen to double and stores round from double, so all arithmetic is FP64
// in both modes.  Flux arrays (the remap fluxes, dm and the resistive EMFs)
// are FP64 in both modes.

// Inclusive range relative to the grid size: x0..nx+dx1, y0..ny+dy1,
// z0..nz+dz1.
struct RSpec {
    int x0, dx1, y0, dy1, z0, dz1;
};

// Physics scalars of one grid (simulationData, shared_data.h:60-226).
struct Phys {
    double gas_gamma, visc1, visc2_norm, mu0, dt_multiplier;
    double eta_background, eta0, j_max;
    int resistive;
    // eta_background + eta0 equals eta_background bit for bit (eta0 = 0):
    // eta_calc's choice between them cannot change eta, so |j| is not needed
    int eta_fixed;
    // resistive with the kinetic-energy correction: the predictor reads the
    // prologue's p_i (lagran.cpp:121-125, 606), which differs from one
    // recomputed from e_i on ghost cells: the correction leaves them stale
    // and the resistive energy fill (lagran.cpp:502) refreshes them between
    int keep_pi;
};

// Device arrays are global memory: accessing them through address-space-1
// pointers lets the compiler emit global (not flat) loads and stores, which
// do not count against the scalar-load wait counter.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> gbl(T *p)
{
    return (gptr<T>)p;
}

// Element of a T-typed array: reads as double, assignment rounds to T.
template <class T>
struct Ref {
    gptr<T> q;
    __device__ __forceinline__ operator double() const { return (double)*q; }
    __device__ __forceinline__ const Ref &operator=(double v) const
    {
        *q = (T)v;
        return *this;
    }
    __device__ __forceinline__ const Ref &operator=(const Ref &o) const { return *this = (double)o; }
    __device__ __forceinline__ const Ref &operator+=(double v) const { return *this = (double)*q + v; }
    __device__ __forceinline__ const Ref &operator-=(double v) const { return *this = (double)*q - v; }
    __device__ __forceinline__ const Ref &operator*=(double v) const { return *this = (double)*q * v; }
    __device__ __forceinline__ const Ref &operator/=(double v) const { return *this = ddiv((double)*q, v); }
};

template <class T>
struct Fld {
    T *p;
    __device__ __forceinline__ Ref<T> operator[](int i) const { return Ref<T>{gbl(p) + i}; }
};

template <class T>
struct ArrE {
    // registered variables (dm is a flux: FP64)
    Fld<T> rho, ee, ei, vx, vy, vz, bx, by, bz, vx1, vy1, vz1;
    double *dm;
    // second copy of rho, e_e, e_i and v: the corrector and each remap sweep
    // read one copy and write the other (corrector A -> B, x B -> A,
    // y A -> B, z B -> A), so a kernel may read neighbours' old values
    // while it writes new ones
    Fld<T> rho2, ee2, ei2, vx2, vy2, vz2;
    // Lagrangian phase
    Fld<T> cs, cs_v, rho_v, alpha1, alpha2, alpha3, visc_heat, fxv, fyv, fzv;
    Fld<T> cv1, bx1, by1, bz1, p_e, p_i, eta, curlb;
    double *flux_x, *flux_y, *flux_z;
    // remap: cell volumes, vertex volume, Lagrangian width, the two
    // constrained-transport fluxes
    Fld<T> cv2, cvc1, db1;
    double *fl[2];
    // B's third copy: the fused y sweep reads B from bx1..bz1 (where the x
    // sweep left it) and writes it here, for the z sweep to read back
    Fld<T> bxy, byy, bzy;
    // the cell volume CV(ix, iy, iz) over the whole box, filled once when the
    // grid is created (the z sweep's cv2: its passes make the Lagrangian
    // volume after the sweep the cell volume itself)
    Fld<T> vol;
    // the remap kinetic-energy correction's delta_ke (FP64; only with rke)
    double *dke;
    // 1D axis deltas, index i stored at [i + 2]
    double *dxb, *dyb, *dzb, *dxc, *dyc, *dzc;
    Scal *s;
};

template <class T>
struct EDev {
    Geo g;
    ArrE<T> a;
    Phys p;
};

// One launch over a group: the members' device views, the first block of
// each member (ngrid entries, ascending), the total block count, the range
// and the workgroup order.
template <class T>
struct Src {
    const EDev<T> *devs;
    const int *boff;
    int ngrid;
    int nblk;
    RSpec r;
    int order;  // 0: chunk fastest, then plane; 1: chunks in 8 bands, band-major
    int st;     // 1: the state is in the second copy (rho2 ...); kernels see it as rho ...
};

// Built-in boundary conditions of one variable along one dimension, for one
// member of a group (see lare_kernels.h BCVar); f64 marks an FP64 array (dm)
// in an FP32-storage grid.
struct BCVarE {
    void *p;
    void *src;  // the other state copy, for a variable just written into p's copy
    int f64;
    int sd, sa, sb;
    int k0, k1;
    double v0, v1;
    int m0[3], m1[3];  // layer tables of map edges (kind 5)
};
struct BCSetE {
    BCVarE v[9];
};
// One variable's fills along all three dimensions in one pass (every ghost
// cell computed from the cells it finally derives from): staggering, kinds
// and clamp values per dimension, and the variable's workgroups (b0: first
// within its grid; nsh: ghost cells).
struct BCVar3 {
    void *p;
    int f64;
    int s[3];
    int k0[3], k1[3];
    double v0[3], v1[3];
    int m0[3][3], m1[3][3];  // layer tables of map edges (kind 5), per dimension
    int b0, nsh;
};
struct BCSet3 {
    BCVar3 v[9];
    int nv;
};

}  // namespace lare

// Linear index of (i,j,k) in the shared box.
#define IX(i, j, k) (G.base + (i) * G.sx + (j) * G.sy + (k))
// 1D deltas
#define DXB(i) (lare::gbl(A.dxb)[(i) + 2])
#define DYB(i) (lare::gbl(A.dyb)[(i) + 2])
#define DZB(i) (lare::gbl(A.dzb)[(i) + 2])
#define DXC(i) (lare::gbl(A.dxc)[(i) + 2])
#define DYC(i) (lare::gbl(A.dyc)[(i) + 2])
#define DZC(i) (lare::gbl(A.dzc)[(i) + 2])
// metric factors (kernel set 0)
#define HY(i) (A.hy[(i) + 2])
#define HYC(i) (A.hyc[(i) + 2])
#define HZ(i, j) (A.hz[((i) + 2) * (G.ny + 5) + (j) + 2])
#define HZC(i, j) (A.hzc[((i) + 2) * (G.ny + 5) + (j) + 2])
#define HZ1(i, j) (A.hz1[((i) + 2) * (G.ny + 5) + (j) + 2])
#define HZ2(i, j) (A.hz2[((i) + 2) * (G.ny + 5) + (j) + 2])
