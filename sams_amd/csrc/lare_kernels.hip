// gfx950 kernels of the LARE3D Lagrangian-remap step: one kernel per
// reference kernel, over the reference's own index ranges, on the shared
// padded layout of lare_dev.h.  Compiled with -ffp-contract=off: every
// expression keeps the reference's association so that results are
// bit-identical to the reference's CPU path.
//
// Threads map z to threadIdx.x (one 64-lane wavefront covers 64 consecutive
// doubles of a row, so loads and stores are coalesced), y to threadIdx.y and
// x to blockIdx.z.
#include "lare_kernels.h"

#include <float.h>

using lare::Arr;
using lare::Dev;
using lare::Geo;
using lare::R3;

#define KBEGIN                                                           \
    const Geo G = D.g;                                                   \
    const Arr A = D.a;                                                   \
    const int iz = R.z0 + (int)(blockIdx.x * blockDim.x + threadIdx.x); \
    const int iy = R.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y); \
    const int ix = R.x0 + (int)blockIdx.z;                               \
    if (iz > R.z1 || iy > R.y1) return;                                  \
    (void)ix;

#include "lare_math.h"

// ------------------------------------------------------------- utilities

KERNEL void k_fill(double *p, long n, double v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// A field over its reference extent x0..x0+nx-1 (likewise y, z), contiguous
// with z fastest, to / from its place in the padded box; E is the box's
// element type (FP32 storage rounds on the way in and widens on the way out).
template <class E>
KERNEL void k_scatter(E *box, const double *src, Geo G, int x0, int y0, int z0, int nx, int ny, int nz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nx * ny * nz) return;
    const int iz = (int)(i % nz), iy = (int)((i / nz) % ny), ix = (int)(i / ((long)ny * nz));
    box[IX(x0 + ix, y0 + iy, z0 + iz)] = (E)src[i];
}
template <class E>
KERNEL void k_gather(double *dst, const E *box, Geo G, int x0, int y0, int z0, int nx, int ny, int nz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nx * ny * nz) return;
    const int iz = (int)(i % nz), iy = (int)((i / nz) % ny), ix = (int)(i / ((long)ny * nz));
    dst[i] = (double)box[IX(x0 + ix, y0 + iy, z0 + iz)];
}

// ------------------------------------------------- Lagrangian: prologue

// lagran.cpp:113-125 over the centred extent; also resets the CFL minimum.
KERNEL void k_prologue1(Dev D, R3 R, double gas_gamma)
{
    KBEGIN
    if (ix == R.x0 && iy == R.y0 && iz == R.z0) A.s->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;  // DBL_MAX
    const int c = IX(ix, iy, iz);
    A.bx1[c] = 0.5 * (BX[c] + BX[c - G.sx]);
    A.by1[c] = 0.5 * (BY[c] + BY[c - G.sy]);
    A.bz1[c] = 0.5 * (BZ[c] + BZ[c - 1]);
    A.p_e[c] = (gas_gamma - 1.0) * RHO[c] * EE[c];
    A.p_i[c] = (gas_gamma - 1.0) * RHO[c] * EI[c];
    A.pressure[c] = A.p_e[c] + A.p_i[c];
}

#define SUM8(X, Y, c)                                                                                  \
    ((X)[c] * (Y)[c] + (X)[c + G.sx] * (Y)[c + G.sx] + (X)[c + G.sy] * (Y)[c + G.sy] +                 \
     (X)[c + G.sx + G.sy] * (Y)[c + G.sx + G.sy] + (X)[c + 1] * (Y)[c + 1] +                           \
     (X)[c + G.sx + 1] * (Y)[c + G.sx + 1] + (X)[c + G.sy + 1] * (Y)[c + G.sy + 1] +                   \
     (X)[c + G.sx + G.sy + 1] * (Y)[c + G.sx + G.sy + 1])
#define PLAIN8(X, c)                                                                                   \
    ((X)[c] + (X)[c + G.sx] + (X)[c + G.sy] + (X)[c + G.sx + G.sy] + (X)[c + 1] + (X)[c + G.sx + 1] + \
     (X)[c + G.sy + 1] + (X)[c + G.sx + G.sy + 1])

// lagran.cpp:128-153 over (-1..n+1)^3.
KERNEL void k_prologue2(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double sum_rho_cv = SUM8(RHO, A.cv, c);
    double sum_cv = PLAIN8(A.cv, c);
    A.rho_v[c] = ddiv(sum_rho_cv, sum_cv);
    A.cv_v[c] = 0.125 * sum_cv;
}

// ------------------------------------------------- shock viscosity

// lagran.cpp:190-197 over the centred extent.
KERNEL void k_cs(Dev D, R3 R, double gas_gamma)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double rmin = pmax(RHO[c], DBL_EPSILON);
    double b2 = A.bx1[c] * A.bx1[c] + A.by1[c] * A.by1[c] + A.bz1[c] * A.bz1[c];
    double p = A.pressure[c];
    A.cs[c] = sqrt(ddiv(gas_gamma * p + b2, rmin));
}

// lagran.cpp:199-216 over (-1..n+1)^3.
KERNEL void k_cs_v(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double sum = SUM8(A.cs, A.cv, c);
    A.cs_v[c] = ddiv(0.125 * sum, A.cv_v[c]);
}

// edge_viscosity (lagran.cpp:56-93) for the edge between linear indices
// c1 and c2 = c1 + s (s the stride along the edge), with c0 = c1 - s and
// c3 = c2 + s.
__device__ __forceinline__ double edge_viscosity(const Arr &A, double dt, double visc1, double visc2_norm,
                                                 double dvdots, double dx, double dxm, double dxp,
                                                 double cs_edge, int c1, int s)
{
    const int c0 = c1 - s, c2 = c1 + s, c3 = c2 + s;
    dvdots = pmin(0.0, dvdots);
    double rho_edge = ddiv(2.0 * A.rho_v[c1] * A.rho_v[c2], A.rho_v[c1] + A.rho_v[c2]);
    double dvx = VX[c1] - VX[c2];
    double dvy = VY[c1] - VY[c2];
    double dvz = VZ[c1] - VZ[c2];
    double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    double dv = sqrt(dv2);
    double psi = 0.0;
    dvdots = ddiv(dv * dt, dx) < 1.e-14 ? 0.0 : ddiv(dvdots, dv);
    double dvxm = VX[c0] - VX[c1];
    double dvxp = VX[c2] - VX[c3];
    double dvym = VY[c0] - VY[c1];
    double dvyp = VY[c2] - VY[c3];
    double dvzm = VZ[c0] - VZ[c1];
    double dvzp = VZ[c2] - VZ[c3];
    double rl = 1.0, rr = 1.0;
    rl = ddiv(dv * dt, dx) < 1.e-14 ? 1.0 : ddiv((dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx, dxp * dv2);
    rr = ddiv(dv * dt, dx) < 1.e-14 ? 1.0 : ddiv((dvxm * dvx + dvym * dvy + dvzm * dvz) * dx, dxm * dv2);
    psi = pmin_list4(0.5 * (rr + rl), 2.0 * rl, 2.0 * rr, 1.0);
    psi = pmax(0.0, psi);
    double q_k_bar = rho_edge * (visc2_norm * dv + sqrt(visc2_norm * visc2_norm * dv2 + (visc1 * cs_edge) * (visc1 * cs_edge)));
    return q_k_bar * (1.0 - psi) * dvdots;
}

// alpha1 (lagran.cpp:211-232), alpha2 (235-253), alpha3 (256-274): dir
// selects the edge direction.  dt is the previous step's dt.
KERNEL void k_alpha(Dev D, R3 R, int dir, double visc1, double visc2_norm)
{
    KBEGIN
    const double dt = A.s->dt;
    if (dir == 0) {
        const int c1 = IX(ix - 1, iy - 1, iz - 1);
        double dx = DXB(ix), dxp = DXB(ix + 1), dxm = DXB(ix - 1);
        double dvdots = -(VX[c1] - VX[c1 + G.sx]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + G.sx]);
        A.alpha1[IX(ix, iy, iz)] = edge_viscosity(A, dt, visc1, visc2_norm, dvdots, dx, dxm, dxp, cs_edge, c1, G.sx);
    } else if (dir == 1) {
        const int c1 = IX(ix, iy - 1, iz - 1);
        double dx = DYB(iy) * HY(ix), dxp = DYB(iy + 1) * HY(ix), dxm = DYB(iy - 1) * HY(ix);
        double dvdots = -(VY[c1] - VY[c1 + G.sy]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + G.sy]);
        A.alpha2[IX(ix, iy, iz)] = edge_viscosity(A, dt, visc1, visc2_norm, dvdots, dx, dxm, dxp, cs_edge, c1, G.sy);
    } else {
        const int c1 = IX(ix, iy, iz - 1);
        double dx = DZB(iz) * HZ(ix, iy), dxp = DZB(iz + 1) * HZ(ix, iy), dxm = DZB(iz - 1) * HZ(ix, iy);
        double dvdots = -(VZ[c1] - VZ[c1 + 1]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + 1]);
        A.alpha3[IX(ix, iy, iz)] = edge_viscosity(A, dt, visc1, visc2_norm, dvdots, dx, dxm, dxp, cs_edge, c1, 1);
    }
}

// squared velocity jump between two vertices
#define SQD(p, q)                                                                        \
    ((VX[p] - VX[q]) * (VX[p] - VX[q]) + (VY[p] - VY[q]) * (VY[p] - VY[q]) +            \
     (VZ[p] - VZ[q]) * (VZ[p] - VZ[q]))

// p_visc and visc_heat (lagran.cpp:279-323) over (0..n+1)^3; p_visc starts
// from the zero fill of lagran.cpp:187.
KERNEL void k_pvisc(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    // vertex (ixm,iym,izm) etc. as offsets from c
    const int mmm = c - sx - sy - 1, omm = c - sy - 1, oom = c - 1, mom = c - sx - 1;
    const int mmo = c - sx - sy, omo = c - sy, ooo = c, moo = c - sx;
    double a1 = SQD(mmm, omm);
    double a2 = SQD(omm, oom);
    double a3 = SQD(oom, mom);
    double a4 = SQD(mom, mmm);
    double a5 = SQD(mmo, omo);
    double a6 = SQD(omo, ooo);
    double a7 = SQD(ooo, moo);
    double a8 = SQD(moo, mmo);
    double a9 = SQD(oom, ooo);
    double a10 = SQD(mom, moo);
    double a11 = SQD(mmm, mmo);
    double a12 = SQD(omm, omo);
    double pv = 0.0;
    pv = pmax(pv, -A.alpha1[c] * sqrt(a1));
    pv = pmax(pv, -A.alpha2[c] * sqrt(a2));
    pv = pmax(pv, -A.alpha3[c] * sqrt(a9));
    A.p_visc[c] = pv;
    double dx = DXB(ix);
    double dy = DYB(iy) * HYC(ix);
    double dz = DZB(iz) * HZ2(ix, iy);
    double vh =
        -0.25 * dy * dz * A.alpha1[c] * a1
        - 0.25 * dx * dz * A.alpha2[c] * a2
        - 0.25 * dy * dz * A.alpha1[c + sy] * a3
        - 0.25 * dx * dz * A.alpha2[c - sx] * a4
        - 0.25 * dy * dz * A.alpha1[c + 1] * a5
        - 0.25 * dx * dz * A.alpha2[c + 1] * a6
        - 0.25 * dy * dz * A.alpha1[c + sy + 1] * a7
        - 0.25 * dx * dz * A.alpha2[c - sx + 1] * a8
        - 0.25 * dx * dy * A.alpha3[c] * a9
        - 0.25 * dx * dy * A.alpha3[c - sx] * a10
        - 0.25 * dx * dy * A.alpha3[c - sx - sy] * a11
        - 0.25 * dx * dy * A.alpha3[c - sy] * a12;
    A.visc_heat[c] = ddiv(vh, A.cv[c]);
}

// Viscous forces (lagran.cpp:330-376) over (0..n)^3.
KERNEL void k_fvisc(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    double dx = DXB(ix);
    double dy = DYB(iy) * HYC(ix);
    double dz = DZB(iz) * HZ2(ix, iy);
    double a1 = A.alpha1[c + sy + 1] * dy * dz;
    double a2 = A.alpha1[c + sx + sy + 1] * dy * dz;
    double a3 = A.alpha2[c + 1] * dx * dz;
    double a4 = A.alpha2[c + sy + 1] * dx * dz;
    double a5 = A.alpha3[c] * dx * dy;
    double a6 = A.alpha3[c + 1] * dx * dy;
    double cvv = A.cv_v[c];
#define FVISC(V)                                                                                   \
    ddiv(a1 * (V[c] - V[c - sx]) + a2 * (V[c] - V[c + sx]) + a3 * (V[c] - V[c - sy]) +                \
             a4 * (V[c] - V[c + sy]) + a5 * (V[c] - V[c - 1]) + a6 * (V[c] - V[c + 1]),               \
         cvv)
    A.fxv[c] = FVISC(VX);
    A.fyv[c] = FVISC(VY);
    A.fzv[c] = FVISC(VZ);
#undef FVISC
}

// ------------------------------------------------- CFL reduction

// set_dt (lagran.cpp:380-443) over (0..n)^3: per-cell min(dt1..dt4),
// reduced per wavefront with lane shuffles, per block through LDS, and
// across blocks with one 64-bit atomicMin on the bit pattern (order-free
// for the non-negative values the CFL produces).
KERNEL void k_set_dt(Dev D, R3 R, double gas_gamma, double mu0)
{
    const Geo G = D.g;
    const Arr A = D.a;
    const int iz = R.z0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int iy = R.y0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int ix = R.x0 + (int)blockIdx.z;
    double m = DBL_MAX;
    if (iz <= R.z1 && iy <= R.y1) {
        const int c = IX(ix, iy, iz);
        const int sx = G.sx, sy = G.sy;
        double dx = DXB(ix);
        double dy = DYB(iy);
        double dz = DZB(iz);
        double dhx = dx, dhy = dy * HYC(ix), dhz = dz * HZC(ix, iy);
        double rho0 = pmax(RHO[c], DBL_EPSILON);
        double cs2 = ddiv(gas_gamma * A.pressure[c], rho0);
        double w1 = ddiv(ddiv(BX[c] * BX[c] + BY[c] * BY[c] + BZ[c] * BZ[c], mu0), rho0);
        double c_visc2 = ddiv(A.p_visc[c], rho0);
        double length = pmin(pmin(dhx, dhy), dhz);
        double dt1 = ddiv(length, sqrt(c_visc2) + sqrt(cs2 + w1 + c_visc2));
        double ax = 0.25 * A.dxab[c];
        double ay = 0.25 * A.dyab[c];
        double az = 0.25 * A.dzab[c];
        double vxbm = (VX[c - sx] + VX[c - sx - sy] + VX[c - sx - 1] + VX[c - sx - sy - 1]) * ax;
        double vxbp = (VX[c] + VX[c - sy] + VX[c - 1] + VX[c - sy - 1]) * ax;
        double vybm = (VY[c - sy] + VY[c - sx - sy] + VY[c - sy - 1] + VY[c - sx - sy - 1]) * ay;
        double vybp = (VY[c] + VY[c - sx] + VY[c - 1] + VY[c - sx - 1]) * ay;
        double vzbm = (VZ[c - 1] + VZ[c - sx - 1] + VZ[c - sy - 1] + VZ[c - sx - sy - 1]) * az;
        double vzbp = (VZ[c] + VZ[c - sx] + VZ[c - sy] + VZ[c - sx - sy]) * az;
        double dvx = fabs(vxbp - vxbm);
        double dvy = fabs(vybp - vybm);
        double dvz = fabs(vzbp - vzbm);
        double avxm = fabs(vxbm), avxp = fabs(vxbp);
        double avym = fabs(vybm), avyp = fabs(vybp);
        double avzm = fabs(vzbm), avzp = fabs(vzbp);
        double volume = A.cv[c];
        double dt2 = ddiv(volume, pmax(pmax(pmax(avxm, avxp), dvx), 1.0e-10 * volume));
        double dt3 = ddiv(volume, pmax(pmax(pmax(avym, avyp), dvy), 1.0e-10 * volume));
        double dt4 = ddiv(volume, pmax(pmax(pmax(avzm, avzp), dvz), 1.0e-10 * volume));
        m = pmin(pmin(pmin(dt1, dt2), dt3), dt4);
    }
    // wavefront reduction (64 lanes)
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = pmin(m, o);
    }
    __shared__ double wmin[16];
    const int lane = threadIdx.x & 63;
    const int wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    const int nw = (blockDim.x * blockDim.y) >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double b = wmin[0];
        for (int w = 1; w < nw; ++w) b = pmin(b, wmin[w]);
        if (b >= 0.0) atomicMin(&A.s->dtmin_bits, (unsigned long long)__double_as_longlong(b));
    }
}

// dt = dt_multiplier * min; the runner keeps min(dt, DBL_MAX)
// (runner.h:158-169, shared_data.h:428-440); dtr = dt (lagran.cpp:441-442).
KERNEL void k_dt_final(Dev D, double dt_multiplier)
{
    const Arr A = D.a;
    double red = __longlong_as_double((long long)A.s->dtmin_bits);
    double dt = dt_multiplier * red;
    double tdt = DBL_MAX;
    tdt = dt < tdt ? dt : tdt;
    A.s->dt = tdt;
    A.s->dtr = dt;
    A.s->dt_sub = tdt / 2.0;  // resistive substeps: int(dt/dtr)+1 = 2 (lagran.cpp:159-161)
}

// End of a step: count it.
KERNEL void k_step_done(Dev D) { D.a.s->nstep += 1; }

// ------------------------------------------------- resistive path

// eta_calc (lagran.cpp:445-476) over (-1..n+1)^3.
KERNEL void k_eta_calc(Dev D, R3 R, double j_max, double eta_background, double eta0)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    double dxc = DXC(ix), dyc = DYC(iy), dzc = DZC(iz);
    double jx = ddiv(BZ[c + sy] - BZ[c], dyc) - ddiv(BY[c + 1] - BY[c], dzc);
    double jxp = ddiv(BZ[c + sx + sy] - BZ[c + sx], dyc) - ddiv(BY[c + sx + 1] - BY[c + sx], dzc);
    double jy = ddiv(BX[c + 1] - BX[c], dzc) - ddiv(BZ[c + sx] - BZ[c], dxc);
    double jyp = ddiv(BX[c + sy + 1] - BX[c + sy], dzc) - ddiv(BZ[c + sx + sy] - BZ[c + sy], dxc);
    double jz = ddiv(BY[c + sx] - BY[c], dxc) - ddiv(BX[c + sy] - BX[c], dyc);
    double jzp = ddiv(BY[c + sx + 1] - BY[c + 1], dxc) - ddiv(BX[c + sy + 1] - BX[c + 1], dyc);
    jx = (jx + jxp) * 0.5;
    jy = (jy + jyp) * 0.5;
    jz = (jz + jzp) * 0.5;
    double j_local = sqrt(jx * jx + jy * jy + jz * jz);
    A.eta[c] = j_local > j_max ? eta_background + eta0 : eta_background;
}

// Copy of B into bx1..bz1 over the centred extent: the reference's
// pw::assign(data.bx1, data.bx) at lagran.cpp:482 throws on the extent
// mismatch; this is its element-wise reading over bx1's extent.
KERNEL void k_res_copy(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    A.bx1[c] = BX[c];
    A.by1[c] = BY[c];
    A.bz1[c] = BZ[c];
}

// rkstep (lagran.cpp:507-535) over (0..n)^3.
KERNEL void k_rkstep(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    double dxc = DXC(ix), dyc = DYC(iy), dzc = DZC(iz);
    double jx1 = ddiv(BZ[c + sy] - BZ[c], dyc) - ddiv(BY[c + 1] - BY[c], dzc);
    double jx2 = ddiv(BZ[c + sx + sy] - BZ[c + sx], dyc) - ddiv(BY[c + sx + 1] - BY[c + sx], dzc);
    double jy1 = ddiv(BX[c + 1] - BX[c], dzc) - ddiv(BZ[c + sx] - BZ[c], dxc);
    double jy2 = ddiv(BX[c + sy + 1] - BX[c + sy], dzc) - ddiv(BZ[c + sx + sy] - BZ[c + sy], dxc);
    double jz1 = ddiv(BY[c + sx] - BY[c], dxc) - ddiv(BX[c + sy] - BX[c], dyc);
    double jz2 = ddiv(BY[c + sx + 1] - BY[c + 1], dxc) - ddiv(BX[c + sy + 1] - BX[c + 1], dyc);
    double jx = 0.5 * (jx1 + jx2);
    double jy = 0.5 * (jy1 + jy2);
    double jz = 0.5 * (jz1 + jz2);
    double eta = A.eta[c];
    A.flux_x[c] = -jx * eta * dxc * 0.5;
    A.flux_y[c] = -jy * eta * dyc * 0.5;
    A.flux_z[c] = -jz * eta * dzc * 0.5;
    A.curlb[c] = eta * (jx * jx + jy * jy + jz * jz);
}

// bstep (lagran.cpp:537-590): comp 0/1/2 updates bx/by/bz over its range.
KERNEL void k_bstep(Dev D, R3 R, int comp)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt_sub;
    const double *fx = A.flux_x, *fy = A.flux_y, *fz = A.flux_z;
    if (comp == 0) {
        double area = DYB(iy) * DZB(iz);
        BX[c] = A.bx1[c] + ddiv((fz[c] - fz[c - sy] + fz[c - 1] - fz[c - sy - 1] - fy[c] + fy[c - 1] - fy[c - sy] + fy[c - sy - 1]) * dt, area);
    } else if (comp == 1) {
        double area = DXB(ix) * DZB(iz);
        BY[c] = A.by1[c] + ddiv((fx[c] - fx[c - 1] + fx[c - sx] - fx[c - sx - 1] - fz[c] + fz[c - sx] - fz[c - 1] + fz[c - sx - 1]) * dt, area);
    } else {
        double area = DXB(ix) * DYB(iy);
        BZ[c] = A.bz1[c] + ddiv((fy[c] - fy[c - sx] + fy[c - sy] - fy[c - sx - sy] - fx[c] + fx[c - sy] - fx[c - sx] + fx[c - sx - sy]) * dt, area);
    }
}

// Ohmic heating of the electrons (lagran.cpp:490-500) over (1..n)^3.
KERNEL void k_res_energy(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt_sub;
    const double *cb = A.curlb;
    double sum_curlb = cb[c] + cb[c - sx] + cb[c - sy] + cb[c - sx - sy] + cb[c - 1] + cb[c - sx - 1] + cb[c - sy - 1] + cb[c - sx - sy - 1];
    EE[c] += ddiv(sum_curlb * dt, 8.0 * RHO[c]);
}

// ------------------------------------------------- predictor

// b_field_and_cv1_update (lagran.cpp:806-882) over (-1..n+2)^3.
KERNEL void k_bfield_cv1(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    // vertex offsets: (i,j,k) o, (i-1) m
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = 0.25 * (VX[ooo] + VX[omo] + VX[oom] + VX[omm]);
    double vxbm = 0.25 * (VX[moo] + VX[mmo] + VX[mom] + VX[mmm]);
    double vyb = 0.25 * (VY[ooo] + VY[moo] + VY[oom] + VY[mom]);
    double vybm = 0.25 * (VY[omo] + VY[mmo] + VY[omm] + VY[mmm]);
    double vzb = 0.25 * (VZ[ooo] + VZ[moo] + VZ[omo] + VZ[mmo]);
    double vzbm = 0.25 * (VZ[oom] + VZ[mom] + VZ[omm] + VZ[mmm]);
    double vol = A.cv[c];
    double dvxdx = ddiv(vxb * A.dxab[c] - vxbm * A.dxab[c - sx], vol);
    double dvydy = ddiv(vyb * A.dyab[c] - vybm * A.dyab[c - sy], vol);
    double dvzdz = ddiv(vzb * A.dzab[c] - vzbm * A.dzab[c - 1], vol);
    double dv = (dvxdx + dvydy + dvzdz) * dt / 2.0;
    A.cv1[c] = vol * (1.0 + dv);

    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[oom] + VX[mom]);
    vxbm = 0.25 * (VX[omo] + VX[mmo] + VX[omm] + VX[mmm]);
    vyb = 0.25 * (VY[ooo] + VY[omo] + VY[oom] + VY[omm]);
    vybm = 0.25 * (VY[moo] + VY[mmo] + VY[mom] + VY[mmm]);
    double dvxdy = ddiv(vxb * A.dyab[c] - vxbm * A.dyab[c - sy], vol);
    double dvydx = ddiv(vyb * A.dxab[c] - vybm * A.dxab[c - sx], vol);

    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[omo] + VX[mmo]);
    vxbm = 0.25 * (VX[oom] + VX[mom] + VX[omm] + VX[mmm]);
    vzb = 0.25 * (VZ[ooo] + VZ[omo] + VZ[oom] + VZ[omm]);
    vzbm = 0.25 * (VZ[moo] + VZ[mmo] + VZ[mom] + VZ[mmm]);
    double dvxdz = ddiv(vxb * A.dzab[c] - vxbm * A.dzab[c - 1], vol);
    double dvzdx = ddiv(vzb * A.dxab[c] - vzbm * A.dxab[c - sx], vol);

    vyb = 0.25 * (VY[ooo] + VY[moo] + VY[omo] + VY[mmo]);
    vybm = 0.25 * (VY[oom] + VY[mom] + VY[omm] + VY[mmm]);
    vzb = 0.25 * (VZ[ooo] + VZ[moo] + VZ[oom] + VZ[mom]);
    vzbm = 0.25 * (VZ[omo] + VZ[mmo] + VZ[omm] + VZ[mmm]);
    double dvydz = ddiv(vyb * A.dzab[c] - vybm * A.dzab[c - 1], vol);
    double dvzdy = ddiv(vzb * A.dyab[c] - vzbm * A.dyab[c - sy], vol);

    double bx1 = A.bx1[c], by1 = A.by1[c], bz1 = A.bz1[c];
    double w3 = bx1 * dvxdx + by1 * dvxdy + bz1 * dvxdz;
    double w4 = bx1 * dvydx + by1 * dvydy + bz1 * dvydz;
    double w5 = bx1 * dvzdx + by1 * dvzdy + bz1 * dvzdz;
    // bx1 *= cv1 of predictor_step (lagran.cpp:598-601) fused: the same
    // product of the stored value, so bit-identical.
    double nbx = ddiv(bx1 + w3 * dt / 2.0, 1.0 + dv);
    double nby = ddiv(by1 + w4 * dt / 2.0, 1.0 + dv);
    double nbz = ddiv(bz1 + w5 * dt / 2.0, 1.0 + dv);
    double cv1 = A.cv1[c];
    A.bx1[c] = nbx * cv1;
    A.by1[c] = nby * cv1;
    A.bz1[c] = nbz * cv1;
}

// Energy/pressure predictor (lagran.cpp:604-614) over (0..n+1)^3.
KERNEL void k_energy_pred(Dev D, R3 R, double gas_gamma)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const double dt = A.s->dt;
    double rho = RHO[c], cv = A.cv[c], cv1 = A.cv1[c];
    double dv = ddiv(cv1, cv) - 1.0;
    double e1_e = EE[c] - ddiv(A.p_e[c] * dv, rho);
    double e1_i = EI[c] - ddiv(A.p_i[c] * dv, rho);
    e1_i += ddiv(A.visc_heat[c] * dt / 2.0, rho);
    double pe = ddiv(e1_e * (gas_gamma - 1.0) * rho * cv, cv1);
    double pi = ddiv(e1_i * (gas_gamma - 1.0) * rho * cv, cv1);
    A.p_e[c] = pe;
    A.p_i[c] = pi;
    A.pressure[c] = pe + pi;
}

// Forces and half-step velocities (lagran.cpp:619-731) over (0..n)^3.
KERNEL void k_forces(Dev D, R3 R, double mu0, int geometry)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const double *P = A.pressure, *C1 = A.cv1, *B1x = A.bx1, *B1y = A.by1, *B1z = A.bz1;
    double dx = DXC(ix), dy = DYC(iy), dz = DZC(iz);
    const double h2 = HY(ix), h2x = HY(ix + 1);
    const double h3 = HZ(ix, iy), h3x = HZ(ix + 1, iy), h3y = HZ(ix, iy + 1);
    const double h2c = HYC(ix), h3c = HZC(ix, iy);
    double dhy = h2c * dy, dhz = h3c * dz;
    const int x_ = c + sx, y_ = c + sy, z_ = c + 1, xy = c + sx + sy, xz = c + sx + 1, yz = c + sy + 1, xyz = c + sx + sy + 1;
    double pp = P[c], ppx = P[x_], ppy = P[y_], ppz = P[z_], ppxy = P[xy], ppxz = P[xz], ppyz = P[yz], ppxyz = P[xyz];
    double w1 = pp + ppy + ppz + ppyz;
    double w2 = ppx + ppxy + ppxz + ppxyz;
    double fx = ddiv((w1 - w2) * 0.25, dx);
    w1 = pp + ppx + ppz + ppxz;
    w2 = ppy + ppxy + ppyz + ppxyz;
    double fy = ddiv((w1 - w2) * 0.25, dhy);
    w1 = pp + ppx + ppy + ppxy;
    w2 = ppz + ppxz + ppyz + ppxyz;
    double fz = ddiv((w1 - w2) * 0.25, dhz);

    double cvx = C1[c] + C1[y_] + C1[z_] + C1[yz];
    double cvxp = C1[x_] + C1[xy] + C1[xz] + C1[xyz];
    double cvy = C1[c] + C1[x_] + C1[z_] + C1[xz];
    double cvyp = C1[y_] + C1[xy] + C1[yz] + C1[xyz];
    double cvz = C1[c] + C1[x_] + C1[y_] + C1[xy];
    double cvzp = C1[z_] + C1[xz] + C1[yz] + C1[xyz];

    w1 = ddiv((B1z[c] + B1z[x_] + B1z[z_] + B1z[xz]) * h3, cvy);
    w2 = ddiv((B1z[y_] + B1z[xy] + B1z[yz] + B1z[xyz]) * h3y, cvyp);
    double jx = ddiv(ddiv(w2 - w1, dhy), h3c);
    w1 = ddiv(B1y[c] + B1y[x_] + B1y[y_] + B1y[xy], cvz);
    w2 = ddiv(B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz], cvzp);
    jx -= ddiv(w2 - w1, dhz);

    w1 = ddiv((B1z[c] + B1z[y_] + B1z[z_] + B1z[yz]) * h3, cvx);
    w2 = ddiv((B1z[x_] + B1z[xy] + B1z[xz] + B1z[xyz]) * h3x, cvxp);
    double jy = ddiv(ddiv(-(w2 - w1), dx), h3c);
    w1 = ddiv(B1x[c] + B1x[x_] + B1x[y_] + B1x[xy], cvz);
    w2 = ddiv(B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz], cvzp);
    jy += ddiv(w2 - w1, dhz);

    w1 = ddiv((B1y[c] + B1y[y_] + B1y[z_] + B1y[yz]) * h2, cvx);
    w2 = ddiv((B1y[x_] + B1y[xy] + B1y[xz] + B1y[xyz]) * h2x, cvxp);
    double jz = ddiv(ddiv(w2 - w1, dx), h2c);
    w1 = ddiv(B1x[c] + B1x[x_] + B1x[z_] + B1x[xz], cvy);
    w2 = ddiv(B1x[y_] + B1x[xy] + B1x[yz] + B1x[xyz], cvyp);
    jz -= ddiv(w2 - w1, dhy);

    double bxv = ddiv(B1x[c] + B1x[x_] + B1x[y_] + B1x[xy] + B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz], cvx + cvxp);
    double byv = ddiv(B1y[c] + B1y[x_] + B1y[y_] + B1y[xy] + B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz], cvx + cvxp);
    double bzv = ddiv(B1z[c] + B1z[x_] + B1z[y_] + B1z[xy] + B1z[z_] + B1z[xz] + B1z[yz] + B1z[xyz], cvx + cvxp);
    fx += ddiv(jy * bzv - jz * byv, mu0);
    fy += ddiv(jz * bxv - jx * bzv, mu0);
    fz += ddiv(jx * byv - jy * bxv, mu0);
    // geometry corrections (lagran.cpp:711-721); cot(yb) and xb from the host
    if (geometry == 2) {
        const double cotantheta = A.cotv[iy + 2], r = A.xbv[ix + 2], rv = A.rho_v[c];
        const double ux = VX[c], uy = VY[c], uz = VZ[c];
        fx += ddiv(rv * (uy * uy + uz * uz), r);
        fy -= ddiv(rv * (uy * ux - cotantheta * uz * uz), r);
        fz -= ddiv(rv * (uz * ux + cotantheta * uz * uy), r);
    } else if (geometry == 1) {
        const double r = A.xbv[ix + 2], rv = A.rho_v[c];
        const double ux = VX[c], uy = VY[c];
        fx += ddiv(rv * uy * uy, r);
        fy -= ddiv(rv * uy * ux, r);
    }
    A.fx[c] = fx;
    A.fy[c] = fy;
    A.fz[c] = fz;
    double rv = A.rho_v[c];
    VX1[c] = VX[c] + ddiv(dt / 2.0 * (A.fxv[c] + fx), rv);
    VY1[c] = VY[c] + ddiv(dt / 2.0 * (A.fyv[c] + fy), rv);
    VZ1[c] = VZ[c] + ddiv(dt / 2.0 * (A.fzv[c] + fz), rv);
}

// ------------------------------------------------- corrector

#define DOT(p, q)                                                                                   \
    (1 * ((VX[p] - VX[q]) * (VX1[p] - VX1[q]) + (VY[p] - VY[q]) * (VY1[p] - VY1[q]) +               \
          (VZ[p] - VZ[q]) * (VZ1[p] - VZ1[q])))

// shock_heating (lagran.cpp:884-993) over (0..n+1)^3, with the dy*dy
// factors of the a9 and a11 terms (lagran.cpp:987).
KERNEL void k_shock_heating(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int ooo = c, moo = c - sx, omo = c - sy, mmo = c - sx - sy;
    const int oop = c + 1, mop = c - sx + 1, omp = c - sy + 1, mmp = c - sx - sy + 1;
    const int oom = c - 1, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double a1 = DOT(mmo, omo);
    double a2 = DOT(omo, ooo);
    double a3 = DOT(ooo, moo);
    double a4 = DOT(moo, mmo);
    double a5 = DOT(mmp, omp);
    double a6 = DOT(omp, oop);
    double a7 = DOT(oop, mop);
    double a8 = DOT(mop, mmp);
    double a9 = DOT(oom, ooo);
    double a10 = DOT(mom, moo);
    double a11 = DOT(mmm, mmo);
    double a12 = DOT(omm, omo);
    double dx = DXB(ix);
    double dy = DYB(iy) * HYC(ix);
    double dz = DZB(iz) * HZ2(ix, iy);
    double vh =
        (-0.25 * dy * dz * A.alpha1[c] * a1 - 0.25 * dx * dz * A.alpha2[c] * a2 -
         0.25 * dy * dz * A.alpha1[c + sy] * a3 - 0.25 * dx * dz * A.alpha2[c - sx] * a4 -
         0.25 * dy * dz * A.alpha1[c + 1] * a5 - 0.25 * dx * dz * A.alpha2[c + 1] * a6 -
         0.25 * dy * dz * A.alpha1[c + sy + 1] * a7 - 0.25 * dx * dz * A.alpha2[c - sx + 1] * a8 -
         0.25 * dy * dy * A.alpha3[c] * a9 - 0.25 * dx * dy * A.alpha3[c - sx] * a10 -
         0.25 * dy * dy * A.alpha3[c - sx - sy] * a11 - 0.25 * dx * dy * A.alpha3[c - sy] * a12);
    A.visc_heat[c] = pmax(ddiv(vh, A.cv[c]), 0.0);
}

// Velocity correction (lagran.cpp:754-758) over (0..n)^3.
KERNEL void k_vel_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const double dt = A.s->dt;
    double rv = A.rho_v[c];
    VX[c] += ddiv(dt * (A.fxv[c] + A.fx[c]), rv);
    VY[c] += ddiv(dt * (A.fyv[c] + A.fy[c]), rv);
    VZ[c] += ddiv(dt * (A.fzv[c] + A.fz[c]), rv);
}

// Density and energy correction (lagran.cpp:762-796) over (1..n)^3.
KERNEL void k_rho_e_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = 0.25 * (VX1[ooo] + VX1[omo] + VX1[oom] + VX1[omm]);
    double vxbm = 0.25 * (VX1[moo] + VX1[mmo] + VX1[mom] + VX1[mmm]);
    double vyb = 0.25 * (VY1[ooo] + VY1[moo] + VY1[oom] + VY1[mom]);
    double vybm = 0.25 * (VY1[omo] + VY1[mmo] + VY1[omm] + VY1[mmm]);
    double vzb = 0.25 * (VZ1[ooo] + VZ1[moo] + VZ1[omo] + VZ1[mmo]);
    double vzbm = 0.25 * (VZ1[oom] + VZ1[mom] + VZ1[omm] + VZ1[mmm]);
    double vol = A.cv[c];
    double dvxdx = ddiv(vxb * A.dxab[c] - vxbm * A.dxab[c - sx], vol);
    double dvydy = ddiv(vyb * A.dyab[c] - vybm * A.dyab[c - sy], vol);
    double dvzdz = ddiv(vzb * A.dzab[c] - vzbm * A.dzab[c - 1], vol);
    double dv = (dvxdx + dvydy + dvzdz) * dt;
    A.cv1[c] = vol * (1.0 + dv);
    double rho = RHO[c];
    EE[c] -= ddiv(dv * A.p_e[c], rho);
    EI[c] += ddiv(dt * A.visc_heat[c] - dv * A.p_i[c], rho);
    RHO[c] = ddiv(rho, 1.0 + dv);
}

// ------------------------------------------------- remap

// B *= area (remap.cpp:46-50) and B /= (area + none_zero) (91-95) over each
// component's full extent.
KERNEL void k_b_area(Dev D, R3 R, int comp, int divide)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double *B = fsel(A, 6 + comp);
    const double *ar = comp == 0 ? A.dxab : comp == 1 ? A.dyab : A.dzab;
    if (divide)
        B[c] = ddiv(B[c], ar[c] + DBL_EPSILON);
    else
        B[c] *= ar[c];
}

// Cell volumes before/after the remap of direction d, db1, and the
// per-sweep resets dm = 0 and rho1 = rho (xremap.cpp:24-70,
// yremap.cpp:23-76, zremap.cpp:25-79) over the centred extent.
KERNEL void k_remap_prep(Dev D, R3 R, int d, double xpass, double ypass, double zpass)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = (VX1[ooo] + VX1[omo] + VX1[oom] + VX1[omm]) * 0.25;
    double vxbm = (VX1[moo] + VX1[mmo] + VX1[mom] + VX1[mmm]) * 0.25;
    double vyb = (VY1[ooo] + VY1[moo] + VY1[oom] + VY1[mom]) * 0.25;
    double vybm = (VY1[omo] + VY1[mmo] + VY1[omm] + VY1[mmm]) * 0.25;
    double vzb = (VZ1[ooo] + VZ1[moo] + VZ1[omo] + VZ1[mmo]) * 0.25;
    double vzbm = (VZ1[oom] + VZ1[mom] + VZ1[omm] + VZ1[mmm]) * 0.25;
    double vol = A.cv[c];
    double dv;
    if (d == 0) {
        double dvxdx = ddiv(vxb * A.dxab[c] - vxbm * A.dxab[c - sx], vol);
        double dvydy = ddiv(ypass * (vyb * A.dyab[c] - vybm * A.dyab[c - sy]), vol);
        double dvzdz = ddiv(zpass * (vzb * A.dzab[c] - vzbm * A.dzab[c - 1]), vol);
        dv = (dvydy + dvzdz) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvxdx * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = DXB(ix) + (vxb - vxbm) * dt;
    } else if (d == 1) {
        double dvxdx = ddiv(xpass * (vxb * A.dxab[c] - vxbm * A.dxab[c - sx]), vol);
        double dvydy = ddiv(vyb * A.dyab[c] - vybm * A.dyab[c - sy], vol);
        double dvzdz = ddiv(zpass * (vzb * A.dzab[c] - vzbm * A.dzab[c - 1]), vol);
        dv = (dvxdx + dvzdz) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvydy * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = HYC(ix) * DYB(iy) + (vyb - vybm) * dt;
    } else {
        double dvxdx = ddiv(xpass * (vxb * A.dxab[c] - vxbm * A.dxab[c - sx]), vol);
        double dvydy = ddiv(ypass * (vyb * A.dyab[c] - vybm * A.dyab[c - sy]), vol);
        double dvzdz = ddiv(vzb * A.dzab[c] - vzbm * A.dzab[c - 1], vol);
        dv = (dvxdx + dvydy) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvzdz * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = HZC(ix, iy) * DZB(iz) + (vzb - vzbm) * dt;
    }
    DM[c] = 0.0;
    A.rho1[c] = RHO[c];
}

// cvc1 (xremap.cpp:74-81 and twins) over (-1..n+1)^3.
KERNEL void k_cvc1(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    A.cvc1[c] = 0.125 * PLAIN8(A.cv1, c);
}

// Constrained-transport flux of a face magnetic flux B along direction d
// (stride sd): v_advect = mean of w at c and c+ow, db1 averaged with c+tw.
// vx_by_flux, vx_bz_flux (xremap.cpp:268-371), vy_bx_flux, vy_bz_flux
// (yremap.cpp:291-393), vz_bx_flux, vz_by_flux (zremap.cpp:250-346).
KERNEL void k_b_flux(Dev D, R3 R, int bcomp, int wcomp, int sd, int ow, int tw)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const double dt = A.s->dt;
    const double *B = fsel(A, 6 + bcomp);
    const double *w = fsel(A, 9 + wcomp);
    const double *db1 = A.db1;
    double v_advect = (w[c] + w[c + ow]) * 0.5;
    double o_v = v_advect * dt;
    double dbx = (db1[c] + db1[c + tw]) * 0.5;
    double dbxp = (db1[c + sd] + db1[c + sd + tw]) * 0.5;
    double dbxp2 = (db1[c + 2 * sd] + db1[c + 2 * sd + tw]) * 0.5;
    double dbxm = (db1[c - sd] + db1[c - sd + tw]) * 0.5;
    double fm = ddiv(B[c - sd], dbxm);
    double fi = ddiv(B[c], dbx);
    double fp = ddiv(B[c + sd], dbxp);
    double fp2 = ddiv(B[c + 2 * sd], dbxp2);
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = A.cvc1[c];
    double dxcu = A.cvc1[c - sd] * vad_p + A.cvc1[c + sd] * vad_m;
    double dxbu = A.cv1[c] * vad_p + A.cv1[c + sd] * vad_m;
    double dxu = dbx * vad_p + dbxp * vad_m;
    double phi = ddiv(fabs(o_v), dxu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    A.rflux[c] = (fu + Di * (1.0 - phi)) * o_v;
}

// B[c] = B[c] + sgn*flux[c] - sgn*flux[c - so]: the CT updates that follow
// each B flux (xremap.cpp:87-114 and twins).  sgn = -1 computes
// B - flux + flux(m), sgn = +1 computes B + flux - flux(m).
KERNEL void k_b_ct(Dev D, R3 R, int bcomp, int so, int sgn)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double *B = fsel(A, 6 + bcomp);
    const double *fl = A.rflux;
    if (sgn < 0)
        B[c] = B[c] - fl[c] + fl[c - so];
    else
        B[c] = B[c] + fl[c] - fl[c - so];
}

// x_mass_flux (xremap.cpp:373-422), y_mass_flux (yremap.cpp:395-444),
// z_mass_flux (zremap.cpp:349-397).
KERNEL void k_mass_flux(Dev D, R3 R, int d)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int sd = d == 0 ? sx : d == 1 ? sy : 1;
    const double dt = A.s->dt;
    double area, v_advect;
    if (d == 0) {
        area = A.dxab[c];
        v_advect = (VX1[c] + VX1[c - sy] + VX1[c - 1] + VX1[c - sy - 1]) * 0.25;
    } else if (d == 1) {
        area = A.dyab[c];
        v_advect = (VY1[c - sx] + VY1[c] + VY1[c - sx - 1] + VY1[c - 1]) * 0.25;
    } else {
        area = A.dzab[c];
        v_advect = (VZ1[c - sx] + VZ1[c] + VZ1[c - sx - sy] + VZ1[c - sy]) * 0.25;
    }
    double o_v = v_advect * dt * area;
    double fm = RHO[c - sd], fi = RHO[c], fp = RHO[c + sd], fp2 = RHO[c + 2 * sd];
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = A.cvc1[c];
    double dxcu = A.cvc1[c - sd] * vad_p + A.cvc1[c + sd] * vad_m;
    double dxbu = A.cv1[c] * vad_p + A.cv1[c + sd] * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    DM[c] = (fu + Di * (1.0 - phi)) * o_v;
}

// rho = (rho1*cv1 + dm(m) - dm) / cv2 over (1..n)^3.
KERNEL void k_rho_remap(Dev D, R3 R, int sd)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    RHO[c] = ddiv(A.rho1[c] * A.cv1[c] + DM[c - sd] - DM[c], A.cv2[c]);
}

// x/y/z_energy_flux (xremap.cpp:429-483, yremap.cpp:452-506,
// zremap.cpp:405-459) for energy e (1 electron, 2 ion).
KERNEL void k_energy_flux(Dev D, R3 R, int d, int e)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int sd = d == 0 ? sx : d == 1 ? sy : 1;
    const double dt = A.s->dt;
    const double *E = fsel(A, e);
    double area, v_advect;
    if (d == 0) {
        area = A.dxab[c];
        v_advect = (VX1[c] + VX1[c - sy] + VX1[c - 1] + VX1[c - sy - 1]) * 0.25;
    } else if (d == 1) {
        area = A.dyab[c];
        v_advect = (VY1[c - sx] + VY1[c] + VY1[c - sx - 1] + VY1[c - 1]) * 0.25;
    } else {
        area = A.dzab[c];
        v_advect = (VZ1[c - sx] + VZ1[c] + VZ1[c - sx - sy] + VZ1[c - sy]) * 0.25;
    }
    double o_v = v_advect * dt * area;
    double fm = E[c - sd], fi = E[c], fp = E[c + sd], fp2 = E[c + 2 * sd];
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = A.cvc1[c];
    double dxcu = A.cvc1[c - sd] * vad_p + A.cvc1[c + sd] * vad_m;
    double dxbu = A.cv1[c] * vad_p + A.cv1[c + sd] * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double rhou = A.rho1[c] * vad_p + A.rho1[c + sd] * vad_m;
    double dmu = ddiv(ddiv(fabs(DM[c]), dxbu), rhou);
    A.rflux[c] = (fu + Di * (1.0 - dmu)) * DM[c];
}

// e = (e*cv1*rho1 + flux(m) - flux) / (cv2*rho) over (1..n)^3.
KERNEL void k_energy_remap(Dev D, R3 R, int sd, int e)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double *E = fsel(A, e);
    E[c] = ddiv(E[c] * A.cv1[c] * A.rho1[c] + A.rflux[c - sd] - A.rflux[c], A.cv2[c] * RHO[c]);
}

// Vertex density before remap (xremap.cpp:159-176, yremap.cpp:163-183,
// zremap.cpp:150-168); x divides (s*0.125)/cvc1, y and z multiply by
// (0.125/cvc1).  Also the vertex cv2 (into cv2v) where inside (0..n)^3.
KERNEL void k_vertex_rho(Dev D, R3 R, int d)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double s = SUM8(A.rho1, A.cv1, c);
    if (d == 0)
        A.rrho_v[c] = ddiv(s * 0.125, A.cvc1[c]);
    else
        A.rrho_v[c] = s * ddiv(0.125, A.cvc1[c]);
}

// Vertex cv2 (xremap.cpp:180-190 and twins) over (0..n)^3 into cv2v.
KERNEL void k_vertex_cv2(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    A.cv2v[c] = PLAIN8(A.cv2, c) * 0.125;
}

// Face-centred sweep velocity (xremap.cpp:196-203 and twins) into wsh.
KERNEL void k_vshift(Dev D, R3 R, int wcomp, int sd)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const double *w = fsel(A, 9 + wcomp);
    A.wsh[c] = (w[c] + w[c + sd]) * 0.5;
}

// The shifted sweep velocity written back into v1 (xremap.cpp:203 and
// twins), over k_vshift's range, after every reader of the unshifted value.
KERNEL void k_wback(Dev D, R3 R, int wcomp)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double *w = fsel(A, 9 + wcomp);
    w[c] = A.wsh[c];
}

// Vertex mass flux (xremap.cpp:209-219 and twins) into dmv.
KERNEL void k_dm_shift(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    A.dmv[c] = PLAIN8(DM, c) * 0.125;
}

// Vertex density after remap (xremap.cpp:225-233 and twins) over (0..n)^3.
KERNEL void k_rho_v1(Dev D, R3 R, int sd)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    A.rho_v1[c] = ddiv(A.rrho_v[c] * A.cvc1[c] + A.dmv[c - sd] - A.dmv[c], A.cv2v[c]);
}

// x/y/z_mom_flux (xremap.cpp:485-566, yremap.cpp:508-588, zremap.cpp:461-541)
// for velocity component comp, multiplied by dm as in their second kernel.
// The z flux takes dyac as its area (zremap.cpp:475) and divides by the
// product (dzbu*rhou) (zremap.cpp:505).
KERNEL void k_mom_flux(Dev D, R3 R, int d, int comp)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sd = d == 0 ? G.sx : d == 1 ? G.sy : 1;
    const double dt = A.s->dt;
    const double *f = fsel(A, 3 + comp);
    double area = d == 0 ? A.dxac[c] : A.dyac[c];
    double v_advect = A.wsh[c];
    double o_v = v_advect * dt * area;
    double fm = f[c - sd], fi = f[c], fp = f[c + sd], fp2 = f[c + 2 * sd];
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = A.cv1[c + sd];
    double dxcu = A.cv1[c] * vad_p + A.cv1[c + 2 * sd] * vad_m;
    double dxbu = A.cvc1[c] * vad_p + A.cvc1[c + sd] * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double rhou = A.rrho_v[c] * vad_p + A.rrho_v[c + sd] * vad_m;
    double dm = A.dmv[c];
    double dmu = d == 2 ? ddiv(fabs(dm), dxbu * rhou) : ddiv(ddiv(fabs(dm), dxbu), rhou);
    double fl = fu + Di * (1.0 - dmu);
    A.rflux[c] = fl * dm;
}

// v = (rho_v*v*cvc1 + flux(m) - flux) / (cv2*rho_v1) over (0..n)^3.
KERNEL void k_mom_remap(Dev D, R3 R, int sd, int comp)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    double *v = fsel(A, 3 + comp);
    v[c] = ddiv(A.rrho_v[c] * v[c] * A.cvc1[c] + A.rflux[c - sd] - A.rflux[c], A.cv2v[c] * A.rho_v1[c]);
}

// ------------------------------------------------- boundary conditions

// One dimension of the built-in BCs of up to nine variables (blockIdx.z):
// each thread owns one line along dim and applies the lower edge then the
// upper edge to it, exactly as the reference's successive slice operations
// (variableDef.h:539-605, builtInBoundaryConditions.h:54-300) leave it.
// Lines span the other two dimensions' full extents including ghosts.
__device__ __forceinline__ void bc_edge(double *p, int c0, int sd, int n, int s, int edge, int kind, double val)
{
    if (kind == 0) return;
    const int LB = s ? 0 : 1, UB = n;
    const int g0 = edge ? n + 1 - s : (s ? -2 : -1);
    const int g1 = edge ? n + 2 : 0;
    const int first = edge ? g0 : g1;
    const int last = edge ? UB : LB;
    for (int g = g0; g <= g1; ++g) {
        double out;
        if (kind == 1) {
            out = p[c0 + (edge ? g - n : g + n) * sd];
        } else if (kind == 2) {
            out = p[c0 + last * sd];
        } else if (kind == 3) {
            const double bv = p[c0 + last * sd];
            const double boundary = (bv + bv) / 2.0;
            const double delta = p[c0 + (last - (g - first)) * sd] - boundary;
            out = boundary - delta;
        } else {
            out = val;
        }
        p[c0 + g * sd] = out;
    }
}

KERNEL void k_bc(Dev D, lare::BCSet S, int dim)
{
    const Geo G = D.g;
    lare::BCVar V;
    switch (blockIdx.z) {
    case 0: V = S.v[0]; break;
    case 1: V = S.v[1]; break;
    case 2: V = S.v[2]; break;
    case 3: V = S.v[3]; break;
    case 4: V = S.v[4]; break;
    case 5: V = S.v[5]; break;
    case 6: V = S.v[6]; break;
    case 7: V = S.v[7]; break;
    default: V = S.v[8]; break;
    }
    // the two other dimensions, the faster (in memory) one second
    const int na = dim == 0 ? G.ny : G.nx;
    const int nb = dim == 2 ? G.ny : G.nz;
    const int la = V.sa ? -2 : -1, ua = na + 2;
    const int lb = V.sb ? -2 : -1, ub = nb + 2;
    const int wb = ub - lb + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (ua - la + 1) * wb) return;
    const int ia = la + t / wb, ib = lb + t % wb;
    int c0, sd, n;
    if (dim == 0) {
        c0 = IX(0, ia, ib);
        sd = G.sx;
        n = G.nx;
    } else if (dim == 1) {
        c0 = IX(ia, 0, ib);
        sd = G.sy;
        n = G.ny;
    } else {
        c0 = IX(ia, ib, 0);
        sd = 1;
        n = G.nz;
    }
    bc_edge(V.p, c0, sd, n, V.sd, 0, V.k0, V.v0);
    bc_edge(V.p, c0, sd, n, V.sd, 1, V.k1, V.v1);
}

// =================================================================== host
// Orchestration of the three LARE3D events: the kernels above in the
// reference's order, with the boundary fills at the reference's call sites
// (boundary.cpp:27-96).

namespace lare {

namespace {

// Launch kernel k over range r.  bpc is the kernel's algorithmic HBM
// traffic per cell (distinct arrays read and written, 8 bytes each), used
// only by the optional profiler.
template <typename... KA, typename... Args>
void launch3_(void (*k)(Dev, R3, KA...), const char *name, double bpc, Grid &g, R3 r, Args... args)
{
    const int nz = r.z1 - r.z0 + 1, ny = r.y1 - r.y0 + 1, nx = r.x1 - r.x0 + 1;
    if (nz <= 0 || ny <= 0 || nx <= 0) return;
    dim3 blk(64, 4, 1);
    dim3 grd((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nx);
    Prof::Rec rec{name, bpc * (double)nz * ny * nx, nullptr, nullptr};
    if (g.prof.on) {
        (void)hipEventCreate(&rec.a);
        (void)hipEventCreate(&rec.b);
        (void)hipEventRecord(rec.a, g.stream);
    }
    hipLaunchKernelGGL(k, grd, blk, 0, g.stream, g.dev, r, static_cast<KA>(args)...);
    if (g.prof.on) {
        (void)hipEventRecord(rec.b, g.stream);
        g.prof.recs.push_back(rec);
    }
}
#define launch3(k, bpc, ...) launch3_(k, #k, bpc, __VA_ARGS__)

inline R3 rng(int x0, int x1, int y0, int y1, int z0, int z1) { return R3{x0, x1, y0, y1, z0, z1}; }

// Stagger per dimension of each registered variable (setup.cpp:45-69).
const int kStag[13][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 0, 0},
                          {0, 1, 0}, {0, 0, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {0, 0, 0}};

// variableRegistry::applyBoundaryConditions for a group of variables:
// dims 0, 1, 2 in turn (variableDef.h:601-605), or one dim (dim >= 0).
void bcs(Grid &g, const int *vars, int nvar, int dim = -1)
{
    const int n3[3] = {g.dev.g.nx, g.dev.g.ny, g.dev.g.nz};
    for (int d = 0; d < 3; ++d) {
        if (dim >= 0 && d != dim) continue;
        const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
        BCSet S{};
        int nv = 0;
        for (int i = 0; i < nvar; ++i) {
            const int v = vars[i];
            if (g.bck[v][d][0] == 0 && g.bck[v][d][1] == 0) continue;
            BCVar &b = S.v[nv++];
            b.p = g.dev.a.f[v];
            b.sd = kStag[v][d];
            b.sa = kStag[v][da];
            b.sb = kStag[v][db];
            b.k0 = g.bck[v][d][0];
            b.k1 = g.bck[v][d][1];
            b.v0 = g.bcv[v][d][0];
            b.v1 = g.bcv[v][d][1];
        }
        if (nv == 0) continue;
        // lines: the product of the other two full extents (at most (n+5)^2)
        const int lines = (n3[da] + 5) * (n3[db] + 5);
        dim3 blk(256, 1, 1);
        dim3 grd((unsigned)((lines + 255) / 256), 1, (unsigned)nv);
        Prof::Rec rec{"k_bc", 0.0, nullptr, nullptr};
        if (g.prof.on) {
            (void)hipEventCreate(&rec.a);
            (void)hipEventCreate(&rec.b);
            (void)hipEventRecord(rec.a, g.stream);
        }
        hipLaunchKernelGGL(k_bc, grd, blk, 0, g.stream, g.dev, S, d);
        if (g.prof.on) {
            (void)hipEventRecord(rec.b, g.stream);
            g.prof.recs.push_back(rec);
        }
    }
}

const int kB[3] = {6, 7, 8};
const int kE[2] = {1, 2};
const int kRho[1] = {0};
const int kV[3] = {3, 4, 5};
const int kV1[3] = {9, 10, 11};
const int kDm[1] = {12};

void bfield_bcs(Grid &g) { bcs(g, kB, 3); }
void energy_bcs(Grid &g) { bcs(g, kE, 2); }
void density_bcs(Grid &g) { bcs(g, kRho, 1); }
void velocity_bcs(Grid &g) { bcs(g, kV, 3); }
void remap_v_bcs(Grid &g) { bcs(g, kV1, 3); }
// bfield_bcs, energy_bcs, density_bcs, velocity_bcs (boundary.cpp:27-33):
// the nine variables are independent, so their fills share launches.
const int kAll[9] = {6, 7, 8, 1, 2, 0, 3, 4, 5};
void boundary_conditions(Grid &g) { bcs(g, kAll, 9); }

void remap_sweep(Grid &g, int d, double xpass, double ypass, double zpass)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const int sx = g.dev.g.sx, sy = g.dev.g.sy;
    const int sd = d == 0 ? sx : d == 1 ? sy : 1;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    const R3 all0 = rng(0, nx, 0, ny, 0, nz);
    const R3 in1 = rng(1, nx, 1, ny, 1, nz);
    launch3(k_remap_prep, 104, g, C, d, xpass, ypass, zpass);
    launch3(k_cvc1, 16, g, rng(-1, nx + 1, -1, ny + 1, -1, nz + 1));
    // constrained transport of the two transverse B components
    if (d == 0) {
        launch3(k_b_flux, 48, g, all0, 1, 0, sx, -1, sy);  // vx_by_flux
        launch3(k_b_ct, 24, g, rng(1, nx, 0, ny, 1, nz), 1, sx, -1);
        launch3(k_b_ct, 24, g, rng(0, nx, 1, ny, 1, nz), 0, sy, +1);
        launch3(k_b_flux, 48, g, all0, 2, 0, sx, -sy, 1);  // vx_bz_flux
        launch3(k_b_ct, 24, g, rng(1, nx, 1, ny, 0, nz), 2, sx, -1);
        launch3(k_b_ct, 24, g, rng(0, nx, 1, ny, 1, nz), 0, 1, +1);
    } else if (d == 1) {
        launch3(k_b_flux, 48, g, all0, 0, 1, sy, -1, sx);  // vy_bx_flux
        launch3(k_b_ct, 24, g, rng(0, nx, 1, ny, 1, nz), 0, sy, -1);
        launch3(k_b_ct, 24, g, rng(1, nx, 0, ny, 1, nz), 1, sx, +1);
        launch3(k_b_flux, 48, g, all0, 2, 1, sy, -sx, 1);  // vy_bz_flux
        launch3(k_b_ct, 24, g, rng(1, nx, 1, ny, 0, nz), 2, sy, -1);
        launch3(k_b_ct, 24, g, rng(1, nx, 0, ny, 1, nz), 1, 1, +1);
    } else {
        launch3(k_b_flux, 48, g, all0, 0, 2, 1, -sy, sx);  // vz_bx_flux
        launch3(k_b_ct, 24, g, rng(0, nx, 1, ny, 1, nz), 0, 1, -1);
        launch3(k_b_ct, 24, g, rng(1, nx, 1, ny, 0, nz), 2, sx, +1);
        launch3(k_b_flux, 48, g, all0, 1, 2, 1, -sx, sy);  // vz_by_flux
        launch3(k_b_ct, 24, g, rng(1, nx, 0, ny, 1, nz), 1, 1, -1);
        launch3(k_b_ct, 24, g, rng(1, nx, 1, ny, 0, nz), 2, sy, +1);
    }
    // mass flux on the faces of direction d, its single-dimension BC
    // (dm_x_bcs / dm_y_bcs / dm_z_bcs), then density
    launch3(k_mass_flux, 48, g, rng(0, d == 0 ? nx : nx + 1, 0, d == 1 ? ny : ny + 1, 0, d == 2 ? nz : nz + 1), d);
    bcs(g, kDm, 1, d);
    launch3(k_rho_remap, 40, g, in1, sd);
    for (int e = 1; e <= 2; ++e) {
        launch3(k_energy_flux, 64, g, all0, d, e);
        launch3(k_energy_remap, 56, g, in1, sd, e);
    }
    // vertex-cell quantities
    launch3(k_vertex_rho, 32, g,
            rng(d == 0 ? -1 : 0, d == 0 ? nx + 1 : nx, d == 1 ? -1 : 0, d == 1 ? ny + 1 : ny, d == 2 ? -1 : 0,
                d == 2 ? nz + 1 : nz),
            d);
    launch3(k_vertex_cv2, 16, g, all0);
    launch3(k_vshift, 16, g,
            rng(d == 0 ? -2 : 0, d == 0 ? nx + 1 : nx, d == 1 ? -2 : 0, d == 1 ? ny + 1 : ny, d == 2 ? -2 : 0,
                d == 2 ? nz + 1 : nz),
            d, sd);
    const R3 mom = rng(d == 0 ? -1 : 0, nx, d == 1 ? -1 : 0, ny, d == 2 ? -1 : 0, nz);
    launch3(k_dm_shift, 16, g, mom);
    launch3(k_rho_v1, 40, g, all0, sd);
    for (int comp = 0; comp < 3; ++comp) {
        launch3(k_mom_flux, 64, g, mom, d, comp);
        launch3(k_mom_remap, 56, g, all0, sd, comp);
    }
    launch3(k_wback, 16, g,
            rng(d == 0 ? -2 : 0, d == 0 ? nx + 1 : nx, d == 1 ? -2 : 0, d == 1 ? ny + 1 : ny, d == 2 ? -2 : 0,
                d == 2 ? nz + 1 : nz),
            d);
    boundary_conditions(g);
}

}  // namespace

// The resistive substeps of lagran.cpp:157-170 (int(dt/dtr)+1 of them, two
// when the runner keeps set_dt's dt, with dt_sub = dt/substeps already in
// the device scalars), each eta_calc + resistive_effects (478-505).
void resistive_substeps(Grid &g, int substeps)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const Params &p = g.prm;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    const R3 V = rng(-1, nx + 1, -1, ny + 1, -1, nz + 1);
    for (int i = 0; i < substeps; ++i) {
        launch3(k_eta_calc, 32, g, V, p.j_max, p.eta_background, p.eta0);
        launch3(k_res_copy, 48, g, C);
        launch3(k_rkstep, 64, g, rng(0, nx, 0, ny, 0, nz));
        launch3(k_bstep, 32, g, rng(0, nx, 1, ny, 1, nz), 0);
        launch3(k_bstep, 32, g, rng(1, nx, 0, ny, 1, nz), 1);
        launch3(k_bstep, 32, g, rng(1, nx, 1, ny, 0, nz), 2);
        bfield_bcs(g);
        launch3(k_res_energy, 32, g, rng(1, nx, 1, ny, 1, nz));
        energy_bcs(g);
        // the second rkstep of resistive_effects (lagran.cpp:504) only
        // rewrites flux_x..z and curlb, which nothing reads before the
        // next rkstep recomputes them
    }
}

// LARE3D::lagrangian_step (lagran.cpp:95-173).
void step_start(Grid &g)
{
    step_start_a(g);
    step_start_b(g, 2);
}

// Prologue, shock viscosity and the CFL reduction (lagran.cpp:113-156).
void step_start_a(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const Params &p = g.prm;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    const R3 V = rng(-1, nx + 1, -1, ny + 1, -1, nz + 1);
    launch3(k_prologue1, 96, g, C, p.gas_gamma);
    launch3(k_prologue2, 32, g, V);
    // shock_viscosity (lagran.cpp:175-378)
    launch3(k_cs, 48, g, C, p.gas_gamma);
    launch3(k_cs_v, 32, g, V);
    launch3(k_alpha, 48, g, rng(0, nx + 1, 0, ny + 2, 0, nz + 2), 0, p.visc1, p.visc2_norm);
    launch3(k_alpha, 48, g, rng(-1, nx + 1, 0, ny + 1, 0, nz + 2), 1, p.visc1, p.visc2_norm);
    launch3(k_alpha, 48, g, rng(-1, nx + 1, -1, ny + 1, 0, nz + 1), 2, p.visc1, p.visc2_norm);
    launch3(k_pvisc, 72, g, rng(0, nx + 1, 0, ny + 1, 0, nz + 1));
    launch3(k_fvisc, 80, g, rng(0, nx, 0, ny, 0, nz));
    // controlFunctions::calculateTimestep -> set_dt (lagran.cpp:156, 380-443)
    // the x range starts at 1 outside Cartesian geometry (lagran.cpp:384)
    launch3(k_set_dt, 104, g, rng(p.geometry == 0 ? 0 : 1, nx, 0, ny, 0, nz), p.gas_gamma, p.mu0);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1), 0, g.stream, g.dev, p.dt_multiplier);
}

// Resistive substeps and predictor_step (lagran.cpp:157-173, 592-737).
void step_start_b(Grid &g, int substeps)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const Params &p = g.prm;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    if (p.resistive) resistive_substeps(g, substeps);
    launch3(k_bfield_cv1, 112, g, C);
    launch3(k_energy_pred, 88, g, rng(0, nx + 1, 0, ny + 1, 0, nz + 1), p.gas_gamma);
    launch3(k_forces, 144, g, rng(0, nx, 0, ny, 0, nz), p.mu0, p.geometry);
    remap_v_bcs(g);
}

// LARE3D::corrector_step (lagran.cpp:739-804).  The division of bx1..bz1 by
// cv1 at lagran.cpp:746-749 is not performed: nothing reads bx1..bz1 before
// the next step's prologue overwrites them.  The velocity fills run twice,
// as in the reference (lagran.cpp:759 and 803): a second fill changes the
// result where a staggered line has a periodic lower and a clamped upper
// edge (the lower fill then copies the clamped boundary point).
void step_half(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    launch3(k_shock_heating, 88, g, rng(0, nx + 1, 0, ny + 1, 0, nz + 1));
    launch3(k_vel_update, 104, g, rng(0, nx, 0, ny, 0, nz));
    velocity_bcs(g);
    launch3(k_rho_e_update, 136, g, rng(1, nx, 1, ny, 1, nz));
    energy_bcs(g);
    density_bcs(g);
    velocity_bcs(g);
}

// LARE3D::endOfTimestep (shared_data.h:413-419): eulerian_remap
// (remap.cpp:23-111) with the x, y, z order of case 0.
void step_end(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    launch3(k_b_area, 24, g, rng(-2, nx + 2, -1, ny + 2, -1, nz + 2), 0, 0);
    launch3(k_b_area, 24, g, rng(-1, nx + 2, -2, ny + 2, -1, nz + 2), 1, 0);
    launch3(k_b_area, 24, g, rng(-1, nx + 2, -1, ny + 2, -2, nz + 2), 2, 0);
    remap_sweep(g, 0, 1.0, 1.0, 1.0);
    remap_sweep(g, 1, 0.0, 1.0, 1.0);
    remap_sweep(g, 2, 0.0, 0.0, 1.0);
    launch3(k_b_area, 24, g, rng(-2, nx + 2, -1, ny + 2, -1, nz + 2), 0, 1);
    launch3(k_b_area, 24, g, rng(-1, nx + 2, -2, ny + 2, -1, nz + 2), 1, 1);
    launch3(k_b_area, 24, g, rng(-1, nx + 2, -1, ny + 2, -2, nz + 2), 2, 1);
    bfield_bcs(g);
    hipLaunchKernelGGL(k_step_done, dim3(1), dim3(1), 0, g.stream, g.dev);
}

void box_copy(hipStream_t st, void *box, bool f32, double *lin, const Geo &G, const int lb[3], const int ub[3],
              bool to_box)
{
    const int nx = ub[0] - lb[0] + 1, ny = ub[1] - lb[1] + 1, nz = ub[2] - lb[2] + 1;
    const long total = (long)nx * ny * nz;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    if (to_box) {
        if (f32)
            hipLaunchKernelGGL(k_scatter<float>, dim3(blocks), dim3(256), 0, st, (float *)box, (const double *)lin, G,
                               lb[0], lb[1], lb[2], nx, ny, nz);
        else
            hipLaunchKernelGGL(k_scatter<double>, dim3(blocks), dim3(256), 0, st, (double *)box, (const double *)lin,
                               G, lb[0], lb[1], lb[2], nx, ny, nz);
    } else {
        if (f32)
            hipLaunchKernelGGL(k_gather<float>, dim3(blocks), dim3(256), 0, st, lin, (const float *)box, G, lb[0],
                               lb[1], lb[2], nx, ny, nz);
        else
            hipLaunchKernelGGL(k_gather<double>, dim3(blocks), dim3(256), 0, st, lin, (const double *)box, G, lb[0],
                               lb[1], lb[2], nx, ny, nz);
    }
}

}  // namespace lare

// ================================================================ set-up

// Metric of LARE3D::grid (setup.cpp:337-390), Cartesian, over the extents
// the reference fills: dxab/dyab/dzab over their staggered extents, cv and
// the vertex areas over the centred extent.
// Areas and volumes of LARE3D::grid (setup.cpp:337-390) with the metric
// factors, including its separate y = -2 and z = -2 layers (setup.cpp:356-373:
// dzab's z = -2 layer takes hyc where the others take hy).
KERNEL void k_metric(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const bool cx = ix >= -1, cy = iy >= -1, cz = iz >= -1;
    if (cy && cz) A.dxab[c] = DYB(iy) * DZB(iz) * HY(ix) * HZ1(ix, iy);
    if (cx && cz) A.dyab[c] = DXB(ix) * DZB(iz) * HZ2(ix, iy);
    if (cx && cy) A.dzab[c] = DXB(ix) * DYB(iy) * (iz == -2 ? HYC(ix) : HY(ix));
    if (cx && cy && cz) {
        double dydz = DYB(iy) * DZB(iz);
        A.cv[c] = DXB(ix) * dydz * HYC(ix) * HZC(ix, iy);
        double dz = DZC(iz);
        A.dxac[c] = DYB(iy) * dz * HYC(ix) * HZ2(ix, iy);
        A.dyac[c] = DXB(ix) * dz * HZ1(ix, iy);
        A.dzac[c] = DXB(ix) * DYB(iy) * HY(ix);
    }
}

namespace lare {
void build_metric(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    launch3(k_metric, 0, g, rng(-2, nx + 2, -2, ny + 2, -2, nz + 2));
}
}  // namespace lare
