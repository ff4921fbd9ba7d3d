// Fused remap sweep: the reference's ~33 kernels of one directional sweep
// (xremap.cpp:16-266, yremap.cpp:16-289, zremap.cpp:17-247) merged into
// seven launches wherever no stage reads a neighbour's result of the same
// launch.  Every value is computed by the same expression as the
// one-kernel-per-reference-kernel set in lare_kernels.hip (and so as the
// reference); only where it is stored changes:
//  * the metric (dxab, cv, dxac, ...) is recomputed from the 1D deltas;
//  * rho, energies and velocities are updated in place by the last launch,
//    after every reader of their old values has run, so no rho1 copy is
//    needed;
//  * the shifted sweep velocity and the vertex cv2 are recomputed where used.
// Direction D is a template parameter: strides and stencils are constants.
#include "lare_kernels.h"
#include "lare_math.h"

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
    if (iz > R.z1 || iy > R.y1) return;

template <int DIR>
__device__ __forceinline__ int sdir(const Geo &G)
{
    return DIR == 0 ? G.sx : DIR == 1 ? G.sy : 1;
}

// R1: cell volumes before (cv1) and after (cv2) the remap of DIR and the
// Lagrangian cell width db1 (xremap.cpp:26-70 and twins), centred extent.
template <int DIR>
KERNEL void f_prep(Dev D, R3 R, double xpass, double ypass, double zpass)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1,
              omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = (VX1[ooo] + VX1[omo] + VX1[oom] + VX1[omm]) * 0.25;
    double vxbm = (VX1[moo] + VX1[mmo] + VX1[mom] + VX1[mmm]) * 0.25;
    double vyb = (VY1[ooo] + VY1[moo] + VY1[oom] + VY1[mom]) * 0.25;
    double vybm = (VY1[omo] + VY1[mmo] + VY1[omm] + VY1[mmm]) * 0.25;
    double vzb = (VZ1[ooo] + VZ1[moo] + VZ1[omo] + VZ1[mmo]) * 0.25;
    double vzbm = (VZ1[oom] + VZ1[mom] + VZ1[omm] + VZ1[mmm]) * 0.25;
    const double vol = CV(ix, iy, iz);
    const double axp = DXAB(iy, iz), axm = DXAB(iy, iz);
    const double ayp = DYAB(ix, iz), aym = DYAB(ix, iz);
    const double azp = DZAB(ix, iy), azm = DZAB(ix, iy);
    double dv;
    if (DIR == 0) {
        double dvxdx = (vxb * axp - vxbm * axm) / vol;
        double dvydy = ypass * (vyb * ayp - vybm * aym) / vol;
        double dvzdz = zpass * (vzb * azp - vzbm * azm) / vol;
        dv = (dvydy + dvzdz) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvxdx * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = DXB(ix) + (vxb - vxbm) * dt;
    } else if (DIR == 1) {
        double dvxdx = xpass * (vxb * axp - vxbm * axm) / vol;
        double dvydy = (vyb * ayp - vybm * aym) / vol;
        double dvzdz = zpass * (vzb * azp - vzbm * azm) / vol;
        dv = (dvxdx + dvzdz) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvydy * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = DYB(iy) + (vyb - vybm) * dt;
    } else {
        double dvxdx = xpass * (vxb * axp - vxbm * axm) / vol;
        double dvydy = ypass * (vyb * ayp - vybm * aym) / vol;
        double dvzdz = (vzb * azp - vzbm * azm) / vol;
        dv = (dvxdx + dvydy) * dt;
        A.cv2[c] = vol * (1.0 + dv);
        dv = dv + dvzdz * dt;
        A.cv1[c] = vol * (1.0 + dv);
        A.db1[c] = DZB(iz) + (vzb - vzbm) * dt;
    }
}

// R2: vertex volume before remap, cvc1 = 0.125 * (8 cells of cv1).
KERNEL void f_cvc1(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double *C1 = A.cv1;
    A.cvc1[c] = 0.125 * (C1[c] + C1[c + sx] + C1[c + sy] + C1[c + sx + sy] + C1[c + 1] + C1[c + sx + 1] +
                         C1[c + sy + 1] + C1[c + sx + sy + 1]);
}

// Constrained-transport flux of face magnetic flux B along DIR: v_advect is
// the mean of w at c and c+ow, db1 is averaged with c+tw.
template <int DIR>
__device__ __forceinline__ double ct_flux(const Arr &A, const Geo &G, int c, const double *B, const double *w,
                                          int ow, int tw, double dt)
{
    const int sd = sdir<DIR>(G);
    const double *db1 = A.db1;
    double v_advect = (w[c] + w[c + ow]) * 0.5;
    double o_v = v_advect * dt;
    double dbx = (db1[c] + db1[c + tw]) * 0.5;
    double dbxp = (db1[c + sd] + db1[c + sd + tw]) * 0.5;
    double dbxp2 = (db1[c + 2 * sd] + db1[c + 2 * sd + tw]) * 0.5;
    double dbxm = (db1[c - sd] + db1[c - sd + tw]) * 0.5;
    double fm = B[c - sd] / dbxm;
    double fi = B[c] / dbx;
    double fp = B[c + sd] / dbxp;
    double fp2 = B[c + 2 * sd] / dbxp2;
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
    double phi = fabs(o_v) / dxu;
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    return (fu + Di * (1.0 - phi)) * o_v;
}

// R3: both constrained-transport fluxes of the sweep, over (0..n)^3
// (vx_by_flux + vx_bz_flux, vy_bx_flux + vy_bz_flux, vz_bx_flux + vz_by_flux).
// The second flux reads the B component the first flux's updates leave
// untouched, so both can be computed before any update.
template <int DIR>
KERNEL void f_ct_fluxes(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    if (DIR == 0) {
        A.fl[0][c] = ct_flux<0>(A, G, c, BY, VX1, -1, sy, dt);
        A.fl[1][c] = ct_flux<0>(A, G, c, BZ, VX1, -sy, 1, dt);
    } else if (DIR == 1) {
        A.fl[0][c] = ct_flux<1>(A, G, c, BX, VY1, -1, sx, dt);
        A.fl[1][c] = ct_flux<1>(A, G, c, BZ, VY1, -sx, 1, dt);
    } else {
        A.fl[0][c] = ct_flux<2>(A, G, c, BX, VZ1, -sy, sx, dt);
        A.fl[1][c] = ct_flux<2>(A, G, c, BY, VZ1, -sx, sy, dt);
    }
}

// R4: the four CT updates of the sweep in their reference order, each over
// its own range (xremap.cpp:85-114, yremap.cpp:93-127, zremap.cpp:97-117).
template <int DIR>
KERNEL void f_ct_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const double *fa = A.fl[0], *fb = A.fl[1];
    const bool x1 = ix >= 1 && ix <= nx, x0 = ix <= nx;
    const bool y1 = iy >= 1 && iy <= ny, y0 = iy <= ny;
    const bool z1 = iz >= 1 && iz <= nz, z0 = iz <= nz;
    if (DIR == 0) {
        if (x1 && y0 && z1) BY[c] = BY[c] - fa[c] + fa[c - sx];
        double bx = BX[c];
        const bool bxr = x0 && y1 && z1;
        if (bxr) bx = bx + fa[c] - fa[c - sy];
        if (x1 && y1 && z0) BZ[c] = BZ[c] - fb[c] + fb[c - sx];
        if (bxr) BX[c] = bx + fb[c] - fb[c - 1];
    } else if (DIR == 1) {
        const bool bxr = x0 && y1 && z1;
        if (bxr) BX[c] = BX[c] - fa[c] + fa[c - sy];
        double by = BY[c];
        const bool byr = x1 && y0 && z1;
        if (byr) by = by + fa[c] - fa[c - sx];
        if (x1 && y1 && z0) BZ[c] = BZ[c] - fb[c] + fb[c - sy];
        if (byr) BY[c] = by + fb[c] - fb[c - 1];
    } else {
        if (x0 && y1 && z1) BX[c] = BX[c] - fa[c] + fa[c - 1];
        double bz = BZ[c];
        const bool bzr = x1 && y1 && z0;
        if (bzr) bz = bz + fa[c] - fa[c - sx];
        if (x1 && y0 && z1) BY[c] = BY[c] - fb[c] + fb[c - 1];
        if (bzr) BZ[c] = bz + fb[c] - fb[c - sy];
    }
}

// Face-averaged sweep velocity of the mass and energy fluxes.
template <int DIR>
__device__ __forceinline__ double face_v(const Arr &A, const Geo &G, int c)
{
    const int sx = G.sx, sy = G.sy;
    if (DIR == 0) return (VX1[c] + VX1[c - sy] + VX1[c - 1] + VX1[c - sy - 1]) * 0.25;
    if (DIR == 1) return (VY1[c - sx] + VY1[c] + VY1[c - sx - 1] + VY1[c - 1]) * 0.25;
    return (VZ1[c - sx] + VZ1[c] + VZ1[c - sx - sy] + VZ1[c - sy]) * 0.25;
}

template <int DIR>
__device__ __forceinline__ double face_area(const Arr &A, int ix, int iy, int iz)
{
    if (DIR == 0) return DXAB(iy, iz);
    if (DIR == 1) return DYAB(ix, iz);
    return DZAB(ix, iy);
}

// R5: mass flux dm on the faces of DIR (x_mass_flux, xremap.cpp:373-422 and
// twins) over its reference range; along DIR the range is widened to the
// ghost faces, where dm is set to the zero of the reference's pw::assign(dm,
// 0.0) before the single-dimension boundary fill overwrites them.
template <int DIR>
KERNEL void f_mass_flux(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sd = sdir<DIR>(G);
    const int id = DIR == 0 ? ix : DIR == 1 ? iy : iz;
    const int nd = DIR == 0 ? G.nx : DIR == 1 ? G.ny : G.nz;
    if (id < 0 || id > nd) {
        DM[c] = 0.0;
        return;
    }
    const double dt = A.s->dt;
    double area = face_area<DIR>(A, ix, iy, iz);
    double v_advect = face_v<DIR>(A, G, c);
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
    double phi = fabs(o_v) / dxbu;
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    DM[c] = (fu + Di * (1.0 - phi)) * o_v;
}

// Energy flux of specific energy E on face c (x_energy_flux,
// xremap.cpp:429-483 and twins); rho is still the pre-remap density.
template <int DIR>
__device__ __forceinline__ double energy_flux(const Arr &A, const Geo &G, int c, const double *E, double v_advect,
                                              double o_v)
{
    const int sd = sdir<DIR>(G);
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
    double phi = fabs(o_v) / dxbu;
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double rhou = RHO[c] * vad_p + RHO[c + sd] * vad_m;
    double dm = DM[c];
    double dmu = fabs(dm) / dxbu / rhou;
    return (fu + Di * (1.0 - dmu)) * dm;
}

// R6: both energy fluxes on faces (0..n)^3, the vertex density before remap
// (range -1..n+1 along DIR, 0..n across) and the vertex mass flux dmv
// (range -1..n along DIR, 0..n across).  All read only pre-update values.
template <int DIR>
KERNEL void f_energy_vertex(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int id = DIR == 0 ? ix : DIR == 1 ? iy : iz;
    const int nd = DIR == 0 ? nx : DIR == 1 ? ny : nz;
    const bool face = id >= 0 && id <= nd;
    if (face) {
        const double dt = A.s->dt;
        double area = face_area<DIR>(A, ix, iy, iz);
        double v_advect = face_v<DIR>(A, G, c);
        double o_v = v_advect * dt * area;
        A.fl[0][c] = energy_flux<DIR>(A, G, c, EE, v_advect, o_v);
        A.fl[1][c] = energy_flux<DIR>(A, G, c, EI, v_advect, o_v);
    }
    // vertex density before remap: x divides (s*0.125)/cvc1, y and z
    // multiply by (0.125/cvc1) (xremap.cpp:159-176, yremap.cpp:163-183)
    const double *R1 = RHO, *C1 = A.cv1;
    double s = R1[c] * C1[c] + R1[c + sx] * C1[c + sx] + R1[c + sy] * C1[c + sy] +
               R1[c + sx + sy] * C1[c + sx + sy] + R1[c + 1] * C1[c + 1] + R1[c + sx + 1] * C1[c + sx + 1] +
               R1[c + sy + 1] * C1[c + sy + 1] + R1[c + sx + sy + 1] * C1[c + sx + sy + 1];
    if (DIR == 0)
        A.rrho_v[c] = s * 0.125 / A.cvc1[c];
    else
        A.rrho_v[c] = s * (0.125 / A.cvc1[c]);
    if (id <= nd) {
        const double *M = DM;
        double q = M[c] + M[c + sx] + M[c + sy] + M[c + sx + sy] + M[c + 1] + M[c + sx + 1] + M[c + sy + 1] +
                   M[c + sx + sy + 1];
        A.dmv[c] = DIR == 0 ? q * 0.125 : 0.125 * q;
    }
}

// Momentum flux of velocity component f on vertex face c (x_mom_flux,
// xremap.cpp:485-566 and twins), multiplied by dm.  z takes dyac as its area
// (zremap.cpp:475) and divides by (dzbu*rhou) (zremap.cpp:505).
template <int DIR>
__device__ __forceinline__ double mom_flux(const Arr &A, const Geo &G, int c, const double *f, double o_v,
                                           double sign_v, double dxci, double dxcu, double dxbu, double dmu_base,
                                           double rhou, double dm)
{
    const int sd = sdir<DIR>(G);
    (void)dmu_base;
    double fm = f[c - sd], fi = f[c], fp = f[c + sd], fp2 = f[c + 2 * sd];
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double phi = fabs(o_v) / dxbu;
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double dmu = DIR == 2 ? fabs(dm) / (dxbu * rhou) : fabs(dm) / dxbu / rhou;
    double fl = fu + Di * (1.0 - dmu);
    (void)o_v;
    return fl * dm;
}

// R7: the three momentum fluxes over (-1..n along DIR, 0..n across).
template <int DIR>
KERNEL void f_mom_fluxes(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sd = sdir<DIR>(G);
    const double dt = A.s->dt;
    const double *w = DIR == 0 ? VX1 : DIR == 1 ? VY1 : VZ1;
    double area = DIR == 0 ? DXAC(iy, iz) : DYAC(ix, iz);
    double v_advect = (w[c] + w[c + sd]) * 0.5;
    double o_v = v_advect * dt * area;
    double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double dxci = A.cv1[c + sd];
    double dxcu = A.cv1[c] * vad_p + A.cv1[c + 2 * sd] * vad_m;
    double dxbu = A.cvc1[c] * vad_p + A.cvc1[c + sd] * vad_m;
    double rhou = A.rrho_v[c] * vad_p + A.rrho_v[c + sd] * vad_m;
    double dm = A.dmv[c];
    A.fl[2][c] = mom_flux<DIR>(A, G, c, VX, o_v, sign_v, dxci, dxcu, dxbu, 0.0, rhou, dm);
    A.fl[3][c] = mom_flux<DIR>(A, G, c, VY, o_v, sign_v, dxci, dxcu, dxbu, 0.0, rhou, dm);
    A.fl[4][c] = mom_flux<DIR>(A, G, c, VZ, o_v, sign_v, dxci, dxcu, dxbu, 0.0, rhou, dm);
}

// R8: the in-place updates over (0..n)^3: density and both energies on cells
// 1..n (xremap.cpp:121-154), vertex density after remap and the three
// velocities on vertices 0..n (xremap.cpp:180-262).
template <int DIR>
KERNEL void f_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int sd = sdir<DIR>(G);
    const double *C2 = A.cv2;
    // vertex cv2 (read before any cv2 could change: cv2 is not written here)
    double s2 = C2[c] + C2[c + sx] + C2[c + sy] + C2[c + sx + sy] + C2[c + 1] + C2[c + sx + 1] + C2[c + sy + 1] +
                C2[c + sx + sy + 1];
    const double cv2v = DIR == 0 ? s2 * 0.125 : 0.125 * s2;
    const double cvc1 = A.cvc1[c];
    const double rv = A.rrho_v[c];
    const double rho_v1 = (rv * cvc1 + A.dmv[c - sd] - A.dmv[c]) / cv2v;
    const double den = cv2v * rho_v1;
    VX[c] = (rv * VX[c] * cvc1 + A.fl[2][c - sd] - A.fl[2][c]) / den;
    VY[c] = (rv * VY[c] * cvc1 + A.fl[3][c - sd] - A.fl[3][c]) / den;
    VZ[c] = (rv * VZ[c] * cvc1 + A.fl[4][c - sd] - A.fl[4][c]) / den;
    if (ix >= 1 && iy >= 1 && iz >= 1) {
        const double rho1 = RHO[c], cv1 = A.cv1[c], cv2 = C2[c];
        const double rho = (rho1 * cv1 + DM[c - sd] - DM[c]) / cv2;
        RHO[c] = rho;
        EE[c] = (EE[c] * cv1 * rho1 + A.fl[0][c - sd] - A.fl[0][c]) / (cv2 * rho);
        EI[c] = (EI[c] * cv1 * rho1 + A.fl[1][c - sd] - A.fl[1][c]) / (cv2 * rho);
    }
}

namespace lare {

namespace {

template <typename... KA, typename... Args>
void flaunch_(void (*k)(Dev, R3, KA...), const char *name, double bpc, Grid &g, R3 r, Args... args)
{
    const int nz = r.z1 - r.z0 + 1, ny = r.y1 - r.y0 + 1, nx = r.x1 - r.x0 + 1;
    if (nz <= 0 || ny <= 0 || nx <= 0) return;
    dim3 blk(64, 4, 1);
    dim3 grd((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nx);
    Prof::Rec rec{name, bpc * (double)nz * ny * nx, nullptr, nullptr};
    if (g.prof.on && g.prof.dir >= 0) {
        const size_t at = rec.name.find("<DIR>");
        if (at != std::string::npos) rec.name.replace(at, 5, "<" + std::to_string(g.prof.dir) + ">");
    }
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
#define flaunch(k, bpc, ...) flaunch_(k, #k, bpc, __VA_ARGS__)

inline R3 rng(int x0, int x1, int y0, int y1, int z0, int z1) { return R3{x0, x1, y0, y1, z0, z1}; }

template <int DIR>
void fused_sweep_t(Grid &g, double xpass, double ypass, double zpass)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    const R3 all0 = rng(0, nx, 0, ny, 0, nz);
    g.prof.dir = DIR;
    // bytes per cell: distinct arrays read and written, 8 bytes each
    flaunch(f_prep<DIR>, 48, g, C, xpass, ypass, zpass);
    flaunch(f_cvc1, 16, g, rng(-1, nx + 1, -1, ny + 1, -1, nz + 1));
    flaunch(f_ct_fluxes<DIR>, 72, g, all0);
    flaunch(f_ct_update<DIR>, 64, g, rng(0, nx, 0, ny, 0, nz));
    // mass flux: its reference range, widened to the ghost faces along DIR
    flaunch(f_mass_flux<DIR>, 40, g,
            rng(DIR == 0 ? -1 : 0, DIR == 0 ? nx + 2 : nx + 1, DIR == 1 ? -1 : 0, DIR == 1 ? ny + 2 : ny + 1,
                DIR == 2 ? -1 : 0, DIR == 2 ? nz + 2 : nz + 1));
    dm_bcs(g, DIR);
    flaunch(f_energy_vertex<DIR>, 88, g,
            rng(DIR == 0 ? -1 : 0, DIR == 0 ? nx + 1 : nx, DIR == 1 ? -1 : 0, DIR == 1 ? ny + 1 : ny,
                DIR == 2 ? -1 : 0, DIR == 2 ? nz + 1 : nz));
    flaunch(f_mom_fluxes<DIR>, 88, g, rng(DIR == 0 ? -1 : 0, nx, DIR == 1 ? -1 : 0, ny, DIR == 2 ? -1 : 0, nz));
    flaunch(f_update<DIR>, 184, g, all0);
    g.prof.dir = -1;
    sweep_bcs(g);
}

}  // namespace

void fused_sweep(Grid &g, int d, double xpass, double ypass, double zpass)
{
    if (d == 0)
        fused_sweep_t<0>(g, xpass, ypass, zpass);
    else if (d == 1)
        fused_sweep_t<1>(g, xpass, ypass, zpass);
    else
        fused_sweep_t<2>(g, xpass, ypass, zpass);
}

}  // namespace lare

// ======================================================= fused Lagrangian
// lagrangian_step + corrector_step (lagran.cpp:95-173, 592-804, 806-993) in
// nine launches.  Quantities the reference stores and reads back on the
// same cell (bx1..bz1 as face averages, p_e, p_i, pressure, cv, cv_v, the
// metric) are recomputed with the same expression where they are read; a
// value stored for a neighbour's use stays stored.  With the resistive path
// on, the predictor must see the prologue's p_e and the face-copied bx1..bz1
// (lagran.cpp:482), so those are stored then (template flag RES).

#define PE_OF(c) ((gas_gamma - 1.0) * RHO[c] * EE[c])
#define PI_OF(c) ((gas_gamma - 1.0) * RHO[c] * EI[c])

// L1 (centred extent): sound speed (lagran.cpp:190-197) from the prologue
// values of lagran.cpp:113-125; with RES, also stores p_e (the predictor
// reads the prologue value after the resistive heating changed e_e).
template <int RES>
KERNEL void g_cs(Dev D, R3 R, double gas_gamma)
{
    KBEGIN
    if (ix == R.x0 && iy == R.y0 && iz == R.z0) A.s->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
    const int c = IX(ix, iy, iz);
    double bx1 = 0.5 * (BX[c] + BX[c - G.sx]);
    double by1 = 0.5 * (BY[c] + BY[c - G.sy]);
    double bz1 = 0.5 * (BZ[c] + BZ[c - 1]);
    double pe = PE_OF(c), pi = PI_OF(c);
    double p = pe + pi;
    if (RES) A.p_e[c] = pe;
    double rmin = pmax(RHO[c], DBL_EPSILON);
    double b2 = bx1 * bx1 + by1 * by1 + bz1 * bz1;
    A.cs[c] = sqrt((gas_gamma * p + b2) / rmin);
}

#define CV8SUM(ix, iy, iz)                                                                               \
    (CV(ix, iy, iz) + CV(ix + 1, iy, iz) + CV(ix, iy + 1, iz) + CV(ix + 1, iy + 1, iz) + CV(ix, iy, iz + 1) + \
     CV(ix + 1, iy, iz + 1) + CV(ix, iy + 1, iz + 1) + CV(ix + 1, iy + 1, iz + 1))

// L2 (-1..n+1)^3: vertex density rho_v (lagran.cpp:128-153) and the vertex
// sound speed cs_v (lagran.cpp:199-216).
KERNEL void g_vertex(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double c000 = CV(ix, iy, iz), c100 = CV(ix + 1, iy, iz), c010 = CV(ix, iy + 1, iz),
                 c110 = CV(ix + 1, iy + 1, iz), c001 = CV(ix, iy, iz + 1), c101 = CV(ix + 1, iy, iz + 1),
                 c011 = CV(ix, iy + 1, iz + 1), c111 = CV(ix + 1, iy + 1, iz + 1);
    const double *r = RHO, *s = A.cs;
    double sum_rho_cv = r[c] * c000 + r[c + sx] * c100 + r[c + sy] * c010 + r[c + sx + sy] * c110 + r[c + 1] * c001 +
                        r[c + sx + 1] * c101 + r[c + sy + 1] * c011 + r[c + sx + sy + 1] * c111;
    double sum_cv = c000 + c100 + c010 + c110 + c001 + c101 + c011 + c111;
    A.rho_v[c] = sum_rho_cv / sum_cv;
    double cv_v = 0.125 * sum_cv;
    double sum = s[c] * c000 + s[c + sx] * c100 + s[c + sy] * c010 + s[c + sx + sy] * c110 + s[c + 1] * c001 +
                 s[c + sx + 1] * c101 + s[c + sy + 1] * c011 + s[c + sx + sy + 1] * c111;
    A.cs_v[c] = 0.125 * sum / cv_v;
}

// edge_viscosity (lagran.cpp:56-93) between c1 and c1 + s.
__device__ __forceinline__ double g_edge_visc(const Arr &A, double dt, double visc1, double visc2_norm,
                                              double dvdots, double dx, double dxm, double dxp, double cs_edge,
                                              int c1, int s)
{
    const int c0 = c1 - s, c2 = c1 + s, c3 = c2 + s;
    dvdots = pmin(0.0, dvdots);
    double rho_edge = 2.0 * A.rho_v[c1] * A.rho_v[c2] / (A.rho_v[c1] + A.rho_v[c2]);
    double dvx = VX[c1] - VX[c2];
    double dvy = VY[c1] - VY[c2];
    double dvz = VZ[c1] - VZ[c2];
    double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    double dv = sqrt(dv2);
    double psi = 0.0;
    const bool small = dv * dt / dx < 1.e-14;
    dvdots = small ? 0.0 : dvdots / dv;
    double dvxm = VX[c0] - VX[c1];
    double dvxp = VX[c2] - VX[c3];
    double dvym = VY[c0] - VY[c1];
    double dvyp = VY[c2] - VY[c3];
    double dvzm = VZ[c0] - VZ[c1];
    double dvzp = VZ[c2] - VZ[c3];
    double rl = small ? 1.0 : (dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);
    double rr = small ? 1.0 : (dvxm * dvx + dvym * dvy + dvzm * dvz) * dx / (dxm * dv2);
    psi = pmin_list4(0.5 * (rr + rl), 2.0 * rl, 2.0 * rr, 1.0);
    psi = pmax(0.0, psi);
    double q_k_bar = rho_edge * (visc2_norm * dv + sqrt(visc2_norm * visc2_norm * dv2 + (visc1 * cs_edge) * (visc1 * cs_edge)));
    return q_k_bar * (1.0 - psi) * dvdots;
}

// L3: alpha1, alpha2, alpha3 (lagran.cpp:211-274), each over its own range
// inside the union (-1..n+1, -1..n+2, 0..n+2).  dt is the previous step's.
KERNEL void g_alpha(Dev D, R3 R, double visc1, double visc2_norm)
{
    KBEGIN
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const double dt = A.s->dt;
    const int c = IX(ix, iy, iz);
    if (ix >= 0 && ix <= nx + 1 && iy >= 0 && iy <= ny + 2 && iz >= 0 && iz <= nz + 2) {
        const int c1 = c - sx - sy - 1;
        double dvdots = -(VX[c1] - VX[c1 + sx]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + sx]);
        A.alpha1[c] = g_edge_visc(A, dt, visc1, visc2_norm, dvdots, DXB(ix), DXB(ix - 1), DXB(ix + 1), cs_edge, c1, sx);
    }
    if (ix <= nx + 1 && iy >= 0 && iy <= ny + 1 && iz >= 0 && iz <= nz + 2) {
        const int c1 = c - sy - 1;
        double dvdots = -(VY[c1] - VY[c1 + sy]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + sy]);
        A.alpha2[c] = g_edge_visc(A, dt, visc1, visc2_norm, dvdots, DYB(iy), DYB(iy - 1), DYB(iy + 1), cs_edge, c1, sy);
    }
    if (ix <= nx + 1 && iy <= ny + 1 && iz >= 0 && iz <= nz + 1) {
        const int c1 = c - 1;
        double dvdots = -(VZ[c1] - VZ[c1 + 1]);
        double cs_edge = pmin(A.cs_v[c1], A.cs_v[c1 + 1]);
        A.alpha3[c] = g_edge_visc(A, dt, visc1, visc2_norm, dvdots, DZB(iz), DZB(iz - 1), DZB(iz + 1), cs_edge, c1, 1);
    }
}

#define SQD(p, q) \
    ((VX[p] - VX[q]) * (VX[p] - VX[q]) + (VY[p] - VY[q]) * (VY[p] - VY[q]) + (VZ[p] - VZ[q]) * (VZ[p] - VZ[q]))

// L4 over (0..n+1)^3: p_visc and visc_heat (lagran.cpp:279-323); on (0..n)^3
// also the viscous forces (330-376), stored as fx_visc, and the CFL minimum
// (set_dt, 380-443) reduced per wavefront, per block and by atomicMin.
KERNEL void g_visc_dt(Dev D, R3 R, double gas_gamma, double mu0)
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
        const int mmm = c - sx - sy - 1, omm = c - sy - 1, oom = c - 1, mom = c - sx - 1;
        const int mmo = c - sx - sy, omo = c - sy, ooo = c, moo = c - sx;
        double a1 = SQD(mmm, omm), a2 = SQD(omm, oom), a3 = SQD(oom, mom), a4 = SQD(mom, mmm);
        double a5 = SQD(mmo, omo), a6 = SQD(omo, ooo), a7 = SQD(ooo, moo), a8 = SQD(moo, mmo);
        double a9 = SQD(oom, ooo), a10 = SQD(mom, moo), a11 = SQD(mmm, mmo), a12 = SQD(omm, omo);
        double pv = 0.0;
        pv = pmax(pv, -A.alpha1[c] * sqrt(a1));
        pv = pmax(pv, -A.alpha2[c] * sqrt(a2));
        pv = pmax(pv, -A.alpha3[c] * sqrt(a9));
        double dx = DXB(ix), dy = DYB(iy), dz = DZB(iz);
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
        const double cv = CV(ix, iy, iz);
        A.visc_heat[c] = vh / cv;
        if (ix <= G.nx && iy <= G.ny && iz <= G.nz) {
            // viscous forces
            double b1 = A.alpha1[c + sy + 1] * dy * dz;
            double b2 = A.alpha1[c + sx + sy + 1] * dy * dz;
            double b3 = A.alpha2[c + 1] * dx * dz;
            double b4 = A.alpha2[c + sy + 1] * dx * dz;
            double b5 = A.alpha3[c] * dx * dy;
            double b6 = A.alpha3[c + 1] * dx * dy;
            double cvv = 0.125 * CV8SUM(ix, iy, iz);
#define FV(V)                                                                                   \
    ((b1 * (V[c] - V[c - sx]) + b2 * (V[c] - V[c + sx]) + b3 * (V[c] - V[c - sy]) +            \
      b4 * (V[c] - V[c + sy]) + b5 * (V[c] - V[c - 1]) + b6 * (V[c] - V[c + 1])) / cvv)
            A.fxv[c] = FV(VX);
            A.fyv[c] = FV(VY);
            A.fzv[c] = FV(VZ);
#undef FV
            // set_dt
            double rho0 = pmax(RHO[c], DBL_EPSILON);
            double press = PE_OF(c) + PI_OF(c);
            double cs2 = gas_gamma * press / rho0;
            double w1 = (BX[c] * BX[c] + BY[c] * BY[c] + BZ[c] * BZ[c]) / mu0 / rho0;
            double c_visc2 = pv / rho0;
            double length = pmin(pmin(dx, dy), dz);
            double dt1 = length / (sqrt(c_visc2) + sqrt(cs2 + w1 + c_visc2));
            double ax = 0.25 * DXAB(iy, iz);
            double ay = 0.25 * DYAB(ix, iz);
            double az = 0.25 * DZAB(ix, iy);
            double vxbm = (VX[c - sx] + VX[c - sx - sy] + VX[c - sx - 1] + VX[c - sx - sy - 1]) * ax;
            double vxbp = (VX[c] + VX[c - sy] + VX[c - 1] + VX[c - sy - 1]) * ax;
            double vybm = (VY[c - sy] + VY[c - sx - sy] + VY[c - sy - 1] + VY[c - sx - sy - 1]) * ay;
            double vybp = (VY[c] + VY[c - sx] + VY[c - 1] + VY[c - sx - 1]) * ay;
            double vzbm = (VZ[c - 1] + VZ[c - sx - 1] + VZ[c - sy - 1] + VZ[c - sx - sy - 1]) * az;
            double vzbp = (VZ[c] + VZ[c - sx] + VZ[c - sy] + VZ[c - sx - sy]) * az;
            double dvx = fabs(vxbp - vxbm), dvy = fabs(vybp - vybm), dvz = fabs(vzbp - vzbm);
            double dt2 = cv / pmax(pmax(pmax(fabs(vxbm), fabs(vxbp)), dvx), 1.0e-10 * cv);
            double dt3 = cv / pmax(pmax(pmax(fabs(vybm), fabs(vybp)), dvy), 1.0e-10 * cv);
            double dt4 = cv / pmax(pmax(pmax(fabs(vzbm), fabs(vzbp)), dvz), 1.0e-10 * cv);
            m = pmin(pmin(pmin(dt1, dt2), dt3), dt4);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = pmin(m, __shfl_xor(m, off, 64));
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

// L6 over the centred extent: b_field_and_cv1_update (lagran.cpp:806-882)
// with bx1 *= cv1 (598-601); on (0..n+1)^3 also the energy/pressure
// predictor (604-614), storing the predicted p_e and p_i.
template <int RES>
KERNEL void g_predict(Dev D, R3 R, double gas_gamma)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1,
              omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = 0.25 * (VX[ooo] + VX[omo] + VX[oom] + VX[omm]);
    double vxbm = 0.25 * (VX[moo] + VX[mmo] + VX[mom] + VX[mmm]);
    double vyb = 0.25 * (VY[ooo] + VY[moo] + VY[oom] + VY[mom]);
    double vybm = 0.25 * (VY[omo] + VY[mmo] + VY[omm] + VY[mmm]);
    double vzb = 0.25 * (VZ[ooo] + VZ[moo] + VZ[omo] + VZ[mmo]);
    double vzbm = 0.25 * (VZ[oom] + VZ[mom] + VZ[omm] + VZ[mmm]);
    const double vol = CV(ix, iy, iz);
    const double axx = DXAB(iy, iz), ayy = DYAB(ix, iz), azz = DZAB(ix, iy);
    double dvxdx = (vxb * axx - vxbm * axx) / vol;
    double dvydy = (vyb * ayy - vybm * ayy) / vol;
    double dvzdz = (vzb * azz - vzbm * azz) / vol;
    double dv = (dvxdx + dvydy + dvzdz) * dt / 2.0;
    const double cv1 = vol * (1.0 + dv);
    A.cv1[c] = cv1;
    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[oom] + VX[mom]);
    vxbm = 0.25 * (VX[omo] + VX[mmo] + VX[omm] + VX[mmm]);
    vyb = 0.25 * (VY[ooo] + VY[omo] + VY[oom] + VY[omm]);
    vybm = 0.25 * (VY[moo] + VY[mmo] + VY[mom] + VY[mmm]);
    double dvxdy = (vxb * ayy - vxbm * ayy) / vol;
    double dvydx = (vyb * axx - vybm * axx) / vol;
    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[omo] + VX[mmo]);
    vxbm = 0.25 * (VX[oom] + VX[mom] + VX[omm] + VX[mmm]);
    double vzb2 = 0.25 * (VZ[ooo] + VZ[omo] + VZ[oom] + VZ[omm]);
    double vzbm2 = 0.25 * (VZ[moo] + VZ[mmo] + VZ[mom] + VZ[mmm]);
    double dvxdz = (vxb * azz - vxbm * azz) / vol;
    double dvzdx = (vzb2 * axx - vzbm2 * axx) / vol;
    double vyb3 = 0.25 * (VY[ooo] + VY[moo] + VY[omo] + VY[mmo]);
    double vybm3 = 0.25 * (VY[oom] + VY[mom] + VY[omm] + VY[mmm]);
    double vzb3 = 0.25 * (VZ[ooo] + VZ[moo] + VZ[oom] + VZ[mom]);
    double vzbm3 = 0.25 * (VZ[omo] + VZ[mmo] + VZ[omm] + VZ[mmm]);
    double dvydz = (vyb3 * azz - vybm3 * azz) / vol;
    double dvzdy = (vzb3 * ayy - vzbm3 * ayy) / vol;
    double bx1, by1, bz1;
    if (RES) {
        bx1 = A.bx1[c];
        by1 = A.by1[c];
        bz1 = A.bz1[c];
    } else {
        bx1 = 0.5 * (BX[c] + BX[c - sx]);
        by1 = 0.5 * (BY[c] + BY[c - sy]);
        bz1 = 0.5 * (BZ[c] + BZ[c - 1]);
    }
    double w3 = bx1 * dvxdx + by1 * dvxdy + bz1 * dvxdz;
    double w4 = bx1 * dvydx + by1 * dvydy + bz1 * dvydz;
    double w5 = bx1 * dvzdx + by1 * dvzdy + bz1 * dvzdz;
    A.bx1[c] = (bx1 + w3 * dt / 2.0) / (1.0 + dv) * cv1;
    A.by1[c] = (by1 + w4 * dt / 2.0) / (1.0 + dv) * cv1;
    A.bz1[c] = (bz1 + w5 * dt / 2.0) / (1.0 + dv) * cv1;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix <= G.nx + 1 && iy <= G.ny + 1 && iz <= G.nz + 1) {
        double rho = RHO[c];
        double dvp = cv1 / vol - 1.0;
        double pe0 = RES ? A.p_e[c] : PE_OF(c);
        double pi0 = PI_OF(c);
        double e1_e = EE[c] - pe0 * dvp / rho;
        double e1_i = EI[c] - pi0 * dvp / rho;
        e1_i += A.visc_heat[c] * dt / 2.0 / rho;
        A.p_e[c] = e1_e * (gas_gamma - 1.0) * rho * vol / cv1;
        A.p_i[c] = e1_i * (gas_gamma - 1.0) * rho * vol / cv1;
    }
}

// L7 over (0..n)^3: forces and half-step velocities (lagran.cpp:619-731);
// fx_visc + fx is stored back into fx_visc for the corrector.
KERNEL void g_forces(Dev D, R3 R, double mu0)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const double *PE = A.p_e, *PI = A.p_i, *C1 = A.cv1, *B1x = A.bx1, *B1y = A.by1, *B1z = A.bz1;
    double dx = DXC(ix), dy = DYC(iy), dz = DZC(iz);
    double dhy = dy, dhz = dz;
    const int x_ = c + sx, y_ = c + sy, z_ = c + 1, xy = c + sx + sy, xz = c + sx + 1, yz = c + sy + 1,
              xyz = c + sx + sy + 1;
    double pp = PE[c] + PI[c], ppx = PE[x_] + PI[x_], ppy = PE[y_] + PI[y_], ppz = PE[z_] + PI[z_];
    double ppxy = PE[xy] + PI[xy], ppxz = PE[xz] + PI[xz], ppyz = PE[yz] + PI[yz], ppxyz = PE[xyz] + PI[xyz];
    double w1 = pp + ppy + ppz + ppyz;
    double w2 = ppx + ppxy + ppxz + ppxyz;
    double fx = (w1 - w2) * 0.25 / dx;
    w1 = pp + ppx + ppz + ppxz;
    w2 = ppy + ppxy + ppyz + ppxyz;
    double fy = (w1 - w2) * 0.25 / dhy;
    w1 = pp + ppx + ppy + ppxy;
    w2 = ppz + ppxz + ppyz + ppxyz;
    double fz = (w1 - w2) * 0.25 / dhz;
    double cvx = C1[c] + C1[y_] + C1[z_] + C1[yz];
    double cvxp = C1[x_] + C1[xy] + C1[xz] + C1[xyz];
    double cvy = C1[c] + C1[x_] + C1[z_] + C1[xz];
    double cvyp = C1[y_] + C1[xy] + C1[yz] + C1[xyz];
    double cvz = C1[c] + C1[x_] + C1[y_] + C1[xy];
    double cvzp = C1[z_] + C1[xz] + C1[yz] + C1[xyz];
    w1 = (B1z[c] + B1z[x_] + B1z[z_] + B1z[xz]) / cvy;
    w2 = (B1z[y_] + B1z[xy] + B1z[yz] + B1z[xyz]) / cvyp;
    double jx = (w2 - w1) / dhy;
    w1 = (B1y[c] + B1y[x_] + B1y[y_] + B1y[xy]) / cvz;
    w2 = (B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz]) / cvzp;
    jx -= (w2 - w1) / dhz;
    w1 = (B1z[c] + B1z[y_] + B1z[z_] + B1z[yz]) / cvx;
    w2 = (B1z[x_] + B1z[xy] + B1z[xz] + B1z[xyz]) / cvxp;
    double jy = -(w2 - w1) / dx;
    w1 = (B1x[c] + B1x[x_] + B1x[y_] + B1x[xy]) / cvz;
    w2 = (B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz]) / cvzp;
    jy += (w2 - w1) / dhz;
    w1 = (B1y[c] + B1y[y_] + B1y[z_] + B1y[yz]) / cvx;
    w2 = (B1y[x_] + B1y[xy] + B1y[xz] + B1y[xyz]) / cvxp;
    double jz = (w2 - w1) / dx;
    w1 = (B1x[c] + B1x[x_] + B1x[z_] + B1x[xz]) / cvy;
    w2 = (B1x[y_] + B1x[xy] + B1x[yz] + B1x[xyz]) / cvyp;
    jz -= (w2 - w1) / dhy;
    double bxv = (B1x[c] + B1x[x_] + B1x[y_] + B1x[xy] + B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz]) / (cvx + cvxp);
    double byv = (B1y[c] + B1y[x_] + B1y[y_] + B1y[xy] + B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz]) / (cvx + cvxp);
    double bzv = (B1z[c] + B1z[x_] + B1z[y_] + B1z[xy] + B1z[z_] + B1z[xz] + B1z[yz] + B1z[xyz]) / (cvx + cvxp);
    fx += (jy * bzv - jz * byv) / mu0;
    fy += (jz * bxv - jx * bzv) / mu0;
    fz += (jx * byv - jy * bxv) / mu0;
    double rv = A.rho_v[c];
    double sxv = A.fxv[c] + fx, syv = A.fyv[c] + fy, szv = A.fzv[c] + fz;
    A.fxv[c] = sxv;
    A.fyv[c] = syv;
    A.fzv[c] = szv;
    VX1[c] = VX[c] + dt / 2.0 * sxv / rv;
    VY1[c] = VY[c] + dt / 2.0 * syv / rv;
    VZ1[c] = VZ[c] + dt / 2.0 * szv / rv;
}

#define DOT(p, q)                                                                                   \
    (1 * ((VX[p] - VX[q]) * (VX1[p] - VX1[q]) + (VY[p] - VY[q]) * (VY1[p] - VY1[q]) +               \
          (VZ[p] - VZ[q]) * (VZ1[p] - VZ1[q])))

// C1 over (0..n+1)^3: shock_heating (lagran.cpp:884-993) with the dy*dy
// factors of a9 and a11 (lagran.cpp:987).
KERNEL void g_shock_heating(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int ooo = c, moo = c - sx, omo = c - sy, mmo = c - sx - sy;
    const int oop = c + 1, mop = c - sx + 1, omp = c - sy + 1, mmp = c - sx - sy + 1;
    const int oom = c - 1, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double a1 = DOT(mmo, omo), a2 = DOT(omo, ooo), a3 = DOT(ooo, moo), a4 = DOT(moo, mmo);
    double a5 = DOT(mmp, omp), a6 = DOT(omp, oop), a7 = DOT(oop, mop), a8 = DOT(mop, mmp);
    double a9 = DOT(oom, ooo), a10 = DOT(mom, moo), a11 = DOT(mmm, mmo), a12 = DOT(omm, omo);
    double dx = DXB(ix), dy = DYB(iy), dz = DZB(iz);
    double vh =
        (-0.25 * dy * dz * A.alpha1[c] * a1 - 0.25 * dx * dz * A.alpha2[c] * a2 -
         0.25 * dy * dz * A.alpha1[c + sy] * a3 - 0.25 * dx * dz * A.alpha2[c - sx] * a4 -
         0.25 * dy * dz * A.alpha1[c + 1] * a5 - 0.25 * dx * dz * A.alpha2[c + 1] * a6 -
         0.25 * dy * dz * A.alpha1[c + sy + 1] * a7 - 0.25 * dx * dz * A.alpha2[c - sx + 1] * a8 -
         0.25 * dy * dy * A.alpha3[c] * a9 - 0.25 * dx * dy * A.alpha3[c - sx] * a10 -
         0.25 * dy * dy * A.alpha3[c - sx - sy] * a11 - 0.25 * dx * dy * A.alpha3[c - sy] * a12);
    A.visc_heat[c] = pmax(vh / CV(ix, iy, iz), 0.0);
}

// C2 over (0..n)^3: velocity correction (lagran.cpp:754-758) with the sum
// fx_visc + fx stored by g_forces.
KERNEL void g_vel_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const double dt = A.s->dt;
    double rv = A.rho_v[c];
    VX[c] += dt * A.fxv[c] / rv;
    VY[c] += dt * A.fyv[c] / rv;
    VZ[c] += dt * A.fzv[c] / rv;
}

// C3 over (1..n)^3: density and energy correction (lagran.cpp:762-796).  The
// cv1 it also writes in the reference is recomputed by the remap before use.
KERNEL void g_rho_e_update(Dev D, R3 R)
{
    KBEGIN
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = A.s->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1,
              omm = c - sy - 1, mmm = c - sx - sy - 1;
    double vxb = 0.25 * (VX1[ooo] + VX1[omo] + VX1[oom] + VX1[omm]);
    double vxbm = 0.25 * (VX1[moo] + VX1[mmo] + VX1[mom] + VX1[mmm]);
    double vyb = 0.25 * (VY1[ooo] + VY1[moo] + VY1[oom] + VY1[mom]);
    double vybm = 0.25 * (VY1[omo] + VY1[mmo] + VY1[omm] + VY1[mmm]);
    double vzb = 0.25 * (VZ1[ooo] + VZ1[moo] + VZ1[omo] + VZ1[mmo]);
    double vzbm = 0.25 * (VZ1[oom] + VZ1[mom] + VZ1[omm] + VZ1[mmm]);
    const double vol = CV(ix, iy, iz);
    const double axx = DXAB(iy, iz), ayy = DYAB(ix, iz), azz = DZAB(ix, iy);
    double dvxdx = (vxb * axx - vxbm * axx) / vol;
    double dvydy = (vyb * ayy - vybm * ayy) / vol;
    double dvzdz = (vzb * azz - vzbm * azz) / vol;
    double dv = (dvxdx + dvydy + dvzdz) * dt;
    double rho = RHO[c];
    EE[c] -= dv * A.p_e[c] / rho;
    EI[c] += (dt * A.visc_heat[c] - dv * A.p_i[c]) / rho;
    RHO[c] = rho / (1.0 + dv);
}

namespace lare {

void fused_start_a(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const Params &p = g.prm;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    const R3 V = rng(-1, nx + 1, -1, ny + 1, -1, nz + 1);
    if (p.resistive)
        flaunch(g_cs<1>, 56, g, C, p.gas_gamma);
    else
        flaunch(g_cs<0>, 56, g, C, p.gas_gamma);
    flaunch(g_vertex, 32, g, V);
    flaunch(g_alpha, 64, g, rng(-1, nx + 1, -1, ny + 2, 0, nz + 2), p.visc1, p.visc2_norm);
    flaunch(g_visc_dt, 128, g, rng(0, nx + 1, 0, ny + 1, 0, nz + 1), p.gas_gamma, p.mu0);
    dt_final(g);
}

void fused_start_b(Grid &g, int substeps)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    const Params &p = g.prm;
    const R3 C = rng(-1, nx + 2, -1, ny + 2, -1, nz + 2);
    if (p.resistive) resistive_substeps(g, substeps);
    if (p.resistive)
        flaunch(g_predict<1>, 152, g, C, p.gas_gamma);
    else
        flaunch(g_predict<0>, 128, g, C, p.gas_gamma);
    flaunch(g_forces, 152, g, rng(0, nx, 0, ny, 0, nz), p.mu0);
    remap_v_fill(g);
}

void fused_half(Grid &g)
{
    const int nx = g.dev.g.nx, ny = g.dev.g.ny, nz = g.dev.g.nz;
    flaunch(g_shock_heating, 80, g, rng(0, nx + 1, 0, ny + 1, 0, nz + 1));
    flaunch(g_vel_update, 64, g, rng(0, nx, 0, ny, 0, nz));
    velocity_fill(g);
    flaunch(g_rho_e_update, 96, g, rng(1, nx, 1, ny, 1, nz));
    energy_density_fill(g);
}

}  // namespace lare
