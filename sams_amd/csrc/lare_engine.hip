// Grid-group engine kernels: the LARE3D step (lagran.cpp, remap.cpp,
// xremap.cpp, yremap.cpp, zremap.cpp) as 21 launch shapes, each covering
// every grid of a group.
//
// Work decomposition.  A launch has one 256-lane workgroup per 256 cells of
// one x-plane of one grid's range: lanes run along the flattened (y, z)
// plane, z fastest, so a wavefront reads 64 consecutive doubles of a row
// (or the tail of one row and the head of the next).  Workgroup b of the
// launch belongs to the grid whose first block is the largest boff <= b.
// The hardware hands workgroups to the 8 XCDs round-robin; xcd_swizzle
// renumbers them so that each XCD walks one contiguous run of x-planes and
// the planes a stencil shares stay in that XCD's L2.
//
// Arithmetic.  Every value is computed by the expression, association and
// operand order of the reference (and of the one-kernel-per-reference-kernel
// set in lare_kernels.hip); -ffp-contract=off.  In FP32 storage the same
// expressions run in FP64 on values widened from FP32.
//
// Fusion (what differs from the reference's kernel boundaries; none changes
// a value):
//  * Lagrangian phase in 9 launches: quantities the reference stores and
//    reads back on the same cell (bx1..bz1 face averages, p_e, p_i,
//    pressure, cv, cv_v, the metric) are recomputed where read;
//  * each remap sweep in 7 launches plus its boundary fills: rho, energies
//    and velocities are updated in place by the last launch after every
//    reader of their old values has run; the shifted sweep velocity and the
//    vertex cv2 are recomputed where used;
//  * the metric (dxab, cv, dxac, ...) is always the product of the 1D
//    deltas, as LARE3D::grid stores it (setup.cpp:337-390, Cartesian h = 1).
#include "lare_engine.h"
#include "lare_num.h"

#include <float.h>

#include <algorithm>
#include <mutex>
#include <cstdlib>
#include <cstring>

using lare::ArrE;
using lare::EDev;
using lare::Fld;
using lare::Geo;
using lare::Phys;
using lare::R3;
using lare::Scal;
using lare::Src;

// Lower bound on resident waves per SIMD for the register-heavy kernels (the
// compiler then keeps them within the matching register budget; the bounds
// are the highest that compile without scratch spills).
#define OCC(n) __attribute__((amdgpu_waves_per_eu(n, 8)))

#define RHO A.rho
#define EE A.ee
#define EI A.ei
#define VX A.vx
#define VY A.vy
#define VZ A.vz
#define BX A.bx
#define BY A.by
#define BZ A.bz
#define VX1 A.vx1
#define VY1 A.vy1
#define VZ1 A.vz1
#define DM lare::gbl(A.dm)

// Workgroup b in dispatch order -> logical workgroup: XCD b % 8 takes the
// contiguous logical run [start(x), start(x) + count(x)).
__device__ __forceinline__ int xcd_swizzle(int b, int nblk)
{
    const int per = nblk >> 3, rem = nblk & 7;
    const int x = b & 7, k = b >> 3;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}

// The arrays with the state copies named for this launch: rho .. vz the
// copy holding the current state (the second copy when bit 0 of st is set),
// rho2 .. vz2 the other one; with bit 1 set, bx .. bz name bx1 .. bz1 (B's
// second copy during the remap) and bx1 .. bz1 name B's own arrays; bits 2
// and 3 name the y sweep's cv1, cv2, db1 (alpha1..3, from f_prep_all) or
// the z sweep's (alpha2, the cell volume, fzv); bit 4 (applied after bit 1)
// names B's third copy bxy .. bzy as bx1 .. bz1.
template <class T>
__device__ __forceinline__ ArrE<T> state_view(const ArrE<T> &a, int st)
{
    ArrE<T> v = a;
    if (st & 12) {  // prep set of the y (4) or z (8) sweep
        const bool y = (st & 12) == 4;
        v.cv1 = y ? a.alpha1 : a.alpha2;
        v.cv2 = y ? a.alpha2 : a.vol;
        v.db1 = y ? a.alpha3 : a.fzv;
    }
    if (st & 2) {
        v.bx = a.bx1;
        v.by = a.by1;
        v.bz = a.bz1;
        v.bx1 = a.bx;
        v.by1 = a.by;
        v.bz1 = a.bz;
    }
    if (st & 16) {  // B's third copy, written by the fused y sweep
        v.bx1 = a.bxy;
        v.by1 = a.byy;
        v.bz1 = a.bzy;
    }
    if (st & 1) {
        v.rho = a.rho2;
        v.ee = a.ee2;
        v.ei = a.ei2;
        v.vx = a.vx2;
        v.vy = a.vy2;
        v.vz = a.vz2;
        v.rho2 = a.rho;
        v.ee2 = a.ee;
        v.ei2 = a.ei;
        v.vx2 = a.vx;
        v.vy2 = a.vy;
        v.vz2 = a.vz;
    }
    return v;
}

// Largest g with boff[g] <= b (boff ascending, boff[0] = 0).
__device__ __forceinline__ int grid_of(const int *boff, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Work of this workgroup: grid g_ (found from the first-block offsets) and
// one chunk of 256 cells of the flattened (y, z) plane px_ of range R (lane
// -> (iy_, iz_), z fastest; ok_ is false past the end of the plane).  Order
// within a grid (S.order): 0 chunk fastest, then plane; 1 the chunks of a
// plane split into 8 bands, band-major, so that with the XCD renumbering
// each XCD walks one band of every plane and the x- and y-neighbours a
// stencil reads stay in its L2.
#define EGRID                                                                                        \
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);                                           \
    const int g_ = grid_of(S.boff, S.ngrid, b_);                                                   \
    const EDev<T> &D_ = S.devs[g_];                                                                \
    const Geo G = D_.g;                                                                            \
    const ArrE<T> A = state_view(D_.a, S.st);                                                      \
    const Phys P = D_.p;                                                                           \
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};         \
    const int nzr_ = R.z1 - R.z0 + 1;                                                              \
    const int plane_ = (R.y1 - R.y0 + 1) * nzr_;                                                   \
    const int nxr_ = R.x1 - R.x0 + 1;                                                              \
    const int lb_ = b_ - S.boff[g_];                                                               \
    const int bpp_ = (plane_ + 255) >> 8;                                                          \
    int px_, ch_;                                                                                  \
    if (S.order) {                                                                                 \
        const int bs_ = (bpp_ + 7) >> 3;                                                           \
        const int band_ = lb_ / (nxr_ * bs_);                                                      \
        const int rem_ = lb_ - band_ * (nxr_ * bs_);                                               \
        px_ = rem_ / bs_;                                                                          \
        ch_ = band_ * bs_ + (rem_ - px_ * bs_);                                                     \
    } else {                                                                                       \
        px_ = lb_ / bpp_;                                                                          \
        ch_ = lb_ - px_ * bpp_;                                                                    \
    }                                                                                              \
    const int t_ = ch_ * 256 + (int)threadIdx.x;                                                   \
    const bool ok_ = ch_ < bpp_ && t_ < plane_;                                                    \
    const int ix_ = R.x0 + px_;                                                                    \
    const int iy_ = R.y0 + t_ / nzr_;                                                              \
    const int iz_ = R.z0 + t_ % nzr_;                                                              \
    (void)P;
// Run cell(ix, iy, iz) for the cell of this lane.
#define ECELL(cell) \
    if (ok_) cell(ix_, iy_, iz_);

template <int DIR>
__device__ __forceinline__ int sdir(const Geo &G)
{
    return DIR == 0 ? G.sx : DIR == 1 ? G.sy : 1;
}

// =================================================================== remap

// R1 at one cell from the eight corners (x-1..x, y-1..y, z-1..z) of v1:
// V[k][p] is component k at corner p (P_OOO ... P_MMM); out = cv1, cv2, db1.
enum { P_OOO, P_MOO, P_OMO, P_OOM, P_MMO, P_MOM, P_OMM, P_MMM };
__device__ __forceinline__ void prep_avg(const double V[3][8], double avg[6]);
template <int DIR, class T>
__device__ __forceinline__ void prep_out(const ArrE<T> &A, int ix, int iy, int iz, const double avg[6], double dt,
                                         double xpass, double ypass, double zpass, double out[3]);

template <int DIR, class T>
__device__ __forceinline__ void prep_v(const ArrE<T> &A, int ix, int iy, int iz, const double V[3][8], double dt,
                                       double xpass, double ypass, double zpass, double out[3])
{
    double avg[6];
    prep_avg(V, avg);
    prep_out<DIR>(A, ix, iy, iz, avg, dt, xpass, ypass, zpass, out);
}

// The six face-averaged v1 of R1 (vxb, vxbm, vyb, vybm, vzb, vzbm).
__device__ __forceinline__ void prep_avg(const double V[3][8], double avg[6])
{
    const double *X = V[0], *Y = V[1], *Zv = V[2];
    avg[0] = (X[P_OOO] + X[P_OMO] + X[P_OOM] + X[P_OMM]) * 0.25;
    avg[1] = (X[P_MOO] + X[P_MMO] + X[P_MOM] + X[P_MMM]) * 0.25;
    avg[2] = (Y[P_OOO] + Y[P_MOO] + Y[P_OOM] + Y[P_MOM]) * 0.25;
    avg[3] = (Y[P_OMO] + Y[P_MMO] + Y[P_OMM] + Y[P_MMM]) * 0.25;
    avg[4] = (Zv[P_OOO] + Zv[P_MOO] + Zv[P_OMO] + Zv[P_MMO]) * 0.25;
    avg[5] = (Zv[P_OOM] + Zv[P_MOM] + Zv[P_OMM] + Zv[P_MMM]) * 0.25;
}

// R1's outputs of sweep DIR from the face averages.
template <int DIR, class T>
__device__ __forceinline__ void prep_out(const ArrE<T> &A, int ix, int iy, int iz, const double avg[6], double dt,
                                         double xpass, double ypass, double zpass, double out[3])
{
    const double vxb = avg[0], vxbm = avg[1], vyb = avg[2], vybm = avg[3], vzb = avg[4], vzbm = avg[5];
    const double vol = CV(ix, iy, iz);
    const double axp = DXAB(iy, iz), axm = DXAB(iy, iz);
    const double ayp = DYAB(ix, iz), aym = DYAB(ix, iz);
    const double azp = DZAB(ix, iy), azm = DZAB(ix, iy);
    double dv;
    if (DIR == 0) {
        double dvxdx = ddiv(vxb * axp - vxbm * axm, vol);
        double dvydy = ddiv(ypass * (vyb * ayp - vybm * aym), vol);
        double dvzdz = ddiv(zpass * (vzb * azp - vzbm * azm), vol);
        dv = (dvydy + dvzdz) * dt;
        out[1] = vol * (1.0 + dv);
        dv = dv + dvxdx * dt;
        out[0] = vol * (1.0 + dv);
        out[2] = DXB(ix) + (vxb - vxbm) * dt;
    } else if (DIR == 1) {
        double dvxdx = ddiv(xpass * (vxb * axp - vxbm * axm), vol);
        double dvydy = ddiv(vyb * ayp - vybm * aym, vol);
        double dvzdz = ddiv(zpass * (vzb * azp - vzbm * azm), vol);
        dv = (dvxdx + dvzdz) * dt;
        out[1] = vol * (1.0 + dv);
        dv = dv + dvydy * dt;
        out[0] = vol * (1.0 + dv);
        out[2] = DYB(iy) + (vyb - vybm) * dt;
    } else {
        double dvxdx = ddiv(xpass * (vxb * axp - vxbm * axm), vol);
        double dvydy = ddiv(ypass * (vyb * ayp - vybm * aym), vol);
        double dvzdz = ddiv(vzb * azp - vzbm * azm, vol);
        dv = (dvxdx + dvydy) * dt;
        out[1] = vol * (1.0 + dv);
        dv = dv + dvzdz * dt;
        out[0] = vol * (1.0 + dv);
        out[2] = DZB(iz) + (vzb - vzbm) * dt;
    }
}

// R1: cell volumes before (cv1) and after (cv2) the remap of DIR and the
// Lagrangian cell width db1 (xremap.cpp:26-70 and twins), centred extent.
template <class T, int DIR>
KERNEL void f_prep(const Src<T> S, double xpass, double ypass, double zpass)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
    const int ooo = c, moo = c - sx, omo = c - sy, oom = c - 1, mmo = c - sx - sy, mom = c - sx - 1,
              omm = c - sy - 1, mmm = c - sx - sy - 1;
    const int pos[8] = {ooo, moo, omo, oom, mmo, mom, omm, mmm};
    double V[3][8];
    for (int p = 0; p < 8; ++p) {
        V[0][p] = VX1[pos[p]];
        V[1][p] = VY1[pos[p]];
        V[2][p] = VZ1[pos[p]];
    }
    double out[3];
    prep_v<DIR>(A, ix, iy, iz, V, dt, xpass, ypass, zpass, out);
    A.cv1[c] = out[0];
    A.cv2[c] = out[1];
    A.db1[c] = out[2];
    };
    ECELL(cell)
}

// R1 of all three sweeps at once, in the x sweep: each sweep's prep reads
// the same v1 face averages (the only v1 a sweep writes back is its own
// component, which later sweeps' preps weight by a pass of 0: xpass in y and
// z, ypass in z), with the passes of the x, y and z sweeps (1,1,1), (0,1,1),
// (0,0,1).  The y sweep's cv1, cv2, db1 go into alpha1..3 and the z sweep's
// db1 into fzv, which nothing reads between the corrector and the next
// step's viscosity; their kernels see them through the prep-set bits of the
// view.  With the z passes the x and y divergences are signed zeros, so the
// z sweep's cv2 is vol * (1 + 0) = vol, the cell volume, and its cv1,
// vol * (1 + dvzdz * dt), is the y sweep's cv2 with the same operands
// (zpass = 1 multiplies exactly, and a signed zero added to dvzdz leaves
// 1 + dv unchanged): neither is stored, the view names the volume array and
// alpha2 for them.
template <class T>
KERNEL void f_prep_all(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
    const int pos[8] = {c, c - sx, c - sy, c - 1, c - sx - sy, c - sx - 1, c - sy - 1, c - sx - sy - 1};
    double V[3][8];
    for (int p = 0; p < 8; ++p) {
        V[0][p] = VX1[pos[p]];
        V[1][p] = VY1[pos[p]];
        V[2][p] = VZ1[pos[p]];
    }
    double avg[6], o[3];
    prep_avg(V, avg);
    prep_out<0>(A, ix, iy, iz, avg, dt, 1.0, 1.0, 1.0, o);
    A.cv1[c] = o[0];
    A.cv2[c] = o[1];
    A.db1[c] = o[2];
    prep_out<1>(A, ix, iy, iz, avg, dt, 0.0, 1.0, 1.0, o);
    A.alpha1[c] = o[0];
    A.alpha2[c] = o[1];
    A.alpha3[c] = o[2];
    prep_out<2>(A, ix, iy, iz, avg, dt, 0.0, 0.0, 1.0, o);
    A.fzv[c] = o[2];
    };
    ECELL(cell)
}

// R2: vertex volume before remap, cvc1 = 0.125 * (8 cells of cv1).
template <class T>
KERNEL void f_cvc1(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const Fld<T> C1 = A.cv1;
    A.cvc1[c] = 0.125 * (C1[c] + C1[c + sx] + C1[c + sy] + C1[c + sx + sy] + C1[c + 1] + C1[c + sx + 1] +
                         C1[c + sy + 1] + C1[c + sx + sy + 1]);
    };
    ECELL(cell)
}

// Face area of B component SC (remap.cpp:46-50): dxab, dyab, dzab.
template <int SC, class T>
__device__ __forceinline__ double b_area(const ArrE<T> &A, int ix, int iy, int iz)
{
    return SC == 0 ? DXAB(iy, iz) : SC == 1 ? DYAB(ix, iz) : DZAB(ix, iy);
}

// The constrained-transport flux from the values along DIR: f = B / db at
// c-1..c+2, the averaged widths db at c and c+1, the advection velocity, and
// cvc1 at c-1..c+1, cv1 at c, c+1.
__device__ __forceinline__ double ct_flux_v(double fm, double fi, double fp, double fp2, double dbx, double dbxp,
                                            double v_advect, double dt, double cvc_m, double cvc_0, double cvc_p,
                                            double cv1_0, double cv1_p)
{
    double o_v = v_advect * dt;
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = cvc_0;
    double dxcu = cvc_m * vad_p + cvc_p * vad_m;
    double dxbu = cv1_0 * vad_p + cv1_p * vad_m;
    double dxu = dbx * vad_p + dbxp * vad_m;
    double phi = ddiv(fabs(o_v), dxu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    return (fu + Di * (1.0 - phi)) * o_v;
}

// Constrained-transport flux of face magnetic flux B along DIR: v_advect is
// the mean of w at c and c+ow, db1 is averaged with c+tw (vx_by_flux,
// xremap.cpp:269-319, and its five twins).
// SC >= 0: B is read unscaled and multiplied here by the area of
// component SC at each position, as remap.cpp:46-50 stores it (the x sweep's
// flux kernel runs before that product is stored).
template <int DIR, int SC = -1, class T>
__device__ __forceinline__ double ct_flux(const ArrE<T> &A, const Geo &G, int c, const Fld<T> Bf, const Fld<T> w,
                                          int ow, int tw, double dt, int ix = 0, int iy = 0, int iz = 0)
{
    const int sd = sdir<DIR>(G);
    const Fld<T> db1 = A.db1;
    auto B = [&](int k) -> double {
        double b = Bf[c + k * sd];
        if (SC >= 0) b = b * b_area<SC>(A, ix + (DIR == 0 ? k : 0), iy + (DIR == 1 ? k : 0), iz + (DIR == 2 ? k : 0));
        return b;
    };
    double v_advect = (w[c] + w[c + ow]) * 0.5;
    double dbx = (db1[c] + db1[c + tw]) * 0.5;
    double dbxp = (db1[c + sd] + db1[c + sd + tw]) * 0.5;
    double dbxp2 = (db1[c + 2 * sd] + db1[c + 2 * sd + tw]) * 0.5;
    double dbxm = (db1[c - sd] + db1[c - sd + tw]) * 0.5;
    double fm = ddiv(B(-1), dbxm);
    double fi = ddiv(B(0), dbx);
    double fp = ddiv(B(1), dbxp);
    double fp2 = ddiv(B(2), dbxp2);
    return ct_flux_v(fm, fi, fp, fp2, dbx, dbxp, v_advect, dt, A.cvc1[c - sd], A.cvc1[c], A.cvc1[c + sd], A.cv1[c],
                     A.cv1[c + sd]);
}

// R3: both constrained-transport fluxes of the sweep, over (0..n)^3
// (vx_by_flux + vx_bz_flux, vy_bx_flux + vy_bz_flux, vz_bx_flux + vz_by_flux).
// The second flux reads the B component the first flux's updates leave
// untouched, so both can be computed before any update.
template <class T, int DIR>
KERNEL void f_ct_fluxes(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
    if (DIR == 0) {
        lare::gbl(A.fl[0])[c] = ct_flux<0, 1>(A, G, c, BY, VX1, -1, sy, dt, ix, iy, iz);
        lare::gbl(A.fl[1])[c] = ct_flux<0, 2>(A, G, c, BZ, VX1, -sy, 1, dt, ix, iy, iz);
    } else if (DIR == 1) {
        lare::gbl(A.fl[0])[c] = ct_flux<1>(A, G, c, BX, VY1, -1, sx, dt);
        lare::gbl(A.fl[1])[c] = ct_flux<1>(A, G, c, BZ, VY1, -sx, 1, dt);
    } else {
        lare::gbl(A.fl[0])[c] = ct_flux<2>(A, G, c, BX, VZ1, -sy, sx, dt);
        lare::gbl(A.fl[1])[c] = ct_flux<2>(A, G, c, BY, VZ1, -sx, sy, dt);
    }
    };
    ECELL(cell)
}

// R4: the four CT updates of the sweep in their reference order, each over
// its own range (xremap.cpp:85-114, yremap.cpp:93-127, zremap.cpp:97-117).
// The x sweep's kernel runs over the union of the B extents and also stores
// B *= area (remap.cpp:46-50) on every cell of each component's extent; the
// z sweep's stores B /= (area + none_zero) (remap.cpp:91-95) likewise.  No
// kernel reads B between those products and these stores.
template <class T, int DIR>
KERNEL void f_ct_update(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const auto fa = lare::gbl(A.fl[0]), fb = lare::gbl(A.fl[1]);
    const bool x1 = ix >= 1 && ix <= nx, x0 = ix >= 0 && ix <= nx;
    const bool y1 = iy >= 1 && iy <= ny, y0 = iy >= 0 && iy <= ny;
    const bool z1 = iz >= 1 && iz <= nz, z0 = iz >= 0 && iz <= nz;
    // extents: bx (-2.., -1.., -1..), by (-1.., -2.., -1..), bz (-1.., -1.., -2..)
    const bool ebx = iy >= -1 && iz >= -1, eby = ix >= -1 && iz >= -1, ebz = ix >= -1 && iy >= -1;
    if (DIR == 0) {
        double bx = ebx ? BX[c] * b_area<0>(A, ix, iy, iz) : 0.0;
        double by = eby ? BY[c] * b_area<1>(A, ix, iy, iz) : 0.0;
        double bz = ebz ? BZ[c] * b_area<2>(A, ix, iy, iz) : 0.0;
        if (x1 && y0 && z1) by = by - fa[c] + fa[c - sx];
        const bool bxr = x0 && y1 && z1;
        if (bxr) bx = bx + fa[c] - fa[c - sy];
        if (x1 && y1 && z0) bz = bz - fb[c] + fb[c - sx];
        if (bxr) bx = bx + fb[c] - fb[c - 1];
        if (ebx) BX[c] = bx;
        if (eby) BY[c] = by;
        if (ebz) BZ[c] = bz;
    } else if (DIR == 1) {
        const bool bxr = x0 && y1 && z1;
        if (bxr) BX[c] = BX[c] - fa[c] + fa[c - sy];
        double by = BY[c];
        const bool byr = x1 && y0 && z1;
        if (byr) by = by + fa[c] - fa[c - sx];
        if (x1 && y1 && z0) BZ[c] = BZ[c] - fb[c] + fb[c - sy];
        if (byr) BY[c] = by + fb[c] - fb[c - 1];
    } else {
        double bx = ebx ? (double)BX[c] : 0.0;
        double by = eby ? (double)BY[c] : 0.0;
        double bz = ebz ? (double)BZ[c] : 0.0;
        if (x0 && y1 && z1) bx = bx - fa[c] + fa[c - 1];
        const bool bzr = x1 && y1 && z0;
        if (bzr) bz = bz + fa[c] - fa[c - sx];
        if (x1 && y0 && z1) by = by - fb[c] + fb[c - 1];
        if (bzr) bz = bz + fb[c] - fb[c - sy];
        if (ebx) BX[c] = ddiv(bx, b_area<0>(A, ix, iy, iz) + DBL_EPSILON);
        if (eby) BY[c] = ddiv(by, b_area<1>(A, ix, iy, iz) + DBL_EPSILON);
        if (ebz) BZ[c] = ddiv(bz, b_area<2>(A, ix, iy, iz) + DBL_EPSILON);
    }
    };
    ECELL(cell)
}

// Face-averaged sweep velocity of the mass and energy fluxes.
template <int DIR, class T>
__device__ __forceinline__ double face_v(const ArrE<T> &A, const Geo &G, int c)
{
    const int sx = G.sx, sy = G.sy;
    if (DIR == 0) return (VX1[c] + VX1[c - sy] + VX1[c - 1] + VX1[c - sy - 1]) * 0.25;
    if (DIR == 1) return (VY1[c - sx] + VY1[c] + VY1[c - sx - 1] + VY1[c - 1]) * 0.25;
    return (VZ1[c - sx] + VZ1[c] + VZ1[c - sx - sy] + VZ1[c - sy]) * 0.25;
}

template <int DIR, class T>
__device__ __forceinline__ double face_area(const ArrE<T> &A, int ix, int iy, int iz)
{
    if (DIR == 0) return DXAB(iy, iz);
    if (DIR == 1) return DYAB(ix, iz);
    return DZAB(ix, iy);
}

// Mass flux through a face from the values along DIR: rho at c-1..c+2, the
// face velocity and o_v, cvc1 at c-1..c+1, cv1 at c, c+1.
__device__ __forceinline__ double mass_flux_v(double fm, double fi, double fp, double fp2, double v_advect,
                                              double o_v, double cvc_m, double cvc_0, double cvc_p, double cv1_0,
                                              double cv1_p)
{
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = cvc_0;
    double dxcu = cvc_m * vad_p + cvc_p * vad_m;
    double dxbu = cv1_0 * vad_p + cv1_p * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    return (fu + Di * (1.0 - phi)) * o_v;
}

// R5: mass flux dm on the faces of DIR (x_mass_flux, xremap.cpp:373-422 and
// twins) over its reference range; along DIR the range is widened to the
// ghost faces, where dm is set to the zero of the reference's pw::assign(dm,
// 0.0) before the single-dimension boundary fill overwrites them.
template <class T, int DIR>
KERNEL void f_mass_flux(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sd = sdir<DIR>(G);
    const int id = DIR == 0 ? ix : DIR == 1 ? iy : iz;
    const int nd = DIR == 0 ? G.nx : DIR == 1 ? G.ny : G.nz;
    if (id < 0 || id > nd) {
        DM[c] = 0.0;
        return;
    }
    const double dt = lare::gbl(A.s)->dt;
    double area = face_area<DIR>(A, ix, iy, iz);
    double v_advect = face_v<DIR>(A, G, c);
    double o_v = v_advect * dt * area;
    DM[c] = mass_flux_v(RHO[c - sd], RHO[c], RHO[c + sd], RHO[c + 2 * sd], v_advect, o_v, A.cvc1[c - sd], A.cvc1[c],
                        A.cvc1[c + sd], A.cv1[c], A.cv1[c + sd]);
    };
    ECELL(cell)
}

// Energy flux of specific energy E on face c (x_energy_flux,
// xremap.cpp:429-483 and twins); rho is still the pre-remap density.
// The same from the values along DIR: E at c-sd..c+2sd, cvc1 at c-sd..c+sd,
// cv1 and rho at c, c+sd, and the face's mass flux dm.
__device__ __forceinline__ double energy_flux_v(double fm, double fi, double fp, double fp2, double v_advect,
                                                double o_v, double cvc_m, double cvc_0, double cvc_p, double cv1_0,
                                                double cv1_p, double rho_0, double rho_p, double dm)
{
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double dxci = cvc_0;
    double dxcu = cvc_m * vad_p + cvc_p * vad_m;
    double dxbu = cv1_0 * vad_p + cv1_p * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double rhou = rho_0 * vad_p + rho_p * vad_m;
    double dmu = ddiv(ddiv(fabs(dm), dxbu), rhou);
    return (fu + Di * (1.0 - dmu)) * dm;
}

template <int DIR, class T>
__device__ __forceinline__ double energy_flux(const ArrE<T> &A, const Geo &G, int c, const Fld<T> E,
                                              double v_advect, double o_v)
{
    const int sd = sdir<DIR>(G);
    return energy_flux_v(E[c - sd], E[c], E[c + sd], E[c + 2 * sd], v_advect, o_v, A.cvc1[c - sd], A.cvc1[c],
                         A.cvc1[c + sd], A.cv1[c], A.cv1[c + sd], RHO[c], RHO[c + sd], DM[c]);
}

// Energy flux of E through cell face f = (fx, fy, fz) along DIR with the
// face's own sweep velocity and area (f_energy: the same v_advect and o_v as
// the mass flux of that face).
template <int DIR, class T>
__device__ __forceinline__ double energy_flux_face(const ArrE<T> &A, const Geo &G, int f, int fx, int fy, int fz,
                                                   const Fld<T> E, double dt)
{
    double area = face_area<DIR>(A, fx, fy, fz);
    double v_advect = face_v<DIR>(A, G, f);
    double o_v = v_advect * dt * area;
    return energy_flux<DIR>(A, G, f, E, v_advect, o_v);
}

// Momentum flux of velocity component f on vertex face c (x_mom_flux,
// xremap.cpp:485-566 and twins), multiplied by dm.  z takes dyac as its area
// (zremap.cpp:475) and divides by (dzbu*rhou) (zremap.cpp:505).
// mom_fl_v is the flux before its multiplication by dm (the value the
// kinetic-energy correction reads, xremap.cpp:534-558).
template <int DIR>
__device__ __forceinline__ double mom_fl_v(double fm, double fi, double fp, double fp2, double o_v, double sign_v,
                                           double dxci, double dxcu, double dxbu, double rhou, double dm)
{
    double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double fu = fi * vad_p + fp * vad_m;
    double dfu = dfm * vad_p + dfp * vad_m;
    double phi = ddiv(fabs(o_v), dxbu);
    double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
    double dmu = DIR == 2 ? ddiv(fabs(dm), dxbu * rhou) : ddiv(ddiv(fabs(dm), dxbu), rhou);
    return fu + Di * (1.0 - dmu);
}
template <int DIR>
__device__ __forceinline__ double mom_flux_v(double fm, double fi, double fp, double fp2, double o_v, double sign_v,
                                             double dxci, double dxcu, double dxbu, double rhou, double dm)
{
    return mom_fl_v<DIR>(fm, fi, fp, fp2, o_v, sign_v, dxci, dxcu, dxbu, rhou, dm) * dm;
}

template <int DIR, class T>
__device__ __forceinline__ double mom_flux(const Geo &G, int c, const Fld<T> f, double o_v, double sign_v,
                                           double dxci, double dxcu, double dxbu, double rhou, double dm)
{
    const int sd = sdir<DIR>(G);
    return mom_flux_v<DIR>(f[c - sd], f[c], f[c + sd], f[c + 2 * sd], o_v, sign_v, dxci, dxcu, dxbu, rhou, dm);
}

// Vertex density before remap at vertex c (xremap.cpp:159-176,
// yremap.cpp:163-183, zremap.cpp:160-177): x divides (s*0.125)/cvc1, y and z
// multiply by (0.125/cvc1).
template <int DIR, class T>
__device__ __forceinline__ double rrho_v_at(const ArrE<T> &A, const Geo &G, int c)
{
    const int sx = G.sx, sy = G.sy;
    const Fld<T> R1 = RHO, C1 = A.cv1;
    double s = R1[c] * C1[c] + R1[c + sx] * C1[c + sx] + R1[c + sy] * C1[c + sy] +
               R1[c + sx + sy] * C1[c + sx + sy] + R1[c + 1] * C1[c + 1] + R1[c + sx + 1] * C1[c + sx + 1] +
               R1[c + sy + 1] * C1[c + sy + 1] + R1[c + sx + sy + 1] * C1[c + sx + sy + 1];
    if (DIR == 0) return ddiv(s * 0.125, A.cvc1[c]);
    return s * ddiv(0.125, A.cvc1[c]);
}

// Vertex mass flux at vertex c (xremap.cpp:209-219 and twins).
template <int DIR, class T>
__device__ __forceinline__ double dmv_at(const ArrE<T> &A, const Geo &G, int c)
{
    const int sx = G.sx, sy = G.sy;
    const auto M = DM;
    double q = M[c] + M[c + sx] + M[c + sy] + M[c + sx + sy] + M[c + 1] + M[c + sx + 1] + M[c + sy + 1] +
               M[c + sx + sy + 1];
    return DIR == 0 ? q * 0.125 : 0.125 * q;
}

// The three momentum fluxes through vertex face f = (fx, fy, fz) along DIR
// (x_mom_flux, xremap.cpp:485-566 and twins), given the vertex densities
// at f and f + sd.  z takes dyac as its area (zremap.cpp:475).
template <int DIR, class T, bool TIMES_DM = true>
__device__ __forceinline__ void mom_fluxes_at(const ArrE<T> &A, const Geo &G, int f, int fx, int fy, int fz,
                                              double rv0, double rv1, double dm, double dt, double out[3])
{
    const int sd = sdir<DIR>(G);
    const Fld<T> w = DIR == 0 ? VX1 : DIR == 1 ? VY1 : VZ1;
    double area = DIR == 0 ? DXAC(fy, fz) : DYAC(fx, fz);
    double v_advect = (w[f] + w[f + sd]) * 0.5;
    double o_v = v_advect * dt * area;
    double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
    double vad_p = (sign_v + 1.0) * 0.5;
    double vad_m = 1.0 - vad_p;
    double dxci = A.cv1[f + sd];
    double dxcu = A.cv1[f] * vad_p + A.cv1[f + 2 * sd] * vad_m;
    double dxbu = A.cvc1[f] * vad_p + A.cvc1[f + sd] * vad_m;
    double rhou = rv0 * vad_p + rv1 * vad_m;
    const Fld<T> V3[3] = {VX, VY, VZ};
    for (int k = 0; k < 3; ++k) {
        const Fld<T> v = V3[k];
        const double fl = mom_fl_v<DIR>(v[f - sd], v[f], v[f + sd], v[f + 2 * sd], o_v, sign_v, dxci, dxcu, dxbu,
                                        rhou, dm);
        out[k] = TIMES_DM ? fl * dm : fl;
    }
}

// R6 over (0..n)^3: the whole transport of the sweep from the current state
// copy into the other.  Each lane computes the fluxes of both faces of its
// cell along DIR itself (the reference stores them and the next kernel
// reads two of them; the values are the same expressions on the same
// inputs): on cells 1..n the mass (dm, stored), energy fluxes and the
// density and energy update (xremap.cpp:121-154 and x_energy_flux
// 429-483); on vertices 0..n the vertex densities, vertex mass fluxes,
// momentum fluxes and the velocity update (xremap.cpp:159-262, x_mom_flux
// 485-566).  Nothing here is written where anything is read, so no lane
// sees another's new values.
template <class T, int DIR>
KERNEL void f_update_all(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int sd = sdir<DIR>(G);
    const double dt = lare::gbl(A.s)->dt;
    const int mx = DIR == 0 ? ix - 1 : ix, my = DIR == 1 ? iy - 1 : iy, mz = DIR == 2 ? iz - 1 : iz;
    // vertex quantities at c - sd, c, c + sd
    const double rvm = rrho_v_at<DIR>(A, G, c - sd), rv = rrho_v_at<DIR>(A, G, c), rvp = rrho_v_at<DIR>(A, G, c + sd);
    const double dmvm = dmv_at<DIR>(A, G, c - sd), dmv = dmv_at<DIR>(A, G, c);
    double flm[3], flc[3];
    mom_fluxes_at<DIR>(A, G, c - sd, mx, my, mz, rvm, rv, dmvm, dt, flm);
    mom_fluxes_at<DIR>(A, G, c, ix, iy, iz, rv, rvp, dmv, dt, flc);
    const Fld<T> C2 = A.cv2;
    // vertex cv2
    double s2 = C2[c] + C2[c + sx] + C2[c + sy] + C2[c + sx + sy] + C2[c + 1] + C2[c + sx + 1] + C2[c + sy + 1] +
                C2[c + sx + sy + 1];
    const double cv2v = DIR == 0 ? s2 * 0.125 : 0.125 * s2;
    const double cvc1 = A.cvc1[c];
    const double rho_v1 = ddiv(rv * cvc1 + dmvm - dmv, cv2v);
    const double den = cv2v * rho_v1;
    A.vx2[c] = ddiv(rv * VX[c] * cvc1 + flm[0] - flc[0], den);
    A.vy2[c] = ddiv(rv * VY[c] * cvc1 + flm[1] - flc[1], den);
    A.vz2[c] = ddiv(rv * VZ[c] * cvc1 + flm[2] - flc[2], den);
    if (ix >= 1 && iy >= 1 && iz >= 1) {
        const double em[2] = {energy_flux_face<DIR>(A, G, c - sd, mx, my, mz, EE, dt),
                              energy_flux_face<DIR>(A, G, c - sd, mx, my, mz, EI, dt)};
        const double ec[2] = {energy_flux_face<DIR>(A, G, c, ix, iy, iz, EE, dt),
                              energy_flux_face<DIR>(A, G, c, ix, iy, iz, EI, dt)};
        const double rho1 = RHO[c], cv1 = A.cv1[c], cv2 = C2[c];
        const double rho = ddiv(rho1 * cv1 + DM[c - sd] - DM[c], cv2);
        A.rho2[c] = rho;
        A.ee2[c] = ddiv(EE[c] * cv1 * rho1 + em[0] - ec[0], cv2 * rho);
        A.ei2[c] = ddiv(EI[c] * cv1 * rho1 + em[1] - ec[1], cv2 * rho);
    }
    };
    ECELL(cell)
}

// ------------------------------------------ remap kinetic-energy correction
//
// With rke the reference adds, after each momentum-flux kernel and before
// its multiplication by dm, a kinetic-energy term dk of every vertex face to
// four delta_ke cells with atomic adds (xremap.cpp:534-558,
// yremap.cpp:558-584, zremap.cpp:514-537), and energy_correction
// (diagnostics.cpp:125-135) heats the cells after the remap.  Here the sum
// is a gather in a fixed order: r_ke_dk stores dk of the three velocity
// components per vertex face, r_ke_add adds each cell's contributions in
// the order one thread of the reference adds them (component by component,
// the faces in ascending (ix, iy, iz)).  Atomic adds of a multi-threaded
// reference run sum in an arbitrary order; a one-thread run sums in this one.

// dk of the vertex face c = (ix, iy, iz) along DIR for the velocity
// components (lagged in the state copy the transport read), over
// (0..n-1 along DIR, 0..n across), into flux_x, flux_y, flux_z (free during
// the remap).  m, mp are the vertex masses after the remap at c, c + sd;
// the fluxes at c - sd, c, c + sd are the momentum fluxes before their
// multiplication by dm, each from the same expressions as the transport.
template <class T, int DIR>
KERNEL void r_ke_dk(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int sd = sdir<DIR>(G);
    const double dt = lare::gbl(A.s)->dt;
    const int ox = DIR == 0, oy = DIR == 1, oz = DIR == 2;
    double rv[4], dmv[3], cv2v[2];
    for (int k = 0; k < 4; ++k) rv[k] = rrho_v_at<DIR>(A, G, c + (k - 1) * sd);
    for (int k = 0; k < 3; ++k) dmv[k] = dmv_at<DIR>(A, G, c + (k - 1) * sd);
    const Fld<T> C2 = A.cv2;
    for (int k = 0; k < 2; ++k) {
        const int q = c + k * sd;
        double s2 = C2[q] + C2[q + sx] + C2[q + sy] + C2[q + sx + sy] + C2[q + 1] + C2[q + sx + 1] + C2[q + sy + 1] +
                    C2[q + sx + sy + 1];
        cv2v[k] = DIR == 0 ? s2 * 0.125 : 0.125 * s2;
    }
    const double rho_v1_0 = ddiv(rv[1] * A.cvc1[c] + dmv[0] - dmv[1], cv2v[0]);
    const double rho_v1_p = ddiv(rv[2] * A.cvc1[c + sd] + dmv[1] - dmv[2], cv2v[1]);
    const double m = rho_v1_0 * cv2v[0];
    const double mp = rho_v1_p * cv2v[1];
    double flm[3], fl0[3], flp[3];
    mom_fluxes_at<DIR, T, false>(A, G, c - sd, ix - ox, iy - oy, iz - oz, rv[0], rv[1], dmv[0], dt, flm);
    mom_fluxes_at<DIR, T, false>(A, G, c, ix, iy, iz, rv[1], rv[2], dmv[1], dt, fl0);
    mom_fluxes_at<DIR, T, false>(A, G, c + sd, ix + ox, iy + oy, iz + oz, rv[2], rv[3], dmv[2], dt, flp);
    const Fld<T> V3[3] = {VX, VY, VZ};
    double *const out[3] = {A.flux_x, A.flux_y, A.flux_z};
    for (int k = 0; k < 3; ++k) {
        const double f0 = V3[k][c], fp = V3[k][c + sd];
        double ai = ddiv((f0 - flm[k]) * dmv[0], m) - ddiv((f0 - fl0[k]) * dmv[1], m);
        double aip = ddiv((fp - fl0[k]) * dmv[1], mp) - ddiv((fp - flp[k]) * dmv[2], mp);
        double dk = (fp - f0) * (fl0[k] - 0.5 * (fp + f0)) - 0.5 * ai * (f0 - fl0[k]) + 0.5 * aip * (fp - fl0[k]);
        dk = dk * dmv[1] * 0.5;
        lare::gbl(out[k])[c] = dk;
    }
    };
    ECELL(cell)
}

// delta_ke of cell (ix, iy, iz) += the dk of the faces that reach it: along
// DIR the face at index - 1, across it the four faces at offsets -1 and 0,
// in ascending order, each if inside r_ke_dk's range; component by
// component.
template <class T, int DIR>
KERNEL void r_ke_add(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    // the two axes across DIR, in loop order (x before y before z)
    const int s1 = DIR == 0 ? sy : sx, s2 = DIR == 2 ? sy : 1;
    const int i1 = DIR == 0 ? iy : ix, i2 = DIR == 2 ? iy : iz;
    const int n1 = DIR == 0 ? G.ny : G.nx, n2 = DIR == 2 ? G.ny : G.nz;
    const int f = c - sdir<DIR>(G);
    auto dke = lare::gbl(A.dke);
    double acc = dke[c];
    const double *const dk3[3] = {A.flux_x, A.flux_y, A.flux_z};
    for (int k = 0; k < 3; ++k) {
        const auto dk = lare::gbl(dk3[k]);
        for (int a = -1; a <= 0; ++a)
            for (int b = -1; b <= 0; ++b) {
                const int j1 = i1 + a, j2 = i2 + b;
                if (j1 >= 0 && j1 <= n1 && j2 >= 0 && j2 <= n2) acc = acc + dk[f + a * s1 + b * s2];
            }
    }
    dke[c] = acc;
    };
    ECELL(cell)
}

// delta_ke = 0 before the first sweep (remap.cpp:38-41).
template <class T>
KERNEL void r_ke_zero(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) { lare::gbl(A.dke)[IX(ix, iy, iz)] = 0.0; };
    ECELL(cell)
}

// energy_correction (diagnostics.cpp:125-135) over (1..n)^3 after the remap:
// half of max(-delta_ke, 0) / (rho cv) to each energy.
template <class T>
KERNEL void r_ke_energy(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const double dke = ddiv(pmax(-lare::gbl(A.dke)[c], 0.0), RHO[c] * CV(ix, iy, iz));
    EE[c] += 0.5 * dke;
    EI[c] += 0.5 * dke;
    };
    ECELL(cell)
}

// Sum of the eight corner values of vertex c in the reference's order (c,
// c+sx, c+sy, c+sx+sy, c+1, c+sx+1, c+sy+1, c+sx+sy+1), given the four
// corners on the lower plane along DIR (lo) and on the upper one (hi), each
// indexed j = (offset across the other non-z axis) + 2 * (offset in z).
template <int DIR>
__device__ __forceinline__ double sum8(const double lo[4], const double hi[4])
{
    if (DIR == 0)
        return lo[0] + hi[0] + lo[1] + hi[1] + lo[2] + hi[2] + lo[3] + hi[3];
    return lo[0] + lo[1] + hi[0] + hi[1] + lo[2] + lo[3] + hi[2] + hi[3];
}

// R6' for the x and y sweeps: the transport of f_update_all with each lane
// walking K consecutive cells along DIR (one warm-up position before them
// writes nothing).  Every value that consecutive cells share is loaded once
// and carried in registers: the line values along DIR, the corner planes of
// the vertex sums, the face and vertex fluxes and vertex densities.  The
// arithmetic of each flux and update is the same expression on the same
// operands as in f_update_all, so the results are identical.  The plane
// across DIR is flattened z fastest: (y, z) for x, (x, z) for y.
template <class T, int DIR, int K>
KERNEL void f_update_march(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int ps0 = DIR == 0 ? R.y0 : R.x0, ps1 = DIR == 0 ? R.y1 : R.x1;
    const int nzr = R.z1 - R.z0 + 1;
    const int plane = (ps1 - ps0 + 1) * nzr;
    const int bpp = (plane + 255) >> 8;
    const int lb = b_ - S.boff[g_];
    const int mb = lb / bpp;
    const int t = (lb - mb * bpp) * 256 + (int)threadIdx.x;
    if (t >= plane) return;
    const int ps = ps0 + t / nzr, iz = R.z0 + t % nzr;
    const int dlo = DIR == 0 ? R.x0 : R.y0, dhi = DIR == 0 ? R.x1 : R.y1;
    const int d0 = dlo + mb * K;
    const int sd = sdir<DIR>(G);
    const int so = DIR == 0 ? G.sy : G.sx;  // stride of the other non-z axis
    const int tj[4] = {0, so, 1, so + 1};   // corner offsets within a plane
    const double dt = lare::gbl(A.s)->dt;
    const int ix0 = DIR == 0 ? d0 - 1 : ps, iy0 = DIR == 0 ? ps : d0 - 1;
    const bool tr1 = ps >= 1 && iz >= 1;  // transverse indices >= 1: density/energy cells
    // face areas are constant along DIR
    const double area_e = DIR == 0 ? DXAB(ps, iz) : DYAB(ps, iz);
    const double area_m = DIR == 0 ? DXAC(ps, iz) : DYAC(ps, iz);
    const Fld<T> W = DIR == 0 ? VX1 : VY1;
    const Fld<T> R1 = RHO, C1 = A.cv1, C2 = A.cv2;
    const auto M = DM;
    // the window enters the warm-up position c = (d0 - 1)
    int c = IX(ix0, iy0, iz);
    double E[2][3], F[3][3];
    for (int e = 0; e < 3; ++e) {
        E[0][e] = EE[c + (e - 1) * sd];
        E[1][e] = EI[c + (e - 1) * sd];
        F[0][e] = VX[c + (e - 1) * sd];
        F[1][e] = VY[c + (e - 1) * sd];
        F[2][e] = VZ[c + (e - 1) * sd];
    }
    double cvc_m = A.cvc1[c - sd], cvc_0 = A.cvc1[c];
    double cv1_0 = C1[c], cv1_p = C1[c + sd];
    double rho_0 = R1[c], rho_p = R1[c + sd];
    double w_0 = W[c];
    double dm_m = 0.0;
    double DML[4], C2L[4], PL[4], P0[4];
    for (int j = 0; j < 4; ++j) {
        DML[j] = M[c + tj[j]];
        C2L[j] = 0.0;
        P0[j] = R1[c + tj[j]] * C1[c + tj[j]];
        PL[j] = R1[c + sd + tj[j]] * C1[c + sd + tj[j]];
    }
    double rv_c = DIR == 0 ? ddiv(sum8<DIR>(P0, PL) * 0.125, cvc_0) : sum8<DIR>(P0, PL) * ddiv(0.125, cvc_0);
    double dmv_m = 0.0, flm[3] = {0.0, 0.0, 0.0}, em[2] = {0.0, 0.0};
    for (int k = -1; k < K; ++k) {
        const int d = d0 + k;
        if (d > dhi) break;
        // new values: line at c+sd / c+2sd, corner planes d+1 (DM, cv2) and d+2 (rho cv1)
        double Ep[2] = {EE[c + 2 * sd], EI[c + 2 * sd]};
        double Fp[3] = {VX[c + 2 * sd], VY[c + 2 * sd], VZ[c + 2 * sd]};
        const double cvc_p = A.cvc1[c + sd];
        const double w_p = W[c + sd];
        const double fv = DIR == 0 ? (w_0 + W[c - so] + W[c - 1] + W[c - so - 1]) * 0.25
                                   : (W[c - so] + w_0 + W[c - so - 1] + W[c - 1]) * 0.25;
        double DMH[4], C2H[4], PH[4];
        double cv1_pp = 0.0, rho_pp = 0.0;
        for (int j = 0; j < 4; ++j) {
            DMH[j] = M[c + sd + tj[j]];
            C2H[j] = C2[c + sd + tj[j]];
            const double r = R1[c + 2 * sd + tj[j]], v = C1[c + 2 * sd + tj[j]];
            if (j == 0) {
                rho_pp = r;
                cv1_pp = v;
            }
            PH[j] = r * v;
        }
        // vertex density at c+sd, vertex mass flux at c
        const double sp = sum8<DIR>(PL, PH);
        const double rv_p = DIR == 0 ? ddiv(sp * 0.125, cvc_p) : sp * ddiv(0.125, cvc_p);
        const double q = sum8<DIR>(DML, DMH);
        const double dmv_c = DIR == 0 ? q * 0.125 : 0.125 * q;
        // momentum fluxes through vertex face c
        double flc[3];
        {
            double v_advect = (w_0 + w_p) * 0.5;
            double o_v = v_advect * dt * area_m;
            double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
            double vad_p = (sign_v + 1.0) * 0.5;
            double vad_m = 1.0 - vad_p;
            double dxci = cv1_p;
            double dxcu = cv1_0 * vad_p + cv1_pp * vad_m;
            double dxbu = cvc_0 * vad_p + cvc_p * vad_m;
            double rhou = rv_c * vad_p + rv_p * vad_m;
            for (int f = 0; f < 3; ++f)
                flc[f] = mom_flux_v<DIR>(F[f][0], F[f][1], F[f][2], Fp[f], o_v, sign_v, dxci, dxcu, dxbu, rhou,
                                         dmv_c);
        }
        if (k >= 0) {
            const double s2 = sum8<DIR>(C2L, C2H);
            const double cv2v = DIR == 0 ? s2 * 0.125 : 0.125 * s2;
            const double rho_v1 = ddiv(rv_c * cvc_0 + dmv_m - dmv_c, cv2v);
            const double den = cv2v * rho_v1;
            A.vx2[c] = ddiv(rv_c * F[0][1] * cvc_0 + flm[0] - flc[0], den);
            A.vy2[c] = ddiv(rv_c * F[1][1] * cvc_0 + flm[1] - flc[1], den);
            A.vz2[c] = ddiv(rv_c * F[2][1] * cvc_0 + flm[2] - flc[2], den);
        }
        // energy fluxes through cell face c, and the cell update
        double ec[2] = {0.0, 0.0};
        if (tr1 && d >= 0) {
            const double o_v = fv * dt * area_e;
            for (int e = 0; e < 2; ++e)
                ec[e] = energy_flux_v(E[e][0], E[e][1], E[e][2], Ep[e], fv, o_v, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p,
                                      rho_0, rho_p, DML[0]);
            if (k >= 0 && d >= 1) {
                const double cv2 = C2L[0];
                const double rho = ddiv(rho_0 * cv1_0 + dm_m - DML[0], cv2);
                A.rho2[c] = rho;
                A.ee2[c] = ddiv(E[0][1] * cv1_0 * rho_0 + em[0] - ec[0], cv2 * rho);
                A.ei2[c] = ddiv(E[1][1] * cv1_0 * rho_0 + em[1] - ec[1], cv2 * rho);
            }
        }
        // slide the window one cell along DIR
        for (int e = 0; e < 2; ++e) {
            E[e][0] = E[e][1];
            E[e][1] = E[e][2];
            E[e][2] = Ep[e];
            em[e] = ec[e];
        }
        for (int f = 0; f < 3; ++f) {
            F[f][0] = F[f][1];
            F[f][1] = F[f][2];
            F[f][2] = Fp[f];
            flm[f] = flc[f];
        }
        cvc_m = cvc_0;
        cvc_0 = cvc_p;
        cv1_0 = cv1_p;
        cv1_p = cv1_pp;
        rho_0 = rho_p;
        rho_p = rho_pp;
        w_0 = w_p;
        dm_m = DML[0];
        for (int j = 0; j < 4; ++j) {
            DML[j] = DMH[j];
            C2L[j] = C2H[j];
            PL[j] = PH[j];
        }
        rv_c = rv_p;
        dmv_m = dmv_c;
        c += sd;
    }
}

// Value of v held by the previous (lane - 1) or next (lane + 1) lane of the
// wavefront (DPP wave_shr:1 / wave_shl:1 on both halves); the end lanes
// (0 for lane_prev, 63 for lane_next) get 0.  Those are halo lanes: no
// lane that writes reads a value derived from their exchange.
__device__ __forceinline__ double lane_prev(double v)
{
    const int2 h = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_mov_dpp(h.x, 0x138, 0xf, 0xf, true),
                                                __builtin_amdgcn_mov_dpp(h.y, 0x138, 0xf, 0xf, true)));
}
__device__ __forceinline__ double lane_next(double v)
{
    const int2 h = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_mov_dpp(h.x, 0x130, 0xf, 0xf, true),
                                                __builtin_amdgcn_mov_dpp(h.y, 0x130, 0xf, 0xf, true)));
}

// A 1D delta along x for the positions a wavefront marches over: lane l
// holds d[i0 + l] (indices clamped to the array's -2..imax), and
// x_at(v, i - i0) returns the value of position i to every lane.  The x
// deltas of a march are loaded once, before its loop, instead of once per
// position (a uniform load per position would be waited for at once, to move
// it to a scalar register, before the loads issued ahead of it complete).
__device__ __forceinline__ double x_row(const double *d, int i0, int imax)
{
    const int i = i0 + ((int)threadIdx.x & 63);
    return lare::gbl(d)[(i < imax ? i : imax) + 2];
}
__device__ __forceinline__ double x_at(double v, int l)
{
    const int2 h = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double,
                              make_int2(__builtin_amdgcn_readlane(h.x, l), __builtin_amdgcn_readlane(h.y, l)));
}

// Lanes of one wavefront per z-row segment: a wave covers 64 consecutive
// positions of a z-row flattened over (y, z), with two halo positions
// beyond each end of the row (z0-2 .. z1+2), and successive waves overlap by
// four lanes; lanes 2..61 write.
const int kZLine = 2000;

// Lanes of a z-line x-march kernel: each wavefront covers 64 consecutive
// positions of the z-rows of one (y, z) plane flattened z fastest, with H
// halo positions beyond each end of a row; successive wavefronts overlap by
// 2H lanes and lanes H..63-H write.  Each lane walks K cells along x, after
// one warm-up position (x0 - 1) that writes nothing.  Values at z +- 1 come
// from the neighbouring lanes and the values of plane x - 1 from the
// previous position, so each value is loaded once.
const int kZX = 3000;  // table code kZX + 100 * H + K

struct ZXPos {
    int ix0, ix1;  // first (warm-up) and last x of the lane
    int iy, iz;
    bool wz;       // lane writes (inside, not a halo lane, z in range)
};

template <int H, int K, class T>
__device__ __forceinline__ ZXPos zx_map(const Src<T> &S, const Geo &G, int g_, int b_)
{
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int nyr = R.y1 - R.y0 + 1, nzr = R.z1 - R.z0 + 1;
    const int L = nzr + 2 * H, P = nyr * L;
    const int per = 4 * (64 - 2 * H);
    const int bpp = (P + per - 1) / per;
    const int lb = b_ - S.boff[g_];
    const int xc = lb / bpp;
    const int lane = (int)threadIdx.x & 63;
    const int w = (lb - xc * bpp) * 4 + ((int)threadIdx.x >> 6);
    int u = (64 - 2 * H) * w + lane - H;
    const bool inside = u >= 0 && u < P;
    u = u < 0 ? 0 : (u >= P ? P - 1 : u);
    const int row = u / L;
    ZXPos z;
    z.iy = R.y0 + row;
    z.iz = R.z0 - H + (u - row * L);
    z.wz = inside && lane >= H && lane < 64 - H && z.iz >= R.z0 && z.iz <= R.z1;
    // a halo position beyond the box (a range reaching z = -2 or n+2) loads
    // at the box edge; only a lane beyond it would read its values, and
    // none writes
    if (z.iz > G.nz + 2) z.iz = G.nz + 2;
    if (z.iz < -2) z.iz = -2;
    z.ix0 = R.x0 + xc * K - 1;
    z.ix1 = z.ix0 + K < R.x1 ? z.ix0 + K : R.x1;
    return z;
}

// R6' for the z sweep: the transport of f_update_all with lanes along z,
// each loading only its own z (the values at z-1, z+1, z+2 and the fluxes of
// the face below come from the neighbouring lanes, two halo lanes at each
// end of a wavefront's z-row segment, so each face flux, vertex density and
// vertex mass flux is computed once), and each lane walking K positions
// along x.  The arithmetic of each flux and update is the same expression
// on the same operands as in f_update_all; every lane stays active through
// the exchanges, and lanes outside the range work on a clamped position and
// write nothing.
// The columns (x+1, y) and (x+1, y+1) of rho, cv1, dm and cv2 loaded at one
// position are the columns (x, y) and (x, y+1) of the next, and w at (x, y)
// and (x, y-1) is w at (x-1, ...) there, so each value is loaded once per
// lane instead of re-read from the neighbouring plane.
template <class T, int K>
KERNEL void f_update_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZXPos Z = zx_map<2, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> R1 = RHO, C1 = A.cv1, C2 = A.cv2, W = VZ1;
    const auto M = DM;
    // column x (carried): rho and cv1 of row y, rho cv1 of rows y and y+1,
    // dm and cv2 of rows y and y+1; w of column x - 1, rows y and y-1
    double rho_x = 0.0, cv1_x = 0.0, p_x[2] = {0.0, 0.0}, d_x[2] = {0.0, 0.0}, s_x[2] = {0.0, 0.0};
    double w_xm = 0.0, w_xmm = 0.0;
    // face areas: DYAC(ix, iz) = DXB(ix) DZC(iz), DZAB(ix, iy) = DXB(ix) DYB(iy),
    // with DXB along the march in lane registers (lane l: DXB(ix0 + l))
    const double dzc = DZC(iz), dyb = DYB(iy);
    const double dxb_row = x_row(A.dxb, Z.ix0, G.nx + 2);
    int c = IX(Z.ix0, iy, iz);
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        // every load of the position, issued before any is used:
        // column x + 1, rows y and y + 1; w at (x, y) and (x, y - 1); the
        // cell's own cvc1, v and energies
        const double rho_a = R1[c + sx], cv1_a = C1[c + sx];
        const double r_b = R1[c + sx + sy], c_b = C1[c + sx + sy];
        const double d_a[2] = {M[c + sx], M[c + sx + sy]};
        const double s_a[2] = {C2[c + sx], C2[c + sx + sy]};
        const double w_0 = W[c], w_0m = W[c - sy];
        const double cvc_c = A.cvc1[c];
        const double F[3] = {VX[c], VY[c], VZ[c]};
        const double E[2] = {EE[c], EI[c]};
        __builtin_amdgcn_sched_barrier(0);
        const double p_a[2] = {rho_a * cv1_a, r_b * c_b};
        const double dxb = x_at(dxb_row, ix - Z.ix0);
        if (ix > Z.ix0) {
            const double P4[4] = {p_x[0], p_a[0], p_x[1], p_a[1]};
            const double D4[4] = {d_x[0], d_a[0], d_x[1], d_a[1]};
            const double S4[4] = {s_x[0], s_a[0], s_x[1], s_a[1]};
            const double rho_0 = rho_x, cv1_0 = cv1_x, cvc_0 = cvc_c;
            const double fv = (w_xm + w_0 + w_xmm + w_0m) * 0.25;
            const double area_m = dxb * dzc, area_e = dxb * dyb;
            double sp = P4[0] + P4[1] + P4[2] + P4[3];
            double q = D4[0] + D4[1] + D4[2] + D4[3];
            double s2 = S4[0] + S4[1] + S4[2] + S4[3];
            for (int j = 0; j < 4; ++j) {
                sp = sp + lane_next(P4[j]);
                q = q + lane_next(D4[j]);
                s2 = s2 + lane_next(S4[j]);
            }
            const double rv = sp * ddiv(0.125, cvc_0);
            const double dmv = 0.125 * q;
            const double cv2v = 0.125 * s2;
            const double rv_p = lane_next(rv), w_p = lane_next(w_0), rho_p = lane_next(rho_0);
            const double cv1_p = lane_next(cv1_0), cv1_pp = lane_next(cv1_p);
            const double cvc_p = lane_next(cvc_0), cvc_m = lane_prev(cvc_0);
            double flc[3];
            {
                double v_advect = (w_0 + w_p) * 0.5;
                double o_v = v_advect * dt * area_m;
                double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
                double vad_p = (sign_v + 1.0) * 0.5;
                double vad_m = 1.0 - vad_p;
                double dxci = cv1_p;
                double dxcu = cv1_0 * vad_p + cv1_pp * vad_m;
                double dxbu = cvc_0 * vad_p + cvc_p * vad_m;
                double rhou = rv * vad_p + rv_p * vad_m;
                for (int f = 0; f < 3; ++f) {
                    const double fp = lane_next(F[f]);
                    flc[f] = mom_flux_v<2>(lane_prev(F[f]), F[f], fp, lane_next(fp), o_v, sign_v, dxci, dxcu, dxbu,
                                           rhou, dmv);
                }
            }
            double ec[2];
            {
                const double o_v = fv * dt * area_e;
                for (int e = 0; e < 2; ++e) {
                    const double ep = lane_next(E[e]);
                    ec[e] = energy_flux_v(lane_prev(E[e]), E[e], ep, lane_next(ep), fv, o_v, cvc_m, cvc_0, cvc_p,
                                          cv1_0, cv1_p, rho_0, rho_p, D4[0]);
                }
            }
            const double flm[3] = {lane_prev(flc[0]), lane_prev(flc[1]), lane_prev(flc[2])};
            const double em[2] = {lane_prev(ec[0]), lane_prev(ec[1])};
            const double dmv_m = lane_prev(dmv), dm_m = lane_prev(D4[0]);
            if (Z.wz) {
                const double rho_v1 = ddiv(rv * cvc_0 + dmv_m - dmv, cv2v);
                const double den = cv2v * rho_v1;
                A.vx2[c] = ddiv(rv * F[0] * cvc_0 + flm[0] - flc[0], den);
                A.vy2[c] = ddiv(rv * F[1] * cvc_0 + flm[1] - flc[1], den);
                A.vz2[c] = ddiv(rv * F[2] * cvc_0 + flm[2] - flc[2], den);
                if (ix >= 1 && iy >= 1 && iz >= 1) {
                    const double cv2 = S4[0];
                    const double rho = ddiv(rho_0 * cv1_0 + dm_m - D4[0], cv2);
                    A.rho2[c] = rho;
                    A.ee2[c] = ddiv(E[0] * cv1_0 * rho_0 + em[0] - ec[0], cv2 * rho);
                    A.ei2[c] = ddiv(E[1] * cv1_0 * rho_0 + em[1] - ec[1], cv2 * rho);
                }
            }
        }
        rho_x = rho_a;
        cv1_x = cv1_a;
        for (int j = 0; j < 2; ++j) {
            p_x[j] = p_a[j];
            d_x[j] = d_a[j];
            s_x[j] = s_a[j];
        }
        w_xm = w_0;
        w_xmm = w_0m;
    }
}

// R2 (f_cvc1) as a z-line x-march (H = 1): cv1 of the columns (x+1, y) and
// (x+1, y+1) is loaded once and carried as the columns at x of the next
// position; the values at z + 1 come from the next lane.  The eight terms
// are summed in f_cvc1's order.
template <class T, int K>
KERNEL void f_cvc1_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const Fld<T> C1 = A.cv1;
    double o0 = 0.0, o1 = 0.0, o0z = 0.0, o1z = 0.0;  // column x, rows y and y+1, at z and z+1
    int c = IX(Z.ix0, Z.iy, Z.iz);
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        const double n0 = C1[c + sx], n1 = C1[c + sx + sy];
        const double n0z = lane_next(n0), n1z = lane_next(n1);
        if (Z.wz && ix > Z.ix0) A.cvc1[c] = 0.125 * (o0 + n0 + o1 + n1 + o0z + n0z + o1z + n1z);
        o0 = n0;
        o1 = n1;
        o0z = n0z;
        o1z = n1z;
    }
}

// R3 for the x and y sweeps with each lane walking K vertices along DIR (the
// mapping of f_update_march): B / db along DIR, cvc1 and cv1 are carried as
// a window, so each is loaded once; the same arithmetic as f_ct_fluxes.
// The x sweep reads B unscaled and multiplies by the component's area at
// each position (remap.cpp:46-50), as ct_flux<0, SC> does.
template <class T, int DIR, int K>
KERNEL void f_ct_march(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int ps0 = DIR == 0 ? R.y0 : R.x0, ps1 = DIR == 0 ? R.y1 : R.x1;
    const int nzr = R.z1 - R.z0 + 1;
    const int plane = (ps1 - ps0 + 1) * nzr;
    const int bpp = (plane + 255) >> 8;
    const int lb = b_ - S.boff[g_];
    const int mb = lb / bpp;
    const int t = (lb - mb * bpp) * 256 + (int)threadIdx.x;
    if (t >= plane) return;
    const int ps = ps0 + t / nzr, iz = R.z0 + t % nzr;
    const int dlo = DIR == 0 ? R.x0 : R.y0, dhi = DIR == 0 ? R.x1 : R.y1;
    const int d0 = dlo + mb * K;
    const int sx = G.sx, sy = G.sy;
    const int sd = sdir<DIR>(G);
    const double dt = lare::gbl(A.s)->dt;
    // flux a: by (x sweep) or bx (y sweep); flux b: bz
    const Fld<T> Ba = DIR == 0 ? BY : BX, Bb = BZ;
    const Fld<T> W = DIR == 0 ? VX1 : VY1;
    const int owa = -1, owb = DIR == 0 ? -sy : -sx;
    const int twa = DIR == 0 ? sy : sx, twb = 1;
    const Fld<T> db1 = A.db1;
    // B area of the x sweep at position d along x: dyab(d, iz), dzab(d, iy)
    auto area_a = [&](int d) -> double { return DIR == 0 ? DXB(d) * DZB(iz) : 1.0; };
    auto area_b = [&](int d) -> double { return DIR == 0 ? DXB(d) * DYB(ps) : 1.0; };
    auto fval = [&](const Fld<T> Bf, int p, int tw, int d, bool a, double &db) -> double {
        db = (db1[p] + db1[p + tw]) * 0.5;
        double b = Bf[p];
        if (DIR == 0) b = b * (a ? area_a(d) : area_b(d));
        return ddiv(b, db);
    };
    int c = DIR == 0 ? IX(d0, ps, iz) : IX(ps, d0, iz);
    double fa[3], fb[3], da[2], db[2], dtmp;
    fa[0] = fval(Ba, c - sd, twa, d0 - 1, true, dtmp);
    fa[1] = fval(Ba, c, twa, d0, true, da[0]);
    fa[2] = fval(Ba, c + sd, twa, d0 + 1, true, da[1]);
    fb[0] = fval(Bb, c - sd, twb, d0 - 1, false, dtmp);
    fb[1] = fval(Bb, c, twb, d0, false, db[0]);
    fb[2] = fval(Bb, c + sd, twb, d0 + 1, false, db[1]);
    double cvc_m = A.cvc1[c - sd], cvc_0 = A.cvc1[c], cv1_0 = A.cv1[c];
    const auto FA = lare::gbl(A.fl[0]), FB = lare::gbl(A.fl[1]);
    for (int k = 0; k < K; ++k) {
        const int d = d0 + k;
        if (d > dhi) break;
        double da2, db2;
        const double fa3 = fval(Ba, c + 2 * sd, twa, d + 2, true, da2);
        const double fb3 = fval(Bb, c + 2 * sd, twb, d + 2, false, db2);
        const double cvc_p = A.cvc1[c + sd], cv1_p = A.cv1[c + sd];
        const double w0 = W[c];
        const double va = (w0 + W[c + owa]) * 0.5, vb = (w0 + W[c + owb]) * 0.5;
        FA[c] = ct_flux_v(fa[0], fa[1], fa[2], fa3, da[0], da[1], va, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
        FB[c] = ct_flux_v(fb[0], fb[1], fb[2], fb3, db[0], db[1], vb, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
        fa[0] = fa[1];
        fa[1] = fa[2];
        fa[2] = fa3;
        fb[0] = fb[1];
        fb[1] = fb[2];
        fb[2] = fb3;
        da[0] = da[1];
        da[1] = da2;
        db[0] = db[1];
        db[1] = db2;
        cvc_m = cvc_0;
        cvc_0 = cvc_p;
        cv1_0 = cv1_p;
        c += sd;
    }
}

// R3 and R5 of the y sweep in one march (the mapping of f_update_march, each
// lane walking K positions along y) over the union of their ranges (x
// 0..n+1, y -1..n+2, z 0..n+1): the CT fluxes of f_ct_march<1> on (0..n)^3
// and the y mass flux of f_mass_flux<1> (zero on the ghost faces), sharing
// the cvc1 and cv1 windows; rho (the current state copy) is a window along
// y too.  Positions beyond the box along y are clamped; they feed only
// fluxes that are not stored.
template <class T, int K>
KERNEL OCC(5) void f_cty_mass(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int nzr = R.z1 - R.z0 + 1;
    const int plane = (R.x1 - R.x0 + 1) * nzr;
    const int bpp = (plane + 255) >> 8;
    const int lb = b_ - S.boff[g_];
    const int mb = lb / bpp;
    const int t = (lb - mb * bpp) * 256 + (int)threadIdx.x;
    if (t >= plane) return;
    const int ix = R.x0 + t / nzr, iz = R.z0 + t % nzr;
    const int d0 = R.y0 + mb * K;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int sx = G.sx;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> Ba = BX, Bb = BZ, W = VY1, db1 = A.db1;
    auto cy = [&](int y) { return y > ny + 2 ? ny + 2 : y; };
    auto at = [&](int y) { return IX(ix, cy(y), iz); };
    // f = B / db at y (bx flux: tw = sx; bz flux: tw = 1)
    auto fval = [&](const Fld<T> Bf, int p, int tw, double &db) -> double {
        db = (db1[p] + db1[p + tw]) * 0.5;
        return ddiv((double)Bf[p], db);
    };
    double fa[3], fb[3], da[2], db[2], dtmp;
    fa[0] = fval(Ba, at(d0 - 1), sx, dtmp);
    fa[1] = fval(Ba, at(d0), sx, da[0]);
    fa[2] = fval(Ba, at(d0 + 1), sx, da[1]);
    fb[0] = fval(Bb, at(d0 - 1), 1, dtmp);
    fb[1] = fval(Bb, at(d0), 1, db[0]);
    fb[2] = fval(Bb, at(d0 + 1), 1, db[1]);
    double cvc_m = A.cvc1[at(d0 - 1)], cvc_0 = A.cvc1[at(d0)], cv1_0 = A.cv1[at(d0)];
    double ro[3] = {RHO[at(d0 - 1)], RHO[at(d0)], RHO[at(d0 + 1)]};
    const double area_m = DYAB(ix, iz);  // face area of the y mass flux
    const bool ct_xz = ix <= nx && iz <= nz;
    const bool mf_xz = ix <= nx + 1 && iz <= nz + 1;
    const auto FA = lare::gbl(A.fl[0]), FB = lare::gbl(A.fl[1]);
    for (int k = 0; k < K; ++k) {
        const int d = d0 + k;
        if (d > R.y1) break;
        const int c = at(d);
        // every load of the position, issued before any is used (ix >= 0
        // and iz >= 0 over this kernel's range, so c - sx - 1 is in the box)
        const int p2 = at(d + 2), p1 = at(d + 1);
        const double ba2 = Ba[p2], dba2 = db1[p2], dba2t = db1[p2 + sx], bb2 = Bb[p2], dbb2t = db1[p2 + 1];
        const double cvc_p = A.cvc1[p1], cv1_p = A.cv1[p1];
        const double ro3 = RHO[p2];
        const double w0 = W[c], w_z = W[c - 1], w_x = W[c - sx], w_xz = W[c - sx - 1];
        __builtin_amdgcn_sched_barrier(0);
        // fval at d + 2
        const double da2 = (dba2 + dba2t) * 0.5, db2 = (dba2 + dbb2t) * 0.5;
        const double fa3 = ddiv(ba2, da2), fb3 = ddiv(bb2, db2);
        if (ct_xz && d >= 0 && d <= ny) {
            const double va = (w0 + w_z) * 0.5, vb = (w0 + w_x) * 0.5;
            FA[c] = ct_flux_v(fa[0], fa[1], fa[2], fa3, da[0], da[1], va, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
            FB[c] = ct_flux_v(fb[0], fb[1], fb[2], fb3, db[0], db[1], vb, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
        }
        if (mf_xz) {
            double dmf = 0.0;
            if (d >= 0 && d <= ny) {
                const double v_advect = (w_x + w0 + w_xz + w_z) * 0.25;  // face_v<1>
                const double o_v = v_advect * dt * area_m;
                dmf = mass_flux_v(ro[0], ro[1], ro[2], ro3, v_advect, o_v, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
            }
            DM[c] = dmf;
        }
        fa[0] = fa[1];
        fa[1] = fa[2];
        fa[2] = fa3;
        fb[0] = fb[1];
        fb[1] = fb[2];
        fb[2] = fb3;
        da[0] = da[1];
        da[1] = da2;
        db[0] = db[1];
        db[1] = db2;
        ro[0] = ro[1];
        ro[1] = ro[2];
        ro[2] = ro3;
        cvc_m = cvc_0;
        cvc_0 = cvc_p;
        cv1_0 = cv1_p;
    }
}

// R3 for the z sweep with the z-line lane mapping (table code kZLine): each
// lane loads its own z and takes B / db, db, cvc1 and cv1 at z-1, z+1, z+2
// from its neighbour lanes; the same arithmetic as f_ct_fluxes<2>.
template <class T>
KERNEL void f_ct_zline(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int nyr = R.y1 - R.y0 + 1, nzr = R.z1 - R.z0 + 1;
    const int L = nzr + 4, P = nyr * L;
    const int bpp = (P + 239) / 240;
    const int lb = b_ - S.boff[g_];
    const int xp = lb / bpp;
    const int lane = (int)threadIdx.x & 63;
    const int w = (lb - xp * bpp) * 4 + ((int)threadIdx.x >> 6);
    int u = 60 * w + lane - 2;
    const bool inside = u >= 0 && u < P;
    u = u < 0 ? 0 : (u >= P ? P - 1 : u);
    const int row = u / L;
    const int ix = R.x0 + xp, iy = R.y0 + row, iz = R.z0 - 2 + (u - row * L);
    const bool writer = inside && lane >= 2 && lane < 62 && iz >= R.z0 && iz <= R.z1;
    const int sx = G.sx, sy = G.sy;
    const int c = IX(ix, iy, iz);
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> db1 = A.db1, W = VZ1;
    const double d0 = db1[c];
    // flux a: bx (tw = sx, ow = -sy); flux b: by (tw = sy, ow = -sx)
    const double dba = (d0 + db1[c + sx]) * 0.5, dbb = (d0 + db1[c + sy]) * 0.5;
    const double fa = ddiv(BX[c], dba), fb = ddiv(BY[c], dbb);
    const double w0 = W[c];
    const double va = (w0 + W[c - sy]) * 0.5, vb = (w0 + W[c - sx]) * 0.5;
    const double cvc_0 = A.cvc1[c], cv1_0 = A.cv1[c];
    const double cvc_m = lane_prev(cvc_0), cvc_p = lane_next(cvc_0), cv1_p = lane_next(cv1_0);
    const double dbap = lane_next(dba), dbbp = lane_next(dbb);
    const double fam = lane_prev(fa), fap = lane_next(fa), fap2 = lane_next(fap);
    const double fbm = lane_prev(fb), fbp = lane_next(fb), fbp2 = lane_next(fbp);
    if (!writer) return;
    lare::gbl(A.fl[0])[c] = ct_flux_v(fam, fa, fap, fap2, dba, dbap, va, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
    lare::gbl(A.fl[1])[c] = ct_flux_v(fbm, fb, fbp, fbp2, dbb, dbbp, vb, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
}

// R3 + R4 of the x sweep in one z-line x-march (H = 1) over the union of the
// B extents (-2..n+2): B is read from its own arrays and written, times its
// face area and CT-updated, into bx1..bz1 (B's second copy during the
// remap), so no lane reads a B another lane has written.  The by flux is
// evaluated on this lane's column (y) and on (y-1) (for bx), the bz flux on
// (y) only; their values at x-1 are carried and the bz flux at z-1 comes
// from the lane below.  Each flux and update is the expression of
// ct_flux<0, SC> and f_ct_update<0>.  Loads beyond the box are clamped;
// they feed only fluxes outside (0..n)^3, which no update reads.
template <class T, int K>
KERNEL OCC(3) void f_ctx_fused(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    auto clx = [&](int x) { return x < -2 ? -2 : (x > nx + 2 ? nx + 2 : x); };
    const int ym = iy - 1 < -2 ? -2 : iy - 1, yp = iy + 1 > ny + 2 ? ny + 2 : iy + 1;
    const int sy = G.sy, dyp = (yp - iy) * sy, dym = (ym - iy) * sy;
    const Fld<T> db1 = A.db1;
    const double dzb = DZB(iz), dyb = DYB(iy), axx = dyb * dzb;  // DXAB(iy, iz), b_area<0>
    // DXB along the march: lane l holds DXB(clx(ix0 - 1 + l))
    const int xr0 = Z.ix0 - 1;
    const double dxb_row = lare::gbl(A.dxb)[clx(xr0 + ((int)threadIdx.x & 63)) + 2];
    // f = B * area / db and db along x at x+k, for the by flux on rows y
    // (a) and y-1 (m) and the bz flux on row y (b); area: dyab = dxb * dzb,
    // dzab = dxb * dyb (remap.cpp:46-50).  The loads of a column (colv) are
    // issued together with the position's other loads, before any of them
    // is used.
    struct ColV {
        double d0, dp, dm, by, bym, bz;
    };
    auto col_load = [&](int xq) {
        const int c = IX(clx(xq), iy, iz);
        return ColV{db1[c], db1[c + dyp], db1[c + dym], BY[c], BY[c + dym], BZ[c]};
    };
    auto col = [&](const ColV &v, int xq, double &fa, double &da, double &sa, double &fm, double &dm, double &fb,
                   double &db, double &sb) {
        const double d0 = v.d0, dbz = lane_next(d0);
        da = (d0 + v.dp) * 0.5;
        dm = (v.dm + d0) * 0.5;
        db = (d0 + dbz) * 0.5;
        const double dxb = x_at(dxb_row, xq - xr0);
        sa = v.by * (dxb * dzb);
        fa = ddiv(sa, da);
        fm = ddiv(v.bym * (dxb * dzb), dm);
        sb = v.bz * (dxb * dyb);
        fb = ddiv(sb, db);
    };
    double fa[3], da[2], sa[2], fm[3], dmm[2], fb[3], dbb[2], sb[2], t[5];
    {
        const ColV v0 = col_load(Z.ix0 - 1), v1 = col_load(Z.ix0), v2 = col_load(Z.ix0 + 1);
        col(v0, Z.ix0 - 1, fa[0], t[0], t[1], fm[0], t[2], fb[0], t[3], t[4]);  // distinct outputs: col writes them in turn
        col(v1, Z.ix0, fa[1], da[0], sa[0], fm[1], dmm[0], fb[1], dbb[0], sb[0]);
        col(v2, Z.ix0 + 1, fa[2], da[1], sa[1], fm[2], dmm[1], fb[2], dbb[1], sb[1]);
    }
    auto ca = [&](int x, int y) { return A.cvc1[IX(clx(x), y, iz)]; };
    auto cv = [&](int x, int y) { return A.cv1[IX(clx(x), y, iz)]; };
    double cva_m = ca(Z.ix0 - 1, iy), cva_0 = ca(Z.ix0, iy), cvm_m = ca(Z.ix0 - 1, ym), cvm_0 = ca(Z.ix0, ym);
    double c1a_0 = cv(Z.ix0, iy), c1m_0 = cv(Z.ix0, ym);
    double Fa_x = 0.0, Fb_x = 0.0;  // fluxes at x-1
    // rho (current state copy) along x at x-1, x, x+1 for the mass flux
    double ro[3] = {RHO[IX(clx(Z.ix0 - 1), iy, iz)], RHO[IX(clx(Z.ix0), iy, iz)], RHO[IX(clx(Z.ix0 + 1), iy, iz)]};
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix) {
        const int c = IX(ix, iy, iz);
        // every load of the position, issued before any is used
        const ColV v3 = col_load(ix + 2);
        const double cva_p = ca(ix + 1, iy), cvm_p = ca(ix + 1, ym), c1a_p = cv(ix + 1, iy), c1m_p = cv(ix + 1, ym);
        const int xc = clx(ix);
        const double w0 = VX1[IX(xc, iy, iz)], wm = VX1[IX(xc, ym, iz)];
        const double ro3 = RHO[IX(clx(ix + 2), iy, iz)];
        const double bxc = BX[IX(xc, iy, iz)];  // the warm-up position may lie below the box
        __builtin_amdgcn_sched_barrier(0);
        double fa3, da2, sa2, fm3, dm2, fb3, db2, sb2;
        col(v3, ix + 2, fa3, da2, sa2, fm3, dm2, fb3, db2, sb2);
        const double w0z = lane_prev(w0), wmz = lane_prev(wm);
        // by flux (x_by: v = (vx1[c] + vx1[c-1]) / 2) at (y) and (y-1);
        // bz flux (v = (vx1[c] + vx1[c-sy]) / 2) at (y)
        const double Fa = ct_flux_v(fa[0], fa[1], fa[2], fa3, da[0], da[1], (w0 + w0z) * 0.5, dt, cva_m, cva_0, cva_p,
                                    c1a_0, c1a_p);
        const double Fm = ct_flux_v(fm[0], fm[1], fm[2], fm3, dmm[0], dmm[1], (wm + wmz) * 0.5, dt, cvm_m, cvm_0,
                                    cvm_p, c1m_0, c1m_p);
        const double Fb = ct_flux_v(fb[0], fb[1], fb[2], fb3, dbb[0], dbb[1], (w0 + wm) * 0.5, dt, cva_m, cva_0,
                                    cva_p, c1a_0, c1a_p);
        const double Fbz = lane_prev(Fb);
        if (Z.wz && ix > Z.ix0) {
            // R5: the x mass flux over its range (dm = 0 on the ghost faces)
            if (ix >= -1 && ix <= nx + 2 && iy >= 0 && iy <= ny + 1 && iz >= 0 && iz <= nz + 1) {
                double dmf = 0.0;
                if (ix >= 0 && ix <= nx) {
                    const double v_advect = (w0 + wm + w0z + wmz) * 0.25;  // face_v<0>
                    const double o_v = v_advect * dt * axx;
                    dmf = mass_flux_v(ro[0], ro[1], ro[2], ro3, v_advect, o_v, cva_m, cva_0, cva_p, c1a_0, c1a_p);
                }
                DM[c] = dmf;
            }
            const bool x1 = ix >= 1 && ix <= nx, x0 = ix >= 0 && ix <= nx;
            const bool y1 = iy >= 1 && iy <= ny, y0 = iy >= 0 && iy <= ny;
            const bool z1 = iz >= 1 && iz <= nz, z0 = iz >= 0 && iz <= nz;
            const bool ebx = iy >= -1 && iz >= -1, eby = ix >= -1 && iz >= -1, ebz = ix >= -1 && iy >= -1;
            double bx = ebx ? bxc * axx : 0.0;
            double by = eby ? sa[0] : 0.0;
            double bz = ebz ? sb[0] : 0.0;
            if (x1 && y0 && z1) by = by - Fa + Fa_x;
            const bool bxr = x0 && y1 && z1;
            if (bxr) bx = bx + Fa - Fm;
            if (x1 && y1 && z0) bz = bz - Fb + Fb_x;
            if (bxr) bx = bx + Fb - Fbz;
            if (ebx) A.bx1[c] = bx;
            if (eby) A.by1[c] = by;
            if (ebz) A.bz1[c] = bz;
        }
        Fa_x = Fa;
        Fb_x = Fb;
        ro[0] = ro[1];
        ro[1] = ro[2];
        ro[2] = ro3;
        fa[0] = fa[1];
        fa[1] = fa[2];
        fa[2] = fa3;
        fm[0] = fm[1];
        fm[1] = fm[2];
        fm[2] = fm3;
        fb[0] = fb[1];
        fb[1] = fb[2];
        fb[2] = fb3;
        da[0] = da[1];
        da[1] = da2;
        dmm[0] = dmm[1];
        dmm[1] = dm2;
        dbb[0] = dbb[1];
        dbb[1] = db2;
        sa[0] = sa[1];
        sa[1] = sa2;
        sb[0] = sb[1];
        sb[1] = sb2;
        cva_m = cva_0;
        cva_0 = cva_p;
        cvm_m = cvm_0;
        cvm_0 = cvm_p;
        c1a_0 = c1a_p;
        c1m_0 = c1m_p;
    }
}

// Lanes of a z-line y-march kernel: as zx_map with the roles of x and y
// exchanged, each wavefront covering 64 consecutive positions of the z-rows
// of one (x, z) plane flattened z fastest (H halo positions beyond each end
// of a row), each lane walking K cells along y after one warm-up position.
const int kZY = 4000;  // table code kZY + 100 * H + K

struct ZYPos {
    int iy0, iy1;  // first (warm-up) and last y of the lane
    int ix, iz;
    bool wz;       // lane writes (inside, not a halo lane, z in range)
};

template <int H, int K, class T>
__device__ __forceinline__ ZYPos zy_map(const Src<T> &S, const Geo &G, int g_, int b_)
{
    const R3 R = {S.r.x0, G.nx + S.r.dx1, S.r.y0, G.ny + S.r.dy1, S.r.z0, G.nz + S.r.dz1};
    const int nxr = R.x1 - R.x0 + 1, nzr = R.z1 - R.z0 + 1;
    const int L = nzr + 2 * H, P = nxr * L;
    const int per = 4 * (64 - 2 * H);
    const int bpp = (P + per - 1) / per;
    const int lb = b_ - S.boff[g_];
    const int yc = lb / bpp;
    const int lane = (int)threadIdx.x & 63;
    const int w = (lb - yc * bpp) * 4 + ((int)threadIdx.x >> 6);
    int u = (64 - 2 * H) * w + lane - H;
    const bool inside = u >= 0 && u < P;
    u = u < 0 ? 0 : (u >= P ? P - 1 : u);
    const int row = u / L;
    ZYPos z;
    z.ix = R.x0 + row;
    z.iz = R.z0 - H + (u - row * L);
    z.wz = inside && lane >= H && lane < 64 - H && z.iz >= R.z0 && z.iz <= R.z1;
    if (z.iz > G.nz + 2) z.iz = G.nz + 2;
    if (z.iz < -2) z.iz = -2;
    z.iy0 = R.y0 + yc * K - 1;
    z.iy1 = z.iy0 + K < R.y1 ? z.iy0 + K : R.y1;
    return z;
}

// R3, R4 and R5 of the y sweep in one z-line y-march (H = 1) over the union
// of the B extents: the CT fluxes of f_cty_mass (the bx flux on this lane's
// column and on column x - 1, the bz flux, whose value at z - 1 comes from
// the lane below), the update of f_ct_update<1> and the y mass flux on
// f_cty_mass's range.  B is read from the copy the x sweep wrote (view bit
// 1, st value 2) and written into the third copy (bxy..bzy, view bit 4, st
// value 16), unchanged where no flux updates it, so no lane reads a B that
// another lane has updated.  Each flux, update and mass flux is the expression of f_cty_mass
// and f_ct_update<1> on the same operands; positions beyond the box are
// clamped and feed only values that are not stored.
template <class T, int K>
KERNEL OCC(3) void f_cty_fused(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZYPos Z = zy_map<1, K>(S, G, g_, b_);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int sx = G.sx;
    const int ix = Z.ix, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> Ba = BX, Bb = BZ, W = VY1, db1 = A.db1;
    const Fld<T> OX = A.bx1, OY = A.by1, OZ = A.bz1;
    // columns x - 1 and x + 1, clamped into the box
    const int oxm = (ix - 1 < -2 ? 0 : -1) * sx, oxp = (ix + 1 > nx + 2 ? 0 : 1) * sx;
    auto cy = [&](int y) { return y > ny + 2 ? ny + 2 : (y < -2 ? -2 : y); };
    auto at = [&](int y) { return IX(ix, cy(y), iz); };
    // f = B / db at y with db the average with the column across (bx flux:
    // x + 1, from the column loads; bz flux: z + 1, from the next lane)
    struct Col {
        double ba, bam, bb, d0, dm, dp;
    };
    auto col = [&](int y) {
        const int p = at(y);
        return Col{Ba[p], Ba[p + oxm], Bb[p], db1[p], db1[p + oxm], db1[p + oxp]};
    };
    auto fvals = [&](const Col &v, double &fa, double &da, double &fam, double &dam, double &fb, double &db) {
        da = (v.d0 + v.dp) * 0.5;
        fa = ddiv(v.ba, da);
        dam = (v.dm + v.d0) * 0.5;
        fam = ddiv(v.bam, dam);
        db = (v.d0 + lane_next(v.d0)) * 0.5;
        fb = ddiv(v.bb, db);
    };
    double fa[3], fam[3], fb[3], da[2], dam[2], db[2], t0, t1, t2;
    {
        const Col v0 = col(Z.iy0 - 1), v1 = col(Z.iy0), v2 = col(Z.iy0 + 1);
        fvals(v0, fa[0], t0, fam[0], t1, fb[0], t2);
        fvals(v1, fa[1], da[0], fam[1], dam[0], fb[1], db[0]);
        fvals(v2, fa[2], da[1], fam[2], dam[1], fb[2], db[1]);
    }
    double cvc_m = A.cvc1[at(Z.iy0 - 1)], cvc_0 = A.cvc1[at(Z.iy0)], cv1_0 = A.cv1[at(Z.iy0)];
    double cvcm_m = A.cvc1[at(Z.iy0 - 1) + oxm], cvcm_0 = A.cvc1[at(Z.iy0) + oxm], cv1m_0 = A.cv1[at(Z.iy0) + oxm];
    double ro[3] = {RHO[at(Z.iy0 - 1)], RHO[at(Z.iy0)], RHO[at(Z.iy0 + 1)]};
    const double area_m = DYAB(ix, iz);  // face area of the y mass flux
    const bool x0 = ix >= 0 && ix <= nx, x1 = ix >= 1 && ix <= nx;
    const bool z0 = iz >= 0 && iz <= nz, z1 = iz >= 1 && iz <= nz;
    const bool ebx = iz >= -1, eby = ix >= -1 && iz >= -1, ebz = ix >= -1;  // B extents but y
    const bool mf_xz = ix >= 0 && ix <= nx + 1 && iz >= 0 && iz <= nz + 1;
    double FA_y = 0.0, FB_y = 0.0;  // fluxes at y - 1
    for (int d = Z.iy0; d <= Z.iy1; ++d) {
        const int c = at(d);  // the warm-up position may lie below the box
        // every load of the position, issued before any is used
        const Col v3 = col(d + 2);
        const int p1 = at(d + 1);
        const double cvc_p = A.cvc1[p1], cv1_p = A.cv1[p1], cvcm_p = A.cvc1[p1 + oxm], cv1m_p = A.cv1[p1 + oxm];
        const double ro3 = RHO[at(d + 2)];
        const double w0 = W[c], wxm = W[c + oxm];
        const double bx = BX[c], by = BY[c], bz = BZ[c];
        __builtin_amdgcn_sched_barrier(0);
        double fa3, da2, fam3, dam2, fb3, db2;
        fvals(v3, fa3, da2, fam3, dam2, fb3, db2);
        const double w0z = lane_prev(w0), wxmz = lane_prev(wxm);
        const double va = (w0 + w0z) * 0.5, vb = (w0 + wxm) * 0.5, vam = (wxm + wxmz) * 0.5;
        const double FA = ct_flux_v(fa[0], fa[1], fa[2], fa3, da[0], da[1], va, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
        const double FB = ct_flux_v(fb[0], fb[1], fb[2], fb3, db[0], db[1], vb, dt, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
        const double FAm = ct_flux_v(fam[0], fam[1], fam[2], fam3, dam[0], dam[1], vam, dt, cvcm_m, cvcm_0, cvcm_p,
                                     cv1m_0, cv1m_p);
        const double FBz = lane_prev(FB);
        if (Z.wz && d > Z.iy0) {
            // R5 on f_cty_mass's range (x 0..n+1, y -1..n+2, z 0..n+1)
            if (mf_xz && d >= -1 && d <= ny + 2) {
                double dmf = 0.0;
                if (d >= 0 && d <= ny) {
                    const double v_advect = (wxm + w0 + wxmz + w0z) * 0.25;  // face_v<1>
                    const double o_v = v_advect * dt * area_m;
                    dmf = mass_flux_v(ro[0], ro[1], ro[2], ro3, v_advect, o_v, cvc_m, cvc_0, cvc_p, cv1_0, cv1_p);
                }
                DM[c] = dmf;
            }
            // R4 (f_ct_update<1>), into the third copy
            const bool y0 = d >= 0 && d <= ny, y1 = d >= 1 && d <= ny;
            double bxo = bx, byo = by, bzo = bz;
            if (x0 && y1 && z1) bxo = bx - FA + FA_y;
            if (x1 && y0 && z1) {
                byo = by + FA - FAm;
                byo = byo + FB - FBz;
            }
            if (x1 && y1 && z0) bzo = bz - FB + FB_y;
            if (ebx && d >= -1) OX[c] = bxo;
            if (eby) OY[c] = byo;
            if (ebz && d >= -1) OZ[c] = bzo;
        }
        FA_y = FA;
        FB_y = FB;
        ro[0] = ro[1];
        ro[1] = ro[2];
        ro[2] = ro3;
        fa[0] = fa[1];
        fa[1] = fa[2];
        fa[2] = fa3;
        fam[0] = fam[1];
        fam[1] = fam[2];
        fam[2] = fam3;
        fb[0] = fb[1];
        fb[1] = fb[2];
        fb[2] = fb3;
        da[0] = da[1];
        da[1] = da2;
        dam[0] = dam[1];
        dam[1] = dam2;
        db[0] = db[1];
        db[1] = db2;
        cvc_m = cvc_0;
        cvc_0 = cvc_p;
        cv1_0 = cv1_p;
        cvcm_m = cvcm_0;
        cvcm_0 = cvcm_p;
        cv1m_0 = cv1m_p;
    }
}

// R3 + R4 of the z sweep in one z-line x-march (H = 2) over the union of the
// B extents: B is read from bx1..bz1 (where the x and y sweeps left it),
// CT-updated and divided by its face area (remap.cpp:91-95) into its own
// arrays.  The bx flux is evaluated on this lane's column, the by flux on
// (y) and (y-1); z neighbours come from the lanes around, the bx flux at
// x-1 is carried.  Each flux and update is the expression of ct_flux<2>
// and f_ct_update<2>.
template <class T, int K>
KERNEL OCC(4) void f_ctz_fused(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZXPos Z = zx_map<2, K>(S, G, g_, b_);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    auto clx = [&](int x) { return x < -2 ? -2 : (x > nx + 2 ? nx + 2 : x); };
    const int ym = iy - 1 < -2 ? -2 : iy - 1, yp = iy + 1 > ny + 2 ? ny + 2 : iy + 1;
    const Fld<T> db1 = A.db1, W = VZ1, B1x = A.bx1, B1y = A.by1, B1z = A.bz1;
    // the z flux of f = B / db at this lane (ct_flux<2>: fm..fp2 and db,
    // dbp from the lanes at z-1..z+2, v from w and its offset column)
    auto zflux = [&](double b, double d, double w, double wo, double cvc, double cv1) -> double {
        const double f = ddiv(b, d);
        const double fp = lane_next(f);
        return ct_flux_v(lane_prev(f), f, fp, lane_next(fp), d, lane_next(d), (w + wo) * 0.5, dt, lane_prev(cvc), cvc,
                         lane_next(cvc), cv1, lane_next(cv1));
    };
    const double dyb = DYB(iy), dzb = DZB(iz), axx = dyb * dzb;  // b_area<0>
    const double dxb_row = lare::gbl(A.dxb)[clx(Z.ix0 + ((int)threadIdx.x & 63)) + 2];  // DXB(clx(ix0 + lane))
    double w_xm = W[IX(clx(Z.ix0 - 1), iy, iz)], wm_xm = W[IX(clx(Z.ix0 - 1), ym, iz)];
    double Fa_x = 0.0;
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix) {
        const int x = clx(ix);
        const int c = IX(x, iy, iz);
        const int cm = IX(x, ym, iz);
        // every load of the position, issued before any is used
        const double d0 = db1[c], d_xp = db1[IX(clx(ix + 1), iy, iz)], d_yp = db1[IX(x, yp, iz)], d_m = db1[cm];
        const double w0 = W[c], wm = W[cm];
        const double cvc0 = A.cvc1[c], cv10 = A.cv1[c], cvcm = A.cvc1[cm], cv1m = A.cv1[cm];
        const double bx0 = B1x[c], by0 = B1y[c], bym = B1y[cm], bz0 = B1z[c];
        const double r0 = RHO[c];
        __builtin_amdgcn_sched_barrier(0);
        // bx flux (tw = sx, ow = -sy); by flux (tw = sy, ow = -sx) at (y)
        const double Fa = zflux(bx0, (d0 + d_xp) * 0.5, w0, wm, cvc0, cv10);
        const double Fb = zflux(by0, (d0 + d_yp) * 0.5, w0, w_xm, cvc0, cv10);
        // by flux at (y-1)
        const double Fbm = zflux(bym, (d_m + d0) * 0.5, wm, wm_xm, cvcm, cv1m);
        const double Faz = lane_prev(Fa), Fbz = lane_prev(Fb);
        // R5: the z mass flux from rho (current state copy) at z-1..z+2
        const double rp = lane_next(r0);
        const double rm = lane_prev(r0), rp2 = lane_next(rp);
        const double cvc_m = lane_prev(cvc0), cvc_p = lane_next(cvc0), cv1_p = lane_next(cv10);
        const double dxb = x_at(dxb_row, ix - Z.ix0);  // DXB(ix) wherever a lane writes
        if (Z.wz && ix > Z.ix0) {
            if (ix >= 0 && ix <= nx + 1 && iy >= 0 && iy <= ny + 1 && iz >= -1 && iz <= nz + 2) {
                double dmf = 0.0;
                if (iz >= 0 && iz <= nz) {
                    const double v_advect = (w_xm + w0 + wm_xm + wm) * 0.25;  // face_v<2>
                    const double o_v = v_advect * dt * (dxb * dyb);  // DZAB(ix, iy)
                    dmf = mass_flux_v(rm, r0, rp, rp2, v_advect, o_v, cvc_m, cvc0, cvc_p, cv10, cv1_p);
                }
                DM[c] = dmf;
            }
            const bool x1 = ix >= 1 && ix <= nx, x0 = ix >= 0 && ix <= nx;
            const bool y1 = iy >= 1 && iy <= ny, y0 = iy >= 0 && iy <= ny;
            const bool z1 = iz >= 1 && iz <= nz, z0 = iz >= 0 && iz <= nz;
            const bool ebx = iy >= -1 && iz >= -1, eby = ix >= -1 && iz >= -1, ebz = ix >= -1 && iy >= -1;
            double bx = ebx ? bx0 : 0.0;
            double by = eby ? by0 : 0.0;
            double bz = ebz ? bz0 : 0.0;
            if (x0 && y1 && z1) bx = bx - Fa + Faz;
            const bool bzr = x1 && y1 && z0;
            if (bzr) bz = bz + Fa - Fa_x;
            if (x1 && y0 && z1) by = by - Fb + Fbz;
            if (bzr) bz = bz + Fb - Fbm;
            // b_area<0..2>: DXAB(iy, iz), DYAB(ix, iz), DZAB(ix, iy)
            if (ebx) BX[c] = ddiv(bx, axx + DBL_EPSILON);
            if (eby) BY[c] = ddiv(by, dxb * dzb + DBL_EPSILON);
            if (ebz) BZ[c] = ddiv(bz, dxb * dyb + DBL_EPSILON);
        }
        Fa_x = Fa;
        w_xm = w0;
        wm_xm = wm;
    }
}

// The shifted sweep velocity the reference writes back into the registered
// v1 component at the end of the sweep's vertex stage (xremap.cpp:193-203,
// yremap.cpp:197-207, zremap.cpp:194-204), over (-2..n+1 along DIR, 0..n
// across): BACK = 0 stages (w[c] + w[c+sd]) * 0.5 in fl[0], BACK = 1 copies
// it into w.  Needed only where a v1 edge along DIR has no boundary
// condition: otherwise every shifted cell is rewritten (by the next
// predictor or by remap_v_bcs) before anything but a zero-weighted term of
// a later sweep reads it.
template <class T, int DIR, int BACK>
KERNEL void f_wshift(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const Fld<T> w = DIR == 0 ? VX1 : DIR == 1 ? VY1 : VZ1;
    if (BACK)
        w[c] = lare::gbl(A.fl[0])[c];
    else
        lare::gbl(A.fl[0])[c] = (w[c] + w[c + sdir<DIR>(G)]) * 0.5;
    };
    ECELL(cell)
}

// ============================================================= Lagrangian

#define PE_OF(c) ((gas_gamma - 1.0) * RHO[c] * EE[c])
#define PI_OF(c) ((gas_gamma - 1.0) * RHO[c] * EI[c])

// L1 (centred extent): sound speed (lagran.cpp:190-197) from the prologue
// values of lagran.cpp:113-125; resets the CFL minimum.  With resistive MHD
// it also stores p_e: the predictor reads the prologue value after the
// resistive heating changed e_e; and p_i when the kinetic-energy correction
// is on too (the resistive energy fill then changes e_i on ghost cells).
template <class T>
KERNEL void g_cs(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const double gas_gamma = P.gas_gamma;
    if (ix == R.x0 && iy == R.y0 && iz == R.z0) lare::gbl(A.s)->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
    const int c = IX(ix, iy, iz);
    double bx1 = 0.5 * (BX[c] + BX[c - G.sx]);
    double by1 = 0.5 * (BY[c] + BY[c - G.sy]);
    double bz1 = 0.5 * (BZ[c] + BZ[c - 1]);
    double pe = PE_OF(c), pi = PI_OF(c);
    double p = pe + pi;
    if (P.resistive) A.p_e[c] = pe;
    if (P.keep_pi) A.p_i[c] = pi;
    double rmin = pmax(RHO[c], DBL_EPSILON);
    double b2 = bx1 * bx1 + by1 * by1 + bz1 * bz1;
    A.cs[c] = sqrt(ddiv(gas_gamma * p + b2, rmin));
    };
    ECELL(cell)
}

#define CV8SUM(ix, iy, iz)                                                                               \
    (CV(ix, iy, iz) + CV(ix + 1, iy, iz) + CV(ix, iy + 1, iz) + CV(ix + 1, iy + 1, iz) + CV(ix, iy, iz + 1) + \
     CV(ix + 1, iy, iz + 1) + CV(ix, iy + 1, iz + 1) + CV(ix + 1, iy + 1, iz + 1))

// L2 (-1..n+1)^3: vertex density rho_v (lagran.cpp:128-153) and the vertex
// sound speed cs_v (lagran.cpp:199-216).
template <class T>
KERNEL void g_vertex(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double c000 = CV(ix, iy, iz), c100 = CV(ix + 1, iy, iz), c010 = CV(ix, iy + 1, iz),
                 c110 = CV(ix + 1, iy + 1, iz), c001 = CV(ix, iy, iz + 1), c101 = CV(ix + 1, iy, iz + 1),
                 c011 = CV(ix, iy + 1, iz + 1), c111 = CV(ix + 1, iy + 1, iz + 1);
    const Fld<T> r = RHO, s = A.cs;
    double sum_rho_cv = r[c] * c000 + r[c + sx] * c100 + r[c + sy] * c010 + r[c + sx + sy] * c110 + r[c + 1] * c001 +
                        r[c + sx + 1] * c101 + r[c + sy + 1] * c011 + r[c + sx + sy + 1] * c111;
    double sum_cv = c000 + c100 + c010 + c110 + c001 + c101 + c011 + c111;
    A.rho_v[c] = ddiv(sum_rho_cv, sum_cv);
    double cv_v = 0.125 * sum_cv;
    double sum = s[c] * c000 + s[c + sx] * c100 + s[c + sy] * c010 + s[c + sx + sy] * c110 + s[c + 1] * c001 +
                 s[c + sx + 1] * c101 + s[c + sy + 1] * c011 + s[c + sx + sy + 1] * c111;
    A.cs_v[c] = ddiv(0.125 * sum, cv_v);
    };
    ECELL(cell)
}

// L1 and L2 in one z-line x-march (one halo lane each end) over the
// centred extent (-1..n+2)^3: at each x a lane computes the sound speed of
// g_cs on its own cell column (y) and on column y+1, carries both to the
// next x, and forms the vertex density and vertex sound speed of g_vertex
// at vertex (x-1, y, z) from the two x planes (the values at z+1 from the
// lane above).  It stores p_e / p_i on its own cell as g_cs does and resets
// the CFL minimum; the sound speed is rounded to the storage type as g_cs's
// store rounds it, and never stored.  Each value is the expression of g_cs
// and g_vertex on the same operands; positions beyond the box are clamped
// and feed only values that are not stored.
template <class T, int K>
KERNEL OCC(4) void g_csv_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const Phys P = D_.p;
    const double gas_gamma = P.gas_gamma;
    const bool store_pe = P.resistive, store_pi = P.keep_pi;
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int sx = G.sx, sy = G.sy;
    const int iy = Z.iy, iz = Z.iz;  // iy >= -1 over this range, so y-1 stays in the box
    const int yn = iy + 1 > ny + 2 ? ny + 2 : iy + 1, zn = iz + 1 > nz + 2 ? nz + 2 : iz + 1;
    const int dn = (yn - iy) * sy;  // column y+1
    const bool vyz = iy <= ny + 1 && iz <= nz + 1;  // vertex range along y and z
    // CV(x, y', z') = DXB(x) * (DYB(y') * DZB(z')) for y' in y, y+1 and z' in z, z+1
    const double dy0 = DYB(iy), dy1 = DYB(yn), dz0 = DZB(iz), dz1 = DZB(zn);
    const double a00 = dy0 * dz0, a10 = dy1 * dz0, a01 = dy0 * dz1, a11 = dy1 * dz1;
    const double dx_row = x_row(A.dxb, Z.ix0, nx + 2);  // lane l: DXB(ix0 + l)
    auto cs_of = [&](double bx, double bxm, double by, double bym, double bz, double bzm, double r, double ee,
                     double ei, double &pe, double &pi) {
        const double bx1 = 0.5 * (bx + bxm), by1 = 0.5 * (by + bym), bz1 = 0.5 * (bz + bzm);
        pe = (gas_gamma - 1.0) * r * ee;
        pi = (gas_gamma - 1.0) * r * ei;
        const double p = pe + pi;
        const double rmin = pmax(r, DBL_EPSILON);
        const double b2 = bx1 * bx1 + by1 * by1 + bz1 * bz1;
        return (double)(T)sqrt(ddiv(gas_gamma * p + b2, rmin));
    };
    // plane x-1: bx of both columns, rho and cs of both columns and their
    // values at z+1
    const int xm0 = Z.ix0 - 1 < -2 ? -2 : Z.ix0 - 1;
    double bxm_o = BX[IX(xm0, iy, iz)], bxm_n = BX[IX(xm0, yn, iz)];
    double rm_o = 0.0, rm_n = 0.0, rm_oz = 0.0, rm_nz = 0.0, sm_o = 0.0, sm_n = 0.0, sm_oz = 0.0, sm_nz = 0.0;
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix) {
        const int c = IX(ix, iy, iz);  // ix >= -2 over this range
        // every load of the position, issued before any is used
        const double bxo = BX[c], bxn = BX[c + dn];
        const double bym = BY[c - sy], byo = BY[c], byn = BY[c + dn];
        const double bzo = BZ[c], bzn = BZ[c + dn];
        const double ro = RHO[c], rn = RHO[c + dn];
        const double eo = EE[c], en = EE[c + dn];
        const double io = EI[c], in = EI[c + dn];
        __builtin_amdgcn_sched_barrier(0);
        const double bzo_m = lane_prev(bzo), bzn_m = lane_prev(bzn);
        double pe, pi, pe_n, pi_n;
        const double so = cs_of(bxo, bxm_o, byo, bym, bzo, bzo_m, ro, eo, io, pe, pi);
        const double sn = cs_of(bxn, bxm_n, byn, byo, bzn, bzn_m, rn, en, in, pe_n, pi_n);
        const double ro_z = lane_next(ro), rn_z = lane_next(rn), so_z = lane_next(so), sn_z = lane_next(sn);
        if (Z.wz && ix > Z.ix0) {
            if (store_pe) A.p_e[c] = pe;
            if (store_pi) A.p_i[c] = pi;
            if (ix == S.r.x0 && iy == S.r.y0 && iz == S.r.z0) lare::gbl(A.s)->dtmin_bits = 0x7FEFFFFFFFFFFFFFull;
            if (vyz && ix >= 0 && ix <= nx + 2) {
                // vertex (x-1, y, z): corners c000 .. c111 of g_vertex
                const double dxm = x_at(dx_row, ix - 1 - Z.ix0), dx0 = x_at(dx_row, ix - Z.ix0);
                const double c000 = dxm * a00, c100 = dx0 * a00, c010 = dxm * a10, c110 = dx0 * a10,
                             c001 = dxm * a01, c101 = dx0 * a01, c011 = dxm * a11, c111 = dx0 * a11;
                const double sum_rho_cv = rm_o * c000 + ro * c100 + rm_n * c010 + rn * c110 + rm_oz * c001 +
                                          ro_z * c101 + rm_nz * c011 + rn_z * c111;
                const double sum_cv = c000 + c100 + c010 + c110 + c001 + c101 + c011 + c111;
                A.rho_v[c - sx] = ddiv(sum_rho_cv, sum_cv);
                const double cv_v = 0.125 * sum_cv;
                const double sum = sm_o * c000 + so * c100 + sm_n * c010 + sn * c110 + sm_oz * c001 + so_z * c101 +
                                   sm_nz * c011 + sn_z * c111;
                A.cs_v[c - sx] = ddiv(0.125 * sum, cv_v);
            }
        }
        bxm_o = bxo;
        bxm_n = bxn;
        rm_o = ro;
        rm_n = rn;
        rm_oz = ro_z;
        rm_nz = rn_z;
        sm_o = so;
        sm_n = sn;
        sm_oz = so_z;
        sm_nz = sn_z;
    }
}

// edge_viscosity (lagran.cpp:56-93) between c1 and c1 + s.
// The same from values: the velocity at the four points c0..c3 of the edge
// line (v[k][0..2] = vx, vy, vz) and rho_v at c1, c2.
__device__ __forceinline__ double edge_visc_v(double dt, double visc1, double visc2_norm, double dvdots, double dx,
                                              double dxm, double dxp, double cs_edge, double rv1, double rv2,
                                              const double *v0, const double *v1, const double *v2,
                                              const double *v3)
{
    dvdots = pmin(0.0, dvdots);
    double rho_edge = ddiv(2.0 * rv1 * rv2, rv1 + rv2);
    double dvx = v1[0] - v2[0];
    double dvy = v1[1] - v2[1];
    double dvz = v1[2] - v2[2];
    double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    double dv = sqrt(dv2);
    double psi = 0.0;
    const bool small = ddiv(dv * dt, dx) < 1.e-14;
    dvdots = small ? 0.0 : ddiv(dvdots, dv);
    double dvxm = v0[0] - v1[0];
    double dvxp = v2[0] - v3[0];
    double dvym = v0[1] - v1[1];
    double dvyp = v2[1] - v3[1];
    double dvzm = v0[2] - v1[2];
    double dvzp = v2[2] - v3[2];
    double rl = small ? 1.0 : ddiv((dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx, dxp * dv2);
    double rr = small ? 1.0 : ddiv((dvxm * dvx + dvym * dvy + dvzm * dvz) * dx, dxm * dv2);
    psi = pmin_list4(0.5 * (rr + rl), 2.0 * rl, 2.0 * rr, 1.0);
    psi = pmax(0.0, psi);
    double q_k_bar = rho_edge * (visc2_norm * dv + sqrt(visc2_norm * visc2_norm * dv2 + (visc1 * cs_edge) * (visc1 * cs_edge)));
    return q_k_bar * (1.0 - psi) * dvdots;
}

template <class T>
__device__ __forceinline__ double g_edge_visc(const ArrE<T> &A, double dt, double visc1, double visc2_norm,
                                              double dvdots, double dx, double dxm, double dxp, double cs_edge,
                                              int c1, int s)
{
    const int c0 = c1 - s, c2 = c1 + s, c3 = c2 + s;
    const double v0[3] = {VX[c0], VY[c0], VZ[c0]}, v1[3] = {VX[c1], VY[c1], VZ[c1]};
    const double v2[3] = {VX[c2], VY[c2], VZ[c2]}, v3[3] = {VX[c3], VY[c3], VZ[c3]};
    return edge_visc_v(dt, visc1, visc2_norm, dvdots, dx, dxm, dxp, cs_edge, A.rho_v[c1], A.rho_v[c2], v0, v1, v2,
                       v3);
}

// L3: alpha1, alpha2, alpha3 (lagran.cpp:211-274), each over its own range
// inside the union (-1..n+1, -1..n+2, 0..n+2).  dt is the previous step's.
template <class T>
KERNEL void g_alpha(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const double visc1 = P.visc1, visc2_norm = P.visc2_norm;
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const double dt = lare::gbl(A.s)->dt;
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
    };
    ECELL(cell)
}

#define SQD(p, q) \
    ((VX[p] - VX[q]) * (VX[p] - VX[q]) + (VY[p] - VY[q]) * (VY[p] - VY[q]) + (VZ[p] - VZ[q]) * (VZ[p] - VZ[q]))

// L4 over (0..n+1)^3: p_visc and visc_heat (lagran.cpp:279-323); on (0..n)^3
// also the viscous forces (330-376), stored as fx_visc, and the CFL minimum
// (set_dt, 380-443) reduced per wavefront, per workgroup and by a 64-bit
// atomicMin on the bit pattern (non-negative doubles order as integers; the
// minimum is exact, so the result does not depend on the order).
template <class T>
KERNEL OCC(4) void g_visc_dt(const Src<T> S)
{
    EGRID
    const double gas_gamma = P.gas_gamma, mu0 = P.mu0;
    auto cell = [&](const int ix, const int iy, const int iz) -> double {
        double m = DBL_MAX;
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
        A.visc_heat[c] = ddiv(vh, cv);
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
    ddiv(b1 * (V[c] - V[c - sx]) + b2 * (V[c] - V[c + sx]) + b3 * (V[c] - V[c - sy]) +              \
             b4 * (V[c] - V[c + sy]) + b5 * (V[c] - V[c - 1]) + b6 * (V[c] - V[c + 1]),             \
         cvv)
            A.fxv[c] = FV(VX);
            A.fyv[c] = FV(VY);
            A.fzv[c] = FV(VZ);
#undef FV
            // set_dt
            double rho0 = pmax(RHO[c], DBL_EPSILON);
            double press = PE_OF(c) + PI_OF(c);
            double cs2 = ddiv(gas_gamma * press, rho0);
            double w1 = ddiv(ddiv(BX[c] * BX[c] + BY[c] * BY[c] + BZ[c] * BZ[c], mu0), rho0);
            double c_visc2 = ddiv(pv, rho0);
            double length = pmin(pmin(dx, dy), dz);
            double dt1 = ddiv(length, sqrt(c_visc2) + sqrt(cs2 + w1 + c_visc2));
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
            double dt2 = ddiv(cv, pmax(pmax(pmax(fabs(vxbm), fabs(vxbp)), dvx), 1.0e-10 * cv));
            double dt3 = ddiv(cv, pmax(pmax(pmax(fabs(vybm), fabs(vybp)), dvy), 1.0e-10 * cv));
            double dt4 = ddiv(cv, pmax(pmax(pmax(fabs(vzbm), fabs(vzbp)), dvz), 1.0e-10 * cv));
            m = pmin(pmin(pmin(dt1, dt2), dt3), dt4);
        }
        return m;
    };
    double m = DBL_MAX;
    if (ok_) m = cell(ix_, iy_, iz_);
    for (int off = 32; off > 0; off >>= 1) m = pmin(m, __shfl_xor(m, off, 64));
    __shared__ double wmin[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = pmin(pmin(pmin(wmin[0], wmin[1]), wmin[2]), wmin[3]);
        if (b >= 0.0 && b < DBL_MAX) atomicMin(&A.s->dtmin_bits, (unsigned long long)__double_as_longlong(b));
    }
}

// L4 as a z-line x-march (H = 1): the same arithmetic as g_visc_dt.  Each
// lane carries the velocity columns (x, y) and (x, y-1) of the previous
// position and loads those of x + 1 ahead (the viscous force reads x + 1);
// the squared differences of the corner ring at z - 1 are those of the lane
// below at its z.
template <class T, int K>
KERNEL OCC(3) void g_visc_dt_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const Phys P = D_.p;
    const double gas_gamma = P.gas_gamma, mu0 = P.mu0;
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const int iy = Z.iy, iz = Z.iz;
    const Fld<T> V3[3] = {VX, VY, VZ};
    const double dy = DYB(iy), dz = DZB(iz), dyp = DYB(iy + 1), dzp = DZB(iz + 1);
    const bool fz = iy <= G.ny && iz <= G.nz;  // force and dt range, y and z
    const double ax = 0.25 * (dy * dz);  // DXAB(iy, iz)
    // DXB along the march: lane l holds DXB(ix0 + l)
    const double dx_row = x_row(A.dxb, Z.ix0, G.nx + 2);
    int c = IX(Z.ix0, iy, iz);
    // columns at own z: O = (x, y), Om = (x, y-1), M = (x-1, y), Mm = (x-1, y-1)
    double M[3] = {0, 0, 0}, Mm[3] = {0, 0, 0};
    double pM[3] = {0, 0, 0}, pMm[3] = {0, 0, 0};
    double a6_m = 0.0, a9_m = 0.0, a12_m = 0.0, vxbp_m = 0.0;
    double al2_m = 0.0, al2p_m = 0.0, al3_m = 0.0, al3y_m = 0.0;
    double al1y_n = A.alpha1[c + sy];  // alpha1 at (x, y+1), loaded ahead
    double m = DBL_MAX;
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        double O[3], Om[3], N[3], Yp[3], pO[3], pOm[3];
        for (int k = 0; k < 3; ++k) {
            O[k] = V3[k][c];
            Om[k] = V3[k][c - sy];
            N[k] = V3[k][c + sx];
            Yp[k] = V3[k][c + sy];
            pO[k] = lane_prev(O[k]);
            pOm[k] = lane_prev(Om[k]);
        }
        const double al1 = A.alpha1[c], al1y = al1y_n, al2 = A.alpha2[c], al2y = A.alpha2[c + sy];
        const double al3 = A.alpha3[c], al3y = A.alpha3[c - sy];
        al1y_n = A.alpha1[c + sx + sy];
        __builtin_amdgcn_sched_barrier(0);
        const double al1p = lane_next(al1), al1yp = lane_next(al1y), al2p = lane_next(al2);
        const double al1ynp = lane_next(al1y_n), al2yp = lane_next(al2y), al3p = lane_next(al3);
#define ZSQD(p, q) \
    (((p)[0] - (q)[0]) * ((p)[0] - (q)[0]) + ((p)[1] - (q)[1]) * ((p)[1] - (q)[1]) + ((p)[2] - (q)[2]) * ((p)[2] - (q)[2]))
        const double a5 = ZSQD(Mm, Om), a6 = ZSQD(Om, O), a7 = ZSQD(O, M), a8 = a6_m;
        const double a9 = ZSQD(pO, O), a12 = ZSQD(pOm, Om), a10 = a9_m, a11 = a12_m;
#undef ZSQD
        const double a1 = lane_prev(a5), a2 = lane_prev(a6), a3 = lane_prev(a7), a4 = lane_prev(a8);
        const double vxbp = (O[0] + Om[0] + pO[0] + pOm[0]) * ax;
        const double fO[3] = {lane_next(O[0]), lane_next(O[1]), lane_next(O[2])};
        const double dx = x_at(dx_row, ix - Z.ix0), dxp = x_at(dx_row, ix + 1 - Z.ix0);
        if (Z.wz && ix > Z.ix0) {
            double pv = 0.0;
            pv = pmax(pv, -al1 * sqrt(a1));
            pv = pmax(pv, -al2 * sqrt(a2));
            pv = pmax(pv, -al3 * sqrt(a9));
            double vh = -0.25 * dy * dz * al1 * a1 - 0.25 * dx * dz * al2 * a2 - 0.25 * dy * dz * al1y * a3 -
                        0.25 * dx * dz * al2_m * a4 - 0.25 * dy * dz * al1p * a5 - 0.25 * dx * dz * al2p * a6 -
                        0.25 * dy * dz * al1yp * a7 - 0.25 * dx * dz * al2p_m * a8 - 0.25 * dx * dy * al3 * a9 -
                        0.25 * dx * dy * al3_m * a10 - 0.25 * dx * dy * al3y_m * a11 - 0.25 * dx * dy * al3y * a12;
            const double cv = dx * (dy * dz);  // CV(ix, iy, iz)
            A.visc_heat[c] = ddiv(vh, cv);
            if (fz && ix <= G.nx) {
                // the cell's state for the time-step bound, loaded before the
                // viscous forces and first used after them
                const double rho_c = RHO[c], ee_c = EE[c], ei_c = EI[c], bx_c = BX[c], by_c = BY[c], bz_c = BZ[c];
                __builtin_amdgcn_sched_barrier(0);
                double b1 = al1yp * dy * dz;
                double b2 = al1ynp * dy * dz;
                double b3 = al2p * dx * dz;
                double b4 = al2yp * dx * dz;
                double b5 = al3 * dx * dy;
                double b6 = al3p * dx * dy;
                // CV8SUM(ix, iy, iz)
                double cvv = 0.125 * (dx * (dy * dz) + dxp * (dy * dz) + dx * (dyp * dz) + dxp * (dyp * dz) +
                                      dx * (dy * dzp) + dxp * (dy * dzp) + dx * (dyp * dzp) + dxp * (dyp * dzp));
                double fv[3];
                for (int k = 0; k < 3; ++k)
                    fv[k] = ddiv(b1 * (O[k] - M[k]) + b2 * (O[k] - N[k]) + b3 * (O[k] - Om[k]) + b4 * (O[k] - Yp[k]) +
                             b5 * (O[k] - pO[k]) + b6 * (O[k] - fO[k]), cvv);
                A.fxv[c] = fv[0];
                A.fyv[c] = fv[1];
                A.fzv[c] = fv[2];
                __builtin_amdgcn_sched_barrier(0);
                double rho0 = pmax(rho_c, DBL_EPSILON);
                double press = (gas_gamma - 1.0) * rho_c * ee_c + (gas_gamma - 1.0) * rho_c * ei_c;  // PE_OF + PI_OF
                double cs2 = ddiv(gas_gamma * press, rho0);
                double w1 = ddiv(ddiv(bx_c * bx_c + by_c * by_c + bz_c * bz_c, mu0), rho0);
                double c_visc2 = ddiv(pv, rho0);
                double length = pmin(pmin(dx, dy), dz);
                double dt1 = ddiv(length, sqrt(c_visc2) + sqrt(cs2 + w1 + c_visc2));
                double ay = 0.25 * (dx * dz);  // DYAB(ix, iz)
                double az = 0.25 * (dx * dy);  // DZAB(ix, iy)
                double vxbm = vxbp_m;
                double vybm = (Om[1] + Mm[1] + pOm[1] + pMm[1]) * ay;
                double vybp = (O[1] + M[1] + pO[1] + pM[1]) * ay;
                double vzbm = (pO[2] + pM[2] + pOm[2] + pMm[2]) * az;
                double vzbp = (O[2] + M[2] + Om[2] + Mm[2]) * az;
                double dvx = fabs(vxbp - vxbm), dvy = fabs(vybp - vybm), dvz = fabs(vzbp - vzbm);
                double dt2 = ddiv(cv, pmax(pmax(pmax(fabs(vxbm), fabs(vxbp)), dvx), 1.0e-10 * cv));
                double dt3 = ddiv(cv, pmax(pmax(pmax(fabs(vybm), fabs(vybp)), dvy), 1.0e-10 * cv));
                double dt4 = ddiv(cv, pmax(pmax(pmax(fabs(vzbm), fabs(vzbp)), dvz), 1.0e-10 * cv));
                m = pmin(m, pmin(pmin(pmin(dt1, dt2), dt3), dt4));
            }
        }
        for (int k = 0; k < 3; ++k) {
            M[k] = O[k];
            Mm[k] = Om[k];
            pM[k] = pO[k];
            pMm[k] = pOm[k];
        }
        a6_m = a6;
        a9_m = a9;
        a12_m = a12;
        vxbp_m = vxbp;
        al2_m = al2;
        al2p_m = al2p;
        al3_m = al3;
        al3y_m = al3y;
    }
    for (int off = 32; off > 0; off >>= 1) m = pmin(m, __shfl_xor(m, off, 64));
    __shared__ double wmin[4];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = pmin(pmin(pmin(wmin[0], wmin[1]), wmin[2]), wmin[3]);
        if (b >= 0.0 && b < DBL_MAX) atomicMin(&A.s->dtmin_bits, (unsigned long long)__double_as_longlong(b));
    }
}

// L5 per grid: dt = dt_multiplier * min (lagran.cpp:380-443), the runner's
// minimum over packages (runner.h:158-169: one package here), dtr = dt and
// the resistive substep dt = dt / (int(dt/dtr)+1) = dt / 2 (159-161).
template <class T>
KERNEL void g_dt_final(const EDev<T> *devs, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    Scal *s = devs[i].a.s;
    double red = __longlong_as_double((long long)s->dtmin_bits);
    double dt = devs[i].p.dt_multiplier * red;
    double tdt = DBL_MAX;
    tdt = dt < tdt ? dt : tdt;
    s->dt = tdt;
    s->dtr = dt;
    s->dt_sub = tdt / 2.0;
}

// End of a step: count it.
template <class T>
KERNEL void g_step_done(const EDev<T> *devs, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) devs[i].a.s->nstep += 1;
}

// L6 over the centred extent: b_field_and_cv1_update (lagran.cpp:806-882)
// with bx1 *= cv1 (598-601); on (0..n+1)^3 also the energy/pressure
// predictor (604-614), storing the predicted p_e and p_i.  With resistive
// MHD, bx1..bz1 hold the copy of B made by the resistive substeps
// (lagran.cpp:482) and p_e the prologue value.
template <class T>
KERNEL void g_predict(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const double gas_gamma = P.gas_gamma;
    const bool res = P.resistive != 0;
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
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
    double dvxdx = ddiv(vxb * axx - vxbm * axx, vol);
    double dvydy = ddiv(vyb * ayy - vybm * ayy, vol);
    double dvzdz = ddiv(vzb * azz - vzbm * azz, vol);
    double dv = (dvxdx + dvydy + dvzdz) * dt / 2.0;
    const double cv1 = vol * (1.0 + dv);
    A.cv1[c] = cv1;
    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[oom] + VX[mom]);
    vxbm = 0.25 * (VX[omo] + VX[mmo] + VX[omm] + VX[mmm]);
    vyb = 0.25 * (VY[ooo] + VY[omo] + VY[oom] + VY[omm]);
    vybm = 0.25 * (VY[moo] + VY[mmo] + VY[mom] + VY[mmm]);
    double dvxdy = ddiv(vxb * ayy - vxbm * ayy, vol);
    double dvydx = ddiv(vyb * axx - vybm * axx, vol);
    vxb = 0.25 * (VX[ooo] + VX[moo] + VX[omo] + VX[mmo]);
    vxbm = 0.25 * (VX[oom] + VX[mom] + VX[omm] + VX[mmm]);
    double vzb2 = 0.25 * (VZ[ooo] + VZ[omo] + VZ[oom] + VZ[omm]);
    double vzbm2 = 0.25 * (VZ[moo] + VZ[mmo] + VZ[mom] + VZ[mmm]);
    double dvxdz = ddiv(vxb * azz - vxbm * azz, vol);
    double dvzdx = ddiv(vzb2 * axx - vzbm2 * axx, vol);
    double vyb3 = 0.25 * (VY[ooo] + VY[moo] + VY[omo] + VY[mmo]);
    double vybm3 = 0.25 * (VY[oom] + VY[mom] + VY[omm] + VY[mmm]);
    double vzb3 = 0.25 * (VZ[ooo] + VZ[moo] + VZ[oom] + VZ[mom]);
    double vzbm3 = 0.25 * (VZ[omo] + VZ[mmo] + VZ[omm] + VZ[mmm]);
    double dvydz = ddiv(vyb3 * azz - vybm3 * azz, vol);
    double dvzdy = ddiv(vzb3 * ayy - vzbm3 * ayy, vol);
    double bx1, by1, bz1;
    if (res) {
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
    A.bx1[c] = ddiv(bx1 + w3 * dt / 2.0, 1.0 + dv) * cv1;
    A.by1[c] = ddiv(by1 + w4 * dt / 2.0, 1.0 + dv) * cv1;
    A.bz1[c] = ddiv(bz1 + w5 * dt / 2.0, 1.0 + dv) * cv1;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix <= G.nx + 1 && iy <= G.ny + 1 && iz <= G.nz + 1) {
        double rho = RHO[c];
        double dvp = ddiv(cv1, vol) - 1.0;
        double pe0 = res ? (double)A.p_e[c] : PE_OF(c);
        double pi0 = P.keep_pi ? (double)A.p_i[c] : PI_OF(c);
        double e1_e = EE[c] - ddiv(pe0 * dvp, rho);
        double e1_i = EI[c] - ddiv(pi0 * dvp, rho);
        e1_i += ddiv(A.visc_heat[c] * dt / 2.0, rho);
        A.p_e[c] = ddiv(e1_e * (gas_gamma - 1.0) * rho * vol, cv1);
        A.p_i[c] = ddiv(e1_i * (gas_gamma - 1.0) * rho * vol, cv1);
    }
    };
    ECELL(cell)
}

// b_field_and_cv1_update and the energy predictor of g_predict at one cell,
// in two parts: predict_grad, the velocity gradients from the eight v
// corners W[k][p] (p: ooo, moo, omo, oom, mmo, mom, omm, mmm), dv and cv1;
// predict_rest, the B and energy predictor from them, the cell's
// face-averaged B (bx1..bz1: the resistive copy or the average of its faces),
// its rho, e_e, e_i, the predictor's p_e and p_i inputs pe0, pi0 and
// visc_heat.  vol is the cell volume and axx, ayy, azz its face areas (CV,
// DXAB, DYAB, DZAB).  out = cv1, bx1..bz1 (times cv1), p_e, p_i.
struct PredGrad {
    double dvxdx, dvydy, dvzdz, dvxdy, dvydx, dvxdz, dvzdx, dvydz, dvzdy, dv, cv1;
};
__device__ __forceinline__ PredGrad predict_grad(double vol, double axx, double ayy, double azz,
                                                 const double W[3][8], double dt)
{
    enum { ooo, moo, omo, oom, mmo, mom, omm, mmm };
    const double *X = W[0], *Y = W[1], *Zv = W[2];
    PredGrad g;
    double vxb = 0.25 * (X[ooo] + X[omo] + X[oom] + X[omm]);
    double vxbm = 0.25 * (X[moo] + X[mmo] + X[mom] + X[mmm]);
    double vyb = 0.25 * (Y[ooo] + Y[moo] + Y[oom] + Y[mom]);
    double vybm = 0.25 * (Y[omo] + Y[mmo] + Y[omm] + Y[mmm]);
    double vzb = 0.25 * (Zv[ooo] + Zv[moo] + Zv[omo] + Zv[mmo]);
    double vzbm = 0.25 * (Zv[oom] + Zv[mom] + Zv[omm] + Zv[mmm]);
    g.dvxdx = ddiv(vxb * axx - vxbm * axx, vol);
    g.dvydy = ddiv(vyb * ayy - vybm * ayy, vol);
    g.dvzdz = ddiv(vzb * azz - vzbm * azz, vol);
    g.dv = (g.dvxdx + g.dvydy + g.dvzdz) * dt / 2.0;
    g.cv1 = vol * (1.0 + g.dv);
    vxb = 0.25 * (X[ooo] + X[moo] + X[oom] + X[mom]);
    vxbm = 0.25 * (X[omo] + X[mmo] + X[omm] + X[mmm]);
    vyb = 0.25 * (Y[ooo] + Y[omo] + Y[oom] + Y[omm]);
    vybm = 0.25 * (Y[moo] + Y[mmo] + Y[mom] + Y[mmm]);
    g.dvxdy = ddiv(vxb * ayy - vxbm * ayy, vol);
    g.dvydx = ddiv(vyb * axx - vybm * axx, vol);
    vxb = 0.25 * (X[ooo] + X[moo] + X[omo] + X[mmo]);
    vxbm = 0.25 * (X[oom] + X[mom] + X[omm] + X[mmm]);
    double vzb2 = 0.25 * (Zv[ooo] + Zv[omo] + Zv[oom] + Zv[omm]);
    double vzbm2 = 0.25 * (Zv[moo] + Zv[mmo] + Zv[mom] + Zv[mmm]);
    g.dvxdz = ddiv(vxb * azz - vxbm * azz, vol);
    g.dvzdx = ddiv(vzb2 * axx - vzbm2 * axx, vol);
    double vyb3 = 0.25 * (Y[ooo] + Y[moo] + Y[omo] + Y[mmo]);
    double vybm3 = 0.25 * (Y[oom] + Y[mom] + Y[omm] + Y[mmm]);
    double vzb3 = 0.25 * (Zv[ooo] + Zv[moo] + Zv[oom] + Zv[mom]);
    double vzbm3 = 0.25 * (Zv[omo] + Zv[mmo] + Zv[omm] + Zv[mmm]);
    g.dvydz = ddiv(vyb3 * azz - vybm3 * azz, vol);
    g.dvzdy = ddiv(vzb3 * ayy - vzbm3 * ayy, vol);
    return g;
}
__device__ __forceinline__ void predict_rest(double gas_gamma, double vol, const PredGrad &g, double bx1,
                                             double by1, double bz1, double rho, double ee, double ei, double pe0,
                                             double pi0, double vh, double dt, double out[6])
{
    double w3 = bx1 * g.dvxdx + by1 * g.dvxdy + bz1 * g.dvxdz;
    double w4 = bx1 * g.dvydx + by1 * g.dvydy + bz1 * g.dvydz;
    double w5 = bx1 * g.dvzdx + by1 * g.dvzdy + bz1 * g.dvzdz;
    const double cv1 = g.cv1, dv = g.dv;
    out[0] = cv1;
    out[1] = ddiv(bx1 + w3 * dt / 2.0, 1.0 + dv) * cv1;
    out[2] = ddiv(by1 + w4 * dt / 2.0, 1.0 + dv) * cv1;
    out[3] = ddiv(bz1 + w5 * dt / 2.0, 1.0 + dv) * cv1;
    double dvp = ddiv(cv1, vol) - 1.0;
    double e1_e = ee - ddiv(pe0 * dvp, rho);
    double e1_i = ei - ddiv(pi0 * dvp, rho);
    e1_i += ddiv(vh * dt / 2.0, rho);
    out[4] = ddiv(e1_e * (gas_gamma - 1.0) * rho * vol, cv1);
    out[5] = ddiv(e1_i * (gas_gamma - 1.0) * rho * vol, cv1);
}

// L6 as a z-line x-march (H = 1): the same arithmetic as g_predict.  The
// eight corners (x-1..x, y-1..y, z-1..z) of the velocities are this lane's
// columns (x, y) and (x, y-1), the previous position's, and the lane below.
// The velocities of a position are loaded one position ahead (the last
// position reloads its own), so their latency overlaps the previous
// position's arithmetic; the cell's other operands are loaded at the start
// of its position and first used after the velocity gradients.
template <class T, int K>
KERNEL OCC(3) void g_predict_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const Phys P = D_.p;
    const double gas_gamma = P.gas_gamma;
    const bool res = P.resistive != 0;
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> V3[3] = {VX, VY, VZ};
    enum { ooo, moo, omo, oom, mmo, mom, omm, mmm };
    // face-averaged B: the resistive copy bx1..bz1, or the average of B's
    // faces (x - 1 from the previous position)
    const Fld<T> B1x = res ? A.bx1 : BX, B1y = res ? A.by1 : BY, B1z = res ? A.bz1 : BZ;
    const int zm = res ? 0 : 1;
    const double dyb = DYB(iy), dzb = DZB(iz), axx = dyb * dzb;
    const double dxb_row = x_row(A.dxb, Z.ix0, G.nx + 2);
    double M[3] = {0, 0, 0}, Mm[3] = {0, 0, 0}, pM[3] = {0, 0, 0}, pMm[3] = {0, 0, 0};
    double bx_m = 0.0;
    int c = IX(Z.ix0, iy, iz);
    double nO[3], nOm[3];
    for (int k = 0; k < 3; ++k) {
        nO[k] = V3[k][c];
        nOm[k] = V3[k][c - sy];
    }
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        double O[3], Om[3], pO[3], pOm[3];
        for (int k = 0; k < 3; ++k) {
            O[k] = nO[k];
            Om[k] = nOm[k];
        }
        const double x1 = B1x[c], y1 = B1y[c], y1m = B1y[c - sy], z1 = B1z[c], z1m = B1z[c - zm];
        const double rho = RHO[c], ee = EE[c], ei = EI[c], pe_a = A.p_e[c], pi_a = A.p_i[c], vh = A.visc_heat[c];
        {
            const int n = ix < Z.ix1 ? c + sx : c;
            for (int k = 0; k < 3; ++k) {
                nO[k] = V3[k][n];
                nOm[k] = V3[k][n - sy];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        for (int k = 0; k < 3; ++k) {
            pO[k] = lane_prev(O[k]);
            pOm[k] = lane_prev(Om[k]);
        }
        const double dxb = x_at(dxb_row, ix - Z.ix0);
        if (ix > Z.ix0 && Z.wz) {  // the warm-up position only loads
            double W[3][8];
            for (int k = 0; k < 3; ++k) {
                W[k][ooo] = O[k];
                W[k][moo] = M[k];
                W[k][omo] = Om[k];
                W[k][oom] = pO[k];
                W[k][mmo] = Mm[k];
                W[k][mom] = pM[k];
                W[k][omm] = pOm[k];
                W[k][mmm] = pMm[k];
            }
            const double vol = dxb * axx, ayy = dxb * dzb, azz = dxb * dyb;
            const PredGrad gr = predict_grad(vol, axx, ayy, azz, W, dt);
            // the cell's own operands are first used here: their loads have
            // had the gradients' time to land
            __builtin_amdgcn_sched_barrier(0);
            double bx1, by1, bz1;
            if (res) {
                bx1 = x1;
                by1 = y1;
                bz1 = z1;
            } else {
                bx1 = 0.5 * (x1 + bx_m);
                by1 = 0.5 * (y1 + y1m);
                bz1 = 0.5 * (z1 + z1m);
            }
            // the energy predictor covers (0..n+1)^3; elsewhere p_e and p_i
            // are stored back unchanged
            const bool energy = ix >= 0 && iy >= 0 && iz >= 0 && ix <= G.nx + 1 && iy <= G.ny + 1 && iz <= G.nz + 1;
            const double pe0 = res ? pe_a : (gas_gamma - 1.0) * rho * ee;
            const double pi0 = P.keep_pi ? pi_a : (gas_gamma - 1.0) * rho * ei;
            double o[6];
            predict_rest(gas_gamma, vol, gr, bx1, by1, bz1, rho, ee, ei, pe0, pi0, vh, dt, o);
            A.cv1[c] = o[0];
            A.bx1[c] = o[1];
            A.by1[c] = o[2];
            A.bz1[c] = o[3];
            A.p_e[c] = energy ? o[4] : pe_a;
            A.p_i[c] = energy ? o[5] : pi_a;
        }
        for (int k = 0; k < 3; ++k) {
            M[k] = O[k];
            Mm[k] = Om[k];
            pM[k] = pO[k];
            pMm[k] = pOm[k];
        }
        bx_m = x1;
    }
}

// L7 over (0..n)^3: forces and half-step velocities (lagran.cpp:619-731),
// and the corrector's final velocities (lagran.cpp:754-758: v + dt (fx_visc
// + fx) / rho_v, into the second state copy), which read nothing the
// corrector changes before them: the same values, without storing fx_visc +
// fx for the corrector to read back.  fx_visc..fz_visc are dead after this.
template <class T>
KERNEL void g_forces(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const double mu0 = P.mu0;
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> PE = A.p_e, PI = A.p_i, C1 = A.cv1, B1x = A.bx1, B1y = A.by1, B1z = A.bz1;
    double dx = DXC(ix), dy = DYC(iy), dz = DZC(iz);
    double dhy = dy, dhz = dz;
    const int x_ = c + sx, y_ = c + sy, z_ = c + 1, xy = c + sx + sy, xz = c + sx + 1, yz = c + sy + 1,
              xyz = c + sx + sy + 1;
    double pp = PE[c] + PI[c], ppx = PE[x_] + PI[x_], ppy = PE[y_] + PI[y_], ppz = PE[z_] + PI[z_];
    double ppxy = PE[xy] + PI[xy], ppxz = PE[xz] + PI[xz], ppyz = PE[yz] + PI[yz], ppxyz = PE[xyz] + PI[xyz];
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
    w1 = ddiv(B1z[c] + B1z[x_] + B1z[z_] + B1z[xz], cvy);
    w2 = ddiv(B1z[y_] + B1z[xy] + B1z[yz] + B1z[xyz], cvyp);
    double jx = ddiv(w2 - w1, dhy);
    w1 = ddiv(B1y[c] + B1y[x_] + B1y[y_] + B1y[xy], cvz);
    w2 = ddiv(B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz], cvzp);
    jx -= ddiv(w2 - w1, dhz);
    w1 = ddiv(B1z[c] + B1z[y_] + B1z[z_] + B1z[yz], cvx);
    w2 = ddiv(B1z[x_] + B1z[xy] + B1z[xz] + B1z[xyz], cvxp);
    double jy = ddiv(-(w2 - w1), dx);
    w1 = ddiv(B1x[c] + B1x[x_] + B1x[y_] + B1x[xy], cvz);
    w2 = ddiv(B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz], cvzp);
    jy += ddiv(w2 - w1, dhz);
    w1 = ddiv(B1y[c] + B1y[y_] + B1y[z_] + B1y[yz], cvx);
    w2 = ddiv(B1y[x_] + B1y[xy] + B1y[xz] + B1y[xyz], cvxp);
    double jz = ddiv(w2 - w1, dx);
    w1 = ddiv(B1x[c] + B1x[x_] + B1x[z_] + B1x[xz], cvy);
    w2 = ddiv(B1x[y_] + B1x[xy] + B1x[yz] + B1x[xyz], cvyp);
    jz -= ddiv(w2 - w1, dhy);
    double bxv = ddiv(B1x[c] + B1x[x_] + B1x[y_] + B1x[xy] + B1x[z_] + B1x[xz] + B1x[yz] + B1x[xyz], cvx + cvxp);
    double byv = ddiv(B1y[c] + B1y[x_] + B1y[y_] + B1y[xy] + B1y[z_] + B1y[xz] + B1y[yz] + B1y[xyz], cvx + cvxp);
    double bzv = ddiv(B1z[c] + B1z[x_] + B1z[y_] + B1z[xy] + B1z[z_] + B1z[xz] + B1z[yz] + B1z[xyz], cvx + cvxp);
    fx += ddiv(jy * bzv - jz * byv, mu0);
    fy += ddiv(jz * bxv - jx * bzv, mu0);
    fz += ddiv(jx * byv - jy * bxv, mu0);
    double rv = A.rho_v[c];
    double sxv = A.fxv[c] + fx, syv = A.fyv[c] + fy, szv = A.fzv[c] + fz;
    const double vx = VX[c], vy = VY[c], vz = VZ[c];
    VX1[c] = vx + ddiv(dt / 2.0 * sxv, rv);
    VY1[c] = vy + ddiv(dt / 2.0 * syv, rv);
    VZ1[c] = vz + ddiv(dt / 2.0 * szv, rv);
    // the sums as the storage type holds them (the corrector's operands)
    sxv = (double)(T)sxv;
    syv = (double)(T)syv;
    szv = (double)(T)szv;
    A.vx2[c] = vx + ddiv(dt * sxv, rv);
    A.vy2[c] = vy + ddiv(dt * syv, rv);
    A.vz2[c] = vz + ddiv(dt * szv, rv);
    };
    ECELL(cell)
}

#define DOT(p, q)                                                                                   \
    (1 * ((VX[p] - VX[q]) * (VX1[p] - VX1[q]) + (VY[p] - VY[q]) * (VY1[p] - VY1[q]) +               \
          (VZ[p] - VZ[q]) * (VZ1[p] - VZ1[q])))

// C over (0..n)^3: corrector_step (lagran.cpp:739-804) into the second state
// copy, on cells 1..n: shock_heating (884-993, with the dy*dy factors of a9
// and a11 at 987) and the density and energy correction (762-796); the
// velocity correction (754-758) is g_forces'.  The corrector's visc_heat is read only by the energy
// correction of the same cell, so it stays in a register; over the rest of
// shock_heating's range (0..n+1) nothing reads it before the next step's
// viscosity recomputes it.  The cv1 the reference also writes here is
// recomputed by the remap before use.
template <class T>
KERNEL OCC(2) void g_correct(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const double dt = lare::gbl(A.s)->dt;
    if (ix < 1 || iy < 1 || iz < 1) return;
    const int ooo = c, moo = c - sx, omo = c - sy, mmo = c - sx - sy;
    const int oop = c + 1, mop = c - sx + 1, omp = c - sy + 1, mmp = c - sx - sy + 1;
    const int oom = c - 1, mom = c - sx - 1, omm = c - sy - 1, mmm = c - sx - sy - 1;
    double visc_heat;
    {
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
        visc_heat = pmax(ddiv(vh, CV(ix, iy, iz)), 0.0);
    }
    double vxb = 0.25 * (VX1[ooo] + VX1[omo] + VX1[oom] + VX1[omm]);
    double vxbm = 0.25 * (VX1[moo] + VX1[mmo] + VX1[mom] + VX1[mmm]);
    double vyb = 0.25 * (VY1[ooo] + VY1[moo] + VY1[oom] + VY1[mom]);
    double vybm = 0.25 * (VY1[omo] + VY1[mmo] + VY1[omm] + VY1[mmm]);
    double vzb = 0.25 * (VZ1[ooo] + VZ1[moo] + VZ1[omo] + VZ1[mmo]);
    double vzbm = 0.25 * (VZ1[oom] + VZ1[mom] + VZ1[omm] + VZ1[mmm]);
    const double vol = CV(ix, iy, iz);
    const double axx = DXAB(iy, iz), ayy = DYAB(ix, iz), azz = DZAB(ix, iy);
    double dvxdx = ddiv(vxb * axx - vxbm * axx, vol);
    double dvydy = ddiv(vyb * ayy - vybm * ayy, vol);
    double dvzdz = ddiv(vzb * azz - vzbm * azz, vol);
    double dv = (dvxdx + dvydy + dvzdz) * dt;
    double rho = RHO[c];
    A.ee2[c] = EE[c] - ddiv(dv * A.p_e[c], rho);
    A.ei2[c] = EI[c] + ddiv(dt * visc_heat - dv * A.p_i[c], rho);
    A.rho2[c] = ddiv(rho, 1.0 + dv);
    };
    ECELL(cell)
}

// C (corrector) as a z-line x-march (H = 1): the same arithmetic as g_correct.
// Corner values of the shock heating at (x, y, z) come from this lane (z),
// the lanes at z -+ 1 and the previous position (x - 1); each DOT product
// is formed once and reused by the cells that share it (a DOT of two
// points is the same value in either order: both differences change sign).
template <class T, int K>
KERNEL OCC(3) void g_correct_zx(const Src<T> S)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt;
    const Fld<T> V6[6] = {VX, VY, VZ, VX1, VY1, VZ1};
    const double dy = DYB(iy), dz = DZB(iz), dyz = dy * dz;
    const double dx_row = x_row(A.dxb, Z.ix0, G.nx + 2);
    const bool cyz = iy >= 1 && iz >= 1;
    // plane x - 1: column (x-1, y) = M, (x-1, y-1) = Mm at z; their z-1
    // values of VY1, VZ1; the carried products and coefficients
    double M[6] = {0, 0, 0, 0, 0, 0}, Mm[6] = {0, 0, 0, 0, 0, 0};
    double pM_vy1 = 0.0, pM_vz1 = 0.0, pMm_vy1 = 0.0, pMm_vz1 = 0.0;
    double a2_m = 0.0, a9_m = 0.0, a12_m = 0.0, vxb_m = 0.0;
    double al2_m = 0.0, al2p_m = 0.0, al3_m = 0.0, al3y_m = 0.0;
    int c = IX(Z.ix0, iy, iz);
    // the velocities of a position are loaded one position ahead (the last
    // position reloads its own); the alphas and the cell's energies are
    // loaded at the start of its position, the energies first used after
    // its viscous heating and velocity gradients
    double nO[6], nOm[6];
    auto fetch = [&](int n) {
        for (int k = 0; k < 6; ++k) {
            nO[k] = V6[k][n];
            nOm[k] = V6[k][n - sy];
        }
    };
    fetch(c);
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        double O[6], Om[6], pO[6], pOm[6];
        for (int k = 0; k < 6; ++k) {
            O[k] = nO[k];
            Om[k] = nOm[k];
        }
        const double al1 = A.alpha1[c], al1y = A.alpha1[c + sy], al2 = A.alpha2[c], al3 = A.alpha3[c],
                     al3y = A.alpha3[c - sy];
        const double rho = RHO[c], ee = EE[c], ei = EI[c], pe = A.p_e[c], pi = A.p_i[c];
        fetch(ix < Z.ix1 ? c + sx : c);
        __builtin_amdgcn_sched_barrier(0);
        for (int k = 0; k < 6; ++k) {
            pO[k] = lane_prev(O[k]);
            pOm[k] = lane_prev(Om[k]);
        }
        const double al1p = lane_next(al1), al1yp = lane_next(al1y), al2p = lane_next(al2);
#define ZDOT(p, q) \
    (1 * (((p)[0] - (q)[0]) * ((p)[3] - (q)[3]) + ((p)[1] - (q)[1]) * ((p)[4] - (q)[4]) + ((p)[2] - (q)[2]) * ((p)[5] - (q)[5])))
        const double a1 = ZDOT(Mm, Om), a2 = ZDOT(Om, O), a3 = ZDOT(O, M), a4 = a2_m;
        const double a9 = ZDOT(pO, O), a12 = ZDOT(pOm, Om), a10 = a9_m, a11 = a12_m;
#undef ZDOT
        const double a5 = lane_next(a1), a6 = lane_next(a2), a7 = lane_next(a3), a8 = lane_next(a4);
        const double vxb = 0.25 * (O[3] + Om[3] + pO[3] + pOm[3]);
        const double dx = x_at(dx_row, ix - Z.ix0);
        if (Z.wz && ix > Z.ix0 && ix >= 1 && cyz) {
            const double vh =
                (-0.25 * dy * dz * al1 * a1 - 0.25 * dx * dz * al2 * a2 - 0.25 * dy * dz * al1y * a3 -
                 0.25 * dx * dz * al2_m * a4 - 0.25 * dy * dz * al1p * a5 - 0.25 * dx * dz * al2p * a6 -
                 0.25 * dy * dz * al1yp * a7 - 0.25 * dx * dz * al2p_m * a8 - 0.25 * dy * dy * al3 * a9 -
                 0.25 * dx * dy * al3_m * a10 - 0.25 * dy * dy * al3y_m * a11 - 0.25 * dx * dy * al3y * a12);
            // CV, DXAB, DYAB, DZAB from the 1D deltas
            const double vol = dx * dyz;
            const double visc_heat = pmax(ddiv(vh, vol), 0.0);
            const double vxbm = vxb_m;
            const double vyb = 0.25 * (O[4] + M[4] + pO[4] + pM_vy1);
            const double vybm = 0.25 * (Om[4] + Mm[4] + pOm[4] + pMm_vy1);
            const double vzb = 0.25 * (O[5] + M[5] + Om[5] + Mm[5]);
            const double vzbm = 0.25 * (pO[5] + pM_vz1 + pOm[5] + pMm_vz1);
            const double axx = dyz, ayy = dx * dz, azz = dx * dy;
            double dvxdx = ddiv(vxb * axx - vxbm * axx, vol);
            double dvydy = ddiv(vyb * ayy - vybm * ayy, vol);
            double dvzdz = ddiv(vzb * azz - vzbm * azz, vol);
            double dv = (dvxdx + dvydy + dvzdz) * dt;
            // the energies are first used here
            __builtin_amdgcn_sched_barrier(0);
            A.ee2[c] = ee - ddiv(dv * pe, rho);
            A.ei2[c] = ei + ddiv(dt * visc_heat - dv * pi, rho);
            A.rho2[c] = ddiv(rho, 1.0 + dv);
        }
        for (int k = 0; k < 6; ++k) {
            M[k] = O[k];
            Mm[k] = Om[k];
        }
        pM_vy1 = pO[4];
        pM_vz1 = pO[5];
        pMm_vy1 = pOm[4];
        pMm_vz1 = pOm[5];
        a2_m = a2;
        a9_m = a9;
        a12_m = a12;
        vxb_m = vxb;
        al2_m = al2;
        al2p_m = al2p;
        al3_m = al3;
        al3y_m = al3y;
    }
}

// ------------------------------------------------- resistive substeps

// The copy of B into bx1..bz1 over the centred extent at the start of a
// resistive substep: the reference's pw::assign(data.bx1, data.bx) at
// lagran.cpp:482 throws on the extent mismatch; this is its element-wise
// reading over bx1's extent.  Only the last substep's copy is read later (by
// the predictor): within a substep bstep reads bx1 only at the cell it
// writes, where it still equals B, so earlier substeps skip the copy.
template <class T>
KERNEL void r_copy(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    A.bx1[c] = BX[c];
    A.by1[c] = BY[c];
    A.bz1[c] = BZ[c];
    };
    ECELL(cell)
}

// eta_calc (lagran.cpp:445-476) and rkstep (lagran.cpp:507-535) over
// (0..n)^3.  rkstep reads eta only at its own cell, and eta_calc's current
// there is rkstep's own (the same differences; (j + jp) * 0.5 and
// 0.5 * (j1 + j2) are the same product), so eta is computed in place rather
// than stored over eta_calc's wider range (-1..n+1), which nothing else reads.
template <class T>
KERNEL void r_rkstep(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const QDiv qx = qdiv_of(DXC(ix)), qy = qdiv_of(DYC(iy)), qz = qdiv_of(DZC(iz));
    double jx1 = qdiv(BZ[c + sy] - BZ[c], qy) - qdiv(BY[c + 1] - BY[c], qz);
    double jx2 = qdiv(BZ[c + sx + sy] - BZ[c + sx], qy) - qdiv(BY[c + sx + 1] - BY[c + sx], qz);
    double jy1 = qdiv(BX[c + 1] - BX[c], qz) - qdiv(BZ[c + sx] - BZ[c], qx);
    double jy2 = qdiv(BX[c + sy + 1] - BX[c + sy], qz) - qdiv(BZ[c + sx + sy] - BZ[c + sy], qx);
    double jz1 = qdiv(BY[c + sx] - BY[c], qx) - qdiv(BX[c + sy] - BX[c], qy);
    double jz2 = qdiv(BY[c + sx + 1] - BY[c + 1], qx) - qdiv(BX[c + sy + 1] - BX[c + 1], qy);
    double jx = 0.5 * (jx1 + jx2);
    double jy = 0.5 * (jy1 + jy2);
    double jz = 0.5 * (jz1 + jz2);
    double eta = P.eta_background;
    if (!P.eta_fixed) {
        double j_local = sqrt(jx * jx + jy * jy + jz * jz);
        eta = j_local > P.j_max ? P.eta_background + P.eta0 : P.eta_background;
    }
    lare::gbl(A.flux_x)[c] = -jx * eta * qx.d * 0.5;
    lare::gbl(A.flux_y)[c] = -jy * eta * qy.d * 0.5;
    lare::gbl(A.flux_z)[c] = -jz * eta * qz.d * 0.5;
    A.curlb[c] = eta * (jx * jx + jy * jy + jz * jz);
    };
    ECELL(cell)
}

// bstep's three components (lagran.cpp:537-590) with the Ohmic heating of
// lagran.cpp:490-500, over (0..n)^3: both read only rkstep's outputs; the B
// fill between them in the reference touches neither curlb, rho nor
// energy_electron.  bstep's bx1 (B at the substep's start) is read from B at
// the same cell before it is overwritten.
template <class T>
KERNEL void r_b_energy(const Src<T> S)
{
    EGRID
    auto cell = [&](const int ix, const int iy, const int iz) {
    const int c = IX(ix, iy, iz);
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const double dt = lare::gbl(A.s)->dt_sub;
    const auto fx = lare::gbl(A.flux_x), fy = lare::gbl(A.flux_y), fz = lare::gbl(A.flux_z);
    if (iy >= 1 && iz >= 1) {
        double area = DYB(iy) * DZB(iz);
        BX[c] = BX[c] + ddiv((fz[c] - fz[c - sy] + fz[c - 1] - fz[c - sy - 1] - fy[c] + fy[c - 1] - fy[c - sy] + fy[c - sy - 1]) * dt, area);
    }
    if (ix >= 1 && iz >= 1) {
        double area = DXB(ix) * DZB(iz);
        BY[c] = BY[c] + ddiv((fx[c] - fx[c - 1] + fx[c - sx] - fx[c - sx - 1] - fz[c] + fz[c - sx] - fz[c - 1] + fz[c - sx - 1]) * dt, area);
    }
    if (ix >= 1 && iy >= 1) {
        double area = DXB(ix) * DYB(iy);
        BZ[c] = BZ[c] + ddiv((fy[c] - fy[c - sx] + fy[c - sy] - fy[c - sx - sy] - fx[c] + fx[c - sy] - fx[c - sx] + fx[c - sx - sy]) * dt, area);
    }
    if (ix >= 1 && iy >= 1 && iz >= 1 && ix <= nx && iy <= ny && iz <= nz) {
        const Fld<T> cb = A.curlb;
        double sum_curlb = cb[c] + cb[c - sx] + cb[c - sy] + cb[c - sx - sy] + cb[c - 1] + cb[c - sx - 1] + cb[c - sy - 1] + cb[c - sx - sy - 1];
        EE[c] += ddiv(sum_curlb * dt, 8.0 * RHO[c]);
    }
    };
    ECELL(cell)
}

// rkstep's current, resistivity, flux and Ohmic term (lagran.cpp:445-535) at
// one vertex from the B values around it: bzAB / byAB / bxAB is the
// component at (x + A, y + B), the z suffix its value at z + 1.  The curlb
// returned is rounded to the storage type, as stored by r_rkstep.
// ETA_VAR: some grid of the launch has eta_calc's current-dependent branch
// (eta0 changes eta); otherwise eta is eta_background everywhere and the
// current's magnitude is not formed.
template <class T, bool FAST, bool ETA_VAR>
__device__ __forceinline__ void rk_vertex_q(double bz00, double bz01, double bz10, double bz11, double by00,
                                            double by00z, double by10, double by10z, double bx00, double bx00z,
                                            double bx01, double bx01z, const QDiv &qx, const QDiv &qy,
                                            const QDiv &qz, const Phys &P, double f[4], bool &slow)
{
    auto D = [&](double a, const QDiv &R) -> double {
        if constexpr (FAST)
            return qdiv_f(a, R, slow);
        else
            return ddiv(a, R.d);
    };
    double jx1 = D(bz01 - bz00, qy) - D(by00z - by00, qz);
    double jx2 = D(bz11 - bz10, qy) - D(by10z - by10, qz);
    double jy1 = D(bx00z - bx00, qz) - D(bz10 - bz00, qx);
    double jy2 = D(bx01z - bx01, qz) - D(bz11 - bz01, qx);
    double jz1 = D(by10 - by00, qx) - D(bx01 - bx00, qy);
    double jz2 = D(by10z - by00z, qx) - D(bx01z - bx00z, qy);
    double jx = 0.5 * (jx1 + jx2);
    double jy = 0.5 * (jy1 + jy2);
    double jz = 0.5 * (jz1 + jz2);
    double eta = P.eta_background;
    if (ETA_VAR && !P.eta_fixed) {
        double j_local = sqrt(jx * jx + jy * jy + jz * jz);
        eta = j_local > P.j_max ? P.eta_background + P.eta0 : P.eta_background;
    }
    f[0] = -jx * eta * qx.d * 0.5;
    f[1] = -jy * eta * qy.d * 0.5;
    f[2] = -jz * eta * qz.d * 0.5;
    f[3] = (double)(T)(eta * (jx * jx + jy * jy + jz * jz));
}
template <class T, bool ETA_VAR>
__device__ __forceinline__ void rk_vertex(double bz00, double bz01, double bz10, double bz11, double by00,
                                          double by00z, double by10, double by10z, double bx00, double bx00z,
                                          double bx01, double bx01z, const QDiv &qx, const QDiv &qy,
                                          const QDiv &qz, const Phys &P, double f[4])
{
    bool slow = false;
    rk_vertex_q<T, true, ETA_VAR>(bz00, bz01, bz10, bz11, by00, by00z, by10, by10z, bx00, bx00z, bx01, bx01z, qx, qy,
                                  qz, P, f, slow);
    if (__builtin_expect(slow, 0))
        rk_vertex_q<T, false, ETA_VAR>(bz00, bz01, bz10, bz11, by00, by00z, by10, by10z, bx00, bx00z, bx01, bx01z,
                                       qx, qy, qz, P, f, slow);
}

// One resistive substep (r_rkstep then r_b_energy, lagran.cpp:490-590) as a
// z-line x-march (H = 1) over (0..n)^3, B read from one copy and written to
// the other (to_bx1: B -> bx1..bz1, else bx1..bz1 -> B), so no lane reads a
// B another lane has updated.  Each lane computes the rkstep fluxes and
// Ohmic terms of its vertex column (x, y) and of (x, y-1); those at x - 1
// are carried and those at z - 1 come from the lane below.  The same
// arithmetic as r_rkstep and r_b_energy; B outside each component's update
// range is left to the boundary fill that follows.
template <class T, int K, bool ETA_VAR>
KERNEL OCC(3) void r_substep_zx(const Src<T> S, int to_bx1)
{
    const int b_ = xcd_swizzle((int)blockIdx.x, S.nblk);
    const int g_ = grid_of(S.boff, S.ngrid, b_);
    const EDev<T> &D_ = S.devs[g_];
    const Geo G = D_.g;
    const ArrE<T> A = state_view(D_.a, S.st);
    const Phys P = D_.p;
    const ZXPos Z = zx_map<1, K>(S, G, g_, b_);
    const int sx = G.sx, sy = G.sy;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int iy = Z.iy, iz = Z.iz;
    const double dt = lare::gbl(A.s)->dt_sub;
    const Fld<T> IX_ = to_bx1 ? BX : A.bx1, IY_ = to_bx1 ? BY : A.by1, IZ_ = to_bx1 ? BZ : A.bz1;
    const Fld<T> OX_ = to_bx1 ? A.bx1 : BX, OY_ = to_bx1 ? A.by1 : BY, OZ_ = to_bx1 ? A.bz1 : BZ;
    const QDiv qy0 = qdiv_of(DYC(iy)), qym = qdiv_of(DYC(iy - 1)), qz = qdiv_of(DZC(iz));
    const double dyb = DYB(iy), dzb = DZB(iz), area_x = dyb * dzb;
    int c = IX(Z.ix0, iy, iz);
    // plane x: BY at (x, y-1), (x, y) and their z+1; BZ at (x, y-1..y+1)
    double by_m = IY_[c - sy], by_0 = IY_[c];
    double by_mz = lane_next(by_m), by_0z = lane_next(by_0);
    double bz_m = IZ_[c - sy], bz_0 = IZ_[c], bz_p = IZ_[c + sy];
    // rkstep terms (fx, fy, fz, curlb) of plane x - 1 at (y) and (y-1), and
    // of the lane below
    double F0x[4] = {0, 0, 0, 0}, Fmx[4] = {0, 0, 0, 0}, pF0x[4] = {0, 0, 0, 0}, pFmx[4] = {0, 0, 0, 0};
    // The loads of a position are issued one position ahead (n*; the last
    // position reloads its own), so their latency overlaps the previous
    // position's arithmetic; the x deltas of the march are lane values.
    double nbx_m = IX_[c - sy], nbx_0 = IX_[c], nbx_p = IX_[c + sy];
    double nby_m = IY_[c + sx - sy], nby_0 = IY_[c + sx];
    double nbz_m = IZ_[c + sx - sy], nbz_0 = IZ_[c + sx], nbz_p = IZ_[c + sx + sy];
    const double dxc_row = x_row(A.dxc, Z.ix0, nx + 2), dxb_row = x_row(A.dxb, Z.ix0, nx + 2);
    for (int ix = Z.ix0; ix <= Z.ix1; ++ix, c += sx) {
        const double bx_m = nbx_m, bx_0 = nbx_0, bx_p = nbx_p;
        const double byN_m = nby_m, byN_0 = nby_0;
        const double bzN_m = nbz_m, bzN_0 = nbz_0, bzN_p = nbz_p;
        const double dxc = x_at(dxc_row, ix - Z.ix0), dxb = x_at(dxb_row, ix - Z.ix0);
        const double rho = RHO[c], ee = EE[c];
        {
            const int n = ix < Z.ix1 ? c + sx : c;
            nbx_m = IX_[n - sy];
            nbx_0 = IX_[n];
            nbx_p = IX_[n + sy];
            nby_m = IY_[n + sx - sy];
            nby_0 = IY_[n + sx];
            nbz_m = IZ_[n + sx - sy];
            nbz_0 = IZ_[n + sx];
            nbz_p = IZ_[n + sx + sy];
        }
        const double bx_mz = lane_next(bx_m), bx_0z = lane_next(bx_0), bx_pz = lane_next(bx_p);
        const double byN_mz = lane_next(byN_m), byN_0z = lane_next(byN_0);
        const QDiv qx = qdiv_of(dxc);
        double F0[4], Fm[4];
        rk_vertex<T, ETA_VAR>(bz_0, bz_p, bzN_0, bzN_p, by_0, by_0z, byN_0, byN_0z, bx_0, bx_0z, bx_p, bx_pz, qx, qy0,
                              qz, P, F0);
        rk_vertex<T, ETA_VAR>(bz_m, bz_0, bzN_m, bzN_0, by_m, by_mz, byN_m, byN_mz, bx_m, bx_mz, bx_0, bx_0z, qx, qym,
                              qz, P, Fm);
        double pF0[4], pFm[4];
        for (int q = 0; q < 4; ++q) {
            pF0[q] = lane_prev(F0[q]);
            pFm[q] = lane_prev(Fm[q]);
        }
        if (Z.wz && ix > Z.ix0) {
            // fx = [0], fy = [1], fz = [2], curlb = [3]; c, c-sy, c-1, c-sy-1 at
            // this x (F0, Fm, pF0, pFm) and at x - 1 (F0x, Fmx, pF0x, pFmx)
            if (iy >= 1 && iz >= 1)
                OX_[c] = bx_0 + ddiv((F0[2] - Fm[2] + pF0[2] - pFm[2] - F0[1] + pF0[1] - Fm[1] + pFm[1]) * dt, area_x);
            if (ix >= 1 && iz >= 1) {
                double area = dxb * dzb;
                OY_[c] = by_0 + ddiv((F0[0] - pF0[0] + F0x[0] - pF0x[0] - F0[2] + F0x[2] - pF0[2] + pF0x[2]) * dt, area);
            }
            if (ix >= 1 && iy >= 1) {
                double area = dxb * dyb;
                OZ_[c] = bz_0 + ddiv((F0[1] - F0x[1] + Fm[1] - Fmx[1] - F0[0] + Fm[0] - F0x[0] + Fmx[0]) * dt, area);
            }
            if (ix >= 1 && iy >= 1 && iz >= 1 && ix <= nx && iy <= ny && iz <= nz) {
                double sum_curlb = F0[3] + F0x[3] + Fm[3] + Fmx[3] + pF0[3] + pF0x[3] + pFm[3] + pFmx[3];
                EE[c] = ee + ddiv(sum_curlb * dt, 8.0 * rho);
            }
        }
        for (int q = 0; q < 4; ++q) {
            F0x[q] = F0[q];
            Fmx[q] = Fm[q];
            pF0x[q] = pF0[q];
            pFmx[q] = pFm[q];
        }
        by_m = byN_m;
        by_0 = byN_0;
        by_mz = byN_mz;
        by_0z = byN_0z;
        bz_m = bzN_m;
        bz_0 = bzN_0;
        bz_p = bzN_p;
    }
}

// ------------------------------------------------- boundary conditions

// One edge of one built-in boundary condition along a line of a variable
// (variableDef.h:539-605, builtInBoundaryConditions.h:54-300): ghost layers
// -1..0 (centred) or -2..0 (staggered, the boundary point included) at the
// lower edge, n+1..n+2 or n..n+2 at the upper.
// Map edges (kind 5; LARE_BC_MAP) give each ghost layer g0 + k a source
// layer m[k] (read before the edge writes), a constant (kMapValue) or
// nothing (kMapKeep).
constexpr int kMapKeep = -1000, kMapValue = -1001;

template <class E>
__device__ __forceinline__ void bc_edge_e(E *p, int c0, int sd, int n, int s, int edge, int kind, double val,
                                          const int *m)
{
    if (kind == 0) return;
    const int LB = s ? 0 : 1, UB = n;
    const int g0 = edge ? n + 1 - s : (s ? -2 : -1);
    const int g1 = edge ? n + 2 : 0;
    if (kind == 5) {
        double out[3];
        for (int g = g0; g <= g1; ++g) {
            const int q = m[g - g0];
            out[g - g0] = q == kMapValue ? val : q == kMapKeep ? 0.0 : (double)p[c0 + q * sd];
        }
        for (int g = g0; g <= g1; ++g)
            if (m[g - g0] != kMapKeep) p[c0 + g * sd] = (E)out[g - g0];
        return;
    }
    const int first = edge ? g0 : g1;
    const int last = edge ? UB : LB;
    for (int g = g0; g <= g1; ++g) {
        double out;
        if (kind == 1) {
            out = (double)p[c0 + (edge ? g - n : g + n) * sd];
        } else if (kind == 2) {
            out = (double)p[c0 + last * sd];
        } else if (kind == 3) {
            const double bv = (double)p[c0 + last * sd];
            const double boundary = (bv + bv) / 2.0;
            const double delta = (double)p[c0 + (last - (g - first)) * sd] - boundary;
            out = boundary - delta;
        } else {
            out = val;
        }
        p[c0 + g * sd] = (E)out;
    }
}

// Both edges of a line at once, with every value the fills read loaded
// before any is written, so the loads of a lane overlap instead of waiting
// on the stores before them.  The values are those of bc_edge_e applied to
// the lower edge and then the upper edge: every source of either fill is an
// unfilled cell (n >= 3) except one, the upper periodic fill of a staggered
// line at g = n, which copies index 0 as the lower fill has just set it.
template <class E>
__device__ __forceinline__ void bc_line_e(E *p, int c0, int sd, int n, int s, int k0, double v0, int k1, double v1,
                                          const int *m0, const int *m1)
{
    if (k0 == 5 || k1 == 5) {  // map edges: the two edges in turn
        bc_edge_e(p, c0, sd, n, s, 0, k0, v0, m0);
        bc_edge_e(p, c0, sd, n, s, 1, k1, v1, m1);
        return;
    }
    const int gl = s ? -2 : -1, nl = 1 - gl;  // lower ghosts gl..0
    const int gh = n + 1 - s, nh = n + 3 - gh;  // upper ghosts gh..n+2
    const int LB = s ? 0 : 1, UB = n;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (k0 == 3 || k0 == 2) {
        const double bv = (double)p[c0 + LB * sd];
        const double boundary = (bv + bv) / 2.0;
        for (int j = 0; j < nl; ++j) {
            const int g = gl + j;
            if (k0 == 2) {
                lo[j] = bv;
            } else {
                const double delta = (double)p[c0 + (LB - g) * sd] - boundary;
                lo[j] = boundary - delta;
            }
        }
    } else if (k0 == 1) {
        for (int j = 0; j < nl; ++j) lo[j] = (double)p[c0 + (gl + j + n) * sd];
    } else if (k0 != 0) {  // clamp
        for (int j = 0; j < nl; ++j) lo[j] = v0;
    }
    if (k1 == 3 || k1 == 2) {
        const double bv = (double)p[c0 + UB * sd];
        const double boundary = (bv + bv) / 2.0;
        for (int j = 0; j < nh; ++j) {
            const int g = gh + j;
            if (k1 == 2) {
                hi[j] = bv;
            } else {
                const double delta = (double)p[c0 + (UB - (g - gh)) * sd] - boundary;
                hi[j] = boundary - delta;
            }
        }
    } else if (k1 == 1) {
        for (int j = 0; j < nh; ++j) {
            const int src = gh + j - n;
            hi[j] = (k0 != 0 && src <= 0) ? (double)(E)lo[src - gl] : (double)p[c0 + src * sd];
        }
    } else if (k1 != 0) {  // clamp
        for (int j = 0; j < nh; ++j) hi[j] = v1;
    }
    if (k0)
        for (int j = 0; j < nl; ++j) p[c0 + (gl + j) * sd] = (E)lo[j];
    if (k1)
        for (int j = 0; j < nh; ++j) p[c0 + (gh + j) * sd] = (E)hi[j];
}

// Workgroups per grid: nvar variables x ceil(lines / 256); each lane owns
// one line along dim (spanning the other dimensions' full extents) and
// applies the lower edge, then the upper edge.
template <class T>
KERNEL void e_bc(const EDev<T> *devs, const lare::BCSetE *sets, const int *boff, int ngrid, int nblk, int dim,
                 int nvar)
{
    const int b = xcd_swizzle((int)blockIdx.x, nblk);
    const int g = grid_of(boff, ngrid, b);
    const Geo G = devs[g].g;
    const int na = dim == 0 ? G.ny : G.nx;
    const int nb = dim == 2 ? G.ny : G.nz;
    const int bpl = ((na + 5) * (nb + 5) + 255) >> 8;
    const int lb = b - boff[g];
    const int vi = lb / bpl;
    if (vi >= nvar) return;
    const lare::BCVarE V = sets[g].v[vi];
    if (V.k0 == 0 && V.k1 == 0 && !V.src) return;
    const int la = V.sa ? -2 : -1, ua = na + 2;
    const int lbb = V.sb ? -2 : -1, ub = nb + 2;
    const int wb = ub - lbb + 1;
    const int t = (lb - vi * bpl) * 256 + (int)threadIdx.x;
    if (t >= (ua - la + 1) * wb) return;
    const int ia = la + t / wb, ib = lbb + t % wb;
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
    if (V.src) {
        // The variable was just written into the other state copy over its
        // update range (1..n centred, 0..n staggered, per dimension).  Its
        // remaining cells take the values of the copy it was read from, so
        // that the fills below see what an in-place update would have left.
        const int loa = V.sa ? 0 : 1, lob = V.sb ? 0 : 1, lo = V.sd ? 0 : 1;
        const bool shell = ia < loa || ia > na || ib < lob || ib > nb;
        T *q = (T *)V.p;
        const T *r = (const T *)V.src;
        if (shell) {
            for (int g = V.sd ? -2 : -1; g <= n + 2; ++g) q[c0 + g * sd] = r[c0 + g * sd];
        } else {
            for (int g = V.sd ? -2 : -1; g < lo; ++g) q[c0 + g * sd] = r[c0 + g * sd];
            for (int g = n + 1; g <= n + 2; ++g) q[c0 + g * sd] = r[c0 + g * sd];
        }
    }
    // z lines (one scattered row end per lane) gain from overlapping their
    // loads; x and y lines are coalesced across lanes already.  Lines of
    // two cells (n = 2) keep the edges' own order: there a fill can read a
    // layer the other edge has just written.
    if (V.src || dim != 2 || n < 3) {  // in the fills' own order (after the carry-over)
        if (V.f64) {
            bc_edge_e((double *)V.p, c0, sd, n, V.sd, 0, V.k0, V.v0, V.m0);
            bc_edge_e((double *)V.p, c0, sd, n, V.sd, 1, V.k1, V.v1, V.m1);
        } else {
            bc_edge_e((T *)V.p, c0, sd, n, V.sd, 0, V.k0, V.v0, V.m0);
            bc_edge_e((T *)V.p, c0, sd, n, V.sd, 1, V.k1, V.v1, V.m1);
        }
    } else if (V.f64) {
        bc_line_e((double *)V.p, c0, sd, n, V.sd, V.k0, V.v0, V.k1, V.v1, V.m0, V.m1);
    } else {
        bc_line_e((T *)V.p, c0, sd, n, V.sd, V.k0, V.v0, V.k1, V.v1, V.m0, V.m1);
    }
}

// The fills of dims 0, 1, 2 in turn (variableDef.h:601-605) composed: the
// value at a cell after the fills of dims 0..D-1, computed from the values
// before any fill.  The fill of dim d rewrites the cells whose index along d
// lies in a ghost range of an edge with a condition (lower gl..0, upper
// n+1-s..n+2), each from cells of the same line, over the full extents of
// the other dims; so a cell's final value is its edge rule applied to
// values after the fills of the dims before d, at the source indices the
// rule names, with every stage rounded to the storage type as a store
// would.  Rules as bc_edge_e: periodic (1), zero-gradient (2), mirror (3),
// clamp (4).  The one source a fill reads that the same fill has written is
// index 0 for the upper periodic edge of a staggered line (the lower edge's
// value, recomputed here).
template <class E, int D>
__device__ __forceinline__ double bc3_val(const lare::BCVar3 &V, const E *p, const Geo &G, int i0, int i1, int i2)
{
    if constexpr (D == 0) {
        return (double)p[IX(i0, i1, i2)];
    } else {
        constexpr int d = D - 1;
        const int x = d == 0 ? i0 : d == 1 ? i1 : i2;
        const int n = d == 0 ? G.nx : d == 1 ? G.ny : G.nz;
        const int s = V.s[d], k0 = V.k0[d], k1 = V.k1[d];
        // the rule as a form (0: copy of q1; 1: mirror about q1 of q2; 2:
        // the constant c) and its source indices
        int form = 0, q1 = x, q2 = x;
        double c = 0.0;
        auto mapped = [&](int q, double value) {  // map layer: source q, the constant, or kept
            if (q == kMapValue) {
                form = 2;
                c = value;
            } else if (q != kMapKeep) {
                q1 = q;
            }
        };
        auto lower = [&](int g) {
            const int LB = s ? 0 : 1;
            if (k0 == 5) {
                mapped(V.m0[d][g - (s ? -2 : -1)], V.v0[d]);
            } else if (k0 == 1) {
                q1 = g + n;
            } else if (k0 == 2) {
                q1 = LB;
            } else if (k0 == 3) {
                form = 1;
                q1 = LB;
                q2 = LB - g;
            } else {
                form = 2;
                c = V.v0[d];
            }
        };
        const int gh = n + 1 - s;
        if (k1 == 5 && x >= gh) {
            mapped(V.m1[d][x - gh], V.v1[d]);
        } else if (k0 && x <= 0) {
            lower(x);
        } else if (k1 && x >= gh) {
            if (k1 == 1) {
                const int src = x - n;
                if (src <= 0 && k0)
                    lower(src);
                else
                    q1 = src;
            } else if (k1 == 2) {
                q1 = n;
            } else if (k1 == 3) {
                form = 1;
                q1 = n;
                q2 = n - (x - gh);
            } else {
                form = 2;
                c = V.v1[d];
            }
        }
        if (form == 2) return (double)(E)c;
        const double a = bc3_val<E, D - 1>(V, p, G, d == 0 ? q1 : i0, d == 1 ? q1 : i1, d == 2 ? q1 : i2);
        if (form == 0) return a;
        const double boundary = (a + a) / 2.0;
        const double delta = bc3_val<E, D - 1>(V, p, G, d == 0 ? q2 : i0, d == 1 ? q2 : i1, d == 2 ? q2 : i2) - boundary;
        return (double)(E)(boundary - delta);
    }
}

// Workgroups per grid: each variable's ghost cells (the box lo..n+2 of
// every dim less the cells 1..n-s of all three), 256 kBC3Cells per
// workgroup, kBC3Cells per lane: first the x slabs (all y, z), then the y slabs (x
// inside), then the z slabs (x, y inside), z fastest.  A cell whose every
// index lies outside the ghost ranges of edges with a condition is left
// as it is.  Every value read is of a cell no fill rewrites, or one whose
// rewrite stores the value it holds (the boundary point of a staggered
// line under zero-gradient or mirror, or under periodic at both edges);
// the host takes the per-dim fills for a staggered line with a periodic
// lower and a clamped upper edge, whose lower fill reads the point n that
// the upper fill then changes.
constexpr int kBC3Cells = 4;  // ghost cells per lane of e_bc3
template <class T>
KERNEL void e_bc3(const EDev<T> *devs, const lare::BCSet3 *sets, const int *boff, int ngrid, int nblk)
{
    const int b = xcd_swizzle((int)blockIdx.x, nblk);
    const int g = grid_of(boff, ngrid, b);
    const Geo G = devs[g].g;
    const int lb = b - boff[g];
    const lare::BCSet3 &S = sets[g];
    int vi = 0;
    while (vi + 1 < S.nv && S.v[vi + 1].b0 <= lb) ++vi;
    const lare::BCVar3 &V = S.v[vi];
    const int n3[3] = {G.nx, G.ny, G.nz};
    int lo[3], ext[3], wl[3], wd[3], in[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = V.s[d] ? -2 : -1;
        ext[d] = n3[d] + 3 - lo[d];
        wl[d] = 1 - lo[d];
        wd[d] = wl[d] + 2 + V.s[d];
        in[d] = n3[d] - V.s[d];
    }
    auto slab = [&](int d, int a) { return a < wl[d] ? lo[d] + a : n3[d] + 1 - V.s[d] + (a - wl[d]); };
    const int cx = wd[0] * ext[1] * ext[2], cy = in[0] * wd[1] * ext[2];
    // kBC3Cells ghost cells per lane, 256 apart, so the workgroup's set-up
    // (its grid, variable and rules) serves 256 kBC3Cells cells
    const int tb = (lb - V.b0) * (256 * kBC3Cells) + (int)threadIdx.x;
    for (int r = 0; r < kBC3Cells; ++r) {
    const int t = tb + 256 * r;
    if (t >= V.nsh) break;
    int x, y, z;
    if (t < cx) {
        const int a = t / (ext[1] * ext[2]), r = t - a * (ext[1] * ext[2]);
        x = slab(0, a);
        y = lo[1] + r / ext[2];
        z = lo[2] + r % ext[2];
    } else if (t < cx + cy) {
        const int u = t - cx;
        const int a = u / (wd[1] * ext[2]), r = u - a * (wd[1] * ext[2]);
        x = 1 + a;
        y = slab(1, r / ext[2]);
        z = lo[2] + r % ext[2];
    } else {
        const int u = t - cx - cy;
        const int a = u / (in[1] * wd[2]), r = u - a * (in[1] * wd[2]);
        x = 1 + a;
        y = 1 + r / wd[2];
        z = slab(2, r % wd[2]);
    }
    const int pos[3] = {x, y, z};
    bool filled = false;
    for (int d = 0; d < 3; ++d)
        filled = filled || (V.k0[d] && pos[d] <= 0) || (V.k1[d] && pos[d] >= n3[d] + 1 - V.s[d]);
    if (!filled) continue;
    if (V.f64) {
        double *p = (double *)V.p;
        p[IX(x, y, z)] = bc3_val<double, 3>(V, p, G, x, y, z);
    } else {
        T *p = (T *)V.p;
        p[IX(x, y, z)] = (T)bc3_val<T, 3>(V, p, G, x, y, z);
    }
    }
}

// ------------------------------------------------- energy diagnostic

// Per-workgroup partial sums over the domain cells (1..n)^3 of one grid:
// internal rho (e_e + e_i) cv, kinetic 1/2 rho_v |v|^2 cv_v at the upper
// vertex of the cell, magnetic |B_c|^2 / (2 mu0) cv with B_c the cell average
// of the face values.  Fixed tree order: the result is reproducible.
template <class T>
KERNEL void e_energy_part(const Src<T> S, double *part)
{
    EGRID
    double e[3] = {0.0, 0.0, 0.0};
    auto cell = [&](const int ix, const int iy, const int iz) {
        const int c = IX(ix, iy, iz);
        const int sx = G.sx, sy = G.sy;
        const double cv = CV(ix, iy, iz);
        e[0] += RHO[c] * (EE[c] + EI[c]) * cv;
        const double c000 = cv, c100 = CV(ix + 1, iy, iz), c010 = CV(ix, iy + 1, iz),
                     c110 = CV(ix + 1, iy + 1, iz), c001 = CV(ix, iy, iz + 1), c101 = CV(ix + 1, iy, iz + 1),
                     c011 = CV(ix, iy + 1, iz + 1), c111 = CV(ix + 1, iy + 1, iz + 1);
        const Fld<T> r = RHO;
        double src = r[c] * c000 + r[c + sx] * c100 + r[c + sy] * c010 + r[c + sx + sy] * c110 + r[c + 1] * c001 +
                     r[c + sx + 1] * c101 + r[c + sy + 1] * c011 + r[c + sx + sy + 1] * c111;
        double scv = c000 + c100 + c010 + c110 + c001 + c101 + c011 + c111;
        const double v2 = VX[c] * VX[c] + VY[c] * VY[c] + VZ[c] * VZ[c];
        e[1] += 0.5 * ddiv(src, scv) * v2 * (0.125 * scv);
        const double bx = 0.5 * (BX[c] + BX[c - sx]), by = 0.5 * (BY[c] + BY[c - sy]), bz = 0.5 * (BZ[c] + BZ[c - 1]);
        e[2] += ddiv(bx * bx + by * by + bz * bz, 2.0 * P.mu0) * cv;
    };
    ECELL(cell)
    __shared__ double sh[3][256];
    for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = e[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[3 * b_ + k] = sh[k][0];
}

// Sum of each grid's partials in block order (one workgroup per grid).
KERNEL void e_energy_sum(const double *part, const int *boff, int ngrid, int nblk, double *out)
{
    const int g = (int)blockIdx.x;
    if (g >= ngrid || threadIdx.x != 0) return;
    const int b1 = g + 1 < ngrid ? boff[g + 1] : nblk;
    double e[3] = {0.0, 0.0, 0.0};
    for (int b = boff[g]; b < b1; ++b)
        for (int k = 0; k < 3; ++k) e[k] += part[3 * b + k];
    out[4 * g + 0] = e[0];
    out[4 * g + 1] = e[1];
    out[4 * g + 2] = e[2];
    out[4 * g + 3] = e[0] + e[1] + e[2];
}

// Scratch of the allocation-free energy diagnostic (group_energy_noalloc):
// 3 block partial sums per workgroup of the launch, then 4 totals per member.
// 16 MiB covers one grid of up to about 560^3 cells.  Each GPU the library
// runs on holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
static __device__ __align__(256) double scratch[(16u << 20) / sizeof(double)];

// ================================================================== host

namespace lare {

namespace {

// Stagger per dimension of each registered variable (setup.cpp:45-69).
const int kStagE[13][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 0, 0},
                           {0, 1, 0}, {0, 0, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {0, 0, 0}};

inline RSpec rs(int x0, int dx1, int y0, int dy1, int z0, int dz1) { return RSpec{x0, dx1, y0, dy1, z0, dz1}; }

// Whether the fill of edge e of variable u along d writes every ghost layer
// of the edge (a map edge may keep some).
bool edge_writes_all(const Grid &g, int u, int d, int e)
{
    const int k = g.bck[u][d][e];
    if (k == 0) return false;
    if (k != 5) return true;
    const int nl = kStagE[u][d] ? 3 : 2;
    for (int j = 0; j < nl; ++j)
        if (g.bcm[u][d][e][j] == kMapKeep) return false;
    return true;
}

// Whether a map edge of u along d reads a layer that a fill of d rewrites
// (the all-dimension fill reads values before any fill of the launch).
bool map_reads_written(const Grid &g, int u, int d, int n)
{
    const int s = kStagE[u][d], gl = s ? -2 : -1, gh = n + 1 - s;
    auto written = [&](int q) {
        for (int e = 0; e < 2; ++e) {
            const int k = g.bck[u][d][e], g0 = e ? gh : gl, g1 = e ? n + 2 : 0;
            if (k == 0 || q < g0 || q > g1) continue;
            if (k != 5 || g.bcm[u][d][e][q - g0] != kMapKeep) return true;
        }
        return false;
    };
    for (int e = 0; e < 2; ++e) {
        if (g.bck[u][d][e] != 5) continue;
        for (int j = 0; j < (s ? 3 : 2); ++j) {
            const int q = g.bcm[u][d][e][j];
            if (q != kMapKeep && q != kMapValue && written(q)) return true;
        }
    }
    return false;
}

// Variable ids of the boundary fills: 0..12 the registered variables (state
// in the first copy), 13..18 the second copy of rho, e_e, e_i, vx, vy, vz,
// 19..21 bx1..bz1 as the second copy of B (resistive substeps and the
// remap), 22..24 bxy..bzy as B's third copy (the fused y sweep).
const int kSecond = 13, kThirdB = 22;
inline int base_var(int v) { return v >= kThirdB ? v - kThirdB + 6 : v >= kSecond ? v - kSecond : v; }

template <class T>
void *field_ptr_t(const ArrE<T> &a, int f)
{
    const Fld<T> fs[25] = {a.rho, a.ee,  a.ei,  a.vx,  a.vy,  a.vz,  a.bx,  a.by,  a.bz,
                           a.vx1, a.vy1, a.vz1, {nullptr}, a.rho2, a.ee2, a.ei2, a.vx2, a.vy2,
                           a.vz2, a.bx1, a.by1, a.bz1, a.bxy, a.byy, a.bzy};
    return f == 12 ? (void *)a.dm : (void *)fs[f].p;
}
void *field_ptr(const Grid &g, int f) { return g.fp32 ? field_ptr_t(g.e32.a, f) : field_ptr_t(g.e64.a, f); }

// Block offsets of a launch over `subset` with range r (built outside graph
// capture, by group_prepare or on first use).
// Table order codes: 0 plane-major, 1 band-major, kMarch + 100 * DIR + K a
// lane walking K cells along DIR (f_update_march).
const int kMarch = 1000;

Group::Tab &table(Group &gp, int subset, const RSpec &r, int order)
{
    const std::array<int, 7> key{r.x0, r.dx1, r.y0, r.dy1, r.z0, r.dz1, order};
    auto it = gp.tabs.find({subset, key});
    if (it != gp.tabs.end()) return it->second;
    Group::Tab t;
    std::vector<int> boff;
    long total = 0;
    for (size_t m = 0; m < gp.mem.size(); ++m) {
        if (subset == 1 && std::find(gp.res.begin(), gp.res.end(), (int)m) == gp.res.end()) continue;
        if (subset == 2 && std::find(gp.rke.begin(), gp.rke.end(), (int)m) == gp.rke.end()) continue;
        const Geo &G = gp.mem[m]->dev.g;
        const long nxr = (long)G.nx + r.dx1 - r.x0 + 1, nyr = (long)G.ny + r.dy1 - r.y0 + 1,
                   nzr = (long)G.nz + r.dz1 - r.z0 + 1;
        boff.push_back((int)total);
        t.units += (double)G.nx * G.ny * G.nz;
        if (nxr > 0 && nyr > 0 && nzr > 0) {
            if (order >= kZY) {
                const int H = (order - kZY) / 100, K = (order - kZY) % 100;
                const long per = 4 * (64 - 2 * H);
                total += ((nyr + K - 1) / K) * ((nxr * (nzr + 2 * H) + per - 1) / per);
            } else if (order >= kZX) {
                const int H = (order - kZX) / 100, K = (order - kZX) % 100;
                const long per = 4 * (64 - 2 * H);
                total += ((nxr + K - 1) / K) * ((nyr * (nzr + 2 * H) + per - 1) / per);
            } else if (order == kZLine) {
                total += nxr * ((nyr * (nzr + 4) + 239) / 240);
            } else if (order >= kMarch) {
                // a lane walks K cells along DIR; the plane across DIR is
                // (y, z) for DIR 0, (x, z) for DIR 1
                const int mdir = (order - kMarch) / 100, K = (order - kMarch) % 100;
                const long nd = mdir == 0 ? nxr : nyr, ns = mdir == 0 ? nyr : nxr;
                total += ((ns * nzr + 255) / 256) * ((nd + K - 1) / K);
            } else {
                const long bpp = (nyr * nzr + 255) / 256;
                total += nxr * (order ? 8 * ((bpp + 7) / 8) : bpp);
            }
            t.cells += (double)(nxr * nyr * nzr);
        }
    }
    t.nblk = (int)total;
    t.order = order;
    if (!boff.empty()) {
        (void)hipMalloc(&t.boff, boff.size() * sizeof(int));
        (void)hipMemcpyAsync(t.boff, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice, gp.stream);
        (void)hipStreamSynchronize(gp.stream);
    }
    return gp.tabs.emplace(std::make_pair(subset, key), t).first->second;
}

// "(f_prep<T, DIR>)" -> "f_prep<DIR>", "g_cs<T>" -> "g_cs"
std::string kname(const char *k)
{
    std::string n;
    for (const char *p = k; *p; ++p)
        if (*p != '(' && *p != ')') n += *p;
    for (const char *pat : {"<T>", "T, "}) {
        size_t at;
        while ((at = n.find(pat)) != std::string::npos) n.erase(at, std::strlen(pat));
    }
    return n;
}

template <class T, typename... KA, typename... Args>
void elaunch_(void (*k)(Src<T>, KA...), const char *name, double bpc, Group &gp, int subset, RSpec r, int st,
              int order, Args... args)
{
    Group::Tab &t = table(gp, subset, r, order);
    if (gp.planning || t.nblk == 0) return;
    Src<T> s{(const EDev<T> *)gp.devs[subset], t.boff, gp.n[subset], t.nblk, r, t.order, st};
    Prof *pf = gp.prof;
    // algorithmic bytes: per cell-update (bpc) times the members' domain cells
    Prof::Rec rec{kname(name), bpc * t.units, nullptr, nullptr};
    if (pf && pf->on) {
        if (pf->dir >= 0) {
            const size_t at = rec.name.find("<DIR>");
            if (at != std::string::npos) rec.name.replace(at, 5, "<" + std::to_string(pf->dir) + ">");
        }
        (void)hipEventCreate(&rec.a);
        (void)hipEventCreate(&rec.b);
        (void)hipEventRecord(rec.a, gp.stream);
    }
    hipLaunchKernelGGL(k, dim3((unsigned)t.nblk), dim3(256), 0, gp.stream, s, static_cast<KA>(args)...);
    if (pf && pf->on) {
        (void)hipEventRecord(rec.b, gp.stream);
        pf->recs.push_back(rec);
    }
}
// launch on the state copy 0 / on state copy st
#define EL(k, bpc, gp, subset, r, ...) elaunch_(k, #k, bpc, gp, subset, r, 0, gp.order, ##__VA_ARGS__)
#define ELS(k, bpc, gp, subset, r, st, ...) elaunch_(k, #k, bpc, gp, subset, r, st, gp.order, ##__VA_ARGS__)

// Per-grid kernels (one lane per member).
template <class T>
void per_grid(void (*k)(const EDev<T> *, int), const char *name, Group &gp)
{
    if (gp.planning || gp.n[0] == 0) return;
    Prof *pf = gp.prof;
    Prof::Rec rec{name, 0.0, nullptr, nullptr};
    if (pf && pf->on) {
        (void)hipEventCreate(&rec.a);
        (void)hipEventCreate(&rec.b);
        (void)hipEventRecord(rec.a, gp.stream);
    }
    hipLaunchKernelGGL(k, dim3((unsigned)((gp.n[0] + 255) / 256)), dim3(256), 0, gp.stream,
                       (const EDev<T> *)gp.devs[0], gp.n[0]);
    if (pf && pf->on) {
        (void)hipEventRecord(rec.b, gp.stream);
        pf->recs.push_back(rec);
    }
}

// Variable groups of the boundary fills (boundary.cpp:27-96): the variables
// filled and, for a variable just written into the other state copy, the
// copy it was read from (-1: updated in place).
enum { VB, VE, VRHO, VV, VV1, VDM, VALL, VERHO, VBE, VCORR_2, VALL_2, VALL_1, VBE_1, VBE_0, VALL_2B, VALL_1B, VV_2,
       VALL_2C, VALL_1C, NVG };
struct VarGroup {
    int n;
    int dst[9];
    int src[9];
};
const VarGroup kGroups[NVG] = {
    {3, {6, 7, 8}, {-1, -1, -1}},
    {2, {1, 2}, {-1, -1}},
    {1, {0}, {-1}},
    {3, {3, 4, 5}, {-1, -1, -1}},
    {3, {9, 10, 11}, {-1, -1, -1}},
    {1, {12}, {-1}},
    {9, {6, 7, 8, 1, 2, 0, 3, 4, 5}, {-1, -1, -1, -1, -1, -1, -1, -1, -1}},
    {3, {1, 2, 0}, {-1, -1, -1}},
    {5, {6, 7, 8, 1, 2}, {-1, -1, -1, -1, -1}},
    // velocity_bcs, energy_bcs and density_bcs after the corrector wrote the
    // second copy
    {6, {16, 17, 18, 14, 15, 13}, {3, 4, 5, 1, 2, 0}},
    // boundary_conditions after a sweep that wrote the second copy
    {9, {6, 7, 8, 14, 15, 13, 16, 17, 18}, {-1, -1, -1, 1, 2, 0, 3, 4, 5}},
    // boundary_conditions after a sweep that wrote the first copy
    {9, {6, 7, 8, 1, 2, 0, 3, 4, 5}, {-1, -1, -1, 14, 15, 13, 16, 17, 18}},
    // the resistive substep's bfield_bcs and energy_bcs after it wrote B
    // into bx1..bz1, and after it wrote B back
    {5, {19, 20, 21, 1, 2}, {6, 7, 8, -1, -1}},
    {5, {6, 7, 8, 1, 2}, {19, 20, 21, -1, -1}},
    // boundary_conditions after the x / y sweep with B in bx1..bz1 (fused
    // CT kernels); the state as in VALL_2 / VALL_1
    {9, {19, 20, 21, 14, 15, 13, 16, 17, 18}, {-1, -1, -1, 1, 2, 0, 3, 4, 5}},
    {9, {19, 20, 21, 1, 2, 0, 3, 4, 5}, {-1, -1, -1, 14, 15, 13, 16, 17, 18}},
    // the corrector's first velocity_bcs, into the second copy
    {3, {16, 17, 18}, {3, 4, 5}},
    // boundary_conditions after the fused y sweep, B in bxy..bzy; the state
    // as in VALL_2 / VALL_1
    {9, {22, 23, 24, 14, 15, 13, 16, 17, 18}, {-1, -1, -1, 1, 2, 0, 3, 4, 5}},
    {9, {22, 23, 24, 1, 2, 0, 3, 4, 5}, {-1, -1, -1, 14, 15, 13, 16, 17, 18}},
};

void drop_btabs(Group &gp)
{
    for (auto &kv : gp.btabs) {
        if (kv.second.sets) (void)hipFree(kv.second.sets);
        if (kv.second.sets3) (void)hipFree(kv.second.sets3);
        if (kv.second.boff) (void)hipFree(kv.second.boff);
    }
    gp.btabs.clear();
}

// Boundary-fill table of variable group vg along dim over a member subset
// (0 all, 1 resistive): each member's BC descriptors and block offsets.
Group::BTab &btable(Group &gp, int vg, int dim, int subset)
{
    const int key = (vg * 3 + dim) * 2 + subset;
    auto it = gp.btabs.find({key, 0});
    if (it != gp.btabs.end()) return it->second;
    Group::BTab t;
    std::vector<const Grid *> mem;
    for (size_t m = 0; m < gp.mem.size(); ++m)
        if (subset == 0 || gp.mem[m]->prm.resistive) mem.push_back(gp.mem[m]);
    std::vector<BCSetE> sets(mem.size());
    std::vector<int> boff(mem.size());
    long total = 0;
    const int da = dim == 0 ? 1 : 0, db = dim == 2 ? 1 : 2;
    for (size_t m = 0; m < mem.size(); ++m) {
        const Grid &g = *mem[m];
        BCSetE &S = sets[m];
        std::memset(&S, 0, sizeof(S));
        for (int i = 0; i < kGroups[vg].n; ++i) {
            const int v = kGroups[vg].dst[i], u = base_var(v), src = kGroups[vg].src[i];
            BCVarE &b = S.v[i];
            b.p = field_ptr(g, v);
            // The carry-over of the other copy (first pass only) matters only
            // for a variable with an edge without a condition: when every
            // edge has one, each cell outside the update range is rewritten
            // by the fill of the last dimension in whose ghost layers it
            // lies, from values that derive from the updated cells alone.
            bool none_edge = false;
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 2; ++e) none_edge = none_edge || !edge_writes_all(g, u, d, e);
            b.src = (src >= 0 && dim == 0 && none_edge) ? field_ptr(g, src) : nullptr;
            b.f64 = (u == 12 || !g.fp32) ? 1 : 0;
            b.sd = kStagE[u][dim];
            b.sa = kStagE[u][da];
            b.sb = kStagE[u][db];
            b.k0 = g.bck[u][dim][0];
            b.k1 = g.bck[u][dim][1];
            b.v0 = g.bcv[u][dim][0];
            b.v1 = g.bcv[u][dim][1];
            for (int j = 0; j < 3; ++j) {
                b.m0[j] = g.bcm[u][dim][0][j];
                b.m1[j] = g.bcm[u][dim][1][j];
            }
            if (b.k0 || b.k1 || b.src) t.any = true;
        }
        const int n3[3] = {g.dev.g.nx, g.dev.g.ny, g.dev.g.nz};
        boff[m] = (int)total;
        total += (long)kGroups[vg].n * (((n3[da] + 5) * (n3[db] + 5) + 255) / 256);
    }
    t.nblk = (int)total;
    if (!mem.empty()) {
        (void)hipMalloc(&t.sets, sets.size() * sizeof(BCSetE));
        (void)hipMemcpyAsync(t.sets, sets.data(), sets.size() * sizeof(BCSetE), hipMemcpyHostToDevice, gp.stream);
        (void)hipMalloc(&t.boff, boff.size() * sizeof(int));
        (void)hipMemcpyAsync(t.boff, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice, gp.stream);
        (void)hipStreamSynchronize(gp.stream);
    }
    return gp.btabs.emplace(std::make_pair(key, 0), t).first->second;
}

// All-dimension fill table of variable group vg over a member subset, or
// one with nblk = -1 when some member needs the per-dim fills: a variable
// whose fill carries values over from the other state copy (an edge
// without a condition), or a staggered line with a periodic lower and a
// clamped upper edge.
Group::BTab &btable3(Group &gp, int vg, int subset, int only_dim = -1)
{
    const int key = (vg * 2 + subset) * 4 + (only_dim + 1);
    auto it = gp.btabs.find({key, 3});
    if (it != gp.btabs.end()) return it->second;
    Group::BTab t;
    std::vector<const Grid *> mem;
    for (size_t m = 0; m < gp.mem.size(); ++m)
        if (subset == 0 || gp.mem[m]->prm.resistive) mem.push_back(gp.mem[m]);
    std::vector<BCSet3> sets(mem.size());
    std::vector<int> boff(mem.size());
    long total = 0;
    bool ok = true;
    for (size_t m = 0; m < mem.size(); ++m) {
        const Grid &g = *mem[m];
        BCSet3 &S = sets[m];
        std::memset(&S, 0, sizeof(S));
        S.nv = kGroups[vg].n;
        const int n3[3] = {g.dev.g.nx, g.dev.g.ny, g.dev.g.nz};
        int nb = 0;
        for (int i = 0; i < kGroups[vg].n; ++i) {
            const int v = kGroups[vg].dst[i], u = base_var(v), src = kGroups[vg].src[i];
            BCVar3 &b = S.v[i];
            b.p = field_ptr(g, v);
            b.f64 = (u == 12 || !g.fp32) ? 1 : 0;
            bool none_edge = false;
            for (int d = 0; d < 3; ++d) {
                b.s[d] = kStagE[u][d];
                // one dimension's fill: the other dimensions have no condition
                const bool on = only_dim < 0 || d == only_dim;
                b.k0[d] = on ? g.bck[u][d][0] : 0;
                b.k1[d] = on ? g.bck[u][d][1] : 0;
                b.v0[d] = g.bcv[u][d][0];
                b.v1[d] = g.bcv[u][d][1];
                for (int j = 0; j < 3; ++j) {
                    b.m0[d][j] = g.bcm[u][d][0][j];
                    b.m1[d][j] = g.bcm[u][d][1][j];
                }
                if (!on) continue;
                none_edge = none_edge || !edge_writes_all(g, u, d, 0) || !edge_writes_all(g, u, d, 1);
                if (b.s[d] && b.k0[d] == 1 && b.k1[d] == 4) ok = false;
                if (map_reads_written(g, u, d, n3[d])) ok = false;
                if (n3[d] < 3) ok = false;  // two-cell lines: fills read layers the other edge writes
                if (b.k0[d] || b.k1[d]) t.any = true;
            }
            if (src >= 0 && none_edge) ok = false;
            long nsh = 1, inner = 1;
            for (int d = 0; d < 3; ++d) {
                nsh *= n3[d] + 3 - (b.s[d] ? -2 : -1);
                inner *= n3[d] - b.s[d];
            }
            b.nsh = (int)(nsh - inner);
            b.b0 = nb;
            nb += (b.nsh + 256 * kBC3Cells - 1) / (256 * kBC3Cells);
        }
        boff[m] = (int)total;
        total += nb;
    }
    t.nblk = ok ? (int)total : -1;
    if (!mem.empty() && ok) {
        (void)hipMalloc(&t.sets3, sets.size() * sizeof(BCSet3));
        (void)hipMemcpyAsync(t.sets3, sets.data(), sets.size() * sizeof(BCSet3), hipMemcpyHostToDevice, gp.stream);
        (void)hipMalloc(&t.boff, boff.size() * sizeof(int));
        (void)hipMemcpyAsync(t.boff, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice, gp.stream);
        (void)hipStreamSynchronize(gp.stream);
    }
    return gp.btabs.emplace(std::make_pair(key, 3), t).first->second;
}

// variableRegistry::applyBoundaryConditions for a variable group: dims 0, 1,
// 2 in turn (variableDef.h:601-605), or one dim.
template <class T>
void ebcs(Group &gp, int vg, int only_dim = -1, int subset = 0)
{
    // the fill of all three dimensions, or of one, as one pass over the ghost
    // cells (e_bc3) where that pass equals the fills one dimension and edge
    // at a time; otherwise those fills (e_bc)
    if (gp.bc_all) {
        Group::BTab &t = btable3(gp, vg, subset, only_dim);
        if (t.nblk >= 0) {
            if (gp.planning || !t.any || t.nblk == 0) return;
            Prof *pf = gp.prof;
            Prof::Rec rec{only_dim < 0 ? std::string("e_bc3") : std::string("e_bc<") + char('0' + only_dim) + ">", 0.0,
                          nullptr, nullptr};
            if (pf && pf->on) {
                (void)hipEventCreate(&rec.a);
                (void)hipEventCreate(&rec.b);
                (void)hipEventRecord(rec.a, gp.stream);
            }
            hipLaunchKernelGGL(e_bc3<T>, dim3((unsigned)t.nblk), dim3(256), 0, gp.stream,
                               (const EDev<T> *)gp.devs[subset], (const BCSet3 *)t.sets3, (const int *)t.boff,
                               gp.n[subset], t.nblk);
            if (pf && pf->on) {
                (void)hipEventRecord(rec.b, gp.stream);
                pf->recs.push_back(rec);
            }
            return;
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (only_dim >= 0 && d != only_dim) continue;
        Group::BTab &t = btable(gp, vg, d, subset);
        if (gp.planning || !t.any || t.nblk == 0) continue;
        Prof *pf = gp.prof;
        Prof::Rec rec{std::string("e_bc<") + char('0' + d) + ">", 0.0, nullptr, nullptr};
        if (pf && pf->on) {
            (void)hipEventCreate(&rec.a);
            (void)hipEventCreate(&rec.b);
            (void)hipEventRecord(rec.a, gp.stream);
        }
        hipLaunchKernelGGL(e_bc<T>, dim3((unsigned)t.nblk), dim3(256), 0, gp.stream,
                           (const EDev<T> *)gp.devs[subset], (const BCSetE *)t.sets, (const int *)t.boff,
                           gp.n[subset], t.nblk, d, kGroups[vg].n);
        if (pf && pf->on) {
            (void)hipEventRecord(rec.b, gp.stream);
            pf->recs.push_back(rec);
        }
    }
}

// Rebuild the boundary tables if a member's conditions changed.
void check_bcs(Group &gp)
{
    bool changed = gp.bc_seen.size() != gp.mem.size();
    for (size_t m = 0; !changed && m < gp.mem.size(); ++m) changed = gp.bc_seen[m] != gp.mem[m]->bc_version;
    if (!changed) return;
    drop_btabs(gp);
    gp.bc_seen.resize(gp.mem.size());
    for (size_t m = 0; m < gp.mem.size(); ++m) gp.bc_seen[m] = gp.mem[m]->bc_version;
}

// ------------------------------------------------------------ the events

// Cells per launch of subset (0 all members, 1 resistive ones).  The lane-
// walking kernels (x/y march, z-line, z-line x-march) give each lane several
// cells, so below a few hundred thousand cells a launch would not fill the chip and
// the one-cell-per-lane kernels are used instead.
double subset_cells(const Group &gp, int subset)
{
    double n = 0.0;
    for (const Grid *g : gp.mem)
        if (subset == 0 || (subset == 1 && g->prm.resistive) || (subset == 2 && g->prm.rke))
            n += (double)g->dev.g.nx * g->dev.g.ny * g->dev.g.nz;
    return n;
}
bool walk_ok(const Group &gp, int subset)
{
    if (gp.walk_mode) return gp.walk_mode == 1;
    return subset_cells(gp, subset) >= gp.walk_min_cells;
}
bool use_zx(const Group &gp, int subset) { return gp.zx_k > 0 && walk_ok(gp, subset); }
bool use_walk(const Group &gp) { return walk_ok(gp, 0); }
int zxk(const Group &gp) { return gp.zx_k >= 16 ? 16 : 8; }
bool ct_fused(const Group &gp) { return gp.ct_fuse != 0 && use_zx(gp, 0); }


template <class T>
void start_a_t(Group &gp)
{
    // bytes per cell: distinct arrays read and written, sizeof(T) each
    // (FP64 flux arrays 8)
    const double w = (double)sizeof(T);
    const RSpec C = rs(-1, 2, -1, 2, -1, 2);
    const RSpec V = rs(-1, 1, -1, 1, -1, 1);
    if (gp.csv_fused && use_zx(gp, 0)) {
        // g_cs and g_vertex in one z-line x-march over the centred extent
        const int K = zxk(gp);
        elaunch_(K == 16 ? g_csv_zx<T, 16> : g_csv_zx<T, 8>, "g_csv", 8 * w, gp, 0, C, 0, kZX + 100 + K);
    } else {
        EL(g_cs<T>, 7 * w, gp, 0, C);
        EL(g_vertex<T>, 4 * w, gp, 0, V);
    }
    EL(g_alpha<T>, 8 * w, gp, 0, rs(-1, 1, -1, 2, 0, 2));
    if (use_zx(gp, 0))
        elaunch_(gp.zx_k >= 16 ? g_visc_dt_zx<T, 16> : g_visc_dt_zx<T, 8>, "g_visc_dt", 16 * w, gp, 0,
                 rs(0, 1, 0, 1, 0, 1), 0, kZX + 100 + zxk(gp));
    else
        EL(g_visc_dt<T>, 16 * w, gp, 0, rs(0, 1, 0, 1, 0, 1));
    per_grid<T>(g_dt_final<T>, "g_dt_final", gp);
}

template <class T>
void start_b_t(Group &gp, int substeps)
{
    const double w = (double)sizeof(T);
    const RSpec C = rs(-1, 2, -1, 2, -1, 2);
    if (gp.n[1] > 0 && use_zx(gp, 1)) {
        // B alternates between its arrays and bx1..bz1 and ends in its own;
        // the last substep reads bx1..bz1, which then hold B at its start
        // (the copy of lagran.cpp:482).  An odd count starts from that copy.
        if (substeps % 2) EL(r_copy<T>, 6 * w, gp, 1, C);
        // eta_calc's current-dependent branch only where some grid has it
        bool eta_var = false;
        for (int i : gp.res) {
            const Grid &g = *gp.mem[i];
            eta_var = eta_var || !(g.fp32 ? g.e32.p.eta_fixed : g.e64.p.eta_fixed);
        }
        for (int i = 0; i < substeps; ++i) {
            const int to_bx1 = (substeps - 1 - i) % 2;
            void (*k)(Src<T>, int) = eta_var ? (gp.zx_k >= 16 ? r_substep_zx<T, 16, true> : r_substep_zx<T, 8, true>)
                                             : (gp.zx_k >= 16 ? r_substep_zx<T, 16, false> : r_substep_zx<T, 8, false>);
            elaunch_(k, "r_substep", 9 * w, gp, 1, rs(0, 0, 0, 0, 0, 0), 0, kZX + 100 + zxk(gp), to_bx1);
            ebcs<T>(gp, to_bx1 ? VBE_1 : VBE_0, -1, 1);
        }
    } else if (gp.n[1] > 0) {
        for (int i = 0; i < substeps; ++i) {
            if (i == substeps - 1) EL(r_copy<T>, 6 * w, gp, 1, C);
            EL(r_rkstep<T>, 4 * w + 3 * 8, gp, 1, rs(0, 0, 0, 0, 0, 0));
            EL(r_b_energy<T>, 10 * w + 3 * 8, gp, 1, rs(0, 0, 0, 0, 0, 0));
            ebcs<T>(gp, VBE, -1, 1);
        }
    }
    if (use_zx(gp, 0))
        elaunch_(gp.zx_k >= 16 ? g_predict_zx<T, 16> : g_predict_zx<T, 8>, "g_predict", 16 * w, gp, 0, C, 0,
                 kZX + 100 + zxk(gp));
    else
        EL(g_predict<T>, 16 * w, gp, 0, C);
    EL(g_forces<T>, 19 * w, gp, 0, rs(0, 0, 0, 0, 0, 0));
    ebcs<T>(gp, VV1);
}

// The corrector writes rho, e_e, e_i, v into the second state copy; the
// x sweep moves them back to the first, y to the second, z to the first.
template <class T>
void half_t(Group &gp)
{
    const double w = (double)sizeof(T);
    // reads v v1 alpha1..3 rho e_e e_i p_e p_i; writes rho e_e e_i
    if (use_zx(gp, 0))
        elaunch_(gp.zx_k >= 16 ? g_correct_zx<T, 16> : g_correct_zx<T, 8>, "g_correct", 17 * w, gp, 0,
                 rs(0, 0, 0, 0, 0, 0), 0, kZX + 100 + zxk(gp));
    else
        EL(g_correct<T>, 17 * w, gp, 0, rs(0, 0, 0, 0, 0, 0));
    // velocity_bcs, energy_bcs, density_bcs (independent variables).  The
    // reference fills the velocities twice (lagran.cpp:759 and 803); the
    // second fill repeats the first except on a staggered line with a
    // periodic lower and a clamped upper edge, where the lower fill then
    // copies the clamped boundary point, so only then is the first issued.
    bool twice = false;
    for (const Grid *g : gp.mem)
        for (int v = 3; v <= 5; ++v)
            for (int d = 0; d < 3; ++d)
                twice = twice || (kStagE[v][d] && g->bck[v][d][0] == 1 && g->bck[v][d][1] == 4);
    if (twice) ebcs<T>(gp, VV_2);
    ebcs<T>(gp, VCORR_2);
    for (Grid *g : gp.mem) g->st = 1;
}

template <class T, int DIR>
void sweep_t(Group &gp, double xpass, double ypass, double zpass, int st)
{
    const double w = (double)sizeof(T);
    const RSpec all0 = rs(0, 0, 0, 0, 0, 0);
    if (gp.prof) gp.prof->dir = DIR;
    // R1 of the three sweeps at once in the x sweep (f_prep_all); the y and
    // z sweeps' kernels see their sets through view bits 4 and 8
    const int pv = gp.prep_all ? DIR << 2 : 0;
    if (!gp.prep_all)
        EL((f_prep<T, DIR>), 6 * w, gp, 0, rs(-1, 2, -1, 2, -1, 2), xpass, ypass, zpass);
    else if (DIR == 0)
        elaunch_(f_prep_all<T>, "f_prep_all", 10 * w, gp, 0, rs(-1, 2, -1, 2, -1, 2), 0, gp.order);
    // (8 cells per lane whatever zx_k: 211 against 226 us per step at 256^3
    // with 16, a kernel of two loads and one store per position)
    if (use_zx(gp, 0))
        elaunch_(f_cvc1_zx<T, 8>, "f_cvc1", 2 * w, gp, 0, rs(-1, 1, -1, 1, -1, 1), pv, kZX + 100 + 8);
    else
        elaunch_(f_cvc1<T>, "f_cvc1", 2 * w, gp, 0, rs(-1, 1, -1, 1, -1, 1), pv, gp.order);
    // fused CT (flux and update in one kernel): B moves to bx1..bz1 in the x
    // sweep, from there into bxy..bzy in the y sweep (or is updated in place
    // by the y sweep's two kernels, B view bit 1) and comes back in the z
    // sweep (reading bxy..bzy through view bit 4 after the fused y sweep)
    const bool fuse = ct_fused(gp);
    const bool yfuse = fuse && gp.ct_fuse >= 2;
    const int bst = fuse && DIR == 1 ? 2 : 0;
    const RSpec ext = rs(-2, 2, -2, 2, -2, 2);
    if (fuse && (DIR != 1 || yfuse)) {
        const int K = zxk(gp);
        // also the sweep's mass flux (rho from the current state copy: st)
        if (DIR == 0)
            elaunch_(K == 16 ? f_ctx_fused<T, 16> : f_ctx_fused<T, 8>, "f_ct<0>", 11 * w + 8, gp, 0, ext, st | pv,
                     kZX + 100 + K);
        else if (DIR == 1)
            elaunch_(K == 16 ? f_cty_fused<T, 16> : f_cty_fused<T, 8>, "f_ct<1>", 11 * w + 8, gp, 0, ext,
                     2 | 16 | st | pv, kZY + 100 + K);
        else
            elaunch_(K == 16 ? f_ctz_fused<T, 16> : f_ctz_fused<T, 8>, "f_ct<2>", 11 * w + 8, gp, 0, ext,
                     st | pv | (yfuse ? 16 : 0), kZX + 200 + K);
    } else {
        if (DIR == 2 && gp.zline && use_walk(gp)) {
            elaunch_(f_ct_zline<T>, "f_ct_fluxes<2>", 7 * w + 2 * 8, gp, 0, all0, bst | pv, kZLine);
        } else if (DIR == 1 && gp.march_k > 0 && use_walk(gp)) {
            // with the y mass flux, over the union of the two ranges
            const int K = gp.march_k >= 16 ? 16 : 8;
            elaunch_(K == 16 ? f_cty_mass<T, 16> : f_cty_mass<T, 8>, "f_ct_mass<1>", 9 * w + 3 * 8, gp, 0,
                     rs(0, 1, -1, 2, 0, 1), bst | st | pv, kMarch + 100 * DIR + K);
        } else if (DIR == 0 && gp.march_k > 0 && use_walk(gp)) {
            const int K = gp.march_k >= 16 ? 16 : 8;
            elaunch_(K == 16 ? f_ct_march<T, 0, 16> : f_ct_march<T, 0, 8>, "f_ct_fluxes<0>", 7 * w + 2 * 8, gp, 0,
                     all0, bst | pv, kMarch + K);
        } else {
            elaunch_(f_ct_fluxes<T, DIR>, DIR == 0 ? "f_ct_fluxes<0>" : DIR == 1 ? "f_ct_fluxes<1>" : "f_ct_fluxes<2>",
                     7 * w + 2 * 8, gp, 0, all0, bst | pv, gp.order);
        }
        const char *un = DIR == 0 ? "f_ct_update<0>" : DIR == 1 ? "f_ct_update<1>" : "f_ct_update<2>";
        if (DIR == 1)
            elaunch_(f_ct_update<T, DIR>, un, 6 * w + 2 * 8, gp, 0, all0, bst | pv, gp.order);
        else  // also the B area product (x) or quotient (z) over the B extents
            elaunch_(f_ct_update<T, DIR>, un, 6 * w + 2 * 8, gp, 0, ext, bst | pv, gp.order);
    }
    // mass flux: its reference range, widened to the ghost faces along DIR
    // (the fused CT kernels of the x and z sweeps compute it themselves)
    const bool mass_done = (fuse && (DIR != 1 || yfuse)) || (DIR == 1 && gp.march_k > 0 && use_walk(gp));
    if (!mass_done)
        ELS((f_mass_flux<T, DIR>), 4 * w + 8, gp, 0,
            rs(DIR == 0 ? -1 : 0, DIR == 0 ? 2 : 1, DIR == 1 ? -1 : 0, DIR == 1 ? 2 : 1, DIR == 2 ? -1 : 0,
               DIR == 2 ? 2 : 1),
            st | pv);
    ebcs<T>(gp, VDM, DIR);
    // reads rho e_e e_i v cv1 cv2 cvc1 VX1..VZ1 (sweep component) dm;
    // writes rho e_e e_i v
    if (DIR == 2 && gp.zline && use_zx(gp, 0)) {
        const int K = zxk(gp);
        elaunch_(K == 16 ? f_update_zx<T, 16> : f_update_zx<T, 8>, "f_update_zline", 16 * w + 8, gp, 0, all0,
                 st | pv, kZX + 200 + K);
    } else if (DIR == 2 || gp.march_k <= 0 || !use_walk(gp)) {
        ELS((f_update_all<T, DIR>), 16 * w + 8, gp, 0, all0, st | pv);
    } else {
        constexpr int MD = DIR == 2 ? 0 : DIR;
        const int K = gp.march_k >= 16 ? 16 : gp.march_k >= 8 ? 8 : 4;
        void (*k)(Src<T>) = K == 16  ? f_update_march<T, MD, 16>
                            : K == 8 ? f_update_march<T, MD, 8>
                                     : f_update_march<T, MD, 4>;
        elaunch_(k, DIR == 0 ? "f_update_march<0>" : "f_update_march<1>", 16 * w + 8, gp, 0, all0, st | pv,
                 kMarch + 100 * DIR + K);
    }
    if (gp.n[2]) {
        // the kinetic-energy correction's sums, before v1 is shifted
        const RSpec faces = rs(0, DIR == 0 ? -1 : 0, 0, DIR == 1 ? -1 : 0, 0, DIR == 2 ? -1 : 0);
        const RSpec cells = rs(DIR == 0 ? 1 : 0, DIR == 0 ? 0 : 1, DIR == 1 ? 1 : 0, DIR == 1 ? 0 : 1,
                               DIR == 2 ? 1 : 0, DIR == 2 ? 0 : 1);
        ELS((r_ke_dk<T, DIR>), 14 * w + 3 * 8, gp, 2, faces, st | pv);
        ELS((r_ke_add<T, DIR>), 4 * 8, gp, 2, cells, st | pv);
    }
    bool wshift = false;
    for (const Grid *g : gp.mem)
        wshift = wshift || !edge_writes_all(*g, 9 + DIR, DIR, 0) || !edge_writes_all(*g, 9 + DIR, DIR, 1);
    if (wshift) {
        const RSpec sh = rs(DIR == 0 ? -2 : 0, DIR == 0 ? 1 : 0, DIR == 1 ? -2 : 0, DIR == 1 ? 1 : 0,
                            DIR == 2 ? -2 : 0, DIR == 2 ? 1 : 0);
        ELS((f_wshift<T, DIR, 0>), w + 8, gp, 0, sh, pv);
        ELS((f_wshift<T, DIR, 1>), w + 8, gp, 0, sh, pv);
    }
    if (gp.prof) gp.prof->dir = -1;
    if (fuse && DIR == 1 && yfuse)
        ebcs<T>(gp, st ? VALL_1C : VALL_2C);
    else if (fuse && DIR != 2)
        ebcs<T>(gp, st ? VALL_1B : VALL_2B);
    else
        ebcs<T>(gp, st ? VALL_1 : VALL_2);
    for (Grid *g : gp.mem) g->st = 1 - st;
}

template <class T>
void end_t(Group &gp)
{
    const double w = (double)sizeof(T);
    if (gp.n[2]) EL(r_ke_zero<T>, 8, gp, 2, rs(-1, 2, -1, 2, -1, 2));
    sweep_t<T, 0>(gp, 1.0, 1.0, 1.0, 1);
    sweep_t<T, 1>(gp, 0.0, 1.0, 1.0, 0);
    sweep_t<T, 2>(gp, 0.0, 0.0, 1.0, 1);
    // bfield_bcs after the B quotient (remap.cpp:91-95): the z sweep's fill of
    // its nine variables has just filled B from the same values, and the
    // all-dimension fill is idempotent (it reads only cells that no fill
    // changes), so the repeat is needed only where the per-dimension fills ran
    // (a staggered line with a periodic lower and a clamped upper edge, map
    // edges that read written layers, two-cell dimensions)
    if (!gp.bc_all || btable3(gp, VALL_1, 0).nblk < 0) ebcs<T>(gp, VB);
    // energy_correction after the remap's last fill (shared_data.h:415-417);
    // no fill follows it in the reference either
    if (gp.n[2]) EL(r_ke_energy<T>, 4 * w + 8, gp, 2, rs(1, 0, 1, 0, 1, 0));
    per_grid<T>(g_step_done<T>, "g_step_done", gp);
}

}  // namespace

int group_init(Group &gp, const std::vector<Grid *> &members, hipStream_t stream, Prof *prof, std::string &err)
{
    group_free(gp);
    gp.mem = members;
    gp.stream = stream;
    gp.prof = prof;
    if (const char *o = std::getenv("LARE_ORDER")) gp.order = std::atoi(o) != 0;
    if (const char *o = std::getenv("LARE_MARCH")) gp.march_k = std::atoi(o);
    if (const char *o = std::getenv("LARE_ZLINE")) gp.zline = std::atoi(o) != 0;
    if (const char *o = std::getenv("LARE_WALK_MIN_CELLS")) gp.walk_min_cells = std::atof(o);
    if (const char *o = std::getenv("LARE_ZX")) gp.zx_k = std::atoi(o);
    if (const char *o = std::getenv("LARE_CTFUSE")) gp.ct_fuse = std::atoi(o);
    if (const char *o = std::getenv("LARE_PREPALL")) gp.prep_all = std::atoi(o) != 0;
    if (const char *o = std::getenv("LARE_BCALL")) gp.bc_all = std::atoi(o) != 0;
    if (const char *o = std::getenv("LARE_CSVFUSE")) gp.csv_fused = std::atoi(o) != 0;
    if (members.empty()) return 0;
    gp.fp32 = members[0]->fp32;
    for (size_t m = 0; m < members.size(); ++m) {
        if (members[m]->fp32 != gp.fp32) {
            err = "a group holds grids of one storage precision";
            return -1;
        }
        if (members[m]->prm.resistive) gp.res.push_back((int)m);
        if (members[m]->prm.rke) gp.rke.push_back((int)m);
    }
    const size_t es = gp.fp32 ? sizeof(EDev<float>) : sizeof(EDev<double>);
    for (int sub = 0; sub < 3; ++sub) {
        std::vector<char> host;
        int n = 0;
        for (size_t m = 0; m < members.size(); ++m) {
            if (sub == 1 && !members[m]->prm.resistive) continue;
            if (sub == 2 && !members[m]->prm.rke) continue;
            const char *src = gp.fp32 ? (const char *)&members[m]->e32 : (const char *)&members[m]->e64;
            host.insert(host.end(), src, src + es);
            ++n;
        }
        gp.n[sub] = n;
        if (n == 0) continue;
        hipError_t e = hipMalloc(&gp.devs[sub], host.size());
        if (e == hipSuccess) e = hipMemcpyAsync(gp.devs[sub], host.data(), host.size(), hipMemcpyHostToDevice, gp.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(gp.stream);
        if (e != hipSuccess) {
            err = std::string("group device views: ") + hipGetErrorString(e);
            group_free(gp);
            return -3;
        }
    }
    return 0;
}

void group_drop_graph(Group &gp)
{
    if (gp.graph_exec) (void)hipGraphExecDestroy(gp.graph_exec);
    if (gp.graph) (void)hipGraphDestroy(gp.graph);
    gp.graph_exec = nullptr;
    gp.graph = nullptr;
}

void group_free(Group &gp)
{
    group_drop_graph(gp);
    for (auto &kv : gp.tabs)
        if (kv.second.boff) (void)hipFree(kv.second.boff);
    gp.tabs.clear();
    drop_btabs(gp);
    gp.bc_seen.clear();
    for (int s = 0; s < 3; ++s) {
        if (gp.devs[s]) (void)hipFree(gp.devs[s]);
        gp.devs[s] = nullptr;
        gp.n[s] = 0;
    }
    gp.mem.clear();
    gp.res.clear();
    gp.rke.clear();
}

void group_start_a(Group &gp)
{
    check_bcs(gp);
    if (gp.fp32)
        start_a_t<float>(gp);
    else
        start_a_t<double>(gp);
}

void group_start_b(Group &gp, int substeps)
{
    check_bcs(gp);
    if (gp.fp32)
        start_b_t<float>(gp, substeps);
    else
        start_b_t<double>(gp, substeps);
}

void group_half(Group &gp)
{
    check_bcs(gp);
    if (gp.fp32)
        half_t<float>(gp);
    else
        half_t<double>(gp);
}

void group_end(Group &gp)
{
    check_bcs(gp);
    if (gp.fp32)
        end_t<float>(gp);
    else
        end_t<double>(gp);
}

void group_prepare(Group &gp)
{
    gp.planning = true;
    group_step(gp);
    gp.planning = false;
}

namespace {

// cells 1..n of every member: the range of the energy sums
const RSpec kEnergyRange{1, 0, 1, 0, 1, 0};

// The two energy kernels of a group over partial sums `part` (3 per block of
// table t) into `res` (4 per member).
hipError_t launch_energy(Group &gp, const Group::Tab &t, double *part, double *res)
{
    if (gp.fp32) {
        Src<float> s{(const EDev<float> *)gp.devs[0], t.boff, gp.n[0], t.nblk, kEnergyRange, 0, 0};
        hipLaunchKernelGGL(e_energy_part<float>, dim3((unsigned)t.nblk), dim3(256), 0, gp.stream, s, part);
    } else {
        Src<double> s{(const EDev<double> *)gp.devs[0], t.boff, gp.n[0], t.nblk, kEnergyRange, 0, 0};
        hipLaunchKernelGGL(e_energy_part<double>, dim3((unsigned)t.nblk), dim3(256), 0, gp.stream, s, part);
    }
    hipLaunchKernelGGL(e_energy_sum, dim3((unsigned)gp.n[0]), dim3(64), 0, gp.stream, (const double *)part,
                       (const int *)t.boff, gp.n[0], t.nblk, res);
    return hipGetLastError();
}

// per-device address of `scratch` (hipGetSymbolAddress on that device)
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_dev;

}  // namespace

size_t energy_scratch_bytes() { return sizeof(scratch); }

int energy_scratch_address(double **out, std::string &err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0) {
        err = std::string("energy scratch: no current HIP device: ") + hipGetErrorString(e);
        return -3;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch_dev.size()) g_scratch_dev.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("energy scratch: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return -3;
        }
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

void group_energy_prepare(Group &gp)
{
    if (!gp.mem.empty()) (void)table(gp, 0, kEnergyRange, 0);
}

int group_energy_noalloc(Group &gp, double *scratch_dev, double *out, std::string &err)
{
    if (gp.mem.empty()) return 0;
    if (!scratch_dev) {
        err = "energy diagnostic without allocation: no scratch address for this device";
        return -1;
    }
    const RSpec &r = kEnergyRange;
    auto it = gp.tabs.find({0, std::array<int, 7>{r.x0, r.dx1, r.y0, r.dy1, r.z0, r.dz1, 0}});
    if (it == gp.tabs.end()) {
        err = "energy diagnostic without allocation: the launch table was not prepared at set-up";
        return -1;
    }
    const Group::Tab &t = it->second;
    // partial sums, then the totals on a 256-byte boundary
    const size_t npart = ((size_t)std::max(t.nblk, 1) * 3 + 31) / 32 * 32;
    const size_t need = (npart + 4 * gp.mem.size()) * sizeof(double);
    if (need > energy_scratch_bytes()) {
        err = "energy diagnostic without allocation: the grid needs " + std::to_string(need) +
              " bytes of scratch, more than the " + std::to_string(energy_scratch_bytes()) +
              " bytes it has (at most about 560^3 cells)";
        return -1;
    }
    double *part = scratch_dev, *res = scratch_dev + npart;
    hipError_t e = launch_energy(gp, t, part, res);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, res, 4 * gp.mem.size() * sizeof(double), hipMemcpyDeviceToHost, gp.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(gp.stream);
    if (e != hipSuccess) {
        err = std::string("energy diagnostic: ") + hipGetErrorString(e);
        return -3;
    }
    return 0;
}

int group_energy(Group &gp, std::vector<double> &out, std::string &err)
{
    out.assign(4 * gp.mem.size(), 0.0);
    if (gp.mem.empty()) return 0;
    Group::Tab &t = table(gp, 0, kEnergyRange, 0);
    double *part = nullptr, *res = nullptr;
    hipError_t e = hipMalloc(&part, (size_t)std::max(t.nblk, 1) * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&res, out.size() * sizeof(double));
    if (e == hipSuccess) e = launch_energy(gp, t, part, res);
    if (e == hipSuccess) e = hipMemcpyAsync(out.data(), res, out.size() * sizeof(double), hipMemcpyDeviceToHost, gp.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(gp.stream);
    if (part) (void)hipFree(part);
    if (res) (void)hipFree(res);
    if (e != hipSuccess) {
        err = std::string("energy diagnostic: ") + hipGetErrorString(e);
        return -3;
    }
    return 0;
}

// The cell volume CV over the whole box of one grid (ix, iy, iz in -2..n+2).
template <class T>
KERNEL void k_fill_volume(const EDev<T> D)
{
    const Geo G = D.g;
    const ArrE<T> &A = D.a;
    const int ly = G.ny + 5, lz = G.nz + 5;
    const long n = (long)(G.nx + 5) * ly * lz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ix = (int)(t / ((long)ly * lz)), r = (int)(t - (long)ix * ly * lz);
    const int iy = r / lz, iz = r - iy * lz;
    A.vol[IX(ix - 2, iy - 2, iz - 2)] = CV(ix - 2, iy - 2, iz - 2);
}

int grid_fill_volume(Grid &g, std::string &err)
{
    const Geo &G = g.fp32 ? g.e32.g : g.e64.g;
    const long n = (long)(G.nx + 5) * (G.ny + 5) * (G.nz + 5);
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (g.fp32)
        hipLaunchKernelGGL(k_fill_volume<float>, dim3(nb), dim3(256), 0, g.stream, g.e32);
    else
        hipLaunchKernelGGL(k_fill_volume<double>, dim3(nb), dim3(256), 0, g.stream, g.e64);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) {
        err = std::string("cell-volume fill: ") + hipGetErrorString(e);
        return -3;
    }
    return 0;
}

// Arithmetic self-check: q[i] = a[i] / d[i] as the kernels form it, by ddiv
// (mode 0) or by a shared-divisor QDiv (mode 1), or q[i] = sqrt(a[i]) (mode 2).
KERNEL void k_div_check(const double *a, const double *d, double *q, long n, int mode)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    q[i] = mode == 0 ? ddiv(a[i], d[i]) : mode == 1 ? qdiv(a[i], qdiv_of(d[i])) : sqrt(a[i]);
}

int div_check(const double *a, const double *d, double *q, long n, int mode, std::string &err)
{
    if (n <= 0) return 0;
    const size_t bytes = (size_t)n * sizeof(double);
    double *da = nullptr, *dd = nullptr, *dq = nullptr;
    hipError_t e = hipMalloc(&da, bytes);
    if (e == hipSuccess) e = hipMalloc(&dd, bytes);
    if (e == hipSuccess) e = hipMalloc(&dq, bytes);
    if (e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dd, d, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_div_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (const double *)da,
                           (const double *)dd, dq, n, mode);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(q, dq, bytes, hipMemcpyDeviceToHost);
    if (da) (void)hipFree(da);
    if (dd) (void)hipFree(dd);
    if (dq) (void)hipFree(dq);
    if (e != hipSuccess) {
        err = std::string("division check: ") + hipGetErrorString(e);
        return -3;
    }
    return 0;
}

}  // namespace lare
