/*
 * lare_oracle.c -- CPU restatement of the reference LARE3D Lagrangian-remap
 * step.  TEST INFRASTRUCTURE ONLY (see lare_oracle.h): the checker for the
 * MI355X kernels and the "port" CPU baseline; never part of the product path.
 *
 * Every function below cites the reference function it restates.  Ranges,
 * array extents, the order of kernels and boundary fills, and the
 * floating-point association of every expression follow the reference
 * source; the library is compiled with -ffp-contract=off so that no
 * multiply-add is fused.  Geometry (shared_data.h:38-43) is Cartesian,
 * cylindrical or spherical: the metric factors hy, hyc, hz, hzc, hz1, hz2 of
 * setup.cpp:276-335 (1.0 in Cartesian geometry, where every product with
 * them is exact) enter the metric, the edge viscosities, the viscous heating
 * and forces, the CFL lengths, the predictor's forces (with the geometric
 * terms of lagran.cpp:711-721) and the y and z sweeps' widths exactly where
 * the reference multiplies or divides by them.  Positions x,y,z
 * (lagran.cpp:724-726, remap.cpp:100-103) are dead and omitted.
 *
 * min/max follow include/pp/utils.h:277-335: two-argument forms are the
 * ternaries a<b?a:b / a>b?a:b, variadic forms fold left, initializer-list
 * forms scan from the first element with a strict comparison.
 */
#include "lare_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double *p; long l0, l1, l2, n0, n1, n2; } a3;
typedef struct { double *p; long l, n; } a1;

/* Row-major, last index fastest, offsets from the array's own lower bounds
 * (include/pp/array.h:72-95).  No bounds check: an index one below a bound
 * addresses the previous row exactly as the reference's flat offset does. */
#ifdef LO_BOUNDS_CHECK
#include <stdio.h>
static long lo_chk(const char *nm, long l, long n, long i, int line)
{
    if (i < l || i >= l + n) fprintf(stderr, "OOB %s idx %ld range [%ld,%ld] line %d\n", nm, i, l, l + n - 1, line);
    return i;
}
#define A(a, i, j, k)                                                                                        \
    ((a).p[((lo_chk(#a, (a).l0, (a).n0, (long)(i), __LINE__) - (a).l0) * (a).n1 +                              \
            (lo_chk(#a, (a).l1, (a).n1, (long)(j), __LINE__) - (a).l1)) * (a).n2 +                             \
           (lo_chk(#a, (a).l2, (a).n2, (long)(k), __LINE__) - (a).l2)])
#else
#define A(a, i, j, k) ((a).p[(((long)(i) - (a).l0) * (a).n1 + ((long)(j) - (a).l1)) * (a).n2 + ((long)(k) - (a).l2)])
#endif
#define L(a, i) ((a).p[(long)(i) - (a).l])
/* Element s cells along the sweep direction (di,dj,dk) from (ix,iy,iz). */
#define AP(a, s) A(a, ix + (s) * di, iy + (s) * dj, iz + (s) * dk)

#define FOR3(ix, x0, x1, iy, y0, y1, iz, z0, z1)                              \
    _Pragma("omp parallel for schedule(static)")                              \
    for (long ix = (x0); ix <= (x1); ++ix)                                    \
        for (long iy = (y0); iy <= (y1); ++iy)                                \
            for (long iz = (z0); iz <= (z1); ++iz)

static const double sixth = 1.0 / 6.0; /* constants.h:46 */

static inline double pmin(double a, double b) { return a < b ? a : b; }
static inline double pmax(double a, double b) { return a > b ? a : b; }
static inline double pmin_list3(double a, double b, double c)
{
    double m = a;
    if (a < m) m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}
static inline double pmin_list4(double a, double b, double c, double d)
{
    double m = a;
    if (a < m) m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    if (d < m) m = d;
    return m;
}

struct lo_ctx {
    long nx, ny, nz;
    double gas_gamma, visc1, visc2, visc2_norm, mu0, dt_multiplier;
    double none_zero, largest_number;
    int resistive, rke, geometry;
    double eta_background, eta0, j_max;
    double dt, dtr;
    double xpass, ypass, zpass;
    int bck[LO_NVAR][3][2];
    double bcv[LO_NVAR][3][2];
    int bcm[LO_NVAR][3][2][3]; /* layer tables of LO_BC_MAP edges */

    a1 xb, yb, zb, xc, yc, zc, dxb, dyb, dzb, dxc, dyc, dzc;
    /* metric factors (setup.cpp:276-335): hy(-2..nx+2), hyc(-1..nx+2);
     * hz(-2..nx+2, -2..ny+2), hzc(-1..nx+2, -1..ny+2), hz1(-2..nx+2,
     * -1..ny+2), hz2(-1..nx+2, -2..ny+2), 2D arrays with a unit z extent */
    a1 hy, hyc;
    a3 hz, hzc, hz1, hz2;
    /* registered variables (setup.cpp:38-70) */
    a3 var[LO_NVAR];
    /* package arrays (setup.cpp:170-212) */
    a3 p_visc, eta, dxab, dyab, dzab, dxac, dyac, dzac, cv, cv1, cvc;
    a3 bx1, by1, bz1, alpha1, alpha2, alpha3, visc_heat, pressure, p_e, p_i;
    a3 rho_v, cv_v, fx, fy, fz, fx_visc, fy_visc, fz_visc;
    a3 flux_x, flux_y, flux_z, curlb;
    /* shock_viscosity temporaries (lagran.cpp:185-186) */
    a3 cs, cs_v;
    /* remap scratch (remap.cpp:30-35, xremap.cpp:22, yremap.cpp:22, zremap.cpp:23) */
    a3 rho1, cv2, cvc1, db1, rrho_v, rho_v1, rflux_x, rflux_y, rflux_z;
    /* remap kinetic-energy correction (setup.cpp:249-252), when rke */
    a3 delta_ke;
};

#define RHO(c) ((c)->var[LO_RHO])
#define EE(c) ((c)->var[LO_EE])
#define EI(c) ((c)->var[LO_EI])
#define VX(c) ((c)->var[LO_VX])
#define VY(c) ((c)->var[LO_VY])
#define VZ(c) ((c)->var[LO_VZ])
#define BX(c) ((c)->var[LO_BX])
#define BY(c) ((c)->var[LO_BY])
#define BZ(c) ((c)->var[LO_BZ])
#define VX1(c) ((c)->var[LO_VX1])
#define VY1(c) ((c)->var[LO_VY1])
#define VZ1(c) ((c)->var[LO_VZ1])
#define DM(c) ((c)->var[LO_DM])

static a3 a3_new(long l0, long u0, long l1, long u1, long l2, long u2, double fill)
{
    a3 a;
    a.l0 = l0; a.l1 = l1; a.l2 = l2;
    a.n0 = u0 - l0 + 1; a.n1 = u1 - l1 + 1; a.n2 = u2 - l2 + 1;
    size_t n = (size_t)a.n0 * (size_t)a.n1 * (size_t)a.n2;
    a.p = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) a.p[i] = fill;
    return a;
}
static a1 a1_new(long l, long u)
{
    a1 a;
    a.l = l; a.n = u - l + 1;
    a.p = (double *)calloc((size_t)a.n, sizeof(double));
    return a;
}
static void a3_fill(a3 a, double v)
{
    size_t n = (size_t)a.n0 * (size_t)a.n1 * (size_t)a.n2;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) a.p[i] = v;
}
static void a3_copy(a3 dst, a3 src)
{
    size_t n = (size_t)dst.n0 * (size_t)dst.n1 * (size_t)dst.n2;
    memcpy(dst.p, src.p, n * sizeof(double));
}

/* ------------------------------------------------------------------ axes */

/* axisRegistry::setDomain (axisRegistry.h:973-984) with createLinearAxis
 * (393-400), createCentredAxis (408-431) and the delta construction of
 * createPositionAxisCore (176-201): non-periodic axes copy the first delta
 * from its neighbour. */
static void make_axis(long n, double mn, double mx, a1 *xb, a1 *xc, a1 *dxb, a1 *dxc)
{
    *xb = a1_new(-2, n + 2);
    *xc = a1_new(-1, n + 2);
    *dxb = a1_new(-2, n + 2);
    *dxc = a1_new(-1, n + 2);
    double delta = (mx - mn) / (double)(n + 1 - 1);
    for (long i = -2; i <= n + 2; ++i) L(*xb, i) = mn + delta * (double)i;
    for (long i = -1; i <= n + 2; ++i) L(*dxb, i) = L(*xb, i) - L(*xb, i - 1);
    L(*dxb, -2) = L(*dxb, -1);
    for (long i = -1; i <= n + 2; ++i) L(*xc, i) = 0.5 * (L(*xb, i) + L(*xb, i - 1));
    for (long i = 0; i <= n + 2; ++i) L(*dxc, i) = L(*xc, i) - L(*xc, i - 1);
    L(*dxc, -1) = L(*dxc, 0);
}

#define H2(a, i, j) A(a, i, j, 0)

/* Metric factors of LARE3D::grid (setup.cpp:276-335).  Cylindrical and
 * spherical: hy = |xb|, hyc = |xc| (the reference's loop also writes
 * hyc(-2), outside hyc's extent, from xc(-2), outside xc's; nothing reads
 * it, so it is left out); spherical: hz = hy |sin yb|, hz2 = hyc |sin yb|
 * over y -2..ny+2 and hz1 = hy |sin yc|, hzc = hyc |sin yc| over y
 * -1..ny+2; 1 elsewhere.  The node volume factor hyv only enters cvc, which
 * nothing reads. */
static void make_geometry(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny;
    c->hy = a1_new(-2, nx + 2);
    c->hyc = a1_new(-1, nx + 2);
    c->hz = a3_new(-2, nx + 2, -2, ny + 2, 0, 0, 1.0);
    c->hzc = a3_new(-1, nx + 2, -1, ny + 2, 0, 0, 1.0);
    c->hz1 = a3_new(-2, nx + 2, -1, ny + 2, 0, 0, 1.0);
    c->hz2 = a3_new(-1, nx + 2, -2, ny + 2, 0, 0, 1.0);
    for (long ix = -2; ix <= nx + 2; ++ix) {
        L(c->hy, ix) = c->geometry == 0 ? 1.0 : fabs(L(c->xb, ix));
        if (ix >= -1) L(c->hyc, ix) = c->geometry == 0 ? 1.0 : fabs(L(c->xc, ix));
    }
    if (c->geometry != 2) return;
    for (long iy = -2; iy <= ny + 2; ++iy) {
        double s = fabs(sin(L(c->yb, iy)));
        for (long ix = -2; ix <= nx + 2; ++ix) H2(c->hz, ix, iy) = L(c->hy, ix) * s;
        for (long ix = -1; ix <= nx + 2; ++ix) H2(c->hz2, ix, iy) = L(c->hyc, ix) * s;
    }
    for (long iy = -1; iy <= ny + 2; ++iy) {
        double sc = fabs(sin(L(c->yc, iy)));
        for (long ix = -2; ix <= nx + 2; ++ix) H2(c->hz1, ix, iy) = L(c->hy, ix) * sc;
        for (long ix = -1; ix <= nx + 2; ++ix) H2(c->hzc, ix, iy) = L(c->hyc, ix) * sc;
    }
}

/* LARE3D::grid (setup.cpp:337-390): areas and volumes with the metric
 * factors, including the reference's separate y = -2 and z = -2 layers
 * (setup.cpp:356-373: dzab's z = -2 layer takes hyc, not hy). */
static void make_metric(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    FOR3(ix, -2, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        double dydz = L(c->dyb, iy) * L(c->dzb, iz);
        A(c->dxab, ix, iy, iz) = dydz * L(c->hy, ix) * H2(c->hz1, ix, iy);
    }
    FOR3(ix, -1, nx + 2, iy, -2, ny + 2, iz, -1, nz + 2)
    {
        A(c->dyab, ix, iy, iz) = L(c->dxb, ix) * L(c->dzb, iz) * H2(c->hz2, ix, iy);
    }
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -2, nz + 2)
    {
        if (iz == -2)
            A(c->dzab, ix, iy, iz) = L(c->dxb, ix) * L(c->dyb, iy) * L(c->hyc, ix);
        else
            A(c->dzab, ix, iy, iz) = L(c->dxb, ix) * L(c->dyb, iy) * L(c->hy, ix);
    }
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        double dx = L(c->dxb, ix);
        double dydz = L(c->dyb, iy) * L(c->dzb, iz);
        A(c->cv, ix, iy, iz) = dx * dydz * L(c->hyc, ix) * H2(c->hzc, ix, iy);
    }
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        double dx = L(c->dxb, ix);
        double dy = L(c->dyb, iy);
        double dz = L(c->dzc, iz);
        double dydz = dy * dz;
        A(c->dxac, ix, iy, iz) = dydz * L(c->hyc, ix) * H2(c->hz2, ix, iy);
        A(c->dyac, ix, iy, iz) = dx * dz * H2(c->hz1, ix, iy);
        A(c->dzac, ix, iy, iz) = dx * dy * L(c->hy, ix);
        A(c->cvc, ix, iy, iz) = dx * dydz;
    }
}

/* --------------------------------------------------------------- set-up */

/* Stagger per dimension of each registered variable (setup.cpp:45-69):
 * 0 = CENTRED (-1..n+2, domain 1..n), 1 = HALF_CELL (-2..n+2, domain 0..n). */
static const int var_stag[LO_NVAR][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0},             /* rho, energy_electron, energy_ion */
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1},             /* vx vy vz */
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},             /* bx by bz */
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1},             /* vx1 vy1 vz1 */
    {0, 0, 0}};                                  /* dm */

lo_ctx *lo_create(const lo_desc *d)
{
    lo_ctx *c = (lo_ctx *)calloc(1, sizeof(lo_ctx));
    const long nx = d->nx, ny = d->ny, nz = d->nz;
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->gas_gamma = d->gas_gamma;
    c->visc1 = d->visc1;
    c->visc2 = d->visc2;
    c->mu0 = d->mu0;
    c->dt_multiplier = d->dt_multiplier;
    c->none_zero = DBL_EPSILON;   /* shared_data.h:92 */
    c->largest_number = DBL_MAX;  /* shared_data.h:93 */
    c->resistive = d->resistive;
    c->rke = d->rke;
    c->geometry = d->geometry;
    c->eta_background = d->eta_background;
    c->eta0 = d->eta0;
    c->j_max = d->j_max;
    c->dt = 0.0; /* simulationData is value-initialised: step 1 sees dt = 0 */
    c->dtr = 0.0;
    make_axis(nx, d->x_min, d->x_max, &c->xb, &c->xc, &c->dxb, &c->dxc);
    make_axis(ny, d->y_min, d->y_max, &c->yb, &c->yc, &c->dyb, &c->dyc);
    make_axis(nz, d->z_min, d->z_max, &c->zb, &c->zc, &c->dzb, &c->dzc);

    const long n3[3] = {nx, ny, nz};
    for (int v = 0; v < LO_NVAR; ++v) {
        long lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = var_stag[v][k] ? -2 : -1;
            hi[k] = n3[k] + 2;
        }
        c->var[v] = a3_new(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 0.0);
    }
    const double NaN = NAN;
#define CCC -1, nx + 2, -1, ny + 2, -1, nz + 2
    c->p_visc = a3_new(CCC, NaN);
    c->eta = a3_new(CCC, NaN);
    c->dxab = a3_new(-2, nx + 2, -1, ny + 2, -1, nz + 2, NaN);
    c->dyab = a3_new(-1, nx + 2, -2, ny + 2, -1, nz + 2, NaN);
    c->dzab = a3_new(-1, nx + 2, -1, ny + 2, -2, nz + 2, NaN);
    c->dxac = a3_new(CCC, NaN);
    c->dyac = a3_new(CCC, NaN);
    c->dzac = a3_new(CCC, NaN);
    c->cv = a3_new(CCC, NaN);
    c->cv1 = a3_new(CCC, NaN);
    c->cvc = a3_new(CCC, NaN);
    c->bx1 = a3_new(CCC, NaN);
    c->by1 = a3_new(CCC, NaN);
    c->bz1 = a3_new(CCC, NaN);
    c->alpha1 = a3_new(0, nx + 1, 0, ny + 2, 0, nz + 2, NaN);
    c->alpha2 = a3_new(-1, nx + 1, 0, ny + 1, 0, nz + 2, NaN);
    c->alpha3 = a3_new(-1, nx + 2, -1, ny + 2, 0, nz + 1, NaN);
    c->visc_heat = a3_new(0, nx + 1, 0, ny + 1, 0, nz + 1, NaN);
    c->pressure = a3_new(CCC, NaN);
    c->p_e = a3_new(CCC, NaN);
    c->p_i = a3_new(CCC, NaN);
    c->rho_v = a3_new(-1, nx + 1, -1, ny + 1, -1, nz + 1, NaN);
    c->cv_v = a3_new(-1, nx + 1, -1, ny + 1, -1, nz + 1, NaN);
#define D0N 0, nx, 0, ny, 0, nz
    c->fx = a3_new(D0N, NaN);
    c->fy = a3_new(D0N, NaN);
    c->fz = a3_new(D0N, NaN);
    c->fx_visc = a3_new(D0N, NaN);
    c->fy_visc = a3_new(D0N, NaN);
    c->fz_visc = a3_new(D0N, NaN);
    c->flux_x = a3_new(D0N, NaN);
    c->flux_y = a3_new(D0N, NaN);
    c->flux_z = a3_new(D0N, NaN);
    c->curlb = a3_new(D0N, NaN);
    c->cs = a3_new(CCC, NaN);
    c->cs_v = a3_new(-1, nx + 1, -1, ny + 1, -1, nz + 1, NaN);
    c->rho1 = a3_new(CCC, NaN);
    c->cv2 = a3_new(CCC, NaN);
    c->cvc1 = a3_new(CCC, NaN);
    c->db1 = a3_new(CCC, NaN);
    c->rrho_v = a3_new(CCC, NaN);
    c->rho_v1 = a3_new(CCC, NaN);
    c->rflux_x = a3_new(-2, nx + 2, -1, ny + 2, -1, nz + 2, NaN);
    c->rflux_y = a3_new(-1, nx + 2, -2, ny + 2, -1, nz + 2, NaN);
    c->rflux_z = a3_new(-1, nx + 2, -1, ny + 2, -2, nz + 2, NaN);
    /* zeroed by LARE3D::defaultVariables (control.cpp:48-51) */
    if (c->rke) c->delta_ke = a3_new(CCC, 0.0);
#undef CCC
#undef D0N
    make_geometry(c);
    make_metric(c);
    return c;
}

void lo_destroy(lo_ctx *c)
{
    if (!c) return;
    a3 *all[] = {&c->p_visc, &c->eta, &c->dxab, &c->dyab, &c->dzab, &c->dxac, &c->dyac, &c->dzac,
                 &c->cv, &c->cv1, &c->cvc, &c->bx1, &c->by1, &c->bz1, &c->alpha1, &c->alpha2,
                 &c->alpha3, &c->visc_heat, &c->pressure, &c->p_e, &c->p_i, &c->rho_v, &c->cv_v,
                 &c->fx, &c->fy, &c->fz, &c->fx_visc, &c->fy_visc, &c->fz_visc, &c->flux_x,
                 &c->flux_y, &c->flux_z, &c->curlb, &c->cs, &c->cs_v, &c->rho1, &c->cv2, &c->cvc1,
                 &c->db1, &c->rrho_v, &c->rho_v1, &c->rflux_x, &c->rflux_y, &c->rflux_z};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]->p);
    if (c->rke) free(c->delta_ke.p);
    for (int v = 0; v < LO_NVAR; ++v) free(c->var[v].p);
    a1 *ax[] = {&c->xb, &c->yb, &c->zb, &c->xc, &c->yc, &c->zc, &c->dxb, &c->dyb, &c->dzb,
                &c->dxc, &c->dyc, &c->dzc, &c->hy, &c->hyc};
    for (size_t i = 0; i < sizeof(ax) / sizeof(ax[0]); ++i) free(ax[i]->p);
    free(c->hz.p);
    free(c->hzc.p);
    free(c->hz1.p);
    free(c->hz2.p);
    free(c);
}

static double *a3_out(a3 *a, long lb[3], long ub[3])
{
    if (lb) { lb[0] = a->l0; lb[1] = a->l1; lb[2] = a->l2; }
    if (ub) { ub[0] = a->l0 + a->n0 - 1; ub[1] = a->l1 + a->n1 - 1; ub[2] = a->l2 + a->n2 - 1; }
    return a->p;
}

double *lo_field(lo_ctx *c, int id, long lb[3], long ub[3])
{
    if (id < 0 || id >= LO_NVAR) return NULL;
    return a3_out(&c->var[id], lb, ub);
}

double *lo_array(lo_ctx *c, const char *name, long lb[3], long ub[3])
{
    struct { const char *n; a3 *a; } t[] = {
        {"p_visc", &c->p_visc}, {"eta", &c->eta}, {"dxab", &c->dxab}, {"dyab", &c->dyab},
        {"dzab", &c->dzab}, {"dxac", &c->dxac}, {"dyac", &c->dyac}, {"dzac", &c->dzac},
        {"cv", &c->cv}, {"cv1", &c->cv1}, {"cvc", &c->cvc}, {"bx1", &c->bx1}, {"by1", &c->by1},
        {"bz1", &c->bz1}, {"alpha1", &c->alpha1}, {"alpha2", &c->alpha2}, {"alpha3", &c->alpha3},
        {"visc_heat", &c->visc_heat}, {"pressure", &c->pressure}, {"p_e", &c->p_e}, {"p_i", &c->p_i},
        {"rho_v", &c->rho_v}, {"cv_v", &c->cv_v}, {"fx", &c->fx}, {"fy", &c->fy}, {"fz", &c->fz},
        {"fx_visc", &c->fx_visc}, {"fy_visc", &c->fy_visc}, {"fz_visc", &c->fz_visc},
        {"cs", &c->cs}, {"cs_v", &c->cs_v}, {"cv2", &c->cv2}, {"cvc1", &c->cvc1}, {"db1", &c->db1},
        {"delta_ke", c->rke ? &c->delta_ke : NULL}};
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
        if (strcmp(t[i].n, name) == 0) return t[i].a ? a3_out(t[i].a, lb, ub) : NULL;
    return NULL;
}

double *lo_axis(lo_ctx *c, const char *name, long *lb, long *ub)
{
    struct { const char *n; a1 *a; } t[] = {
        {"xb", &c->xb}, {"yb", &c->yb}, {"zb", &c->zb}, {"xc", &c->xc}, {"yc", &c->yc},
        {"zc", &c->zc}, {"dxb", &c->dxb}, {"dyb", &c->dyb}, {"dzb", &c->dzb}, {"dxc", &c->dxc},
        {"dyc", &c->dyc}, {"dzc", &c->dzc}};
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
        if (strcmp(t[i].n, name) == 0) {
            if (lb) *lb = t[i].a->l;
            if (ub) *ub = t[i].a->l + t[i].a->n - 1;
            return t[i].a->p;
        }
    return NULL;
}

int lo_set_bc(lo_ctx *c, int var, int dim, int edge, int kind, double value)
{
    if (var < 0 || var >= LO_NVAR || dim < 0 || dim > 2 || edge < 0 || edge > 1) return -1;
    if (kind < LO_BC_NONE || kind > LO_BC_CLAMP) return -1;
    c->bck[var][dim][edge] = kind;
    c->bcv[var][dim][edge] = value;
    return 0;
}

int lo_set_bc_map(lo_ctx *c, int var, int dim, int edge, const int src[3], double value)
{
    if (var < 0 || var >= LO_NVAR || dim < 0 || dim > 2 || edge < 0 || edge > 1) return -1;
    for (int k = 0; k < 3; ++k) c->bcm[var][dim][edge][k] = src[k];
    c->bck[var][dim][edge] = LO_BC_MAP;
    c->bcv[var][dim][edge] = value;
    return 0;
}

double lo_get_dt(const lo_ctx *c) { return c->dt; }
void lo_set_dt(lo_ctx *c, double dt) { c->dt = dt; }

void lo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* --------------------------------------------------- boundary conditions */

/* One built-in BC on one edge of one dimension of a registered variable,
 * with the other two dimensions spanning the full local range including
 * ghosts (builtInBoundaryConditions.h:18-46).  The edge ranges come from
 * dimension.h:1047-1197: a CENTRED dimension's lower/upper non-domain ranges
 * are -1..0 and n+1..n+2; a HALF_CELL dimension's (ghostAdjust 1,
 * staggerRegistry.h:59) are -2..0 and n..n+2, i.e. they include the
 * boundary point itself.
 *   simplePeriodicBC  (251-300): lower dst <- Range(UB-count+1, UB),
 *                      upper dst <- Range(LB, LB+count-1), i.e. g+n / g-n
 *   simpleZeroGradientBC (181-249): ghost <- last domain point (LB or UB)
 *   simpleMirrorBC (92-172): b = (a+a)/2 at the last domain point, and
 *                      ghost g = b - (A(last - (g - firstGhost)) - b)
 *   simpleClamp (54-84): ghost = value */
static void bc_apply_edge(lo_ctx *c, int v, int dim, int edge)
{
    int kind = c->bck[v][dim][edge];
    if (kind == LO_BC_NONE) return;
    a3 a = c->var[v];
    const long n3[3] = {c->nx, c->ny, c->nz};
    const long s = var_stag[v][dim];
    const long n = n3[dim];
    const long LB = s ? 0 : 1, UB = n;
    const long g0 = edge ? n + 1 - s : (s ? -2 : -1);
    const long g1 = edge ? n + 2 : 0;
    const long first = edge ? g0 : g1; /* ghost point nearest the domain */
    const long last = edge ? UB : LB;  /* domain point mirrored about */
    long lo[3] = {a.l0, a.l1, a.l2};
    long hi[3] = {a.l0 + a.n0 - 1, a.l1 + a.n1 - 1, a.l2 + a.n2 - 1};
    lo[dim] = g0;
    hi[dim] = g1;
    const double val = c->bcv[v][dim][edge];
    if (kind == LO_BC_MAP) {
        /* a LAREBoundaryConditions function (lareBoundaryClass.h:10-34)
         * written as its effect on one edge: per ghost layer g0 + t a source
         * layer, the constant or nothing; every source is read before the
         * edge writes (the rules of src/InitialConditions/LARE/boundary.cpp
         * never read a layer they write) */
        const int *m = c->bcm[v][dim][edge];
        lo[dim] = hi[dim] = 0;
        for (long i = lo[0]; i <= hi[0]; ++i)
            for (long j = lo[1]; j <= hi[1]; ++j)
                for (long k = lo[2]; k <= hi[2]; ++k) {
                    double out[3] = {0.0, 0.0, 0.0};
                    for (long g = g0; g <= g1; ++g) {
                        long s3[3] = {i, j, k};
                        s3[dim] = m[g - g0];
                        if (m[g - g0] == LO_BC_MAP_VALUE) out[g - g0] = val;
                        else if (m[g - g0] != LO_BC_MAP_KEEP) out[g - g0] = A(a, s3[0], s3[1], s3[2]);
                    }
                    for (long g = g0; g <= g1; ++g) {
                        long t3[3] = {i, j, k};
                        t3[dim] = g;
                        if (m[g - g0] != LO_BC_MAP_KEEP) A(a, t3[0], t3[1], t3[2]) = out[g - g0];
                    }
                }
        return;
    }
    for (long i = lo[0]; i <= hi[0]; ++i)
        for (long j = lo[1]; j <= hi[1]; ++j)
            for (long k = lo[2]; k <= hi[2]; ++k) {
                long s3[3] = {i, j, k};
                const long g = s3[dim];
                double out;
                if (kind == LO_BC_PERIODIC) {
                    s3[dim] = edge ? g - n : g + n;
                    out = A(a, s3[0], s3[1], s3[2]);
                } else if (kind == LO_BC_ZEROGRAD) {
                    s3[dim] = last;
                    out = A(a, s3[0], s3[1], s3[2]);
                } else if (kind == LO_BC_MIRROR) {
                    long b3[3] = {i, j, k};
                    b3[dim] = last;
                    const double bv = A(a, b3[0], b3[1], b3[2]);
                    const double boundary = (bv + bv) / 2.0;
                    s3[dim] = last - (g - first);
                    const double delta = A(a, s3[0], s3[1], s3[2]) - boundary;
                    out = boundary - delta;
                } else {
                    out = val;
                }
                A(a, i, j, k) = out;
            }
}

/* variableDef::applyBoundaryConditions (variableDef.h:539-605): per dim
 * lower then upper; all dims 0..2 when dim < 0. */
static void bcs(lo_ctx *c, int v, int dim)
{
    if (dim >= 0) {
        bc_apply_edge(c, v, dim, 0);
        bc_apply_edge(c, v, dim, 1);
        return;
    }
    for (int d = 0; d < 3; ++d) {
        bc_apply_edge(c, v, d, 0);
        bc_apply_edge(c, v, d, 1);
    }
}

/* boundary.cpp:27-96 */
static void bfield_bcs(lo_ctx *c) { bcs(c, LO_BX, -1); bcs(c, LO_BY, -1); bcs(c, LO_BZ, -1); }
static void energy_bcs(lo_ctx *c) { bcs(c, LO_EE, -1); bcs(c, LO_EI, -1); }
static void density_bcs(lo_ctx *c) { bcs(c, LO_RHO, -1); }
static void velocity_bcs(lo_ctx *c) { bcs(c, LO_VX, -1); bcs(c, LO_VY, -1); bcs(c, LO_VZ, -1); }
static void remap_v_bcs(lo_ctx *c) { bcs(c, LO_VX1, -1); bcs(c, LO_VY1, -1); bcs(c, LO_VZ1, -1); }
static void boundary_conditions(lo_ctx *c)
{
    bfield_bcs(c);
    energy_bcs(c);
    density_bcs(c);
    velocity_bcs(c);
}

/* ------------------------------------------------------- Lagrangian step */

/* edge_viscosity (lagran.cpp:56-93). */
static inline double edge_viscosity(const lo_ctx *c, double dvdots, double dx, double dxm, double dxp,
                                    double cs_edge, long i0, long i1, long i2, long i3, long j0, long j1,
                                    long j2, long j3, long k0, long k1, long k2, long k3)
{
    const a3 rho_v = c->rho_v, vx = c->var[LO_VX], vy = c->var[LO_VY], vz = c->var[LO_VZ];
    dvdots = pmin(0.0, dvdots);
    double rho_edge = 2.0 * A(rho_v, i1, j1, k1) * A(rho_v, i2, j2, k2) /
                      (A(rho_v, i1, j1, k1) + A(rho_v, i2, j2, k2));
    double dvx = A(vx, i1, j1, k1) - A(vx, i2, j2, k2);
    double dvy = A(vy, i1, j1, k1) - A(vy, i2, j2, k2);
    double dvz = A(vz, i1, j1, k1) - A(vz, i2, j2, k2);
    double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    double dv = sqrt(dv2);
    double psi = 0.0;
    dvdots = dv * c->dt / dx < 1.e-14 ? 0.0 : dvdots / dv;
    double dvxm = A(vx, i0, j0, k0) - A(vx, i1, j1, k1);
    double dvxp = A(vx, i2, j2, k2) - A(vx, i3, j3, k3);
    double dvym = A(vy, i0, j0, k0) - A(vy, i1, j1, k1);
    double dvyp = A(vy, i2, j2, k2) - A(vy, i3, j3, k3);
    double dvzm = A(vz, i0, j0, k0) - A(vz, i1, j1, k1);
    double dvzp = A(vz, i2, j2, k2) - A(vz, i3, j3, k3);
    double rl = 1.0, rr = 1.0;
    rl = dv * c->dt / dx < 1.e-14 ? 1.0 : (dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);
    rr = dv * c->dt / dx < 1.e-14 ? 1.0 : (dvxm * dvx + dvym * dvy + dvzm * dvz) * dx / (dxm * dv2);
    psi = pmin_list4(0.5 * (rr + rl), 2.0 * rl, 2.0 * rr, 1.0);
    psi = pmax(0.0, psi);
    double q_k_bar = rho_edge * (c->visc2_norm * dv +
                                 sqrt(c->visc2_norm * c->visc2_norm * dv2 + (c->visc1 * cs_edge) * (c->visc1 * cs_edge)));
    return q_k_bar * (1.0 - psi) * dvdots;
}

/* shock_viscosity (lagran.cpp:175-378). */
static void shock_viscosity(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 rho = RHO(c), vx = VX(c), vy = VY(c), vz = VZ(c);
    const a3 cs = c->cs, cs_v = c->cs_v, cv = c->cv;
    c->visc2_norm = 0.25 * (c->gas_gamma + 1.0) * c->visc2;
    a3_fill(c->p_visc, 0.0);
    a3_fill(c->visc_heat, 0.0);

    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        double rmin = pmax(A(rho, ix, iy, iz), c->none_zero);
        double b2 = A(c->bx1, ix, iy, iz) * A(c->bx1, ix, iy, iz) + A(c->by1, ix, iy, iz) * A(c->by1, ix, iy, iz) +
                    A(c->bz1, ix, iy, iz) * A(c->bz1, ix, iy, iz);
        double p = A(c->pressure, ix, iy, iz);
        A(cs, ix, iy, iz) = sqrt((c->gas_gamma * p + b2) / rmin);
    }
    FOR3(ix, -1, nx + 1, iy, -1, ny + 1, iz, -1, nz + 1)
    {
        long izp = iz + 1, iyp = iy + 1, ixp = ix + 1;
        double sum = A(cs, ix, iy, iz) * A(cv, ix, iy, iz) + A(cs, ixp, iy, iz) * A(cv, ixp, iy, iz) +
                     A(cs, ix, iyp, iz) * A(cv, ix, iyp, iz) + A(cs, ixp, iyp, iz) * A(cv, ixp, iyp, iz) +
                     A(cs, ix, iy, izp) * A(cv, ix, iy, izp) + A(cs, ixp, iy, izp) * A(cv, ixp, iy, izp) +
                     A(cs, ix, iyp, izp) * A(cv, ix, iyp, izp) + A(cs, ixp, iyp, izp) * A(cv, ixp, iyp, izp);
        A(cs_v, ix, iy, iz) = 0.125 * sum / A(c->cv_v, ix, iy, iz);
    }
    /* alpha1 (lagran.cpp:211-232) */
    FOR3(ix, 0, nx + 1, iy, 0, ny + 2, iz, 0, nz + 2)
    {
        long ixm = ix - 1, iym = iy - 1, izm = iz - 1, ixp = ix + 1;
        long i1 = ixm, j1 = iym, k1 = izm;
        long i2 = ix, j2 = iym, k2 = izm;
        long i0 = i1 - 1, j0 = j1, k0 = k1;
        long i3 = i2 + 1, j3 = j2, k3 = k2;
        double dx = L(c->dxb, ix), dxp = L(c->dxb, ixp), dxm = L(c->dxb, ixm);
        double dvdots = -(A(vx, i1, j1, k1) - A(vx, i2, j2, k2));
        double cs_edge = pmin(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));
        A(c->alpha1, ix, iy, iz) =
            edge_viscosity(c, dvdots, dx, dxm, dxp, cs_edge, i0, i1, i2, i3, j0, j1, j2, j3, k0, k1, k2, k3);
    }
    /* alpha2 (lagran.cpp:235-253) */
    FOR3(ix, -1, nx + 1, iy, 0, ny + 1, iz, 0, nz + 2)
    {
        long izm = iz - 1, iym = iy - 1, iyp = iy + 1;
        long i1 = ix, j1 = iym, k1 = izm;
        long i2 = ix, j2 = iy, k2 = izm;
        long i0 = i1, j0 = j1 - 1, k0 = k1;
        long i3 = i2, j3 = j2 + 1, k3 = k2;
        double dx = L(c->dyb, iy) * L(c->hy, ix), dxp = L(c->dyb, iyp) * L(c->hy, ix),
               dxm = L(c->dyb, iym) * L(c->hy, ix);
        double dvdots = -(A(vy, i1, j1, k1) - A(vy, i2, j2, k2));
        double cs_edge = pmin(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));
        A(c->alpha2, ix, iy, iz) =
            edge_viscosity(c, dvdots, dx, dxm, dxp, cs_edge, i0, i1, i2, i3, j0, j1, j2, j3, k0, k1, k2, k3);
    }
    /* alpha3 (lagran.cpp:256-274) */
    FOR3(ix, -1, nx + 1, iy, -1, ny + 1, iz, 0, nz + 1)
    {
        long izm = iz - 1, izp = iz + 1;
        long i1 = ix, j1 = iy, k1 = izm;
        long i2 = ix, j2 = iy, k2 = iz;
        long i0 = i1, j0 = j1, k0 = k1 - 1;
        long i3 = i2, j3 = j2, k3 = k2 + 1;
        double dx = L(c->dzb, iz) * H2(c->hz, ix, iy), dxp = L(c->dzb, izp) * H2(c->hz, ix, iy),
               dxm = L(c->dzb, izm) * H2(c->hz, ix, iy);
        double dvdots = -(A(vz, i1, j1, k1) - A(vz, i2, j2, k2));
        double cs_edge = pmin(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));
        A(c->alpha3, ix, iy, iz) =
            edge_viscosity(c, dvdots, dx, dxm, dxp, cs_edge, i0, i1, i2, i3, j0, j1, j2, j3, k0, k1, k2, k3);
    }
    /* p_visc and visc_heat (lagran.cpp:279-323) */
    FOR3(ix, 0, nx + 1, iy, 0, ny + 1, iz, 0, nz + 1)
    {
        long izm = iz - 1, izp = iz + 1, iym = iy - 1, iyp = iy + 1, ixm = ix - 1;
#define SQD(i, j, k, I, J, K)                                                                                   \
    ((A(vx, i, j, k) - A(vx, I, J, K)) * (A(vx, i, j, k) - A(vx, I, J, K)) +                                    \
     (A(vy, i, j, k) - A(vy, I, J, K)) * (A(vy, i, j, k) - A(vy, I, J, K)) +                                    \
     (A(vz, i, j, k) - A(vz, I, J, K)) * (A(vz, i, j, k) - A(vz, I, J, K)))
        double a1 = SQD(ixm, iym, izm, ix, iym, izm);
        double a2 = SQD(ix, iym, izm, ix, iy, izm);
        double a3v = SQD(ix, iy, izm, ixm, iy, izm);
        double a4 = SQD(ixm, iy, izm, ixm, iym, izm);
        double a5 = SQD(ixm, iym, iz, ix, iym, iz);
        double a6 = SQD(ix, iym, iz, ix, iy, iz);
        double a7 = SQD(ix, iy, iz, ixm, iy, iz);
        double a8 = SQD(ixm, iy, iz, ixm, iym, iz);
        double a9 = SQD(ix, iy, izm, ix, iy, iz);
        double a10 = SQD(ixm, iy, izm, ixm, iy, iz);
        double a11 = SQD(ixm, iym, izm, ixm, iym, iz);
        double a12 = SQD(ix, iym, izm, ix, iym, iz);
#undef SQD
        A(c->p_visc, ix, iy, iz) = pmax(A(c->p_visc, ix, iy, iz), -A(c->alpha1, ix, iy, iz) * sqrt(a1));
        A(c->p_visc, ix, iy, iz) = pmax(A(c->p_visc, ix, iy, iz), -A(c->alpha2, ix, iy, iz) * sqrt(a2));
        A(c->p_visc, ix, iy, iz) = pmax(A(c->p_visc, ix, iy, iz), -A(c->alpha3, ix, iy, iz) * sqrt(a9));
        double dx = L(c->dxb, ix);
        double dy = L(c->dyb, iy) * L(c->hyc, ix);
        double dz = L(c->dzb, iz) * H2(c->hz2, ix, iy);
        A(c->visc_heat, ix, iy, iz) =
            -0.25 * dy * dz * A(c->alpha1, ix, iy, iz) * a1
            - 0.25 * dx * dz * A(c->alpha2, ix, iy, iz) * a2
            - 0.25 * dy * dz * A(c->alpha1, ix, iyp, iz) * a3v
            - 0.25 * dx * dz * A(c->alpha2, ixm, iy, iz) * a4
            - 0.25 * dy * dz * A(c->alpha1, ix, iy, izp) * a5
            - 0.25 * dx * dz * A(c->alpha2, ix, iy, izp) * a6
            - 0.25 * dy * dz * A(c->alpha1, ix, iyp, izp) * a7
            - 0.25 * dx * dz * A(c->alpha2, ixm, iy, izp) * a8
            - 0.25 * dx * dy * A(c->alpha3, ix, iy, iz) * a9
            - 0.25 * dx * dy * A(c->alpha3, ixm, iy, iz) * a10
            - 0.25 * dx * dy * A(c->alpha3, ixm, iym, iz) * a11
            - 0.25 * dx * dy * A(c->alpha3, ix, iym, iz) * a12;
        A(c->visc_heat, ix, iy, iz) = A(c->visc_heat, ix, iy, iz) / A(cv, ix, iy, iz);
    }
    a3_fill(c->fx_visc, 0.0);
    a3_fill(c->fy_visc, 0.0);
    a3_fill(c->fz_visc, 0.0);
    /* viscous forces (lagran.cpp:330-376) */
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        long izm = iz - 1, izp = iz + 1, iym = iy - 1, iyp = iy + 1, ixm = ix - 1, ixp = ix + 1;
        double dx = L(c->dxb, ix);
        double dy = L(c->dyb, iy) * L(c->hyc, ix);
        double dz = L(c->dzb, iz) * H2(c->hz2, ix, iy);
        double a1 = A(c->alpha1, ix, iyp, izp) * dy * dz;
        double a2 = A(c->alpha1, ixp, iyp, izp) * dy * dz;
        double a3v = A(c->alpha2, ix, iy, izp) * dx * dz;
        double a4 = A(c->alpha2, ix, iyp, izp) * dx * dz;
        double a5 = A(c->alpha3, ix, iy, iz) * dx * dy;
        double a6 = A(c->alpha3, ix, iy, izp) * dx * dy;
#define FVISC(v)                                                                                                \
    ((a1 * (A(v, ix, iy, iz) - A(v, ixm, iy, iz)) + a2 * (A(v, ix, iy, iz) - A(v, ixp, iy, iz)) +              \
      a3v * (A(v, ix, iy, iz) - A(v, ix, iym, iz)) + a4 * (A(v, ix, iy, iz) - A(v, ix, iyp, iz)) +             \
      a5 * (A(v, ix, iy, iz) - A(v, ix, iy, izm)) + a6 * (A(v, ix, iy, iz) - A(v, ix, iy, izp))) /             \
     A(c->cv_v, ix, iy, iz))
        A(c->fx_visc, ix, iy, iz) = FVISC(vx);
        A(c->fy_visc, ix, iy, iz) = FVISC(vy);
        A(c->fz_visc, ix, iy, iz) = FVISC(vz);
#undef FVISC
    }
}

/* LARE3D::set_dt (lagran.cpp:380-443) with the OpenMP min-reduction of
 * applyReduction (OpenMPBackend.h:182-327): min is order independent. */
static void set_dt(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 rho = RHO(c), vx = VX(c), vy = VY(c), vz = VZ(c), bx = BX(c), by = BY(c), bz = BZ(c);
    double red = c->largest_number;
    /* lagran.cpp:384: the x range starts at 1 outside Cartesian geometry */
    const long i0 = c->geometry == 0 ? 0 : 1;
#pragma omp parallel for schedule(static) reduction(min : red)
    for (long ix = i0; ix <= nx; ++ix)
        for (long iy = 0; iy <= ny; ++iy)
            for (long iz = 0; iz <= nz; ++iz) {
                long izm = iz - 1, iym = iy - 1, ixm = ix - 1;
                double dx = L(c->dxb, ix);
                double dy = L(c->dyb, iy);
                double dz = L(c->dzb, iz);
                double dhx = dx, dhy = dy * L(c->hyc, ix), dhz = dz * H2(c->hzc, ix, iy);
                double rho0 = pmax(A(rho, ix, iy, iz), c->none_zero);
                double cs2 = c->gas_gamma * A(c->pressure, ix, iy, iz) / rho0;
                double w1 = (A(bx, ix, iy, iz) * A(bx, ix, iy, iz) + A(by, ix, iy, iz) * A(by, ix, iy, iz) +
                             A(bz, ix, iy, iz) * A(bz, ix, iy, iz)) / c->mu0 / rho0;
                double c_visc2 = A(c->p_visc, ix, iy, iz) / rho0;
                double length = pmin(pmin(dhx, dhy), dhz);
                double dt1 = length / (sqrt(c_visc2) + sqrt(cs2 + w1 + c_visc2));
                double ax = 0.25 * A(c->dxab, ix, iy, iz);
                double ay = 0.25 * A(c->dyab, ix, iy, iz);
                double az = 0.25 * A(c->dzab, ix, iy, iz);
                double vxbm = (A(vx, ixm, iy, iz) + A(vx, ixm, iym, iz) + A(vx, ixm, iy, izm) + A(vx, ixm, iym, izm)) * ax;
                double vxbp = (A(vx, ix, iy, iz) + A(vx, ix, iym, iz) + A(vx, ix, iy, izm) + A(vx, ix, iym, izm)) * ax;
                double vybm = (A(vy, ix, iym, iz) + A(vy, ixm, iym, iz) + A(vy, ix, iym, izm) + A(vy, ixm, iym, izm)) * ay;
                double vybp = (A(vy, ix, iy, iz) + A(vy, ixm, iy, iz) + A(vy, ix, iy, izm) + A(vy, ixm, iy, izm)) * ay;
                double vzbm = (A(vz, ix, iy, izm) + A(vz, ixm, iy, izm) + A(vz, ix, iym, izm) + A(vz, ixm, iym, izm)) * az;
                double vzbp = (A(vz, ix, iy, iz) + A(vz, ixm, iy, iz) + A(vz, ix, iym, iz) + A(vz, ixm, iym, iz)) * az;
                double dvx = fabs(vxbp - vxbm);
                double dvy = fabs(vybp - vybm);
                double dvz = fabs(vzbp - vzbm);
                double avxm = fabs(vxbm), avxp = fabs(vxbp);
                double avym = fabs(vybm), avyp = fabs(vybp);
                double avzm = fabs(vzbm), avzp = fabs(vzbp);
                double volume = A(c->cv, ix, iy, iz);
                double dt2 = volume / pmax(pmax(pmax(avxm, avxp), dvx), 1.0e-10 * volume);
                double dt3 = volume / pmax(pmax(pmax(avym, avyp), dvy), 1.0e-10 * volume);
                double dt4 = volume / pmax(pmax(pmax(avzm, avzp), dvz), 1.0e-10 * volume);
                double m = pmin(pmin(pmin(dt1, dt2), dt3), dt4);
                red = pmin(red, m);
            }
    c->dt = c->dt_multiplier * red;
    c->dtr = c->dt;
}

/* runner::calculateTimestep (runner.h:158-169) + LARE3D::calculateTimestep /
 * getTimestep (shared_data.h:428-440). */
static void calculate_timestep(lo_ctx *c)
{
    double tdt = DBL_MAX;
    set_dt(c);
    tdt = c->dt < tdt ? c->dt : tdt;
    c->dt = tdt;
}

/* eta_calc (lagran.cpp:445-476). */
static void eta_calc(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 bx = BX(c), by = BY(c), bz = BZ(c);
    FOR3(ix, -1, nx + 1, iy, -1, ny + 1, iz, -1, nz + 1)
    {
        long izp = iz + 1, iyp = iy + 1, ixp = ix + 1;
        double dxc = L(c->dxc, ix), dyc = L(c->dyc, iy), dzc = L(c->dzc, iz);
        double jx = (A(bz, ix, iyp, iz) - A(bz, ix, iy, iz)) / dyc - (A(by, ix, iy, izp) - A(by, ix, iy, iz)) / dzc;
        double jxp = (A(bz, ixp, iyp, iz) - A(bz, ixp, iy, iz)) / dyc - (A(by, ixp, iy, izp) - A(by, ixp, iy, iz)) / dzc;
        double jy = (A(bx, ix, iy, izp) - A(bx, ix, iy, iz)) / dzc - (A(bz, ixp, iy, iz) - A(bz, ix, iy, iz)) / dxc;
        double jyp = (A(bx, ix, iyp, izp) - A(bx, ix, iyp, iz)) / dzc - (A(bz, ixp, iyp, iz) - A(bz, ix, iyp, iz)) / dxc;
        double jz = (A(by, ixp, iy, iz) - A(by, ix, iy, iz)) / dxc - (A(bx, ix, iyp, iz) - A(bx, ix, iy, iz)) / dyc;
        double jzp = (A(by, ixp, iy, izp) - A(by, ix, iy, izp)) / dxc - (A(bx, ix, iyp, izp) - A(bx, ix, iy, izp)) / dyc;
        jx = (jx + jxp) * 0.5;
        jy = (jy + jyp) * 0.5;
        jz = (jz + jzp) * 0.5;
        double j_local = sqrt(jx * jx + jy * jy + jz * jz);
        A(c->eta, ix, iy, iz) = j_local > c->j_max ? c->eta_background + c->eta0 : c->eta_background;
    }
}

/* rkstep (lagran.cpp:507-535). */
static void rkstep(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 bx = BX(c), by = BY(c), bz = BZ(c);
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        long izp = iz + 1, iyp = iy + 1, ixp = ix + 1;
        double dxc = L(c->dxc, ix), dyc = L(c->dyc, iy), dzc = L(c->dzc, iz);
        double jx1 = (A(bz, ix, iyp, iz) - A(bz, ix, iy, iz)) / dyc - (A(by, ix, iy, izp) - A(by, ix, iy, iz)) / dzc;
        double jx2 = (A(bz, ixp, iyp, iz) - A(bz, ixp, iy, iz)) / dyc - (A(by, ixp, iy, izp) - A(by, ixp, iy, iz)) / dzc;
        double jy1 = (A(bx, ix, iy, izp) - A(bx, ix, iy, iz)) / dzc - (A(bz, ixp, iy, iz) - A(bz, ix, iy, iz)) / dxc;
        double jy2 = (A(bx, ix, iyp, izp) - A(bx, ix, iyp, iz)) / dzc - (A(bz, ixp, iyp, iz) - A(bz, ix, iyp, iz)) / dxc;
        double jz1 = (A(by, ixp, iy, iz) - A(by, ix, iy, iz)) / dxc - (A(bx, ix, iyp, iz) - A(bx, ix, iy, iz)) / dyc;
        double jz2 = (A(by, ixp, iy, izp) - A(by, ix, iy, izp)) / dxc - (A(bx, ix, iyp, izp) - A(bx, ix, iy, izp)) / dyc;
        double jx = 0.5 * (jx1 + jx2);
        double jy = 0.5 * (jy1 + jy2);
        double jz = 0.5 * (jz1 + jz2);
        double eta = A(c->eta, ix, iy, iz);
        A(c->flux_x, ix, iy, iz) = -jx * eta * dxc * 0.5;
        A(c->flux_y, ix, iy, iz) = -jy * eta * dyc * 0.5;
        A(c->flux_z, ix, iy, iz) = -jz * eta * dzc * 0.5;
        A(c->curlb, ix, iy, iz) = eta * (jx * jx + jy * jy + jz * jz);
    }
}

/* bstep (lagran.cpp:537-590). */
static void bstep(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 fx = c->flux_x, fy = c->flux_y, fz = c->flux_z;
    const double dt = c->dt;
    FOR3(ix, 0, nx, iy, 1, ny, iz, 1, nz)
    {
        long izm = iz - 1, iym = iy - 1;
        double area = L(c->dyb, iy) * L(c->dzb, iz);
        A(BX(c), ix, iy, iz) = A(c->bx1, ix, iy, iz) +
            (A(fz, ix, iy, iz) - A(fz, ix, iym, iz) + A(fz, ix, iy, izm) - A(fz, ix, iym, izm) -
             A(fy, ix, iy, iz) + A(fy, ix, iy, izm) - A(fy, ix, iym, iz) + A(fy, ix, iym, izm)) * dt / area;
    }
    FOR3(ix, 1, nx, iy, 0, ny, iz, 1, nz)
    {
        long izm = iz - 1, ixm = ix - 1;
        double area = L(c->dxb, ix) * L(c->dzb, iz);
        A(BY(c), ix, iy, iz) = A(c->by1, ix, iy, iz) +
            (A(fx, ix, iy, iz) - A(fx, ix, iy, izm) + A(fx, ixm, iy, iz) - A(fx, ixm, iy, izm) -
             A(fz, ix, iy, iz) + A(fz, ixm, iy, iz) - A(fz, ix, iy, izm) + A(fz, ixm, iy, izm)) * dt / area;
    }
    FOR3(ix, 1, nx, iy, 1, ny, iz, 0, nz)
    {
        long iym = iy - 1, ixm = ix - 1;
        double area = L(c->dxb, ix) * L(c->dyb, iy);
        A(BZ(c), ix, iy, iz) = A(c->bz1, ix, iy, iz) +
            (A(fy, ix, iy, iz) - A(fy, ixm, iy, iz) + A(fy, ix, iym, iz) - A(fy, ixm, iym, iz) -
             A(fx, ix, iy, iz) + A(fx, ix, iym, iz) - A(fx, ixm, iy, iz) + A(fx, ixm, iym, iz)) * dt / area;
    }
    bfield_bcs(c);
}

/* resistive_effects (lagran.cpp:478-505).  The reference's first statement,
 * pw::assign(data.bx1, data.bx), throws because the extents differ
 * (algorithm.h:76-78); this restatement copies bx into bx1 over bx1's own
 * extent (likewise by, bz), the smallest change that lets the path run. */
static void resistive_effects(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        A(c->bx1, ix, iy, iz) = A(BX(c), ix, iy, iz);
        A(c->by1, ix, iy, iz) = A(BY(c), ix, iy, iz);
        A(c->bz1, ix, iy, iz) = A(BZ(c), ix, iy, iz);
    }
    rkstep(c);
    bstep(c);
    FOR3(ix, 1, nx, iy, 1, ny, iz, 1, nz)
    {
        long izm = iz - 1, iym = iy - 1, ixm = ix - 1;
        const a3 cb = c->curlb;
        double sum_curlb = A(cb, ix, iy, iz) + A(cb, ixm, iy, iz) + A(cb, ix, iym, iz) + A(cb, ixm, iym, iz) +
                           A(cb, ix, iy, izm) + A(cb, ixm, iy, izm) + A(cb, ix, iym, izm) + A(cb, ixm, iym, izm);
        A(EE(c), ix, iy, iz) += sum_curlb * c->dt / (8.0 * A(RHO(c), ix, iy, iz));
    }
    energy_bcs(c);
    rkstep(c);
}

/* b_field_and_cv1_update (lagran.cpp:806-882). */
static void b_field_and_cv1_update(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 vx = VX(c), vy = VY(c), vz = VZ(c);
    const double dt = c->dt;
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        long izm = iz - 1, iym = iy - 1, ixm = ix - 1;
        double vxb = 0.25 * (A(vx, ix, iy, iz) + A(vx, ix, iym, iz) + A(vx, ix, iy, izm) + A(vx, ix, iym, izm));
        double vxbm = 0.25 * (A(vx, ixm, iy, iz) + A(vx, ixm, iym, iz) + A(vx, ixm, iy, izm) + A(vx, ixm, iym, izm));
        double vyb = 0.25 * (A(vy, ix, iy, iz) + A(vy, ixm, iy, iz) + A(vy, ix, iy, izm) + A(vy, ixm, iy, izm));
        double vybm = 0.25 * (A(vy, ix, iym, iz) + A(vy, ixm, iym, iz) + A(vy, ix, iym, izm) + A(vy, ixm, iym, izm));
        double vzb = 0.25 * (A(vz, ix, iy, iz) + A(vz, ixm, iy, iz) + A(vz, ix, iym, iz) + A(vz, ixm, iym, iz));
        double vzbm = 0.25 * (A(vz, ix, iy, izm) + A(vz, ixm, iy, izm) + A(vz, ix, iym, izm) + A(vz, ixm, iym, izm));
        double vol = A(c->cv, ix, iy, iz);
        double dvxdx = (vxb * A(c->dxab, ix, iy, iz) - vxbm * A(c->dxab, ixm, iy, iz)) / vol;
        double dvydy = (vyb * A(c->dyab, ix, iy, iz) - vybm * A(c->dyab, ix, iym, iz)) / vol;
        double dvzdz = (vzb * A(c->dzab, ix, iy, iz) - vzbm * A(c->dzab, ix, iy, izm)) / vol;
        double dv = (dvxdx + dvydy + dvzdz) * dt / 2.0;
        A(c->cv1, ix, iy, iz) = vol * (1.0 + dv);

        vxb = 0.25 * (A(vx, ix, iy, iz) + A(vx, ixm, iy, iz) + A(vx, ix, iy, izm) + A(vx, ixm, iy, izm));
        vxbm = 0.25 * (A(vx, ix, iym, iz) + A(vx, ixm, iym, iz) + A(vx, ix, iym, izm) + A(vx, ixm, iym, izm));
        vyb = 0.25 * (A(vy, ix, iy, iz) + A(vy, ix, iym, iz) + A(vy, ix, iy, izm) + A(vy, ix, iym, izm));
        vybm = 0.25 * (A(vy, ixm, iy, iz) + A(vy, ixm, iym, iz) + A(vy, ixm, iy, izm) + A(vy, ixm, iym, izm));
        double dvxdy = (vxb * A(c->dyab, ix, iy, iz) - vxbm * A(c->dyab, ix, iym, iz)) / vol;
        double dvydx = (vyb * A(c->dxab, ix, iy, iz) - vybm * A(c->dxab, ixm, iy, iz)) / vol;

        vxb = 0.25 * (A(vx, ix, iy, iz) + A(vx, ixm, iy, iz) + A(vx, ix, iym, iz) + A(vx, ixm, iym, iz));
        vxbm = 0.25 * (A(vx, ix, iy, izm) + A(vx, ixm, iy, izm) + A(vx, ix, iym, izm) + A(vx, ixm, iym, izm));
        vzb = 0.25 * (A(vz, ix, iy, iz) + A(vz, ix, iym, iz) + A(vz, ix, iy, izm) + A(vz, ix, iym, izm));
        vzbm = 0.25 * (A(vz, ixm, iy, iz) + A(vz, ixm, iym, iz) + A(vz, ixm, iy, izm) + A(vz, ixm, iym, izm));
        double dvxdz = (vxb * A(c->dzab, ix, iy, iz) - vxbm * A(c->dzab, ix, iy, izm)) / vol;
        double dvzdx = (vzb * A(c->dxab, ix, iy, iz) - vzbm * A(c->dxab, ixm, iy, iz)) / vol;

        vyb = 0.25 * (A(vy, ix, iy, iz) + A(vy, ixm, iy, iz) + A(vy, ix, iym, iz) + A(vy, ixm, iym, iz));
        vybm = 0.25 * (A(vy, ix, iy, izm) + A(vy, ixm, iy, izm) + A(vy, ix, iym, izm) + A(vy, ixm, iym, izm));
        vzb = 0.25 * (A(vz, ix, iy, iz) + A(vz, ixm, iy, iz) + A(vz, ix, iy, izm) + A(vz, ixm, iy, izm));
        vzbm = 0.25 * (A(vz, ix, iym, iz) + A(vz, ixm, iym, iz) + A(vz, ix, iym, izm) + A(vz, ixm, iym, izm));
        double dvydz = (vyb * A(c->dzab, ix, iy, iz) - vybm * A(c->dzab, ix, iy, izm)) / vol;
        double dvzdy = (vzb * A(c->dyab, ix, iy, iz) - vzbm * A(c->dyab, ix, iym, iz)) / vol;

        double bx1 = A(c->bx1, ix, iy, iz), by1 = A(c->by1, ix, iy, iz), bz1 = A(c->bz1, ix, iy, iz);
        double w3 = bx1 * dvxdx + by1 * dvxdy + bz1 * dvxdz;
        double w4 = bx1 * dvydx + by1 * dvydy + bz1 * dvydz;
        double w5 = bx1 * dvzdx + by1 * dvzdy + bz1 * dvzdz;
        A(c->bx1, ix, iy, iz) = (bx1 + w3 * dt / 2.0) / (1.0 + dv);
        A(c->by1, ix, iy, iz) = (by1 + w4 * dt / 2.0) / (1.0 + dv);
        A(c->bz1, ix, iy, iz) = (bz1 + w5 * dt / 2.0) / (1.0 + dv);
    }
}

/* predictor_step (lagran.cpp:592-737). */
static void predictor_step(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt, g = c->gas_gamma;
    b_field_and_cv1_update(c);
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        A(c->bx1, ix, iy, iz) *= A(c->cv1, ix, iy, iz);
        A(c->by1, ix, iy, iz) *= A(c->cv1, ix, iy, iz);
        A(c->bz1, ix, iy, iz) *= A(c->cv1, ix, iy, iz);
    }
    FOR3(ix, 0, nx + 1, iy, 0, ny + 1, iz, 0, nz + 1)
    {
        double rho = A(RHO(c), ix, iy, iz), cv = A(c->cv, ix, iy, iz), cv1 = A(c->cv1, ix, iy, iz);
        double dv = cv1 / cv - 1.0;
        double e1_e = A(EE(c), ix, iy, iz) - A(c->p_e, ix, iy, iz) * dv / rho;
        double e1_i = A(EI(c), ix, iy, iz) - A(c->p_i, ix, iy, iz) * dv / rho;
        e1_i += A(c->visc_heat, ix, iy, iz) * dt / 2.0 / rho;
        A(c->p_e, ix, iy, iz) = e1_e * (g - 1.0) * rho * cv / cv1;
        A(c->p_i, ix, iy, iz) = e1_i * (g - 1.0) * rho * cv / cv1;
        A(c->pressure, ix, iy, iz) = A(c->p_e, ix, iy, iz) + A(c->p_i, ix, iy, iz);
    }
    const a3 P = c->pressure, C1 = c->cv1, B1x = c->bx1, B1y = c->by1, B1z = c->bz1;
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        long izp = iz + 1, iyp = iy + 1, ixp = ix + 1;
        double dx = L(c->dxc, ix), dy = L(c->dyc, iy), dz = L(c->dzc, iz);
        double h2 = L(c->hy, ix), h2x = L(c->hy, ixp);
        double h3 = H2(c->hz, ix, iy), h3x = H2(c->hz, ixp, iy), h3y = H2(c->hz, ix, iyp);
        double h2c = L(c->hyc, ix), h3c = H2(c->hzc, ix, iy);
        double dhy = h2c * dy, dhz = h3c * dz;
        double pp = A(P, ix, iy, iz), ppx = A(P, ixp, iy, iz), ppy = A(P, ix, iyp, iz), ppz = A(P, ix, iy, izp);
        double ppxy = A(P, ixp, iyp, iz), ppxz = A(P, ixp, iy, izp), ppyz = A(P, ix, iyp, izp);
        double ppxyz = A(P, ixp, iyp, izp);
        double w1 = pp + ppy + ppz + ppyz;
        double w2 = ppx + ppxy + ppxz + ppxyz;
        double fx = (w1 - w2) * 0.25 / dx;
        w1 = pp + ppx + ppz + ppxz;
        w2 = ppy + ppxy + ppyz + ppxyz;
        double fy = (w1 - w2) * 0.25 / dhy;
        w1 = pp + ppx + ppy + ppxy;
        w2 = ppz + ppxz + ppyz + ppxyz;
        double fz = (w1 - w2) * 0.25 / dhz;

        double cvx = A(C1, ix, iy, iz) + A(C1, ix, iyp, iz) + A(C1, ix, iy, izp) + A(C1, ix, iyp, izp);
        double cvxp = A(C1, ixp, iy, iz) + A(C1, ixp, iyp, iz) + A(C1, ixp, iy, izp) + A(C1, ixp, iyp, izp);
        double cvy = A(C1, ix, iy, iz) + A(C1, ixp, iy, iz) + A(C1, ix, iy, izp) + A(C1, ixp, iy, izp);
        double cvyp = A(C1, ix, iyp, iz) + A(C1, ixp, iyp, iz) + A(C1, ix, iyp, izp) + A(C1, ixp, iyp, izp);
        double cvz = A(C1, ix, iy, iz) + A(C1, ixp, iy, iz) + A(C1, ix, iyp, iz) + A(C1, ixp, iyp, iz);
        double cvzp = A(C1, ix, iy, izp) + A(C1, ixp, iy, izp) + A(C1, ix, iyp, izp) + A(C1, ixp, iyp, izp);

        w1 = (A(B1z, ix, iy, iz) + A(B1z, ixp, iy, iz) + A(B1z, ix, iy, izp) + A(B1z, ixp, iy, izp)) * h3 / cvy;
        w2 = (A(B1z, ix, iyp, iz) + A(B1z, ixp, iyp, iz) + A(B1z, ix, iyp, izp) + A(B1z, ixp, iyp, izp)) * h3y / cvyp;
        double jx = (w2 - w1) / dhy / h3c;
        w1 = (A(B1y, ix, iy, iz) + A(B1y, ixp, iy, iz) + A(B1y, ix, iyp, iz) + A(B1y, ixp, iyp, iz)) / cvz;
        w2 = (A(B1y, ix, iy, izp) + A(B1y, ixp, iy, izp) + A(B1y, ix, iyp, izp) + A(B1y, ixp, iyp, izp)) / cvzp;
        jx -= (w2 - w1) / dhz;

        w1 = (A(B1z, ix, iy, iz) + A(B1z, ix, iyp, iz) + A(B1z, ix, iy, izp) + A(B1z, ix, iyp, izp)) * h3 / cvx;
        w2 = (A(B1z, ixp, iy, iz) + A(B1z, ixp, iyp, iz) + A(B1z, ixp, iy, izp) + A(B1z, ixp, iyp, izp)) * h3x / cvxp;
        double jy = -(w2 - w1) / dx / h3c;
        w1 = (A(B1x, ix, iy, iz) + A(B1x, ixp, iy, iz) + A(B1x, ix, iyp, iz) + A(B1x, ixp, iyp, iz)) / cvz;
        w2 = (A(B1x, ix, iy, izp) + A(B1x, ixp, iy, izp) + A(B1x, ix, iyp, izp) + A(B1x, ixp, iyp, izp)) / cvzp;
        jy += (w2 - w1) / dhz;

        w1 = (A(B1y, ix, iy, iz) + A(B1y, ix, iyp, iz) + A(B1y, ix, iy, izp) + A(B1y, ix, iyp, izp)) * h2 / cvx;
        w2 = (A(B1y, ixp, iy, iz) + A(B1y, ixp, iyp, iz) + A(B1y, ixp, iy, izp) + A(B1y, ixp, iyp, izp)) * h2x / cvxp;
        double jz = (w2 - w1) / dx / h2c;
        w1 = (A(B1x, ix, iy, iz) + A(B1x, ixp, iy, iz) + A(B1x, ix, iy, izp) + A(B1x, ixp, iy, izp)) / cvy;
        w2 = (A(B1x, ix, iyp, iz) + A(B1x, ixp, iyp, iz) + A(B1x, ix, iyp, izp) + A(B1x, ixp, iyp, izp)) / cvyp;
        jz -= (w2 - w1) / dhy;

#define BSUM8(B)                                                                                                \
    (A(B, ix, iy, iz) + A(B, ixp, iy, iz) + A(B, ix, iyp, iz) + A(B, ixp, iyp, iz) + A(B, ix, iy, izp) +        \
     A(B, ixp, iy, izp) + A(B, ix, iyp, izp) + A(B, ixp, iyp, izp))
        double bxv = BSUM8(B1x) / (cvx + cvxp);
        double byv = BSUM8(B1y) / (cvx + cvxp);
        double bzv = BSUM8(B1z) / (cvx + cvxp);
#undef BSUM8
        fx += (jy * bzv - jz * byv) / c->mu0;
        fy += (jz * bxv - jx * bzv) / c->mu0;
        fz += (jx * byv - jy * bxv) / c->mu0;
        /* geometry corrections (lagran.cpp:711-721) */
        if (c->geometry == 2) {
            double cotantheta = 1.0 / tan(L(c->yb, iy));
            double rv = A(c->rho_v, ix, iy, iz), ux = A(VX(c), ix, iy, iz), uy = A(VY(c), ix, iy, iz),
                   uz = A(VZ(c), ix, iy, iz), r = L(c->xb, ix);
            fx += rv * (uy * uy + uz * uz) / r;
            fy -= rv * (uy * ux - cotantheta * uz * uz) / r;
            fz -= rv * (uz * ux + cotantheta * uz * uy) / r;
        } else if (c->geometry == 1) {
            double rv = A(c->rho_v, ix, iy, iz), ux = A(VX(c), ix, iy, iz), uy = A(VY(c), ix, iy, iz),
                   r = L(c->xb, ix);
            fx += rv * uy * uy / r;
            fy -= rv * uy * ux / r;
        }
        A(c->fx, ix, iy, iz) = fx;
        A(c->fy, ix, iy, iz) = fy;
        A(c->fz, ix, iy, iz) = fz;
        double rv = A(c->rho_v, ix, iy, iz);
        A(VX1(c), ix, iy, iz) = A(VX(c), ix, iy, iz) + dt / 2.0 * (A(c->fx_visc, ix, iy, iz) + fx) / rv;
        A(VY1(c), ix, iy, iz) = A(VY(c), ix, iy, iz) + dt / 2.0 * (A(c->fy_visc, ix, iy, iz) + fy) / rv;
        A(VZ1(c), ix, iy, iz) = A(VZ(c), ix, iy, iz) + dt / 2.0 * (A(c->fz_visc, ix, iy, iz) + fz) / rv;
    }
    remap_v_bcs(c);
}

/* lagrangian_step prologue (lagran.cpp:113-153): bx1.., p_e, p_i, pressure
 * over the centred extent, then vertex rho_v and cv_v. */
static void lagrangian_prologue(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double g = c->gas_gamma;
    const a3 rho = RHO(c), cv = c->cv;
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        long izm = iz - 1, iym = iy - 1, ixm = ix - 1;
        A(c->bx1, ix, iy, iz) = 0.5 * (A(BX(c), ix, iy, iz) + A(BX(c), ixm, iy, iz));
        A(c->by1, ix, iy, iz) = 0.5 * (A(BY(c), ix, iy, iz) + A(BY(c), ix, iym, iz));
        A(c->bz1, ix, iy, iz) = 0.5 * (A(BZ(c), ix, iy, iz) + A(BZ(c), ix, iy, izm));
        A(c->p_e, ix, iy, iz) = (g - 1.0) * A(rho, ix, iy, iz) * A(EE(c), ix, iy, iz);
        A(c->p_i, ix, iy, iz) = (g - 1.0) * A(rho, ix, iy, iz) * A(EI(c), ix, iy, iz);
        A(c->pressure, ix, iy, iz) = A(c->p_e, ix, iy, iz) + A(c->p_i, ix, iy, iz);
    }
    FOR3(ix, -1, nx + 1, iy, -1, ny + 1, iz, -1, nz + 1)
    {
        long izp = iz + 1, iyp = iy + 1, ixp = ix + 1;
        double sum_rho_cv = A(rho, ix, iy, iz) * A(cv, ix, iy, iz) + A(rho, ixp, iy, iz) * A(cv, ixp, iy, iz) +
                            A(rho, ix, iyp, iz) * A(cv, ix, iyp, iz) + A(rho, ixp, iyp, iz) * A(cv, ixp, iyp, iz) +
                            A(rho, ix, iy, izp) * A(cv, ix, iy, izp) + A(rho, ixp, iy, izp) * A(cv, ixp, iy, izp) +
                            A(rho, ix, iyp, izp) * A(cv, ix, iyp, izp) +
                            A(rho, ixp, iyp, izp) * A(cv, ixp, iyp, izp);
        double sum_cv = A(cv, ix, iy, iz) + A(cv, ixp, iy, iz) + A(cv, ix, iyp, iz) + A(cv, ixp, iyp, iz) +
                        A(cv, ix, iy, izp) + A(cv, ixp, iy, izp) + A(cv, ix, iyp, izp) + A(cv, ixp, iyp, izp);
        A(c->rho_v, ix, iy, iz) = sum_rho_cv / sum_cv;
        A(c->cv_v, ix, iy, iz) = 0.125 * sum_cv;
    }
}

/* LARE3D::lagrangian_step (lagran.cpp:95-173). */
double lo_start_of_timestep(lo_ctx *c)
{
    lagrangian_prologue(c);
    shock_viscosity(c);
    calculate_timestep(c);
    if (c->resistive) {
        double dt_sub = c->dtr;
        int substeps = (int)(c->dt / dt_sub) + 1;
        double actual_dt = c->dt;
        c->dt /= (double)substeps;
        for (int i = 0; i < substeps; ++i) {
            eta_calc(c);
            resistive_effects(c);
        }
        c->dt = actual_dt;
    }
    predictor_step(c);
    return c->dt;
}

double lo_set_dt_only(lo_ctx *c)
{
    set_dt(c);
    return c->dt;
}

/* shock_heating (lagran.cpp:884-993), including the dy*dy factors of the
 * alpha3 terms a9 and a11 (lagran.cpp:987). */
static void shock_heating(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 vx = VX(c), vy = VY(c), vz = VZ(c), wx = VX1(c), wy = VY1(c), wz = VZ1(c);
    FOR3(ix, 0, nx + 1, iy, 0, ny + 1, iz, 0, nz + 1)
    {
        long izm = iz - 1, izp = iz + 1, iym = iy - 1, iyp = iy + 1, ixm = ix - 1;
#define DOT(i, j, k, I, J, K)                                                                                   \
    (1 * ((A(vx, i, j, k) - A(vx, I, J, K)) * (A(wx, i, j, k) - A(wx, I, J, K)) +                               \
          (A(vy, i, j, k) - A(vy, I, J, K)) * (A(wy, i, j, k) - A(wy, I, J, K)) +                               \
          (A(vz, i, j, k) - A(vz, I, J, K)) * (A(wz, i, j, k) - A(wz, I, J, K))))
        double a1 = DOT(ixm, iym, iz, ix, iym, iz);
        double a2 = DOT(ix, iym, iz, ix, iy, iz);
        double a3v = DOT(ix, iy, iz, ixm, iy, iz);
        double a4 = DOT(ixm, iy, iz, ixm, iym, iz);
        double a5 = DOT(ixm, iym, izp, ix, iym, izp);
        double a6 = DOT(ix, iym, izp, ix, iy, izp);
        double a7 = DOT(ix, iy, izp, ixm, iy, izp);
        double a8 = DOT(ixm, iy, izp, ixm, iym, izp);
        double a9 = DOT(ix, iy, izm, ix, iy, iz);
        double a10 = DOT(ixm, iy, izm, ixm, iy, iz);
        double a11 = DOT(ixm, iym, izm, ixm, iym, iz);
        double a12 = DOT(ix, iym, izm, ix, iym, iz);
#undef DOT
        double dx = L(c->dxb, ix);
        double dy = L(c->dyb, iy) * L(c->hyc, ix);
        double dz = L(c->dzb, iz) * H2(c->hz2, ix, iy);
        A(c->visc_heat, ix, iy, iz) =
            (-0.25 * dy * dz * A(c->alpha1, ix, iy, iz) * a1 - 0.25 * dx * dz * A(c->alpha2, ix, iy, iz) * a2 -
             0.25 * dy * dz * A(c->alpha1, ix, iyp, iz) * a3v - 0.25 * dx * dz * A(c->alpha2, ixm, iy, iz) * a4 -
             0.25 * dy * dz * A(c->alpha1, ix, iy, izp) * a5 - 0.25 * dx * dz * A(c->alpha2, ix, iy, izp) * a6 -
             0.25 * dy * dz * A(c->alpha1, ix, iyp, izp) * a7 - 0.25 * dx * dz * A(c->alpha2, ixm, iy, izp) * a8 -
             0.25 * dy * dy * A(c->alpha3, ix, iy, iz) * a9 - 0.25 * dx * dy * A(c->alpha3, ixm, iy, iz) * a10 -
             0.25 * dy * dy * A(c->alpha3, ixm, iym, iz) * a11 - 0.25 * dx * dy * A(c->alpha3, ix, iym, iz) * a12);
        A(c->visc_heat, ix, iy, iz) = pmax(A(c->visc_heat, ix, iy, iz) / A(c->cv, ix, iy, iz), 0.0);
    }
}

/* LARE3D::corrector_step (lagran.cpp:739-804). */
void lo_half_timestep(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt;
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        A(c->bx1, ix, iy, iz) /= A(c->cv1, ix, iy, iz);
        A(c->by1, ix, iy, iz) /= A(c->cv1, ix, iy, iz);
        A(c->bz1, ix, iy, iz) /= A(c->cv1, ix, iy, iz);
    }
    shock_heating(c);
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        double rv = A(c->rho_v, ix, iy, iz);
        A(VX(c), ix, iy, iz) += dt * (A(c->fx_visc, ix, iy, iz) + A(c->fx, ix, iy, iz)) / rv;
        A(VY(c), ix, iy, iz) += dt * (A(c->fy_visc, ix, iy, iz) + A(c->fy, ix, iy, iz)) / rv;
        A(VZ(c), ix, iy, iz) += dt * (A(c->fz_visc, ix, iy, iz) + A(c->fz, ix, iy, iz)) / rv;
    }
    velocity_bcs(c);
    const a3 wx = VX1(c), wy = VY1(c), wz = VZ1(c);
    FOR3(ix, 1, nx, iy, 1, ny, iz, 1, nz)
    {
        long izm = iz - 1, iym = iy - 1, ixm = ix - 1;
        double vxb = 0.25 * (A(wx, ix, iy, iz) + A(wx, ix, iym, iz) + A(wx, ix, iy, izm) + A(wx, ix, iym, izm));
        double vxbm = 0.25 * (A(wx, ixm, iy, iz) + A(wx, ixm, iym, iz) + A(wx, ixm, iy, izm) + A(wx, ixm, iym, izm));
        double vyb = 0.25 * (A(wy, ix, iy, iz) + A(wy, ixm, iy, iz) + A(wy, ix, iy, izm) + A(wy, ixm, iy, izm));
        double vybm = 0.25 * (A(wy, ix, iym, iz) + A(wy, ixm, iym, iz) + A(wy, ix, iym, izm) + A(wy, ixm, iym, izm));
        double vzb = 0.25 * (A(wz, ix, iy, iz) + A(wz, ixm, iy, iz) + A(wz, ix, iym, iz) + A(wz, ixm, iym, iz));
        double vzbm = 0.25 * (A(wz, ix, iy, izm) + A(wz, ixm, iy, izm) + A(wz, ix, iym, izm) + A(wz, ixm, iym, izm));
        double vol = A(c->cv, ix, iy, iz);
        double dvxdx = (vxb * A(c->dxab, ix, iy, iz) - vxbm * A(c->dxab, ixm, iy, iz)) / vol;
        double dvydy = (vyb * A(c->dyab, ix, iy, iz) - vybm * A(c->dyab, ix, iym, iz)) / vol;
        double dvzdz = (vzb * A(c->dzab, ix, iy, iz) - vzbm * A(c->dzab, ix, iy, izm)) / vol;
        double dv = (dvxdx + dvydy + dvzdz) * dt;
        A(c->cv1, ix, iy, iz) = vol * (1.0 + dv);
        double rho = A(RHO(c), ix, iy, iz);
        A(EE(c), ix, iy, iz) -= dv * A(c->p_e, ix, iy, iz) / rho;
        A(EI(c), ix, iy, iz) += (dt * A(c->visc_heat, ix, iy, iz) - dv * A(c->p_i, ix, iy, iz)) / rho;
        A(RHO(c), ix, iy, iz) /= (1.0 + dv);
    }
    energy_bcs(c);
    density_bcs(c);
    velocity_bcs(c);
}

/* ------------------------------------------------------------ remap core */

/* The limited-slope core shared by all fifteen remap flux kernels
 * (xremap.cpp:304-316 and its twins): Da, ss and the minmod-limited Di. */
static inline double vl_di(double sign_v, double dfi, double dfu, double phi, double dxci, double dxcu, double dxbu)
{
    double Da = (2.0 - phi) * fabs(dfi) / dxci + (1.0 + phi) * fabs(dfu) / dxcu;
    Da = Da * sixth;
    double ss = 0.5 * ((dfi >= 0.0 ? 1.0 : -1.0) + (dfu >= 0.0 ? 1.0 : -1.0));
    return sign_v * ss * pmin_list3(fabs(Da) * dxbu, fabs(dfi), fabs(dfu));
}

/* Cell-volume preparation common to the three sweeps (xremap.cpp:26-70,
 * yremap.cpp:25-76, zremap.cpp:27-79): dir selects which divergence is
 * remapped (cv1 before, cv2 after) and which db1 is built. */
static void remap_prep(lo_ctx *c, int dir)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const a3 wx = VX1(c), wy = VY1(c), wz = VZ1(c);
    const double dt = c->dt, xpass = c->xpass, ypass = c->ypass, zpass = c->zpass;
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2)
    {
        long ixm = ix - 1, iym = iy - 1, izm = iz - 1;
        double vxb = (A(wx, ix, iy, iz) + A(wx, ix, iym, iz) + A(wx, ix, iy, izm) + A(wx, ix, iym, izm)) * 0.25;
        double vxbm = (A(wx, ixm, iy, iz) + A(wx, ixm, iym, iz) + A(wx, ixm, iy, izm) + A(wx, ixm, iym, izm)) * 0.25;
        double vyb = (A(wy, ix, iy, iz) + A(wy, ixm, iy, iz) + A(wy, ix, iy, izm) + A(wy, ixm, iy, izm)) * 0.25;
        double vybm = (A(wy, ix, iym, iz) + A(wy, ixm, iym, iz) + A(wy, ix, iym, izm) + A(wy, ixm, iym, izm)) * 0.25;
        double vzb = (A(wz, ix, iy, iz) + A(wz, ixm, iy, iz) + A(wz, ix, iym, iz) + A(wz, ixm, iym, iz)) * 0.25;
        double vzbm = (A(wz, ix, iy, izm) + A(wz, ixm, iy, izm) + A(wz, ix, iym, izm) + A(wz, ixm, iym, izm)) * 0.25;
        double vol = A(c->cv, ix, iy, iz);
        double dv;
        if (dir == 0) {
            double dvxdx = (vxb * A(c->dxab, ix, iy, iz) - vxbm * A(c->dxab, ixm, iy, iz)) / vol;
            double dvydy = ypass * (vyb * A(c->dyab, ix, iy, iz) - vybm * A(c->dyab, ix, iym, iz)) / vol;
            double dvzdz = zpass * (vzb * A(c->dzab, ix, iy, iz) - vzbm * A(c->dzab, ix, iy, izm)) / vol;
            dv = (dvydy + dvzdz) * dt;
            A(c->cv2, ix, iy, iz) = vol * (1.0 + dv);
            dv = dv + dvxdx * dt;
            A(c->cv1, ix, iy, iz) = vol * (1.0 + dv);
            A(c->db1, ix, iy, iz) = L(c->dxb, ix) + (vxb - vxbm) * dt;
        } else if (dir == 1) {
            double dvxdx = xpass * (vxb * A(c->dxab, ix, iy, iz) - vxbm * A(c->dxab, ixm, iy, iz)) / vol;
            double dvydy = (vyb * A(c->dyab, ix, iy, iz) - vybm * A(c->dyab, ix, iym, iz)) / vol;
            double dvzdz = zpass * (vzb * A(c->dzab, ix, iy, iz) - vzbm * A(c->dzab, ix, iy, izm)) / vol;
            dv = (dvxdx + dvzdz) * dt;
            A(c->cv2, ix, iy, iz) = vol * (1.0 + dv);
            dv = dv + dvydy * dt;
            A(c->cv1, ix, iy, iz) = vol * (1.0 + dv);
            A(c->db1, ix, iy, iz) = L(c->hyc, ix) * L(c->dyb, iy) + (vyb - vybm) * dt;
        } else {
            double dvxdx = xpass * (vxb * A(c->dxab, ix, iy, iz) - vxbm * A(c->dxab, ixm, iy, iz)) / vol;
            double dvydy = ypass * (vyb * A(c->dyab, ix, iy, iz) - vybm * A(c->dyab, ix, iym, iz)) / vol;
            double dvzdz = (vzb * A(c->dzab, ix, iy, iz) - vzbm * A(c->dzab, ix, iy, izm)) / vol;
            dv = (dvxdx + dvydy) * dt;
            A(c->cv2, ix, iy, iz) = vol * (1.0 + dv);
            dv = dv + dvzdz * dt;
            A(c->cv1, ix, iy, iz) = vol * (1.0 + dv);
            A(c->db1, ix, iy, iz) = H2(c->hzc, ix, iy) * L(c->dzb, iz) + (vzb - vzbm) * dt;
        }
    }
    /* cvc1: vertex control volume before remap */
    const a3 C1 = c->cv1;
    FOR3(ix, -1, nx + 1, iy, -1, ny + 1, iz, -1, nz + 1)
    {
        long ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        A(c->cvc1, ix, iy, iz) =
            0.125 * (A(C1, ix, iy, iz) + A(C1, ixp, iy, iz) + A(C1, ix, iyp, iz) + A(C1, ixp, iyp, iz) +
                     A(C1, ix, iy, izp) + A(C1, ixp, iy, izp) + A(C1, ix, iyp, izp) + A(C1, ixp, iyp, izp));
    }
}

/* Van Leer flux of a face-centred magnetic flux B along sweep direction d.
 * v_advect is the mean of w at (i,j,k) and at the offset (oi,oj,ok); db1 is
 * averaged with the transverse offset (ti,tj,tk).  Restates vx_by_flux
 * (xremap.cpp:268-319), vx_bz_flux (322-371), vy_bx_flux (yremap.cpp:291-341),
 * vy_bz_flux (344-393), vz_bx_flux (zremap.cpp:250-297), vz_by_flux (300-346). */
static void b_flux(lo_ctx *c, int d, a3 B, a3 w, int oi, int oj, int ok, int ti, int tj, int tk, a3 flux)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const a3 db1 = c->db1, cvc1 = c->cvc1, cv1 = c->cv1;
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
#define DB(s) ((AP(db1, s) + A(db1, ix + (s) * di + ti, iy + (s) * dj + tj, iz + (s) * dk + tk)) * 0.5)
        double v_advect = (A(w, ix, iy, iz) + A(w, ix + oi, iy + oj, iz + ok)) * 0.5;
        double o_v = v_advect * dt;
        double dbx = DB(0), dbxp = DB(1), dbxp2 = DB(2), dbxm = DB(-1);
        double fm = AP(B, -1) / dbxm;
        double fi = AP(B, 0) / dbx;
        double fp = AP(B, 1) / dbxp;
        double fp2 = AP(B, 2) / dbxp2;
        double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
        double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
        double vad_p = (sign_v + 1.0) * 0.5;
        double vad_m = 1.0 - vad_p;
        double fu = fi * vad_p + fp * vad_m;
        double dfu = dfm * vad_p + dfp * vad_m;
        double dxci = A(cvc1, ix, iy, iz);
        double dxcu = AP(cvc1, -1) * vad_p + AP(cvc1, 1) * vad_m;
        double dxbu = A(cv1, ix, iy, iz) * vad_p + AP(cv1, 1) * vad_m;
        double dxu = dbx * vad_p + dbxp * vad_m;
        double phi = fabs(o_v) / dxu;
        double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
        A(flux, ix, iy, iz) = (fu + Di * (1.0 - phi)) * o_v;
#undef DB
    }
}

/* x_mass_flux (xremap.cpp:373-422), y_mass_flux (yremap.cpp:395-444),
 * z_mass_flux (zremap.cpp:349-397): dm on the faces of direction d. */
static void mass_flux(lo_ctx *c, int d)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const a3 rho = RHO(c), cvc1 = c->cvc1, cv1 = c->cv1, dm = DM(c);
    const a3 wx = VX1(c), wy = VY1(c), wz = VZ1(c);
    const long x1 = d == 0 ? nx : nx + 1, y1 = d == 1 ? ny : ny + 1, z1 = d == 2 ? nz : nz + 1;
    FOR3(ix, 0, x1, iy, 0, y1, iz, 0, z1)
    {
        long ixm = ix - 1, iym = iy - 1, izm = iz - 1;
        double area, v_advect;
        if (d == 0) {
            area = A(c->dxab, ix, iy, iz);
            v_advect = (A(wx, ix, iy, iz) + A(wx, ix, iym, iz) + A(wx, ix, iy, izm) + A(wx, ix, iym, izm)) * 0.25;
        } else if (d == 1) {
            area = A(c->dyab, ix, iy, iz);
            v_advect = (A(wy, ixm, iy, iz) + A(wy, ix, iy, iz) + A(wy, ixm, iy, izm) + A(wy, ix, iy, izm)) * 0.25;
        } else {
            area = A(c->dzab, ix, iy, iz);
            v_advect = (A(wz, ixm, iy, iz) + A(wz, ix, iy, iz) + A(wz, ixm, iym, iz) + A(wz, ix, iym, iz)) * 0.25;
        }
        double o_v = v_advect * dt * area;
        double fm = AP(rho, -1), fi = AP(rho, 0), fp = AP(rho, 1), fp2 = AP(rho, 2);
        double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
        double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
        double vad_p = (sign_v + 1.0) * 0.5;
        double vad_m = 1.0 - vad_p;
        double fu = fi * vad_p + fp * vad_m;
        double dfu = dfm * vad_p + dfp * vad_m;
        double dxci = A(cvc1, ix, iy, iz);
        double dxcu = AP(cvc1, -1) * vad_p + AP(cvc1, 1) * vad_m;
        double dxbu = A(cv1, ix, iy, iz) * vad_p + AP(cv1, 1) * vad_m;
        double phi = fabs(o_v) / dxbu;
        double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
        A(dm, ix, iy, iz) = (fu + Di * (1.0 - phi)) * o_v;
    }
}

/* x_energy_flux (xremap.cpp:429-483), y_energy_flux (yremap.cpp:452-506),
 * z_energy_flux (zremap.cpp:405-459) for one specific energy e. */
static void energy_flux(lo_ctx *c, int d, a3 e, a3 flux)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const a3 cvc1 = c->cvc1, cv1 = c->cv1, dm = DM(c), rho1 = c->rho1;
    const a3 wx = VX1(c), wy = VY1(c), wz = VZ1(c);
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        long ixm = ix - 1, iym = iy - 1, izm = iz - 1;
        double area, v_advect;
        if (d == 0) {
            area = A(c->dxab, ix, iy, iz);
            v_advect = (A(wx, ix, iy, iz) + A(wx, ix, iym, iz) + A(wx, ix, iy, izm) + A(wx, ix, iym, izm)) * 0.25;
        } else if (d == 1) {
            area = A(c->dyab, ix, iy, iz);
            v_advect = (A(wy, ixm, iy, iz) + A(wy, ix, iy, iz) + A(wy, ixm, iy, izm) + A(wy, ix, iy, izm)) * 0.25;
        } else {
            area = A(c->dzab, ix, iy, iz);
            v_advect = (A(wz, ixm, iy, iz) + A(wz, ix, iy, iz) + A(wz, ixm, iym, iz) + A(wz, ix, iym, iz)) * 0.25;
        }
        double o_v = v_advect * dt * area;
        double fm = AP(e, -1), fi = AP(e, 0), fp = AP(e, 1), fp2 = AP(e, 2);
        double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
        double sign_v = (v_advect >= 0.0) ? 1.0 : -1.0;
        double vad_p = (sign_v + 1.0) * 0.5;
        double vad_m = 1.0 - vad_p;
        double fu = fi * vad_p + fp * vad_m;
        double dfu = dfm * vad_p + dfp * vad_m;
        double dxci = A(cvc1, ix, iy, iz);
        double dxcu = AP(cvc1, -1) * vad_p + AP(cvc1, 1) * vad_m;
        double dxbu = A(cv1, ix, iy, iz) * vad_p + AP(cv1, 1) * vad_m;
        double phi = fabs(o_v) / dxbu;
        double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
        double rhou = A(rho1, ix, iy, iz) * vad_p + AP(rho1, 1) * vad_m;
        double dmu = fabs(A(dm, ix, iy, iz)) / dxbu / rhou;
        A(flux, ix, iy, iz) = (fu + Di * (1.0 - dmu)) * A(dm, ix, iy, iz);
    }
}

/* The remap kinetic-energy correction's accumulation for one velocity
 * component f, between a momentum-flux kernel and its multiplication by dm
 * (xremap.cpp:534-558, yremap.cpp:558-584, zremap.cpp:514-537): dk of the
 * vertex face (ix, iy, iz) is added to four delta_ke cells.  The loops are
 * sequential, in the order one thread runs the reference's loop (ix, iy, iz
 * ascending), so each cell's sum has that order; the reference's atomic
 * adds make the order of a multi-threaded run arbitrary (y and z: two
 * ix iterations add to one cell), and that of one thread is this one. */
static void rke_accumulate(lo_ctx *c, int d, a3 f, a3 flux)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const a3 rho_v1 = c->rho_v1, cv2 = c->cv2, dm = DM(c), ke = c->delta_ke;
    const long x1 = d == 0 ? nx - 1 : nx, y1 = d == 1 ? ny - 1 : ny, z1 = d == 2 ? nz - 1 : nz;
    for (long ix = 0; ix <= x1; ++ix)
        for (long iy = 0; iy <= y1; ++iy)
            for (long iz = 0; iz <= z1; ++iz) {
                const long ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
                double m = A(rho_v1, ix, iy, iz) * A(cv2, ix, iy, iz);
                double mp = AP(rho_v1, 1) * AP(cv2, 1);
                double f0 = A(f, ix, iy, iz), fp = AP(f, 1);
                double flm = AP(flux, -1), fl0 = A(flux, ix, iy, iz), flp = AP(flux, 1);
                double dmm = AP(dm, -1), dm0 = A(dm, ix, iy, iz), dmp = AP(dm, 1);
                double ai = (f0 - flm) * dmm / m - (f0 - fl0) * dm0 / m;
                double aip = (fp - fl0) * dm0 / mp - (fp - flp) * dmp / mp;
                double e = (fp - f0) * (fl0 - 0.5 * (fp + f0)) - 0.5 * ai * (f0 - fl0) + 0.5 * aip * (fp - fl0);
                e = e * dm0 * 0.5;
                if (d == 0) {
                    A(ke, ixp, iy, iz) += e;
                    A(ke, ixp, iyp, iz) += e;
                    A(ke, ixp, iyp, izp) += e;
                    A(ke, ixp, iy, izp) += e;
                } else if (d == 1) {
                    A(ke, ixp, iyp, iz) += e;
                    A(ke, ix, iyp, iz) += e;
                    A(ke, ixp, iyp, izp) += e;
                    A(ke, ix, iyp, izp) += e;
                } else {
                    A(ke, ix, iy, izp) += e;
                    A(ke, ixp, iy, izp) += e;
                    A(ke, ix, iyp, izp) += e;
                    A(ke, ixp, iyp, izp) += e;
                }
            }
}

/* LARE3D::energy_correction (diagnostics.cpp:125-135). */
static void energy_correction(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    FOR3(ix, 1, nx, iy, 1, ny, iz, 1, nz)
    {
        double dke = pmax(-A(c->delta_ke, ix, iy, iz), 0.0) / (A(RHO(c), ix, iy, iz) * A(c->cv, ix, iy, iz));
        A(EE(c), ix, iy, iz) += 0.5 * dke;
        A(EI(c), ix, iy, iz) += 0.5 * dke;
    }
}

/* x_mom_flux (xremap.cpp:485-566), y_mom_flux (yremap.cpp:508-588),
 * z_mom_flux (zremap.cpp:461-541) for one velocity component f; the flux is
 * multiplied by dm afterwards as in the reference's second kernel.
 * Reproduces: z uses dyac as its area (zremap.cpp:475), and z's dmu divides
 * by the product (dzbu*rhou) (zremap.cpp:505). */
static void mom_flux(lo_ctx *c, int d, a3 f, a3 flux)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const double dt = c->dt;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const a3 cvc1 = c->cvc1, cv1 = c->cv1, dm = DM(c), rho_v = c->rrho_v;
    const a3 w = d == 0 ? VX1(c) : d == 1 ? VY1(c) : VZ1(c);
    const a3 area_a = d == 0 ? c->dxac : c->dyac;
    const long x0 = d == 0 ? -1 : 0, y0 = d == 1 ? -1 : 0, z0 = d == 2 ? -1 : 0;
    FOR3(ix, x0, nx, iy, y0, ny, iz, z0, nz)
    {
        double area = A(area_a, ix, iy, iz);
        double v_advect = A(w, ix, iy, iz);
        double o_v = v_advect * dt * area;
        double fm = AP(f, -1), fi = AP(f, 0), fp = AP(f, 1), fp2 = AP(f, 2);
        double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
        double sign_v = (o_v >= 0.0) ? 1.0 : -1.0;
        double vad_p = (sign_v + 1.0) * 0.5;
        double vad_m = 1.0 - vad_p;
        double fu = fi * vad_p + fp * vad_m;
        double dfu = dfm * vad_p + dfp * vad_m;
        double dxci = AP(cv1, 1);
        double dxcu = A(cv1, ix, iy, iz) * vad_p + AP(cv1, 2) * vad_m;
        double dxbu = A(cvc1, ix, iy, iz) * vad_p + AP(cvc1, 1) * vad_m;
        double phi = fabs(o_v) / dxbu;
        double Di = vl_di(sign_v, dfi, dfu, phi, dxci, dxcu, dxbu);
        double rhou = A(rho_v, ix, iy, iz) * vad_p + AP(rho_v, 1) * vad_m;
        double dmu = d == 2 ? fabs(A(dm, ix, iy, iz)) / (dxbu * rhou) : fabs(A(dm, ix, iy, iz)) / dxbu / rhou;
        A(flux, ix, iy, iz) = fu + Di * (1.0 - dmu);
    }
    if (c->rke) rke_accumulate(c, d, f, flux);
    FOR3(ix, x0, nx, iy, y0, ny, iz, z0, nz)
    {
        A(flux, ix, iy, iz) *= A(dm, ix, iy, iz);
    }
}

/* Density and energy update after mass_flux/energy_flux; the vertex-cell
 * redefinitions; and the momentum updates.  Restates the bodies of remap_x
 * (xremap.cpp:118-262), remap_y (yremap.cpp:131-285), remap_z
 * (zremap.cpp:121-241) for direction d. */
static void remap_sweep(lo_ctx *c, int d)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    const int di = d == 0, dj = d == 1, dk = d == 2;
    a3 flux = d == 0 ? c->rflux_x : d == 1 ? c->rflux_y : c->rflux_z;
    a3 rho = RHO(c), dm = DM(c), cv1 = c->cv1, cv2 = c->cv2, rho1 = c->rho1, cvc1 = c->cvc1;
    a3 rho_v = c->rrho_v, rho_v1 = c->rho_v1;
    a3_fill(dm, 0.0);
    a3_copy(rho1, rho);
    remap_prep(c, d);

    /* constrained-transport B fluxes */
    if (d == 0) {
        b_flux(c, 0, BY(c), VX1(c), 0, 0, -1, 0, 1, 0, flux); /* vx_by_flux */
        FOR3(ix, 1, nx, iy, 0, ny, iz, 1, nz) { A(BY(c), ix, iy, iz) = A(BY(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix - 1, iy, iz); }
        FOR3(ix, 0, nx, iy, 1, ny, iz, 1, nz) { A(BX(c), ix, iy, iz) = A(BX(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix, iy - 1, iz); }
        b_flux(c, 0, BZ(c), VX1(c), 0, -1, 0, 0, 0, 1, flux); /* vx_bz_flux */
        FOR3(ix, 1, nx, iy, 1, ny, iz, 0, nz) { A(BZ(c), ix, iy, iz) = A(BZ(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix - 1, iy, iz); }
        FOR3(ix, 0, nx, iy, 1, ny, iz, 1, nz) { A(BX(c), ix, iy, iz) = A(BX(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix, iy, iz - 1); }
    } else if (d == 1) {
        b_flux(c, 1, BX(c), VY1(c), 0, 0, -1, 1, 0, 0, flux); /* vy_bx_flux */
        FOR3(ix, 0, nx, iy, 1, ny, iz, 1, nz) { A(BX(c), ix, iy, iz) = A(BX(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix, iy - 1, iz); }
        FOR3(ix, 1, nx, iy, 0, ny, iz, 1, nz) { A(BY(c), ix, iy, iz) = A(BY(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix - 1, iy, iz); }
        b_flux(c, 1, BZ(c), VY1(c), -1, 0, 0, 0, 0, 1, flux); /* vy_bz_flux */
        FOR3(ix, 1, nx, iy, 1, ny, iz, 0, nz) { A(BZ(c), ix, iy, iz) = A(BZ(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix, iy - 1, iz); }
        FOR3(ix, 1, nx, iy, 0, ny, iz, 1, nz) { A(BY(c), ix, iy, iz) = A(BY(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix, iy, iz - 1); }
    } else {
        b_flux(c, 2, BX(c), VZ1(c), 0, -1, 0, 1, 0, 0, flux); /* vz_bx_flux */
        FOR3(ix, 0, nx, iy, 1, ny, iz, 1, nz) { A(BX(c), ix, iy, iz) = A(BX(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix, iy, iz - 1); }
        FOR3(ix, 1, nx, iy, 1, ny, iz, 0, nz) { A(BZ(c), ix, iy, iz) = A(BZ(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix - 1, iy, iz); }
        b_flux(c, 2, BY(c), VZ1(c), -1, 0, 0, 0, 1, 0, flux); /* vz_by_flux */
        FOR3(ix, 1, nx, iy, 0, ny, iz, 1, nz) { A(BY(c), ix, iy, iz) = A(BY(c), ix, iy, iz) - A(flux, ix, iy, iz) + A(flux, ix, iy, iz - 1); }
        FOR3(ix, 1, nx, iy, 1, ny, iz, 0, nz) { A(BZ(c), ix, iy, iz) = A(BZ(c), ix, iy, iz) + A(flux, ix, iy, iz) - A(flux, ix, iy - 1, iz); }
    }

    /* mass */
    mass_flux(c, d);
    bcs(c, LO_DM, d); /* dm_x_bcs / dm_y_bcs / dm_z_bcs */
    FOR3(ix, 1, nx, iy, 1, ny, iz, 1, nz)
    {
        A(rho, ix, iy, iz) = (A(rho1, ix, iy, iz) * A(cv1, ix, iy, iz) + A(dm, ix - di, iy - dj, iz - dk) -
                              A(dm, ix, iy, iz)) / A(cv2, ix, iy, iz);
    }
    /* energies */
    for (int which = 0; which < 2; ++which) {
        a3 e = which == 0 ? EE(c) : EI(c);
        energy_flux(c, d, e, flux);
        FOR3(ix, 1, nx, iy, 1, ny, iz, 1, nz)
        {
            A(e, ix, iy, iz) = (A(e, ix, iy, iz) * A(cv1, ix, iy, iz) * A(rho1, ix, iy, iz) +
                                A(flux, ix - di, iy - dj, iz - dk) - A(flux, ix, iy, iz)) /
                               (A(cv2, ix, iy, iz) * A(rho, ix, iy, iz));
        }
    }
    /* vertex density before remap */
    {
        const long x0 = d == 0 ? -1 : 0, x1 = d == 0 ? nx + 1 : nx;
        const long y0 = d == 1 ? -1 : 0, y1 = d == 1 ? ny + 1 : ny;
        const long z0 = d == 2 ? -1 : 0, z1 = d == 2 ? nz + 1 : nz;
        FOR3(ix, x0, x1, iy, y0, y1, iz, z0, z1)
        {
            long ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
            double s = A(rho1, ix, iy, iz) * A(cv1, ix, iy, iz) + A(rho1, ixp, iy, iz) * A(cv1, ixp, iy, iz) +
                       A(rho1, ix, iyp, iz) * A(cv1, ix, iyp, iz) + A(rho1, ixp, iyp, iz) * A(cv1, ixp, iyp, iz) +
                       A(rho1, ix, iy, izp) * A(cv1, ix, iy, izp) + A(rho1, ixp, iy, izp) * A(cv1, ixp, iy, izp) +
                       A(rho1, ix, iyp, izp) * A(cv1, ix, iyp, izp) + A(rho1, ixp, iyp, izp) * A(cv1, ixp, iyp, izp);
            if (d == 0) {
                A(rho_v, ix, iy, iz) = s * 0.125 / A(cvc1, ix, iy, iz);
            } else {
                A(rho_v, ix, iy, iz) = s;
                A(rho_v, ix, iy, iz) *= 0.125 / A(cvc1, ix, iy, iz);
            }
        }
    }
    /* vertex cv2 through the flux temporary */
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
    {
        long ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        double s = A(cv2, ix, iy, iz) + A(cv2, ixp, iy, iz) + A(cv2, ix, iyp, iz) + A(cv2, ixp, iyp, iz) +
                   A(cv2, ix, iy, izp) + A(cv2, ixp, iy, izp) + A(cv2, ix, iyp, izp) + A(cv2, ixp, iyp, izp);
        A(flux, ix, iy, iz) = d == 0 ? s * 0.125 : 0.125 * s;
    }
    FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz) { A(cv2, ix, iy, iz) = A(flux, ix, iy, iz); }
    /* shift the sweep velocity to the face */
    {
        a3 w = d == 0 ? VX1(c) : d == 1 ? VY1(c) : VZ1(c);
        const long x0 = d == 0 ? -2 : 0, x1 = d == 0 ? nx + 1 : nx;
        const long y0 = d == 1 ? -2 : 0, y1 = d == 1 ? ny + 1 : ny;
        const long z0 = d == 2 ? -2 : 0, z1 = d == 2 ? nz + 1 : nz;
        FOR3(ix, x0, x1, iy, y0, y1, iz, z0, z1)
        {
            A(flux, ix, iy, iz) = (A(w, ix, iy, iz) + A(w, ix + di, iy + dj, iz + dk)) * 0.5;
        }
        FOR3(ix, x0, x1, iy, y0, y1, iz, z0, z1) { A(w, ix, iy, iz) = A(flux, ix, iy, iz); }
    }
    /* shift the mass flux to the vertex cell */
    {
        const long x0 = d == 0 ? -1 : 0, y0 = d == 1 ? -1 : 0, z0 = d == 2 ? -1 : 0;
        FOR3(ix, x0, nx, iy, y0, ny, iz, z0, nz)
        {
            long ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
            double s = A(dm, ix, iy, iz) + A(dm, ixp, iy, iz) + A(dm, ix, iyp, iz) + A(dm, ixp, iyp, iz) +
                       A(dm, ix, iy, izp) + A(dm, ixp, iy, izp) + A(dm, ix, iyp, izp) + A(dm, ixp, iyp, izp);
            A(flux, ix, iy, iz) = d == 0 ? s * 0.125 : 0.125 * s;
        }
        FOR3(ix, x0, nx, iy, y0, ny, iz, z0, nz) { A(dm, ix, iy, iz) = A(flux, ix, iy, iz); }
    }
    /* vertex density after remap; z covers -1..nz+1 (zremap.cpp:222) */
    {
        const long z0 = d == 2 ? -1 : 0, z1 = d == 2 ? nz + 1 : nz;
        FOR3(ix, 0, nx, iy, 0, ny, iz, z0, z1)
        {
            A(rho_v1, ix, iy, iz) = (A(rho_v, ix, iy, iz) * A(cvc1, ix, iy, iz) +
                                     A(dm, ix - di, iy - dj, iz - dk) - A(dm, ix, iy, iz)) / A(cv2, ix, iy, iz);
        }
    }
    /* momentum; y_mom_flux clears its flux array first (yremap.cpp:511) */
    for (int comp = 0; comp < 3; ++comp) {
        a3 v = comp == 0 ? VX(c) : comp == 1 ? VY(c) : VZ(c);
        if (d == 1) a3_fill(flux, 0.0);
        mom_flux(c, d, v, flux);
        FOR3(ix, 0, nx, iy, 0, ny, iz, 0, nz)
        {
            A(v, ix, iy, iz) = (A(rho_v, ix, iy, iz) * A(v, ix, iy, iz) * A(cvc1, ix, iy, iz) +
                                A(flux, ix - di, iy - dj, iz - dk) - A(flux, ix, iy, iz)) /
                               (A(cv2, ix, iy, iz) * A(rho_v1, ix, iy, iz));
        }
    }
    boundary_conditions(c);
    if (d == 0) c->xpass = 0.0;
    else if (d == 1) c->ypass = 0.0;
    else c->zpass = 0.0;
}

/* LARE3D::eulerian_remap (remap.cpp:23-111): Strang order is always
 * x, y, z because simulationData::step is never advanced (remap.cpp:52). */
void lo_end_of_timestep(lo_ctx *c)
{
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    if (c->rke) a3_fill(c->delta_ke, 0.0); /* remap.cpp:38-41 */
    c->xpass = 1.0;
    c->ypass = 1.0;
    c->zpass = 1.0;
    FOR3(ix, -2, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2) { A(BX(c), ix, iy, iz) *= A(c->dxab, ix, iy, iz); }
    FOR3(ix, -1, nx + 2, iy, -2, ny + 2, iz, -1, nz + 2) { A(BY(c), ix, iy, iz) *= A(c->dyab, ix, iy, iz); }
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -2, nz + 2) { A(BZ(c), ix, iy, iz) *= A(c->dzab, ix, iy, iz); }
    remap_sweep(c, 0);
    remap_sweep(c, 1);
    remap_sweep(c, 2);
    FOR3(ix, -2, nx + 2, iy, -1, ny + 2, iz, -1, nz + 2) { A(BX(c), ix, iy, iz) /= (A(c->dxab, ix, iy, iz) + c->none_zero); }
    FOR3(ix, -1, nx + 2, iy, -2, ny + 2, iz, -1, nz + 2) { A(BY(c), ix, iy, iz) /= (A(c->dyab, ix, iy, iz) + c->none_zero); }
    FOR3(ix, -1, nx + 2, iy, -1, ny + 2, iz, -2, nz + 2) { A(BZ(c), ix, iy, iz) /= (A(c->dzab, ix, iy, iz) + c->none_zero); }
    bfield_bcs(c);
    if (c->rke) energy_correction(c); /* shared_data.h:415-417 */
    /* eta_calc (shared_data.h:418) only feeds the resistive path, which
     * recomputes it before use; it runs here for fidelity when resistive. */
    if (c->resistive) eta_calc(c);
}

/* One runner iteration (runner.h:564-586): startOfTimestep, halfTimestep,
 * endOfTimestep. */
double lo_step(lo_ctx *c)
{
    double dt = lo_start_of_timestep(c);
    lo_half_timestep(c);
    lo_end_of_timestep(c);
    return dt;
}
