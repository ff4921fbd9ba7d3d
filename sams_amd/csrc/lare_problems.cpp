// Initial conditions and boundary conditions of the benchmark problems
// (include/lare_hip.h, lare_problem_*): host code that plays the role of
// the reference's initial-condition packages.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lare_hip.h"

extern "C" const char *lare_last_error(void);

namespace {

thread_local std::string p_err;

// splitmix64 finaliser; one counter-based stream per (seed, stream, index)
inline uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t lin)
{
    uint64_t h = sm64(sm64(seed + 0x632BE59BD9B4E019ull * stream) ^ lin);
    return (double)(h >> 11) * 0x1.0p-53;
}
inline double gauss(uint64_t seed, uint64_t comp, uint64_t lin)
{
    double u1 = uniform01(seed, 16 + 2 * comp, lin);
    double u2 = uniform01(seed, 17 + 2 * comp, lin);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
}

const int kStag[LARE_NFIELDS][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 1, 1}, {0, 0, 0}};

// Linear axes of axisRegistry::setDomain (axisRegistry.h:973-984).
void axes(long n, double mn, double mx, std::vector<double> &xb, std::vector<double> &xc)
{
    xb.assign(n + 5, 0.0);
    xc.assign(n + 5, 0.0);
    const double delta = (mx - mn) / (double)(n + 1 - 1);
    for (long i = -2; i <= n + 2; ++i) xb[i + 2] = mn + delta * (double)i;
    for (long i = -1; i <= n + 2; ++i) xc[i + 2] = 0.5 * (xb[i + 2] + xb[i + 1]);
}

}  // namespace

extern "C" int lare_problem_desc(const lare_problem *p, lare_desc *d)
{
    if (!p || !d) return -1;
    std::memset(d, 0, sizeof(*d));
    d->nx = p->nx;
    d->ny = p->ny;
    d->nz = p->nz;
    d->visc1 = 0.1;
    d->visc2 = 1.0;
    d->mu0 = 1.0;
    d->dt_multiplier = 0.8;
    d->precision = p->precision;
    d->rke = p->rke != 0;
    d->j_max = 1.0e30;
    if (p->kind == LARE_PROBLEM_ORSZAG_TANG) {
        d->x_min = 0.0;
        d->x_max = 1.0;
        d->y_min = 0.0;
        d->y_max = 1.0;
        d->z_min = 0.0;
        d->z_max = 1.0;
        d->gas_gamma = 5.0 / 3.0;
    } else if (p->kind == LARE_PROBLEM_BRIO_WU) {
        d->x_min = -1.0;
        d->x_max = 1.0;
        d->y_min = 0.0;
        d->y_max = (d->x_max - d->x_min) * (double)p->ny / (double)p->nx;
        d->z_min = 0.0;
        d->z_max = (d->x_max - d->x_min) * (double)p->nz / (double)p->nx;
        d->gas_gamma = 2.0;
    } else {
        return -1;
    }
    d->x_min += p->x0;
    d->x_max += p->x0;
    d->y_min += p->y0;
    d->y_max += p->y0;
    if (p->geometry < LARE_CARTESIAN || p->geometry > LARE_SPHERICAL) return -1;
    d->geometry = p->geometry;
    if (p->eta < 0.0 || p->eta0 < 0.0) return -1;
    if (p->eta > 0.0 || p->eta0 > 0.0) {
        d->resistive = 1;
        d->eta_background = p->eta;
        d->eta0 = p->eta0;
        if (p->j_max > 0.0) d->j_max = p->j_max;
    }
    return 0;
}

extern "C" int lare_problem_fill(const lare_problem *p, int field, double *host)
{
    lare_desc d;
    if (!host || field < 0 || field >= LARE_NFIELDS || lare_problem_desc(p, &d)) return -1;
    const long nx = p->nx, ny = p->ny, nz = p->nz;
    long lb[3], ub[3];
    const long n3[3] = {nx, ny, nz};
    for (int k = 0; k < 3; ++k) {
        lb[k] = kStag[field][k] ? -2 : -1;
        ub[k] = n3[k] + 2;
    }
    const long nyr = ub[1] - lb[1] + 1, nzr = ub[2] - lb[2] + 1;
    std::memset(host, 0, sizeof(double) * (size_t)(ub[0] - lb[0] + 1) * nyr * nzr);
    std::vector<double> xb, xc, yb, yc, zb, zc;
    axes(nx, d.x_min, d.x_max, xb, xc);
    axes(ny, d.y_min, d.y_max, yb, yc);
    axes(nz, d.z_min, d.z_max, zb, zc);
    const double gam = d.gas_gamma;
    const long ny4 = ny + 4, nz4 = nz + 4;
    // x planes split over host threads (every value depends on its own
    // indices only, so the result does not depend on the split)
    auto planes = [&](long x0, long x1) {
    for (long ix = x0; ix <= x1; ++ix)
        for (long iy = -1; iy <= ny + 2; ++iy)
            for (long iz = -1; iz <= nz + 2; ++iz) {
                const uint64_t lin = (uint64_t)(((ix + 1) * ny4 + (iy + 1)) * nz4 + (iz + 1));
                double v = 0.0;
                if (p->kind == LARE_PROBLEM_ORSZAG_TANG) {
                    const double r = 25.0 / (36.0 * M_PI);
                    double pr = 5.0 / (12.0 * M_PI);
                    if (p->beta > 0.0) pr = p->beta * (1.0 / (4.0 * M_PI)) / 2.0;
                    const double b0 = std::sqrt(4.0 * M_PI);
                    switch (field) {
                    case LARE_RHO: v = r; break;
                    case LARE_ENERGY_ELECTRON:
                    case LARE_ENERGY_ION: v = pr / ((gam - 1.0) * r) / 2.0; break;
                    case LARE_VX:
                        v = -std::sin(2.0 * M_PI * yb[iy + 2]);
                        if (p->amp != 0.0) v = v + p->amp * gauss(p->seed, 0, lin);
                        break;
                    case LARE_VY:
                        v = std::sin(2.0 * M_PI * xb[ix + 2]);
                        if (p->amp != 0.0) v = v + p->amp * gauss(p->seed, 1, lin);
                        break;
                    case LARE_VZ:
                        v = 0.0;
                        if (p->amp != 0.0) v = v + p->amp * gauss(p->seed, 2, lin);
                        break;
                    case LARE_BX: v = -std::sin(2.0 * M_PI * yc[iy + 2]) / b0; break;
                    case LARE_BY: v = std::sin(4.0 * M_PI * xc[ix + 2]) / b0; break;
                    default: v = 0.0;
                    }
                } else {
                    double r, pr, b;
                    if (xc[ix + 2] < 0.5 * (d.x_min + d.x_max)) {
                        r = 1.0;
                        b = 1.0;
                        pr = 1.0;
                    } else {
                        r = 0.125;
                        b = -1.0;
                        pr = 0.1;
                    }
                    if (p->noise != 0.0) r = r * (1.0 + p->noise * (2.0 * uniform01(p->seed, 3, lin) - 1.0));
                    switch (field) {
                    case LARE_RHO: v = r; break;
                    case LARE_ENERGY_ELECTRON:
                    case LARE_ENERGY_ION: v = pr / ((gam - 1.0) * r) / 2.0; break;
                    case LARE_BX: v = 0.75; break;
                    case LARE_BY: v = b; break;
                    default: v = 0.0;
                    }
                }
                host[((ix - lb[0]) * nyr + (iy - lb[1])) * nzr + (iz - lb[2])] = v;
            }
    };
    const long np = nx + 4;
    const long work = np * (ny + 4) * (nz + 4);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long nt = std::min<long>((long)hw, std::max<long>(1, work / 65536));
    if (nt <= 1) {
        planes(-1, nx + 2);
        return 0;
    }
    std::vector<std::thread> th;
    for (long t = 0; t < nt; ++t) {
        const long a = -1 + np * t / nt, b = -1 + np * (t + 1) / nt - 1;
        if (a <= b) th.emplace_back(planes, a, b);
    }
    for (auto &t : th) t.join();
    return 0;
}

extern "C" int lare_problem_bcs(lare_ctx *c, const lare_problem *p)
{
    if (!c || !p) return -1;
    int bx = LARE_BC_PERIODIC;
    if (p->kind == LARE_PROBLEM_BRIO_WU) bx = p->bcx == 0 ? LARE_BC_ZERO_GRADIENT : p->bcx;
    for (int f = 0; f < LARE_NFIELDS; ++f) {
        int k = bx;
        const bool vel = f == LARE_VX || f == LARE_VY || f == LARE_VZ || f == LARE_VX1 || f == LARE_VY1 || f == LARE_VZ1;
        if (bx == LARE_BC_CLAMP && !vel) k = LARE_BC_ZERO_GRADIENT;
        for (int e = 0; e < 2; ++e) {
            if (lare_set_bc(c, f, 0, e, k, 0.0)) return -1;
            if (lare_set_bc(c, f, 1, e, LARE_BC_PERIODIC, 0.0)) return -1;
            if (lare_set_bc(c, f, 2, e, LARE_BC_PERIODIC, 0.0)) return -1;
        }
    }
    return 0;
}

extern "C" int lare_problem_init(lare_ctx *c, const lare_problem *p)
{
    if (!c || !p) return -1;
    if (lare_problem_bcs(c, p)) return -1;
    std::vector<double> buf((size_t)(p->nx + 5) * (p->ny + 5) * (p->nz + 5));
    for (int f = 0; f < 9; ++f) {
        if (lare_problem_fill(p, f, buf.data())) return -1;
        int rc = lare_upload_field(c, f, buf.data());
        if (rc) return rc;
    }
    return 0;
}
