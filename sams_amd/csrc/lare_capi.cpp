// C-ABI of the MI355X LARE3D step (include/lare_hip.h): context set-up, the
// device layout, host<->device transfers in the reference's array layout,
// the boundary-condition table and the step entry points.
#include "lare_hip.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lare_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(-3, std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Stagger per dimension of each registered variable (setup.cpp:45-69).
const int kStag[LARE_NFIELDS][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 1, 1}, {0, 0, 0}};

}  // namespace

struct lare_ctx {
    lare::Grid g;
    lare_desc desc;
    int device;
    struct KStat {
        std::string name;
        int calls;
        double ms;
        double bytes;
    };
    std::vector<KStat> stats;
};

extern "C" const char *lare_last_error(void) { return g_err.c_str(); }

namespace {

// The captured step bakes in the boundary table and the kernel set; any
// change to either drops it and the next lare_steps captures afresh.
void drop_graph(lare::Grid &g)
{
    if (g.graph_exec) (void)hipGraphExecDestroy(g.graph_exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.graph_exec = nullptr;
    g.graph = nullptr;
}

// Capture one step (the kernels read dt and every other scalar from device
// memory, so one capture serves every later step).  No memset or host copy
// is recorded in it.
hipError_t capture_step(lare::Grid &g)
{
    drop_graph(g);
    hipError_t e = hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    lare::step_start(g);
    lare::step_half(g);
    lare::step_end(g);
    e = hipStreamEndCapture(g.stream, &g.graph);
    if (e != hipSuccess) return e;
    return hipGraphInstantiate(&g.graph_exec, g.graph, nullptr, nullptr, 0);
}

}  // namespace

namespace {

// Linear axis and deltas of axisRegistry::setDomain (axisRegistry.h:973-984,
// 393-431, 176-201): xb(i) = min + ((max-min)/n) * i on -2..n+2, centres
// 0.5*(xb(i)+xb(i-1)) on -1..n+2, backward differences with the first
// delta copied from its neighbour.  Stored at [i+2].
void make_axis(long n, double mn, double mx, std::vector<double> &db, std::vector<double> &dc)
{
    std::vector<double> xb(n + 5), xc(n + 5);
    db.assign(n + 5, 0.0);
    dc.assign(n + 5, 0.0);
    const double delta = (mx - mn) / (double)(n + 1 - 1);
    for (long i = -2; i <= n + 2; ++i) xb[i + 2] = mn + delta * (double)i;
    for (long i = -1; i <= n + 2; ++i) db[i + 2] = xb[i + 2] - xb[i + 1];
    db[0] = db[1];
    for (long i = -1; i <= n + 2; ++i) xc[i + 2] = 0.5 * (xb[i + 2] + xb[i + 1]);
    for (long i = 0; i <= n + 2; ++i) dc[i + 2] = xc[i + 2] - xc[i + 1];
    dc[1] = dc[2];
}

}  // namespace

extern "C" int lare_ctx_create(const lare_desc *d, lare_ctx **out)
{
    if (!d || !out) return fail(-1, "lare_ctx_create: null argument");
    *out = nullptr;
    if (d->nx < 3 || d->ny < 3 || d->nz < 3)
        return fail(-1, "lare_ctx_create: each of nx, ny, nz must be at least 3");
    if (d->precision != LARE_FP64)
        return fail(-2, "lare_ctx_create: only LARE_FP64 storage is implemented");
    const long pz = ((d->nz + 5) + 7) / 8 * 8;
    const long sx = (d->ny + 5) * pz;
    const long box = (d->nx + 5) * sx;
    if (box >= (1L << 31)) return fail(-1, "lare_ctx_create: grid too large for 32-bit indexing");

    lare_ctx *c = new (std::nothrow) lare_ctx();
    if (!c) return fail(-4, "lare_ctx_create: out of host memory");
    c->desc = *d;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(-3, "lare_ctx_create: no HIP device");
    }
    lare::Grid &g = c->g;
    g.dev.g.nx = (int)d->nx;
    g.dev.g.ny = (int)d->ny;
    g.dev.g.nz = (int)d->nz;
    g.dev.g.sy = (int)pz;
    g.dev.g.sx = (int)sx;
    g.dev.g.base = (int)(2 * sx + 2 * pz + 2);
    g.box = (size_t)box;
    g.prm.gas_gamma = d->gas_gamma;
    g.prm.visc1 = d->visc1;
    g.prm.visc2 = d->visc2;
    g.prm.visc2_norm = 0.25 * (d->gas_gamma + 1.0) * d->visc2;  // lagran.cpp:179
    g.prm.mu0 = d->mu0;
    g.prm.dt_multiplier = d->dt_multiplier;
    g.prm.resistive = d->resistive;
    g.prm.eta_background = d->eta_background;
    g.prm.eta0 = d->eta0;
    g.prm.j_max = d->j_max;
    g.kernel_set = 1;
    std::memset(g.bck, 0, sizeof(g.bck));
    std::memset(g.bcv, 0, sizeof(g.bcv));

    // one pool: 3D arrays, then the six 1D delta arrays, then the scalars
    const int n3d = 13 + 28 + 6 + 10 + 5;
    const size_t lmax = (size_t)std::max(d->nx, std::max(d->ny, d->nz)) + 5;
    const size_t bytes3 = box * sizeof(double);
    g.pool_bytes = n3d * bytes3 + 6 * lmax * sizeof(double) + 256;
    hipError_t e = hipStreamCreate(&g.stream);
    if (e != hipSuccess) {
        delete c;
        return fail(-3, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipMalloc(&g.pool, g.pool_bytes);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(g.stream);
        delete c;
        return fail(-4, std::string("hipMalloc of the grid pool: ") + hipGetErrorString(e));
    }
    char *p = (char *)g.pool;
    auto take3 = [&]() {
        double *r = (double *)p;
        p += bytes3;
        return r;
    };
    lare::Arr &a = g.dev.a;
    for (int f = 0; f < 13; ++f) a.f[f] = take3();
    double **lag[] = {&a.p_visc, &a.eta, &a.cv, &a.cv1, &a.bx1, &a.by1, &a.bz1, &a.alpha1, &a.alpha2, &a.alpha3,
                      &a.visc_heat, &a.pressure, &a.p_e, &a.p_i, &a.rho_v, &a.cv_v, &a.fx, &a.fy, &a.fz, &a.fxv,
                      &a.fyv, &a.fzv, &a.flux_x, &a.flux_y, &a.flux_z, &a.curlb, &a.cs, &a.cs_v};
    for (double **q : lag) *q = take3();
    double **met[] = {&a.dxab, &a.dyab, &a.dzab, &a.dxac, &a.dyac, &a.dzac};
    for (double **q : met) *q = take3();
    double **rem[] = {&a.rho1, &a.cv2, &a.cvc1, &a.db1, &a.rrho_v, &a.rho_v1, &a.rflux, &a.cv2v, &a.wsh, &a.dmv};
    for (double **q : rem) *q = take3();
    for (int k = 0; k < 5; ++k) a.fl[k] = take3();
    double **ax1[] = {&a.dxb, &a.dyb, &a.dzb, &a.dxc, &a.dyc, &a.dzc};
    for (double **q : ax1) {
        *q = (double *)p;
        p += lmax * sizeof(double);
    }
    a.s = (lare::Scal *)p;

    int rc = 0;
    do {
        e = hipMemsetAsync(g.pool, 0, g.pool_bytes, g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) { rc = fail(-3, std::string("pool clear: ") + hipGetErrorString(e)); break; }
        std::vector<double> db[3], dc[3];
        make_axis(d->nx, d->x_min, d->x_max, db[0], dc[0]);
        make_axis(d->ny, d->y_min, d->y_max, db[1], dc[1]);
        make_axis(d->nz, d->z_min, d->z_max, db[2], dc[2]);
        double *dstb[3] = {a.dxb, a.dyb, a.dzb}, *dstc[3] = {a.dxc, a.dyc, a.dzc};
        for (int k = 0; k < 3 && rc == 0; ++k) {
            e = hipMemcpy(dstb[k], db[k].data(), db[k].size() * sizeof(double), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(dstc[k], dc[k].data(), dc[k].size() * sizeof(double), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = fail(-3, std::string("axis upload: ") + hipGetErrorString(e));
        }
        if (rc) break;
        lare::build_metric(g);
        e = hipStreamSynchronize(g.stream);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) rc = fail(-3, std::string("grid set-up: ") + hipGetErrorString(e));
    } while (0);
    if (rc) {
        (void)hipFree(g.pool);
        (void)hipStreamDestroy(g.stream);
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

extern "C" int lare_ctx_destroy(lare_ctx *c)
{
    if (!c) return 0;
    (void)hipStreamSynchronize(c->g.stream);
    drop_graph(c->g);
    (void)hipFree(c->g.pool);
    (void)hipStreamDestroy(c->g.stream);
    delete c;
    return 0;
}

extern "C" int lare_field_bounds(const lare_ctx *c, int field, int64_t lb[3], int64_t ub[3])
{
    if (!c || !lb || !ub) return fail(-1, "lare_field_bounds: null argument");
    if (field < 0 || field >= LARE_NFIELDS) return fail(-1, "lare_field_bounds: unknown field");
    const int64_t n3[3] = {c->desc.nx, c->desc.ny, c->desc.nz};
    for (int k = 0; k < 3; ++k) {
        lb[k] = kStag[field][k] ? -2 : -1;
        ub[k] = n3[k] + 2;
    }
    return 0;
}

namespace {

// host array (reference extent) <-> device box, one 2D copy per x plane
int copy_field(lare_ctx *c, int field, double *host, bool up)
{
    int64_t lb[3], ub[3];
    int rc = lare_field_bounds(c, field, lb, ub);
    if (rc) return rc;
    if (!host) return fail(-1, "lare field copy: null host pointer");
    const lare::Geo &G = c->g.dev.g;
    const size_t nyr = (size_t)(ub[1] - lb[1] + 1), nzr = (size_t)(ub[2] - lb[2] + 1);
    double *dev = c->g.dev.a.f[field];
    HIPCHK(hipStreamSynchronize(c->g.stream));
    for (int64_t ix = lb[0]; ix <= ub[0]; ++ix) {
        double *d = dev + (G.base + ix * G.sx + lb[1] * G.sy + lb[2]);
        double *h = host + (size_t)(ix - lb[0]) * nyr * nzr;
        if (up)
            HIPCHK(hipMemcpy2D(d, G.sy * sizeof(double), h, nzr * sizeof(double), nzr * sizeof(double), nyr,
                               hipMemcpyHostToDevice));
        else
            HIPCHK(hipMemcpy2D(h, nzr * sizeof(double), d, G.sy * sizeof(double), nzr * sizeof(double), nyr,
                               hipMemcpyDeviceToHost));
    }
    return 0;
}

}  // namespace

extern "C" int lare_upload_field(lare_ctx *c, int field, const double *host)
{
    if (!c) return fail(-1, "lare_upload_field: null context");
    return copy_field(c, field, const_cast<double *>(host), true);
}

extern "C" int lare_download_field(lare_ctx *c, int field, double *host)
{
    if (!c) return fail(-1, "lare_download_field: null context");
    return copy_field(c, field, host, false);
}

extern "C" int lare_set_bc(lare_ctx *c, int field, int dim, int edge, int kind, double value)
{
    if (!c) return fail(-1, "lare_set_bc: null context");
    if (field < 0 || field >= LARE_NFIELDS || dim < 0 || dim > 2 || edge < 0 || edge > 1)
        return fail(-1, "lare_set_bc: field, dim or edge out of range");
    if (kind < LARE_BC_NONE || kind > LARE_BC_CLAMP) return fail(-1, "lare_set_bc: unknown boundary kind");
    c->g.bck[field][dim][edge] = kind;
    c->g.bcv[field][dim][edge] = value;
    drop_graph(c->g);
    return 0;
}

extern "C" int lare_get_dt(lare_ctx *c, double *dt)
{
    if (!c || !dt) return fail(-1, "lare_get_dt: null argument");
    HIPCHK(hipMemcpyAsync(dt, &c->g.dev.a.s->dt, sizeof(double), hipMemcpyDeviceToHost, c->g.stream));
    HIPCHK(hipStreamSynchronize(c->g.stream));
    return 0;
}

extern "C" int lare_set_dt(lare_ctx *c, double dt)
{
    if (!c) return fail(-1, "lare_set_dt: null context");
    HIPCHK(hipMemcpyAsync(&c->g.dev.a.s->dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->g.stream));
    HIPCHK(hipStreamSynchronize(c->g.stream));
    return 0;
}

namespace {
int after_launch(lare_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-3, std::string(what) + ": " + hipGetErrorString(e));
    (void)c;
    return 0;
}
}  // namespace

extern "C" int lare_start_of_timestep(lare_ctx *c, double *dt_out)
{
    if (!c) return fail(-1, "lare_start_of_timestep: null context");
    lare::step_start(c->g);
    int rc = after_launch(c, "lare_start_of_timestep");
    if (rc) return rc;
    if (dt_out) return lare_get_dt(c, dt_out);
    return 0;
}

extern "C" int lare_half_timestep(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_half_timestep: null context");
    lare::step_half(c->g);
    return after_launch(c, "lare_half_timestep");
}

extern "C" int lare_end_of_timestep(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_end_of_timestep: null context");
    lare::step_end(c->g);
    return after_launch(c, "lare_end_of_timestep");
}

extern "C" int lare_step(lare_ctx *c, double *dt_out)
{
    if (!c) return fail(-1, "lare_step: null context");
    lare::step_start(c->g);
    lare::step_half(c->g);
    lare::step_end(c->g);
    int rc = after_launch(c, "lare_step");
    if (rc) return rc;
    if (dt_out) return lare_get_dt(c, dt_out);
    return 0;
}

extern "C" int lare_steps(lare_ctx *c, int32_t nsteps, double *dt_seq)
{
    if (!c) return fail(-1, "lare_steps: null context");
    if (nsteps < 0) return fail(-1, "lare_steps: negative step count");
    lare::Grid &g = c->g;
    const bool graph = g.use_graph && !g.prof.on && nsteps > 0;
    if (graph && !g.graph_exec) {
        hipError_t e = capture_step(g);
        if (e != hipSuccess) {
            drop_graph(g);
            return fail(-3, std::string("lare_steps: capturing the step graph: ") + hipGetErrorString(e));
        }
    }
    double *pinned = nullptr;
    if (dt_seq && nsteps > 0) HIPCHK(hipHostMalloc((void **)&pinned, nsteps * sizeof(double), hipHostMallocDefault));
    int rc = 0;
    for (int i = 0; i < nsteps && !rc; ++i) {
        if (graph) {
            hipError_t e = hipGraphLaunch(g.graph_exec, g.stream);
            if (e != hipSuccess) rc = fail(-3, std::string("lare_steps: hipGraphLaunch: ") + hipGetErrorString(e));
        } else {
            lare::step_start(g);
            lare::step_half(g);
            lare::step_end(g);
        }
        // the step's dt survives in device memory until the next step's
        // start overwrites it; this copy is ordered before that on the stream
        if (pinned && !rc)
            (void)hipMemcpyAsync(pinned + i, &g.dev.a.s->dt, sizeof(double), hipMemcpyDeviceToHost, g.stream);
    }
    if (!rc) rc = after_launch(c, "lare_steps");
    hipError_t e = hipStreamSynchronize(g.stream);
    if (!rc && e != hipSuccess) rc = fail(-3, std::string("lare_steps: ") + hipGetErrorString(e));
    if (pinned) {
        if (!rc) std::memcpy(dt_seq, pinned, nsteps * sizeof(double));
        (void)hipHostFree(pinned);
    }
    return rc;
}

extern "C" int lare_set_graph(lare_ctx *c, int32_t on)
{
    if (!c) return fail(-1, "lare_set_graph: null context");
    c->g.use_graph = on != 0;
    if (!on) drop_graph(c->g);
    return 0;
}

extern "C" int lare_sync(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_sync: null context");
    HIPCHK(hipStreamSynchronize(c->g.stream));
    return 0;
}

extern "C" int lare_profile_begin(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_profile_begin: null context");
    HIPCHK(hipStreamSynchronize(c->g.stream));
    c->g.prof.recs.clear();
    c->g.prof.on = true;
    return 0;
}

extern "C" int lare_profile_end(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_profile_end: null context");
    c->g.prof.on = false;
    HIPCHK(hipStreamSynchronize(c->g.stream));
    c->stats.clear();
    for (auto &r : c->g.prof.recs) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        bool found = false;
        for (auto &k : c->stats)
            if (k.name == r.name) {
                k.calls += 1;
                k.ms += ms;
                k.bytes += r.bytes;
                found = true;
                break;
            }
        if (!found) c->stats.push_back({r.name, 1, (double)ms, r.bytes});
    }
    c->g.prof.recs.clear();
    return 0;
}

extern "C" int lare_profile_count(lare_ctx *c, int32_t *n)
{
    if (!c || !n) return fail(-1, "lare_profile_count: null argument");
    *n = (int32_t)c->stats.size();
    return 0;
}

extern "C" int lare_profile_get(lare_ctx *c, int32_t i, char *name, int32_t name_len, int32_t *calls,
                                double *total_ms, double *bytes)
{
    if (!c || i < 0 || (size_t)i >= c->stats.size()) return fail(-1, "lare_profile_get: index out of range");
    const auto &k = c->stats[i];
    if (name && name_len > 0) {
        std::strncpy(name, k.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (calls) *calls = k.calls;
    if (total_ms) *total_ms = k.ms;
    if (bytes) *bytes = k.bytes;
    return 0;
}

extern "C" int lare_set_kernel_set(lare_ctx *c, int32_t kernel_set)
{
    if (!c) return fail(-1, "lare_set_kernel_set: null context");
    if (kernel_set < 0 || kernel_set > 1) return fail(-1, "lare_set_kernel_set: unknown kernel set");
    HIPCHK(hipStreamSynchronize(c->g.stream));
    c->g.kernel_set = kernel_set;
    drop_graph(c->g);
    return 0;
}

extern "C" int lare_viscosity_dt(lare_ctx *c, double *dt_out, double *dtr_out)
{
    if (!c) return fail(-1, "lare_viscosity_dt: null context");
    lare::step_start_a(c->g);
    int rc = after_launch(c, "lare_viscosity_dt");
    if (rc) return rc;
    lare::Scal s;
    HIPCHK(hipMemcpyAsync(&s, c->g.dev.a.s, sizeof(s), hipMemcpyDeviceToHost, c->g.stream));
    HIPCHK(hipStreamSynchronize(c->g.stream));
    if (dt_out) *dt_out = s.dt;
    if (dtr_out) *dtr_out = s.dtr;
    return 0;
}

extern "C" int lare_lagrange(lare_ctx *c, double dt)
{
    if (!c) return fail(-1, "lare_lagrange: null context");
    lare::Scal s;
    HIPCHK(hipMemcpyAsync(&s, c->g.dev.a.s, sizeof(s), hipMemcpyDeviceToHost, c->g.stream));
    HIPCHK(hipStreamSynchronize(c->g.stream));
    // lagran.cpp:159-162: substeps = int(dt/dtr) + 1, dt_sub = dt/substeps
    int substeps = 1;
    if (c->g.prm.resistive) {
        double q = dt / s.dtr;
        if (!(q >= 0.0) || q > 1.0e6) return fail(-1, "lare_lagrange: dt/dtr out of range for the resistive substeps");
        substeps = (int)q + 1;
    }
    s.dt = dt;
    s.dt_sub = dt / (double)substeps;
    HIPCHK(hipMemcpyAsync(c->g.dev.a.s, &s, sizeof(s), hipMemcpyHostToDevice, c->g.stream));
    lare::step_start_b(c->g, substeps);
    return after_launch(c, "lare_lagrange");
}
