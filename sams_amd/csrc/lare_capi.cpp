// C-ABI of the MI355X LARE3D step (include/lare_hip.h): context set-up, the
// device layout, host<->device transfers in the reference's array layout,
// the boundary-condition table and the step entry points.
#include "lare_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lare_engine.h"

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

// Metric factors of LARE3D::grid (setup.cpp:276-335) and the inputs of the
// geometric forces (lagran.cpp:711-721), on the host (glibc sin and tan, as
// the reference calls them): h1 = hy, hyc, xb, cot(yb) at [i + 2]; h2 = hz,
// hzc, hz1, hz2 at [(ix + 2) (ny + 5) + iy + 2].  Cartesian: every factor 1.
// Cylindrical and spherical: hy = |xb|, hyc = |xc| (the reference's loop also
// writes hyc(-2) from xc(-2), both outside their extents; nothing reads it);
// spherical: hz = hy |sin yb|, hz2 = hyc |sin yb| over y -2..ny+2 and
// hz1 = hy |sin yc|, hzc = hyc |sin yc| over y -1..ny+2.
void make_geometry(const lare_desc &d, const std::vector<double> &xb, const std::vector<double> &xc,
                   const std::vector<double> &yb, const std::vector<double> &yc, std::vector<double> h1[4],
                   std::vector<double> h2[4])
{
    const long nx = d.nx, ny = d.ny;
    const size_t lmax = (size_t)std::max(d.nx, std::max(d.ny, d.nz)) + 5;
    for (int k = 0; k < 4; ++k) {
        h1[k].assign(lmax, 1.0);
        h2[k].assign((size_t)(nx + 5) * (size_t)(ny + 5), 1.0);
    }
    auto P = [&](long ix, long iy) { return (size_t)(ix + 2) * (size_t)(ny + 5) + (size_t)(iy + 2); };
    for (long ix = -2; ix <= nx + 2; ++ix) h1[2][ix + 2] = xb[ix + 2];
    for (long iy = -2; iy <= ny + 2; ++iy) h1[3][iy + 2] = 1.0 / std::tan(yb[iy + 2]);
    if (d.geometry == LARE_CARTESIAN) return;
    for (long ix = -2; ix <= nx + 2; ++ix) {
        h1[0][ix + 2] = std::abs(xb[ix + 2]);
        if (ix >= -1) h1[1][ix + 2] = std::abs(xc[ix + 2]);
    }
    if (d.geometry != LARE_SPHERICAL) return;
    for (long iy = -2; iy <= ny + 2; ++iy) {
        const double s = std::abs(std::sin(yb[iy + 2]));
        for (long ix = -2; ix <= nx + 2; ++ix) h2[0][P(ix, iy)] = h1[0][ix + 2] * s;
        for (long ix = -1; ix <= nx + 2; ++ix) h2[3][P(ix, iy)] = h1[1][ix + 2] * s;
    }
    for (long iy = -1; iy <= ny + 2; ++iy) {
        const double sc = std::abs(std::sin(yc[iy + 2]));
        for (long ix = -2; ix <= nx + 2; ++ix) h2[2][P(ix, iy)] = h1[0][ix + 2] * sc;
        for (long ix = -1; ix <= nx + 2; ++ix) h2[1][P(ix, iy)] = h1[1][ix + 2] * sc;
    }
}

// Stagger per dimension of each registered variable (setup.cpp:45-69).
const int kStag[LARE_NFIELDS][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1},
                                    {1, 1, 1}, {1, 1, 1}, {0, 0, 0}};

}  // namespace

struct lare_ctx {
    lare::Grid g;
    lare_desc desc;
    int device;
    // the grid as a group of one for the engine (kernel set 1)
    lare::Group solo;
    bool solo_ready = false;
    // arrays only the one-kernel-per-reference-kernel set reads
    void *pool0 = nullptr;
    // this device's copy of the engine's energy scratch (lare_energy_noalloc)
    double *scratch = nullptr;
    // device staging of field transfers (allocated on first use)
    void *stage = nullptr;
    size_t stage_bytes = 0;
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

void drop_graphs(lare_ctx *c)
{
    drop_graph(c->g);
    lare::group_drop_graph(c->solo);
}

int ensure_solo(lare_ctx *c)
{
    if (!c->solo_ready) {
        std::string err;
        int rc = lare::group_init(c->solo, {&c->g}, c->g.stream, &c->g.prof, err);
        if (rc) return fail(rc, "grid group: " + err);
        c->solo_ready = true;
    }
    // kernel set 2: lane-walking kernels always; 3: never; 1: by launch
    // size or as tuned (lare_set_tune, lare_autotune)
    if (c->g.kernel_set == 2) c->solo.walk_mode = 1;
    if (c->g.kernel_set == 3) c->solo.walk_mode = 2;
    return 0;
}

// Launch configuration of a group (the fields of lare_tune).
lare_tune default_tune()
{
    lare::Group d;
    return lare_tune{d.walk_mode, d.march_k, d.zx_k, d.zline ? 1 : 0, d.order, d.ct_fuse, 0.0};
}
lare_tune get_tune(const lare::Group &gp)
{
    return lare_tune{gp.walk_mode, gp.march_k, gp.zx_k, gp.zline ? 1 : 0, gp.order, gp.ct_fuse, 0.0};
}
bool valid_tune(const lare_tune &t)
{
    return t.walk >= 0 && t.walk <= 2 && (t.march_k == 8 || t.march_k == 16) && (t.zx_k == 8 || t.zx_k == 16) &&
           (t.zline == 0 || t.zline == 1) && (t.order == 0 || t.order == 1) && t.ct_fuse >= 0 && t.ct_fuse <= 2;
}
void apply_tune(lare::Group &gp, const lare_tune &t)
{
    gp.walk_mode = t.walk;
    gp.march_k = t.march_k;
    gp.zx_k = t.zx_k;
    gp.zline = t.zline != 0;
    gp.order = t.order;
    gp.ct_fuse = t.ct_fuse;
    lare::group_drop_graph(gp);
}

// Autotuner: coordinate descent over the launch configuration of group gp,
// each candidate timed as `steps` replays of one captured step (after one
// warm-up replay) with HIP events on `stream`.  Every member's device pool
// (its whole state, scalars included) is copied aside first and restored at
// the end, so the members continue exactly where they were.
int tune_group(lare::Group &gp, hipStream_t stream, int steps, lare_tune *best_out, const char *who)
{
    const std::string w = who;
    if (steps < 1) return fail(-1, w + ": steps must be at least 1");
    if (gp.prof && gp.prof->on) return fail(-1, w + ": not while profiling");
    for (lare::Grid *g : gp.mem)
        if (g->st != 0) return fail(-1, w + ": only between steps");
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<void *> snap(gp.mem.size(), nullptr);
    auto release = [&]() {
        for (void *p : snap)
            if (p) (void)hipFree(p);
    };
    for (size_t m = 0; m < gp.mem.size(); ++m) {
        hipError_t e = hipMalloc(&snap[m], gp.mem[m]->pool_bytes);
        if (e != hipSuccess) {
            snap[m] = nullptr;
            release();
            return fail(-4, w + ": no device memory for the copy of the state: " + hipGetErrorString(e));
        }
        e = hipMemcpyAsync(snap[m], gp.mem[m]->pool, gp.mem[m]->pool_bytes, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) {
            release();
            return fail(-3, w + ": saving the state: " + hipGetErrorString(e));
        }
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    (void)hipEventCreate(&ea);
    (void)hipEventCreate(&eb);
    auto measure = [&](const lare_tune &t, double &ms) -> hipError_t {
        apply_tune(gp, t);
        lare::group_prepare(gp);
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        lare::group_step(gp);
        e = hipStreamEndCapture(stream, &gp.graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gp.graph_exec, gp.graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(gp.graph_exec, stream);
        if (e == hipSuccess) e = hipEventRecord(ea, stream);
        for (int i = 0; i < steps && e == hipSuccess; ++i) e = hipGraphLaunch(gp.graph_exec, stream);
        if (e == hipSuccess) e = hipEventRecord(eb, stream);
        if (e == hipSuccess) e = hipEventSynchronize(eb);
        float f = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, ea, eb);
        ms = (double)f / steps;
        lare::group_drop_graph(gp);
        return e;
    };
    lare_tune best = get_tune(gp);
    double best_ms = 0.0;
    hipError_t e = measure(best, best_ms);
    // one field at a time, keeping whatever is fastest so far
    const int walks[2] = {1, 2}, ks[2] = {8, 16}, flags[2] = {0, 1}, fuses[3] = {0, 1, 2};
    for (int field = 0; field < 6 && e == hipSuccess; ++field) {
        for (int v = 0; v < (field == 5 ? 3 : 2) && e == hipSuccess; ++v) {
            lare_tune t = best;
            if (field == 0) t.walk = walks[v];
            if (field == 1) t.march_k = ks[v];
            if (field == 2) t.zx_k = ks[v];
            if (field == 3) t.zline = flags[v];
            if (field == 4) t.order = flags[v];
            if (field == 5) t.ct_fuse = fuses[v];
            if (t.walk == best.walk && t.march_k == best.march_k && t.zx_k == best.zx_k && t.zline == best.zline &&
                t.order == best.order && t.ct_fuse == best.ct_fuse)
                continue;
            double ms = 0.0;
            e = measure(t, ms);
            if (e == hipSuccess && ms < best_ms) {
                best = t;
                best_ms = ms;
            }
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    apply_tune(gp, best);
    for (size_t m = 0; m < gp.mem.size(); ++m)
        (void)hipMemcpyAsync(gp.mem[m]->pool, snap[m], gp.mem[m]->pool_bytes, hipMemcpyDeviceToDevice, stream);
    hipError_t e2 = hipStreamSynchronize(stream);
    release();
    if (e != hipSuccess) return fail(-3, w + ": " + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(-3, w + ": restoring the state: " + hipGetErrorString(e2));
    best.ms_per_step = best_ms;
    if (best_out) *best_out = best;
    return 0;
}

// The three events on the selected kernel set.
int ev_start_a(lare_ctx *c)
{
    if (c->g.kernel_set == 0) {
        lare::step_start_a(c->g);
        return 0;
    }
    int rc = ensure_solo(c);
    if (!rc) lare::group_start_a(c->solo);
    return rc;
}
int ev_start_b(lare_ctx *c, int substeps)
{
    if (c->g.kernel_set == 0) {
        lare::step_start_b(c->g, substeps);
        return 0;
    }
    int rc = ensure_solo(c);
    if (!rc) lare::group_start_b(c->solo, substeps);
    return rc;
}
int ev_half(lare_ctx *c)
{
    if (c->g.kernel_set == 0) {
        lare::step_half(c->g);
        return 0;
    }
    int rc = ensure_solo(c);
    if (!rc) lare::group_half(c->solo);
    return rc;
}
int ev_end(lare_ctx *c)
{
    if (c->g.kernel_set == 0) {
        lare::step_end(c->g);
        return 0;
    }
    int rc = ensure_solo(c);
    if (!rc) lare::group_end(c->solo);
    return rc;
}
int ev_step(lare_ctx *c)
{
    int rc = ev_start_a(c);
    if (!rc) rc = ev_start_b(c, 2);
    if (!rc) rc = ev_half(c);
    if (!rc) rc = ev_end(c);
    return rc;
}

// Capture one step (the kernels read dt and every other scalar from device
// memory, so one capture serves every later step).  No memset or host copy
// is recorded in it; every table a launch needs exists before capture.
hipError_t capture_step(lare_ctx *c, hipGraphExec_t *exec)
{
    drop_graphs(c);
    hipGraph_t *graph = &c->g.graph;
    hipGraphExec_t *gexec = &c->g.graph_exec;
    if (c->g.kernel_set != 0) {
        if (ensure_solo(c)) return hipErrorInvalidValue;
        lare::group_prepare(c->solo);
        graph = &c->solo.graph;
        gexec = &c->solo.graph_exec;
    }
    hipError_t e = hipStreamBeginCapture(c->g.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    (void)ev_step(c);
    e = hipStreamEndCapture(c->g.stream, graph);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(gexec, *graph, nullptr, nullptr, 0);
    *exec = *gexec;
    return e;
}

hipGraphExec_t captured(lare_ctx *c) { return c->g.kernel_set != 0 ? c->solo.graph_exec : c->g.graph_exec; }

}  // namespace

namespace {

// Linear axis and deltas of axisRegistry::setDomain (axisRegistry.h:973-984,
// 393-431, 176-201): xb(i) = min + ((max-min)/n) * i on -2..n+2, centres
// 0.5*(xb(i)+xb(i-1)) on -1..n+2, backward differences with the first
// delta copied from its neighbour.  Stored at [i+2].
void make_axis(long n, double mn, double mx, std::vector<double> &db, std::vector<double> &dc,
               std::vector<double> *pb = nullptr, std::vector<double> *pc = nullptr)
{
    std::vector<double> xb(n + 5), xc(n + 5, 0.0);
    db.assign(n + 5, 0.0);
    dc.assign(n + 5, 0.0);
    const double delta = (mx - mn) / (double)(n + 1 - 1);
    for (long i = -2; i <= n + 2; ++i) xb[i + 2] = mn + delta * (double)i;
    for (long i = -1; i <= n + 2; ++i) db[i + 2] = xb[i + 2] - xb[i + 1];
    db[0] = db[1];
    for (long i = -1; i <= n + 2; ++i) xc[i + 2] = 0.5 * (xb[i + 2] + xb[i + 1]);
    for (long i = 0; i <= n + 2; ++i) dc[i + 2] = xc[i + 2] - xc[i + 1];
    dc[1] = dc[2];
    if (pb) *pb = xb;  // positions at [i + 2]: xb over -2..n+2, xc over -1..n+2
    if (pc) *pc = xc;
}

}  // namespace

namespace {

// Bytes between consecutive 3D arrays of the pool: the array rounded up to
// `align` bytes plus `skew` (array layout knobs, LARE_ARRAY_ALIGN and
// LARE_ARRAY_SKEW; defaults 256 and 0).
struct PoolLayout {
    size_t align = 256, skew = 0;
    size_t stride(size_t bytes) const { return (bytes + align - 1) / align * align + skew; }
};
PoolLayout pool_layout()
{
    PoolLayout L;
    if (const char *v = std::getenv("LARE_ARRAY_ALIGN")) L.align = std::max<size_t>(256, std::strtoull(v, nullptr, 10));
    if (const char *v = std::getenv("LARE_ARRAY_SKEW")) L.skew = std::strtoull(v, nullptr, 10) / 256 * 256;
    return L;
}

// Engine arrays (lare::ArrE) of one grid in storage type T: 43 T arrays and
// 6 FP64 flux arrays carved from the pool at p.
template <class T>
void carve_engine(lare::ArrE<T> &a, char *&p, size_t box, const PoolLayout &L)
{
    auto tk = [&]() {
        lare::Fld<T> r{(T *)p};
        p += L.stride(box * sizeof(T));
        return r;
    };
    auto dk = [&]() {
        double *r = (double *)p;
        p += L.stride(box * sizeof(double));
        return r;
    };
    lare::Fld<T> *ts[] = {&a.rho, &a.ee, &a.ei, &a.vx, &a.vy, &a.vz, &a.bx, &a.by, &a.bz, &a.vx1, &a.vy1,
                          &a.vz1, &a.cs, &a.cs_v, &a.rho_v, &a.alpha1, &a.alpha2, &a.alpha3, &a.visc_heat,
                          &a.fxv, &a.fyv, &a.fzv, &a.cv1, &a.bx1, &a.by1, &a.bz1, &a.p_e, &a.p_i, &a.eta,
                          &a.curlb, &a.cv2, &a.cvc1, &a.db1, &a.rho2, &a.ee2, &a.ei2, &a.vx2, &a.vy2, &a.vz2,
                          &a.bxy, &a.byy, &a.bzy, &a.vol};
    for (auto *q : ts) *q = tk();
    double **ds[] = {&a.dm, &a.flux_x, &a.flux_y, &a.flux_z, &a.fl[0], &a.fl[1]};
    for (double **q : ds) *q = dk();
}
const int kEngineT = 43, kEngineD = 6;

// The one-kernel-per-reference-kernel set's view of the FP64 engine arrays.
void alias_set0(lare::Arr &a, const lare::ArrE<double> &e)
{
    const lare::Fld<double> f12[12] = {e.rho, e.ee, e.ei, e.vx, e.vy, e.vz, e.bx, e.by, e.bz, e.vx1, e.vy1, e.vz1};
    for (int k = 0; k < 12; ++k) a.f[k] = f12[k].p;
    a.f[12] = e.dm;
    a.eta = e.eta.p;
    a.cv1 = e.cv1.p;
    a.bx1 = e.bx1.p;
    a.by1 = e.by1.p;
    a.bz1 = e.bz1.p;
    a.alpha1 = e.alpha1.p;
    a.alpha2 = e.alpha2.p;
    a.alpha3 = e.alpha3.p;
    a.visc_heat = e.visc_heat.p;
    a.p_e = e.p_e.p;
    a.p_i = e.p_i.p;
    a.rho_v = e.rho_v.p;
    a.fxv = e.fxv.p;
    a.fyv = e.fyv.p;
    a.fzv = e.fzv.p;
    a.flux_x = e.flux_x;
    a.flux_y = e.flux_y;
    a.flux_z = e.flux_z;
    a.curlb = e.curlb.p;
    a.cs = e.cs.p;
    a.cs_v = e.cs_v.p;
    a.cv2 = e.cv2.p;
    a.cvc1 = e.cvc1.p;
    a.db1 = e.db1.p;
    for (int k = 0; k < 2; ++k) a.fl[k] = e.fl[k];
}

// Arrays read only by the one-kernel-per-reference-kernel set (the stored
// metric, cv, pressure, cv_v, the forces, p_visc and the remap
// intermediates), allocated when that set is first selected.
const int kSet0Arrays = 20;
int ensure_set0(lare_ctx *c)
{
    if (c->pool0) return 0;
    if (c->g.fp32) return fail(-2, "the one-kernel-per-reference-kernel set runs FP64 storage only");
    const size_t bytes3 = c->g.box * sizeof(double);
    HIPCHK(hipStreamSynchronize(c->g.stream));
    HIPCHK(hipMalloc(&c->pool0, kSet0Arrays * bytes3));
    char *p = (char *)c->pool0;
    lare::Arr &a = c->g.dev.a;
    double **q[kSet0Arrays] = {&a.p_visc, &a.cv,   &a.pressure, &a.cv_v,  &a.fx,   &a.fy,     &a.fz,
                               &a.dxab,   &a.dyab, &a.dzab,     &a.dxac,  &a.dyac, &a.dzac,   &a.rho1,
                               &a.rho_v1, &a.rflux, &a.cv2v,    &a.wsh,   &a.rrho_v, &a.dmv};
    for (double **r : q) {
        *r = (double *)p;
        p += bytes3;
    }
    HIPCHK(hipMemsetAsync(c->pool0, 0, kSet0Arrays * bytes3, c->g.stream));
    lare::build_metric(c->g);
    HIPCHK(hipStreamSynchronize(c->g.stream));
    HIPCHK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" int lare_ctx_create(const lare_desc *d, lare_ctx **out)
{
    if (!d || !out) return fail(-1, "lare_ctx_create: null argument");
    *out = nullptr;
    if (d->nx < 2 || d->ny < 2 || d->nz < 2)
        return fail(-1, "lare_ctx_create: each of nx, ny, nz must be at least 2");
    if (d->precision != LARE_FP64 && d->precision != LARE_FP32_STORAGE)
        return fail(-1, "lare_ctx_create: unknown storage precision");
    if (d->geometry < LARE_CARTESIAN || d->geometry > LARE_SPHERICAL)
        return fail(-1, "lare_ctx_create: unknown geometry");
    if (d->geometry != LARE_CARTESIAN && (d->precision != LARE_FP64 || d->rke))
        return fail(-1, "lare_ctx_create: cylindrical and spherical grids run the one-kernel-per-reference-kernel "
                        "set, which has FP64 storage and no kinetic-energy correction");
    long zp = 8;  // row pitch rounded to this many elements (LARE_ZPITCH)
    if (const char *v = std::getenv("LARE_ZPITCH")) zp = std::max(1L, std::atol(v));
    const long pz = ((d->nz + 5) + zp - 1) / zp * zp;
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
    g.fp32 = d->precision == LARE_FP32_STORAGE;
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
    g.prm.rke = d->rke != 0;
    g.prm.geometry = d->geometry;
    g.kernel_set = 1;
    std::memset(g.bck, 0, sizeof(g.bck));
    std::memset(g.bcv, 0, sizeof(g.bcv));

    // one pool: the engine's 3D arrays, the six 1D delta arrays, the scalars
    const size_t tsz = g.fp32 ? sizeof(float) : sizeof(double);
    const size_t lmax = (size_t)std::max(d->nx, std::max(d->ny, d->nz)) + 5;
    const PoolLayout lay = pool_layout();
    const size_t plane = (size_t)(d->nx + 5) * (size_t)(d->ny + 5);
    g.pool_bytes = kEngineT * lay.stride(box * tsz) + (kEngineD + (g.prm.rke ? 1 : 0)) * lay.stride(box * sizeof(double)) +
                   (6 + 4) * lmax * sizeof(double) + 4 * plane * sizeof(double) + 256;
    hipError_t e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking);
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
    lare::Arr &a = g.dev.a;
    double *d1[6];
    if (g.fp32) {
        carve_engine(g.e32.a, p, g.box, lay);
    } else {
        carve_engine(g.e64.a, p, g.box, lay);
        alias_set0(a, g.e64.a);
    }
    double *dke = nullptr;  // delta_ke of the kinetic-energy correction
    if (g.prm.rke) {
        dke = (double *)p;
        p += lay.stride(box * sizeof(double));
    }
    for (int k = 0; k < 6; ++k) {
        d1[k] = (double *)p;
        p += lmax * sizeof(double);
    }
    // metric factors of the one-kernel-per-reference-kernel set
    double *geo1[4], *geo2[4];
    for (int k = 0; k < 4; ++k) {
        geo1[k] = (double *)p;
        p += lmax * sizeof(double);
    }
    for (int k = 0; k < 4; ++k) {
        geo2[k] = (double *)p;
        p += plane * sizeof(double);
    }
    a.hy = geo1[0];
    a.hyc = geo1[1];
    a.xbv = geo1[2];
    a.cotv = geo1[3];
    a.hz = geo2[0];
    a.hzc = geo2[1];
    a.hz1 = geo2[2];
    a.hz2 = geo2[3];
    a.dxb = d1[0];
    a.dyb = d1[1];
    a.dzb = d1[2];
    a.dxc = d1[3];
    a.dyc = d1[4];
    a.dzc = d1[5];
    a.s = (lare::Scal *)p;
    const double eta_sum = g.prm.eta_background + g.prm.eta0;
    const int eta_fixed = std::memcmp(&eta_sum, &g.prm.eta_background, sizeof(double)) == 0;
    lare::Phys ph{g.prm.gas_gamma, g.prm.visc1, g.prm.visc2_norm, g.prm.mu0, g.prm.dt_multiplier,
                  g.prm.eta_background, g.prm.eta0, g.prm.j_max, g.prm.resistive, eta_fixed,
                  g.prm.resistive && g.prm.rke};
    auto finish_view = [&](auto &ev) {
        ev.g = g.dev.g;
        ev.p = ph;
        ev.a.dxb = a.dxb;
        ev.a.dyb = a.dyb;
        ev.a.dzb = a.dzb;
        ev.a.dxc = a.dxc;
        ev.a.dyc = a.dyc;
        ev.a.dzc = a.dzc;
        ev.a.dke = dke;
        ev.a.s = a.s;
    };
    if (g.fp32)
        finish_view(g.e32);
    else
        finish_view(g.e64);

    int rc = 0;
    do {
        e = hipMemsetAsync(g.pool, 0, g.pool_bytes, g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) { rc = fail(-3, std::string("pool clear: ") + hipGetErrorString(e)); break; }
        std::vector<double> db[3], dc[3], pxb, pxc, pyb, pyc;
        make_axis(d->nx, d->x_min, d->x_max, db[0], dc[0], &pxb, &pxc);
        make_axis(d->ny, d->y_min, d->y_max, db[1], dc[1], &pyb, &pyc);
        make_axis(d->nz, d->z_min, d->z_max, db[2], dc[2]);
        double *dstb[3] = {a.dxb, a.dyb, a.dzb}, *dstc[3] = {a.dxc, a.dyc, a.dzc};
        for (int k = 0; k < 3 && rc == 0; ++k) {
            e = hipMemcpyAsync(dstb[k], db[k].data(), db[k].size() * sizeof(double), hipMemcpyHostToDevice, g.stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dstc[k], dc[k].data(), dc[k].size() * sizeof(double), hipMemcpyHostToDevice, g.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
            if (e != hipSuccess) rc = fail(-3, std::string("axis upload: ") + hipGetErrorString(e));
        }
        if (rc) break;
        std::vector<double> h1[4], h2[4];
        make_geometry(*d, pxb, pxc, pyb, pyc, h1, h2);
        for (int k = 0; k < 4 && e == hipSuccess; ++k) {
            e = hipMemcpyAsync(geo1[k], h1[k].data(), h1[k].size() * sizeof(double), hipMemcpyHostToDevice, g.stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(geo2[k], h2[k].data(), h2[k].size() * sizeof(double), hipMemcpyHostToDevice,
                                   g.stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) {
            rc = fail(-3, std::string("metric factor upload: ") + hipGetErrorString(e));
            break;
        }
        std::string err;
        const int vrc = lare::grid_fill_volume(g, err);
        if (vrc) rc = fail(vrc, err);
    } while (0);
    if (rc) {
        (void)hipFree(g.pool);
        (void)hipStreamDestroy(g.stream);
        delete c;
        return rc;
    }
    // cylindrical and spherical grids run the one-kernel-per-reference-kernel
    // set, whose stored metric carries the metric factors
    if (d->geometry != LARE_CARTESIAN) {
        rc = ensure_set0(c);
        if (rc) {
            lare_ctx_destroy(c);
            return rc;
        }
        g.kernel_set = 0;
    }
    // set-up of the allocation-free energy diagnostic: the grid as a group
    // of one with its energy launch table, and this device's scratch address
    rc = ensure_solo(c);
    if (rc == 0) {
        lare::group_energy_prepare(c->solo);
        std::string err;
        const int erc = lare::energy_scratch_address(&c->scratch, err);
        if (erc) rc = fail(erc, err);
    }
    if (rc) {
        lare_ctx_destroy(c);
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
    lare::group_free(c->solo);
    if (c->pool0) (void)hipFree(c->pool0);
    if (c->stage) (void)hipFree(c->stage);
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

// Device array of field f where it currently lives (rho, e_e, e_i and v are
// in the engine's second state copy between the corrector and the remap).
template <class T>
void *field_dev_ptr_t(const lare::ArrE<T> &a, int f, int st)
{
    const lare::Fld<T> fs[12] = {a.rho, a.ee, a.ei, a.vx, a.vy, a.vz, a.bx, a.by, a.bz, a.vx1, a.vy1, a.vz1};
    const lare::Fld<T> s2[6] = {a.rho2, a.ee2, a.ei2, a.vx2, a.vy2, a.vz2};
    if (f == LARE_DM) return (void *)a.dm;
    return (st && f < 6) ? (void *)s2[f].p : (void *)fs[f].p;
}
void *field_dev_ptr(lare_ctx *c, int f)
{
    if (c->g.fp32) return field_dev_ptr_t(c->g.e32.a, f, c->g.st);
    return field_dev_ptr_t(c->g.e64.a, f, c->g.st);
}

// host array (reference extent) <-> device box: one contiguous copy through
// the context's staging buffer and one scatter / gather kernel (FP32-storage
// fields are rounded on the way in and widened on the way out)
int copy_field(lare_ctx *c, int field, double *host, bool up)
{
    int64_t lb[3], ub[3];
    int rc = lare_field_bounds(c, field, lb, ub);
    if (rc) return rc;
    if (!host) return fail(-1, "lare field copy: null host pointer");
    const size_t count = (size_t)(ub[0] - lb[0] + 1) * (ub[1] - lb[1] + 1) * (ub[2] - lb[2] + 1);
    if (!c->stage) {
        // the largest field extent, (n+5)^3 doubles
        c->stage_bytes = (size_t)(c->desc.nx + 5) * (c->desc.ny + 5) * (c->desc.nz + 5) * sizeof(double);
        HIPCHK(hipMalloc(&c->stage, c->stage_bytes));
    }
    if (count * sizeof(double) > c->stage_bytes) return fail(-1, "lare field copy: field larger than staging");
    const int l3[3] = {(int)lb[0], (int)lb[1], (int)lb[2]}, u3[3] = {(int)ub[0], (int)ub[1], (int)ub[2]};
    const bool f32 = c->g.fp32 && field != LARE_DM;
    double *st = (double *)c->stage;
    if (up) {
        HIPCHK(hipMemcpyAsync(st, host, count * sizeof(double), hipMemcpyHostToDevice, c->g.stream));
        lare::box_copy(c->g.stream, field_dev_ptr(c, field), f32, st, c->g.dev.g, l3, u3, true);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->g.stream));
    } else {
        lare::box_copy(c->g.stream, field_dev_ptr(c, field), f32, st, c->g.dev.g, l3, u3, false);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host, st, count * sizeof(double), hipMemcpyDeviceToHost, c->g.stream));
        HIPCHK(hipStreamSynchronize(c->g.stream));
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
    c->g.bc_version += 1;
    drop_graphs(c);
    return 0;
}

extern "C" int lare_set_bc_map(lare_ctx *c, int field, int dim, int edge, const int32_t src[3], double value)
{
    if (!c || !src) return fail(-1, "lare_set_bc_map: null argument");
    if (field < 0 || field >= LARE_NFIELDS || dim < 0 || dim > 2 || edge < 0 || edge > 1)
        return fail(-1, "lare_set_bc_map: field, dim or edge out of range");
    if (c->g.kernel_set == 0) return fail(-1, "lare_set_bc_map: map conditions run on kernel sets 1-3 only");
    const int64_t n3[3] = {c->desc.nx, c->desc.ny, c->desc.nz};
    const int s = kStag[field][dim];
    const int nl = s ? 3 : 2;
    for (int k = 0; k < 3; ++k) {
        int m = k < nl ? src[k] : LARE_BC_MAP_KEEP;
        if (m != LARE_BC_MAP_KEEP && m != LARE_BC_MAP_VALUE && (m < (s ? -2 : -1) || m > n3[dim] + 2))
            return fail(-1, "lare_set_bc_map: source layer outside the field's extent");
        c->g.bcm[field][dim][edge][k] = m;
    }
    c->g.bck[field][dim][edge] = LARE_BC_MAP;
    c->g.bcv[field][dim][edge] = value;
    c->g.bc_version += 1;
    drop_graphs(c);
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
    int rc = ev_start_a(c);
    if (!rc) rc = ev_start_b(c, 2);
    if (rc) return rc;
    rc = after_launch(c, "lare_start_of_timestep");
    if (rc) return rc;
    if (dt_out) return lare_get_dt(c, dt_out);
    return 0;
}

extern "C" int lare_half_timestep(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_half_timestep: null context");
    int rc = ev_half(c);
    return rc ? rc : after_launch(c, "lare_half_timestep");
}

extern "C" int lare_end_of_timestep(lare_ctx *c)
{
    if (!c) return fail(-1, "lare_end_of_timestep: null context");
    int rc = ev_end(c);
    return rc ? rc : after_launch(c, "lare_end_of_timestep");
}

extern "C" int lare_step(lare_ctx *c, double *dt_out)
{
    if (!c) return fail(-1, "lare_step: null context");
    int rc = ev_step(c);
    if (rc) return rc;
    rc = after_launch(c, "lare_step");
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
    hipGraphExec_t exec = captured(c);
    if (graph && !exec) {
        hipError_t e = capture_step(c, &exec);
        if (e != hipSuccess) {
            drop_graphs(c);
            return fail(-3, std::string("lare_steps: capturing the step graph: ") + hipGetErrorString(e));
        }
    }
    double *pinned = nullptr;
    if (dt_seq && nsteps > 0) HIPCHK(hipHostMalloc((void **)&pinned, nsteps * sizeof(double), hipHostMallocDefault));
    int rc = 0;
    for (int i = 0; i < nsteps && !rc; ++i) {
        if (graph) {
            hipError_t e = hipGraphLaunch(exec, g.stream);
            if (e != hipSuccess) rc = fail(-3, std::string("lare_steps: hipGraphLaunch: ") + hipGetErrorString(e));
        } else {
            rc = ev_step(c);
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
    if (!on) drop_graphs(c);
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
    if (kernel_set < 0 || kernel_set > 3) return fail(-1, "lare_set_kernel_set: unknown kernel set");
    if (kernel_set != 0 && c->desc.geometry != LARE_CARTESIAN)
        return fail(-1, "lare_set_kernel_set: cylindrical and spherical grids run kernel set 0 only");
    HIPCHK(hipStreamSynchronize(c->g.stream));
    if (c->g.st != 0) return fail(-1, "lare_set_kernel_set: only between steps");
    if (kernel_set == 0) {
        if (c->g.prm.rke)
            return fail(-1, "lare_set_kernel_set: the kinetic-energy correction (rke) runs on kernel sets 1-3 only");
        for (int f = 0; f < LARE_NFIELDS; ++f)
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 2; ++e)
                    if (c->g.bck[f][d][e] == LARE_BC_MAP)
                        return fail(-1, "lare_set_kernel_set: map boundary conditions run on kernel sets 1-3 only");
        int rc = ensure_set0(c);
        if (rc) return rc;
    }
    c->g.kernel_set = kernel_set;
    if (c->solo_ready) apply_tune(c->solo, default_tune());
    drop_graphs(c);
    return 0;
}

extern "C" int lare_viscosity_dt(lare_ctx *c, double *dt_out, double *dtr_out)
{
    if (!c) return fail(-1, "lare_viscosity_dt: null context");
    int rc = ev_start_a(c);
    if (rc) return rc;
    rc = after_launch(c, "lare_viscosity_dt");
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
    int rc = ev_start_b(c, substeps);
    return rc ? rc : after_launch(c, "lare_lagrange");
}

// ------------------------------------------------------------------ batches

struct lare_batch {
    std::vector<lare_ctx *> members;
    lare::Group gp;
    lare::Prof prof;
    hipStream_t stream = nullptr;
    bool use_graph = true;
};

extern "C" int lare_batch_create(lare_ctx *const *members, int32_t n, lare_batch **out)
{
    if (!members || !out || n <= 0) return fail(-1, "lare_batch_create: need at least one member");
    *out = nullptr;
    lare_batch *b = new (std::nothrow) lare_batch();
    if (!b) return fail(-4, "lare_batch_create: out of host memory");
    std::vector<lare::Grid *> grids;
    for (int i = 0; i < n; ++i) {
        if (!members[i]) {
            delete b;
            return fail(-1, "lare_batch_create: null member");
        }
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) {
                delete b;
                return fail(-1, "lare_batch_create: a context appears twice");
            }
        if (members[i]->desc.geometry != LARE_CARTESIAN) {
            delete b;
            return fail(-1, "lare_batch_create: batches hold Cartesian grids (cylindrical and spherical grids "
                            "run kernel set 0 alone)");
        }
        b->members.push_back(members[i]);
        grids.push_back(&members[i]->g);
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(-3, std::string("lare_batch_create: hipStreamCreate: ") + hipGetErrorString(e));
    }
    std::string err;
    int rc = lare::group_init(b->gp, grids, b->stream, &b->prof, err);
    if (rc) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(rc, "lare_batch_create: " + err);
    }
    *out = b;
    return 0;
}

extern "C" int lare_batch_destroy(lare_batch *b)
{
    if (!b) return 0;
    (void)hipStreamSynchronize(b->stream);
    lare::group_free(b->gp);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

extern "C" int lare_batch_set_graph(lare_batch *b, int32_t on)
{
    if (!b) return fail(-1, "lare_batch_set_graph: null batch");
    b->use_graph = on != 0;
    if (!on) lare::group_drop_graph(b->gp);
    return 0;
}

extern "C" int lare_batch_steps(lare_batch *b, int32_t nsteps)
{
    if (!b) return fail(-1, "lare_batch_steps: null batch");
    if (nsteps < 0) return fail(-1, "lare_batch_steps: negative step count");
    if (nsteps == 0) return 0;
    // work a caller enqueued on a member's own stream comes first
    for (lare_ctx *c : b->members) HIPCHK(hipStreamSynchronize(c->g.stream));
    lare::Group &gp = b->gp;
    // a boundary-condition change since the capture invalidates it
    bool stale = gp.bc_seen.size() != gp.mem.size();
    for (size_t m = 0; !stale && m < gp.mem.size(); ++m) stale = gp.bc_seen[m] != gp.mem[m]->bc_version;
    if (stale) lare::group_drop_graph(gp);
    const bool graph = b->use_graph && !b->prof.on;
    if (graph && !gp.graph_exec) {
        lare::group_prepare(gp);
        hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            lare::group_step(gp);
            e = hipStreamEndCapture(b->stream, &gp.graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&gp.graph_exec, gp.graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            lare::group_drop_graph(gp);
            return fail(-3, std::string("lare_batch_steps: capturing the step graph: ") + hipGetErrorString(e));
        }
    }
    for (int i = 0; i < nsteps; ++i) {
        if (graph) {
            hipError_t e = hipGraphLaunch(gp.graph_exec, b->stream);
            if (e != hipSuccess) return fail(-3, std::string("lare_batch_steps: hipGraphLaunch: ") + hipGetErrorString(e));
        } else {
            lare::group_step(gp);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-3, std::string("lare_batch_steps: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int lare_batch_sync(lare_batch *b)
{
    if (!b) return fail(-1, "lare_batch_sync: null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int lare_batch_energy(lare_batch *b, double *out)
{
    if (!b || !out) return fail(-1, "lare_batch_energy: null argument");
    std::vector<double> e;
    std::string err;
    int rc = lare::group_energy(b->gp, e, err);
    if (rc) return fail(rc, err);
    std::memcpy(out, e.data(), e.size() * sizeof(double));
    return 0;
}

extern "C" int lare_energy(lare_ctx *c, double *out)
{
    if (!c || !out) return fail(-1, "lare_energy: null argument");
    if (c->desc.geometry != LARE_CARTESIAN) return fail(-1, "lare_energy: Cartesian grids only");
    int rc = ensure_solo(c);
    if (rc) return rc;
    std::vector<double> e;
    std::string err;
    rc = lare::group_energy(c->solo, e, err);
    if (rc) return fail(rc, err);
    std::memcpy(out, e.data(), 4 * sizeof(double));
    return 0;
}

extern "C" int lare_energy_noalloc(lare_ctx *c, double *out)
{
    if (!c || !out) return fail(-1, "lare_energy_noalloc: null argument");
    if (!c->solo_ready) return fail(-1, "lare_energy_noalloc: context not set up");
    if (c->desc.geometry != LARE_CARTESIAN) return fail(-1, "lare_energy_noalloc: Cartesian grids only");
    std::string err;
    const int rc = lare::group_energy_noalloc(c->solo, c->scratch, out, err);
    if (rc) return fail(rc, "lare_energy_noalloc: " + err);
    return 0;
}

extern "C" int lare_div_check(const double *a, const double *d, double *q, int64_t n, int32_t mode)
{
    if (n < 0 || (n > 0 && (!a || !d || !q))) return fail(-1, "lare_div_check: null argument");
    if (mode < 0 || mode > 2) return fail(-1, "lare_div_check: mode must be 0, 1 or 2");
    if (n > (int64_t(1) << 31)) return fail(-1, "lare_div_check: at most 2^31 values");
    std::string err;
    int rc = lare::div_check(a, d, q, (long)n, mode, err);
    if (rc) return fail(rc, err);
    return 0;
}

extern "C" int lare_autotune(lare_ctx *c, int32_t steps, lare_tune *out)
{
    if (!c) return fail(-1, "lare_autotune: null context");
    if (c->g.kernel_set != 1) return fail(-1, "lare_autotune: only on kernel set 1");
    int rc = ensure_solo(c);
    if (rc) return rc;
    drop_graphs(c);
    return tune_group(c->solo, c->g.stream, steps, out, "lare_autotune");
}

extern "C" int lare_get_tune(lare_ctx *c, lare_tune *out)
{
    if (!c || !out) return fail(-1, "lare_get_tune: null argument");
    int rc = ensure_solo(c);
    if (rc) return rc;
    *out = get_tune(c->solo);
    return 0;
}

extern "C" int lare_set_tune(lare_ctx *c, const lare_tune *t)
{
    if (!c || !t) return fail(-1, "lare_set_tune: null argument");
    if (!valid_tune(*t)) return fail(-1, "lare_set_tune: invalid configuration");
    if (c->g.kernel_set != 1) return fail(-1, "lare_set_tune: only on kernel set 1");
    int rc = ensure_solo(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->g.stream));
    drop_graphs(c);
    apply_tune(c->solo, *t);
    return 0;
}

extern "C" int lare_batch_autotune(lare_batch *b, int32_t steps, lare_tune *out)
{
    if (!b) return fail(-1, "lare_batch_autotune: null batch");
    for (lare_ctx *c : b->members) HIPCHK(hipStreamSynchronize(c->g.stream));
    return tune_group(b->gp, b->stream, steps, out, "lare_batch_autotune");
}

extern "C" int lare_batch_get_tune(lare_batch *b, lare_tune *out)
{
    if (!b || !out) return fail(-1, "lare_batch_get_tune: null argument");
    *out = get_tune(b->gp);
    return 0;
}

extern "C" int lare_batch_set_tune(lare_batch *b, const lare_tune *t)
{
    if (!b || !t) return fail(-1, "lare_batch_set_tune: null argument");
    if (!valid_tune(*t)) return fail(-1, "lare_batch_set_tune: invalid configuration");
    HIPCHK(hipStreamSynchronize(b->stream));
    apply_tune(b->gp, *t);
    return 0;
}

extern "C" int lare_batch_profile_begin(lare_batch *b)
{
    if (!b) return fail(-1, "lare_batch_profile_begin: null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->prof.recs.clear();
    b->prof.on = true;
    return 0;
}

extern "C" int lare_batch_profile_end(lare_batch *b, double *total_ms, int32_t *launches)
{
    if (!b) return fail(-1, "lare_batch_profile_end: null batch");
    b->prof.on = false;
    HIPCHK(hipStreamSynchronize(b->stream));
    double ms_sum = 0.0;
    for (auto &r : b->prof.recs) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        ms_sum += ms;
    }
    if (total_ms) *total_ms = ms_sum;
    if (launches) *launches = (int32_t)b->prof.recs.size();
    b->prof.recs.clear();
    return 0;
}
