// LARE::LARE3D on the MI355X: the package's step events implemented over
// the C-ABI of include/lare_hip.h.
//
// This file replaces, in a SAMS build, the reference package's
// src/PhysicsPackages/LARE/{lagran,remap,xremap,yremap,zremap,diagnostics}.cpp.
// Everything else stays the reference's: the class and data declarations
// (shared_data.h), setup.cpp (registration, allocation, grid), control.cpp,
// boundary.cpp, the runner, the harness and every initial-condition package.
// The harness arrays remain the host mirror: the state is uploaded at the
// first step and downloaded before output.
//
// Event mapping (shared_data.h:397-419):
//   startOfTimestep -> lagrangian_step: lare_viscosity_dt, then
//       controlFunctions::calculateTimestep (runner.h:158-169, which calls
//       back set_dt and getTimestep), then lare_lagrange with the runner's dt
//   halfTimestep    -> corrector_step: lare_half_timestep
//   endOfTimestep   -> eulerian_remap: lare_end_of_timestep, with the
//       kinetic-energy correction when rke is set (energy_correction is then
//       a no-op); eta_calc is recomputed on the device inside the resistive
//       substeps
//
// The device context of a simulationData is rebuilt when the simulationData
// at that address has other dimensions or other harness arrays than the
// context was built from (a new simulation reusing the address), and every
// context is released when the process exits (or by lare3d_hip_release).
// Once the device holds the state, the package-private host arrays of
// setup.cpp that only the step reads (metric, Lagrangian and remap work
// arrays) are released (LARE3D_HIP_KEEP_HOST_ARRAYS=1 keeps them); the
// registered variables and the axes stay.  LARE3D_HIP_PRECISION=fp32 selects
// FP32 storage (Cartesian grids).
//
// Host writes between events: after the first upload the harness arrays of
// the 13 registered variables are mapped read-only (every whole page inside
// each array).  A write from another package faults once, is recorded and
// the array is made writable again, so the write completes.  At the next
// event the written arrays are uploaded when the harness held the current
// state at the time (after an output or a download), and otherwise the step
// stops with an error naming the array: the device then held a newer state
// than the one the package modified (LARE3D_HIP_HOST_SYNC=1 keeps the
// harness current after every event for such packages).
//
// Boundary conditions: the built-in BC objects attached to each variable are
// private to variableDef (variableDef.h:56).  They are identified once, at
// the first step, by applying each variable's conditions along one dimension
// (variableRegistry::applyBoundaryConditions(name, dim)) to a marked copy of
// the array and reading which kind of fill each edge performed; the array is
// restored afterwards.  Conditions that match none of the built-in fills
// (simplePeriodicBC, simpleZeroGradientBC, simpleMirrorBC, simpleClamp) are
// identified layer by layer as copies of other layers of the line, constants
// or untouched layers (the LARE-style conditions of the LARE initial-condition
// package); any other effect, and a condition whose effect differs from one
// line of the face to another, is reported as an error.
#include <csignal>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/mman.h>
#include <unistd.h>

#include "shared_data.h"
#include "lare_hip.h"
#if defined(USE_HDF5)
#include "io/writerHDF5.h"
#else
#include "io/writerSimple.h"
#endif

namespace LARE
{
    namespace pw = portableWrapper;

    namespace
    {
        // ---- host-write guard (see the file comment)
        struct guardRegion
        {
            char *lo, *hi;              // whole pages inside one harness array
            volatile sig_atomic_t hit;  // written since armed
            volatile sig_atomic_t used;
        };
        constexpr int kMaxRegions = 1024;
        guardRegion g_regions[kMaxRegions];
        struct sigaction g_prevSegv;
        bool g_handlerInstalled = false;

        void segvHandler(int sig, siginfo_t *si, void *uc)
        {
            char *a = static_cast<char *>(si->si_addr);
            for (int i = 0; i < kMaxRegions; ++i)
            {
                guardRegion &r = g_regions[i];
                if (r.used && a >= r.lo && a < r.hi)
                {
                    r.hit = 1;
                    mprotect(r.lo, static_cast<size_t>(r.hi - r.lo), PROT_READ | PROT_WRITE);
                    return;
                }
            }
            // not a guarded array: restore the previous disposition and let
            // the faulting instruction run again under it
            sigaction(SIGSEGV, &g_prevSegv, nullptr);
            (void)sig;
            (void)uc;
        }

        void installGuardHandler()
        {
            if (g_handlerInstalled)
                return;
            struct sigaction sa;
            std::memset(&sa, 0, sizeof(sa));
            sa.sa_sigaction = segvHandler;
            sa.sa_flags = SA_SIGINFO;
            sigemptyset(&sa.sa_mask);
            if (sigaction(SIGSEGV, &sa, &g_prevSegv) != 0)
                throw std::runtime_error("LARE3D (MI355X): cannot install the host-write guard");
            g_handlerInstalled = true;
        }

        // Device grid bound to one simulationData.
        struct deviceGrid
        {
            lare_ctx *ctx = nullptr;
            // what the context was built from: a simulationData at the same
            // address with other sizes or other harness arrays is a new one
            long nx = 0, ny = 0, nz = 0;
            const void *arrays[LARE_NFIELDS] = {};
            int region[LARE_NFIELDS] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};  // guard slot per field
            bool guarded = false;
            bool hostCurrent = true; // harness arrays hold the latest state
            double dtNew = 0.0;      // dt of the last lare_viscosity_dt
            double dtrNew = 0.0;
            double devDt = 0.0;      // previous dt held by the device
            // host-synchronised mode (LARE3D_HIP_HOST_SYNC=1): the state is
            // downloaded after every event and uploaded before the next, so
            // that packages which write it on the host between LARE's events
            // (the obstacles of KarmanVortex and EmeryWindTunnel) are seen
            bool hostSync = false;
        };

        std::map<const simulationData *, deviceGrid> &grids()
        {
            static std::map<const simulationData *, deviceGrid> g;
            return g;
        }

        void unguard(deviceGrid &g)
        {
            if (!g.guarded)
                return;
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                const int i = g.region[f];
                if (i < 0)
                    continue;
                mprotect(g_regions[i].lo, static_cast<size_t>(g_regions[i].hi - g_regions[i].lo),
                         PROT_READ | PROT_WRITE);
            }
        }
        void reguard(deviceGrid &g)
        {
            if (!g.guarded)
                return;
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                const int i = g.region[f];
                if (i < 0)
                    continue;
                g_regions[i].hit = 0;
                if (mprotect(g_regions[i].lo, static_cast<size_t>(g_regions[i].hi - g_regions[i].lo), PROT_READ) != 0)
                    throw std::runtime_error("LARE3D (MI355X): mprotect of a harness array failed");
            }
        }
        void releaseGuard(deviceGrid &g)
        {
            unguard(g);
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                if (g.region[f] >= 0)
                    g_regions[g.region[f]].used = 0;
                g.region[f] = -1;
            }
            g.guarded = false;
        }

        void releaseGrid(deviceGrid &g)
        {
            releaseGuard(g);
            if (g.ctx)
                lare_ctx_destroy(g.ctx);
            g.ctx = nullptr;
        }

        // every context, at process exit (registered at the first bind, after
        // the HIP runtime is initialised, so this runs before its teardown)
        void releaseAll()
        {
            for (auto &kv : grids())
                releaseGrid(kv.second);
            grids().clear();
        }

        void check(int rc, const char *what)
        {
            if (rc != 0)
                throw std::runtime_error(std::string("LARE3D (MI355X): ") + what + ": " + lare_last_error());
        }

        const char *kVarNames[LARE_NFIELDS] = {"rho", "energy_electron", "energy_ion", "vx", "vy", "vz",
                                               "bx", "by", "bz", "LARE/vx1", "LARE/vy1", "LARE/vz1",
                                               "LARE/dm"};

        volumeArray *stateArray(simulationData &d, int f)
        {
            switch (f)
            {
            case LARE_RHO: return &d.rho;
            case LARE_ENERGY_ELECTRON: return &d.energy_electron;
            case LARE_ENERGY_ION: return &d.energy_ion;
            case LARE_VX: return &d.vx;
            case LARE_VY: return &d.vy;
            case LARE_VZ: return &d.vz;
            case LARE_BX: return &d.bx;
            case LARE_BY: return &d.by;
            case LARE_BZ: return &d.bz;
            case LARE_VX1: return &d.vx1;
            case LARE_VY1: return &d.vy1;
            case LARE_VZ1: return &d.vz1;
            default: return &d.dm;
            }
        }

        // harness array <-> contiguous buffer in the reference layout
        void toBuffer(volumeArray &a, std::vector<double> &buf)
        {
            auto r0 = a.getRange(0), r1 = a.getRange(1), r2 = a.getRange(2);
            buf.resize((size_t)(r0.upper_bound - r0.lower_bound + 1) * (r1.upper_bound - r1.lower_bound + 1) *
                       (r2.upper_bound - r2.lower_bound + 1));
            size_t p = 0;
            for (long i = r0.lower_bound; i <= r0.upper_bound; ++i)
                for (long j = r1.lower_bound; j <= r1.upper_bound; ++j)
                    for (long k = r2.lower_bound; k <= r2.upper_bound; ++k)
                        buf[p++] = a(i, j, k);
        }

        void fromBuffer(volumeArray &a, const std::vector<double> &buf)
        {
            auto r0 = a.getRange(0), r1 = a.getRange(1), r2 = a.getRange(2);
            size_t p = 0;
            for (long i = r0.lower_bound; i <= r0.upper_bound; ++i)
                for (long j = r1.lower_bound; j <= r1.upper_bound; ++j)
                    for (long k = r2.lower_bound; k <= r2.upper_bound; ++k)
                        a(i, j, k) = buf[p++];
        }

        // Identify the built-in boundary condition on one edge of dimension
        // dim of a variable by its effect on a marked array.  Edges are
        // applied one at a time: the upper periodic fill of a staggered
        // variable reads index 0, which the lower fill writes.
        struct edgeBC
        {
            int field, dim, edge, kind;
            double value;
            int32_t src[3]; // layer table of a map edge
        };

        edgeBC probeEdge(SAMS::harness &h, simulationData &d, int f, int dim, int edge)
        {
            volumeArray &a = *stateArray(d, f);
            const long n = dim == 0 ? d.nx : dim == 1 ? d.ny : d.nz;
            const int s = a.getRange(dim).lower_bound == -2 ? 1 : 0; // HALF_CELL extent starts at -2
            // marks: exact values that depend only on the index along dim,
            // quadratic so that no reflection, shift or copy of them equals
            // the marks themselves (a linear mark is its own mirror image
            // about a point)
            auto mark = [](long g) {
                const double x = static_cast<double>(g);
                return 4096.0 + 16.0 * x + x * x + 0.5;
            };
            auto r0 = a.getRange(0), r1 = a.getRange(1), r2 = a.getRange(2);
            for (long i = r0.lower_bound; i <= r0.upper_bound; ++i)
                for (long j = r1.lower_bound; j <= r1.upper_bound; ++j)
                    for (long k = r2.lower_bound; k <= r2.upper_bound; ++k)
                    {
                        const long idx[3] = {i, j, k};
                        a(i, j, k) = mark(idx[dim]);
                    }
            h.variableRegistry.applyBoundaryConditions(kVarNames[f], dim,
                                                       edge ? SAMS::domain::edges::upper : SAMS::domain::edges::lower);
            // one line through the middle of the other two dimensions
            auto at = [&](long g) {
                long idx[3] = {d.nx / 2, d.ny / 2, d.nz / 2};
                idx[dim] = g;
                return a(idx[0], idx[1], idx[2]);
            };
            // ghost layers of this edge (variableDef.h:539-605 with the
            // dimension's ghost adjustment for HALF_CELL)
            const long g0 = edge ? n + 1 - s : (s ? -2 : -1), g1 = edge ? n + 2 : 0;
            const long first = edge ? g0 : g1;                // ghost nearest the domain
            const long last = edge ? n : (s ? 0 : 1);         // domain point of the edge
            auto all = [&](auto expect) {
                for (long g = g0; g <= g1; ++g)
                    if (at(g) != expect(g))
                        return false;
                return true;
            };
            // a value equal to some layer's mark is a copy of that layer, not a constant
            auto isMark = [&](double v) {
                for (long t = s ? -2 : -1; t <= n + 2; ++t)
                    if (v == mark(t))
                        return true;
                return false;
            };
            // the condition must act alike on every line of the face: the
            // marks do not vary across it, so every line's ghost layers must
            // equal the middle line's
            {
                const int o1 = dim == 0 ? 1 : 0, o2 = dim == 2 ? 1 : 2;
                const auto ra = a.getRange(o1), rb = a.getRange(o2);
                for (long p = ra.lower_bound; p <= ra.upper_bound; ++p)
                    for (long q = rb.lower_bound; q <= rb.upper_bound; ++q)
                        for (long g = g0; g <= g1; ++g)
                        {
                            long idx[3];
                            idx[dim] = g;
                            idx[o1] = p;
                            idx[o2] = q;
                            const double v = a(idx[0], idx[1], idx[2]), w = at(g);
                            if (!(v == w || (v != v && w != w)))
                                throw std::runtime_error(
                                    std::string("LARE3D (MI355X): the boundary condition of ") + kVarNames[f] +
                                    " on the " + (edge ? "upper" : "lower") + " edge of dimension " +
                                    std::to_string(dim) +
                                    " varies across the face; conditions that act alike on every line of a face "
                                    "are implemented");
                        }
            }
            int kind;
            double value = 0.0;
            if (all([&](long g) { return mark(g); }))
                kind = LARE_BC_NONE;
            else if (all([&](long g) { return mark(edge ? g - n : g + n); }))
                kind = LARE_BC_PERIODIC;
            else if (all([&](long) { return mark(last); }))
                kind = LARE_BC_ZERO_GRADIENT;
            else if (all([&](long g) { return mark(last) - (mark(last - (g - first)) - mark(last)); }))
                kind = LARE_BC_MIRROR;
            else if (all([&](long) { return at(g0); }) && !isMark(at(g0)))
            {
                kind = LARE_BC_CLAMP;
                value = at(g0);
            }
            else
            {
                // a LARE-style condition (LAREBoundaryConditions,
                // src/InitialConditions/LARE/lareBoundaryClass.h:10-34): each
                // ghost layer a copy of one layer of the line, a constant, or
                // left as it was; the marks name the source layer
                edgeBC b{f, dim, edge, LARE_BC_MAP, 0.0, {LARE_BC_MAP_KEEP, LARE_BC_MAP_KEEP, LARE_BC_MAP_KEEP}};
                bool haveValue = false;
                for (long g = g0; g <= g1; ++g)
                {
                    const double v = at(g);
                    int32_t src = LARE_BC_MAP_VALUE;
                    if (v == mark(g))
                        src = LARE_BC_MAP_KEEP;
                    else
                        for (long t = s ? -2 : -1; t <= n + 2; ++t)
                            if (v == mark(t))
                                src = static_cast<int32_t>(t);
                    if (src == LARE_BC_MAP_VALUE)
                    {
                        if (haveValue && v != b.value)
                            throw std::runtime_error(std::string("LARE3D (MI355X): the boundary condition of ") +
                                                     kVarNames[f] + " on the " + (edge ? "upper" : "lower") +
                                                     " edge of dimension " + std::to_string(dim) +
                                                     " sets two ghost layers to different constants");
                        b.value = v;
                        haveValue = true;
                    }
                    b.src[g - g0] = src;
                }
                return b;
            }
            return edgeBC{f, dim, edge, kind, value, {0, 0, 0}};
        }

        bool sameSimulation(const deviceGrid &g, simulationData &d)
        {
            if (g.nx != d.nx || g.ny != d.ny || g.nz != d.nz)
                return false;
            for (int f = 0; f < LARE_NFIELDS; ++f)
                if (g.arrays[f] != static_cast<const void *>(stateArray(d, f)->data()))
                    return false;
            return true;
        }

        // The package-private host arrays of setup.cpp:170-252 that only the
        // step reads: the device computes its own, so their host copies go.
        void releaseHostWorkArrays(LARE3D &pkg, simulationData &d)
        {
            auto &m = pkg.manager;
            m.deallocate(d.p_visc);
            m.deallocate(d.eta);
            m.deallocate(d.dxab);
            m.deallocate(d.dyab);
            m.deallocate(d.dzab);
            m.deallocate(d.dxac);
            m.deallocate(d.dyac);
            m.deallocate(d.dzac);
            m.deallocate(d.cv);
            m.deallocate(d.cv1);
            m.deallocate(d.cvc);
            m.deallocate(d.bx1);
            m.deallocate(d.by1);
            m.deallocate(d.bz1);
            m.deallocate(d.alpha1);
            m.deallocate(d.alpha2);
            m.deallocate(d.alpha3);
            m.deallocate(d.visc_heat);
            m.deallocate(d.pressure);
            m.deallocate(d.p_e);
            m.deallocate(d.p_i);
            m.deallocate(d.rho_v);
            m.deallocate(d.cv_v);
            m.deallocate(d.fx);
            m.deallocate(d.fy);
            m.deallocate(d.fz);
            m.deallocate(d.fx_visc);
            m.deallocate(d.fy_visc);
            m.deallocate(d.fz_visc);
            m.deallocate(d.flux_x);
            m.deallocate(d.flux_y);
            m.deallocate(d.flux_z);
            m.deallocate(d.curlb);
            m.deallocate(d.x);
            m.deallocate(d.y);
            m.deallocate(d.z);
            m.deallocate(d.xp);
            m.deallocate(d.yp);
            m.deallocate(d.zp);
            if (d.rke)
                m.deallocate(d.delta_ke);
        }

        // Arm the host-write guard over the registered variables' arrays.
        void armGuard(deviceGrid &g, simulationData &d)
        {
            installGuardHandler();
            const uintptr_t page = static_cast<uintptr_t>(sysconf(_SC_PAGESIZE));
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                g.region[f] = -1;
                volumeArray &a = *stateArray(d, f);
                const uintptr_t b = reinterpret_cast<uintptr_t>(a.data());
                const uintptr_t e = b + a.getElements() * sizeof(T_dataType);
                const uintptr_t lo = (b + page - 1) / page * page, hi = e / page * page;
                if (hi <= lo)
                    continue;  // an array within one page is not guarded
                for (int i = 0; i < kMaxRegions; ++i)
                    if (!g_regions[i].used)
                    {
                        g_regions[i].lo = reinterpret_cast<char *>(lo);
                        g_regions[i].hi = reinterpret_cast<char *>(hi);
                        g_regions[i].hit = 0;
                        g_regions[i].used = 1;
                        g.region[f] = i;
                        break;
                    }
            }
            g.guarded = true;
            reguard(g);
        }

        deviceGrid &bind(LARE3D &pkg, SAMS::harness &h, simulationData &d)
        {
            deviceGrid &g = grids()[&d];
            if (g.ctx && sameSimulation(g, d))
                return g;
            releaseGrid(g);  // a new simulation at this address
            static bool atexitRegistered = false;
            if (!atexitRegistered)
            {
                std::atexit(releaseAll);
                atexitRegistered = true;
            }
            lare_desc desc{};
            desc.nx = d.nx;
            desc.ny = d.ny;
            desc.nz = d.nz;
            desc.x_min = d.x_min;
            desc.x_max = d.x_max;
            desc.y_min = d.y_min;
            desc.y_max = d.y_max;
            desc.z_min = d.z_min;
            desc.z_max = d.z_max;
            desc.gas_gamma = d.gas_gamma;
            desc.visc1 = d.visc1;
            desc.visc2 = d.visc2;
            desc.mu0 = d.mu0;
            desc.dt_multiplier = d.dt_multiplier;
            desc.resistive = d.resistiveMHD ? 1 : 0;
            desc.precision = LARE_FP64;
            if (const char *pr = std::getenv("LARE3D_HIP_PRECISION"))
            {
                if (std::strcmp(pr, "fp32") == 0)
                    desc.precision = LARE_FP32_STORAGE;
                else if (std::strcmp(pr, "fp64") != 0)
                    throw std::runtime_error("LARE3D (MI355X): LARE3D_HIP_PRECISION must be fp64 or fp32");
            }
            desc.eta_background = d.eta_background;
            desc.eta0 = d.eta0;
            desc.j_max = d.j_max;
            desc.rke = d.rke ? 1 : 0;
            desc.geometry = d.geometry == geometryType::Spherical     ? LARE_SPHERICAL
                            : d.geometry == geometryType::Cylindrical ? LARE_CYLINDRICAL
                                                                      : LARE_CARTESIAN;
            // boundary conditions first (host only), so that a condition the
            // device does not implement is reported before any device work
            std::vector<edgeBC> bcs;
            std::vector<double> buf;
            const bool report = std::getenv("LARE3D_HIP_REPORT_BCS") != nullptr;
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                toBuffer(*stateArray(d, f), buf);
                for (int dim = 0; dim < 3; ++dim)
                    for (int edge = 0; edge < 2; ++edge)
                    {
                        bcs.push_back(probeEdge(h, d, f, dim, edge));
                        if (report)
                            std::fprintf(stderr, "LARE3D (MI355X) bc %s %d %d %d %.17g %d %d %d\n", kVarNames[f], dim,
                                         edge, bcs.back().kind, bcs.back().value, bcs.back().src[0],
                                         bcs.back().src[1], bcs.back().src[2]);
                    }
                fromBuffer(*stateArray(d, f), buf); // the probes overwrote the host copy
            }
            check(lare_ctx_create(&desc, &g.ctx), "lare_ctx_create");
            for (const edgeBC &b : bcs)
                if (b.kind == LARE_BC_MAP)
                    check(lare_set_bc_map(g.ctx, b.field, b.dim, b.edge, b.src, b.value), "lare_set_bc_map");
                else
                    check(lare_set_bc(g.ctx, b.field, b.dim, b.edge, b.kind, b.value), "lare_set_bc");
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                toBuffer(*stateArray(d, f), buf);
                check(lare_upload_field(g.ctx, f, buf.data()), "lare_upload_field");
            }
            check(lare_set_dt(g.ctx, d.dt), "lare_set_dt");
            g.devDt = d.dt;
            g.hostCurrent = true;
            g.nx = d.nx;
            g.ny = d.ny;
            g.nz = d.nz;
            for (int f = 0; f < LARE_NFIELDS; ++f)
                g.arrays[f] = stateArray(d, f)->data();
            const char *sync = std::getenv("LARE3D_HIP_HOST_SYNC");
            g.hostSync = sync && std::strcmp(sync, "0") != 0;
            const char *keep = std::getenv("LARE3D_HIP_KEEP_HOST_ARRAYS");
            if (!(keep && std::strcmp(keep, "0") != 0))
                releaseHostWorkArrays(pkg, d);
            const char *guard = std::getenv("LARE3D_HIP_WRITE_GUARD");
            if (!g.hostSync && !(guard && std::strcmp(guard, "0") == 0))
                armGuard(g, d);
            return g;
        }

        // Bring the harness arrays up to date (before output).
        void download(simulationData &d)
        {
            auto it = grids().find(&d);
            if (it == grids().end() || it->second.hostCurrent)
                return;
            std::vector<double> buf;
            unguard(it->second);
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                volumeArray &a = *stateArray(d, f);
                toBuffer(a, buf); // sizes buf
                check(lare_download_field(it->second.ctx, f, buf.data()), "lare_download_field");
                fromBuffer(a, buf);
            }
            reguard(it->second);
            it->second.hostCurrent = true;
        }

        // Host-synchronised mode: before an event, the harness arrays (which
        // other packages may have written since the last event) replace the
        // device state; after it, the device state replaces them.
        void beginEvent(deviceGrid &g, simulationData &d)
        {
            std::vector<double> buf;
            if (g.guarded)
            {
                // arrays written on the host since the last event
                for (int f = 0; f < LARE_NFIELDS; ++f)
                {
                    const int i = g.region[f];
                    if (i < 0 || !g_regions[i].hit)
                        continue;
                    if (!g.hostCurrent)
                        throw std::runtime_error(
                            std::string("LARE3D (MI355X): ") + kVarNames[f] +
                            " was written on the host between LARE3D events while the device held a newer "
                            "state; run with LARE3D_HIP_HOST_SYNC=1 for packages that write the state between "
                            "events");
                    toBuffer(*stateArray(d, f), buf);
                    check(lare_upload_field(g.ctx, f, buf.data()), "lare_upload_field");
                }
                reguard(g);
            }
            if (!g.hostSync || !g.hostCurrent)
                return;
            for (int f = 0; f < LARE_NFIELDS; ++f)
            {
                toBuffer(*stateArray(d, f), buf);
                check(lare_upload_field(g.ctx, f, buf.data()), "lare_upload_field");
            }
        }
        void endEvent(deviceGrid &g, simulationData &d)
        {
            g.hostCurrent = false;
            if (g.hostSync)
                download(d);
        }
    } // namespace

    void LARE3D::lagrangian_step(simulationData &data, SAMS::controlFunctions &controlFns)
    {
        deviceGrid &g = bind(*this, harness, data);
        beginEvent(g, data);
        // the shock viscosity reads the previous dt (lagran.cpp:72): the
        // runner's, which another package may have lowered
        if (data.dt != g.devDt)
            check(lare_set_dt(g.ctx, data.dt), "lare_set_dt");
        check(lare_viscosity_dt(g.ctx, &g.dtNew, &g.dtrNew), "lare_viscosity_dt");
        controlFns.calculateTimestep();
        check(lare_lagrange(g.ctx, data.dt), "lare_lagrange");
        g.devDt = data.dt;
        endEvent(g, data);
    }

    void LARE3D::set_dt(simulationData &data)
    {
        auto it = grids().find(&data);
        if (it == grids().end())
            throw std::runtime_error("LARE3D (MI355X): set_dt outside lagrangian_step");
        data.dt = it->second.dtNew;
        data.dtr = it->second.dtrNew;
    }

    void LARE3D::predictor_step(simulationData &)
    {
        throw std::runtime_error("LARE3D (MI355X): predictor_step runs inside lagrangian_step");
    }

    void LARE3D::corrector_step(simulationData &data)
    {
        deviceGrid &g = bind(*this, harness, data);
        beginEvent(g, data);
        check(lare_half_timestep(g.ctx), "lare_half_timestep");
        endEvent(g, data);
    }

    void LARE3D::eta_calc(simulationData &)
    {
        // eta is recomputed on the device at the start of every resistive
        // substep, before any read (lagran.cpp:163-167)
    }

    void LARE3D::eulerian_remap(simulationData &data, remapData &)
    {
        deviceGrid &g = bind(*this, harness, data);
        beginEvent(g, data);
        check(lare_end_of_timestep(g.ctx), "lare_end_of_timestep");
        endEvent(g, data);
    }

    void LARE3D::remap_x(simulationData &, remapData &)
    {
        throw std::runtime_error("LARE3D (MI355X): remap_x runs inside eulerian_remap");
    }
    void LARE3D::remap_y(simulationData &, remapData &)
    {
        throw std::runtime_error("LARE3D (MI355X): remap_y runs inside eulerian_remap");
    }
    void LARE3D::remap_z(simulationData &, remapData &)
    {
        throw std::runtime_error("LARE3D (MI355X): remap_z runs inside eulerian_remap");
    }

    // endOfTimestep calls this after eulerian_remap when rke is set
    // (shared_data.h:415-417); lare_end_of_timestep has already applied the
    // correction on the device, after the remap's last boundary fill
    void LARE3D::energy_correction(simulationData &)
    {
    }

    // ---- output events (diagnostics.cpp:50-107): the same datasets, read
    // from the harness arrays after the device state is downloaded

    template <typename T_writer>
    void LARE3D::registerOutputMeshes(writer<T_writer> &w, simulationData &data)
    {
        w.template registerRectilinearMesh<T_dataType>("MeshCC", data.nx, data.ny, data.nz);
    }

    template <typename T_writer>
    void LARE3D::registerOutputVariables(writer<T_writer> &w, simulationData &)
    {
        for (const char *n : {"rho", "energy_electron", "energy_ion", "vx", "vy", "vz", "bx", "by", "bz"})
            w.template registerData<T_dataType>(n, "MeshCC");
    }

    template <typename T_writer>
    void LARE3D::writeOutputMeshes(writer<T_writer> &w, simulationData &data)
    {
        w.writeRectilinearMesh("MeshCC", &data.xc_host(1), &data.yc_host(1), &data.zc_host(1));
    }

    template <typename T_writer>
    void LARE3D::writeOutputVariables(writer<T_writer> &w, simulationData &data)
    {
        download(data);
        const char *names[9] = {"rho", "energy_electron", "energy_ion", "vx", "vy", "vz", "bx", "by", "bz"};
        std::vector<double> interior((size_t)data.nx * data.ny * data.nz);
        for (int f = 0; f < 9; ++f)
        {
            volumeArray &a = *stateArray(data, f);
            size_t p = 0;
            for (long i = 1; i <= data.nx; ++i)
                for (long j = 1; j <= data.ny; ++j)
                    for (long k = 1; k <= data.nz; ++k)
                        interior[p++] = a(i, j, k);
            w.writeData(names[f], interior.data());
        }
    }

#if defined(USE_HDF5)
    template void LARE3D::registerOutputMeshes<HDF5File>(writer<HDF5File> &, simulationData &);
    template void LARE3D::registerOutputVariables<HDF5File>(writer<HDF5File> &, simulationData &);
    template void LARE3D::writeOutputMeshes<HDF5File>(writer<HDF5File> &, simulationData &);
    template void LARE3D::writeOutputVariables<HDF5File>(writer<HDF5File> &, simulationData &);
#else
    template void LARE3D::registerOutputMeshes<simpleFile>(writer<simpleFile> &, simulationData &);
    template void LARE3D::registerOutputVariables<simpleFile>(writer<simpleFile> &, simulationData &);
    template void LARE3D::writeOutputMeshes<simpleFile>(writer<simpleFile> &, simulationData &);
    template void LARE3D::writeOutputVariables<simpleFile>(writer<simpleFile> &, simulationData &);
#endif

    // Access for drivers that need the harness arrays current outside an
    // output event (e.g. to compare states).
    void lare3d_hip_download(simulationData &data) { download(data); }

    // Release the device context of a simulationData (its state is lost;
    // download first to keep it).  Contexts are also released at process
    // exit and when a new simulation reuses the address.
    void lare3d_hip_release(simulationData &data)
    {
        auto it = grids().find(&data);
        if (it == grids().end())
            return;
        releaseGrid(it->second);
        grids().erase(it);
    }
} // namespace LARE
