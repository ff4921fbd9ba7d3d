// Reference driver for the LARE3D oracle build (test infrastructure only).
//
// This file is compiled together with the UNMODIFIED reference LARE sources
// (src/PhysicsPackages/LARE/*.cpp) and headers, where they lie, by
// oracle/Makefile; the binary goes to oracle/_ref/ and is never shipped as
// product.  It drives the reference's own runner (include/runner/runner.h:
// initializePackages 548-562, runPackages 564-586) with LARE::LARE3D plus one
// problem package defined here, and dumps the reference's state arrays (full
// extents, ghosts included) and the dt sequence so that the C restatement in
// oracle/lare_oracle.c can be pinned bit for bit.
//
// The problem package follows the shape of the reference's IC packages
// (src/InitialConditions/OrszagTang3D/OrszagTang3D.cpp:28-158,
//  src/InitialConditions/BrioAndWu/BrioAndWu.cpp:29-155): it fills every array
// over rho.getRange() (the centred extent) and attaches built-in boundary
// conditions per variable and dimension.
//
// Usage:
//   ref_driver --problem ot|bw --nx N [--ny N --nz N] --steps K
//              [--dump s1,s2,...] [--out DIR] [--seed S] [--amp A] [--beta B]
//              [--noise F] [--eta E] [--eta0 E0] [--j-max J]
//              [--bcx periodic|zerograd|mirror|clamp] [--time] [--rke] [--lare-bc]
//              [--geometry cartesian|cylindrical|spherical] [--x0 X] [--y0 Y]
//              [--face-varying] [--then-n N]
// --then-n runs a second simulation (the same problem on an N^3 grid) in the
// same process after the first, into OUT/second.
// --eta / --eta0 (either > 0) turn resistive MHD on (shared_data.h:113-117):
// eta_background = E, and eta0 is added where the current exceeds j_max
// (lagran.cpp:473).  The reference's resistive path runs only in the
// ref_driver_res build (oracle/Makefile).  --geometry selects the
// coordinate system (shared_data.h:38-43); --x0 / --y0 shift the x and y
// extents of the problem's domain (a curvilinear grid needs r > 0, and
// theta away from the poles); Brio-Wu's discontinuity stays at the x
// midpoint.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <filesystem>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "harness.h"
#include "runner.h"
#include "shared_data.h"
#include "builtInBoundaryConditions.h"
#include "lareic.h"

#if defined(LARE_DROPIN)
// Built with the MI355X package (sams_amd/host/lare3d_hip.cpp) in place of
// the reference's lagran/remap/x,y,zremap/diagnostics.cpp: the state lives on
// the device and is downloaded into the harness arrays before a dump.
namespace LARE
{
    void lare3d_hip_download(simulationData &data);
}
#endif

namespace refdrv
{
    namespace pw = portableWrapper;

    enum class bcKind { periodic, zerograd, mirror, clamp };

    struct config
    {
        std::string problem = "ot";
        long nx = 16, ny = -1, nz = -1;
        long steps = 10;
        std::set<long> dumps;
        std::string out = ".";
        uint64_t seed = 1;
        double amp = 0.0;   // OT: Gaussian velocity perturbation amplitude
        double beta = -1.0; // OT: plasma beta (<=0: classic 5/(12 pi) pressure)
        double noise = 0.0; // BW: relative uniform density noise
        double eta = 0.0;   // background resistivity (>0 turns resistive MHD on)
        double eta0 = 0.0;  // resistivity added where |j| > j_max (>0 turns resistive MHD on)
        double jMax = 1.0e30;
        bcKind bcx = bcKind::periodic;
        bool bcxSet = false;
        bool time = false;
        bool phases = false;
        bool rke = false;   // remap kinetic-energy correction
        bool lareBC = false; // LARE3DInitialConditions' BC_OTHER conditions on every edge
        bool faceVarying = false; // rho's x edges take halfFaceBC
        long thenN = 0;           // > 0: a second simulation of thenN^3 cells in the same process
        LARE::geometryType geometry = LARE::geometryType::Cartesian;
        double x0 = 0.0, y0 = 0.0; // shifts of the domain's x and y extents
        std::vector<double> dts;
        double elapsed = 0.0;
        std::chrono::steady_clock::time_point t0, tLast;
    };

    config &cfg()
    {
        static config c;
        return c;
    }

    // splitmix64 finaliser: counter-based stream for seeded inputs.
    static inline uint64_t sm64(uint64_t x)
    {
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    }
    static inline double uniform01(uint64_t seed, uint64_t stream, uint64_t lin)
    {
        uint64_t h = sm64(sm64(seed + 0x632BE59BD9B4E019ull * stream) ^ lin);
        return static_cast<double>(h >> 11) * 0x1.0p-53;
    }
    static inline double gauss(uint64_t seed, uint64_t comp, uint64_t lin)
    {
        double u1 = uniform01(seed, 16 + 2 * comp, lin);
        double u2 = uniform01(seed, 17 + 2 * comp, lin);
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
    }

    template <typename A>
    void dumpArray(const std::string &path, A &arr)
    {
        std::ofstream f(path, std::ios::binary);
        auto r0 = arr.getRange(0), r1 = arr.getRange(1), r2 = arr.getRange(2);
        for (long i = r0.lower_bound; i <= r0.upper_bound; ++i)
            for (long j = r1.lower_bound; j <= r1.upper_bound; ++j)
                for (long k = r2.lower_bound; k <= r2.upper_bound; ++k)
                {
                    double v = arr(i, j, k);
                    f.write(reinterpret_cast<const char *>(&v), sizeof(double));
                }
    }

    void dumpExtra(LARE::simulationData &d, const std::string &p)
    {
        dumpArray(p + "vx1.f64", d.vx1);
        dumpArray(p + "vy1.f64", d.vy1);
        dumpArray(p + "vz1.f64", d.vz1);
        dumpArray(p + "cv1.f64", d.cv1);
        dumpArray(p + "p_visc.f64", d.p_visc);
        dumpArray(p + "visc_heat.f64", d.visc_heat);
        dumpArray(p + "pressure.f64", d.pressure);
        dumpArray(p + "fx.f64", d.fx);
        dumpArray(p + "fy.f64", d.fy);
        dumpArray(p + "fz.f64", d.fz);
        dumpArray(p + "fx_visc.f64", d.fx_visc);
        dumpArray(p + "alpha1.f64", d.alpha1);
        dumpArray(p + "alpha2.f64", d.alpha2);
        dumpArray(p + "alpha3.f64", d.alpha3);
        dumpArray(p + "bx1.f64", d.bx1);
        dumpArray(p + "rho_v.f64", d.rho_v);
        dumpArray(p + "dm.f64", d.dm);
    }

    void dumpState(LARE::simulationData &d, long step)
    {
#if defined(LARE_DROPIN)
        LARE::lare3d_hip_download(d);
#endif
        char pre[64];
        std::snprintf(pre, sizeof(pre), "/s%05ld_", step);
        std::string p = cfg().out + pre;
        dumpArray(p + "rho.f64", d.rho);
        dumpArray(p + "energy_electron.f64", d.energy_electron);
        dumpArray(p + "energy_ion.f64", d.energy_ion);
        dumpArray(p + "vx.f64", d.vx);
        dumpArray(p + "vy.f64", d.vy);
        dumpArray(p + "vz.f64", d.vz);
        dumpArray(p + "bx.f64", d.bx);
        dumpArray(p + "by.f64", d.by);
        dumpArray(p + "bz.f64", d.bz);
    }

    // A condition that acts differently on different lines of a face (for
    // the drop-in's check that rejects such conditions): ghost cells of the
    // lower half of the next dimension set to 1, the rest to 2.
    class halfFaceBC : public SAMS::singleVariableBC<SAMS::T_dataType, 3>
    {
    public:
        halfFaceBC(const SAMS::variableDef &vd) : SAMS::singleVariableBC<SAMS::T_dataType, 3>(vd) {}
        void apply(int dimension, SAMS::domain::edges edge) override
        {
            const int e = edge == SAMS::domain::edges::lower ? 0 : 1;
            auto r = this->boundaryRanges[dimension * 2 + e];
            auto r0 = std::get<0>(r), r1 = std::get<1>(r), r2 = std::get<2>(r);
            const int other = dimension == 1 ? 0 : 1;
            auto ro = other == 0 ? r0 : r1;
            const long mid = (ro.lower_bound + ro.upper_bound) / 2;
            for (long i = r0.lower_bound; i <= r0.upper_bound; ++i)
                for (long j = r1.lower_bound; j <= r1.upper_bound; ++j)
                    for (long k = r2.lower_bound; k <= r2.upper_bound; ++k)
                        this->variable(i, j, k) = ((other == 0 ? i : j) <= mid) ? 1.0 : 2.0;
        }
    };

    class RefProblem
    {
        void attach(const std::string &var, SAMS::harness &h, int dim, bcKind kind)
        {
            SAMS::variableDef &vd = h.variableRegistry.getVariable(var);
            std::shared_ptr<SAMS::boundaryConditions> bc;
            switch (kind)
            {
            case bcKind::periodic:
                bc = std::make_shared<SAMS::simplePeriodicBC<SAMS::T_dataType, 3>>(vd);
                break;
            case bcKind::zerograd:
                bc = std::make_shared<SAMS::simpleZeroGradientBC<SAMS::T_dataType, 3>>(vd);
                break;
            case bcKind::mirror:
                bc = std::make_shared<SAMS::simpleMirrorBC<SAMS::T_dataType, 3>>(vd);
                break;
            case bcKind::clamp:
                bc = std::make_shared<SAMS::simpleClamp<SAMS::T_dataType, 3>>(vd, 0.0);
                break;
            }
            vd.addBoundaryCondition(dim, bc);
        }

    public:
        constexpr static std::string_view name = "RefProblem";
        constexpr static bool timeSimulation = false;

        void controlVariables(LARE::simulationData &data)
        {
            auto &c = cfg();
            data.nx = c.nx;
            data.ny = c.ny;
            data.nz = c.nz;
            data.t_end = 1.0e30;
            data.dt_snapshots = 1.0e30;
            data.dt_multiplier = 0.8;
            data.geometry = c.geometry;
            data.visc1 = 0.1;
            data.visc2 = 1.0;
            data.mf = 1.2;
            data.mu0 = 1.0;
            data.rke = c.rke;
            if (c.lareBC)
            {
                data.xbc_min = data.xbc_max = LARE::BCType::BC_OTHER;
                data.ybc_min = data.ybc_max = LARE::BCType::BC_OTHER;
                data.zbc_min = data.zbc_max = LARE::BCType::BC_OTHER;
            }
            data.resistiveMHD = false;
            data.eta_background = 0.0;
            data.j_max = 1.0e30;
            data.eta0 = 0.0;
            if (c.problem == "ot")
            {
                data.x_min = 0.0;
                data.x_max = 1.0;
                data.y_min = 0.0;
                data.y_max = 1.0;
                data.z_min = 0.0;
                data.z_max = 1.0;
                data.gas_gamma = 5.0 / 3.0;
            }
            else
            {
                data.x_min = -1.0;
                data.x_max = 1.0;
                data.y_min = 0.0;
                data.y_max = (data.x_max - data.x_min) * static_cast<double>(c.ny) / static_cast<double>(c.nx);
                data.z_min = 0.0;
                data.z_max = (data.x_max - data.x_min) * static_cast<double>(c.nz) / static_cast<double>(c.nx);
                data.gas_gamma = 2.0;
            }
            data.x_min += c.x0;
            data.x_max += c.x0;
            data.y_min += c.y0;
            data.y_max += c.y0;
            if (c.eta > 0.0 || c.eta0 > 0.0)
            {
                data.resistiveMHD = true;
                data.eta_background = c.eta;
                data.eta0 = c.eta0;
                data.j_max = c.jMax;
            }
        }

        void setDomain(SAMS::harness &harness, LARE::simulationData &data)
        {
            auto &ax = harness.axisRegistry;
            ax.setDomain("X", data.nx, data.x_min, data.x_max);
            ax.setDomain("Y", data.ny, data.y_min, data.y_max);
            ax.setDomain("Z", data.nz, data.z_min, data.z_max);
        }

        void setBoundaryConditions(SAMS::harness &harness, LARE::simulationData &data, SAMS::timeState &time)
        {
            if (cfg().lareBC)
            {
                // the LARE-style conditions of the LARE initial-condition
                // package (src/InitialConditions/LARE/lareic.h:124-146,
                // boundary.cpp), attached exactly as that package does
                LARE::LARE3DInitialConditions lare;
                lare.setBoundaryConditions(harness, data, time);
                return;
            }
            const char *vars[] = {"bx", "by", "bz", "energy_ion", "energy_electron", "rho",
                                  "vx", "vy", "vz", "LARE/vx1", "LARE/vy1", "LARE/vz1", "LARE/dm"};
            for (const char *v : vars)
            {
                // "clamp" walls: velocities clamped to zero, everything else
                // zero-gradient (a clamp of density or energy to zero is not
                // a physical state).
                bcKind kx = cfg().bcx;
                if (kx == bcKind::clamp && v[0] != 'v' && std::strncmp(v, "LARE/v", 6) != 0)
                    kx = bcKind::zerograd;
                if (cfg().faceVarying && std::strcmp(v, "rho") == 0)
                {
                    SAMS::variableDef &vd = harness.variableRegistry.getVariable(v);
                    vd.addBoundaryCondition(0, std::make_shared<halfFaceBC>(vd));
                }
                else
                    attach(v, harness, 0, kx);
                attach(v, harness, 1, bcKind::periodic);
                attach(v, harness, 2, bcKind::periodic);
            }
        }

        void initialConditions(SAMS::harness &harnessRef, LARE::simulationData &data)
        {
            auto &c = cfg();
            pw::portableArray<SAMS::T_dataType, 3> rho, ee, ei, vx, vy, vz, bx, by, bz;
            pw::portableArray<SAMS::T_dataType, 1> xc, yc, zc, xb, yb, zb;
            auto &axr = harnessRef.axisRegistry;
            axr.fillPPLocalAxis("X", xc, SAMS::staggerType::CENTRED);
            axr.fillPPLocalAxis("Y", yc, SAMS::staggerType::CENTRED);
            axr.fillPPLocalAxis("Z", zc, SAMS::staggerType::CENTRED);
            axr.fillPPLocalAxis("X", xb, SAMS::staggerType::HALF_CELL);
            axr.fillPPLocalAxis("Y", yb, SAMS::staggerType::HALF_CELL);
            axr.fillPPLocalAxis("Z", zb, SAMS::staggerType::HALF_CELL);
            auto &vr = harnessRef.variableRegistry;
            vr.fillPPArray("rho", rho);
            vr.fillPPArray("energy_electron", ee);
            vr.fillPPArray("energy_ion", ei);
            vr.fillPPArray("vx", vx);
            vr.fillPPArray("vy", vy);
            vr.fillPPArray("vz", vz);
            vr.fillPPArray("bx", bx);
            vr.fillPPArray("by", by);
            vr.fillPPArray("bz", bz);
            const long ny4 = data.ny + 4, nz4 = data.nz + 4;
            const double gam = data.gas_gamma;
            auto r0 = rho.getRange(0), r1 = rho.getRange(1), r2 = rho.getRange(2);
            for (long ix = r0.lower_bound; ix <= r0.upper_bound; ++ix)
                for (long iy = r1.lower_bound; iy <= r1.upper_bound; ++iy)
                    for (long iz = r2.lower_bound; iz <= r2.upper_bound; ++iz)
                    {
                        const uint64_t lin = static_cast<uint64_t>(((ix + 1) * ny4 + (iy + 1)) * nz4 + (iz + 1));
                        if (c.problem == "ot")
                        {
                            const double r = 25.0 / (36.0 * M_PI);
                            double p = 5.0 / (12.0 * M_PI);
                            if (c.beta > 0.0)
                                p = c.beta * (1.0 / (4.0 * M_PI)) / 2.0;
                            const double b0 = std::sqrt(4.0 * M_PI);
                            rho(ix, iy, iz) = r;
                            double ux = -std::sin(2.0 * M_PI * yb(iy));
                            double uy = std::sin(2.0 * M_PI * xb(ix));
                            double uz = 0.0;
                            if (c.amp != 0.0)
                            {
                                ux = ux + c.amp * gauss(c.seed, 0, lin);
                                uy = uy + c.amp * gauss(c.seed, 1, lin);
                                uz = uz + c.amp * gauss(c.seed, 2, lin);
                            }
                            vx(ix, iy, iz) = ux;
                            vy(ix, iy, iz) = uy;
                            vz(ix, iy, iz) = uz;
                            bx(ix, iy, iz) = -std::sin(2.0 * M_PI * yc(iy)) / b0;
                            by(ix, iy, iz) = std::sin(4.0 * M_PI * xc(ix)) / b0;
                            bz(ix, iy, iz) = 0.0;
                            ee(ix, iy, iz) = p / ((gam - 1.0) * r) / 2.0;
                            ei(ix, iy, iz) = p / ((gam - 1.0) * r) / 2.0;
                        }
                        else
                        {
                            double r, p, b;
                            if (xc(ix) < 0.5 * (data.x_min + data.x_max))
                            {
                                r = 1.0;
                                b = 1.0;
                                p = 1.0;
                            }
                            else
                            {
                                r = 0.125;
                                b = -1.0;
                                p = 0.1;
                            }
                            if (c.noise != 0.0)
                                r = r * (1.0 + c.noise * (2.0 * uniform01(c.seed, 3, lin) - 1.0));
                            rho(ix, iy, iz) = r;
                            by(ix, iy, iz) = b;
                            bx(ix, iy, iz) = 0.75;
                            ee(ix, iy, iz) = p / ((gam - 1.0) * r) / 2.0;
                            ei(ix, iy, iz) = p / ((gam - 1.0) * r) / 2.0;
                        }
                    }
            if (c.dumps.count(0))
                dumpState(data, 0);
        }

        // Phase dumps (debugging aid): state and Lagrangian arrays after
        // startOfTimestep and halfTimestep of the steps listed in --dump.
        void startOfTimestep(LARE::simulationData &data, SAMS::timeState &t)
        {
            if (cfg().phases && cfg().dumps.count(static_cast<long>(t.step) + 1))
            {
                char pre[64];
                std::snprintf(pre, sizeof(pre), "/p%05ld_a_", static_cast<long>(t.step) + 1);
                dumpExtra(data, cfg().out + pre);
                dumpState(data, -1000 - static_cast<long>(t.step));
            }
        }

        void halfTimestep(LARE::simulationData &data, SAMS::timeState &t)
        {
            if (cfg().phases && cfg().dumps.count(static_cast<long>(t.step) + 1))
            {
                char pre[64];
                std::snprintf(pre, sizeof(pre), "/p%05ld_b_", static_cast<long>(t.step) + 1);
                dumpExtra(data, cfg().out + pre);
                dumpState(data, -2000 - static_cast<long>(t.step));
            }
        }

        void endOfTimestep(LARE::simulationData &data, SAMS::timeState &t)
        {
            auto &c = cfg();
            c.dts.push_back(data.dt);
            c.tLast = std::chrono::steady_clock::now();
            if (c.dumps.count(static_cast<long>(t.step)))
                dumpState(data, static_cast<long>(t.step));
        }

        void queryTerminate(bool &terminate, [[maybe_unused]] LARE::simulationData &data, SAMS::timeState &t)
        {
            terminate = static_cast<long>(t.step) >= cfg().steps;
        }

        void queryOutput(bool &shouldOutput, [[maybe_unused]] LARE::simulationData &data, [[maybe_unused]] SAMS::timeState &t)
        {
            shouldOutput = false;
        }
    };
} // namespace refdrv

int main(int argc, char *argv[])
{
    auto &c = refdrv::cfg();
    for (int i = 1; i < argc; ++i)
    {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc)
            {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--problem")
            c.problem = next();
        else if (a == "--nx")
            c.nx = std::stol(next());
        else if (a == "--ny")
            c.ny = std::stol(next());
        else if (a == "--nz")
            c.nz = std::stol(next());
        else if (a == "--steps")
            c.steps = std::stol(next());
        else if (a == "--out")
            c.out = next();
        else if (a == "--seed")
            c.seed = std::stoull(next());
        else if (a == "--amp")
            c.amp = std::stod(next());
        else if (a == "--beta")
            c.beta = std::stod(next());
        else if (a == "--noise")
            c.noise = std::stod(next());
        else if (a == "--eta")
            c.eta = std::stod(next());
        else if (a == "--geometry")
        {
            std::string gname = next();
            if (gname == "cartesian")
                c.geometry = LARE::geometryType::Cartesian;
            else if (gname == "cylindrical")
                c.geometry = LARE::geometryType::Cylindrical;
            else if (gname == "spherical")
                c.geometry = LARE::geometryType::Spherical;
            else
            {
                std::fprintf(stderr, "unknown geometry %s\n", gname.c_str());
                return 2;
            }
        }
        else if (a == "--x0")
            c.x0 = std::stod(next());
        else if (a == "--y0")
            c.y0 = std::stod(next());
        else if (a == "--eta0")
            c.eta0 = std::stod(next());
        else if (a == "--j-max")
            c.jMax = std::stod(next());
        else if (a == "--time")
            c.time = true;
        else if (a == "--rke")
            c.rke = true;
        else if (a == "--lare-bc")
            c.lareBC = true;
        else if (a == "--face-varying")
            c.faceVarying = true;
        else if (a == "--then-n")
            c.thenN = std::stol(next());
        else if (a == "--phases")
        {
#if defined(LARE_DROPIN)
            std::fprintf(stderr, "--phases: the Lagrangian arrays stay on the device in this build\n");
            std::exit(2);
#endif
            c.phases = true;
        }
        else if (a == "--dump")
        {
            std::stringstream ss(next());
            std::string tok;
            while (std::getline(ss, tok, ','))
                c.dumps.insert(std::stol(tok));
        }
        else if (a == "--bcx")
        {
            std::string b = next();
            c.bcxSet = true;
            if (b == "periodic")
                c.bcx = refdrv::bcKind::periodic;
            else if (b == "zerograd")
                c.bcx = refdrv::bcKind::zerograd;
            else if (b == "mirror")
                c.bcx = refdrv::bcKind::mirror;
            else if (b == "clamp")
                c.bcx = refdrv::bcKind::clamp;
            else
            {
                std::fprintf(stderr, "unknown bc %s\n", b.c_str());
                return 2;
            }
        }
        else
        {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (c.ny < 0)
        c.ny = c.nx;
    if (c.nz < 0)
        c.nz = c.nx;
    if (!c.bcxSet && c.problem == "bw")
        c.bcx = refdrv::bcKind::zerograd;

    int rargc = 1;
    char prog[] = "ref_driver";
    char *rargvBuf[] = {prog, nullptr};
    char **rargv = rargvBuf;
    SAMS::MPI::initialize(rargc, rargv);
    portableWrapper::initialize(rargc, rargv);
    auto runOnce = [&]() {
        SAMS::runner<LARE::LARE3D, refdrv::RefProblem> runner;
        runner.initialize(rargc, rargv);
        runner.activatePackage("LARE3D");
        runner.activatePackage("RefProblem");
        runner.initializePackages();
        // Wall time of the step loop only: from entry into runPackages()
        // (runner.h:564-586) to the end of the last endOfTimestep, so that
        // initialisation and the final writeOutput are excluded.
        c.t0 = std::chrono::steady_clock::now();
        runner.runPackages();
        c.elapsed = std::chrono::duration<double>(c.tLast - c.t0).count();
        runner.finalizePackages();
        runner.finalize();

        std::ofstream m(c.out + "/meta.json");
        m << "{\"problem\": \"" << c.problem << "\", \"nx\": " << c.nx << ", \"ny\": " << c.ny << ", \"nz\": " << c.nz
          << ", \"steps\": " << c.steps << ", \"elapsed_s\": " << c.elapsed << ", \"dt\": [";
        for (size_t i = 0; i < c.dts.size(); ++i)
        {
            char buf[64];
            std::snprintf(buf, sizeof(buf), "\"%a\"", c.dts[i]);
            m << (i ? ", " : "") << buf;
        }
        m << "]}\n";
    };
    runOnce();
    if (c.thenN > 0)
    {
        // a second simulation in the same process, n^3, into OUT/second
        c.nx = c.ny = c.nz = c.thenN;
        c.out += "/second";
        std::filesystem::create_directories(c.out);
        c.dts.clear();
        runOnce();
    }
    portableWrapper::finalize();
    return 0;
}
