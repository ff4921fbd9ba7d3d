// Initial-condition driver for the LARE3D checks (test infrastructure only).
//
// Runs the reference's own initial-condition packages (src/InitialConditions/
// SodShockTube, BrioAndWu, MHDRotor, OrszagTang, OrszagTang3D,
// EmeryWindTunnel, KarmanVortex) exactly as the reference's src/main.cpp:10-44
// does -- one runner over LARE::LARE3D and every package, activated by name --
// plus one observer package that stops the run after --steps steps, records
// the dt sequence and dumps the state arrays (full extents, ghosts included)
// at the steps listed in --dump.  oracle/Makefile compiles it with the
// reference's LARE sources where they lie (oracle/_ref/ic_driver) and, with
// -DLARE_DROPIN, with the MI355X package in their place
// (oracle/_ref/ic_dropin).  Outputs only into --out.
//
// Usage: ic_driver --ic <package name> --steps K [--dump s1,s2,...] --out DIR
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "harness.h"
#include "shared_data.h"
#include "SodShockTube.h"
#include "BrioAndWu.h"
#include "MHDRotor.h"
#include "OrszagTang.h"
#include "OrszagTang3D.h"
#include "EmeryWindTunnel.h"
#include "KarmanVortex.h"
#include "builtInBoundaryConditions.h"
#include "runner.h"
#include "lareic.h"

#if defined(LARE_DROPIN)
namespace LARE
{
    void lare3d_hip_download(simulationData &data);
}
#endif

namespace icdrv
{
    struct config
    {
        std::string ic;
        long steps = 10;
        std::set<long> dumps;
        std::string out = ".";
        std::vector<double> dts;
        bool phases = false;  // also dump after startOfTimestep and halfTimestep of the dumped steps
    };
    config &cfg()
    {
        static config c;
        return c;
    }

    template <class A>
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

    // Last package of the runner: its queryTerminate has the final say
    // (runner.h:171-181 passes one flag through every package in order).
    class ICObserver
    {
    public:
        constexpr static std::string_view name = "ICObserver";
        constexpr static bool timeSimulation = false;

        void startOfTimestep(LARE::simulationData &d, SAMS::timeState &t)
        {
            if (cfg().phases && cfg().dumps.count(static_cast<long>(t.step) + 1))
                dump(d, "/a%05ld_", static_cast<long>(t.step) + 1);
        }
        void halfTimestep(LARE::simulationData &d, SAMS::timeState &t)
        {
            if (cfg().phases && cfg().dumps.count(static_cast<long>(t.step) + 1))
                dump(d, "/b%05ld_", static_cast<long>(t.step) + 1);
        }
        void endOfTimestep(LARE::simulationData &d, SAMS::timeState &t)
        {
            auto &c = cfg();
            c.dts.push_back(d.dt);
            const long s = static_cast<long>(t.step);
            if (c.dumps.count(s))
                dump(d, "/s%05ld_", s);
        }

        void dump(LARE::simulationData &d, const char *fmt, long s)
        {
            auto &c = cfg();
#if defined(LARE_DROPIN)
            LARE::lare3d_hip_download(d);
#endif
            char pre[64];
            std::snprintf(pre, sizeof(pre), fmt, s);
            const std::string p = c.out + pre;
            if (c.phases)
            {
                dumpArray(p + "vx1.f64", d.vx1);
                dumpArray(p + "vy1.f64", d.vy1);
                dumpArray(p + "vz1.f64", d.vz1);
            }
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

        void queryTerminate(bool &terminate, [[maybe_unused]] LARE::simulationData &d, SAMS::timeState &t)
        {
            if (static_cast<long>(t.step) >= cfg().steps)
                terminate = true;
        }
    };
} // namespace icdrv

int main(int argc, char *argv[])
{
    auto &c = icdrv::cfg();
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        if (a == "--phases")
        {
            c.phases = true;
            continue;
        }
        if (i + 1 >= argc)
        {
            std::fprintf(stderr, "missing value for %s\n", a.c_str());
            return 2;
        }
        const std::string v = argv[++i];
        if (a == "--ic")
            c.ic = v;
        else if (a == "--steps")
            c.steps = std::stol(v);
        else if (a == "--out")
            c.out = v;
        else if (a == "--dump")
        {
            std::stringstream ss(v);
            std::string tok;
            while (std::getline(ss, tok, ','))
                c.dumps.insert(std::stol(tok));
        }
        else
        {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    int rargc = 1;
    char prog[] = "ic_driver";
    char *rargvBuf[] = {prog, nullptr};
    char **rargv = rargvBuf;
    SAMS::MPI::initialize(rargc, rargv);
    portableWrapper::initialize(rargc, rargv);
    SAMS::runner<LARE::LARE3D, LARE::LARE3DInitialConditions, examples::SodShockTube, examples::BrioAndWu,
                 examples::MHDRotor, examples::OrszagTang, examples::OrszagTang3D, examples::EmeryWindTunnel,
                 examples::KarmanVortex, icdrv::ICObserver>
        runner;
    runner.initialize(rargc, rargv);
    runner.activatePackage("LARE3D");
    runner.activatePackage(c.ic);
    runner.activatePackage("ICObserver");
    runner.initializePackages();
    runner.runPackages();
    runner.finalizePackages();
    runner.finalize();
    portableWrapper::finalize();

    std::ofstream m(c.out + "/meta.json");
    m << "{\"ic\": \"" << c.ic << "\", \"steps\": " << c.steps << ", \"dt\": [";
    for (size_t i = 0; i < c.dts.size(); ++i)
    {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "\"%a\"", c.dts[i]);
        m << (i ? ", " : "") << buf;
    }
    m << "]}\n";
    return 0;
}
