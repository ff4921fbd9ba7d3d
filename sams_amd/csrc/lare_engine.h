// Grid-group engine: the product kernel set of the LARE3D step.
//
// A Group is an ordered set of grids of one storage type that advance
// together: every kernel of the step is ONE launch covering all members
// (uneven sizes included), each workgroup finding its grid by a binary
// search over the members' first-block offsets.  A single grid is a group of
// one.  Results per grid do not depend on the group it runs in: no value
// crosses grids and the only reduction (the CFL minimum) is per grid and
// exact.
#pragma once

#include "lare_kernels.h"

#include <array>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace lare {

struct Group {
    bool fp32 = false;
    std::vector<Grid *> mem;  // members, in order
    std::vector<int> res;     // indices of resistive members
    std::vector<int> rke;     // indices of members with the kinetic-energy correction
    // device arrays of EDev<T>: [0] all members, [1] resistive members,
    // [2] members with the kinetic-energy correction
    void *devs[3] = {nullptr, nullptr, nullptr};
    int n[3] = {0, 0, 0};
    struct Tab {
        int *boff = nullptr;
        int nblk = 0;
        int order = 0;
        double cells = 0.0;  // cells of the launch range
        double units = 0.0;  // domain cells of the members (nx ny nz each): the cell-updates it serves
    };
    // (subset, range spec, order)
    std::map<std::pair<int, std::array<int, 7>>, Tab> tabs;
    struct BTab {
        BCSetE *sets = nullptr;
        BCSet3 *sets3 = nullptr;  // all-dimension fill (dim 3 tables)
        int *boff = nullptr;
        int nblk = 0;
        bool any = false;
    };
    std::map<std::pair<int, int>, BTab> btabs;  // (variable group, dim); dim 3: all at once
    std::vector<long> bc_seen;                  // members' bc_version when btabs were built
    hipStream_t stream = nullptr;
    Prof *prof = nullptr;
    bool planning = false;
    int order = 1;    // workgroup order within a grid (Src::order)
    int march_k = 16;   // cells per lane of the x and y transport kernels (0: one)
    bool zline = true;  // z transport with neighbours exchanged between lanes
    int zx_k = 8;       // cells per lane of the z-line x-march kernels (0: one cell per lane)
    double walk_min_cells = 3.0e5;  // launches over fewer cells use one cell per lane
    int walk_mode = 0;              // lane-walking kernels: 0 by launch size, 1 always, 2 never
    int ct_fuse = 2;                // CT flux and update in one kernel (lane-walking only): 1 the x and z sweeps, 2 all three
    bool prep_all = true;           // the three sweeps' cell volumes and widths in one kernel
    bool bc_all = true;             // a variable group's fills of all three dimensions in one kernel
    bool csv_fused = true;          // sound speed and vertex values in one kernel (lane-walking only)
    // one captured step of the whole group
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

// Build the group's device views (members must outlive it); the group
// launches on `stream` and records into `prof` when profiling is on.
int group_init(Group &gp, const std::vector<Grid *> &members, hipStream_t stream, Prof *prof, std::string &err);
void group_free(Group &gp);
// Drop the captured graph (after a boundary-condition change).
void group_drop_graph(Group &gp);
// Create every table a step needs (outside graph capture).
void group_prepare(Group &gp);

// The three events of the runner (runner.h:564-586) for every member:
// start_a = prologue, shock viscosity and CFL reduction (dt on the device),
// start_b = resistive substeps (substeps of them) and predictor,
// half = corrector, end = remap with its boundary fills and the step count.
void group_start_a(Group &gp);
void group_start_b(Group &gp, int substeps);
void group_half(Group &gp);
void group_end(Group &gp);
inline void group_step(Group &gp)
{
    group_start_a(gp);
    group_start_b(gp, 2);
    group_half(gp);
    group_end(gp);
}

// Total energy of every member, FP64 (the energy-conservation diagnostic):
// out[4*i + 0..3] = internal, kinetic, magnetic, total, summed over the
// domain cells and vertices in a fixed order.
int group_energy(Group &gp, std::vector<double> &out, std::string &err);

// The same diagnostic without any device allocation, for callers that cannot
// allocate: partial sums and totals go to the file-scope device array of
// lare_engine.hip (one copy per GPU, energy_scratch_bytes() long; calls that
// overlap on one GPU race for it).  energy_scratch_address looks the array
// up on the current device (hipGetSymbolAddress; once per device, at set-up),
// group_energy_prepare builds the group's launch table (allocates: set-up
// only); group_energy_noalloc then launches, copies the four totals of each
// member to out and waits.  It fails, allocating nothing, when the table was
// not prepared or the group needs more scratch than the array holds.
size_t energy_scratch_bytes();
int energy_scratch_address(double **out, std::string &err);
void group_energy_prepare(Group &gp);
int group_energy_noalloc(Group &gp, double *scratch_dev, double *out, std::string &err);

// Fill the engine's constant cell-volume array of grid g (after its axes are
// on the device; waits for the fill).
int grid_fill_volume(Grid &g, std::string &err);

// q[i] = a[i] / d[i] for host arrays of n values, computed on the device by
// the kernels' division (mode 0) or shared-divisor quotient (mode 1); mode 2
// gives q[i] = sqrt(a[i]) as the kernels compute it (d unused).
int div_check(const double *a, const double *d, double *q, long n, int mode, std::string &err);

}  // namespace lare
