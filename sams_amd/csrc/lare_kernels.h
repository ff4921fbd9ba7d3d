// Host-visible declarations of the LARE3D device step.
#pragma once

#include "lare_dev.h"

#include <string>
#include <vector>

namespace lare {

// Built-in boundary conditions of one registered variable along one
// dimension (device view of one launch): the stagger along that dimension
// and across it, and the kind/value of the lower and upper edge.
struct BCVar {
    double *p;
    int sd, sa, sb;     // 1 = HALF_CELL: along dim, then the two other dims
    int k0, k1;         // enum lare_bc_kind of the lower / upper edge
    double v0, v1;      // clamp values
};
// Up to nine variables filled by one launch (blockIdx.z selects).
struct BCSet {
    BCVar v[9];
};

// Physics scalars of simulationData that the step reads.
struct Params {
    double gas_gamma, visc1, visc2, visc2_norm, mu0, dt_multiplier;
    int resistive;
    double eta_background, eta0, j_max;
};

// Optional per-kernel timing with HIP events on the grid's stream.
struct Prof {
    bool on = false;
    int dir = -1;  // sweep direction being launched; names "<DIR>" as "<0>".."<2>"
    struct Rec {
        std::string name;
        double bytes;  // algorithmic bytes of this launch
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
};

// One device-resident grid.
struct Grid {
    Dev dev;
    Params prm;
    int bck[13][3][2];
    double bcv[13][3][2];
    hipStream_t stream;
    size_t box;          // doubles per 3D array
    void *pool;          // one allocation holding every device array
    size_t pool_bytes;
    Prof prof;
    int kernel_set;      // 0: one kernel per reference kernel, 1: fused
    // one captured step (start + half + end), replayed by lare_steps
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = true;
};

// Metric arrays from the 1D deltas (LARE3D::grid, setup.cpp:337-390).
void build_metric(Grid &g);

// Boundary fills used by the fused sweep (lare_kernels.hip).
void dm_bcs(Grid &g, int dim);
void sweep_bcs(Grid &g);
void remap_v_fill(Grid &g);
void velocity_fill(Grid &g);
void energy_density_fill(Grid &g);
void dt_final(Grid &g);
void resistive_substeps(Grid &g, int substeps);
// Fused Lagrangian events (lare_fused.hip): start_a runs up to the CFL
// reduction, start_b the resistive substeps and the predictor.
void fused_start_a(Grid &g);
void fused_start_b(Grid &g, int substeps);
void fused_half(Grid &g);
// One fused directional remap sweep (lare_fused.hip).
void fused_sweep(Grid &g, int d, double xpass, double ypass, double zpass);

// The three LARE3D events (shared_data.h:397-419), enqueued on g.stream.
// step_start = step_start_a + step_start_b(2): the CFL reduction of the
// runner's calculateTimestep sits between them (lagran.cpp:156).
void step_start(Grid &g);
void step_start_a(Grid &g);
void step_start_b(Grid &g, int substeps);
void step_half(Grid &g);
void step_end(Grid &g);

}  // namespace lare
