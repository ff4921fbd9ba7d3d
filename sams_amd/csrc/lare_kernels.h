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
    int rke;  // remap kinetic-energy correction
    int geometry;  // enum lare_geometry: 0 Cartesian, 1 cylindrical, 2 spherical
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
    int bcm[13][3][2][3];  // layer tables of LARE_BC_MAP edges
    hipStream_t stream;
    size_t box;          // doubles per 3D array
    void *pool;          // one allocation holding every device array
    size_t pool_bytes;
    Prof prof;
    int kernel_set;      // 0: one kernel per reference kernel, 1: the grid-group engine
    // one captured step (start + half + end), replayed by lare_steps
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = true;
    // storage of the engine's arrays (kernel set 1): FP64, or FP32 with the
    // fluxes in FP64; the engine's device view in that storage type
    bool fp32 = false;
    EDev<double> e64{};
    EDev<float> e32{};
    long bc_version = 0;  // bumped by every boundary-condition change
    // which copy holds rho, e_e, e_i, v between the engine's events: 0 the
    // first (between steps), 1 the second (after the corrector)
    int st = 0;
};

// Metric arrays from the 1D deltas (LARE3D::grid, setup.cpp:337-390).
void build_metric(Grid &g);

// A field over its reference extent lb..ub (inclusive), held contiguous in
// device memory `lin` (z fastest), into (to_box) or out of its padded box;
// f32 marks an FP32-storage box.  Enqueued on st.
void box_copy(hipStream_t st, void *box, bool f32, double *lin, const Geo &G, const int lb[3], const int ub[3],
              bool to_box);



// The three LARE3D events (shared_data.h:397-419), enqueued on g.stream.
// step_start = step_start_a + step_start_b(2): the CFL reduction of the
// runner's calculateTimestep sits between them (lagran.cpp:156).
void step_start(Grid &g);
void step_start_a(Grid &g);
void step_start_b(Grid &g, int substeps);
void step_half(Grid &g);
void step_end(Grid &g);

}  // namespace lare
