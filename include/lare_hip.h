/*
 * lare_hip.h -- C-ABI of the MI355X LARE3D Lagrangian-remap step.
 *
 * The library (sams_amd/liblare_hip.so) owns one device-resident copy of a
 * LARE3D grid per context and advances it with hand-written gfx950 kernels.
 * All entry points return 0 on success and a negative code on failure; the
 * message of the last failure on the calling thread is lare_last_error().
 * A context is thread-compatible: one host thread drives it, on the HIP
 * stream the context owns.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree).  Arrays cross the boundary in the reference's own
 * layout: row-major with z fastest, with the reference's bounds
 * (src/PhysicsPackages/LARE/setup.cpp:38-70, include/pp/array.h:72-95).
 */
#ifndef LARE_HIP_H
#define LARE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Registered variables of the LARE3D package, setup.cpp:45-69. */
enum lare_field {
    LARE_RHO = 0,
    LARE_ENERGY_ELECTRON = 1,
    LARE_ENERGY_ION = 2,
    LARE_VX = 3,
    LARE_VY = 4,
    LARE_VZ = 5,
    LARE_BX = 6,
    LARE_BY = 7,
    LARE_BZ = 8,
    LARE_VX1 = 9,
    LARE_VY1 = 10,
    LARE_VZ1 = 11,
    LARE_DM = 12,
    LARE_NFIELDS = 13
};

/* Built-in boundary conditions,
 * include/boundaryConditions/builtInBoundaryConditions.h:54-300. */
enum lare_bc_kind {
    LARE_BC_NONE = 0,
    LARE_BC_PERIODIC = 1,  /* simplePeriodicBC */
    LARE_BC_ZERO_GRADIENT = 2, /* simpleZeroGradientBC */
    LARE_BC_MIRROR = 3,    /* simpleMirrorBC */
    LARE_BC_CLAMP = 4,     /* simpleClamp(value) */
    LARE_BC_MAP = 5        /* layer by layer copies and constants: lare_set_bc_map */
};

/* Entries of a lare_set_bc_map layer table besides a source index. */
#define LARE_BC_MAP_KEEP (-1000)  /* the layer is left as it is */
#define LARE_BC_MAP_VALUE (-1001) /* the layer is set to the map's value */

/* Storage precision of the state fields. */
enum lare_precision {
    LARE_FP64 = 0,
    LARE_FP32_STORAGE = 1
};

/* Coordinate systems, geometryType of shared_data.h:38-43: x, y, z are
 * (x, y, z), (r, phi, z) or (r, theta, phi). */
enum lare_geometry {
    LARE_CARTESIAN = 0,
    LARE_CYLINDRICAL = 1,
    LARE_SPHERICAL = 2
};

/* Grid and physics scalars of LARE::simulationData
 * (src/PhysicsPackages/LARE/shared_data.h:60-226) that the step reads.
 * The axes are the reference's linear axes built from (n, min, max)
 * (include/harness/axisRegistry.h:973-984). */
typedef struct lare_desc {
    int64_t nx, ny, nz;
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double gas_gamma;
    double visc1, visc2;
    double mu0;
    double dt_multiplier;
    int32_t resistive;      /* resistiveMHD */
    int32_t precision;      /* enum lare_precision */
    double eta_background, eta0, j_max;
    int32_t rke;            /* remap kinetic-energy correction (shared_data.h:118) */
    int32_t geometry;       /* enum lare_geometry (shared_data.h:109); cylindrical and
                               spherical grids run kernel set 0 only */
} lare_desc;

typedef struct lare_ctx lare_ctx;

/* Last error message of the calling thread ("" if none). */
const char *lare_last_error(void);

/* Allocate a context and its device arrays; fields start at zero and the
 * previous dt at 0, as LARE3D::allocate leaves them (setup.cpp:76-255),
 * with the metric of LARE3D::grid (setup.cpp:260-437).
 * Replaces LARE3D::registerVariables / allocate / grid. */
int lare_ctx_create(const lare_desc *desc, lare_ctx **out);
int lare_ctx_destroy(lare_ctx *ctx);

/* Bounds of a field in the reference's indexing (lb/ub inclusive).
 * Replaces variableRegistry::fillPPArray (variableDef.h:611-624). */
int lare_field_bounds(const lare_ctx *ctx, int field, int64_t lb[3], int64_t ub[3]);

/* Copy a host array of the field's full reference extent (lare_field_bounds)
 * to / from the device.  Replaces the host residency of the reference's
 * arrays (initial conditions write them; diagnostics.cpp:36-47 reads them). */
int lare_upload_field(lare_ctx *ctx, int field, const double *host);
int lare_download_field(lare_ctx *ctx, int field, double *host);

/* Attach a built-in boundary condition to one edge (0 lower, 1 upper) of one
 * dimension of a field.  Replaces variableDef::addBoundaryCondition
 * (variableDef.h:402-425) for the built-in classes. */
int lare_set_bc(lare_ctx *ctx, int field, int dim, int edge, int kind, double value);
/* A condition given layer by layer on one edge: the edge's ghost layers, the
 * lower edge's -2..0 (staggered along dim) or -1..0 (centred), the upper
 * edge's n..n+2 or n+1..n+2, in that order, each take src[k]: the value of
 * layer src[k] of the same line (reference indexing along dim, read before
 * the edge writes), LARE_BC_MAP_VALUE, or LARE_BC_MAP_KEEP (src[2] of a
 * centred variable is ignored).  Replaces LAREBoundaryConditions
 * (src/InitialConditions/LARE/lareBoundaryClass.h:10-34) carrying the
 * LARE3DInitialConditions BC_OTHER rules
 * (src/InitialConditions/LARE/boundary.cpp:31-1120): copies such as
 * bx(-2) = bx(2), constants such as vx(-2..0) = 0.  Kernel set 0 does not
 * run it. */
int lare_set_bc_map(lare_ctx *ctx, int field, int dim, int edge, const int32_t src[3], double value);

/* dt of the previous step (the value the next step's shock viscosity reads,
 * lagran.cpp:72,83-84); 0 for a fresh context. */
int lare_get_dt(lare_ctx *ctx, double *dt);
int lare_set_dt(lare_ctx *ctx, double dt);

/* LARE3D::startOfTimestep (shared_data.h:397-399 -> lagran.cpp:95-173),
 * including the CFL reduction of runner::calculateTimestep
 * (runner.h:158-169 -> lagran.cpp:380-443).  If dt_out is not NULL the call
 * waits for the step and returns the new dt there. */
int lare_start_of_timestep(lare_ctx *ctx, double *dt_out);
/* The same event in the two halves the runner's timestep protocol needs
 * (lagran.cpp:95-173): lare_viscosity_dt runs the prologue, the shock
 * viscosity and the CFL reduction and returns dt (already times
 * dt_multiplier) and dtr -- what LARE3D::set_dt hands to
 * runner::calculateTimestep (runner.h:158-169); lare_lagrange then runs the
 * resistive substeps (int(dt/dtr)+1 of them) and predictor_step with the dt
 * the runner settled on. */
int lare_viscosity_dt(lare_ctx *ctx, double *dt_out, double *dtr_out);
int lare_lagrange(lare_ctx *ctx, double dt);

/* LARE3D::halfTimestep (shared_data.h:405-407 -> lagran.cpp:739-804). */
int lare_half_timestep(lare_ctx *ctx);
/* LARE3D::endOfTimestep (shared_data.h:413-419 -> remap.cpp:23-111), with
 * energy_correction (diagnostics.cpp:125-135) when desc.rke is set.  The
 * correction's delta_ke is summed per cell in the order one thread of the
 * reference adds it (its atomic adds order a multi-threaded run
 * arbitrarily); kernel set 0 does not run it. */
int lare_end_of_timestep(lare_ctx *ctx);

/* One runner iteration (runner.h:564-586): the three events in order.
 * Enqueued on the context's stream; dt_out (optional) waits for it. */
int lare_step(lare_ctx *ctx, double *dt_out);
/* nsteps iterations; dt_seq (optional, nsteps doubles) receives every dt. */
int lare_steps(lare_ctx *ctx, int32_t nsteps, double *dt_seq);

/* Wait for all work enqueued on the context. */
int lare_sync(lare_ctx *ctx);

/* Kernel set (no counterpart in the reference): 1 (default) runs the fused
 * kernels, choosing the lane-walking forms (a lane advancing several cells
 * and reusing the values its neighbours share) for launches of 300000
 * cells or more; 2 uses the lane-walking forms on every launch, 3
 * never; 0 runs one kernel per reference kernel over the reference's
 * ranges.  All give bit-identical results; 0, 2 and 3 exist to localise a
 * parity failure and to test every form on small grids. */
int lare_set_kernel_set(lare_ctx *ctx, int32_t kernel_set);

/* Whether lare_steps replays one captured HIP graph of the step (default 1;
 * no counterpart in the reference).  The graph is recaptured after any
 * boundary-condition or kernel-set change. */
int lare_set_graph(lare_ctx *ctx, int32_t on);

/* Per-kernel timing (no counterpart in the reference, whose only timers are
 * per-event host clocks, runnerMacros.h:113-150): between begin and end
 * every kernel launch on the context is bracketed by HIP events on its
 * stream; end aggregates them by kernel name with the kernel's algorithmic
 * HBM bytes: distinct arrays read and written (8 bytes each in FP64, 4 for
 * FP32-storage arrays) per cell-update, times the nx ny nz cells of the
 * grid the launch advances. */
int lare_profile_begin(lare_ctx *ctx);
int lare_profile_end(lare_ctx *ctx);
int lare_profile_count(lare_ctx *ctx, int32_t *n);
int lare_profile_get(lare_ctx *ctx, int32_t i, char *name, int32_t name_len, int32_t *calls,
                     double *total_ms, double *bytes);

/* Total energy of the grid in FP64 (no counterpart in the reference): out[0]
 * internal sum rho (e_e + e_i) cv over the domain cells, out[1] kinetic
 * sum 1/2 rho_v |v|^2 cv_v over the vertices 1..n, out[2] magnetic
 * sum |B_c|^2 / (2 mu0) cv with B_c the cell average of the face values,
 * out[3] their sum.  Summed in a fixed order (reproducible). */
int lare_energy(lare_ctx *ctx, double *out);

/* lare_energy without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same sums, same order, same bits as lare_energy.  The partial sums live in
 * a fixed 16 MiB device array of the library, one per GPU (looked up when a
 * context is created), which limits the grid to about 560^3 cells: a larger
 * grid fails with a clear error and nothing allocated.  Calls that overlap
 * on one GPU (contexts on different streams, other host threads) race for
 * that array; serialise them.  Blocking. */
int lare_energy_noalloc(lare_ctx *ctx, double *out);

/* Quotients as the kernels form them (no counterpart in the reference: a
 * self-check of the device's FP64 division, for tests).  q[i] = a[i] / d[i]
 * rounded to nearest even, for host arrays of n values, computed on the
 * device by the division every kernel uses (mode 0) or by the
 * shared-divisor quotient (mode 1); mode 2 gives q[i] = sqrt(a[i]) as the
 * kernels compute it (d is read but unused).  Blocking; allocates device
 * memory. */
int lare_div_check(const double *a, const double *d, double *q, int64_t n, int32_t mode);

/* ------------------------------------------------------------------------
 * Batches (no counterpart in the reference, whose runner steps one
 * simulationData at a time): contexts of one storage precision, of any
 * sizes, boundary conditions and physics parameters, that advance together.
 * Every kernel of a step is one launch over all members; each member keeps
 * its own state, dt and step count, and its results are bit-identical to
 * stepping it alone.  lare_batch_steps enqueues on the batch's own stream
 * and returns; call lare_batch_sync before using a member on its own
 * (download, lare_get_dt, lare_step ...).  A member must outlive its batch.
 */
typedef struct lare_batch lare_batch;
int lare_batch_create(lare_ctx *const *members, int32_t n, lare_batch **out);
int lare_batch_destroy(lare_batch *batch);
/* nsteps full steps of every member (a captured HIP graph per step unless
 * lare_batch_set_graph(batch, 0)). */
int lare_batch_steps(lare_batch *batch, int32_t nsteps);
int lare_batch_sync(lare_batch *batch);
int lare_batch_set_graph(lare_batch *batch, int32_t on);
/* lare_energy of every member: 4 doubles per member, in member order. */
int lare_batch_energy(lare_batch *batch, double *out);
/* HIP-event timing of every launch between begin and end (eager steps):
 * the summed kernel time and the number of launches. */
int lare_batch_profile_begin(lare_batch *batch);
int lare_batch_profile_end(lare_batch *batch, double *total_ms, int32_t *launches);

/* ------------------------------------------------------------------------
 * Launch configuration and autotuner (no counterpart in the reference).
 * The fused kernel set has several forms of its stencil kernels (lanes
 * walking several cells and reusing the values neighbours share, or one
 * cell per lane) and workgroup orders; every configuration gives
 * bit-identical results, so the fastest can be chosen by measurement.
 */
typedef struct lare_tune {
    int32_t walk;       /* lane-walking kernels: 0 by launch size, 1 always, 2 never */
    int32_t march_k;    /* cells per lane of the x/y transport and CT-flux kernels: 8 or 16 */
    int32_t zx_k;       /* cells per lane of the z-line x-march kernels: 8 or 16 */
    int32_t zline;      /* z-sweep kernels exchanging z-neighbours between lanes: 1, or 0 */
    int32_t order;      /* workgroup order within a grid: 1 band-major, 0 plane-major */
    int32_t ct_fuse;    /* CT flux and update in one kernel (lane-walking): 2 all sweeps, 1 x and z, 0 none */
    double ms_per_step; /* the chosen configuration's measured step time (autotune) */
} lare_tune;
/* Times one captured step of every candidate configuration (coordinate
 * descent over the fields above, `steps` replays each after one warm-up),
 * keeps the fastest and returns it.  The members' device state is saved
 * before and restored after, so tuning changes no result; it needs device
 * memory for one copy of each member's arrays.  Only between steps, on the
 * fused kernel set.  A later lare_set_kernel_set resets the configuration. */
int lare_autotune(lare_ctx *ctx, int32_t steps, lare_tune *out);
int lare_batch_autotune(lare_batch *batch, int32_t steps, lare_tune *out);
int lare_get_tune(lare_ctx *ctx, lare_tune *out);
int lare_set_tune(lare_ctx *ctx, const lare_tune *t);
int lare_batch_get_tune(lare_batch *batch, lare_tune *out);
int lare_batch_set_tune(lare_batch *batch, const lare_tune *t);

/* ------------------------------------------------------------------------
 * Problem set-up (the initial-condition packages' role: controlVariables,
 * setDomain, initialConditions, setBoundaryConditions of
 * src/InitialConditions/OrszagTang3D/OrszagTang3D.cpp:28-158 and
 * src/InitialConditions/BrioAndWu/BrioAndWu.cpp:29-155).  Every array is
 * filled over the centred extent rho.getRange(), as those packages do; the
 * -2 layers of staggered arrays stay zero.
 */
enum lare_problem_kind {
    /* Orszag-Tang on the unit cube, gamma 5/3, all boundaries periodic:
     * rho = 25/(36 pi), p = 5/(12 pi) (or beta/(8 pi) when beta > 0),
     * v = (-sin 2 pi yb, sin 2 pi xb, 0) [+ amp * N(0,1) per component],
     * B = (-sin 2 pi yc, sin 4 pi xc, 0) / sqrt(4 pi),
     * e_e = e_i = p / ((gamma-1) rho) / 2. */
    LARE_PROBLEM_ORSZAG_TANG = 0,
    /* x-directed Brio-Wu on x in [-1,1], y,z in [0, 2 ny/nx], [0, 2 nz/nx],
     * gamma 2: rho,p,By = 1,1,1 for xc below the x midpoint and 0.125,0.1,-1 otherwise,
     * Bx = 0.75, rho *= 1 + noise*(2U-1); x boundaries bcx, y/z periodic. */
    LARE_PROBLEM_BRIO_WU = 1
};

typedef struct lare_problem {
    int32_t kind;       /* enum lare_problem_kind */
    int32_t bcx;        /* Brio-Wu x boundary: LARE_BC_ZERO_GRADIENT (0 also means
                           this), LARE_BC_MIRROR, LARE_BC_PERIODIC, or LARE_BC_CLAMP
                           (walls: velocities clamped to 0, the rest zero-gradient) */
    int64_t nx, ny, nz;
    uint64_t seed;      /* counter-based streams: splitmix64 of (seed, stream, index) */
    double amp;         /* Orszag-Tang: Gaussian velocity perturbation amplitude */
    double beta;        /* Orszag-Tang: plasma beta; <= 0 keeps p = 5/(12 pi) */
    double noise;       /* Brio-Wu: relative uniform density noise */
    double eta;         /* background resistivity (eta_background) */
    int32_t precision;  /* enum lare_precision */
    int32_t rke;        /* remap kinetic-energy correction on */
    double eta0;        /* resistivity added where |j| > j_max (lagran.cpp:473) */
    double j_max;       /* current threshold of eta0; <= 0 means 1e30 (never reached) */
    /* eta > 0 or eta0 > 0 turns resistive MHD on (shared_data.h:113-117) */
    int32_t geometry;   /* enum lare_geometry */
    int32_t reserved;
    double x0, y0;      /* shifts of the domain's x and y extents (a curvilinear grid
                           needs r > 0, and theta away from the poles); Brio-Wu's
                           discontinuity stays at the x midpoint */
} lare_problem;

/* Grid and physics scalars of the problem. */
int lare_problem_desc(const lare_problem *p, lare_desc *out);
/* Initial value of one field over its reference extent into host memory. */
int lare_problem_fill(const lare_problem *p, int field, double *host);
/* Attach the problem's boundary conditions to every registered variable. */
int lare_problem_bcs(lare_ctx *ctx, const lare_problem *p);
/* Fill, upload and attach: the context then holds the problem at t = 0. */
int lare_problem_init(lare_ctx *ctx, const lare_problem *p);

#ifdef __cplusplus
}
#endif
#endif /* LARE_HIP_H */
