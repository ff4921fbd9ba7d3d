/*
 * lare_oracle.h -- CPU restatement of the reference LARE3D step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * kernels: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product path never links or calls it.
 *
 * It restates, kernel by kernel and with the same ranges, array extents,
 * evaluation order and floating-point association, the reference's
 *   src/PhysicsPackages/LARE/{setup,lagran,remap,xremap,yremap,zremap,boundary}.cpp
 *   include/boundaryConditions/builtInBoundaryConditions.h
 *   include/harness/variableDef.h:539-605 (BC order)
 *   include/harness/axisRegistry.h:134-207,393-431,973-984 (axes)
 * for Cartesian geometry.  Build with -ffp-contract=off.
 *
 * Arrays use the reference's own extents (row-major, z fastest); the field
 * pointers returned by lo_field() address those arrays directly.
 */
#ifndef LARE_ORACLE_H
#define LARE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Registered variables, in the order used by the field/BC tables. */
enum lo_field_id {
    LO_RHO = 0, LO_EE, LO_EI, LO_VX, LO_VY, LO_VZ, LO_BX, LO_BY, LO_BZ,
    LO_VX1, LO_VY1, LO_VZ1, LO_DM, LO_NVAR
};

/* Built-in boundary conditions (builtInBoundaryConditions.h). */
enum lo_bc_kind { LO_BC_NONE = 0, LO_BC_PERIODIC, LO_BC_ZEROGRAD, LO_BC_MIRROR, LO_BC_CLAMP, LO_BC_MAP };
/* Entries of a map edge's layer table besides a source layer. */
#define LO_BC_MAP_KEEP (-1000)
#define LO_BC_MAP_VALUE (-1001)

typedef struct lo_desc {
    long nx, ny, nz;
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double gas_gamma, visc1, visc2, mu0, dt_multiplier;
    int resistive;              /* resistiveMHD */
    double eta_background, eta0, j_max;
    int rke;                    /* remap kinetic-energy correction */
    int geometry;               /* geometryType (shared_data.h:38-43): 0 Cartesian, 1 cylindrical, 2 spherical */
} lo_desc;

typedef struct lo_ctx lo_ctx;

lo_ctx *lo_create(const lo_desc *d);
void lo_destroy(lo_ctx *c);

/* Pointer to a registered variable's storage; lb/ub receive its bounds. */
double *lo_field(lo_ctx *c, int id, long lb[3], long ub[3]);
/* Any 3D array by reference name (e.g. "cv1", "p_visc", "alpha1"). */
double *lo_array(lo_ctx *c, const char *name, long lb[3], long ub[3]);
/* 1D axes by name: xb yb zb xc yc zc dxb dyb dzb dxc dyc dzc. */
double *lo_axis(lo_ctx *c, const char *name, long *lb, long *ub);

int lo_set_bc(lo_ctx *c, int var, int dim, int edge, int kind, double value);
/* A map edge: ghost layers g0..g1 of the edge (lower -2..0 or -1..0, upper
 * n..n+2 or n+1..n+2) take src[t]: a layer of the same line (read before the
 * edge writes), LO_BC_MAP_VALUE (value) or LO_BC_MAP_KEEP. */
int lo_set_bc_map(lo_ctx *c, int var, int dim, int edge, const int src[3], double value);

/* dt used by the shock viscosity of the next step (the previous step's dt). */
double lo_get_dt(const lo_ctx *c);
void lo_set_dt(lo_ctx *c, double dt);

/* One full step: startOfTimestep, halfTimestep, endOfTimestep.
 * Returns the dt that the step used. */
double lo_step(lo_ctx *c);

/* The three events separately (shared_data.h:397-419). */
double lo_start_of_timestep(lo_ctx *c);  /* returns the new dt */
void lo_half_timestep(lo_ctx *c);
void lo_end_of_timestep(lo_ctx *c);

/* Only the CFL reduction on the current state (set_dt, lagran.cpp:380-443),
 * after the lagrangian_step prologue that set_dt reads (p_visc is taken as it
 * stands).  Returns dt_multiplier * min. */
double lo_set_dt_only(lo_ctx *c);

void lo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
