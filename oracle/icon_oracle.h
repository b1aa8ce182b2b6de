/*
 * icon_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference's ICON triangle shallow-water tendency
 * (src/FiniteVolumes/ICONShallowWater.jl), RK4 step and diagnostics, used by tests/ to
 * check the HIP path.  Nothing in the product (dmplex_amd/) may call it.
 *
 * Layout is the reference's: tables are k x n column-major (entry (k, j) at k + j*K),
 * indices 0-based int64, padded slots index 0 / weight 0; the state is [u (ne); h (nc)].
 * Every kernel is one OpenMP loop over its entities with the reference's per-entity
 * sequential summation order and no FMA contraction.
 */
#ifndef ICON_ORACLE_H
#define ICON_ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t nc, ne, nv, kv, kr;
    /* ICONShallowWater fields (ICONShallowWater.jl:50-76); cells are triangles (kc = 3) */
    const int64_t* cell_edges;     /* 3 x nc */
    const double* divergence;      /* 3 x nc: n l / A */
    const double* bilinear;        /* 3 x nc: e_bln_c_s */
    const int64_t* edge_cells;     /* 2 x ne */
    const int64_t* edge_vertices;  /* 2 x ne */
    const double* edge_height;     /* 2 x ne: c_lin_e */
    const double* gradient;        /* 2 x ne: coeff_gradekin */
    const double* inverse_dual;    /* ne */
    const int64_t* rbf_edges;      /* kr x ne */
    const double* rbf;             /* kr x ne */
    const double* edge_coriolis;   /* ne */
    const int64_t* vertex_edges;   /* kv x nv */
    const double* curl;            /* kv x nv */
    const int64_t* vertex_cells;   /* kv x nv */
    const double* vertex_height;   /* kv x nv */
    const double* coriolis;        /* nv */
    const double* cell_areas;      /* nc */
    const double* vertex_areas;    /* nv */
    double gravity;
    const double* damping;         /* ne, or NULL without divergence damping */
    /* work vectors */
    double *v, *Ke, *K, *F, *zv, *hv, *qv, *D, *G;
} icon_oracle_model;

/* kinetic_energy!: v, Ke, F at the edges and K at the cells (ICONShallowWater.jl:184-189) */
void icon_oracle_kinetic_energy(const icon_oracle_model* m, const double* u, const double* h);
/* vertex_pv!: zeta_v, h_v, q_v (TRiSKShallowWater.jl:355-360) */
void icon_oracle_vertex_pv(const icon_oracle_model* m, const double* u, const double* h);
void icon_oracle_tendency(const icon_oracle_model* m, const double* y, double* dy);
void icon_oracle_rk4_steps(const icon_oracle_model* m, double* y, double dt, int64_t nsteps, double* k1, double* k2,
                           double* k3, double* k4, double* stage);
double icon_oracle_total_mass(const icon_oracle_model* m, const double* y);
double icon_oracle_total_energy(const icon_oracle_model* m, const double* y);
double icon_oracle_potential_enstrophy(const icon_oracle_model* m, const double* y);
#endif
