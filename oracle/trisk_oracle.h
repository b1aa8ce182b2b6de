/*
 * trisk_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference's TRiSK shallow-water tendency, RK4 step and
 * diagnostics, used by tests/ to check the HIP path and by bench.py as the timed CPU
 * baseline.  Nothing in the product (dmplex_amd/) may call it.
 *
 * Layout is the reference's: tables are k x n column-major (entry (k, j) at k + j*K),
 * indices 0-based int64, padded slots index 0 / weight 0; the state is [u (ne); h (nc)].
 * Every kernel is one OpenMP loop over its entities with the reference's per-entity
 * sequential summation order and no FMA contraction, so the result does not depend on
 * the thread count.
 */
#ifndef TRISK_ORACLE_H
#define TRISK_ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t nc, ne, nv, kc, kv, kt, kg;
    /* TRiSKShallowWater fields (src/FiniteVolumes/TRiSKShallowWater.jl:67-92) */
    const int64_t* cell_edges;     /* kc x nc */
    const double* divergence;      /* kc x nc: n l / A */
    const double* kinetic;         /* kc x nc: alpha l d_ei / 2A */
    const int64_t* edge_cells;     /* 2 x ne */
    const int64_t* edge_vertices;  /* 2 x ne */
    const double* edge_height;     /* 2 x ne: alpha d_ei / d */
    const double* inverse_dual;    /* ne */
    const int64_t* vertex_edges;   /* kv x nv */
    const double* curl;            /* kv x nv: t d / A_v */
    const int64_t* vertex_cells;   /* kv x nv */
    const double* vertex_height;   /* kv x nv: K_iv / A_v */
    const double* coriolis;        /* nv */
    const int64_t* trisk_edges;    /* kt x ne */
    const double* trisk;           /* kt x ne: w l' / d */
    const double* cell_areas;      /* nc */
    const double* vertex_areas;    /* nv */
    double gravity;
    /* vertex kinetic energy (alpha != 1), else NULL */
    const int64_t* cell_vertices;  /* kc x nc */
    const double* vk_kites;        /* kc x nc: (1-alpha) K_iv / A_i */
    const double* vk_weights;      /* kv x nv: A_ev / A_v */
    const double* vk_height;       /* 2 x ne: 2 (1-alpha) A_ev / (l d) */
    /* Gassmann PV flux, else NULL */
    const int64_t* g_edges;        /* kg x ne */
    const int64_t* g_pv_edges;     /* kg x ne */
    const double* g_weights;       /* kg x ne */
    const double* g_rhombus;       /* 2 x ne */
    const double* g_coriolis;      /* ne */
    int g_mean;
    /* work vectors */
    double *K, *B, *F, *qe, *zv, *hv, *qv, *Kv;
} oracle_model;

void oracle_kinetic_energy(const oracle_model* m, const double* u, const double* h);
void oracle_vertex_pv(const oracle_model* m, const double* u, const double* h);
void oracle_mass_flux(const oracle_model* m, const double* u, const double* h);
void oracle_tendency(const oracle_model* m, const double* y, double* dy);
void oracle_rk4_step(const oracle_model* m, double* y, double dt, double* k1, double* k2, double* k3,
                     double* k4, double* stage);
void oracle_rk4_steps(const oracle_model* m, double* y, double dt, int64_t nsteps, double* k1, double* k2,
                      double* k3, double* k4, double* stage);
double oracle_total_mass(const oracle_model* m, const double* y);
double oracle_total_energy(const oracle_model* m, const double* y);
double oracle_potential_enstrophy(const oracle_model* m, const double* y);
void oracle_edge_sum(int64_t n, int64_t k, const int64_t* edges, const double* weights, const double* x,
                     double* out);
/* Adjoint, every PV flux and kinetic-energy form: ybar = (d tendency/dy)^T lam at y; and
 * ybar <- (d y_T / d y0)^T ybar over nsteps rk4_step! from y0.  Both return 0. */
int oracle_tendency_vjp(const oracle_model* m, const double* y, const double* lam, double* ybar);
int oracle_rk4_adjoint(const oracle_model* m, const double* y0, double dt, int64_t nsteps, double* ybar);
int oracle_threads(void);
void oracle_set_threads(int n);
#endif
