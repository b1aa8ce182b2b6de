/*
 * mpas_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the explicit stage tendencies of the reference's MPAS-Gassmann
 * dycore, on flat ground or over orography (slow_tendencies!, src/Nonhydrostatic/MPAS/Gassmann/Stepping.jl:22-56,
 * with the kernels of src/Nonhydrostatic/MPAS/Operators.jl and Gassmann/Tendencies.jl), used
 * by tests/ to check the HIP path.  Nothing in the product (dmplex_amd/) may call it.
 *
 * Layout is the reference's: tables k x n column-major (entry (k, j) at k + j*K) with 0-based
 * indices; the edge-first TRiSK tables trisk (ne x kt) and trisk_index (ne x kt x 3) and all
 * fields (points x levels) with the point index fastest.
 */
#ifndef MPAS_ORACLE_H
#define MPAS_ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t nc, ne, nv, nlev, kc, kv, kt;
    const int64_t* cell_edges;
    const double* kinetic;
    const double* inner;
    const int64_t* edge_cells;
    const int64_t* edge_vertices;
    const double* edge_weight;
    const double* inverse_dual;
    const int64_t* vertex_edges;
    const double* curl;
    const int64_t* vertex_cells;
    const double* vertex_weight;
    const double* coriolis;
    const int32_t* trisk_index;
    const double* trisk;
    const double* rhombus;
    const double* edge_coriolis;
    const double *dz, *dzh, *wlow, *edge_dz, *edge_dzh;
    const int64_t* cell_vertices;  /* NULL: cell kinetic energy only */
    const double* vk_kites;
    const double* vk_weights;
    int32_t iap_pv, vertical, five_band, reserved;
    /* work fields (outputs), points x levels */
    double *theta, *exner, *K, *qv, *zeta, *rhov, *qe, *rhoe, *Fs, *eta, *tw, *tvn, *Kv;
    /* orography: 0 flat, 1 pressure gradient along the levels, 2 at constant height; the
     * metrics z_full, pi_ref (nc x nlev), ground_q (nc x 3), ddxn (ne x nlev), zlev (0-based)
     * and zwt (2 x ne x nlev, entry (s, e, k) at s + 2 (e + ne k)) */
    int32_t orography, reserved2;
    const double *z_full, *pi_ref, *ground_q, *ddxn;
    const int32_t* zlev;
    const double* zwt;
    /* orography work fields: w_c at the full levels and the interfaces, w - w_c, pi', dpi'/dz */
    double *wcf, *wch, *what, *pip, *dpidz;
} mpas_oracle_model;

void mpas_oracle_slow_tendencies(const mpas_oracle_model* m, const double* vn, const double* w, const double* rho,
                                 const double* rhotheta, double pressure);
/* One kernel on the current work fields: 0 diagnose, 1 vertex_kinetic, 2 cell_from_vertex,
 * 3 w_kinetic, 4 vertex_pv, 5 edge_pv, 6 edge_density, 7 eta, 8 w_tendency_vi, 9 w_tendency,
 * 10 vn_tendency, 11 concorr_cell, 12 concorr_interface, 13 contravariant_w,
 * 14 exner_deviation_centred, 15 w_concorr_tendency; -1 for another id. */
int mpas_oracle_kernel(const mpas_oracle_model* m, int id, const double* vn, const double* w, const double* rho,
                       const double* rhotheta, double pressure);
#endif
