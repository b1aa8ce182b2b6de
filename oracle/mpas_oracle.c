/*
 * mpas_oracle.c -- TEST INFRASTRUCTURE ONLY (see mpas_oracle.h).
 *
 * Restates, kernel for kernel, the explicit tendencies of one MPAS-Gassmann stage
 * (slow_tendencies!, src/Nonhydrostatic/MPAS/Gassmann/Stepping.jl:22-56):
 *   diagnose_kernel!          Operators.jl:6-15      vertex_kinetic_kernel!   :18-25
 *   cell_from_vertex_kernel!  :28-35                 w_kinetic_kernel!        :38-42
 *   eta_kernel!               :44-52                 edge_density_kernel!     :55-62
 *   w_tendency_vi_kernel!     :69-85                 vertex_pv_kernel!        :88-101
 *   edge_pv_kernel!           :104-115               w_tendency_kernel!       :118-133
 *   vn_tendency_kernel!, edge_w, pressure_gradient             Gassmann/Tendencies.jl:4-67
 * and over orography (Stepping.jl:26-29,51-54,58-76)
 *   concorr_cell_kernel!      Operators.jl:135-144   w_concorr_tendency_kernel!  :146-152
 *   exner_deviation_centred_kernel!, height_difference, concorr_interface_kernel!,
 *   contravariant_w_kernel!   Common/CGridOperators.jl:93-100,193-214
 * One loop per kernel over (points, levels) with the reference's per-item arithmetic, in the
 * order slow_tendencies! launches them.  Indices are 0-based: layer k of the reference is k-1
 * here, interface j is j-1.
 */
#include "mpas_oracle.h"

#include <math.h>

#define AT(tab, K, k, j) (tab)[(k) + (int64_t)(j) * (K)]
#define F2(f, n, p, k) (f)[(p) + (int64_t)(n) * (k)]

static const double Rd = 287.0, Cpd = 1004.0, P0 = 1.0e5;
#define Cvd (Cpd - Rd)

static double sq(double x) { return x * x; }

static void diagnose(const mpas_oracle_model* m, const double* vn, const double* rho, const double* rt) {
    const int64_t nc = m->nc, ne = m->ne, KC = m->kc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t i = 0; i < nc; ++i) {
            double s = 0.0;
            for (int64_t j = 0; j < KC; ++j) s += AT(m->kinetic, KC, j, i) * sq(F2(vn, ne, AT(m->cell_edges, KC, j, i), k));
            F2(m->K, nc, i, k) = s;
            F2(m->theta, nc, i, k) = F2(rt, nc, i, k) / F2(rho, nc, i, k);
            F2(m->exner, nc, i, k) = pow(Rd * F2(rt, nc, i, k) / P0, Rd / Cvd);
        }
}

static void vertex_kinetic(const mpas_oracle_model* m, const double* vn) {
    const int64_t nv = m->nv, ne = m->ne, KV = m->kv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t v = 0; v < nv; ++v) {
            double s = 0.0;
            for (int64_t j = 0; j < KV; ++j)
                s += AT(m->vk_weights, KV, j, v) * sq(F2(vn, ne, AT(m->vertex_edges, KV, j, v), k));
            F2(m->Kv, nv, v, k) = s;
        }
}

static void cell_from_vertex(const mpas_oracle_model* m) {
    const int64_t nc = m->nc, nv = m->nv, KC = m->kc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t i = 0; i < nc; ++i) {
            double s = 0.0;
            for (int64_t j = 0; j < KC; ++j)
                s += AT(m->vk_kites, KC, j, i) * F2(m->Kv, nv, AT(m->cell_vertices, KC, j, i), k);
            F2(m->K, nc, i, k) += s;
        }
}

static void w_kinetic(const mpas_oracle_model* m, const double* w) {
    const int64_t nc = m->nc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t i = 0; i < nc; ++i)
            F2(m->K, nc, i, k) += (F2(m->dzh, nc, i, k) * sq(F2(w, nc, i, k)) +
                                   F2(m->dzh, nc, i, k + 1) * sq(F2(w, nc, i, k + 1))) / (4 * F2(m->dz, nc, i, k));
}

static void vertex_pv(const mpas_oracle_model* m, const double* vn, const double* rho) {
    const int64_t nv = m->nv, ne = m->ne, nc = m->nc, KV = m->kv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t v = 0; v < nv; ++v) {
            double z = 0.0, eta = 0.0;
            for (int64_t j = 0; j < KV; ++j) z += AT(m->curl, KV, j, v) * F2(vn, ne, AT(m->vertex_edges, KV, j, v), k);
            for (int64_t j = 0; j < KV; ++j)
                eta += AT(m->vertex_weight, KV, j, v) * F2(rho, nc, AT(m->vertex_cells, KV, j, v), k);
            F2(m->zeta, nv, v, k) = z;
            F2(m->rhov, nv, v, k) = eta;
            F2(m->qv, nv, v, k) = (z + m->coriolis[v]) / eta;
        }
}

static void edge_pv(const mpas_oracle_model* m) {
    const int64_t ne = m->ne, nv = m->nv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t e = 0; e < ne; ++e) {
            const int64_t v1 = AT(m->edge_vertices, 2, 0, e), v2 = AT(m->edge_vertices, 2, 1, e);
            if (m->iap_pv) {
                const double r1 = AT(m->rhombus, 2, 0, e), r2 = AT(m->rhombus, 2, 1, e);
                const double z = r1 * F2(m->zeta, nv, v1, k) + r2 * F2(m->zeta, nv, v2, k);
                const double eta = r1 * F2(m->rhov, nv, v1, k) + r2 * F2(m->rhov, nv, v2, k);
                F2(m->qe, ne, e, k) = (z + m->edge_coriolis[e]) / eta;
            } else {
                F2(m->qe, ne, e, k) = (F2(m->qv, nv, v1, k) + F2(m->qv, nv, v2, k)) / 2;
            }
        }
}

static void edge_density(const mpas_oracle_model* m, const double* vn, const double* rho) {
    const int64_t ne = m->ne, nc = m->nc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t e = 0; e < ne; ++e) {
            const double r = AT(m->edge_weight, 2, 0, e) * F2(rho, nc, AT(m->edge_cells, 2, 0, e), k) +
                             AT(m->edge_weight, 2, 1, e) * F2(rho, nc, AT(m->edge_cells, 2, 1, e), k);
            F2(m->rhoe, ne, e, k) = r;
            F2(m->Fs, ne, e, k) = r * F2(vn, ne, e, k);
        }
}

static void eta_kernel(const mpas_oracle_model* m, const double* vn, const double* w) {
    const int64_t ne = m->ne, nc = m->nc, nlev = m->nlev;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j <= nlev; ++j)
        for (int64_t e = 0; e < ne; ++e) {
            if (j >= 1 && j <= nlev - 1)
                F2(m->eta, ne, e, j) =
                    (F2(vn, ne, e, j) - F2(vn, ne, e, j - 1)) / F2(m->edge_dzh, ne, e, j) -
                    (F2(w, nc, AT(m->edge_cells, 2, 1, e), j) - F2(w, nc, AT(m->edge_cells, 2, 0, e), j)) *
                        m->inverse_dual[e];
            else
                F2(m->eta, ne, e, j) = 0.0;
        }
}

static void w_tendency_vi(const mpas_oracle_model* m, const double* vn, int massweighted) {
    const int64_t nc = m->nc, ne = m->ne, nlev = m->nlev, KC = m->kc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j <= nlev; ++j)
        for (int64_t i = 0; i < nc; ++i) {
            if (j >= 1 && j <= nlev - 1) {
                double s = 0.0;
                const double wl = F2(m->wlow, nc, i, j);
                for (int64_t n = 0; n < KC; ++n) {
                    const int64_t e = AT(m->cell_edges, KC, n, i);
                    const double rl = massweighted ? F2(m->rhoe, ne, e, j - 1) : 1.0;
                    const double ru = massweighted ? F2(m->rhoe, ne, e, j) : 1.0;
                    s += AT(m->inner, KC, n, i) * F2(m->eta, ne, e, j) *
                         (rl * F2(vn, ne, e, j - 1) + ru * F2(vn, ne, e, j)) / (2 * (wl * rl + (1 - wl) * ru));
                }
                F2(m->tw, nc, i, j) = s - (F2(m->K, nc, i, j) - F2(m->K, nc, i, j - 1)) / F2(m->dzh, nc, i, j);
            } else {
                F2(m->tw, nc, i, j) = 0.0;
            }
        }
}

static void w_tendency(const mpas_oracle_model* m, const double* vn, const double* w) {
    const int64_t nc = m->nc, ne = m->ne, nlev = m->nlev, KC = m->kc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j <= nlev; ++j)
        for (int64_t i = 0; i < nc; ++i) {
            if (j >= 1 && j <= nlev - 1) {
                double s = 0.0;
                for (int64_t n = 0; n < KC; ++n) {
                    const int64_t e = AT(m->cell_edges, KC, n, i);
                    const double v = (F2(vn, ne, e, j - 1) + F2(vn, ne, e, j)) / 2;
                    s += 2 * AT(m->kinetic, KC, n, i) * v *
                         (F2(w, nc, AT(m->edge_cells, 2, 1, e), j) - F2(w, nc, AT(m->edge_cells, 2, 0, e), j)) *
                         m->inverse_dual[e];
                }
                F2(m->tw, nc, i, j) = -s - F2(w, nc, i, j) * (F2(w, nc, i, j + 1) - F2(w, nc, i, j - 1)) /
                                                (F2(m->dz, nc, i, j - 1) + F2(m->dz, nc, i, j));
            } else {
                F2(m->tw, nc, i, j) = 0.0;
            }
        }
}

/* concorr_cell_kernel! (Operators.jl:135-144) */
static void concorr_cell(const mpas_oracle_model* m, const double* vn) {
    const int64_t nc = m->nc, ne = m->ne, KC = m->kc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < m->nlev; ++k)
        for (int64_t i = 0; i < nc; ++i) {
            double s = 0.0;
            for (int64_t n = 0; n < KC; ++n) {
                const int64_t e = AT(m->cell_edges, KC, n, i);
                s += AT(m->inner, KC, n, i) * F2(vn, ne, e, k) * F2(m->ddxn, ne, e, k);
            }
            F2(m->wcf, nc, i, k) = s;
        }
}

/* concorr_interface_kernel! (CGridOperators.jl:203-208) */
static void concorr_interface(const mpas_oracle_model* m) {
    const int64_t nc = m->nc, nlev = m->nlev;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j <= nlev; ++j)
        for (int64_t i = 0; i < nc; ++i) {
            double wc;
            if (j >= 1 && j <= nlev - 1)
                wc = F2(m->wlow, nc, i, j) * F2(m->wcf, nc, i, j - 1) + (1 - F2(m->wlow, nc, i, j)) * F2(m->wcf, nc, i, j);
            else if (j == 0)
                wc = nlev >= 3 ? F2(m->ground_q, nc, i, 0) * F2(m->wcf, nc, i, 0) +
                                     F2(m->ground_q, nc, i, 1) * F2(m->wcf, nc, i, 1) +
                                     F2(m->ground_q, nc, i, 2) * F2(m->wcf, nc, i, 2)
                               : F2(m->wcf, nc, i, 0);
            else
                wc = 0.0;
            F2(m->wch, nc, i, j) = wc;
        }
}

/* contravariant_w_kernel! (CGridOperators.jl:211-214) */
static void contravariant_w(const mpas_oracle_model* m, const double* w) {
    const int64_t nc = m->nc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j <= m->nlev; ++j)
        for (int64_t i = 0; i < nc; ++i) F2(m->what, nc, i, j) = F2(w, nc, i, j) - F2(m->wch, nc, i, j);
}

/* exner_deviation_centred_kernel! (CGridOperators.jl:93-100) */
static void exner_deviation(const mpas_oracle_model* m) {
    const int64_t nc = m->nc, nlev = m->nlev;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nlev; ++k)
        for (int64_t i = 0; i < nc; ++i) {
            const int64_t kl = k > 0 ? k - 1 : 0, ku = k < nlev - 1 ? k + 1 : nlev - 1;
            F2(m->pip, nc, i, k) = F2(m->exner, nc, i, k) - F2(m->pi_ref, nc, i, k);
            F2(m->dpidz, nc, i, k) = ((F2(m->exner, nc, i, ku) - F2(m->pi_ref, nc, i, ku)) -
                                      (F2(m->exner, nc, i, kl) - F2(m->pi_ref, nc, i, kl))) /
                                     (F2(m->z_full, nc, i, ku) - F2(m->z_full, nc, i, kl));
        }
}

/* w_concorr_tendency_kernel! (Operators.jl:147-152) */
static void w_concorr_tendency(const mpas_oracle_model* m, const double* w) {
    const int64_t nc = m->nc, nlev = m->nlev;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 1; j <= nlev - 1; ++j)
        for (int64_t i = 0; i < nc; ++i)
            F2(m->tw, nc, i, j) += F2(m->wch, nc, i, j) * (F2(w, nc, i, j + 1) - F2(w, nc, i, j - 1)) /
                                   (F2(m->dzh, nc, i, j) + F2(m->dzh, nc, i, j + 1));
}

/* edge_w (Tendencies.jl:4-11) at interface j */
static double edge_w(const mpas_oracle_model* m, const double* w, const double* rho, int64_t e, int64_t l, int64_t r,
                     int64_t j, int massweighted) {
    const int64_t nc = m->nc;
    const double rhl = massweighted ? F2(m->wlow, nc, l, j) * F2(rho, nc, l, j - 1) +
                                          (1 - F2(m->wlow, nc, l, j)) * F2(rho, nc, l, j)
                                    : 1.0;
    const double rhr = massweighted ? F2(m->wlow, nc, r, j) * F2(rho, nc, r, j - 1) +
                                          (1 - F2(m->wlow, nc, r, j)) * F2(rho, nc, r, j)
                                    : 1.0;
    const double w1 = AT(m->edge_weight, 2, 0, e), w2 = AT(m->edge_weight, 2, 1, e);
    const double M = w1 * rhl * F2(w, nc, l, j) + w2 * rhr * F2(w, nc, r, j);
    return M / (w1 * rhl + w2 * rhr);
}

static void vn_tendency(const mpas_oracle_model* m, const double* vn, const double* w, const double* rho,
                        double pressure) {
    const int64_t ne = m->ne, nc = m->nc, nlev = m->nlev, KT = m->kt;
    const int mw = m->vertical == 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nlev; ++k)
        for (int64_t e = 0; e < ne; ++e) {
            const int64_t l = AT(m->edge_cells, 2, 0, e), r = AT(m->edge_cells, 2, 1, e);
            double s = 0.0;
            for (int64_t j = 0; j < KT; ++j) {
                const int64_t a = m->trisk_index[e + ne * (j + KT * 0)];
                const int64_t b = m->trisk_index[e + ne * (j + KT * 1)];
                const int64_t c = m->trisk_index[e + ne * (j + KT * 2)];
                s += m->trisk[e + ne * j] * F2(m->Fs, ne, a, k) * (F2(m->qe, ne, b, k) + F2(m->qe, ne, c, k));
            }
            const double w1 = AT(m->edge_weight, 2, 0, e), w2 = AT(m->edge_weight, 2, 1, e);
            const double te = w1 * F2(m->theta, nc, l, k) + w2 * F2(m->theta, nc, r, k);
            double t = s - m->inverse_dual[e] * (F2(m->K, nc, r, k) - F2(m->K, nc, l, k));
            if (pressure > 0) {
                /* pressure_gradient (Tendencies.jl:19-30): pi' over orography */
                const double* pe = m->orography ? m->pip : m->exner;
                double dpi;
                if (m->orography == 2) {
                    /* height_difference (CGridOperators.jl:194-200) */
                    const int64_t z = 2 * (e + ne * k);
                    const int64_t jl = m->zlev[z], jr = m->zlev[z + 1];
                    const double tl = m->zwt[z], tr = m->zwt[z + 1];
                    dpi = (F2(pe, nc, r, jr) + tr * (F2(pe, nc, r, jr + 1) - F2(pe, nc, r, jr)) - F2(pe, nc, l, jl) -
                           tl * (F2(pe, nc, l, jl + 1) - F2(pe, nc, l, jl))) *
                          m->inverse_dual[e];
                } else {
                    dpi = (F2(pe, nc, r, k) - F2(pe, nc, l, k)) * m->inverse_dual[e];
                    if (m->orography == 1)
                        dpi -= F2(m->ddxn, ne, e, k) * (w1 * F2(m->dpidz, nc, l, k) + w2 * F2(m->dpidz, nc, r, k));
                }
                t -= pressure * (Cpd * te * dpi);
            }
            if (m->vertical > 0) {
                const double Wl = k > 0 ? edge_w(m, w, rho, e, l, r, k, mw) : 0.0;
                const double Wu = k < nlev - 1 ? edge_w(m, w, rho, e, l, r, k + 1, mw) : 0.0;
                F2(m->tvn, ne, e, k) = t - (F2(m->edge_dzh, ne, e, k) * F2(m->eta, ne, e, k) * Wl +
                                            F2(m->edge_dzh, ne, e, k + 1) * F2(m->eta, ne, e, k + 1) * Wu) /
                                               (2 * F2(m->edge_dz, ne, e, k));
            } else {
                const double we = (F2(w, nc, l, k) + F2(w, nc, r, k) + F2(w, nc, l, k + 1) + F2(w, nc, r, k + 1)) / 4;
                const double vlow = k > 0 ? (F2(vn, ne, e, k - 1) + F2(vn, ne, e, k)) / 2 : F2(vn, ne, e, k);
                const double vup = k < nlev - 1 ? (F2(vn, ne, e, k) + F2(vn, ne, e, k + 1)) / 2 : F2(vn, ne, e, k);
                F2(m->tvn, ne, e, k) = t - we * (vup - vlow) / F2(m->edge_dz, ne, e, k);
            }
        }
}

void mpas_oracle_slow_tendencies(const mpas_oracle_model* m, const double* vn, const double* w, const double* rho,
                                 const double* rhotheta, double pressure) {
    diagnose(m, vn, rho, rhotheta);
    if (m->orography) {
        concorr_cell(m, vn);
        concorr_interface(m);
        contravariant_w(m, w);
        if (pressure > 0) exner_deviation(m);
    }
    if (m->cell_vertices) {
        vertex_kinetic(m, vn);
        cell_from_vertex(m);
    }
    vertex_pv(m, vn, rho);
    edge_pv(m);
    edge_density(m, vn, rho);
    if (m->vertical > 0) {
        if (!m->five_band) w_kinetic(m, w);
        eta_kernel(m, vn, w);
        w_tendency_vi(m, vn, m->vertical == 2);
        if (m->five_band) w_kinetic(m, w);
    } else {
        w_tendency(m, vn, w);
    }
    if (m->orography) w_concorr_tendency(m, w);
    vn_tendency(m, vn, m->orography ? m->what : w, rho, pressure);
}

int mpas_oracle_kernel(const mpas_oracle_model* m, int id, const double* vn, const double* w, const double* rho,
                       const double* rhotheta, double pressure) {
    switch (id) {
        case 0: diagnose(m, vn, rho, rhotheta); break;
        case 1: vertex_kinetic(m, vn); break;
        case 2: cell_from_vertex(m); break;
        case 3: w_kinetic(m, w); break;
        case 4: vertex_pv(m, vn, rho); break;
        case 5: edge_pv(m); break;
        case 6: edge_density(m, vn, rho); break;
        case 7: eta_kernel(m, vn, w); break;
        case 8: w_tendency_vi(m, vn, m->vertical == 2); break;
        case 9: w_tendency(m, vn, w); break;
        case 10: vn_tendency(m, vn, w, rho, pressure); break;
        case 11: concorr_cell(m, vn); break;
        case 12: concorr_interface(m); break;
        case 13: contravariant_w(m, w); break;
        case 14: exner_deviation(m); break;
        case 15: w_concorr_tendency(m, w); break;
        default: return -1;
    }
    return 0;
}
