/*
 * trisk_oracle.c -- TEST INFRASTRUCTURE ONLY (see trisk_oracle.h).
 *
 * Restates, kernel for kernel, src/FiniteVolumes/TRiSKShallowWater.jl and
 * src/TimeSteppers/RungeKutta.jl of the reference in C:
 *   kinetic_energy_kernel!         :202-210     vertex_kinetic_energy_kernel!  :212-219
 *   cell_from_vertex_kernel!       :222-230     mass_flux_kernel!              :232-235
 *   mixed_mass_flux_kernel!        :238-244     vertex_pv_kernel!              :246-259
 *   edge_pv_kernel!                :261-264     rhombus_pv_kernel!             :266-273
 *   momentum_kernel!               :275-284     gassmann_momentum_kernel!      :286-296
 *   continuity_kernel!             :298-305     edge_sum_kernel!               :307-314
 *   tendency!                      :379-394     total_mass / total_energy /
 *                                               potential_enstrophy            :397-419
 *   rk4_step!                      RungeKutta.jl:18-28 (one pass per broadcast)
 * Each kernel is a parallel loop with an implicit barrier after it, as run! does on the
 * CPU backend (:322-327).  The diagnostics sum with Neumaier compensation (the reference
 * uses BLAS dot / pairwise sum, whose order is not reproducible).
 */
#include "trisk_oracle.h"

#include <omp.h>
#include <stdlib.h>

#define AT(tab, K, k, j) (tab)[(k) + (int64_t)(j) * (K)]

int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

void oracle_kinetic_energy(const oracle_model* m, const double* u, const double* h) {
    const int64_t KC = m->kc, KV = m->kv;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m->nc; ++i) {
        double s = 0.0;
        for (int64_t k = 0; k < KC; ++k) {
            double ue = u[AT(m->cell_edges, KC, k, i)];
            s += AT(m->kinetic, KC, k, i) * (ue * ue);
        }
        m->K[i] = s;
        m->B[i] = s + m->gravity * h[i];
    }
    if (m->cell_vertices) {
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < m->nv; ++v) {
            double s = 0.0;
            for (int64_t k = 0; k < KV; ++k) {
                double ue = u[AT(m->vertex_edges, KV, k, v)];
                s += AT(m->vk_weights, KV, k, v) * (ue * ue);
            }
            m->Kv[v] = s;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < m->nc; ++i) {
            double s = 0.0;
            for (int64_t k = 0; k < KC; ++k) s += AT(m->vk_kites, KC, k, i) * m->Kv[AT(m->cell_vertices, KC, k, i)];
            m->K[i] += s;
            m->B[i] += s;
        }
    }
}

void oracle_vertex_pv(const oracle_model* m, const double* u, const double* h) {
    const int64_t KV = m->kv;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < m->nv; ++v) {
        double z = 0.0, eta = 0.0;
        for (int64_t k = 0; k < KV; ++k) z += AT(m->curl, KV, k, v) * u[AT(m->vertex_edges, KV, k, v)];
        for (int64_t k = 0; k < KV; ++k) eta += AT(m->vertex_height, KV, k, v) * h[AT(m->vertex_cells, KV, k, v)];
        m->zv[v] = z;
        m->hv[v] = eta;
        m->qv[v] = (z + m->coriolis[v]) / eta;
    }
}

void oracle_mass_flux(const oracle_model* m, const double* u, const double* h) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->ne; ++e) {
        double he = AT(m->edge_height, 2, 0, e) * h[AT(m->edge_cells, 2, 0, e)] +
                    AT(m->edge_height, 2, 1, e) * h[AT(m->edge_cells, 2, 1, e)];
        if (m->vk_height)
            he = he + AT(m->vk_height, 2, 0, e) * m->hv[AT(m->edge_vertices, 2, 0, e)] +
                 AT(m->vk_height, 2, 1, e) * m->hv[AT(m->edge_vertices, 2, 1, e)];
        m->F[e] = he * u[e];
    }
}

static void pv_flux(const oracle_model* m, double* du) {
    if (!m->g_edges) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m->ne; ++e)
            m->qe[e] = (m->qv[AT(m->edge_vertices, 2, 0, e)] + m->qv[AT(m->edge_vertices, 2, 1, e)]) / 2;
        const int64_t KT = m->kt;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m->ne; ++e) {
            double s = 0.0;
            for (int64_t k = 0; k < KT; ++k) {
                int64_t f = AT(m->trisk_edges, KT, k, e);
                s += AT(m->trisk, KT, k, e) * m->F[f] * (m->qe[e] + m->qe[f]);
            }
            du[e] = s / 2 - (m->B[AT(m->edge_cells, 2, 1, e)] - m->B[AT(m->edge_cells, 2, 0, e)]) * m->inverse_dual[e];
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m->ne; ++e) {
            int64_t v1 = AT(m->edge_vertices, 2, 0, e), v2 = AT(m->edge_vertices, 2, 1, e);
            double r1 = AT(m->g_rhombus, 2, 0, e), r2 = AT(m->g_rhombus, 2, 1, e);
            double z = r1 * m->zv[v1] + r2 * m->zv[v2];
            double eta = r1 * m->hv[v1] + r2 * m->hv[v2];
            m->qe[e] = (z + m->g_coriolis[e]) / eta;
        }
        const int64_t KG = m->kg;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m->ne; ++e) {
            double s = 0.0;
            for (int64_t k = 0; k < KG; ++k) {
                int64_t f = AT(m->g_edges, KG, k, e);
                double q = m->g_mean ? (m->qe[e] + m->qe[f]) / 2 : m->qe[AT(m->g_pv_edges, KG, k, e)];
                s += AT(m->g_weights, KG, k, e) * m->F[f] * q;
            }
            du[e] = s - (m->B[AT(m->edge_cells, 2, 1, e)] - m->B[AT(m->edge_cells, 2, 0, e)]) * m->inverse_dual[e];
        }
    }
}

static void continuity(const oracle_model* m, double* dh) {
    const int64_t KC = m->kc;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m->nc; ++i) {
        double s = 0.0;
        for (int64_t k = 0; k < KC; ++k) s += AT(m->divergence, KC, k, i) * m->F[AT(m->cell_edges, KC, k, i)];
        dh[i] = -s;
    }
}

void oracle_tendency(const oracle_model* m, const double* y, double* dy) {
    const double* u = y;
    const double* h = y + m->ne;
    oracle_kinetic_energy(m, u, h);
    oracle_vertex_pv(m, u, h);
    oracle_mass_flux(m, u, h);
    pv_flux(m, dy);
    continuity(m, dy + m->ne);
}

void oracle_rk4_step(const oracle_model* m, double* y, double dt, double* k1, double* k2, double* k3, double* k4,
                     double* stage) {
    const int64_t n = m->ne + m->nc;
    oracle_tendency(m, y, k1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt / 2 * k1[i];
    oracle_tendency(m, stage, k2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt / 2 * k2[i];
    oracle_tendency(m, stage, k3);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt * k3[i];
    oracle_tendency(m, stage, k4);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] = y[i] + dt / 6 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
}

void oracle_rk4_steps(const oracle_model* m, double* y, double dt, int64_t nsteps, double* k1, double* k2,
                      double* k3, double* k4, double* stage) {
    for (int64_t s = 0; s < nsteps; ++s) oracle_rk4_step(m, y, dt, k1, k2, k3, k4, stage);
}

/* Neumaier-compensated sum of n terms, split in fixed blocks so that the result does
 * not depend on the thread count. */
typedef double (*term_fn)(const oracle_model*, const double*, int64_t);

static void neumaier(double* s, double* c, double x) {
    double t = *s + x;
    if (__builtin_fabs(*s) >= __builtin_fabs(x))
        *c += (*s - t) + x;
    else
        *c += (x - t) + *s;
    *s = t;
}

static double comp_sum(const oracle_model* m, const double* y, int64_t n, term_fn f) {
    enum { B = 8192 };
    int64_t nb = (n + B - 1) / B;
    double* ps = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1) * 2);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        double s = 0.0, c = 0.0;
        int64_t hi = (b + 1) * B < n ? (b + 1) * B : n;
        for (int64_t i = b * B; i < hi; ++i) neumaier(&s, &c, f(m, y, i));
        ps[2 * b] = s;
        ps[2 * b + 1] = c;
    }
    double s = 0.0, c = 0.0;
    for (int64_t b = 0; b < nb; ++b) {
        neumaier(&s, &c, ps[2 * b]);
        c += ps[2 * b + 1];
    }
    free(ps);
    return s + c;
}

static double mass_term(const oracle_model* m, const double* y, int64_t i) { return m->cell_areas[i] * y[m->ne + i]; }
static double energy_term(const oracle_model* m, const double* y, int64_t i) {
    double h = y[m->ne + i];
    return m->cell_areas[i] * h * (m->K[i] + m->gravity * h / 2);
}
static double enstrophy_term(const oracle_model* m, const double* y, int64_t v) {
    (void)y;
    return m->vertex_areas[v] * m->hv[v] * (m->qv[v] * m->qv[v]);
}

double oracle_total_mass(const oracle_model* m, const double* y) { return comp_sum(m, y, m->nc, mass_term); }

double oracle_total_energy(const oracle_model* m, const double* y) {
    oracle_kinetic_energy(m, y, y + m->ne);
    return comp_sum(m, y, m->nc, energy_term);
}

double oracle_potential_enstrophy(const oracle_model* m, const double* y) {
    oracle_vertex_pv(m, y, y + m->ne);
    return comp_sum(m, y, m->nv, enstrophy_term) / 2;
}

void oracle_edge_sum(int64_t n, int64_t k, const int64_t* edges, const double* weights, const double* x,
                     double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int64_t q = 0; q < k; ++q) s += AT(weights, k, q, j) * x[AT(edges, k, q, j)];
        out[j] = s;
    }
}

/* ---------------------------------------------------------------------------------
 * Adjoint (vector-Jacobian product) of the tendency and of rk4_step!, for every PV flux
 * (TRiSK, Gassmann, Gassmann mean) and kinetic-energy form (cell, mixed, vertex).  The
 * reference has no adjoint; this is the transpose of the linearisation of oracle_tendency
 * above (TRiSKShallowWater.jl:202-305), written as plain scatter-adds over the reference's
 * tables in one thread (an implementation independent of the device's gather-only
 * kernels).  It is pinned by finite differences in the tests. */

int oracle_tendency_vjp(const oracle_model* m, const double* y, const double* lam, double* ybar) {
    const int64_t ne = m->ne, nc = m->nc, nv = m->nv, KC = m->kc, KV = m->kv, KT = m->kt, KG = m->kg;
    const double* u = y;
    const double* h = y + ne;
    const double* lu = lam;
    const double* lh = lam + ne;
    double* ub = ybar;
    double* hb = ybar + ne;
    const int vk = m->cell_vertices != NULL, gm = m->g_edges != NULL;
    /* forward intermediates: K_v, zeta_v, h_v, q_v, F and the edge PV */
    oracle_kinetic_energy(m, u, h);
    oracle_vertex_pv(m, u, h);
    oracle_mass_flux(m, u, h);
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t v1 = AT(m->edge_vertices, 2, 0, e), v2 = AT(m->edge_vertices, 2, 1, e);
        if (gm) {
            const double r1 = AT(m->g_rhombus, 2, 0, e), r2 = AT(m->g_rhombus, 2, 1, e);
            m->qe[e] = (r1 * m->zv[v1] + r2 * m->zv[v2] + m->g_coriolis[e]) / (r1 * m->hv[v1] + r2 * m->hv[v2]);
        } else {
            m->qe[e] = (m->qv[v1] + m->qv[v2]) / 2;
        }
    }
    double* Fb = (double*)calloc((size_t)ne, sizeof(double));
    double* qeb = (double*)calloc((size_t)ne, sizeof(double));
    double* Bb = (double*)calloc((size_t)nc, sizeof(double));
    double* zb = (double*)calloc((size_t)nv, sizeof(double));
    double* hvb = (double*)calloc((size_t)nv, sizeof(double));
    double* Kvb = (double*)calloc((size_t)nv, sizeof(double));
    for (int64_t j = 0; j < ne + nc; ++j) ybar[j] = 0.0;
    /* du_e = sum_k w F_f q_k - (B_c2 - B_c1) / d_e, q_k the mean (qe_e + qe_f)/2 (TRiSK, Gassmann
     * mean) or the PV of the pv edge (Gassmann) */
    for (int64_t e = 0; e < ne; ++e) {
        if (!gm) {
            for (int64_t k = 0; k < KT; ++k) {
                const int64_t f = AT(m->trisk_edges, KT, k, e);
                const double w = AT(m->trisk, KT, k, e);
                Fb[f] += 0.5 * lu[e] * w * (m->qe[e] + m->qe[f]);
                qeb[e] += 0.5 * lu[e] * w * m->F[f];
                qeb[f] += 0.5 * lu[e] * w * m->F[f];
            }
        } else {
            for (int64_t k = 0; k < KG; ++k) {
                const int64_t f = AT(m->g_edges, KG, k, e);
                const double w = AT(m->g_weights, KG, k, e);
                if (m->g_mean) {
                    Fb[f] += lu[e] * w * (m->qe[e] + m->qe[f]) / 2;
                    qeb[e] += lu[e] * w * m->F[f] / 2;
                    qeb[f] += lu[e] * w * m->F[f] / 2;
                } else {
                    const int64_t pe = AT(m->g_pv_edges, KG, k, e);
                    Fb[f] += lu[e] * w * m->qe[pe];
                    qeb[pe] += lu[e] * w * m->F[f];
                }
            }
        }
        Bb[AT(m->edge_cells, 2, 1, e)] -= lu[e] * m->inverse_dual[e];
        Bb[AT(m->edge_cells, 2, 0, e)] += lu[e] * m->inverse_dual[e];
    }
    /* dh_i = -sum_k div F */
    for (int64_t i = 0; i < nc; ++i)
        for (int64_t k = 0; k < KC; ++k) Fb[AT(m->cell_edges, KC, k, i)] -= AT(m->divergence, KC, k, i) * lh[i];
    /* the edge PV: the mean of q_v = (zeta_v + f_v) / h_v, or the rhombus PV
     * (r1 zeta_1 + r2 zeta_2 + f_e) / (r1 h_1 + r2 h_2) */
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t v1 = AT(m->edge_vertices, 2, 0, e), v2 = AT(m->edge_vertices, 2, 1, e);
        if (gm) {
            const double r1 = AT(m->g_rhombus, 2, 0, e), r2 = AT(m->g_rhombus, 2, 1, e);
            const double eta = r1 * m->hv[v1] + r2 * m->hv[v2];
            const double zbe = qeb[e] / eta, ebe = -qeb[e] * m->qe[e] / eta;
            zb[v1] += r1 * zbe;
            zb[v2] += r2 * zbe;
            hvb[v1] += r1 * ebe;
            hvb[v2] += r2 * ebe;
        } else {
            const double qb = qeb[e] / 2;
            zb[v1] += qb / m->hv[v1];
            zb[v2] += qb / m->hv[v2];
            hvb[v1] -= qb * m->qv[v1] / m->hv[v1];
            hvb[v2] -= qb * m->qv[v2] / m->hv[v2];
        }
    }
    /* F_e = (eh_1 h_c1 + eh_2 h_c2 [+ vkh_1 h_v1 + vkh_2 h_v2]) u_e */
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t c0 = AT(m->edge_cells, 2, 0, e), c1 = AT(m->edge_cells, 2, 1, e);
        const double e0 = AT(m->edge_height, 2, 0, e), e1 = AT(m->edge_height, 2, 1, e);
        double he = e0 * h[c0] + e1 * h[c1];
        hb[c0] += Fb[e] * e0 * u[e];
        hb[c1] += Fb[e] * e1 * u[e];
        if (vk) {
            const int64_t v0 = AT(m->edge_vertices, 2, 0, e), v1 = AT(m->edge_vertices, 2, 1, e);
            const double k0 = AT(m->vk_height, 2, 0, e), k1 = AT(m->vk_height, 2, 1, e);
            he = he + k0 * m->hv[v0] + k1 * m->hv[v1];
            hvb[v0] += Fb[e] * k0 * u[e];
            hvb[v1] += Fb[e] * k1 * u[e];
        }
        ub[e] += Fb[e] * he;
    }
    /* B_i = sum_k kin u^2 [+ sum_k kites K_v] + g h_i, K_v = sum_k vkw u^2 */
    for (int64_t i = 0; i < nc; ++i) {
        hb[i] += m->gravity * Bb[i];
        for (int64_t k = 0; k < KC; ++k) {
            const int64_t e = AT(m->cell_edges, KC, k, i);
            ub[e] += 2 * AT(m->kinetic, KC, k, i) * u[e] * Bb[i];
            if (vk) Kvb[AT(m->cell_vertices, KC, k, i)] += AT(m->vk_kites, KC, k, i) * Bb[i];
        }
    }
    /* zeta_v = sum curl u, h_v = sum vh h [, K_v] */
    for (int64_t v = 0; v < nv; ++v)
        for (int64_t k = 0; k < KV; ++k) {
            const int64_t e = AT(m->vertex_edges, KV, k, v);
            ub[e] += AT(m->curl, KV, k, v) * zb[v];
            if (vk) ub[e] += 2 * AT(m->vk_weights, KV, k, v) * u[e] * Kvb[v];
            hb[AT(m->vertex_cells, KV, k, v)] += AT(m->vertex_height, KV, k, v) * hvb[v];
        }
    free(Fb);
    free(qeb);
    free(Bb);
    free(zb);
    free(hvb);
    free(Kvb);
    return 0;
}

/* ybar (in: the cotangent of y after nsteps rk4_step! from y0; out: the cotangent of y0).
 * Every step's state is stored (small meshes only); each step's stages are recomputed. */
int oracle_rk4_adjoint(const oracle_model* m, const double* y0, double dt, int64_t nsteps, double* ybar) {
    const int64_t n = m->ne + m->nc;
    double* ys = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(nsteps > 0 ? nsteps : 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)n * 11);
    double *k1 = w, *k2 = w + n, *k3 = w + 2 * n, *k4 = w + 3 * n, *st = w + 4 * n;
    double *s2 = w + 5 * n, *s3 = w + 6 * n, *s4 = w + 7 * n, *kb = w + 8 * n, *sb = w + 9 * n, *acc = w + 10 * n;
    for (int64_t j = 0; j < n; ++j) st[j] = y0[j];
    for (int64_t s = 0; s < nsteps; ++s) {
        for (int64_t j = 0; j < n; ++j) ys[s * n + j] = st[j];
        oracle_rk4_step(m, st, dt, k1, k2, k3, k4, s4);
    }
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        const double* y = ys + s * n;
        oracle_tendency(m, y, k1);
        for (int64_t j = 0; j < n; ++j) s2[j] = y[j] + dt / 2 * k1[j];
        oracle_tendency(m, s2, k2);
        for (int64_t j = 0; j < n; ++j) s3[j] = y[j] + dt / 2 * k2[j];
        oracle_tendency(m, s3, k3);
        for (int64_t j = 0; j < n; ++j) s4[j] = y[j] + dt * k3[j];
        /* y' = y + dt/6 (k1 + 2 k2 + 2 k3 + k4); k4 = f(s4), k3 = f(s3), ... */
        for (int64_t j = 0; j < n; ++j) kb[j] = dt / 6 * ybar[j];
        oracle_tendency_vjp(m, s4, kb, sb);
        for (int64_t j = 0; j < n; ++j) {
            acc[j] = ybar[j] + sb[j];
            kb[j] = dt / 3 * ybar[j] + dt * sb[j];
        }
        oracle_tendency_vjp(m, s3, kb, sb);
        for (int64_t j = 0; j < n; ++j) {
            acc[j] += sb[j];
            kb[j] = dt / 3 * ybar[j] + dt / 2 * sb[j];
        }
        oracle_tendency_vjp(m, s2, kb, sb);
        for (int64_t j = 0; j < n; ++j) {
            acc[j] += sb[j];
            kb[j] = dt / 6 * ybar[j] + dt / 2 * sb[j];
        }
        oracle_tendency_vjp(m, y, kb, sb);
        for (int64_t j = 0; j < n; ++j) ybar[j] = acc[j] + sb[j];
    }
    free(ys);
    free(w);
    return 0;
}
