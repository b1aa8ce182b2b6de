/*
 * icon_oracle.c -- TEST INFRASTRUCTURE ONLY (see icon_oracle.h).
 *
 * Restates, kernel for kernel, src/FiniteVolumes/ICONShallowWater.jl of the reference in C:
 *   icon_edge_kernel!        :149-159     icon_momentum_kernel!   :161-171
 *   edge_gradient_kernel!    :173-176     kinetic_energy!         :184-189
 *   tendency!                :191-207     divergence_damping!     :209-217
 * with the shared kernels of src/FiniteVolumes/TRiSKShallowWater.jl (continuity_kernel!
 * :298-305, edge_sum_kernel! :307-314, vertex_pv_kernel! :246-259), the diagnostics
 * (:397-419, Neumaier-compensated here) and rk4_step! (src/TimeSteppers/RungeKutta.jl:18-28).
 * Each kernel is a parallel loop with an implicit barrier after it, as run! does on the
 * CPU backend.
 */
#include "icon_oracle.h"

#include <stdlib.h>

#define AT(tab, K, k, j) (tab)[(k) + (int64_t)(j) * (K)]

/* y_j = sum_k w[k, j] x[idx[k, j]] */
static void edge_sum(int64_t n, int64_t K, const int64_t* idx, const double* w, const double* x, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int64_t k = 0; k < K; ++k) s += AT(w, K, k, j) * x[AT(idx, K, k, j)];
        out[j] = s;
    }
}

/* dh_i = -sum_k divergence[k, i] F[cell_edges[k, i]] */
static void continuity(const icon_oracle_model* m, const double* F, double* dh) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m->nc; ++i) {
        double s = 0.0;
        for (int64_t k = 0; k < 3; ++k) s += AT(m->divergence, 3, k, i) * F[AT(m->cell_edges, 3, k, i)];
        dh[i] = -s;
    }
}

static void edge_gradient(const icon_oracle_model* m, const double* D, double scale, double* G) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->ne; ++e)
        G[e] = scale * (D[AT(m->edge_cells, 2, 1, e)] - D[AT(m->edge_cells, 2, 0, e)]) * m->inverse_dual[e];
}

void icon_oracle_kinetic_energy(const icon_oracle_model* m, const double* u, const double* h) {
    const int64_t KR = m->kr;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->ne; ++e) {
        double s = 0.0;
        for (int64_t k = 0; k < KR; ++k) s += AT(m->rbf, KR, k, e) * u[AT(m->rbf_edges, KR, k, e)];
        m->v[e] = s;
        m->Ke[e] = (u[e] * u[e] + s * s) / 2;
        m->F[e] = (AT(m->edge_height, 2, 0, e) * h[AT(m->edge_cells, 2, 0, e)] +
                   AT(m->edge_height, 2, 1, e) * h[AT(m->edge_cells, 2, 1, e)]) * u[e];
    }
    edge_sum(m->nc, 3, m->cell_edges, m->bilinear, m->Ke, m->K);
}

void icon_oracle_vertex_pv(const icon_oracle_model* m, const double* u, const double* h) {
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

void icon_oracle_tendency(const icon_oracle_model* m, const double* y, double* dy) {
    const int64_t ne = m->ne;
    const double* u = y;
    const double* h = y + ne;
    double* du = dy;
    double* dh = dy + ne;
    icon_oracle_kinetic_energy(m, u, h);
    edge_sum(m->nv, m->kv, m->vertex_edges, m->curl, u, m->zv);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t l = AT(m->edge_cells, 2, 0, e), r = AT(m->edge_cells, 2, 1, e);
        const double zeta = (m->zv[AT(m->edge_vertices, 2, 0, e)] + m->zv[AT(m->edge_vertices, 2, 1, e)]) / 2;
        const double g1 = AT(m->gradient, 2, 0, e), g2 = AT(m->gradient, 2, 1, e);
        const double grad_K = m->Ke[e] * (g1 - g2) + g2 * m->K[r] - g1 * m->K[l];
        du[e] = (m->edge_coriolis[e] + zeta) * m->v[e] - grad_K - m->gravity * (h[r] - h[l]) * m->inverse_dual[e];
    }
    continuity(m, m->F, dh);
    if (m->damping) {
        continuity(m, u, m->D);                   /* -div u */
        edge_gradient(m, m->D, 1.0, m->G);        /* -grad div u */
        continuity(m, m->G, m->D);                /* div grad div u */
        edge_gradient(m, m->D, -1.0, m->G);
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) du[e] = du[e] + m->damping[e] * m->G[e];
    }
}

void icon_oracle_rk4_steps(const icon_oracle_model* m, double* y, double dt, int64_t nsteps, double* k1, double* k2,
                           double* k3, double* k4, double* stage) {
    const int64_t n = m->ne + m->nc;
    for (int64_t step = 0; step < nsteps; ++step) {
        icon_oracle_tendency(m, y, k1);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt / 2 * k1[i];
        icon_oracle_tendency(m, stage, k2);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt / 2 * k2[i];
        icon_oracle_tendency(m, stage, k3);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + dt * k3[i];
        icon_oracle_tendency(m, stage, k4);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) y[i] = y[i] + dt / 6 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    }
}

/* Neumaier-compensated sum in fixed blocks (independent of the thread count). */
static void neumaier(double* s, double* c, double x) {
    double t = *s + x;
    if (__builtin_fabs(*s) >= __builtin_fabs(x))
        *c += (*s - t) + x;
    else
        *c += (x - t) + *s;
    *s = t;
}

typedef double (*term_fn)(const icon_oracle_model*, const double*, int64_t);

static double comp_sum(const icon_oracle_model* m, const double* y, int64_t n, term_fn f) {
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

static double mass_term(const icon_oracle_model* m, const double* y, int64_t i) {
    return m->cell_areas[i] * y[m->ne + i];
}
static double energy_term(const icon_oracle_model* m, const double* y, int64_t i) {
    double h = y[m->ne + i];
    return m->cell_areas[i] * h * (m->K[i] + m->gravity * h / 2);
}
static double enstrophy_term(const icon_oracle_model* m, const double* y, int64_t v) {
    (void)y;
    return m->vertex_areas[v] * m->hv[v] * (m->qv[v] * m->qv[v]);
}

double icon_oracle_total_mass(const icon_oracle_model* m, const double* y) { return comp_sum(m, y, m->nc, mass_term); }

double icon_oracle_total_energy(const icon_oracle_model* m, const double* y) {
    icon_oracle_kinetic_energy(m, y, y + m->ne);
    return comp_sum(m, y, m->nc, energy_term);
}

double icon_oracle_potential_enstrophy(const icon_oracle_model* m, const double* y) {
    icon_oracle_vertex_pv(m, y, y + m->ne);
    return comp_sum(m, y, m->nv, enstrophy_term) / 2;
}
