// TRiSK geometry tables, Hilbert keys and the synthetic streamfunction.
//
// Restates:
//   TRiSKGrid(dm; centers)   src/FiniteVolumes/TRiSKGrid.jl:76-162
//   vertex_edge_areas        src/FiniteVolumes/TRiSKGrid.jl:179-191
//   gassmann_coefficients    src/FiniteVolumes/TRiSKGrid.jl:214-251
//   hilbert_key / hilbert_permutation  src/Topology/Ordering.jl:10-41,85-92
// Tables are k x n column-major with 0-based indices; padded slots hold index 0 and
// weight 0 (padded(), TRiSKGrid.jl:67-74).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "../../../include/dmplex_host.h"
#include "common.hpp"

using namespace dmx;

namespace {

inline int64_t cone_size(const dmx_plex* dm, int64_t p) {
    return dm->cone_offsets[p + 1] - dm->cone_offsets[p];
}

// get_edge_cells (Plex.jl:210-218) on a point e; -1 for a missing side.
inline void edge_cells_of(const dmx_plex* dm, int64_t e, int64_t& left, int64_t& right) {
    left = right = -1;
    for (int64_t s = dm->support_offsets[e]; s < dm->support_offsets[e + 1]; ++s) {
        int64_t c = dm->supports[s];
        for (int64_t k = dm->cone_offsets[c]; k < dm->cone_offsets[c + 1]; ++k)
            if (dm->cones[k] == e) {
                (dm->orientations[k] == 0 ? left : right) = c;
                break;
            }
    }
}

}  // namespace

extern "C" {

int dmx_trisk_grid_widths(const dmx_plex* dm, int64_t widths[3]) {
    const int64_t nc = dm->nc, ne = dm->ne, nv = dm->nv;
    int64_t kc = 0, kv = 0, kt = 0;
    for (int64_t c = 0; c < nc; ++c) kc = std::max(kc, cone_size(dm, c));
    for (int64_t v = 0; v < nv; ++v) {
        int64_t p = nc + ne + v;
        kv = std::max(kv, dm->support_offsets[p + 1] - dm->support_offsets[p]);
    }
    for (int64_t j = 0; j < ne; ++j) {
        int64_t e = nc + j, n = 0;
        for (int64_t s = dm->support_offsets[e]; s < dm->support_offsets[e + 1]; ++s)
            n += cone_size(dm, dm->supports[s]) - 1;
        kt = std::max(kt, n);
    }
    widths[0] = kc;
    widths[1] = kv;
    widths[2] = kt;
    return 0;
}

int dmx_trisk_grid(const dmx_plex* dm, const double* g, dmx_grid_tables* G) {
    const int64_t nc = dm->nc, ne = dm->ne, nv = dm->nv;
    const int64_t KC = G->kc, KV = G->kv, KT = G->kt;
    const double* x = dm->coordinates;
    auto vidx = [&](int64_t p) { return p - nc - ne; };

    // edges (TRiSKGrid.jl:85-106)
    int open = 0;
#pragma omp parallel for schedule(static) reduction(| : open)
    for (int64_t j = 0; j < ne; ++j) {
        int64_t e = nc + j, left, right;
        edge_cells_of(dm, e, left, right);
        if (left < 0 || right < 0) {
            open = 1;
            continue;
        }
        int64_t a = vidx(dm->cones[dm->cone_offsets[e]]), b = vidx(dm->cones[dm->cone_offsets[e] + 1]);
        G->edge_cells[2 * j] = left;
        G->edge_cells[2 * j + 1] = right;
        G->edge_vertices[2 * j] = a;
        G->edge_vertices[2 * j + 1] = b;
        V3 xa = load3(x, a), xb = load3(x, b), gl = load3(g, left), gr = load3(g, right);
        V3 m = cross(cross(xa, xb), cross(gl, gr));
        m = norm(m) > 1e-14 ? normalize(m) : normalize(xa + xb);
        if (dot(m, xa + xb) < 0) m = -m;
        store3(G->edge_points, j, m);
        V3 n = normalize(cross(xb, xa));
        store3(G->edge_normals, j, n);
        G->primal_lengths[j] = arc_length(xa, xb);
        G->dual_lengths[j] = arc_length(gl, gr);
        G->cell_distances[2 * j] = -sgn(dot(gl, n)) * arc_length(gl, m);
        G->cell_distances[2 * j + 1] = sgn(dot(gr, n)) * arc_length(gr, m);
    }
    if (open) return fail(-2, "TRiSKGrid needs a closed mesh");

    // cells: edges counter-clockwise with outward signs, corners with kites (:108-134)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        int64_t lo = dm->cone_offsets[i], n = cone_size(dm, i);
        V3 gi = load3(g, i);
        for (int64_t k = 0; k < KC; ++k) {
            int64_t q = k + i * KC;
            if (k < n) {
                int64_t e = dm->cones[lo + k];
                G->cell_edges[q] = e - nc;
                G->cell_signs[q] = dm->orientations[lo + k] == 0 ? 1.0 : -1.0;
                G->cell_vertices[q] = vidx(dm->cones[dm->cone_offsets[e] + (dm->orientations[lo + k] == 0 ? 0 : 1)]);
            } else {
                G->cell_edges[q] = 0;
                G->cell_signs[q] = 0.0;
                G->cell_vertices[q] = 0;
                G->cell_kites[q] = 0.0;
            }
        }
        double area = 0;
        for (int64_t k = 0; k < n; ++k) {
            int64_t q = k + i * KC;
            int64_t eprev = G->cell_edges[(k + n - 1) % n + i * KC], e = G->cell_edges[q];
            V3 xv = load3(x, G->cell_vertices[q]);
            double kite = signed_triangle_area(gi, load3(G->edge_points, eprev), xv) +
                          signed_triangle_area(gi, xv, load3(G->edge_points, e));
            G->cell_kites[q] = kite;
            area = k == 0 ? kite : area + kite;
        }
        G->cell_areas[i] = area;
    }

    // vertices: cells with kites in cell order, edges with circulation signs (:113-138)
    std::vector<int64_t> fill(nv, 0);
    for (int64_t q = 0; q < KV * nv; ++q) {
        G->vertex_cells[q] = 0;
        G->vertex_kites[q] = 0.0;
    }
    for (int64_t i = 0; i < nc; ++i) {
        int64_t n = cone_size(dm, i);
        for (int64_t k = 0; k < n; ++k) {
            int64_t v = G->cell_vertices[k + i * KC];
            int64_t s = fill[v]++;
            if (s >= KV) return fail(-1, "vertex with more cells than edges");
            G->vertex_cells[s + v * KV] = i;
            G->vertex_kites[s + v * KV] = G->cell_kites[k + i * KC];
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; ++v) {
        double a = 0;
        for (int64_t s = 0; s < fill[v]; ++s) a = s == 0 ? G->vertex_kites[v * KV] : a + G->vertex_kites[s + v * KV];
        G->vertex_areas[v] = a;
        int64_t p = nc + ne + v;
        int64_t lo = dm->support_offsets[p], n = dm->support_offsets[p + 1] - lo;
        V3 xv = load3(x, v);
        for (int64_t k = 0; k < KV; ++k) {
            if (k < n) {
                int64_t e = dm->supports[lo + k] - nc;
                G->vertex_edges[k + v * KV] = e;
                G->vertex_signs[k + v * KV] =
                    sgn(dot(load3(G->edge_normals, e), cross(xv, load3(G->edge_points, e))));
            } else {
                G->vertex_edges[k + v * KV] = 0;
                G->vertex_signs[k + v * KV] = 0.0;
            }
        }
    }

    // TRiSK weights (:140-155): the walk around cell i appends to edge e after the
    // entries of the lower-numbered cell of e, so each cell knows its offset.
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ne; ++j)
        for (int64_t k = 0; k < KT; ++k) {
            G->trisk_edges[k + j * KT] = 0;
            G->trisk_weights[k + j * KT] = 0.0;
        }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; ++i) {
        int64_t n = cone_size(dm, i);
        double R[64];
        for (int64_t k = 0; k < n; ++k) R[k] = G->cell_kites[k + i * KC] / G->cell_areas[i];
        for (int64_t k = 0; k < n; ++k) {
            int64_t e = G->cell_edges[k + i * KC];
            int64_t l = G->edge_cells[2 * e], r = G->edge_cells[2 * e + 1];
            int64_t other = l == i ? r : l;
            int64_t base = i < other ? 0 : cone_size(dm, other) - 1;
            double S = 0.0, sk = G->cell_signs[k + i * KC];
            for (int64_t s = 1; s <= n - 1; ++s) {
                int64_t k2 = (k + s) % n;
                S += R[k2];
                G->trisk_edges[base + s - 1 + e * KT] = G->cell_edges[k2 + i * KC];
                G->trisk_weights[base + s - 1 + e * KT] = sk * G->cell_signs[k2 + i * KC] * (0.5 - S);
            }
        }
    }
    return 0;
}

int dmx_vertex_edge_areas(int64_t nc, int64_t ne, int64_t nv, const dmx_grid_tables* G, const double* centers,
                          const double* vertices, double* by_vertex, double* by_edge) {
    const int64_t KC = G->kc, KV = G->kv;
    auto area = [&](int64_t e, int64_t v) {
        int64_t left = G->edge_cells[2 * e];
        int64_t n = 0;
        for (int64_t k = 0; k < KC; ++k) n += G->cell_signs[k + left * KC] != 0;
        int64_t j = 0;
        while (j < n && G->cell_edges[j + left * KC] != e) ++j;
        double orientation = G->cell_vertices[(j + 1) % n + left * KC] == v ? 1.0 : -1.0;
        return orientation * signed_triangle_area(load3(centers, left), load3(centers, G->edge_cells[2 * e + 1]),
                                                  load3(vertices, v));
    };
    (void)nc;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; ++v)
        for (int64_t k = 0; k < KV; ++k)
            by_vertex[k + v * KV] = G->vertex_signs[k + v * KV] == 0 ? 0.0 : area(G->vertex_edges[k + v * KV], v);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e)
        for (int s = 0; s < 2; ++s) by_edge[s + 2 * e] = area(e, G->edge_vertices[s + 2 * e]);
    return 0;
}

int dmx_gassmann_coefficients(int64_t nc, int64_t ne, int64_t nv, const dmx_grid_tables* G, const double* vertices,
                              int64_t* edges, int64_t* pv_edges, double* weights) {
    (void)nc;
    (void)nv;
    const int64_t KV = G->kv;
    const double *l = G->primal_lengths, *d = G->dual_lengths, *A = G->vertex_areas;
    auto to_vertex = [&](int64_t e, int64_t v) {
        return arc_length(load3(G->edge_points, e), load3(vertices, v));
    };
    auto rhombus = [&](int64_t e) { return A[G->edge_vertices[2 * e]] + A[G->edge_vertices[2 * e + 1]]; };
    auto other = [&](int64_t e, int64_t v) {
        return G->edge_vertices[2 * e] == v ? G->edge_vertices[2 * e + 1] : G->edge_vertices[2 * e];
    };
    auto edges_of = [&](int64_t v, int64_t* out) {
        int n = 0;
        for (int64_t k = 0; k < KV; ++k)
            if (G->vertex_signs[k + v * KV] != 0) out[n++] = G->vertex_edges[k + v * KV];
        return n;
    };
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t e = 0; e < ne; ++e) {
        V3 ne_ = load3(G->edge_normals, e), pe = load3(G->edge_points, e);
        auto metric = [&](int64_t f) {
            return dot(load3(G->edge_normals, f), cross(normalize(pe + load3(G->edge_points, f)), ne_));
        };
        double area = l[e] * d[e];
        int k = 0;
        for (int64_t k2 = 0; k2 < 12; ++k2) {
            edges[k2 + 12 * e] = 0;
            pv_edges[k2 + 12 * e] = 0;
            weights[k2 + 12 * e] = 0.0;
        }
        for (int side = 0; side < 2 && !bad; ++side) {
            int64_t v = G->edge_vertices[side + 2 * e];
            int64_t all[64], inner[64];
            int na = edges_of(v, all), ni = 0;
            for (int q = 0; q < na; ++q)
                if (all[q] != e) inner[ni++] = all[q];
            if (ni != 2) {
                bad = 1;
                break;
            }
            double a = to_vertex(e, v) * d[e] / area;
            for (int j = 0; j < 2; ++j) {
                int64_t f = inner[j], third = inner[1 - j];
                double b = to_vertex(f, v) * d[f];
                edges[k + 12 * e] = f;
                pv_edges[k + 12 * e] = third;
                weights[k + 12 * e] = metric(f) * (-a * b / A[v] + 2.0 / 3.0 * (a * l[f] * d[f] / rhombus(f) +
                                                                                b / rhombus(e) +
                                                                                a * b / rhombus(third)));
                ++k;
                int64_t w = other(f, v);
                int64_t ow[64];
                int no = edges_of(w, ow);
                for (int q = 0; q < no; ++q) {
                    int64_t o = ow[q];
                    if (o == f) continue;
                    if (k >= 12) {
                        bad = 1;
                        break;
                    }
                    edges[k + 12 * e] = o;
                    pv_edges[k + 12 * e] = f;
                    weights[k + 12 * e] = 2.0 / 3.0 * metric(o) * a * to_vertex(o, w) * d[o] / rhombus(f);
                    ++k;
                }
            }
        }
    }
    if (bad) return fail(-2, "Gassmann's reconstruction needs three edges per vertex");
    return 0;
}

// Skilling (2004) transpose-to-axes form of the 3-D Hilbert index.
int dmx_hilbert_keys(int64_t n, const uint64_t* x, int bits, uint64_t* keys) {
    if (bits < 1 || bits > 21) return fail(-1, "bits must be in 1..21");
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        uint64_t X[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
        const uint64_t M = uint64_t(1) << (bits - 1);
        for (uint64_t Q = M; Q > 1; Q >>= 1) {
            uint64_t P = Q - 1;
            for (int i = 0; i < 3; ++i) {
                if (X[i] & Q) {
                    X[0] ^= P;
                } else {
                    uint64_t t = (X[0] ^ X[i]) & P;
                    X[0] ^= t;
                    X[i] ^= t;
                }
            }
        }
        X[1] ^= X[0];
        X[2] ^= X[1];
        uint64_t t = 0;
        for (uint64_t Q = M; Q > 1; Q >>= 1)
            if (X[2] & Q) t ^= Q - 1;
        for (int i = 0; i < 3; ++i) X[i] ^= t;
        uint64_t key = 0;
        for (int b = bits - 1; b >= 0; --b)
            for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1);
        keys[p] = key;
    }
    return 0;
}

int dmx_hilbert_permutation(int64_t n, const double* pts, int bits, int64_t* perm) {
    if (n == 0) return 0;
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) lo[i] = hi[i] = pts[i];
    for (int64_t p = 0; p < n; ++p)
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], pts[3 * p + i]);
            hi[i] = std::max(hi[i], pts[3 * p + i]);
        }
    double scale[3];
    const double eps = 2.220446049250313e-16;
    for (int i = 0; i < 3; ++i) scale[i] = (std::ldexp(1.0, bits) - 1) / std::max(hi[i] - lo[i], eps);
    std::vector<uint64_t> q(3 * n), keys(n);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p)
        for (int i = 0; i < 3; ++i) q[3 * p + i] = uint64_t(std::nearbyint((pts[3 * p + i] - lo[i]) * scale[i]));
    int rc = dmx_hilbert_keys(n, q.data(), bits, keys.data());
    if (rc) return rc;
    std::iota(perm, perm + n, int64_t(0));
    std::stable_sort(perm, perm + n, [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    return 0;
}

// Real orthonormal spherical harmonics Y_lm (no Condon-Shortley phase):
//   Y_l0 = sqrt((2l+1)/4pi) P_l(z),
//   Y_lm = sqrt(2) N_lm P_l^m(z) cos(m phi),  Y_l,-m = sqrt(2) N_lm P_l^m(z) sin(m phi),
// with P_l^m(z) cos(m phi) = Q_lm(z) Re((x + iy)^m), P_l^m = rho^m Q_lm, which is
// smooth at the poles.
int dmx_streamfunction(int lmax, const double* coef, int64_t n, const double* pts, double* psi) {
    if (lmax < 1 || lmax > 32) return fail(-1, "lmax must be in 1..32");
    std::vector<double> norml((lmax + 1) * (lmax + 1));
    for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= l; ++m) {
            double r = 1.0;  // (l-m)!/(l+m)!
            for (int k = l - m + 1; k <= l + m; ++k) r /= k;
            double nlm = std::sqrt((2 * l + 1) / (4 * M_PI) * r);
            norml[l * (lmax + 1) + m] = m == 0 ? nlm : std::sqrt(2.0) * nlm;
        }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
        double s = 0;
        double cr = 1, ci = 0;  // (x + iy)^m
        for (int m = 0; m <= lmax; ++m) {
            // Q_mm = (2m-1)!!, Q_{m+1,m} = (2m+1) z Q_mm, then the three-term recurrence
            double qmm = 1;
            for (int k = 1; k <= m; ++k) qmm *= 2 * k - 1;
            double qa = qmm, qb = (2 * m + 1) * z * qmm;
            for (int l = m; l <= lmax; ++l) {
                double q;
                if (l == m)
                    q = qa;
                else if (l == m + 1)
                    q = qb;
                else {
                    q = ((2 * l - 1) * z * qb - (l + m - 1) * qa) / (l - m);
                    qa = qb;
                    qb = q;
                }
                if (l >= 1) {
                    double nl = norml[l * (lmax + 1) + m];
                    if (m == 0)
                        s += coef[l * l - 1 + l] * nl * q;
                    else
                        s += nl * q * (coef[l * l - 1 + l + m] * cr + coef[l * l - 1 + l - m] * ci);
                }
            }
            double nr = cr * x - ci * y, ni = cr * y + ci * x;
            cr = nr;
            ci = ni;
        }
        psi[p] = s;
    }
    return 0;
}

}  // extern "C"
