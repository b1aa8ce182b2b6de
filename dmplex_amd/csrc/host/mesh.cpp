// Mesh generation and Plex topology of the host library.
//
// Restates, in native code, the mesh pieces the TRiSK path is set up from:
//   create_spherical_mesh  src/Grids/Icosahedral.jl:22-61
//   lloyd! / lloyd_step! / is_delaunay  src/Grids/Icosahedral.jl:126-216
//   create_hexagon_mesh  src/Grids/Hexagonal.jl:11-17 (dual_polygons, sort_ring:
//                        src/Grids/FileGrids.jl:28-42)
//   plex_from_polygons, invert_cones, get_edge_cells, get_cell_vertices, get_star,
//   check_plex  src/Topology/Plex.jl:49-97,142-160,210-262,329-360
//   orthogonality_defect  src/Grids/CellCenters.jl:90-101
// Numbering and orientation conventions are the reference's (0-based here).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "../../../include/dmplex_host.h"
#include "common.hpp"

namespace dmx {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

// ---------------------------------------------------------------------------------
// Triangle stars and Lloyd's algorithm

// Counter-clockwise star of triangles around every vertex, starting at the first
// triangle (in triangle order) that contains it; the triangle after (v, b, w) is the
// one holding the directed edge v -> w (Icosahedral.jl:130-147).
struct Stars {
    std::vector<int64_t> off, tri;         // CSR: star of vertex v
    std::vector<int64_t> inc_off, inc;     // incidences (3*t + corner) in triangle order
};

static int build_stars(int64_t nv, const int64_t* T, int64_t nt, Stars& s) {
    s.inc_off.assign(nv + 1, 0);
    for (int64_t i = 0; i < 3 * nt; ++i) {
        int64_t v = T[i];
        if (v < 0 || v >= nv) return fail(-1, "triangle vertex out of range");
        s.inc_off[v + 1]++;
    }
    for (int64_t v = 0; v < nv; ++v) s.inc_off[v + 1] += s.inc_off[v];
    s.inc.resize(3 * nt);
    std::vector<int64_t> pos(s.inc_off.begin(), s.inc_off.end() - 1);
    for (int64_t i = 0; i < 3 * nt; ++i) s.inc[pos[T[i]]++] = i;
    s.off = s.inc_off;
    s.tri.assign(3 * nt, -1);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t v = 0; v < nv; ++v) {
        int64_t lo = s.inc_off[v], hi = s.inc_off[v + 1];
        if (lo == hi) continue;
        int64_t first = s.inc[lo] / 3, corner = s.inc[lo] % 3;
        int64_t t = first, w = T[3 * t + (corner + 2) % 3];
        int64_t out = lo;
        s.tri[out++] = t;
        while (true) {
            int64_t next = -1, nk = 0;
            for (int64_t q = lo; q < hi; ++q) {
                int64_t tt = s.inc[q] / 3, k = s.inc[q] % 3;
                if (T[3 * tt + (k + 1) % 3] == w) {
                    next = tt;
                    nk = k;
                    break;
                }
            }
            if (next < 0 || out > hi) {
                bad = 1;
                break;
            }
            if (next == first) break;
            if (out >= hi) {
                bad = 1;
                break;
            }
            s.tri[out++] = next;
            w = T[3 * next + (nk + 2) % 3];
        }
        if (out != hi) bad = 1;
    }
    if (bad) return fail(-2, "triangulation is not a closed oriented surface");
    return 0;
}

static void circumcenters(const double* x, const int64_t* T, int64_t nt, std::vector<V3>& c) {
    c.resize(nt);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        V3 a = load3(x, T[3 * t]), b = load3(x, T[3 * t + 1]), d = load3(x, T[3 * t + 2]);
        c[t] = normalize(cross(b - a, d - a));
    }
}

// lloyd_step! (Icosahedral.jl:186-207): f = centroid - x for the free vertices.
static double lloyd_step(double* f, std::vector<V3>& centers, const double* x, const int64_t* T,
                         int64_t nt, const Stars& s, int64_t nv, int64_t nfixed) {
    circumcenters(x, T, nt, centers);
    double residual = 0;
#pragma omp parallel for schedule(static) reduction(max : residual)
    for (int64_t v = 0; v < nv; ++v) {
        if (v < nfixed) {
            store3(f, v, {0, 0, 0});
            continue;
        }
        int64_t lo = s.off[v], n = s.off[v + 1] - lo;
        V3 m{0, 0, 0};
        for (int64_t k = 0; k < n; ++k) {
            V3 p = centers[s.tri[lo + k]], q = centers[s.tri[lo + (k + 1) % n]];
            V3 c = cross(p, q);
            double nc = norm(c);
            m = m + (std::atan2(nc, dot(p, q)) / nc) * c;
        }
        V3 fv = normalize(m) - load3(x, v);
        store3(f, v, fv);
        residual = std::max(residual, norm(fv));
    }
    return residual;
}

// is_delaunay (Icosahedral.jl:210-216) for every triangle.
static bool all_delaunay(const double* x, const int64_t* T, int64_t nt, const Stars& s) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t t = 0; t < nt; ++t) {
        int64_t a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
        V3 xa = load3(x, a);
        V3 center = normalize(cross(load3(x, b) - xa, load3(x, c) - xa));
        int64_t pq[3][2] = {{a, b}, {b, c}, {c, a}};
        for (auto& e : pq) {
            int64_t p = e[0], q = e[1], o = -1;
            // triangle with the directed edge q -> p
            for (int64_t r = s.inc_off[q]; r < s.inc_off[q + 1]; ++r) {
                int64_t tt = s.inc[r] / 3, k = s.inc[r] % 3;
                if (T[3 * tt + (k + 1) % 3] == p) {
                    o = T[3 * tt + (k + 2) % 3];
                    break;
                }
            }
            if (o < 0 || dot(load3(x, o), center) > dot(xa, center) + 1e-14) bad = 1;
        }
    }
    return !bad;
}

// lloyd! (Icosahedral.jl:126-184) with Anderson mixing (type II) over `history` steps.
static int lloyd(double* x, int64_t nv, const int64_t* T, int64_t nt, int64_t nfixed,
                 double tolerance, int history, int64_t max_iterations, int64_t* iterations) {
    if (history < 0 || history > 16) return fail(-1, "history must be in 0..16");
    Stars s;
    int rc = build_stars(nv, T, nt, s);
    if (rc) return rc;
    double h = det_sum(nt, [&](int64_t t) {
                   return norm(load3(x, T[3 * t + 1]) - load3(x, T[3 * t]));
               }) / double(nt);
    const int64_t n3 = 3 * nv;
    std::vector<double> f(n3), fold(n3), xold(n3);
    std::vector<double> dX(history > 0 ? size_t(history) * n3 : 0), dF(dX.size());
    std::vector<V3> centers;
    double N[16 * 16] = {0};
    int m = 0;
    const int64_t B = 4096;
    const int64_t nb = (n3 + B - 1) / B;
    std::vector<double> part(size_t(nb) * 33);
    for (int64_t it = 1; it <= max_iterations; ++it) {
        double residual = lloyd_step(f.data(), centers, x, T, nt, s, nv, nfixed);
        if (residual < tolerance * h) {
            if (iterations) *iterations = it;
            if (!all_delaunay(x, T, nt, s))
                return fail(-3, "Lloyd's algorithm left the triangulation non-Delaunay");
            return 0;
        }
        int k = -1;
        if (history > 0 && it > 1) {
            k = int((it - 2) % history);
            double* cx = &dX[size_t(k) * n3];
            double* cf = &dF[size_t(k) * n3];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n3; ++i) {
                cx[i] = x[i] - xold[i];
                cf[i] = f[i] - fold[i];
            }
            m = std::min(m + 1, history);
        }
        std::memcpy(xold.data(), x, n3 * sizeof(double));
        std::memcpy(fold.data(), f.data(), n3 * sizeof(double));
        if (m > 0) {
            // one pass: the new column's products with all columns, and A' f
            const double* cf = k >= 0 ? &dF[size_t(k) * n3] : nullptr;
#pragma omp parallel for schedule(static)
            for (int64_t b = 0; b < nb; ++b) {
                double acc[33] = {0};
                int64_t lo = b * B, hi = std::min(n3, lo + B);
                for (int j = 0; j < m; ++j) {
                    const double* cj = &dF[size_t(j) * n3];
                    double sf = 0, sk = 0;
                    for (int64_t i = lo; i < hi; ++i) {
                        sf += cj[i] * f[i];
                        if (cf) sk += cj[i] * cf[i];
                    }
                    acc[j] = sf;
                    acc[16 + j] = sk;
                }
                std::memcpy(&part[size_t(b) * 33], acc, sizeof(acc));
            }
            double r[16] = {0}, nk[16] = {0}, g[16] = {0};
            for (int64_t b = 0; b < nb; ++b)
                for (int j = 0; j < m; ++j) {
                    r[j] += part[size_t(b) * 33 + j];
                    nk[j] += part[size_t(b) * 33 + 16 + j];
                }
            if (k >= 0)
                for (int j = 0; j < m; ++j) N[k * 16 + j] = N[j * 16 + k] = nk[j];
            small_solve(m, N, r, g);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n3; ++i) {
                double c = 0;
                for (int j = 0; j < m; ++j) c += g[j] * (dX[size_t(j) * n3 + i] + dF[size_t(j) * n3 + i]);
                x[i] += f[i] - c;
            }
        } else {
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n3; ++i) x[i] += f[i];
        }
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < nv; ++v) store3(x, v, normalize(load3(x, v)));
    }
    if (iterations) *iterations = max_iterations;
    set_error("Lloyd's algorithm did not converge");
    return 1;  // a warning in the reference, not an error
}

// ---------------------------------------------------------------------------------
// Icosahedral bisection (Icosahedral.jl:22-61)

// One bisection of every triangle into four, new vertices at the normalized edge
// midpoints in order of first request (Icosahedral.jl:38-48).
static void bisect(std::vector<double>& x, std::vector<int64_t>& T) {
    int64_t nv = int64_t(x.size() / 3), nt = int64_t(T.size() / 3);
    std::vector<int64_t> cnt(nv + 1, 0);
    for (int64_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            int64_t a = T[3 * t + k], b = T[3 * t + (k + 1) % 3];
            cnt[std::min(a, b) + 1]++;
        }
    for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    std::vector<std::pair<int64_t, int64_t>> slot(cnt[nv]);
    x.reserve(x.size() + 3 * (nv + nt));
    auto midpoint = [&](int64_t a, int64_t b) -> int64_t {
        int64_t lo = std::min(a, b), hi = std::max(a, b);
        for (int64_t q = cnt[lo]; q < fill[lo]; ++q)
            if (slot[q].first == hi) return slot[q].second;
        V3 m = normalize(load3(x.data(), a) + load3(x.data(), b));
        int64_t id = int64_t(x.size() / 3);
        x.push_back(m.x);
        x.push_back(m.y);
        x.push_back(m.z);
        slot[fill[lo]++] = {hi, id};
        return id;
    };
    std::vector<int64_t> R;
    R.reserve(12 * nt);
    for (int64_t t = 0; t < nt; ++t) {
        int64_t a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
        int64_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
        int64_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
        R.insert(R.end(), q, q + 12);
    }
    T.swap(R);
}


static int icosahedral(int subdivisions, int centroidal, std::vector<double>& x,
                       std::vector<int64_t>& T, int64_t* iterations) {
    if (subdivisions < 0) return fail(-1, "subdivisions must be non-negative");
    const double PHI = (1.0 + std::sqrt(5.0)) / 2.0;
    const double raw[12][3] = {{-1, PHI, 0}, {1, PHI, 0},  {-1, -PHI, 0}, {1, -PHI, 0},
                               {0, -1, PHI}, {0, 1, PHI},  {0, -1, -PHI}, {0, 1, -PHI},
                               {PHI, 0, -1}, {PHI, 0, 1},  {-PHI, 0, -1}, {-PHI, 0, 1}};
    const int tri0[20][3] = {{1, 12, 6}, {1, 6, 2},  {1, 2, 8},  {1, 8, 11}, {1, 11, 12},
                             {2, 6, 10}, {6, 12, 5}, {12, 11, 3}, {11, 8, 7}, {8, 2, 9},
                             {4, 10, 5}, {4, 5, 3},  {4, 3, 7},  {4, 7, 9},  {4, 9, 10},
                             {5, 10, 6}, {3, 5, 12}, {7, 3, 11}, {9, 7, 8},  {10, 9, 2}};
    x.clear();
    for (auto& r : raw) {
        V3 v = normalize(V3{r[0], r[1], r[2]});
        x.push_back(v.x);
        x.push_back(v.y);
        x.push_back(v.z);
    }
    T.clear();
    for (auto& t : tri0)
        for (int k = 0; k < 3; ++k) T.push_back(t[k] - 1);
    for (int level = 1; level <= subdivisions; ++level) {
        bisect(x, T);
        if (centroidal) {
            int64_t its = 0;
            int rc = lloyd(x.data(), int64_t(x.size() / 3), T.data(), int64_t(T.size() / 3), 12, 1e-8,
                           16, 10000, &its);
            if (rc < 0) return rc;
            if (iterations) iterations[level - 1] = its;
        } else if (iterations) {
            iterations[level - 1] = 0;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------
// Plex construction (Plex.jl:49-97)

struct PlexBuild {
    int64_t nc = 0, ne = 0, nv = 0;
    std::vector<int64_t> cone_offsets, cones, support_offsets, supports;
    std::vector<int8_t> orient;
};

static void invert_cones(PlexBuild& p) {
    int64_t np = p.nc + p.ne + p.nv;
    p.support_offsets.assign(np + 1, 0);
    for (int64_t q : p.cones) p.support_offsets[q + 1]++;
    for (int64_t i = 0; i < np; ++i) p.support_offsets[i + 1] += p.support_offsets[i];
    p.supports.resize(p.cones.size());
    std::vector<int64_t> pos(p.support_offsets.begin(), p.support_offsets.end() - 1);
    for (int64_t i = 0; i < np; ++i)
        for (int64_t k = p.cone_offsets[i]; k < p.cone_offsets[i + 1]; ++k)
            p.supports[pos[p.cones[k]]++] = i;
}

static int plex_from_polygons(int64_t npoly, const int64_t* off, const int64_t* pv, int64_t nv,
                              const double* coords, PlexBuild& p) {
    p.nc = npoly;
    p.nv = nv;
    // counter-clockwise orientation (orient_ccw, Plex.jl:99-103)
    std::vector<int64_t> poly(off[npoly]);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t c = 0; c < npoly; ++c) {
        int64_t lo = off[c], n = off[c + 1] - lo;
        if (n < 3) {
            bad = 1;
            continue;
        }
        V3 normal{0, 0, 0}, sum{0, 0, 0};
        for (int64_t i = 0; i < n; ++i) {
            int64_t a = pv[lo + i], b = pv[lo + (i + 1) % n];
            if (a < 0 || a >= nv || b < 0 || b >= nv) {
                bad = 1;
                break;
            }
            normal = normal + cross(load3(coords, a), load3(coords, b));
            sum = sum + load3(coords, a);
        }
        bool ccw = dot(normal, normalize(sum)) >= 0;
        for (int64_t i = 0; i < n; ++i) poly[lo + i] = ccw ? pv[lo + i] : pv[lo + n - 1 - i];
    }
    if (bad) return fail(-1, "polygon with fewer than 3 vertices or a vertex out of range");
    // edges by first appearance, looked up in per-lower-vertex slots
    std::vector<int64_t> cnt(nv + 1, 0);
    for (int64_t c = 0; c < npoly; ++c) {
        int64_t lo = off[c], n = off[c + 1] - lo;
        for (int64_t i = 0; i < n; ++i) {
            int64_t a = poly[lo + i], b = poly[lo + (i + 1) % n];
            if (a == b) return fail(-1, "degenerate edge at vertex " + std::to_string(a));
            cnt[std::min(a, b) + 1]++;
        }
    }
    for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    std::vector<std::pair<int64_t, int64_t>> slot(cnt[nv]);
    std::vector<int64_t> edge_lo, edge_hi;
    std::vector<int64_t> cell_edge(off[npoly]);
    for (int64_t c = 0; c < npoly; ++c) {
        int64_t lo = off[c], n = off[c + 1] - lo;
        for (int64_t i = 0; i < n; ++i) {
            int64_t a = poly[lo + i], b = poly[lo + (i + 1) % n];
            int64_t l = std::min(a, b), h = std::max(a, b), id = -1;
            for (int64_t q = cnt[l]; q < fill[l]; ++q)
                if (slot[q].first == h) {
                    id = slot[q].second;
                    break;
                }
            if (id < 0) {
                id = int64_t(edge_lo.size());
                edge_lo.push_back(l);
                edge_hi.push_back(h);
                slot[fill[l]++] = {h, id};
            }
            cell_edge[lo + i] = id;
        }
    }
    p.ne = int64_t(edge_lo.size());
    int64_t np = p.nc + p.ne + p.nv;
    p.cone_offsets.assign(np + 1, 0);
    p.cones.resize(off[npoly] + 2 * p.ne);
    p.orient.resize(p.cones.size());
    for (int64_t c = 0; c < npoly; ++c) {
        int64_t lo = off[c], n = off[c + 1] - lo;
        for (int64_t i = 0; i < n; ++i) {
            int64_t a = poly[lo + i], b = poly[lo + (i + 1) % n];
            p.cones[lo + i] = p.nc + cell_edge[lo + i];
            p.orient[lo + i] = a < b ? 0 : -1;
        }
        p.cone_offsets[c + 1] = off[c + 1];
    }
    int64_t base = off[npoly];
    for (int64_t e = 0; e < p.ne; ++e) {
        p.cones[base + 2 * e] = p.nc + p.ne + edge_lo[e];
        p.cones[base + 2 * e + 1] = p.nc + p.ne + edge_hi[e];
        p.orient[base + 2 * e] = p.orient[base + 2 * e + 1] = 0;
        p.cone_offsets[p.nc + e + 1] = base + 2 * e + 2;
    }
    for (int64_t i = p.nc + p.ne; i < np; ++i) p.cone_offsets[i + 1] = int64_t(p.cones.size());
    invert_cones(p);
    return 0;
}

// get_edge_cells (Plex.jl:210-218): (left, right) cell points, -1 when missing.
static void edge_cells(const dmx_plex* dm, int64_t e, int64_t& left, int64_t& right) {
    left = right = -1;
    for (int64_t s = dm->support_offsets[e]; s < dm->support_offsets[e + 1]; ++s) {
        int64_t c = dm->supports[s];
        for (int64_t k = dm->cone_offsets[c]; k < dm->cone_offsets[c + 1]; ++k)
            if (dm->cones[k] == e) {
                if (dm->orientations[k] == 0)
                    left = c;
                else
                    right = c;
                break;
            }
    }
}

}  // namespace dmx

using namespace dmx;

struct dmx_plex_build {
    PlexBuild p;
};

extern "C" {

const char* dmx_last_error(void) { return g_error.c_str(); }

int dmx_icosahedral(int subdivisions, int centroidal, double* coords, int64_t* triangles,
                    int64_t* iterations) {
    std::vector<double> x;
    std::vector<int64_t> T;
    int rc = icosahedral(subdivisions, centroidal, x, T, iterations);
    if (rc < 0) return rc;
    std::memcpy(coords, x.data(), x.size() * sizeof(double));
    std::memcpy(triangles, T.data(), T.size() * sizeof(int64_t));
    return 0;
}

int dmx_lloyd(double* x, int64_t nv, const int64_t* tri, int64_t nt, int64_t nfixed, double tolerance,
              int history, int64_t max_iterations, int64_t* iterations) {
    return lloyd(x, nv, tri, nt, nfixed, tolerance, history, max_iterations, iterations);
}

int dmx_icosahedron(double* coords, int64_t* triangles) {
    std::vector<double> x;
    std::vector<int64_t> T;
    int rc = icosahedral(0, 0, x, T, nullptr);
    if (rc < 0) return rc;
    std::memcpy(coords, x.data(), x.size() * sizeof(double));
    std::memcpy(triangles, T.data(), T.size() * sizeof(int64_t));
    return 0;
}

int dmx_bisect(int64_t nv, const double* x, int64_t nt, const int64_t* tri, double* x_out, int64_t* tri_out) {
    std::vector<double> X(x, x + 3 * nv);
    std::vector<int64_t> T(tri, tri + 3 * nt);
    bisect(X, T);
    std::memcpy(x_out, X.data(), X.size() * sizeof(double));
    std::memcpy(tri_out, T.data(), T.size() * sizeof(int64_t));
    return 0;
}

int dmx_triangle_stars(int64_t nv, const int64_t* tri, int64_t nt, int64_t* star_offsets, int64_t* stars) {
    Stars s;
    int rc = build_stars(nv, tri, nt, s);
    if (rc) return rc;
    std::memcpy(star_offsets, s.off.data(), (nv + 1) * sizeof(int64_t));
    std::memcpy(stars, s.tri.data(), 3 * nt * sizeof(int64_t));
    return 0;
}

int dmx_delaunay(const double* x, int64_t nv, const int64_t* tri, int64_t nt) {
    Stars s;
    int rc = build_stars(nv, tri, nt, s);
    if (rc) return rc;
    if (!all_delaunay(x, tri, nt, s)) return fail(-3, "Lloyd's algorithm left the triangulation non-Delaunay");
    return 0;
}

int dmx_mean_edge_length(const double* x, const int64_t* tri, int64_t nt, double* h) {
    *h = det_sum(nt, [&](int64_t t) { return norm(load3(x, tri[3 * t + 1]) - load3(x, tri[3 * t])); }) / double(nt);
    return 0;
}

int dmx_lloyd_residual(const double* x, int64_t nv, const int64_t* tri, int64_t nt, int64_t nfixed,
                       double* residual) {
    Stars s;
    int rc = build_stars(nv, tri, nt, s);
    if (rc) return rc;
    std::vector<double> f(3 * nv);
    std::vector<V3> centers;
    *residual = lloyd_step(f.data(), centers, x, tri, nt, s, nv, nfixed);
    return 0;
}

int dmx_plex_from_polygons(int64_t npoly, const int64_t* poly_offsets, const int64_t* poly_vertices,
                           int64_t nv, const double* coords, dmx_plex_build** out, int64_t dims[4]) {
    auto* b = new dmx_plex_build;
    int rc = plex_from_polygons(npoly, poly_offsets, poly_vertices, nv, coords, b->p);
    if (rc) {
        delete b;
        return rc;
    }
    dims[0] = b->p.nc;
    dims[1] = b->p.ne;
    dims[2] = b->p.nv;
    dims[3] = int64_t(b->p.cones.size());
    *out = b;
    return 0;
}

int dmx_plex_export(const dmx_plex_build* b, int64_t* cone_offsets, int64_t* cones, int8_t* orientations,
                    int64_t* support_offsets, int64_t* supports) {
    const PlexBuild& p = b->p;
    std::memcpy(cone_offsets, p.cone_offsets.data(), p.cone_offsets.size() * sizeof(int64_t));
    std::memcpy(cones, p.cones.data(), p.cones.size() * sizeof(int64_t));
    std::memcpy(orientations, p.orient.data(), p.orient.size());
    std::memcpy(support_offsets, p.support_offsets.data(), p.support_offsets.size() * sizeof(int64_t));
    std::memcpy(supports, p.supports.data(), p.supports.size() * sizeof(int64_t));
    return 0;
}

void dmx_plex_free(dmx_plex_build* b) { delete b; }

int dmx_dual_polygons(const dmx_plex* dm, int64_t* poly_offsets, int64_t* poly_vertices, double* centers) {
    const int64_t nc = dm->nc, ne = dm->ne, nv = dm->nv;
    // circumcenters of the (triangle) cells, vertices in get_cell_vertices order
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t c = 0; c < nc; ++c) {
        int64_t lo = dm->cone_offsets[c];
        if (dm->cone_offsets[c + 1] - lo != 3) {
            bad = 1;
            continue;
        }
        V3 p[3];
        for (int i = 0; i < 3; ++i) {
            int64_t e = dm->cones[lo + i];
            int64_t v = dm->cones[dm->cone_offsets[e] + (dm->orientations[lo + i] == 0 ? 0 : 1)];
            p[i] = load3(dm->coordinates, v - nc - ne);
        }
        store3(centers, c, normalize(cross(p[1] - p[0], p[2] - p[0])));
    }
    if (bad) return fail(-1, "circumcenters need a triangle mesh");
    // rings: cells of the star of every vertex in BFS order (get_star, Plex.jl:251-262),
    // sorted counter-clockwise around the vertex (sort_ring, FileGrids.jl:28-32)
    std::vector<int64_t> deg(nv + 1, 0);
    for (int64_t v = 0; v < nv; ++v) {
        int64_t p = nc + ne + v;
        int64_t n = 0;
        for (int64_t s = dm->support_offsets[p]; s < dm->support_offsets[p + 1]; ++s) {
            int64_t e = dm->supports[s];
            n += dm->support_offsets[e + 1] - dm->support_offsets[e];
        }
        deg[v + 1] = n;  // upper bound (each cell appears twice)
    }
    poly_offsets[0] = 0;
    std::vector<int64_t> ring_len(nv);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < nv; ++v) {
        int64_t p = nc + ne + v;
        int64_t ring[64];
        int n = 0;
        for (int64_t s = dm->support_offsets[p]; s < dm->support_offsets[p + 1]; ++s) {
            int64_t e = dm->supports[s];
            for (int64_t t = dm->support_offsets[e]; t < dm->support_offsets[e + 1]; ++t) {
                int64_t c = dm->supports[t];
                bool seen = false;
                for (int q = 0; q < n; ++q) seen |= ring[q] == c;
                if (!seen && n < 64) ring[n++] = c;
            }
        }
        ring_len[v] = n;
    }
    for (int64_t v = 0; v < nv; ++v) poly_offsets[v + 1] = poly_offsets[v] + ring_len[v];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < nv; ++v) {
        int64_t p = nc + ne + v;
        int64_t ring[64];
        int n = 0;
        for (int64_t s = dm->support_offsets[p]; s < dm->support_offsets[p + 1]; ++s) {
            int64_t e = dm->supports[s];
            for (int64_t t = dm->support_offsets[e]; t < dm->support_offsets[e + 1]; ++t) {
                int64_t c = dm->supports[t];
                bool seen = false;
                for (int q = 0; q < n; ++q) seen |= ring[q] == c;
                if (!seen && n < 64) ring[n++] = c;
            }
        }
        V3 axis = normalize(load3(dm->coordinates, v));
        V3 e1 = load3(centers, ring[0]) - axis;
        double ang[64];
        int idx[64];
        for (int q = 0; q < n; ++q) {
            V3 w = load3(centers, ring[q]) - axis;
            ang[q] = std::atan2(dot(cross(e1, w), axis), dot(e1, w));
            idx[q] = q;
        }
        std::stable_sort(idx, idx + n, [&](int a, int b) { return ang[a] < ang[b]; });
        for (int q = 0; q < n; ++q) poly_vertices[poly_offsets[v] + q] = ring[idx[q]];
    }
    return 0;
}

int dmx_orthogonality_defect(const dmx_plex* dm, const double* centers, double* defect) {
    const int64_t nc = dm->nc, ne = dm->ne;
    double d = 0;
#pragma omp parallel for schedule(static) reduction(max : d)
    for (int64_t j = 0; j < ne; ++j) {
        int64_t e = nc + j, l, r;
        edge_cells(dm, e, l, r);
        if (l < 0 || r < 0) continue;
        int64_t lo = dm->cone_offsets[e];
        V3 a = load3(dm->coordinates, dm->cones[lo] - nc - ne);
        V3 b = load3(dm->coordinates, dm->cones[lo + 1] - nc - ne);
        V3 dual = cross(load3(centers, l), load3(centers, r));
        d = std::max(d, std::fabs(dot(normalize(dual), normalize(cross(a, b)))));
    }
    *defect = d;
    return 0;
}

int dmx_check_plex(const dmx_plex* dm) {
    const int64_t nc = dm->nc, ne = dm->ne, nv = dm->nv, np = nc + ne + nv;
    for (int64_t p = 0; p < np; ++p)
        for (int64_t k = dm->cone_offsets[p]; k < dm->cone_offsets[p + 1]; ++k) {
            int64_t q = dm->cones[k];
            bool found = false;
            for (int64_t s = dm->support_offsets[q]; s < dm->support_offsets[q + 1]; ++s)
                found |= dm->supports[s] == p;
            if (!found) return fail(-4, "cone/support mismatch");
        }
    for (int64_t c = 0; c < nc; ++c) {
        int64_t lo = dm->cone_offsets[c], n = dm->cone_offsets[c + 1] - lo;
        if (n < 3) return fail(-4, "cell with fewer than three edges");
        std::vector<int64_t> verts(n);
        for (int64_t i = 0; i < n; ++i) {
            int64_t e = dm->cones[lo + i];
            verts[i] = dm->cones[dm->cone_offsets[e] + (dm->orientations[lo + i] == 0 ? 0 : 1)];
        }
        V3 normal{0, 0, 0}, sum{0, 0, 0};
        for (int64_t i = 0; i < n; ++i) {
            int64_t e = dm->cones[lo + i];
            int64_t stop = dm->cones[dm->cone_offsets[e] + (dm->orientations[lo + i] == 0 ? 1 : 0)];
            if (stop != verts[(i + 1) % n]) return fail(-4, "cell boundary is not a closed loop");
            V3 xi = load3(dm->coordinates, verts[i] - nc - ne);
            V3 xj = load3(dm->coordinates, verts[(i + 1) % n] - nc - ne);
            normal = normal + cross(xi, xj);
            sum = sum + xi;
        }
        if (!(dot(normal, normalize(sum)) > 0)) return fail(-4, "cell is not counter-clockwise");
    }
    for (int64_t j = 0; j < ne; ++j) {
        int64_t e = nc + j, nl = 0, nr = 0;
        for (int64_t s = dm->support_offsets[e]; s < dm->support_offsets[e + 1]; ++s) {
            int64_t c = dm->supports[s];
            for (int64_t k = dm->cone_offsets[c]; k < dm->cone_offsets[c + 1]; ++k)
                if (dm->cones[k] == e) {
                    (dm->orientations[k] == 0 ? nl : nr)++;
                    break;
                }
        }
        if (nl > 1 || nr > 1) return fail(-4, "edge has inconsistent cell orientations");
        if (nl != 1 || nr != 1) return fail(-4, "boundary edge of a closed mesh");
    }
    if (nv - ne + nc != 2) return fail(-4, "Euler characteristic is not 2");
    return 0;
}

}  // extern "C"
