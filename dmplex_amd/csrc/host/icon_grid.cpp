// Setup tables of the ICON triangle shallow-water model: the vector-RBF tangential
// velocity stencils (rbf_coefficients, src/FiniteVolumes/TRiSKGrid.jl:254-281) and the
// bilinear edge-to-cell weights (bilinear_weights, src/FiniteVolumes/ICONShallowWater.jl:
// 79-96), from the TRiSKGrid tables.
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/dmplex_host.h"
#include "common.hpp"

namespace dmx {
namespace {

struct GridView {
    int64_t kc;
    const double *x, *n;          // edge points, normals (ne x 3)
    const int64_t *ce, *ec;       // cell edges (kc x nc), edge cells (2 x ne)
    const double* cs;             // cell signs (kc x nc)
};

// Edges of the two cells of e other than e itself, cell by cell in slot order (the
// reference's `stencil`), appended to out.
void rbf_stencil(const GridView& g, int64_t e, std::vector<int64_t>& out) {
    out.clear();
    for (int s = 0; s < 2; ++s) {
        const int64_t i = g.ec[2 * e + s];
        for (int64_t k = 0; k < g.kc; ++k)
            if (g.cs[k + g.kc * i] != 0 && g.ce[k + g.kc * i] != e) out.push_back(g.ce[k + g.kc * i]);
    }
}

// Solves A c = b for a symmetric positive definite m x m matrix (row-major) by Cholesky
// A = L L^T; returns false if A is not positive definite.
bool cholesky_solve(int m, std::vector<double>& A, std::vector<double>& b) {
    for (int j = 0; j < m; ++j) {
        double d = A[j * m + j];
        for (int k = 0; k < j; ++k) d -= A[j * m + k] * A[j * m + k];
        if (!(d > 0)) return false;
        const double l = std::sqrt(d);
        A[j * m + j] = l;
        for (int i = j + 1; i < m; ++i) {
            double s = A[i * m + j];
            for (int k = 0; k < j; ++k) s -= A[i * m + k] * A[j * m + k];
            A[i * m + j] = s / l;
        }
    }
    for (int i = 0; i < m; ++i) {  // L y = b
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= A[i * m + k] * b[k];
        b[i] = s / A[i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {  // L^T c = y
        double s = b[i];
        for (int k = i + 1; k < m; ++k) s -= A[k * m + i] * b[k];
        b[i] = s / A[i * m + i];
    }
    return true;
}

}  // namespace
}  // namespace dmx

extern "C" {

// rbf_coefficients(g; scale): stencil of edge e = the other edges of its two cells, kernel
// phi(r) = 1/sqrt(1 + (r/scale)^2) of the great-circle distance times n_i . n_j, weights
// from the Cholesky solve normalized to reproduce a constant vector along t = x_e x n_e.
// With edges == NULL only *width (the longest stencil) is computed.  Tables kr x ne,
// column-major, pads index 0 / weight 0.
int dmx_rbf_coefficients(int64_t nc, int64_t ne, const dmx_grid_tables* G, double scale, int64_t kr, int64_t* edges,
                         double* weights, int64_t* width) {
    using namespace dmx;
    if (!G || ne <= 0 || nc <= 0 || !(scale > 0)) return fail(-1, "bad arguments");
    GridView g{G->kc, G->edge_points, G->edge_normals, G->cell_edges, G->edge_cells, G->cell_signs};
    int64_t w = 0;
    {
        std::vector<int64_t> st;
        for (int64_t e = 0; e < ne; ++e) {
            rbf_stencil(g, e, st);
            w = std::max<int64_t>(w, int64_t(st.size()));
        }
    }
    if (width) *width = w;
    if (!edges) return 0;
    if (kr < w) return fail(-1, "rbf table narrower than the longest stencil");
    int bad = 0;
#pragma omp parallel
    {
        std::vector<int64_t> st;
        std::vector<double> A, b;
#pragma omp for schedule(static) reduction(| : bad)
        for (int64_t e = 0; e < ne; ++e) {
            rbf_stencil(g, e, st);
            const int m = int(st.size());
            auto phi = [&](V3 a, V3 c) {
                const double r = arc_length(a, c) / scale;
                return 1 / std::sqrt(1 + r * r);
            };
            A.assign(size_t(m) * m, 0.0);
            b.assign(size_t(m), 0.0);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j)
                    A[i * m + j] = dot(load3(g.n, st[i]), load3(g.n, st[j])) * phi(load3(g.x, st[i]), load3(g.x, st[j]));
            const V3 xe = load3(g.x, e), t = cross(xe, load3(g.n, e));
            for (int j = 0; j < m; ++j) b[j] = dot(t, load3(g.n, st[j])) * phi(xe, load3(g.x, st[j]));
            if (!cholesky_solve(m, A, b)) {
                bad = 1;
                continue;
            }
            double norm_t = 0;
            for (int k = 0; k < m; ++k) norm_t += b[k] * dot(t, load3(g.n, st[k]));
            for (int k = 0; k < kr; ++k) {
                edges[k + kr * e] = k < m ? st[k] : 0;
                weights[k + kr * e] = k < m ? b[k] / norm_t : 0.0;
            }
        }
    }
    return bad ? fail(-2, "an RBF matrix is not positive definite") : 0;
}

// bilinear_weights(g): per triangle i the weights w (3) with sum w_k = 1 and
// sum w_k x_k = 0 in the tangent basis a = normalize(c x (|c_z| < 0.9 ? z : x)),
// b = c x a of the center c, by LU with partial pivoting.  weights: 3 x nc.
int dmx_bilinear_weights(int64_t nc, const dmx_grid_tables* G, const double* centers, double* weights) {
    using namespace dmx;
    if (!G || !centers || !weights || nc <= 0) return fail(-1, "bad arguments");
    if (G->kc != 3) return fail(-3, "bilinear weights need a triangle mesh");
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < nc; ++i) {
        const V3 c = load3(centers, i);
        const V3 a = normalize(cross(c, std::fabs(c.z) < 0.9 ? V3{0, 0, 1} : V3{1, 0, 0}));
        const V3 b = cross(c, a);
        double M[3][4];
        for (int k = 0; k < 3; ++k) {
            if (G->cell_signs[k + 3 * i] == 0) bad = 1;
            const V3 x = load3(G->edge_points, G->cell_edges[k + 3 * i]);
            M[0][k] = 1.0;
            M[1][k] = dot(x, a);
            M[2][k] = dot(x, b);
        }
        M[0][3] = 1.0;
        M[1][3] = 0.0;
        M[2][3] = 0.0;
        for (int col = 0; col < 3; ++col) {
            int p = col;
            for (int r = col + 1; r < 3; ++r)
                if (std::fabs(M[r][col]) > std::fabs(M[p][col])) p = r;
            if (p != col)
                for (int j = 0; j < 4; ++j) std::swap(M[p][j], M[col][j]);
            if (M[col][col] == 0) {
                bad = 1;
                break;
            }
            for (int r = col + 1; r < 3; ++r) {
                const double f = M[r][col] / M[col][col];
                for (int j = col; j < 4; ++j) M[r][j] -= f * M[col][j];
            }
        }
        double w[3] = {0, 0, 0};
        for (int r = 2; r >= 0; --r) {
            double s = M[r][3];
            for (int j = r + 1; j < 3; ++j) s -= M[r][j] * w[j];
            w[r] = M[r][r] != 0 ? s / M[r][r] : 0.0;
        }
        for (int k = 0; k < 3; ++k) weights[k + 3 * i] = w[k];
    }
    return bad ? fail(-3, "bilinear weights need a triangle mesh of non-degenerate cells") : 0;
}

}  // extern "C"
