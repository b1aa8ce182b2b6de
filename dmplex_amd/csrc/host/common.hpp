// Small shared helpers of the host library: 3-vectors, error text, deterministic sums.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace dmx {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 normalize(V3 a) { return a / norm(a); }
inline V3 load3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void store3(double* p, int64_t i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
// sign() with sign(0) = 0, as Julia's.
inline double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

// Great-circle distance on the unit sphere (TRiSKGrid.jl:2).
inline double arc_length(V3 a, V3 b) { return std::atan2(norm(cross(a, b)), dot(a, b)); }
// Signed spherical triangle area, negative when clockwise (TRiSKGrid.jl:5).
inline double signed_triangle_area(V3 a, V3 b, V3 c) {
    return 2 * std::atan2(dot(a, cross(b, c)), 1 + dot(a, b) + dot(b, c) + dot(c, a));
}

// Solves (N + 1e-12 max(diag N) I) g = r for the m x m (m <= 16, row stride 16)
// normal equations of Anderson mixing (Icosahedral.jl:174), by Gaussian elimination
// with partial pivoting.
inline void small_solve(int m, const double* N, const double* r, double* g) {
    double A[16][17];
    double dmax = 0;
    for (int i = 0; i < m; ++i) dmax = std::max(dmax, N[i * 16 + i]);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) A[i][j] = N[i * 16 + j];
        A[i][i] += 1e-12 * dmax;
        A[i][m] = r[i];
    }
    for (int c = 0; c < m; ++c) {
        int p = c;
        for (int i = c + 1; i < m; ++i)
            if (std::fabs(A[i][c]) > std::fabs(A[p][c])) p = i;
        if (p != c)
            for (int j = 0; j <= m; ++j) std::swap(A[p][j], A[c][j]);
        double piv = A[c][c];
        if (piv == 0) continue;
        for (int i = c + 1; i < m; ++i) {
            double fct = A[i][c] / piv;
            for (int j = c; j <= m; ++j) A[i][j] -= fct * A[c][j];
        }
    }
    for (int i = m - 1; i >= 0; --i) {
        double s = A[i][m];
        for (int j = i + 1; j < m; ++j) s -= A[i][j] * g[j];
        g[i] = A[i][i] != 0 ? s / A[i][i] : 0.0;
    }
}

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Sum of n terms term(i) in fixed blocks of 4096, blocks added in order: the result
// does not depend on the number of OpenMP threads.
template <class F>
double det_sum(int64_t n, F term) {
    const int64_t B = 4096;
    int64_t nb = (n + B - 1) / B;
    std::vector<double> part(nb > 0 ? nb : 1, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        double s = 0;
        int64_t hi = std::min(n, (b + 1) * B);
        for (int64_t i = b * B; i < hi; ++i) s += term(i);
        part[b] = s;
    }
    double s = 0;
    for (int64_t b = 0; b < nb; ++b) s += part[b];
    return s;
}

}  // namespace dmx
