// TRiSK tables of ICON-IAP used by the MPAS-Gassmann operators: the PV edges of
// i_cori_method = 4 (trisk_pv_edges, src/FiniteVolumes/TRiSKGrid.jl:318-362) and the TRiSK
// weights from planar kite fractions (iap_trisk_weights, :364-403).  Both walk the cells as
// TRiSKGrid builds its TRiSK stencil, so slot k of edge e matches trisk_edges[k, e].
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/dmplex_host.h"
#include "common.hpp"

namespace dmx {
namespace {

int cell_size(const dmx_grid_tables* G, int64_t i) {
    int n = 0;
    for (int64_t k = 0; k < G->kc; ++k) n += G->cell_signs[k + G->kc * i] != 0;
    return n;
}

inline int64_t mod_n(int64_t a, int64_t n) { return ((a % n) + n) % n; }

}  // namespace
}  // namespace dmx

extern "C" {

int dmx_trisk_pv_edges(int64_t nc, int64_t ne, const dmx_grid_tables* G, int64_t* pv_edges) {
    using namespace dmx;
    const int64_t KC = G->kc, KV = G->kv, KT = G->kt;
    std::vector<int64_t> count(size_t(ne), 0);
    for (int64_t t = 0; t < KT * ne; ++t) pv_edges[t] = -1;
    // the edge of vertex v other than e and e2, if exactly one (else -1)
    auto third_edge = [&](int64_t v, int64_t e, int64_t e2) -> int64_t {
        int64_t found = -1, n = 0;
        for (int64_t j = 0; j < KV; ++j) {
            if (G->vertex_signs[j + KV * v] == 0) continue;
            const int64_t x = G->vertex_edges[j + KV * v];
            if (x == e || x == e2) continue;
            found = x;
            ++n;
        }
        return n == 1 ? found : -1;
    };
    for (int64_t i = 0; i < nc; ++i) {
        const int n = cell_size(G, i);
        const int64_t* edges = G->cell_edges + KC * i;
        const int64_t* corners = G->cell_vertices + KC * i;
        for (int k = 0; k < n; ++k) {
            const int64_t e = edges[k];
            for (int s = 1; s <= n - 1; ++s) {
                const int64_t slot = count[size_t(e)]++;
                const int64_t e2 = edges[mod_n(k + s, n)];
                if (slot >= KT || G->trisk_edges[slot + KT * e] != e2)
                    return fail(-3, "TRiSK edge table out of order");
                int64_t p;
                if (2 * s == n)
                    p = -1;
                else if (s == 1)  // corner k + 1, between edges k and k + 1
                    p = third_edge(corners[mod_n(k + 1, n)], e, e2);
                else if (s == n - 1)  // corner k, between edges k - 1 and k
                    p = third_edge(corners[k], e, e2);
                else if (s == 2)
                    p = edges[mod_n(k + 1, n)];
                else if (s == n - 2)
                    p = edges[mod_n(k - 1, n)];
                else
                    p = -1;
                pv_edges[slot + KT * e] = p;
            }
        }
    }
    return 0;
}

int dmx_iap_trisk_weights(int64_t nc, int64_t ne, const dmx_grid_tables* G, const double* vertices, double* weights) {
    using namespace dmx;
    const int64_t KC = G->kc, KT = G->kt;
    std::vector<int64_t> count(size_t(ne), 0);
    for (int64_t t = 0; t < KT * ne; ++t) weights[t] = 0.0;
    std::vector<double> R;
    for (int64_t i = 0; i < nc; ++i) {
        const int n = cell_size(G, i);
        const int64_t* edges = G->cell_edges + KC * i;
        const double* signs = G->cell_signs + KC * i;
        const int64_t* corners = G->cell_vertices + KC * i;
        auto side = [&](int64_t e) { return G->edge_cells[2 * e] == i ? 0 : 1; };
        auto triangle = [&](int64_t e, int64_t v) {
            return G->cell_distances[side(e) + 2 * e] * arc_length(load3(G->edge_points, e), load3(vertices, v)) / 2;
        };
        // corner k lies between cell edges k - 1 and k
        R.assign(size_t(n), 0.0);
        for (int k = 0; k < n; ++k)
            R[k] = (triangle(edges[mod_n(k - 1, n)], corners[k]) + triangle(edges[k], corners[k])) / G->cell_areas[i];
        for (int k = 0; k < n; ++k) {
            const int64_t e = edges[k];
            double S = 0.0;
            for (int s = 1; s <= n - 1; ++s) {
                const int64_t k2 = mod_n(k + s, n);
                S += R[k2];
                const int64_t slot = count[size_t(e)]++;
                if (slot >= KT || G->trisk_edges[slot + KT * e] != edges[k2])
                    return fail(-3, "TRiSK edge table out of order");
                weights[slot + KT * e] = signs[k] * signs[k2] * (0.5 - S);
            }
        }
    }
    return 0;
}

}  // extern "C"
