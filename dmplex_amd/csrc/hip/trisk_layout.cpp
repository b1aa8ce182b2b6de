// Host-side re-layout of the reference's TRiSK tables for the device.
//
// Input: the fields of TRiSKShallowWater (src/FiniteVolumes/TRiSKShallowWater.jl:67-92)
// as k x n column-major tables.  Output: cells renumbered along the 3-D Hilbert curve of
// `cell_points` (hilbert_permutation, src/Topology/Ordering.jl:85-92), edges by first
// appearance in the new cells' edge lists and vertices by first appearance in the new
// edges' vertex pairs (the rule of hilbert_order, Ordering.jl:54-83), every table
// entity-major with int32 indices.  Per-entity slot order is unchanged, so each
// gather sums its terms in the reference's order.
#include "trisk_layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace trisk {
namespace {

// 3-D Hilbert index of an integer point, Skilling's (2004) transpose form.
uint64_t hilbert_key(uint64_t X0, uint64_t X1, uint64_t X2, int bits) {
    uint64_t X[3] = {X0, X1, X2};
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
    for (auto& x : X) x ^= t;
    uint64_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1);
    return key;
}

void hilbert_cells(int64_t n, const double* pts, std::vector<int64_t>& perm) {
    const int bits = 16;
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) lo[i] = hi[i] = pts[i];
    for (int64_t p = 0; p < n; ++p)
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], pts[3 * p + i]);
            hi[i] = std::max(hi[i], pts[3 * p + i]);
        }
    double scale[3];
    for (int i = 0; i < 3; ++i) scale[i] = (std::ldexp(1.0, bits) - 1) / std::max(hi[i] - lo[i], 2.220446049250313e-16);
    std::vector<uint64_t> keys(n);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        uint64_t q[3];
        for (int i = 0; i < 3; ++i) q[i] = uint64_t(std::nearbyint((pts[3 * p + i] - lo[i]) * scale[i]));
        keys[p] = hilbert_key(q[0], q[1], q[2], bits);
    }
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), int64_t(0));
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
}

bool in_range(const int64_t* t, int64_t count, int64_t n) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < count; ++i) bad |= (t[i] < 0 || t[i] >= n);
    return !bad;
}

// new table[k * n + jnew] = map[old[k + K * perm[jnew]]]
void remap_index(const int64_t* old, int K, int64_t n, const std::vector<int64_t>& perm,
                 const std::vector<int64_t>& map, std::vector<int32_t>& out) {
    out.resize(size_t(K) * n);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int k = 0; k < K; ++k) out[size_t(k) * n + j] = int32_t(map[old[k + K * perm[j]]]);
}

void remap_value(const double* old, int K, int64_t n, const std::vector<int64_t>& perm, std::vector<double>& out) {
    out.resize(size_t(K) * n);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int k = 0; k < K; ++k) out[size_t(k) * n + j] = old[k + K * perm[j]];
}

void invert(const std::vector<int64_t>& perm, std::vector<int64_t>& inv) {
    inv.assign(perm.size(), -1);
    for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = int64_t(i);
}

}  // namespace

// t16 = t32 - base(j) per slot (slot-major K x n), or kWide in slot 0 of an entity whose
// deltas do not all fit; returns the number of such entities
// K-slot table as 16-bit deltas packed two per int32, slot-major: word w of entity j at
// w * n + j holds slots 2w (low half) and 2w + 1 (high half); kWide in the low half of word
// 0 for an entity whose deltas do not all fit.
template <class Base>
static int64_t narrow_table(const std::vector<int32_t>& t32, int K, int64_t n, Base base, std::vector<int32_t>& out,
                            int64_t lim = 32767) {
    const int words = (K + 1) / 2;
    out.assign(size_t(words) * n, 0);
    int64_t wide = 0;
#pragma omp parallel for schedule(static) reduction(+ : wide)
    for (int64_t j = 0; j < n; ++j) {
        const int64_t b = base(j);
        bool ok = true;
        for (int k = 0; k < K && ok; ++k) {
            const int64_t dlt = int64_t(t32[size_t(k) * n + j]) - b;
            ok = dlt >= -lim && dlt <= lim;
        }
        if (!ok) {
            out[j] = int32_t(uint32_t(uint16_t(kWide)));
            ++wide;
            continue;
        }
        for (int k = 0; k < K; ++k) {
            const uint32_t h = uint16_t(int16_t(int64_t(t32[size_t(k) * n + j]) - b));
            out[size_t(k / 2) * n + j] |= int32_t(k % 2 ? h << 16 : h);
        }
    }
    return wide;
}

bool build_layout(const trisk_model_desc* d, Layout& L, std::string& err) {
    if (!d) {
        err = "null model descriptor";
        return false;
    }
    if (d->nc <= 0 || d->ne <= 0 || d->nv <= 0 || d->kc <= 0 || d->kv <= 0 || d->kt <= 0 ||
        d->nc >= (int64_t(1) << 31) || d->ne >= (int64_t(1) << 31) || d->nv >= (int64_t(1) << 31) ||
        d->kc > 16 || d->kv > 16 || d->kt > 32) {
        err = "bad sizes or table widths";
        return false;
    }
    const bool vk = d->cell_vertices != nullptr;
    const bool gm = d->pv_flux != TRISK_PV_TRISK;
    if (d->pv_flux < 0 || d->pv_flux > 2) {
        err = "pv_flux must be TRISK_PV_TRISK, TRISK_PV_GASSMANN or TRISK_PV_GASSMANN_MEAN";
        return false;
    }
    if (!d->cell_edges || !d->divergence || !d->kinetic || !d->edge_cells || !d->edge_vertices ||
        !d->edge_height || !d->inverse_dual || !d->vertex_edges || !d->curl || !d->vertex_cells ||
        !d->vertex_height || !d->coriolis || !d->trisk_edges || !d->trisk || !d->cell_areas ||
        !d->vertex_areas) {
        err = "missing table";
        return false;
    }
    if (vk && (!d->vk_kites || !d->vk_weights || !d->vk_height)) {
        err = "vertex kinetic energy needs cell_vertices, vk_kites, vk_weights and vk_height";
        return false;
    }
    if (gm && (d->kg <= 0 || d->kg > 16 || !d->g_edges || !d->g_pv_edges || !d->g_weights || !d->g_rhombus ||
               !d->g_coriolis)) {
        err = "the Gassmann PV flux needs g_edges, g_pv_edges, g_weights, g_rhombus and g_coriolis";
        return false;
    }
    const int64_t nc = d->nc, ne = d->ne, nv = d->nv;
    const int KC = int(d->kc), KV = int(d->kv), KT = int(d->kt), KG = gm ? int(d->kg) : 0;
    struct Check {
        const int64_t* t;
        int64_t count, n;
        const char* name;
    } checks[] = {{d->cell_edges, KC * nc, ne, "cell_edges"},       {d->edge_cells, 2 * ne, nc, "edge_cells"},
                  {d->edge_vertices, 2 * ne, nv, "edge_vertices"}, {d->vertex_edges, KV * nv, ne, "vertex_edges"},
                  {d->vertex_cells, KV * nv, nc, "vertex_cells"},  {d->trisk_edges, KT * ne, ne, "trisk_edges"},
                  {vk ? d->cell_vertices : nullptr, KC * nc, nv, "cell_vertices"},
                  {gm ? d->g_edges : nullptr, KG * ne, ne, "g_edges"},
                  {gm ? d->g_pv_edges : nullptr, KG * ne, ne, "g_pv_edges"}};
    for (auto& c : checks)
        if (c.t && !in_range(c.t, c.count, c.n)) {
            err = std::string("index out of range in ") + c.name;
            return false;
        }

    L.nc = nc;
    L.ne = ne;
    L.nv = nv;
    L.kc = KC;
    L.kv = KV;
    L.kt = KT;
    L.kg = KG;
    L.vertex_ke = vk;
    L.gassmann = gm;
    L.gassmann_mean = d->pv_flux == TRISK_PV_GASSMANN_MEAN;
    L.gravity = d->gravity;

    // cells along the Hilbert curve; edges and vertices by first appearance
    if (d->cell_points) {
        hilbert_cells(nc, d->cell_points, L.cperm);
    } else {
        L.cperm.resize(nc);
        std::iota(L.cperm.begin(), L.cperm.end(), int64_t(0));
    }
    std::vector<char> seen(ne, 0);
    L.eperm.clear();
    L.eperm.reserve(ne);
    for (int64_t j = 0; j < nc; ++j) {
        int64_t i = L.cperm[j];
        for (int k = 0; k < KC; ++k) {
            if (d->divergence[k + KC * i] == 0.0) continue;  // padded slot
            int64_t e = d->cell_edges[k + KC * i];
            if (!seen[e]) {
                seen[e] = 1;
                L.eperm.push_back(e);
            }
        }
    }
    for (int64_t e = 0; e < ne; ++e)
        if (!seen[e]) L.eperm.push_back(e);
    std::vector<char> vseen(nv, 0);
    L.vperm.clear();
    L.vperm.reserve(nv);
    for (int64_t j = 0; j < ne; ++j)
        for (int s = 0; s < 2; ++s) {
            int64_t v = d->edge_vertices[s + 2 * L.eperm[j]];
            if (!vseen[v]) {
                vseen[v] = 1;
                L.vperm.push_back(v);
            }
        }
    for (int64_t v = 0; v < nv; ++v)
        if (!vseen[v]) L.vperm.push_back(v);
    invert(L.cperm, L.cinv);
    invert(L.eperm, L.einv);
    invert(L.vperm, L.vinv);

    remap_index(d->cell_edges, KC, nc, L.cperm, L.einv, L.ce);
    remap_value(d->divergence, KC, nc, L.cperm, L.div);
    remap_value(d->kinetic, KC, nc, L.cperm, L.kin);
    L.cell_ke_zero = std::all_of(L.kin.begin(), L.kin.end(), [](double w) { return w == 0.0; });
    remap_value(d->cell_areas, 1, nc, L.cperm, L.area_c);
    remap_index(d->edge_cells, 2, ne, L.eperm, L.cinv, L.ec);
    remap_index(d->edge_vertices, 2, ne, L.eperm, L.vinv, L.ev);
    remap_value(d->edge_height, 2, ne, L.eperm, L.eh);
    remap_value(d->inverse_dual, 1, ne, L.eperm, L.invd);
    remap_index(d->trisk_edges, KT, ne, L.eperm, L.einv, L.te);
    remap_value(d->trisk, KT, ne, L.eperm, L.tw);
    remap_index(d->vertex_edges, KV, nv, L.vperm, L.einv, L.ve);
    remap_value(d->curl, KV, nv, L.vperm, L.curl);
    remap_index(d->vertex_cells, KV, nv, L.vperm, L.cinv, L.vc);
    remap_value(d->vertex_height, KV, nv, L.vperm, L.vh);
    remap_value(d->coriolis, 1, nv, L.vperm, L.fv);
    remap_value(d->vertex_areas, 1, nv, L.vperm, L.area_v);
    if (vk) {
        remap_index(d->cell_vertices, KC, nc, L.cperm, L.vinv, L.cv);
        remap_value(d->vk_kites, KC, nc, L.cperm, L.vkk);
        remap_value(d->vk_weights, KV, nv, L.vperm, L.vkw);
        remap_value(d->vk_height, 2, ne, L.eperm, L.vkh);
    }
    if (gm) {
        remap_index(d->g_edges, KG, ne, L.eperm, L.einv, L.ge);
        remap_index(d->g_pv_edges, KG, ne, L.eperm, L.einv, L.gp);
        remap_value(d->g_weights, KG, ne, L.eperm, L.gw);
        remap_value(d->g_rhombus, 2, ne, L.eperm, L.grh);
        remap_value(d->g_coriolis, 1, ne, L.eperm, L.gf);
    }
    int64_t w = 0;
    w += narrow_table(L.ce, KC, nc, [](int64_t j) { return 3 * j; }, L.ce16p);
    w += narrow_table(L.ec, 2, ne, [](int64_t j) { return j / 3; }, L.ec16p);
    w += narrow_table(L.ev, 2, ne, [](int64_t j) { return (2 * j) / 3; }, L.ev16p);
    w += narrow_table(L.te, KT, ne, [](int64_t j) { return j; }, L.te16p);
    w += narrow_table(L.ve, KV, nv, [](int64_t j) { return (3 * j) / 2; }, L.ve16p);
    w += narrow_table(L.vc, KV, nv, [](int64_t j) { return j / 2; }, L.vc16p);
    if (vk) w += narrow_table(L.cv, KC, nc, [](int64_t j) { return 2 * j; }, L.cv16p);
    if (gm) {
        w += narrow_table(L.ge, KG, ne, [](int64_t j) { return j; }, L.ge16p);
        L.gassmann_slots = !L.gassmann_mean && KG == 12;
        for (int64_t e = 0; e < ne && L.gassmann_slots; ++e)
            for (int k = 0; k < 12; ++k)
                if (L.gp[k * ne + e] != L.ge[gassmann_pv_slot(k) * ne + e]) {
                    L.gassmann_slots = false;
                    break;
                }
        if (L.gassmann_slots) {
            L.gp.clear();
        } else {
            w += narrow_table(L.gp, KG, ne, [](int64_t j) { return j; }, L.gp16p);
        }
    }
    L.wide_table_entities = w;
    return true;
}

bool build_icon_layout(const icon_model_desc* d, Layout& L, std::string& err, int32_t narrow_limit) {
    if (!d) {
        err = "null model descriptor";
        return false;
    }
    if (d->nc <= 0 || d->ne <= 0 || d->nv <= 0 || d->kv <= 0 || d->kv > 16 || d->kr <= 0 || d->kr > 32 ||
        d->nc >= (int64_t(1) << 31) || d->ne >= (int64_t(1) << 31) || d->nv >= (int64_t(1) << 31)) {
        err = "bad sizes or table widths";
        return false;
    }
    if (!d->cell_edges || !d->divergence || !d->bilinear || !d->edge_cells || !d->edge_vertices ||
        !d->edge_height || !d->gradient || !d->inverse_dual || !d->rbf_edges || !d->rbf || !d->edge_coriolis ||
        !d->vertex_edges || !d->curl || !d->vertex_cells || !d->vertex_height || !d->coriolis || !d->cell_areas ||
        !d->vertex_areas) {
        err = "missing table";
        return false;
    }
    const int64_t nc = d->nc, ne = d->ne, nv = d->nv;
    const int KV = int(d->kv), KR = int(d->kr);
    struct Check {
        const int64_t* t;
        int64_t count, n;
        const char* name;
    } checks[] = {{d->cell_edges, 3 * nc, ne, "cell_edges"},      {d->edge_cells, 2 * ne, nc, "edge_cells"},
                  {d->edge_vertices, 2 * ne, nv, "edge_vertices"}, {d->rbf_edges, KR * ne, ne, "rbf_edges"},
                  {d->vertex_edges, KV * nv, ne, "vertex_edges"},  {d->vertex_cells, KV * nv, nc, "vertex_cells"}};
    for (auto& c : checks)
        if (!in_range(c.t, c.count, c.n)) {
            err = std::string("index out of range in ") + c.name;
            return false;
        }
    L = Layout();
    L.icon = true;
    L.nc = nc;
    L.ne = ne;
    L.nv = nv;
    L.kc = 3;
    L.kv = KV;
    L.kr = KR;
    L.gravity = d->gravity;
    if (d->cell_points) {
        hilbert_cells(nc, d->cell_points, L.cperm);
    } else {
        L.cperm.resize(nc);
        std::iota(L.cperm.begin(), L.cperm.end(), int64_t(0));
    }
    std::vector<char> seen(ne, 0);
    L.eperm.reserve(ne);
    for (int64_t j = 0; j < nc; ++j)
        for (int k = 0; k < 3; ++k) {
            const int64_t e = d->cell_edges[k + 3 * L.cperm[j]];
            if (!seen[e]) {
                seen[e] = 1;
                L.eperm.push_back(e);
            }
        }
    for (int64_t e = 0; e < ne; ++e)
        if (!seen[e]) L.eperm.push_back(e);
    std::vector<char> vseen(nv, 0);
    L.vperm.reserve(nv);
    for (int64_t j = 0; j < ne; ++j)
        for (int s = 0; s < 2; ++s) {
            const int64_t v = d->edge_vertices[s + 2 * L.eperm[j]];
            if (!vseen[v]) {
                vseen[v] = 1;
                L.vperm.push_back(v);
            }
        }
    for (int64_t v = 0; v < nv; ++v)
        if (!vseen[v]) L.vperm.push_back(v);
    invert(L.cperm, L.cinv);
    invert(L.eperm, L.einv);
    invert(L.vperm, L.vinv);

    remap_index(d->cell_edges, 3, nc, L.cperm, L.einv, L.ce);
    remap_value(d->divergence, 3, nc, L.cperm, L.div);
    remap_value(d->bilinear, 3, nc, L.cperm, L.bln);
    remap_value(d->cell_areas, 1, nc, L.cperm, L.area_c);
    remap_index(d->edge_cells, 2, ne, L.eperm, L.cinv, L.ec);
    remap_index(d->edge_vertices, 2, ne, L.eperm, L.vinv, L.ev);
    remap_value(d->edge_height, 2, ne, L.eperm, L.eh);
    remap_value(d->gradient, 2, ne, L.eperm, L.grad);
    remap_value(d->inverse_dual, 1, ne, L.eperm, L.invd);
    remap_index(d->rbf_edges, KR, ne, L.eperm, L.einv, L.re);
    remap_value(d->rbf, KR, ne, L.eperm, L.rw);
    remap_value(d->edge_coriolis, 1, ne, L.eperm, L.fe);
    if (d->divergence_damping) remap_value(d->divergence_damping, 1, ne, L.eperm, L.damp);
    remap_index(d->vertex_edges, KV, nv, L.vperm, L.einv, L.ve);
    remap_value(d->curl, KV, nv, L.vperm, L.curl);
    remap_index(d->vertex_cells, KV, nv, L.vperm, L.cinv, L.vc);
    remap_value(d->vertex_height, KV, nv, L.vperm, L.vh);
    remap_value(d->coriolis, 1, nv, L.vperm, L.fv);
    remap_value(d->vertex_areas, 1, nv, L.vperm, L.area_v);
    // 16-bit deltas from the linear bases of triangle meshes (ne = 3 nc / 2 = 3 nv): cell i's
    // edges near 3i/2, edge e's cells near 2e/3 and vertices near e/3, its RBF edges near e,
    // vertex v's edges near 3v
    const int64_t lim = narrow_limit > 0 && narrow_limit < 32767 ? narrow_limit : 32767;
    L.wide_cells = narrow_table(L.ce, 3, nc, [](int64_t j) { return (3 * j) / 2; }, L.ce16p, lim);
    L.wide_edges = narrow_table(L.ec, 2, ne, [](int64_t j) { return (2 * j) / 3; }, L.ec16p, lim);
    L.wide_edges += narrow_table(L.ev, 2, ne, [](int64_t j) { return j / 3; }, L.ev16p, lim);
    L.wide_edges += narrow_table(L.re, KR, ne, [](int64_t j) { return j; }, L.re16p, lim);
    L.wide_vertices = narrow_table(L.ve, KV, nv, [](int64_t j) { return 3 * j; }, L.ve16p, lim);
    L.pair_reason = "ICON triangle model: RBF tangential velocity, bilinear kinetic energy";
    return true;
}

bool build_mpas_layout(const mpas_ops_desc* d, MpasLayout& L, std::string& err) {
    if (!d) {
        err = "null descriptor";
        return false;
    }
    if (d->nc <= 0 || d->ne <= 0 || d->nv <= 0 || d->nlev <= 0 || d->nlev > 4096 || d->kc <= 0 || d->kc > 16 ||
        d->kv <= 0 || d->kv > 16 || d->kt <= 0 || d->kt > 32 || d->nc >= (int64_t(1) << 31) ||
        d->ne >= (int64_t(1) << 31) || d->nv >= (int64_t(1) << 31) ||
        (d->nlev + 1) * std::max(d->ne, d->nc) >= (int64_t(1) << 40)) {
        err = "bad sizes or table widths";
        return false;
    }
    if (!d->cell_edges || !d->kinetic || !d->inner || !d->edge_cells || !d->edge_vertices || !d->edge_weight ||
        !d->inverse_dual || !d->vertex_edges || !d->curl || !d->vertex_cells || !d->vertex_weight || !d->coriolis ||
        !d->trisk_index || !d->trisk || !d->rhombus || !d->edge_coriolis || !d->dz || !d->dzh || !d->wlow ||
        !d->edge_dz || !d->edge_dzh) {
        err = "missing table";
        return false;
    }
    if (d->vertical < 0 || d->vertical > 2) {
        err = "vertical must be 0 (advective), 1 (vector invariant) or 2 (ICON-IAP)";
        return false;
    }
    if (d->orography < 0 || d->orography > 2) {
        err = "orography must be 0 (flat), 1 (pressure gradient along the levels) or 2 (at constant height)";
        return false;
    }
    if (d->orography && (!d->z_full || !d->pi_ref || !d->ground_q || !d->ddxn)) {
        err = "orography needs z_full, pi_ref, ground_q and ddxn";
        return false;
    }
    if (d->orography == 2) {
        if (!d->zlev || !d->zwt || d->nlev < 2) {
            err = "the pressure gradient at constant height needs zlev, zwt and at least two levels";
            return false;
        }
        for (int64_t t = 0; t < 2 * d->ne * d->nlev; ++t)
            if (d->zlev[t] < 0 || d->zlev[t] > d->nlev - 2) {
                err = "level out of range in zlev";
                return false;
            }
    }
    const bool vk = d->cell_vertices != nullptr;
    if (vk && (!d->vk_kites || !d->vk_weights)) {
        err = "vertex kinetic energy needs cell_vertices, vk_kites and vk_weights";
        return false;
    }
    const int64_t nc = d->nc, ne = d->ne, nv = d->nv;
    const int KC = int(d->kc), KV = int(d->kv), KT = int(d->kt), NL = int(d->nlev);
    struct Check {
        const int64_t* t;
        int64_t count, n;
        const char* name;
    } checks[] = {{d->cell_edges, KC * nc, ne, "cell_edges"},      {d->edge_cells, 2 * ne, nc, "edge_cells"},
                  {d->edge_vertices, 2 * ne, nv, "edge_vertices"}, {d->vertex_edges, KV * nv, ne, "vertex_edges"},
                  {d->vertex_cells, KV * nv, nc, "vertex_cells"},  {vk ? d->cell_vertices : nullptr, KC * nc, nv,
                                                                     "cell_vertices"}};
    for (auto& c : checks)
        if (c.t && !in_range(c.t, c.count, c.n)) {
            err = std::string("index out of range in ") + c.name;
            return false;
        }
    for (int64_t t = 0; t < 3 * KT * ne; ++t)
        if (d->trisk_index[t] < 0 || d->trisk_index[t] >= ne) {
            err = "index out of range in trisk_index";
            return false;
        }
    L = MpasLayout();
    L.nc = nc;
    L.ne = ne;
    L.nv = nv;
    L.nlev = NL;
    L.kc = KC;
    L.kv = KV;
    L.kt = KT;
    L.iap_pv = d->iap_pv != 0;
    L.vertical = d->vertical;
    L.five_band = d->five_band != 0;
    L.vk = vk;
    if (d->cell_points) {
        hilbert_cells(nc, d->cell_points, L.cperm);
    } else {
        L.cperm.resize(nc);
        std::iota(L.cperm.begin(), L.cperm.end(), int64_t(0));
    }
    std::vector<char> seen(ne, 0);
    L.eperm.reserve(ne);
    for (int64_t j = 0; j < nc; ++j)
        for (int k = 0; k < KC; ++k) {
            if (d->kinetic[k + KC * L.cperm[j]] == 0.0 && d->inner[k + KC * L.cperm[j]] == 0.0) continue;  // pad
            const int64_t e = d->cell_edges[k + KC * L.cperm[j]];
            if (!seen[e]) {
                seen[e] = 1;
                L.eperm.push_back(e);
            }
        }
    for (int64_t e = 0; e < ne; ++e)
        if (!seen[e]) L.eperm.push_back(e);
    std::vector<char> vseen(nv, 0);
    L.vperm.reserve(nv);
    for (int64_t j = 0; j < ne; ++j)
        for (int s = 0; s < 2; ++s) {
            const int64_t v = d->edge_vertices[s + 2 * L.eperm[j]];
            if (!vseen[v]) {
                vseen[v] = 1;
                L.vperm.push_back(v);
            }
        }
    for (int64_t v = 0; v < nv; ++v)
        if (!vseen[v]) L.vperm.push_back(v);
    invert(L.cperm, L.cinv);
    invert(L.eperm, L.einv);
    invert(L.vperm, L.vinv);
    // k x n column-major tables -> slot-major k * n + j
    remap_index(d->cell_edges, KC, nc, L.cperm, L.einv, L.ce);
    remap_value(d->kinetic, KC, nc, L.cperm, L.kin);
    remap_value(d->inner, KC, nc, L.cperm, L.inner);
    remap_index(d->edge_cells, 2, ne, L.eperm, L.cinv, L.ec);
    remap_index(d->edge_vertices, 2, ne, L.eperm, L.vinv, L.ev);
    remap_value(d->edge_weight, 2, ne, L.eperm, L.ew);
    remap_value(d->inverse_dual, 1, ne, L.eperm, L.invd);
    remap_value(d->rhombus, 2, ne, L.eperm, L.rh);
    remap_value(d->edge_coriolis, 1, ne, L.eperm, L.fe);
    remap_index(d->vertex_edges, KV, nv, L.vperm, L.einv, L.ve);
    remap_value(d->curl, KV, nv, L.vperm, L.curl);
    remap_index(d->vertex_cells, KV, nv, L.vperm, L.cinv, L.vc);
    remap_value(d->vertex_weight, KV, nv, L.vperm, L.vw);
    remap_value(d->coriolis, 1, nv, L.vperm, L.fv);
    if (vk) {
        remap_index(d->cell_vertices, KC, nc, L.cperm, L.vinv, L.cv);
        remap_value(d->vk_kites, KC, nc, L.cperm, L.vkk);
        remap_value(d->vk_weights, KV, nv, L.vperm, L.vkw);
    }
    // edge-first tables (entry (e, j) at e + ne * j) and points x levels fields: the point
    // index is renumbered, the slot and level axes stay
    auto points = [](const double* old, int64_t n, int64_t m, const std::vector<int64_t>& perm,
                     std::vector<double>& out) {
        out.resize(size_t(n * m));
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < m; ++k)
            for (int64_t p = 0; p < n; ++p) out[size_t(p + n * k)] = old[perm[p] + n * k];
    };
    points(d->trisk, ne, KT, L.eperm, L.trw);
    L.ti.resize(size_t(3 * KT * ne));
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < 3 * KT; ++m)
        for (int64_t e = 0; e < ne; ++e) L.ti[size_t(e + ne * m)] = int32_t(L.einv[d->trisk_index[L.eperm[e] + ne * m]]);
    // two gathers per TRiSK entry where every entry has b = e or b = c (the reference's tables)
    L.paired = true;
    L.ti2.resize(size_t(2 * KT * ne));
    for (int64_t j = 0; j < KT && L.paired; ++j)
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t a = L.ti[size_t(e + ne * j)], b = L.ti[size_t(e + ne * (KT + j))],
                          c = L.ti[size_t(e + ne * (2 * KT + j))];
            L.ti2[size_t(e + ne * j)] = a;
            if (b == int32_t(e)) {
                L.ti2[size_t(e + ne * (KT + j))] = c;
            } else if (b == c) {
                L.ti2[size_t(e + ne * (KT + j))] = ~c;
            } else {
                L.paired = false;
                break;
            }
        }
    if (!L.paired) L.ti2.clear();
    // the slot pattern of every row (mpas_pv_slot), when all of them have one
    L.pv_slots = L.paired && KT == 10;
    if (L.pv_slots) {
        L.tpat.assign(size_t(ne), 0);
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (int64_t e = 0; e < ne; ++e) {
            int found = -1;
            for (int pt = 0; pt < kMpasPvPatterns && found < 0; ++pt) {
                bool ok = true;
                for (int j = 0; j < 10 && ok; ++j) {
                    const int32_t a = L.ti2[size_t(e + ne * j)], x = L.ti2[size_t(e + ne * (KT + j))];
                    if ((mpas_pv_mean(pt) >> j) & 1)
                        ok = x == a;  // (q_e + q_a)
                    else
                        ok = x < 0 && ~x == L.ti2[size_t(e + ne * mpas_pv_slot(pt, j))];  // (q_x + q_x)
                }
                if (ok) found = pt;
            }
            if (found < 0)
                bad = 1;
            else
                L.tpat[size_t(e)] = uint8_t(found);
        }
        if (bad) {
            L.pv_slots = false;
            L.tpat.clear();
        }
    }
    points(d->dz, nc, NL, L.cperm, L.dz);
    points(d->dzh, nc, NL + 1, L.cperm, L.dzh);
    points(d->wlow, nc, NL + 1, L.cperm, L.wlow);
    points(d->edge_dz, ne, NL, L.eperm, L.edz);
    points(d->edge_dzh, ne, NL + 1, L.eperm, L.edzh);
    L.oro = d->orography;
    if (L.oro) {
        points(d->z_full, nc, NL, L.cperm, L.zf);
        points(d->pi_ref, nc, NL, L.cperm, L.piref);
        points(d->ground_q, nc, 3, L.cperm, L.gq);
        points(d->ddxn, ne, NL, L.eperm, L.ddxn);
    }
    if (L.oro == 2) {
        // entry (s, e, k) at s + 2 (e + ne k): the edge index is renumbered, s and k stay
        L.zlev.resize(size_t(2 * ne * NL));
        L.zwt.resize(size_t(2 * ne * NL));
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < NL; ++k)
            for (int64_t e = 0; e < ne; ++e)
                for (int s2 = 0; s2 < 2; ++s2) {
                    const size_t to = size_t(s2 + 2 * (e + ne * k)), from = size_t(s2 + 2 * (L.eperm[e] + ne * k));
                    L.zlev[to] = d->zlev[from];
                    L.zwt[to] = d->zwt[from];
                }
    }
    return true;
}

void build_flux_form(const trisk_model_desc* d, Layout& L) {
    L.gflux = false;
    L.csign.clear();
    if (!d->primal_lengths || L.kc > 16) return;
    const int64_t nc = L.nc, ne = L.ne;
    const int KC = L.kc;
    const double* l = d->primal_lengths;
    std::vector<uint32_t> sign(nc, 0);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t j = 0; j < nc; ++j) {
        const int64_t i = L.cperm[j];
        uint32_t m = 0;
        for (int k = 0; k < KC; ++k) {
            const double div = d->divergence[k + KC * i];
            if (div == 0.0) continue;  // padded slot
            const double sg = div > 0 ? 1.0 : -1.0;
            const double ref = (sg * l[d->cell_edges[k + KC * i]]) / d->cell_areas[i];
            bad |= !(std::fabs(ref - div) <= 1e-14 * std::fabs(div));
            m |= uint32_t(sg > 0 ? 1 : 2) << (2 * k);
        }
        sign[j] = m;
    }
    for (int64_t e = 0; e < ne; ++e) bad |= !(l[e] > 0.0);
    if (bad) return;
    L.len.resize(ne);
    for (int64_t j = 0; j < ne; ++j) L.len[j] = l[L.eperm[j]];
    L.csign = std::move(sign);
    L.gflux = true;
}

void build_pairs(const trisk_model_desc* d, Layout& L, int32_t narrow_limit) {
    L.pair = false;
    if (!d->primal_lengths || !d->dual_lengths || !d->cell_kites) {
        L.pair_reason = "grid fields (primal_lengths, dual_lengths, cell_kites) not given";
        return;
    }
    if (L.vertex_ke || L.gassmann) {
        L.pair_reason = "vertex kinetic energy and the Gassmann PV flux stream their tables";
        return;
    }
    if (L.kc > kPairSlots) {
        L.pair_reason = "cells with more than six edges stream their tables";
        return;
    }
    const int64_t nc = L.nc, ne = L.ne;
    const int KC = L.kc, KT = L.kt;
    const double* l = d->primal_lengths;
    const double* dl = d->dual_lengths;
    std::vector<int32_t> size(nc), neg(nc);
    std::vector<double> W(size_t(kPairs) * nc, 0.0);
    int bad = 0;
    // sizes, signs, divergence (s l)/A and the pair weights of every cell
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < nc; ++i) {
        int n = 0;
        while (n < KC && d->divergence[n + KC * i] != 0.0) ++n;
        for (int k = n; k < KC; ++k) bad |= d->divergence[k + KC * i] != 0.0;
        if (n < 3) bad |= 1;
        size[i] = n;
        int32_t nb = 0;
        double R[kPairSlots];
        for (int k = 0; k < n; ++k) {
            double div = d->divergence[k + KC * i];
            double sg = div > 0 ? 1.0 : -1.0;
            if (sg < 0) nb |= 1 << k;
            int64_t e = d->cell_edges[k + KC * i];
            double ref = (sg * l[e]) / d->cell_areas[i];
            bad |= std::fabs(ref - div) > 1e-14 * std::fabs(div);
            R[k] = d->cell_kites[k + KC * i] / d->cell_areas[i];
        }
        neg[i] = nb;
        for (int k = 0; k < n; ++k) {
            double S = 0.0, sk = (nb >> k) & 1 ? -1.0 : 1.0;
            for (int k2 = k + 1; k2 < n; ++k2) {
                S += R[k2];
                double s2 = (nb >> k2) & 1 ? -1.0 : 1.0;
                W[pair_index(k, k2) + size_t(kPairs) * i] = sk * s2 * (0.5 - S);
            }
        }
    }
    if (bad) {
        L.pair_reason = "divergence differs from (s l_e)/A_i";
        return;
    }
    // every TRiSK entry of edge e, in the reference's row order, against w l_e'/d_e
    auto weight = [&](int64_t c, int k, int k2) {  // w of the walk from slot k to slot k2 in cell c
        return k < k2 ? W[pair_index(k, k2) + size_t(kPairs) * c] : -W[pair_index(k2, k) + size_t(kPairs) * c];
    };
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t e = 0; e < ne; ++e) {
        bad |= std::fabs(d->inverse_dual[e] - 1.0 / dl[e]) > 1e-14 * std::fabs(d->inverse_dual[e]);
        int64_t c0 = d->edge_cells[2 * e], c1 = d->edge_cells[2 * e + 1];
        int64_t cells[2] = {std::min(c0, c1), std::max(c0, c1)};
        double rowmax = 0;
        for (int q = 0; q < KT; ++q) rowmax = std::max(rowmax, std::fabs(d->trisk[q + KT * e]));
        int pos = 0;
        for (int side = 0; side < 2; ++side) {
            int64_t c = cells[side];
            int n = size[c];
            int k = 0;
            while (k < n && d->cell_edges[k + KC * c] != e) ++k;
            if (k == n) {
                bad |= 1;
                break;
            }
            for (int j = 1; j < n; ++j) {
                int k2 = (k + j) % n;
                int64_t e2 = d->cell_edges[k2 + KC * c];
                double coef = weight(c, k, k2) * l[e2] / dl[e];
                if (pos >= KT || d->trisk_edges[pos + KT * e] != e2 ||
                    std::fabs(coef - d->trisk[pos + KT * e]) > 1e-13 * rowmax) {
                    bad |= 1;
                    break;
                }
                ++pos;
            }
        }
        for (int q = pos; q < KT; ++q) bad |= d->trisk[q + KT * e] != 0.0;
    }
    if (bad) {
        L.pair_reason = "trisk or inverse_dual differ from their recomputation from the grid fields";
        return;
    }
    // corners: corner k starts edge k (get_cell_vertices, Plex.jl:226-230) and the loop closes
    std::vector<int64_t> corner(size_t(KC) * nc, 0);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < nc; ++i) {
        int n = size[i];
        for (int k = 0; k < n; ++k) {
            int64_t e = d->cell_edges[k + KC * i];
            corner[k + size_t(KC) * i] = d->edge_vertices[((neg[i] >> k) & 1) + 2 * e];
        }
        for (int k = 0; k < n; ++k) {
            int64_t e = d->cell_edges[k + KC * i];
            int64_t stop = d->edge_vertices[1 - ((neg[i] >> k) & 1) + 2 * e];
            bad |= stop != corner[(k + 1) % n + size_t(KC) * i];
        }
    }
    if (bad) {
        L.pair_reason = "cell edge lists are not closed loops";
        return;
    }
    // device numbering
    L.cmeta.resize(nc);
    L.ccv.assign(size_t(kPairSlots) * nc, 0);
    L.pr.assign(size_t(kPairSlots) * nc, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; ++j) {
        int64_t i = L.cperm[j];
        for (int k = 0; k < size[i]; ++k)
            L.pr[size_t(k) * nc + j] = d->cell_kites[k + KC * i] / d->cell_areas[i];
        L.cmeta[j] = size[i] | (neg[i] << 4);
        for (int k = 0; k < kPairSlots; ++k) {
            int64_t v = k < size[i] ? corner[k + size_t(KC) * i] : corner[size_t(KC) * i];
            L.ccv[size_t(k) * nc + j] = int32_t(L.vinv[v]);
        }
    }
    L.len.resize(ne);
    L.dual.resize(ne);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ne; ++j) {
        L.len[j] = l[L.eperm[j]];
        L.dual[j] = dl[L.eperm[j]];
    }
    // entity records
    const int KV = L.kv;
    if (KC > 6 || KV != 3) {
        L.pair_reason = "the cell-pair records need cells of at most six edges and vertices of three";
        return;
    }
    L.rc_state.assign(size_t(nc) * 6, 0);
    L.rc_kin.assign(size_t(nc) * 6, 0.0);
    L.rc_pce.assign(size_t(nc) * 6, 0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; ++j) {
        for (int k = 0; k < 6; ++k) {
            int32_t e = k < KC ? L.ce[size_t(k) * nc + j] : L.ce[j];
            L.rc_state[6 * j + k] = e;
            L.rc_kin[6 * j + k] = k < KC ? L.kin[size_t(k) * nc + j] : 0.0;
            L.rc_pce[size_t(k) * nc + j] = e;
        }
    }
    const int64_t nv = L.nv;
    L.rv_idx.assign(size_t(nv) * 6, 0);
    L.rv_coef.assign(size_t(nv) * 6, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; ++v) {
        for (int k = 0; k < 3; ++k) {
            L.rv_idx[6 * v + k] = L.ve[size_t(k) * nv + v];
            L.rv_idx[6 * v + 3 + k] = L.vc[size_t(k) * nv + v];
            L.rv_coef[4 * v + k] = L.curl[size_t(k) * nv + v];
            // vertex heights: the first in the 4-wide record, the others in the 2-wide one
            L.rv_coef[k == 0 ? 4 * v + 3 : 4 * nv + 2 * v + (k - 1)] = L.vh[size_t(k) * nv + v];
        }
    }
    L.re_cells.assign(size_t(ne) * 2, 0);
    L.re_lh.assign(size_t(ne) * 2, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e)
        for (int s2 = 0; s2 < 2; ++s2) {
            L.re_cells[2 * e + s2] = L.ec[size_t(s2) * ne + e];
            L.re_lh[2 * e + s2] = L.len[e] * L.eh[size_t(s2) * ne + e];
        }
    // 16-bit deltas
    const int64_t lim = narrow_limit > 0 && narrow_limit < 32767 ? narrow_limit : 32767;
    auto fits = [lim](int64_t d) { return d >= -lim && d <= lim; };
    L.rc_state16.assign(size_t(nc) * 8, 0);
    L.pc_ev16.assign(size_t(6) * nc, 0);
    int64_t wc = 0, wv = 0, we = 0;
#pragma omp parallel for schedule(static) reduction(+ : wc)
    for (int64_t j = 0; j < nc; ++j) {
        bool ok = true;
        for (int k = 0; k < 6; ++k)
            ok = ok && fits(int64_t(L.rc_pce[size_t(k) * nc + j]) - 3 * j) &&
                 fits(int64_t(L.ccv[size_t(k) * nc + j]) - 2 * j);
        if (!ok) {
            L.rc_state16[8 * j] = kWide;
            L.pc_ev16[j] = uint16_t(kWide);
            ++wc;
            continue;
        }
        for (int k = 0; k < 6; ++k) {
            int16_t de = int16_t(int64_t(L.rc_pce[size_t(k) * nc + j]) - 3 * j);
            int16_t dv = int16_t(int64_t(L.ccv[size_t(k) * nc + j]) - 2 * j);
            L.rc_state16[8 * j + k] = de;
            L.pc_ev16[size_t(k) * nc + j] = int32_t(uint32_t(uint16_t(de)) | (uint32_t(uint16_t(dv)) << 16));
        }
    }
    // per-cell records of the cell kernel: meta, the packed deltas, the FP32 kite fractions,
    // the FP32 kinetic-energy weights
    L.rc_pair.assign(size_t(nc) * kPairRecord, 0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; ++j) {
        int32_t* r = L.rc_pair.data() + kPairRecord * j;
        r[0] = L.cmeta[j];
        for (int k = 0; k < 6; ++k) {
            r[1 + k] = L.pc_ev16[size_t(k) * nc + j];
            const float R = float(L.pr[size_t(k) * nc + j]), W = float(L.rc_kin[6 * j + k]);
            std::memcpy(&r[7 + k], &R, 4);
            std::memcpy(&r[13 + k], &W, 4);
        }
    }
    L.rv_idx16.assign(size_t(nv) * 8, 0);
#pragma omp parallel for schedule(static) reduction(+ : wv)
    for (int64_t v = 0; v < nv; ++v) {
        bool ok = true;
        for (int k = 0; k < 3; ++k)
            ok = ok && fits(int64_t(L.rv_idx[6 * v + k]) - (3 * v) / 2) && fits(int64_t(L.rv_idx[6 * v + 3 + k]) - v / 2);
        // slots 6, 7: the bits of f as FP32 (the records serve the FP32-bulk kernels)
        const float f32 = float(L.fv[v]);
        uint32_t fb;
        std::memcpy(&fb, &f32, 4);
        L.rv_idx16[8 * v + 6] = int16_t(uint16_t(fb & 0xffff));
        L.rv_idx16[8 * v + 7] = int16_t(uint16_t(fb >> 16));
        if (!ok) {
            L.rv_idx16[8 * v] = kWide;
            ++wv;
            continue;
        }
        for (int k = 0; k < 3; ++k) {
            L.rv_idx16[8 * v + k] = int16_t(int64_t(L.rv_idx[6 * v + k]) - (3 * v) / 2);
            L.rv_idx16[8 * v + 3 + k] = int16_t(int64_t(L.rv_idx[6 * v + 3 + k]) - v / 2);
        }
    }
    L.re_cells16.assign(ne, 0);
#pragma omp parallel for schedule(static) reduction(+ : we)
    for (int64_t e = 0; e < ne; ++e) {
        int64_t d0 = int64_t(L.re_cells[2 * e]) - e / 3, d1 = int64_t(L.re_cells[2 * e + 1]) - e / 3;
        if (!fits(d0) || !fits(d1)) {
            L.re_cells16[e] = int32_t(uint32_t(uint16_t(kWide)));
            ++we;
            continue;
        }
        L.re_cells16[e] = int32_t(uint32_t(uint16_t(int16_t(d0))) | (uint32_t(uint16_t(int16_t(d1))) << 16));
    }
    L.wide_cells = wc;
    L.wide_vertices = wv;
    L.wide_edges = we;
    // transposed coefficients (adjoint): the forward coefficient of each (entity, neighbour)
    // pair, stored at the neighbour
    L.adj_ecurl.assign(size_t(ne) * 2, 0.0);
    L.adj_ekin.assign(size_t(ne) * 2, 0.0);
    L.eslot.assign(ne, 0);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t e = 0; e < ne; ++e) {
        bad |= L.ec[e] == L.ec[size_t(ne) + e];  // the half sums of e come from two cells
        for (int s2 = 0; s2 < 2; ++s2) {
            const int64_t v = L.ev[size_t(s2) * ne + e], c = L.ec[size_t(s2) * ne + e];
            int k = 0;
            while (k < 3 && L.ve[size_t(k) * nv + v] != e) ++k;
            bad |= k == 3;
            L.adj_ecurl[2 * e + s2] = k < 3 ? L.curl[size_t(k) * nv + v] : 0.0;
            const int n = L.cmeta[c] & 15;
            k = 0;
            while (k < n && L.ce[size_t(k) * nc + c] != e) ++k;
            bad |= k == n;
            L.adj_ekin[2 * e + s2] = k < n ? L.kin[size_t(k) * nc + c] : 0.0;
            L.eslot[e] |= uint8_t((k < n ? k : 0) << (4 * s2));
        }
    }
    L.adj_cv3.assign(size_t(6) * nc, 0);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < nc; ++i) {
        const int n = L.cmeta[i] & 15;
        for (int k = 0; k < n; ++k) {
            const int64_t v = L.ccv[size_t(k) * nc + i];
            int j = 0;
            while (j < 3 && L.vc[size_t(j) * nv + v] != i) ++j;
            bad |= j == 3;
            L.adj_cv3[size_t(k) * nc + i] = int32_t(3 * v + (j < 3 ? j : 0));
        }
    }
    L.adj_vslot.assign(nv, 0);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t v = 0; v < nv; ++v) {
        int32_t slots = 0;
        for (int j = 0; j < 3; ++j) {
            const int64_t c = L.vc[size_t(j) * nv + v];
            const int n = L.cmeta[c] & 15;
            int k = 0;
            while (k < n && L.ccv[size_t(k) * nc + c] != v) ++k;
            bad |= k == n;
            slots |= int32_t(k < n ? k : 0) << (4 * j);
        }
        L.adj_vslot[v] = slots;
    }
    if (bad) {
        L.pair_reason = "vertex, cell and edge tables are not mutually consistent";
        return;
    }
    L.pair = true;
    L.pair_reason = "cell-pair TRiSK weights (no per-edge TRiSK tables)";
}

}  // namespace trisk
