// Host-side re-layout of the TRiSK model tables for the device (shared by the C ABI in
// trisk_device.hip and the layout builder in trisk_layout.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/trisk_hip.h"

namespace trisk {

// Slot whose edge carries the PV of Gassmann slot k (k < 12): an inner edge takes the other
// inner edge of the shared vertex, an outer edge its inner edge (TRiSKGrid.jl:233-247), which
// gives 3, 0, 0, 0, 3, 3 for the first vertex and 9, 6, 6, 6, 9, 9 for the second.  (constexpr:
// callable in device code too.)
constexpr int gassmann_pv_slot(int k) {
    return (k >= 6 ? 6 : 0) + ((k % 6 == 0 || k % 6 >= 4) ? 3 : 0);
}

// Tables in the device numbering (cells along the Hilbert curve, edges and vertices by
// first appearance), entity-major: slot k of entity j at k * n + j.
struct Layout {
    int64_t nc = 0, ne = 0, nv = 0;
    int kc = 0, kv = 0, kt = 0, kg = 0;
    bool vertex_ke = false, gassmann = false, gassmann_mean = false;
    // every cell kinetic-energy weight is 0 (vertex kinetic energy with alpha = 0): the cell part
    // of K is 0, so the tendency need not gather u for it
    bool cell_ke_zero = false;
    // Gassmann with 12 slots in the order of gassmann_coefficients (per vertex of e: inner f1,
    // the two outer edges of f1, inner f2, the two outer edges of f2): slot k takes the PV of
    // its slot gassmann_pv_slot(k), so the PV-edge table is not needed on the device
    bool gassmann_slots = false;
    double gravity = 0;
    std::vector<int64_t> cperm, eperm, vperm;  // new -> reference
    std::vector<int64_t> cinv, einv, vinv;     // reference -> new
    // cells
    std::vector<int32_t> ce, cv;
    std::vector<double> div, kin, vkk, area_c;
    // edges
    std::vector<int32_t> ec, ev, te, ge, gp;
    std::vector<double> eh, invd, tw, vkh, gw, grh, gf;
    // vertices
    std::vector<int32_t> ve, vc;
    std::vector<double> curl, vh, fv, vkw, area_v;
    // the index tables again as 16-bit deltas from a linear base of the entity (cell i: 3i
    // for edges, 2i for vertices; vertex v: 3v/2 for edges, v/2 for cells; edge e: e/3 for
    // cells, 2e/3 for vertices, e for its TRiSK and Gassmann edges), two per int32 word
    // (slots 2w | 2w+1 << 16), slot-major by word; an entity with any delta outside
    // +-32767 has kWide in the low half of word 0
    std::vector<int32_t> ce16p, cv16p, ec16p, ev16p, te16p, ge16p, gp16p, ve16p, vc16p;
    std::vector<int32_t> re16p;  // ICON: RBF edges (base e)
    int64_t wide_table_entities = 0;
    // flux form of the FP32-bulk table kernels: the edge kernel stores G_e = l_e F_e (edge
    // heights scaled by l_e on upload, momentum weights by 1/l of their edge) and the
    // continuity is dh_i = -(sum_k s_k G_k)/A_i, which reads 2 bits per slot (csign: 1 for
    // s = +1, 2 for s = -1, 0 for a padded slot) instead of the FP64 divergence weights.
    // Needs the edge lengths and the divergence equal to (s l_e)/A_i; len holds l_e.
    bool gflux = false;
    std::vector<uint32_t> csign;  // nc
    // cell-pair path (TRiSK PV flux and cell kinetic energy only): per cell the 15
    // antisymmetric TRiSK weights w_{k,k'} (k < k') of its edge pairs
    bool pair = false;
    std::string pair_reason;             // why the table path is used
    std::vector<int32_t> cmeta;          // per cell: size | negative-sign bits << 4
    std::vector<int32_t> ccv;            // corner k of cell (start of edge k), kc x nc
    std::vector<double> pr;              // 6 x nc kite fractions R_k = kite_k / A_i (pads 0), from
                                         // which the kernels rebuild the pair weights
    std::vector<double> len, dual;       // l_e, d_e
    // the same tables as entity records (AoS) for vector loads
    std::vector<int32_t> rc_state;       // nc x 6: cell edges
    std::vector<double> rc_kin;          // nc x 6: kinetic-energy weights
    std::vector<int32_t> rv_idx;         // nv x 6: vertex edges, vertex cells
    std::vector<double> rv_coef;         // nv x 4: curl (3), vertex height 0; then nv x 2: vertex heights 1, 2
    std::vector<int32_t> re_cells;       // ne x 2: left, right cell
    std::vector<double> re_lh;           // ne x 2: l_e * edge_height
    std::vector<int32_t> rc_pce;         // 6 x nc: cell edges, pads the first edge
    std::vector<uint8_t> eslot;          // ne: slot of e in its left cell | slot in its right cell << 4
    // 16-bit index deltas from a linear base (cell j: edges 3j, corners 2j; vertex v: edges
    // 3v/2, cells v/2; edge e: cells e/3); an entity with a delta outside +-32767 has the
    // sentinel -32768 in its first slot and is read from the int32 tables instead
    std::vector<int16_t> rc_state16;     // nc x 8: edges (6), 0, 0
    std::vector<int16_t> rv_idx16;       // nv x 8: edges (3), cells (3), the FP32 bits of f (2)
    std::vector<int32_t> re_cells16;     // ne: left | right << 16
    std::vector<int32_t> pc_ev16;        // 6 x nc: edge | corner << 16
    std::vector<int32_t> rc_pair;        // nc x kPairRecord: meta, 6 x (edge | corner << 16), 6 x FP32
                                         // bits of R, 6 x FP32 bits of the kinetic-energy weights
    int64_t wide_cells = 0, wide_vertices = 0, wide_edges = 0;
    // transposed coefficients for the adjoint of the cell-pair path
    std::vector<double> adj_ecurl;       // ne x 2: curl coefficient of edge e in vertex ev[s]
    std::vector<double> adj_ekin;        // ne x 2: kinetic-energy weight of edge e in cell ec[s]
    std::vector<int32_t> adj_cv3;        // 6 x nc: 3 v + the slot of cell i among the cells of its corner v
    std::vector<int32_t> adj_vslot;      // nv: corner index of v in its cell j at bits 4j..4j+3
    // ICON triangle model (icon_create): kc = 3; ce, div, ec, ev, eh (c_lin_e), invd, ve,
    // curl, vc, vh, fv and the areas as above, plus
    bool icon = false;
    int kr = 0;                          // RBF stencil width
    std::vector<double> bln;             // 3 x nc: bilinear edge-to-cell weights
    std::vector<double> grad;            // 2 x ne: kinetic-energy gradient coefficients
    std::vector<int32_t> re;             // kr x ne: RBF stencil edges
    std::vector<double> rw;              // kr x ne: RBF weights
    std::vector<double> fe;              // ne: Coriolis parameter at the edges
    std::vector<double> damp;            // ne: divergence damping (empty: none)
};

constexpr int16_t kWide = -32768;

constexpr int kPairSlots = 6;   // cells of up to six edges
constexpr int kPairs = 15;      // pairs k < k' of six slots, in the order of pair_index()
constexpr int kPairRecord = 19; // words of a cell record of the FP32-bulk cell kernel
inline int pair_index(int k, int k2) {  // k < k2 < 6
    return k * (11 - k) / 2 + (k2 - k - 1);
}

// Validates `d` (every index in range, widths positive) and fills `L`.  Returns false
// with a message on a bad table.
bool build_layout(const trisk_model_desc* d, Layout& L, std::string& err);

// Sets L.gflux, L.csign and L.len when d gives the edge lengths and every divergence weight
// is (s l_e)/A_i to 1e-14 (cells of up to 16 edges); otherwise L.gflux stays false.
void build_flux_form(const trisk_model_desc* d, Layout& L);

// Validates an ICON descriptor (triangles, every index in range) and fills the generic
// fields of `L` (sizes, permutations, ce, div, ec, ev, eh, invd, ve, curl, vc, vh, fv,
// areas) and the ICON ones, in the same numbering as build_layout.
bool build_icon_layout(const icon_model_desc* d, Layout& L, std::string& err, int32_t narrow_limit = 32767);

// Slot patterns of the ten-entry TRiSK rows of the MPAS-Gassmann operators on hexagon meshes
// (the edge-first rows of trisk_index with the ICON-IAP PV edges, TRiSKGrid.jl:318-362): in
// pattern p, slot j is a mean entry (q_e + q_a, with a the slot's own flux edge) when bit j of
// mpas_pv_mean(p) is set, and otherwise takes (q_x + q_x) with x the flux edge of slot
// mpas_pv_slot(p, j).  Pattern 0 holds every edge not next to a pentagon, 1 the edges next to
// one, 2 the TRiSK mean flux (every slot a mean entry).
constexpr int kMpasPvPatterns = 3;
constexpr int mpas_pv_slot(int p, int j) {
    constexpr int8_t pos[2][10] = {{9, 0, 2, 4, 5, 4, 5, 7, 9, 0}, {8, 0, 3, 4, 3, 4, 6, 8, 0, 9}};
    return p == 2 ? j : pos[p][j];
}
constexpr uint32_t mpas_pv_mean(int p) {
    return p == 0 ? (1u << 2) | (1u << 7) : p == 1 ? (1u << 6) | (1u << 9) : 0x3ffu;
}

// Tables of the MPAS-Gassmann operators (mpas_ops_create) in the device numbering: cells
// along the Hilbert curve, edges and vertices by first appearance; entity-major tables
// (slot k of entity j at k * n + j, int32 indices) and metric fields point-fastest per level.
struct MpasLayout {
    int64_t nc = 0, ne = 0, nv = 0;
    int nlev = 0, kc = 0, kv = 0, kt = 0;
    int iap_pv = 0, vertical = 0, five_band = 0;
    bool vk = false;
    std::vector<int64_t> cperm, eperm, vperm, cinv, einv, vinv;
    std::vector<int32_t> ce, cv, ec, ev, ti, ve, vc;  // ti: slot c (0..2) of TRiSK entry j at (c * kt + j) * ne + e
    // the TRiSK entries (a, b, c) with b the edge itself or b == c, as two indices: a, and x
    // (b == e: the PV pair is (q_e, q_x)) or ~x (b == c: (q_x, q_x)); at (c * kt + j) * ne + e
    bool paired = false;
    std::vector<int32_t> ti2;
    // every row of ten paired entries in one of the fixed slot patterns (mpas_pv_slot): per
    // edge its pattern; the vn tendency then takes every PV from the (F, q) pairs it gathers
    bool pv_slots = false;
    std::vector<uint8_t> tpat;
    std::vector<double> kin, inner, vkk, dz, dzh, wlow;
    std::vector<double> ew, invd, trw, rh, fe, edz, edzh;
    std::vector<double> curl, vw, fv, vkw;
    // orography: 0 flat, 1 pressure gradient along the levels, 2 at constant height
    int oro = 0;
    std::vector<double> zf, piref, gq, ddxn, zwt;  // points x levels (gq: cells x 3; zwt: 2 x ne x nlev)
    std::vector<int32_t> zlev;                      // 2 x ne x nlev, 0-based levels
};
bool build_mpas_layout(const mpas_ops_desc* d, MpasLayout& L, std::string& err);

// Builds the cell-pair tables from the grid fields: w_{k,k'} = s_k s_k' (1/2 - sum of the
// kite fractions of the corners k+1..k') (TRiSKGrid.jl:140-155), and checks that
// w l_e'/d_e reproduces every entry of the caller's trisk table (to 1e-13 of the row's
// largest entry), the divergence (s l_e)/A_i and 1/d_e; otherwise L.pair stays false and
// L.pair_reason says why.
void build_pairs(const trisk_model_desc* d, Layout& L, int32_t narrow_limit = 32767);

}  // namespace trisk
