// Host-side re-layout of the TRiSK model tables for the device (shared by the C ABI in
// trisk_device.hip and the layout builder in trisk_layout.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/trisk_hip.h"

namespace trisk {

// Tables in the device numbering (cells along the Hilbert curve, edges and vertices by
// first appearance), entity-major: slot k of entity j at k * n + j.
struct Layout {
    int64_t nc = 0, ne = 0, nv = 0;
    int kc = 0, kv = 0, kt = 0, kg = 0;
    bool vertex_ke = false, gassmann = false, gassmann_mean = false;
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
};

// Validates `d` (every index in range, widths positive) and fills `L`.  Returns false
// with a message on a bad table.
bool build_layout(const trisk_model_desc* d, Layout& L, std::string& err);

}  // namespace trisk
