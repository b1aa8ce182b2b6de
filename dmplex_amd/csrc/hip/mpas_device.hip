// The explicit stage tendencies of MPAS-Gassmann on gfx950 (slow_tendencies!, src/Nonhydrostatic/
// MPAS/Gassmann/Stepping.jl:22-56): the TRiSK operators of src/Nonhydrostatic/MPAS/Operators.jl
// and vn_tendency_kernel! of Gassmann/Tendencies.jl with a level dimension, on flat ground or over
// terrain-following levels (the contravariant correction and pi' of Stepping.jl:58-76).
//
// Fields are points x levels with the point index fastest (the reference's (n, nlev) arrays),
// in the device numbering (cells along the Hilbert curve, edges and vertices by first
// appearance).  A thread owns one point and a chunk of levels (8, 32 for the vn tendency): it
// loads the point's table row once and reuses it for every level of its chunk (the
// reference's blocks of 32 points x 8 levels share the rows through the cache instead), and
// at every level the 64 threads of a wave gather neighbours of 64 consecutive points of one
// level plane.  The
// launches, in the reference's order of operations:
//   A  mpas_cell_vertex_kernel       cells: K, theta, pi [+ K_w]; vertices: zeta, rho_v, q_v [, K_v]
//   A2 mpas_cell_from_vertex_kernel  cells: K += vertex part [+ K_w]   (vertex kinetic energy)
//   B  mpas_edge_kernel              edges: q_e, rho_e, F = rho_e vn [, eta at the interfaces]
//   C  mpas_w_tendency_kernel        cells: w tendency at the interfaces (vector invariant or advective)
//   E  mpas_w_kinetic_kernel         cells: K += K_w after C (five-band solver)
//   D  mpas_vn_tendency_kernel       edges: vn tendency
// Every formula keeps the reference's association (no contraction), so the FP64 results equal
// the reference algorithm bit for bit, up to the last-bit differences of pow between libms in
// the Exner pressure.
#include <cmath>
#include <cstring>

#include "trisk_device.hpp"

namespace mpas {

using trisk::DevBuf;
using trisk::kBlock;
using trisk::kXcd;
using trisk::ld;
using trisk::set_err;
using trisk::Tile;
using trisk::tile_of;
using trisk::tiles_per_xcd;

constexpr double Rd = 287.0, Cpd = 1004.0, Cvd = Cpd - Rd, P0 = 1.0e5;
// levels per thread of the cell and edge kernels (1, 2, 4, 16 and 32 measured slower)
constexpr int kLevels = 8;
// The vn tendency reuses its ten-entry TRiSK row over several levels (measured at hexagon 7 x
// 30 levels with the (F, q_e) pair gathers: 0.325 ms per call with 32 levels per thread, 0.326
// with 30, 0.329 with 16, 0.347 with 8).
constexpr int kLevelsVn = 32;

struct Params {
    int64_t nc, ne, nv;
    int nlev, kc, kv, kt;
    int iap_pv, vertical, vk;
    // cells
    const int32_t* ce;
    const int32_t* cv;
    const double* kin;
    const double* inner;
    const double* vkk;
    const double* dz;
    const double* dzh;
    const double* wlow;
    // edges
    const int32_t* ec;
    const int32_t* ev;
    const int32_t* ti;
    const int32_t* ti2;  // paired TRiSK entries (a, x or ~x), or NULL
    const double* ew;
    const double* invd;
    const double* trw;
    const double* rh;
    const double* fe;
    const double* edz;
    const double* edzh;
    const double* edzp;  // edges x nlev: (edz of layer k, edzh of interface k + 1), the vn tendency's
    // vertices
    const int32_t* ve;
    const int32_t* vc;
    const double* curl;
    const double* vw;
    const double* fv;
    const double* vkw;
    // orography: 0 flat, 1 pressure gradient along the levels, 2 at constant height
    int oro;
    const double* zf;     // cells x nlev
    const double* piref;  // cells x nlev
    const double* gq;     // cells x 3
    const double* ddxn;   // edges x nlev
    const int32_t* zlev;  // 2 x ne x nlev
    const double* zwt;    // 2 x ne x nlev
    // the slot pattern of every TRiSK row (trisk::mpas_pv_slot), or NULL
    const uint8_t* tpat;
    // work fields
    double* f[MPAS_FIELDS];
    // Interleaved pairs of work fields that a later kernel gathers together (one 16-byte load
    // each); mpas_ops_field splits them into their MPAS_* fields on request:
    double* fq;  // edges x nlev: (F = rho_e vn, q_e), gathered by the vn tendency
    double* zr;  // vertices x nlev: (zeta_v, rho_v), gathered by the edge kernel
    double* re;  // edges x (nlev + 1): (rho_e of layer j (0 at the top), eta at interface j), the w tendency's
    double* tp;  // cells x nlev: (theta, pi), gathered by the vn tendency
};

__device__ __forceinline__ double sq(double x) { return x * x; }

// K_w of layer k of cell i: (dzh_k w_k^2 + dzh_{k+1} w_{k+1}^2) / 4dz_k (w_kinetic_kernel!, Operators.jl:38-42)
__device__ __forceinline__ double w_kinetic(const Params& p, const double* __restrict__ w, int64_t i, int k) {
    const int64_t a = i + p.nc * k, b = i + p.nc * (k + 1);
    return (p.dzh[a] * sq(w[a]) + p.dzh[b] * sq(w[b])) / (4 * p.dz[a]);
}

// ---------------------------------------------------------------------------------
// A: diagnose_kernel! (Operators.jl:6-15) [+ w_kinetic_kernel!] | vertex_pv_kernel! (:88-101)
// [+ vertex_kinetic_kernel! (:18-25)]

template <int KC, int KV>
__global__ __launch_bounds__(kBlock) void mpas_cell_vertex_kernel(Params p, const double* __restrict__ vn,
                                                                  const double* __restrict__ w,
                                                                  const double* __restrict__ rho,
                                                                  const double* __restrict__ rt, int64_t t0,
                                                                  int64_t t1, int add_kw) {
    const Tile t = tile_of(blockIdx.x, t0, t1);
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    if (t.range == 0) {
        const int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const int kc = KC > 0 ? KC : p.kc;
        int32_t E[KC > 0 ? KC : 1];
        double W[KC > 0 ? KC : 1];
        if constexpr (KC > 0) {
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                E[j] = ld(p.ce + j * p.nc + i);
                W[j] = ld(p.kin + j * p.nc + i);
            }
        }
        for (int k = k0; k < k1; ++k) {
            double s = 0.0;
            if constexpr (KC > 0) {
#pragma unroll
                for (int j = 0; j < KC; ++j) s += W[j] * sq(vn[E[j] + p.ne * k]);
            } else {
                for (int j = 0; j < kc; ++j) s += p.kin[j * p.nc + i] * sq(vn[p.ce[j * p.nc + i] + p.ne * k]);
            }
            const int64_t a = i + p.nc * k;
            double K = s;
            if (add_kw) K = K + w_kinetic(p, w, i, k);
            p.f[MPAS_K][a] = K;
            reinterpret_cast<double2*>(p.tp)[a] = double2{rt[a] / rho[a], pow(Rd * rt[a] / P0, Rd / Cvd)};
        }
    } else {
        const int64_t v = t.first + threadIdx.x;
        if (v >= p.nv) return;
        const int kv = KV > 0 ? KV : p.kv;
        int32_t VE[KV > 0 ? KV : 1], VC[KV > 0 ? KV : 1];
        double Cu[KV > 0 ? KV : 1], Vw[KV > 0 ? KV : 1], Kw[KV > 0 ? KV : 1];
        if constexpr (KV > 0) {
#pragma unroll
            for (int j = 0; j < KV; ++j) {
                VE[j] = ld(p.ve + j * p.nv + v);
                VC[j] = ld(p.vc + j * p.nv + v);
                Cu[j] = ld(p.curl + j * p.nv + v);
                Vw[j] = ld(p.vw + j * p.nv + v);
                Kw[j] = p.vk ? ld(p.vkw + j * p.nv + v) : 0.0;
            }
        }
        const double f = ld(p.fv + v);
        for (int k = k0; k < k1; ++k) {
            double z = 0.0, eta = 0.0, kv_s = 0.0;
            if constexpr (KV > 0) {
#pragma unroll
                for (int j = 0; j < KV; ++j) z += Cu[j] * vn[VE[j] + p.ne * k];
#pragma unroll
                for (int j = 0; j < KV; ++j) eta += Vw[j] * rho[VC[j] + p.nc * k];
                if (p.vk) {
#pragma unroll
                    for (int j = 0; j < KV; ++j) kv_s += Kw[j] * sq(vn[VE[j] + p.ne * k]);
                }
            } else {
                for (int j = 0; j < kv; ++j) z += p.curl[j * p.nv + v] * vn[p.ve[j * p.nv + v] + p.ne * k];
                for (int j = 0; j < kv; ++j) eta += p.vw[j * p.nv + v] * rho[p.vc[j * p.nv + v] + p.nc * k];
                if (p.vk)
                    for (int j = 0; j < kv; ++j) kv_s += p.vkw[j * p.nv + v] * sq(vn[p.ve[j * p.nv + v] + p.ne * k]);
            }
            const int64_t a = v + p.nv * k;
            reinterpret_cast<double2*>(p.zr)[a] = double2{z, eta};
            p.f[MPAS_QV][a] = (z + f) / eta;
            if (p.vk) p.f[MPAS_KV][a] = kv_s;
        }
    }
}

// A2: cell_from_vertex_kernel! (Operators.jl:28-35) [+ w_kinetic_kernel!]
__global__ __launch_bounds__(kBlock) void mpas_cell_from_vertex_kernel(Params p, const double* __restrict__ w,
                                                                       int64_t t0, int add_kw) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    for (int k = k0; k < k1; ++k) {
        double s = 0.0;
        for (int j = 0; j < p.kc; ++j) s += p.vkk[j * p.nc + i] * p.f[MPAS_KV][p.cv[j * p.nc + i] + p.nv * k];
        const int64_t a = i + p.nc * k;
        double K = p.f[MPAS_K][a] + s;
        if (add_kw) K = K + w_kinetic(p, w, i, k);
        p.f[MPAS_K][a] = K;
    }
}

// ---------------------------------------------------------------------------------
// B: edge_pv_kernel! (:104-115), edge_density_kernel! (:55-62), eta_kernel! (:44-52)

__global__ __launch_bounds__(kBlock) void mpas_edge_kernel(Params p, const double* __restrict__ vn,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ rho, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    const int32_t l = ld(p.ec + e), r = ld(p.ec + p.ne + e);
    const int32_t v1 = ld(p.ev + e), v2 = ld(p.ev + p.ne + e);
    const double ew1 = ld(p.ew + e), ew2 = ld(p.ew + p.ne + e);
    const double rh1 = ld(p.rh + e), rh2 = ld(p.rh + p.ne + e);
    const double fe = ld(p.fe + e), invd = ld(p.invd + e);
    for (int k = k0; k < k1; ++k) {
        double q;
        if (p.iap_pv) {
            const double2 a1 = reinterpret_cast<const double2*>(p.zr)[v1 + p.nv * k];
            const double2 a2 = reinterpret_cast<const double2*>(p.zr)[v2 + p.nv * k];
            const double z = rh1 * a1.x + rh2 * a2.x;
            const double eta = rh1 * a1.y + rh2 * a2.y;
            q = (z + fe) / eta;
        } else {
            q = (p.f[MPAS_QV][v1 + p.nv * k] + p.f[MPAS_QV][v2 + p.nv * k]) / 2;
        }
        const int64_t a = e + p.ne * k;
        const double re = ew1 * rho[l + p.nc * k] + ew2 * rho[r + p.nc * k];
        reinterpret_cast<double2*>(p.fq)[a] = double2{re * vn[a], q};
        // (rho_e, eta) at layer and lower interface k; eta only with the vector-invariant coupling
        double eta = 0.0;
        if (p.vertical > 0 && k >= 1)
            eta = (vn[e + p.ne * k] - vn[e + p.ne * (k - 1)]) / ld(p.edzh + e + p.ne * k) -
                  (w[r + p.nc * k] - w[l + p.nc * k]) * invd;
        reinterpret_cast<double2*>(p.re)[a] = double2{re, eta};
    }
    if (k1 == p.nlev) reinterpret_cast<double2*>(p.re)[e + p.ne * p.nlev] = double2{0.0, 0.0};  // the top
}

// ---------------------------------------------------------------------------------
// C: w_tendency_vi_kernel! (:69-85) or w_tendency_kernel! (:118-133) at the interfaces

template <int KC>
__global__ __launch_bounds__(kBlock) void mpas_w_tendency_kernel(Params p, const double* __restrict__ vn,
                                                                 const double* __restrict__ w, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    const int j1 = k1 == p.nlev ? p.nlev + 1 : k1;
    const int kc = KC > 0 ? KC : p.kc;
    constexpr int N = KC > 0 ? KC : 16;
    int32_t E[N], L[N], R[N];
    double Wn[N], Id[N];
    const bool vi = p.vertical > 0;
    for (int n = 0; n < kc; ++n) {
        E[n] = ld(p.ce + n * p.nc + i);
        Wn[n] = vi ? ld(p.inner + n * p.nc + i) : ld(p.kin + n * p.nc + i);
        if (!vi) {
            L[n] = p.ec[E[n]];
            R[n] = p.ec[p.ne + E[n]];
            Id[n] = p.invd[E[n]];
        }
    }
    const bool mw = p.vertical == 2;
    // vn, rho_e of the cell's edges and K of the layer below the interface, carried from one
    // interface to the next (the layer above interface j is the one below interface j + 1)
    double VP[N], RP[N], Kp = 0.0;
    const double2* RE = reinterpret_cast<const double2*>(p.re);
    if (vi && k0 >= 1) {
        for (int n = 0; n < kc; ++n) {
            VP[n] = vn[E[n] + p.ne * (k0 - 1)];
            RP[n] = mw ? RE[E[n] + p.ne * (k0 - 1)].x : 1.0;
        }
        Kp = p.f[MPAS_K][i + p.nc * (k0 - 1)];
    }
#pragma unroll
    for (int q = 0; q <= kLevels; ++q) {
        const int j = k0 + q;
        if (j >= j1) break;
        double tw = 0.0;
        if (vi && j == 0) {
            for (int n = 0; n < kc; ++n) {
                VP[n] = vn[E[n]];
                RP[n] = mw ? RE[E[n]].x : 1.0;
            }
            Kp = p.f[MPAS_K][i];
        }
        if (j >= 1 && j <= p.nlev - 1) {
            double s = 0.0;
            if (vi) {
                const double wl = p.wlow[i + p.nc * j];
                for (int n = 0; n < kc; ++n) {
                    const int64_t up = E[n] + p.ne * j;
                    const double rl = RP[n], vl = VP[n];
                    const double2 rn = RE[up];
                    const double ru = mw ? rn.x : 1.0, vu = vn[up];
                    s += Wn[n] * rn.y * (rl * vl + ru * vu) / (2 * (wl * rl + (1 - wl) * ru));
                    VP[n] = vu;
                    RP[n] = ru;
                }
                const double Ku = p.f[MPAS_K][i + p.nc * j];
                tw = s - (Ku - Kp) / p.dzh[i + p.nc * j];
                Kp = Ku;
            } else {
                for (int n = 0; n < kc; ++n) {
                    const double vm = (vn[E[n] + p.ne * (j - 1)] + vn[E[n] + p.ne * j]) / 2;
                    s += 2 * Wn[n] * vm * (w[R[n] + p.nc * j] - w[L[n] + p.nc * j]) * Id[n];
                }
                tw = -s - w[i + p.nc * j] * (w[i + p.nc * (j + 1)] - w[i + p.nc * (j - 1)]) /
                              (p.dz[i + p.nc * (j - 1)] + p.dz[i + p.nc * j]);
            }
        }
        p.f[MPAS_TW][i + p.nc * j] = tw;
    }
}

// E: K += K_w after the w tendency (the five-band solver keeps K_w out of the w equation)
__global__ __launch_bounds__(kBlock) void mpas_w_kinetic_kernel(Params p, const double* __restrict__ w, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    for (int k = k0; k < k1; ++k) {
        const int64_t a = i + p.nc * k;
        p.f[MPAS_K][a] = p.f[MPAS_K][a] + w_kinetic(p, w, i, k);
    }
}

// ---------------------------------------------------------------------------------
// Orography: the contravariant correction and the Exner deviation (Stepping.jl:58-76)

// concorr_cell_kernel! (Operators.jl:135-144): w_c = sum_n inner[n, i] vn[e, k] ddxn[e, k]
template <int KC>
__global__ __launch_bounds__(kBlock) void mpas_concorr_cell_kernel(Params p, const double* __restrict__ vn,
                                                                   int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    const int kc = KC > 0 ? KC : p.kc;
    constexpr int N = KC > 0 ? KC : 16;
    int32_t E[N];
    double In[N];
    for (int n = 0; n < kc; ++n) {
        E[n] = ld(p.ce + n * p.nc + i);
        In[n] = ld(p.inner + n * p.nc + i);
    }
    for (int k = k0; k < k1; ++k) {
        double s = 0.0;
        for (int n = 0; n < kc; ++n) {
            const int64_t a = E[n] + p.ne * k;
            s += In[n] * vn[a] * p.ddxn[a];
        }
        p.f[MPAS_WCF][i + p.nc * k] = s;
    }
}

// concorr_interface_kernel! (CGridOperators.jl:202-208): w_c at the interfaces, linear inside,
// quadratic to the ground, zero at the lid; and contravariant_w_kernel! (:210-214), w - w_c
__global__ __launch_bounds__(kBlock) void mpas_concorr_interface_kernel(Params p, const double* __restrict__ w,
                                                                        int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    const int j1 = k1 == p.nlev ? p.nlev + 1 : k1;  // the last chunk also has the top interface
    const double* wcf = p.f[MPAS_WCF];
    for (int j = k0; j < j1; ++j) {
        const int64_t a = i + p.nc * j;
        double wc;
        if (j >= 1 && j <= p.nlev - 1)
            wc = p.wlow[a] * wcf[a - p.nc] + (1 - p.wlow[a]) * wcf[a];
        else if (j == 0)
            wc = p.nlev >= 3 ? ld(p.gq + i) * wcf[i] + ld(p.gq + p.nc + i) * wcf[i + p.nc] +
                                   ld(p.gq + 2 * p.nc + i) * wcf[i + 2 * p.nc]
                             : wcf[i];
        else
            wc = 0.0;
        p.f[MPAS_WCH][a] = wc;
        p.f[MPAS_WHAT][a] = w[a] - wc;
    }
}

// exner_deviation_centred_kernel! (CGridOperators.jl:93-100): pi' = pi - pi_ref and its
// centred (one-sided at the ends) vertical derivative
__global__ __launch_bounds__(kBlock) void mpas_exner_deviation_kernel(Params p, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    const double2* TP = reinterpret_cast<const double2*>(p.tp);  // pi at .y
    for (int k = k0; k < k1; ++k) {
        const int kl = k > 0 ? k - 1 : 0, ku = k < p.nlev - 1 ? k + 1 : p.nlev - 1;
        const int64_t a = i + p.nc * k, al = i + p.nc * kl, au = i + p.nc * ku;
        p.f[MPAS_PIP][a] = TP[a].y - p.piref[a];
        p.f[MPAS_DPIDZ][a] = ((TP[au].y - p.piref[au]) - (TP[al].y - p.piref[al])) / (p.zf[au] - p.zf[al]);
    }
}

// w_concorr_tendency_kernel! (Operators.jl:146-152): tw += w_c dw/dz at the interior interfaces
__global__ __launch_bounds__(kBlock) void mpas_w_concorr_kernel(Params p, const double* __restrict__ w, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const int k0 = int(blockIdx.y) * kLevels, k1 = min(k0 + kLevels, p.nlev);
    for (int j = k0 > 1 ? k0 : 1; j < k1; ++j) {
        const int64_t a = i + p.nc * j;
        p.f[MPAS_TW][a] = p.f[MPAS_TW][a] + p.f[MPAS_WCH][a] * (w[a + p.nc] - w[a - p.nc]) / (p.dzh[a] + p.dzh[a + p.nc]);
    }
}

// ---------------------------------------------------------------------------------
// D: vn_tendency_kernel! with edge_w (Gassmann/Tendencies.jl:4-67); over orography w is w - w_c,
// pi the deviation pi' and the pressure gradient that of pressure_gradient (:13-30)

// w~ = M_e / rho_e at interface j (edge_w, Tendencies.jl:4-11); with mass weighting (mw) the
// densities of the layers below (bl, br) and above (al, ar) the interface at the two cells
__device__ __forceinline__ double edge_w(const Params& p, const double* __restrict__ w, int32_t l, int32_t r, int j,
                                         double ew1, double ew2, bool mw, double bl, double al, double br, double ar) {
    const double rhl = mw ? p.wlow[l + p.nc * j] * bl + (1 - p.wlow[l + p.nc * j]) * al : 1.0;
    const double rhr = mw ? p.wlow[r + p.nc * j] * br + (1 - p.wlow[r + p.nc * j]) * ar : 1.0;
    const double M = ew1 * rhl * w[l + p.nc * j] + ew2 * rhr * w[r + p.nc * j];
    return M / (ew1 * rhl + ew2 * rhr);
}

// The PV flux sum of a ten-entry row in slot pattern PT (trisk::mpas_pv_slot): the products
// and their order are those of the row as stored, each PV read from the gathered pairs.
template <int PT>
__device__ __forceinline__ double pattern_sum(const double (&Tw)[10], const double2 (&P)[10], double qs) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        const double qx = P[trisk::mpas_pv_slot(PT, j)].y;
        s += Tw[j] * P[j].x * (((trisk::mpas_pv_mean(PT) >> j) & 1 ? qs : qx) + qx);
    }
    return s;
}

// VERT, PRESS: p.vertical and pressure > 0 fixed at compile time (-1: at run time)
template <int KT, bool PAIRED, int VERT = -1, int PRESS = -1>
__global__ __launch_bounds__(kBlock) void mpas_vn_tendency_kernel(Params p, const double* __restrict__ vn,
                                                                  const double* __restrict__ w,
                                                                  const double* __restrict__ rho, double pressure,
                                                                  int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    const int k0 = int(blockIdx.y) * kLevelsVn, k1 = min(k0 + kLevelsVn, p.nlev);
    const int kt = KT > 0 ? KT : p.kt;
    constexpr int N = KT > 0 ? KT : 32;
    int32_t I0[N], I1[N], I2[N];
    double Tw[N];
    for (int j = 0; j < kt; ++j) {
        Tw[j] = ld(p.trw + j * p.ne + e);
        if constexpr (PAIRED) {
            I0[j] = ld(p.ti2 + j * p.ne + e);
            I1[j] = ld(p.ti2 + (kt + j) * p.ne + e);
        } else {
            I0[j] = ld(p.ti + (0 * kt + j) * p.ne + e);
            I1[j] = ld(p.ti + (1 * kt + j) * p.ne + e);
            I2[j] = ld(p.ti + (2 * kt + j) * p.ne + e);
        }
    }
    const int32_t l = ld(p.ec + e), r = ld(p.ec + p.ne + e);
    const double ew1 = ld(p.ew + e), ew2 = ld(p.ew + p.ne + e), invd = ld(p.invd + e);
    const double2* FQ = reinterpret_cast<const double2*>(p.fq);
    const int pat = PAIRED && KT == 10 && p.tpat ? ld(p.tpat + e) : -1;
    const double2* TP = reinterpret_cast<const double2*>(p.tp);  // (theta, pi)
    const double* pip = p.f[MPAS_PIP];                           // pi' over orography
    const double* K = p.f[MPAS_K];
    const double2* RE = reinterpret_cast<const double2*>(p.re);  // (rho_e, eta): eta at .y
    const int vert = VERT >= 0 ? VERT : p.vertical;
    const bool press = PRESS >= 0 ? PRESS == 1 : pressure > 0;
    const bool mw = vert == 2;
    // w~ at the lower interface of the level; the upper one is the next level's lower
    // the layer densities at the two cells below the next interface, carried from level to level
    double Rl = 0.0, Rr = 0.0;
    if (mw && vert > 0 && k0 > 0) {
        Rl = rho[l + p.nc * (k0 - 1)];
        Rr = rho[r + p.nc * (k0 - 1)];
    }
    double Wnext = 0.0;
    if (vert > 0 && k0 > 0) {
        const double al = mw ? rho[l + p.nc * k0] : 0.0, ar = mw ? rho[r + p.nc * k0] : 0.0;
        Wnext = edge_w(p, w, l, r, k0, ew1, ew2, mw, Rl, al, Rr, ar);
        Rl = al;
        Rr = ar;
    } else if (mw && vert > 0) {
        Rl = rho[l];
        Rr = rho[r];
    }
    // eta at the lower interface of the level, carried like W~ (the upper one is the next level's lower),
    // and so is the layer thickness factor edzh of the interface
    double Enext = vert > 0 && k0 < p.nlev ? RE[e + p.ne * k0].y : 0.0;
    double Dnext = vert > 0 && k0 < p.nlev ? ld(p.edzh + e + p.ne * k0) : 0.0;
#pragma unroll
    for (int q = 0; q < kLevelsVn; ++q) {
        const int k = k0 + q;
        if (k >= k1) break;
        double s = 0.0;
        if constexpr (PAIRED && KT > 0) {
            // (q_b + q_c) with b = e: (q_e + q_x); with b = c: (q_x + q_x).  Every gather of
            // the level is issued before the sum, so they are in flight together.  In a slot
            // pattern every x is the flux edge of a slot of the row, so the (F, q) pairs of
            // the ten flux edges hold every PV: ten 16-byte gathers instead of twenty.
            double2 P[KT];
#pragma unroll
            for (int j = 0; j < KT; ++j) P[j] = FQ[I0[j] + p.ne * k];
            const double qs = FQ[e + p.ne * k].y;
            if (pat >= 0) {
                if (pat == 0)
                    s = pattern_sum<0>(Tw, P, qs);
                else if (pat == 1)
                    s = pattern_sum<1>(Tw, P, qs);
                else
                    s = pattern_sum<2>(Tw, P, qs);
            } else {
                double Qv[KT];
#pragma unroll
                for (int j = 0; j < KT; ++j) Qv[j] = FQ[(I1[j] < 0 ? ~I1[j] : I1[j]) + p.ne * k].y;
#pragma unroll
                for (int j = 0; j < KT; ++j) s += Tw[j] * P[j].x * ((I1[j] < 0 ? Qv[j] : qs) + Qv[j]);
            }
        } else if constexpr (PAIRED) {
            const double qs = FQ[e + p.ne * k].y;
            for (int j = 0; j < kt; ++j) {
                const bool twice = I1[j] < 0;
                const double qx = FQ[(twice ? ~I1[j] : I1[j]) + p.ne * k].y;
                s += Tw[j] * FQ[I0[j] + p.ne * k].x * ((twice ? qx : qs) + qx);
            }
        } else {
            for (int j = 0; j < kt; ++j)
                s += Tw[j] * FQ[I0[j] + p.ne * k].x * (FQ[I1[j] + p.ne * k].y + FQ[I2[j] + p.ne * k].y);
        }
        double tt = s - invd * (K[r + p.nc * k] - K[l + p.nc * k]);
        if (press) {
            const double2 Tl = TP[l + p.nc * k], Tr = TP[r + p.nc * k];
            const double te = ew1 * Tl.x + ew2 * Tr.x;
            double dpi;
            if (p.oro == 2) {
                // height_difference (CGridOperators.jl:193-200): pi' of both cells at the edge height
                const int64_t z = 2 * (e + p.ne * k);
                const int jl = ld(p.zlev + z), jr = ld(p.zlev + z + 1);
                const double tl = ld(p.zwt + z), tr = ld(p.zwt + z + 1);
                const double pr = pip[r + p.nc * jr], pl = pip[l + p.nc * jl];
                dpi = (pr + tr * (pip[r + p.nc * (jr + 1)] - pr) - pl - tl * (pip[l + p.nc * (jl + 1)] - pl)) * invd;
            } else {
                dpi = (p.oro ? pip[r + p.nc * k] - pip[l + p.nc * k] : Tr.y - Tl.y) * invd;
                if (p.oro == 1)
                    dpi -= p.ddxn[e + p.ne * k] *
                           (ew1 * p.f[MPAS_DPIDZ][l + p.nc * k] + ew2 * p.f[MPAS_DPIDZ][r + p.nc * k]);
            }
            tt -= pressure * (Cpd * te * dpi);
        }
        double out;
        if (vert > 0) {
            const double Wl = Wnext;
            double Wu = 0.0;
            if (k < p.nlev - 1) {
                const double al = mw ? rho[l + p.nc * (k + 1)] : 0.0, ar = mw ? rho[r + p.nc * (k + 1)] : 0.0;
                Wu = edge_w(p, w, l, r, k + 1, ew1, ew2, mw, Rl, al, Rr, ar);
                Rl = al;
                Rr = ar;
            }
            Wnext = Wu;
            const double El = Enext, Eu = RE[e + p.ne * (k + 1)].y;
            Enext = Eu;
            // (edz of the layer, edzh of its upper interface) in one 16-byte load
            double dz2[2];
            {
                const trisk::f32x4 v = __builtin_nontemporal_load((trisk::gptr<trisk::f32x4>)(p.edzp + 2 * (e + p.ne * k)));
                __builtin_memcpy(dz2, &v, 16);
            }
            const double Dl = Dnext, Du = dz2[1];
            Dnext = Du;
            out = tt - (Dl * El * Wl + Du * Eu * Wu) / (2 * dz2[0]);
        } else {
            const double we = (w[l + p.nc * k] + w[r + p.nc * k] + w[l + p.nc * (k + 1)] + w[r + p.nc * (k + 1)]) / 4;
            const double vlow = k > 0 ? (vn[e + p.ne * (k - 1)] + vn[e + p.ne * k]) / 2 : vn[e + p.ne * k];
            const double vup = k < p.nlev - 1 ? (vn[e + p.ne * k] + vn[e + p.ne * (k + 1)]) / 2 : vn[e + p.ne * k];
            out = tt - we * (vup - vlow) / ld(p.edz + e + p.ne * k);
        }
        p.f[MPAS_TVN][e + p.ne * k] = out;
    }
}

// a pair plane of n pairs split into its two fields (mpas_ops_field); the first field has
// only the first n1 entries
__global__ __launch_bounds__(kBlock) void mpas_split_pairs_kernel(const double* __restrict__ pairs,
                                                                  double* __restrict__ a, double* __restrict__ b,
                                                                  int64_t n, int64_t n1) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const double2 v = reinterpret_cast<const double2*>(pairs)[i];
    if (i < n1) a[i] = v.x;
    b[i] = v.y;
}

// ---------------------------------------------------------------------------------

enum { K_A = 0, K_A2, K_B, K_C, K_E, K_D, K_CC, K_CI, K_EX, K_WC, K_N };
static const char* kNames[K_N] = {"mpas_cell_vertex_kernel", "mpas_cell_from_vertex_kernel", "mpas_edge_kernel",
                                  "mpas_w_tendency_kernel",  "mpas_w_kinetic_kernel",        "mpas_vn_tendency_kernel",
                                  "mpas_concorr_cell_kernel", "mpas_concorr_interface_kernel",
                                  "mpas_exner_deviation_kernel", "mpas_w_concorr_kernel"};

}  // namespace mpas

struct mpas_ops {
    mpas::Params p{};
    trisk::MpasLayout L;
    int device = 0;
    int five_band = 0;
    std::vector<trisk::DevBuf> owned;
    int64_t field_points[MPAS_FIELDS] = {};
    int field_levels[MPAS_FIELDS] = {};
    bool timing = false;
    std::vector<hipEvent_t> events;  // pairs (start, stop)
    std::vector<int> event_kernel;
    size_t used = 0;
};

namespace mpas {

static int timed_begin(mpas_ops* H, int kernel, hipStream_t s, hipEvent_t* stop) {
    *stop = nullptr;
    if (!H->timing) return 0;
    if (H->used + 2 > H->events.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            H->events.push_back(e);
        }
        H->event_kernel.resize(H->events.size() / 2);
    }
    H->event_kernel[H->used / 2] = kernel;
    HIP_TRY(hipEventRecord(H->events[H->used], s));
    *stop = H->events[H->used + 1];
    H->used += 2;
    return 0;
}

static int timed_end(hipEvent_t stop, hipStream_t s) {
    if (stop) HIP_TRY(hipEventRecord(stop, s));
    return 0;
}

static const std::vector<int64_t>* perm_of(const mpas_ops* h, int kind) {
    return kind == 0 ? &h->L.cperm : kind == 1 ? &h->L.eperm : kind == 2 ? &h->L.vperm : nullptr;
}

static int slow_tendencies(mpas_ops* H, const double* vn, const double* w, const double* rho, const double* rt,
                           double pressure, hipStream_t s) {
    Params& p = H->p;
    const int64_t tc = tiles_per_xcd(p.nc), tv = tiles_per_xcd(p.nv), te = tiles_per_xcd(p.ne);
    const unsigned chunks = unsigned((p.nlev + kLevels - 1) / kLevels);
    const bool hex = p.kc == 6 && p.kv == 3 && p.kt == 10;
    // K_w: with the five-band solver after the w tendency (E), else before it (A or A2)
    const int kw_before = p.vertical > 0 && !H->five_band;
    hipEvent_t stop;
    int rc;
    if ((rc = timed_begin(H, K_A, s, &stop))) return rc;
    const dim3 ga(unsigned(kXcd * (tc + tv)), chunks);
    if (hex)
        hipLaunchKernelGGL((mpas_cell_vertex_kernel<6, 3>), ga, dim3(kBlock), 0, s, p, vn, w, rho, rt, tc, tv,
                           p.vk ? 0 : kw_before);
    else
        hipLaunchKernelGGL((mpas_cell_vertex_kernel<0, 0>), ga, dim3(kBlock), 0, s, p, vn, w, rho, rt, tc, tv,
                           p.vk ? 0 : kw_before);
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;
    const dim3 gc(unsigned(kXcd * tc), chunks);
    if (p.oro) {
        // the contravariant correction of vn, w - w_c, and pi' with its vertical derivative
        if ((rc = timed_begin(H, K_CC, s, &stop))) return rc;
        if (p.kc == 6)
            hipLaunchKernelGGL((mpas_concorr_cell_kernel<6>), gc, dim3(kBlock), 0, s, p, vn, tc);
        else
            hipLaunchKernelGGL((mpas_concorr_cell_kernel<0>), gc, dim3(kBlock), 0, s, p, vn, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
        if ((rc = timed_begin(H, K_CI, s, &stop))) return rc;
        hipLaunchKernelGGL(mpas_concorr_interface_kernel, gc, dim3(kBlock), 0, s, p, w, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
        if (pressure > 0) {
            if ((rc = timed_begin(H, K_EX, s, &stop))) return rc;
            hipLaunchKernelGGL(mpas_exner_deviation_kernel, gc, dim3(kBlock), 0, s, p, tc);
            HIP_TRY(hipGetLastError());
            if ((rc = timed_end(stop, s))) return rc;
        }
    }
    if (p.vk) {
        if ((rc = timed_begin(H, K_A2, s, &stop))) return rc;
        hipLaunchKernelGGL(mpas_cell_from_vertex_kernel, dim3(unsigned(kXcd * tc), chunks), dim3(kBlock), 0, s, p, w,
                           tc, kw_before);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    if ((rc = timed_begin(H, K_B, s, &stop))) return rc;
    hipLaunchKernelGGL(mpas_edge_kernel, dim3(unsigned(kXcd * te), chunks), dim3(kBlock), 0, s, p, vn, w, rho, te);
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;
    if ((rc = timed_begin(H, K_C, s, &stop))) return rc;
    if (p.kc == 6)
        hipLaunchKernelGGL((mpas_w_tendency_kernel<6>), dim3(unsigned(kXcd * tc), chunks), dim3(kBlock), 0, s, p, vn,
                           w, tc);
    else
        hipLaunchKernelGGL((mpas_w_tendency_kernel<0>), dim3(unsigned(kXcd * tc), chunks), dim3(kBlock), 0, s, p, vn,
                           w, tc);
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;
    if (p.vertical > 0 && H->five_band) {
        if ((rc = timed_begin(H, K_E, s, &stop))) return rc;
        hipLaunchKernelGGL(mpas_w_kinetic_kernel, dim3(unsigned(kXcd * tc), chunks), dim3(kBlock), 0, s, p, w, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    if (p.oro) {
        if ((rc = timed_begin(H, K_WC, s, &stop))) return rc;
        hipLaunchKernelGGL(mpas_w_concorr_kernel, gc, dim3(kBlock), 0, s, p, w, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
        w = p.f[MPAS_WHAT];  // the vn tendency couples vertically with w - w_c
    }
    if ((rc = timed_begin(H, K_D, s, &stop))) return rc;
    const unsigned vchunks = unsigned((p.nlev + kLevelsVn - 1) / kLevelsVn);
    const dim3 gd(unsigned(kXcd * te), vchunks);
    const bool press = pressure > 0;
    const dim3 b(kBlock);
    if (p.kt == 10 && p.ti2 && p.vertical == 0 && press)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 0, 1>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10 && p.ti2 && p.vertical == 0)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 0, 0>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10 && p.ti2 && p.vertical == 1 && press)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 1, 1>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10 && p.ti2 && p.vertical == 1)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 1, 0>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10 && p.ti2 && press)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 2, 1>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10 && p.ti2)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, true, 2, 0>), gd, b, 0, s, p, vn, w, rho, pressure, te);
    else if (p.kt == 10)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<10, false>), gd, dim3(kBlock), 0, s, p, vn, w, rho, pressure, te);
    else if (p.ti2)
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<0, true>), gd, dim3(kBlock), 0, s, p, vn, w, rho, pressure, te);
    else
        hipLaunchKernelGGL((mpas_vn_tendency_kernel<0, false>), gd, dim3(kBlock), 0, s, p, vn, w, rho, pressure, te);
    HIP_TRY(hipGetLastError());
    return timed_end(stop, s);
}

}  // namespace mpas

using namespace mpas;

extern "C" {

int mpas_ops_check(const mpas_ops_desc* desc) {
    trisk::MpasLayout L;
    std::string err;
    if (!trisk::build_mpas_layout(desc, L, err)) return set_err(TRISK_EINVAL, err);
    return 0;
}

int mpas_ops_create(const mpas_ops_desc* desc, int32_t device, mpas_ops** out) {
    if (!out) return set_err(TRISK_EINVAL, "null output pointer");
    *out = nullptr;
    auto* H = new mpas_ops;
    std::string err;
    if (!trisk::build_mpas_layout(desc, H->L, err)) {
        delete H;
        return set_err(TRISK_EINVAL, err);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) {
        delete H;
        return set_err(TRISK_EHIP, "no such HIP device");
    }
    H->device = device;
    H->five_band = desc->five_band != 0;
    const trisk::MpasLayout& L = H->L;
    Params& p = H->p;
    p.nc = L.nc;
    p.ne = L.ne;
    p.nv = L.nv;
    p.nlev = L.nlev;
    p.kc = L.kc;
    p.kv = L.kv;
    p.kt = L.kt;
    p.iap_pv = L.iap_pv;
    p.vertical = L.vertical;
    p.vk = L.vk;
    auto& o = H->owned;
    int rc = 0;
    rc |= trisk::upload(o, L.ce, &p.ce);
    rc |= trisk::upload(o, L.kin, &p.kin);
    rc |= trisk::upload(o, L.inner, &p.inner);
    rc |= trisk::upload(o, L.dz, &p.dz);
    rc |= trisk::upload(o, L.dzh, &p.dzh);
    rc |= trisk::upload(o, L.wlow, &p.wlow);
    rc |= trisk::upload(o, L.ec, &p.ec);
    rc |= trisk::upload(o, L.ev, &p.ev);
    rc |= trisk::upload(o, L.ti, &p.ti);
    if (L.paired) rc |= trisk::upload(o, L.ti2, &p.ti2);
    if (L.pv_slots) rc |= trisk::upload(o, L.tpat, &p.tpat);
    rc |= trisk::upload(o, L.ew, &p.ew);
    rc |= trisk::upload(o, L.invd, &p.invd);
    rc |= trisk::upload(o, L.trw, &p.trw);
    rc |= trisk::upload(o, L.rh, &p.rh);
    rc |= trisk::upload(o, L.fe, &p.fe);
    rc |= trisk::upload(o, L.edz, &p.edz);
    rc |= trisk::upload(o, L.edzh, &p.edzh);
    {
        std::vector<double> pz(size_t(2) * L.ne * L.nlev);
        for (int k = 0; k < L.nlev; ++k)
            for (int64_t e = 0; e < L.ne; ++e) {
                pz[2 * (e + L.ne * k)] = L.edz[e + L.ne * k];
                pz[2 * (e + L.ne * k) + 1] = L.edzh[e + L.ne * (k + 1)];
            }
        rc |= trisk::upload(o, pz, &p.edzp);
    }
    rc |= trisk::upload(o, L.ve, &p.ve);
    rc |= trisk::upload(o, L.vc, &p.vc);
    rc |= trisk::upload(o, L.curl, &p.curl);
    rc |= trisk::upload(o, L.vw, &p.vw);
    rc |= trisk::upload(o, L.fv, &p.fv);
    if (L.vk) {
        rc |= trisk::upload(o, L.cv, &p.cv);
        rc |= trisk::upload(o, L.vkk, &p.vkk);
        rc |= trisk::upload(o, L.vkw, &p.vkw);
    }
    p.oro = L.oro;
    if (L.oro) {
        rc |= trisk::upload(o, L.zf, &p.zf);
        rc |= trisk::upload(o, L.piref, &p.piref);
        rc |= trisk::upload(o, L.gq, &p.gq);
        rc |= trisk::upload(o, L.ddxn, &p.ddxn);
    }
    if (L.oro == 2) {
        rc |= trisk::upload(o, L.zlev, &p.zlev);
        rc |= trisk::upload(o, L.zwt, &p.zwt);
    }
    // work fields: points x levels
    const int64_t pts[MPAS_FIELDS] = {L.nc, L.nc, L.nc, L.nv, L.nv, L.nv, L.ne, L.ne, L.ne,
                                      L.ne, L.nc, L.ne, L.nv, L.nc, L.nc, L.nc, L.nc, L.nc};
    const int lev[MPAS_FIELDS] = {L.nlev, L.nlev,     L.nlev,     L.nlev, L.nlev, L.nlev,
                                  L.nlev, L.nlev,     L.nlev,     L.nlev + 1, L.nlev + 1, L.nlev,
                                  L.nlev, L.nlev,     L.nlev + 1, L.nlev + 1, L.nlev, L.nlev};
    for (int f = 0; f < MPAS_FIELDS && !rc; ++f) {
        H->field_points[f] = pts[f];
        H->field_levels[f] = lev[f];
        if (f == MPAS_KV && !L.vk) continue;
        if (f >= MPAS_WCF && !L.oro) continue;  // the orography fields
        void* q = nullptr;
        rc = trisk::alloc(o, size_t(pts[f]) * size_t(lev[f]) * sizeof(double), &q);
        p.f[f] = static_cast<double*>(q);
    }
    if (!rc) {
        void* q = nullptr;
        rc = trisk::alloc(o, size_t(2) * size_t(L.ne) * size_t(L.nlev) * sizeof(double), &q);
        p.fq = static_cast<double*>(q);
    }
    if (!rc) {
        void* q = nullptr;
        rc = trisk::alloc(o, size_t(2) * size_t(L.nv) * size_t(L.nlev) * sizeof(double), &q);
        p.zr = static_cast<double*>(q);
    }
    if (!rc) {
        void* q = nullptr;
        rc = trisk::alloc(o, size_t(2) * size_t(L.nc) * size_t(L.nlev) * sizeof(double), &q);
        p.tp = static_cast<double*>(q);
    }
    if (!rc) {
        void* q = nullptr;
        rc = trisk::alloc(o, size_t(2) * size_t(L.ne) * size_t(L.nlev + 1) * sizeof(double), &q);
        p.re = static_cast<double*>(q);
    }
    if (rc) {
        for (auto& b : o) (void)hipFree(b.p);
        delete H;
        return rc;
    }
    *out = H;
    return 0;
}

int mpas_ops_destroy(mpas_ops* h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    for (auto& e : h->events) (void)hipEventDestroy(e);
    for (auto& b : h->owned) (void)hipFree(b.p);
    delete h;
    return 0;
}

int mpas_ops_permutation(const mpas_ops* h, int64_t* cells, int64_t* edges, int64_t* vertices) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    if (cells) std::memcpy(cells, h->L.cperm.data(), h->L.nc * sizeof(int64_t));
    if (edges) std::memcpy(edges, h->L.eperm.data(), h->L.ne * sizeof(int64_t));
    if (vertices) std::memcpy(vertices, h->L.vperm.data(), h->L.nv * sizeof(int64_t));
    return 0;
}

int mpas_ops_upload_field(mpas_ops* h, int32_t kind, int64_t levels, const double* host, double* dev, void* stream) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    const std::vector<int64_t>* perm = perm_of(h, kind);
    if (!perm || !host || !dev || levels <= 0) return set_err(TRISK_EINVAL, "bad field kind, levels or pointer");
    HIP_TRY(hipSetDevice(h->device));
    const int64_t n = int64_t(perm->size());
    std::vector<double> tmp(size_t(n * levels));
    for (int64_t k = 0; k < levels; ++k)
        for (int64_t q = 0; q < n; ++q) tmp[size_t(q + n * k)] = host[(*perm)[q] + n * k];
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int mpas_ops_download_field(mpas_ops* h, int32_t kind, int64_t levels, const double* dev, double* host,
                            void* stream) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    const std::vector<int64_t>* perm = perm_of(h, kind);
    if (!perm || !host || !dev || levels <= 0) return set_err(TRISK_EINVAL, "bad field kind, levels or pointer");
    HIP_TRY(hipSetDevice(h->device));
    const int64_t n = int64_t(perm->size());
    std::vector<double> tmp(size_t(n * levels));
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int64_t k = 0; k < levels; ++k)
        for (int64_t q = 0; q < n; ++q) host[(*perm)[q] + n * k] = tmp[size_t(q + n * k)];
    return 0;
}

int mpas_slow_tendencies(mpas_ops* h, const double* vn, const double* w, const double* rho, const double* rhotheta,
                         double pressure, void* stream) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    if (!vn || !w || !rho || !rhotheta) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(h->device));
    return slow_tendencies(h, vn, w, rho, rhotheta, pressure, static_cast<hipStream_t>(stream));
}

int mpas_ops_field(mpas_ops* h, int32_t which, double** dev, int64_t* points, int64_t* levels) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    if (which < 0 || which >= MPAS_FIELDS || !h->p.f[which]) return set_err(TRISK_EINVAL, "no such work field");
    // fields that live interleaved in a pair plane: split it after the work in flight
    const double* pairs = nullptr;
    double *first = nullptr, *second = nullptr;
    int64_t n = 0, n1 = 0;
    if (which == MPAS_FS || which == MPAS_QE) {
        pairs = h->p.fq, first = h->p.f[MPAS_FS], second = h->p.f[MPAS_QE];
        n = n1 = h->L.ne * int64_t(h->L.nlev);
    } else if (which == MPAS_ZETA || which == MPAS_RHOV) {
        pairs = h->p.zr, first = h->p.f[MPAS_ZETA], second = h->p.f[MPAS_RHOV];
        n = n1 = h->L.nv * int64_t(h->L.nlev);
    } else if (which == MPAS_THETA || which == MPAS_EXNER) {
        pairs = h->p.tp, first = h->p.f[MPAS_THETA], second = h->p.f[MPAS_EXNER];
        n = n1 = h->L.nc * int64_t(h->L.nlev);
    } else if (which == MPAS_RHOE || which == MPAS_ETA) {
        pairs = h->p.re, first = h->p.f[MPAS_RHOE], second = h->p.f[MPAS_ETA];
        n1 = h->L.ne * int64_t(h->L.nlev);  // rho_e has no top row
        n = n1 + h->L.ne;
    }
    if (pairs) {
        HIP_TRY(hipSetDevice(h->device));
        HIP_TRY(hipDeviceSynchronize());
        hipLaunchKernelGGL(mpas_split_pairs_kernel, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, nullptr,
                           pairs, first, second, n, n1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    if (dev) *dev = h->p.f[which];
    if (points) *points = h->field_points[which];
    if (levels) *levels = h->field_levels[which];
    return 0;
}

int mpas_ops_set_timing(mpas_ops* h, int32_t on) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    h->timing = on != 0;
    h->used = 0;
    return 0;
}

int mpas_ops_kernel_times(mpas_ops* h, int32_t max, const char** names, double* avg_ms, int64_t* launches,
                          int32_t* count) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(h->device));
    double ms[K_N] = {0};
    int64_t n[K_N] = {0};
    for (size_t i = 0; i + 1 < h->used; i += 2) {
        HIP_TRY(hipEventSynchronize(h->events[i + 1]));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, h->events[i], h->events[i + 1]));
        ms[h->event_kernel[i / 2]] += t;
        n[h->event_kernel[i / 2]] += 1;
    }
    int c = 0;
    for (int k = 0; k < K_N && c < max; ++k) {
        if (!n[k]) continue;
        if (names) names[c] = kNames[k];
        if (avg_ms) avg_ms[c] = ms[k] / double(n[k]);
        if (launches) launches[c] = n[k];
        ++c;
    }
    if (count) *count = c;
    return 0;
}

// Algorithmic HBM bytes of one launch: every element of every array a launch must touch
// once (gathered fields at their full size), tables once per point.
int mpas_ops_kernel_bytes(mpas_ops* h, int32_t max, const char** names, double* bytes, int32_t* count) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    const Params& p = h->p;
    const double nc = double(p.nc), ne = double(p.ne), nv = double(p.nv), L = p.nlev, L1 = p.nlev + 1;
    const double KC = p.kc, KV = p.kv, KT = p.kt, ix = 4.0, c = 8.0;
    const bool vi = p.vertical > 0, kwb = vi && !h->five_band;
    double v[K_N] = {0};
    // A: vn, rho, rhotheta (and w with K_w) read; K, theta, pi, zeta, rho_v, q_v [K_v] written; tables
    v[K_A] = ne * L * c + 2 * nc * L * c + (kwb && !p.vk ? nc * (L1 * c + L * c + L1 * c) : 0.0) + 3 * nc * L * c +
             (3 + (p.vk ? 1 : 0)) * nv * L * c + nc * KC * (ix + c) + nv * (KV * 2 * (ix + c) + c + (p.vk ? KV * c : 0));
    v[K_A2] = p.vk ? nc * KC * (ix + c) + nv * L * c + 2 * nc * L * c + (kwb ? nc * (2 * L1 + L) * c : 0.0) : 0.0;
    // B: zeta, rho_v (or q_v), rho, vn [, w] read; q_e, rho_e, F [, eta] written; tables
    // (the pair planes: (F, q_e) per layer, (rho_e, eta) per interface)
    v[K_B] = nv * L * c * (p.iap_pv ? 2 : 1) + nc * L * c + ne * L * c + 2 * ne * L * c + 2 * ne * L1 * c +
             (vi ? nc * L1 * c + ne * L1 * c : 0.0) + ne * (4 * ix + 7 * c);
    // C: eta, vn [rho_e], K, dzh, wlow -> tw (vector invariant); vn, w, dz -> tw (advective)
    v[K_C] = vi ? ne * L1 * c + ne * L * c * (p.vertical == 2 ? 2 : 1) + nc * L * c + 2 * nc * L1 * c + nc * L1 * c +
                      nc * KC * (ix + c)
                : ne * L * c + nc * L1 * c + nc * L * c + nc * L1 * c + nc * KC * (ix + c) + ne * (2 * ix + c);
    v[K_E] = vi && h->five_band ? nc * (L1 * c * 2 + L * c * 3) : 0.0;
    // D: F, q_e, theta, pi, K, [eta, w, rho, wlow, edge dz, dzh | w, vn, edge dz] -> tvn; tables
    // (the values the kernel needs: theta, pi or pi', K; eta, not the rho_e beside it in its pairs)
    v[K_D] = ne * L * c * 2 + 3 * nc * L * c + ne * L * c + ne * (KT * ((p.ti2 ? 2 : 3) * ix + c) + 2 * ix + 3 * c) +
             (vi ? ne * L1 * c * 2 + nc * L1 * c * 2 + nc * L * c + ne * L * c : nc * L1 * c + 2 * ne * L * c) +
             (p.oro == 1 ? ne * L * c + nc * L * c : p.oro == 2 ? 2 * ne * L * (ix + c) : 0.0);
    if (p.oro) {
        // CC: vn, ddxn, the cell rows -> w_c; CI: w_c, wlow, q, w -> w_c and w - w_c at the
        // interfaces; EX: pi, pi_ref, z_full -> pi', dpi'/dz; WC: tw (read, written), w, w_c, dzh
        v[K_CC] = 2 * ne * L * c + nc * KC * (ix + c) + nc * L * c;
        v[K_CI] = nc * L * c + nc * L1 * c * 2 + nc * 3 * c + 2 * nc * L1 * c;
        v[K_EX] = 5 * nc * L * c;  // pi, pi_ref, z_full -> pi', dpi'/dz
        v[K_WC] = 5 * nc * L1 * c;
    }
    int k2 = 0;
    for (int k = 0; k < K_N && k2 < max; ++k) {
        if (v[k] == 0.0) continue;
        if (names) names[k2] = kNames[k];
        bytes[k2] = v[k];
        ++k2;
    }
    if (count) *count = k2;
    return 0;
}

}  // extern "C"
