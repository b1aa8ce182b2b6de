// Adjoint of the TRiSK shallow-water tendency and of the RK4 step (trisk_tendency_vjp,
// trisk_rk4_adjoint in include/trisk_hip.h), on the cell-pair path (below) and on the table
// path (every PV flux and kinetic-energy form; "Table path" further down).
//
// The reference has no adjoint; these kernels transpose the linearisation of the forward
// cell-pair tendency (trisk_device.hip) of tendency! (TRiSKShallowWater.jl:379-394, TRiSK
// PV flux, cell kinetic energy) and of rk4_step! (RungeKutta.jl:18-28).  With cotangents
// (lu, lh) of (du, dh) and the forward pieces
//   B_i = sum_k kin_k u_k^2 + g h_i          q_v = (sum curl u + f_v) / (sum vh h)
//   G_e = (lh_0 h_c0 + lh_1 h_c1) u_e        H_c(k) = sum_k' w_kk' G_k' (qe_k + qe_k')
//   du_e = (H_l(e) + H_r(e)) / (2 d_e) - (B_r - B_l) / d_e,   dh_i = -(sum_k s_k G_k) / A_i
// one vector-Jacobian product is five gather-only launches, each entity summing the
// contributions of its neighbours (no atomics, deterministic):
//   W1   forward B, q_v, G at the stage (pair_state_kernel)
//   A1   cells: Bbar_i = sum_k s_k lu_k / d_k; per cell the cotangents of the G and of the
//        corner PV of its edges (the transpose of the half sums) -> per (edge, side) and
//        per (cell, corner) partials
//   A2   vertices: qbar_v from the corner partials of its three cells; zbar = qbar / h_v,
//        hbar_v = -qbar q_v / h_v
//   A3   edges: Gbar_e from its two sides; ubar_e = Gbar_e h_e + sum curl zbar + 2 u_e sum kin Bbar;
//        and Gbar_e lh_s u_e for each side s
//   A4   cells: hbar_i = g Bbar_i + sum_k Gbar_k lh_k u_k + sum_k vh_k hbar_vk, from the products
//        A3 and A2 formed
// W1 runs only where the sweep has not kept q_v and G of the stage (see enqueue_sweep).
// A3 and A4 end in the RK4 adjoint epilogue, which combines the stage cotangents as the
// transpose of y + dt/6 (k1 + 2 k2 + 2 k3 + k4).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "trisk_device.hpp"

namespace trisk {

// the adjoint's tables and work vectors (device)
template <class P>
struct AdjParams {
    using C = typename P::C;
    using I = typename P::I;
    const C* ecurl;        // ne x 2
    const C* ekin;         // ne x 2
    const int32_t* cv3;    // 6 x nc: 3 v + the slot of cell i among the cells of its corner v
    const int32_t* vslot;  // nv
    double* Bbar;          // nc
    I* Tbar;               // 6 x nc, cell-major: cell i's partial of Gbar at its edge k in Tbar[6 i + k]
    I* QC;                 // nc x 6  (cell, corner) partials of qbar
    I* zbar;               // nv
    I* VH;                 // 3 x nv: vh_j hbar_v, the term of vertex v in the hbar of its cell j
    I* GL;                 // 2 x ne: Gbar_e lh_s u_e, the term of edge e in the hbar of its cell s
};

// Epilogue of A3/A4: the cotangent s of a stage state, combined as the RK4 adjoint.
//   mode 0: out = s (one vector-Jacobian product)
//   mode 4: stage 4 (cotangent dt/6 lam): acc = lam + s, kbar = dt/3 lam + dt s
//   mode 3: stage 3 (cotangent kbar):     acc += s,      kbar = dt/3 lam + dt/2 s
//   mode 2: stage 2 (cotangent kbar):     acc += s,      kbar = dt/6 lam + dt/2 s
//   mode 1: stage 1 (cotangent kbar):     lam = acc + s  (the cotangent of the step's input)
struct AdjRk {
    int mode;
    double* out;
    double* lam;
    double* acc;
    double* kbar;
    double dt;
};

// lam_j and acc_j are loaded first (adj_load), so that these streaming loads do not wait for
// the kernel's gathers; adj_store then applies the update above.
struct AdjIn {
    double lam, acc;
};

__device__ __forceinline__ AdjIn adj_load(const AdjRk& r, int64_t j) {
    AdjIn v{0.0, 0.0};
    if (r.mode >= 2) v.lam = r.lam[j];
    if (r.mode == 1 || r.mode == 2 || r.mode == 3) v.acc = r.acc[j];
    return v;
}

__device__ __forceinline__ void adj_store(const AdjRk& r, int64_t j, double s, const AdjIn& v) {
    switch (r.mode) {
        case 0:
            r.out[j] = s;
            break;
        case 4:
            r.acc[j] = v.lam + s;
            r.kbar[j] = (r.dt / 3) * v.lam + r.dt * s;
            break;
        case 3:
            r.acc[j] = v.acc + s;
            r.kbar[j] = (r.dt / 3) * v.lam + (r.dt / 2) * s;
            break;
        case 2:
            r.acc[j] = v.acc + s;
            r.kbar[j] = (r.dt / 6) * v.lam + (r.dt / 2) * s;
            break;
        default:
            r.lam[j] = v.acc + s;
            break;
    }
}

// A1: cells.  cot = [lu; lh] scaled by c.
template <class P>
__global__ __launch_bounds__(kBlock) void adj_cell_kernel(Params<P> p, AdjParams<P> w, const double* __restrict__ cot,
                                                          double c, int64_t t0) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t i = tile_of(blockIdx.x, t0, 0).first + threadIdx.x;  // XCD tiles
    if (i >= p.nc) return;
    int32_t meta, E[6], V[6];
    C W[15];
    load_cell_pairs<false>(p, i, meta, E, V, W);
    const int n = meta & 15;
    C G[6], Q[6], mu[6], Gb[6], qeb[6], Qb[6], qe[6];
    const double lh = c * cot[p.ne + i];
    const double area = ld(p.area_c + i);
    double bb = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        G[k] = C(p.G[E[k]]);
        Q[k] = C(p.qv[V[k]]);
        const C d = ld(p.dual + E[k]);
        const double s = (meta >> (4 + k)) & 1 ? -1.0 : 1.0;
        const double lu = k < n ? c * cot[E[k]] : 0.0;
        mu[k] = C(lu) / (2 * d);                   // cotangent of the half sum H_c(k)
        if (k < n) bb += s * (lu / double(d));     // of -(B_r - B_l) / d
        Gb[k] = k < n ? C(-s * lh / area) : C(0);  // of -(sum s G) / A
        qeb[k] = 0;
        Qb[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) qe[k] = (Q[k] + (k + 1 < n ? Q[k + 1 < 6 ? k + 1 : 0] : Q[0])) / 2;
    int q = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int k2 = k + 1; k2 < 6; ++k2, ++q) {
            if (k2 < n) {
                // forward: H_k += tq G_k2, H_k2 -= tq G_k with tq = w (qe_k + qe_k2)
                const C tq = W[q] * (qe[k] + qe[k2]);
                Gb[k2] += mu[k] * tq;
                Gb[k] -= mu[k2] * tq;
                const C t = W[q] * (mu[k] * G[k2] - mu[k2] * G[k]);
                qeb[k] += t;
                qeb[k2] += t;
            }
        }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) {
            const int nx = k + 1 < n ? k + 1 : 0;
            Qb[k] += qeb[k] / 2;
            Qb[nx] += qeb[k] / 2;
        }
    }
    // six contiguous values per cell (slots past n hold 0 and are never read)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        w.Tbar[6 * i + k] = I(Gb[k]);
        w.QC[6 * i + k] = I(Qb[k]);
    }
    w.Bbar[i] = bb;
}

// A2: vertices (the stage state recomputes zeta_v, h_v and q_v as the forward does); hbar_v
// leaves as its three products with the vertex-height weights, one per cell
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_vertex_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                            const double* __restrict__ h, int64_t t0) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t v = tile_of(blockIdx.x, t0, 0).first + threadIdx.x;  // XCD tiles
    if (v >= p.nv) return;
    const int32_t slots = ld(w.vslot + v);
    C qb = 0, z = 0, eta = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int64_t cj = ld(p.vc + j * p.nv + v);
        qb += C(w.QC[6 * cj + ((slots >> (4 * j)) & 15)]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) z += ld(p.curl + k * p.nv + v) * C(u[ld(p.ve + k * p.nv + v)]);
#pragma unroll
    for (int k = 0; k < 3; ++k) eta += ld(p.vh + k * p.nv + v) * C(h[ld(p.vc + k * p.nv + v)]);
    const C qv = (z + ld(p.fv + v)) / eta;
    w.zbar[v] = I(qb / eta);
    const C eb = C(I(-qb * qv / eta));
#pragma unroll
    for (int k = 0; k < 3; ++k) w.VH[3 * v + k] = I(ld(p.vh + k * p.nv + v) * eb);
}

// A3: edges
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_edge_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                          const double* __restrict__ h, AdjRk r, int64_t t0) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t e = tile_of(blockIdx.x, t0, 0).first + threadIdx.x;  // XCD tiles
    if (e >= p.ne) return;
    const AdjIn rin = adj_load(r, e);
    int32_t c[2];
    C lh[2], cu[2], kn[2];
    ldrec(p.re_cells + 2 * e, c);
    ldrec(p.re_lh + 2 * e, lh);
    ldrec(w.ecurl + 2 * e, cu);
    ldrec(w.ekin + 2 * e, kn);
    const int32_t v0 = ld(p.ev + e), v1 = ld(p.ev + p.ne + e);
    const uint32_t sl = ld(p.eslot + e);
    const C Gb = C(w.Tbar[6 * int64_t(c[0]) + (sl & 15)]) + C(w.Tbar[6 * int64_t(c[1]) + (sl >> 4)]);
    const C ue = C(u[e]);
    const C he = lh[0] * C(h[c[0]]) + lh[1] * C(h[c[1]]);
    const double s = double(Gb * he) + double(cu[0] * C(w.zbar[v0]) + cu[1] * C(w.zbar[v1])) +
                     2.0 * double(ue) * (double(kn[0]) * w.Bbar[c[0]] + double(kn[1]) * w.Bbar[c[1]]);
    const C gb = C(I(Gb));
    w.GL[2 * e] = I(gb * lh[0] * ue);  // the terms of A4's sum over the edges of each cell
    w.GL[2 * e + 1] = I(gb * lh[1] * ue);
    adj_store(r, e, s, rin);
}

// A4: cells
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_height_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                            AdjRk r, int64_t t0) {
    using C = typename P::C;
    const int64_t i = tile_of(blockIdx.x, t0, 0).first + threadIdx.x;  // XCD tiles
    if (i >= p.nc) return;
    const AdjIn rin = adj_load(r, p.ne + i);
    const int32_t meta = ld(p.cmeta + i);
    const int n = meta & 15;
    double s = p.g * w.Bbar[i];
    // the products were formed by A3 (Gbar lh u) and A2 (vh hbar_v) in the same order
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) {
            const int64_t e = ld(p.ce6 + k * p.nc + i);
            s += double(C(w.GL[2 * e + ((meta >> (4 + k)) & 1)]));
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) s += double(C(w.VH[ld(w.cv3 + k * p.nc + i)]));
    }
    adj_store(r, p.ne + i, s, rin);
}

// ---------------------------------------------------------------------------------
// Table path: the adjoint of the table kernels (trisk_device.hip P1-P3) for every PV flux
// (TRiSK, Gassmann, Gassmann mean; TRiSKShallowWater.jl:261-296) and kinetic-energy form
// (cell, mixed, vertex; :202-244).  With the forward pieces
//   B_i = sum_k kin u_k^2 [+ sum_k kites K_v] + g h_i,  K_v = sum_k vkw u_k^2
//   zeta_v = sum curl u,  h_v = sum vh h,  q_v = (zeta_v + f_v) / h_v
//   F_e = (eh_0 h_c0 + eh_1 h_c1 [+ vkh_0 h_v0 + vkh_1 h_v1]) u_e
//   q_e = (q_v0 + q_v1) / 2  or  (r_0 zeta_v0 + r_1 zeta_v1 + f_e) / (r_0 h_v0 + r_1 h_v1)
//   du_e = sum_k w_k F_fk q_k - (B_c1 - B_c0) / d_e,  dh_i = -sum_k div_k F_k
// (q_k the mean (q_e + q_fk)/2, or the PV of the slot's PV edge for the Gassmann flux) one
// vector-Jacobian product is P1 and P2 of the stage and four gather-only launches:
//   T1 edges: per slot the cotangents cF = lu w q_k of F_fk and cQ of the slot's PV
//   T2 edges: Fbar_f from the slots that hold f (the stencils are symmetric: f is in the
//      stencil of e exactly when e is in that of f; the host lists the (e, slot) pairs of
//      every f, dropping zero weights, which contribute nothing) and from the continuity;
//      qbar_f likewise, then the cotangents of the rhombus numerator and denominator
//      | cells: Bbar_i = sum_k s_k lu_k / d_k
//   T3 vertices: zbar_v, hbar_v (edge PV, mixed mass flux) and Kbar_v (vertex KE)
//   T4 edges: ubar_e | cells: hbar_i, each ending in the RK4 adjoint epilogue
// With the flux form of the FP32-bulk kernels F is G = l F and the continuity weights are
// s / A_i; the adjoint reads the same (scaled) weights as the forward.

template <class P>
struct TabAdj {
    using C = typename P::C;
    using I = typename P::I;
    int K = 0;                 // slots of the PV flux (kt or kg)
    int TW = 0, QW = 0;        // widths of the transposed lists tf, tq
    int CW = 0;                // corners per cell
    const int32_t* nb = nullptr;    // K x ne: the stencil (trisk_edges or g_edges)
    const C* w = nullptr;           // K x ne: its weights, as the forward reads them
    const int32_t* pv = nullptr;    // K x ne: the PV edge of each slot (Gassmann)
    const int32_t* tf_e = nullptr;  // TW x ne: the (edge, slot) pairs whose stencil holds f
    const uint8_t* tf_k = nullptr;  //          (slot 255: none)
    const int32_t* tq_e = nullptr;  // QW x ne: the (edge, slot) pairs whose PV edge is f (Gassmann)
    const uint8_t* tq_k = nullptr;
    const uint32_t* cbits = nullptr;  // nc: bit k: i is the right cell of its edge k; bit 16 + k: slot k is an edge
    const uint32_t* vbits = nullptr;  // nv: bit k: v is the second vertex of its edge k; bits 8 + 4k: slot of v
                                      //     among the cell_vertices of its cell k
    const C* ecurl = nullptr;  // 2 x ne: curl coefficient of e in its vertex s
    const C* ekin = nullptr;   // 2 x ne: kinetic-energy weight of e in its cell s
    const C* evkw = nullptr;   // 2 x ne: vertex kinetic-energy weight of e in its vertex s
    const C* edc = nullptr;    // 2 x ne: continuity weight of e in its cell s
    const int32_t* corner = nullptr;  // CW x nc: the vertices whose vertex_cells hold i (pads: 0)
    const C* cvh = nullptr;           // CW x nc: the vertex_height weight of i there (pads: 0)
    I *cF = nullptr, *cQ = nullptr;   // K x ne
    I *qown = nullptr, *Fb = nullptr, *Zb = nullptr, *Eb = nullptr;  // ne
    double* Bb = nullptr;                                            // nc
    I *zb = nullptr, *hvb = nullptr, *Kvb = nullptr;                 // nv
};

// T1: edges.  GM: the Gassmann form (each slot's PV from its PV edge); else the mean form.
template <class P, bool GM>
__global__ __launch_bounds__(kBlock) void tadj_slot_kernel(Params<P> p, TabAdj<P> a, const double* __restrict__ cot,
                                                           double c, int64_t te) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t e = tile_of(blockIdx.x, te, 0).first + threadIdx.x;  // XCD tiles
    if (e >= p.ne) return;
    const C lu = C(c * cot[e]);
    const C qe = C(p.Fq[e].q);
    C own = 0;
    for (int k = 0; k < a.K; ++k) {
        const int64_t j = k * p.ne + e;
        const Pair<I> f = p.Fq[ld(a.nb + j)];
        const C wl = lu * ld(a.w + j);
        C cf, cq;
        if constexpr (GM) {
            cf = wl * C(p.Fq[ld(a.pv + j)].q);
            cq = wl * C(f.F);
        } else {
            cf = wl * (qe + C(f.q)) / 2;
            cq = wl * C(f.F) / 2;
            own += cq;
        }
        a.cF[j] = I(cf);
        a.cQ[j] = I(cq);
    }
    a.qown[e] = I(own);
}

// T2: edges (Fbar, and the cotangents of the edge PV's parts) | cells (Bbar).  RH: the
// rhombus edge PV (Gassmann fluxes), else the mean of the vertex PV.
template <class P, bool GM, bool RH>
__global__ __launch_bounds__(kBlock) void tadj_flux_kernel(Params<P> p, TabAdj<P> a, const double* __restrict__ cot,
                                                           double c, int64_t t0, int64_t t1) {
    using C = typename P::C;
    using I = typename P::I;
    const Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        const int64_t f = t.first + threadIdx.x;
        if (f >= p.ne) return;
        C Fb = 0, qb = GM ? C(0) : C(a.qown[f]);
        for (int j = 0; j < a.TW; ++j) {
            const int k = ld(a.tf_k + j * p.ne + f);
            if (k == 255) break;
            const int64_t q = k * p.ne + ld(a.tf_e + j * p.ne + f);
            Fb += C(a.cF[q]);
            if constexpr (!GM) qb += C(a.cQ[q]);
        }
        if constexpr (GM) {
            for (int j = 0; j < a.QW; ++j) {
                const int k = ld(a.tq_k + j * p.ne + f);
                if (k == 255) break;
                qb += C(a.cQ[k * p.ne + ld(a.tq_e + j * p.ne + f)]);
            }
        }
        const int32_t c0 = ld(p.ec + f), c1 = ld(p.ec + p.ne + f);
        Fb -= ld(a.edc + f) * C(c * cot[p.ne + c0]) + ld(a.edc + p.ne + f) * C(c * cot[p.ne + c1]);
        a.Fb[f] = I(Fb);
        if constexpr (RH) {
            const int32_t v0 = ld(p.ev + f), v1 = ld(p.ev + p.ne + f);
            const C eta = ld(p.grh + f) * C(p.hv[v0]) + ld(p.grh + p.ne + f) * C(p.hv[v1]);
            a.Zb[f] = I(qb / eta);
            a.Eb[f] = I(-qb * C(p.Fq[f].q) / eta);
        } else {
            a.Zb[f] = I(qb / 2);  // each vertex's share of the mean
        }
    } else {
        const int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const uint32_t bits = ld(a.cbits + i);
        double bb = 0.0;
        for (int k = 0; k < p.kc; ++k) {
            if (!((bits >> (16 + k)) & 1)) continue;
            const int32_t e = ld(p.ce + k * p.nc + i);
            const double x = c * cot[e] * double(ld(p.invd + e));
            bb += (bits >> k) & 1 ? -x : x;
        }
        a.Bb[i] = bb;
    }
}

// T3: vertices (three edges and three cells each)
template <class P, class UIn, bool RH, bool VK>
__global__ __launch_bounds__(kBlock) void tadj_vertex_kernel(Params<P> p, TabAdj<P> a, const UIn* __restrict__ u,
                                                             int64_t tv) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t v = tile_of(blockIdx.x, tv, 0).first + threadIdx.x;  // XCD tiles
    if (v >= p.nv) return;
    const uint32_t bits = ld(a.vbits + v);
    C zb = 0, hb = 0;
    if constexpr (RH) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int32_t e = ld(p.ve + k * p.nv + v);
            const C r = ld(p.grh + ((bits >> k) & 1) * p.ne + e);
            zb += r * C(a.Zb[e]);
            hb += r * C(a.Eb[e]);
        }
    } else {
        C qb = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) qb += C(a.Zb[ld(p.ve + k * p.nv + v)]);
        const C hv = C(p.hv[v]);
        zb = qb / hv;
        hb = -qb * C(p.qv[v]) / hv;
    }
    if constexpr (VK) {
        C kb = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int32_t e = ld(p.ve + k * p.nv + v);
            hb += C(a.Fb[e]) * ld(p.vkh + ((bits >> k) & 1) * p.ne + e) * C(u[e]);
            const int32_t i = ld(p.vc + k * p.nv + v);
            kb += ld(p.vkk + ((bits >> (8 + 4 * k)) & 15) * p.nc + i) * C(a.Bb[i]);
        }
        a.Kvb[v] = I(kb);
    }
    a.zb[v] = I(zb);
    a.hvb[v] = I(hb);
}

// T4: edges (ubar) | cells (hbar), each with the RK4 adjoint epilogue
template <class P, class UIn, bool VK>
__global__ __launch_bounds__(kBlock) void tadj_final_kernel(Params<P> p, TabAdj<P> a, const UIn* __restrict__ u,
                                                            const double* __restrict__ h, AdjRk r, int64_t t0,
                                                            int64_t t1) {
    using C = typename P::C;
    const Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        const int64_t e = t.first + threadIdx.x;
        if (e >= p.ne) return;
        const AdjIn rin = adj_load(r, e);
        const int32_t c0 = ld(p.ec + e), c1 = ld(p.ec + p.ne + e), v0 = ld(p.ev + e), v1 = ld(p.ev + p.ne + e);
        const C ue = C(u[e]);
        C he = ld(p.eh + e) * C(h[c0]) + ld(p.eh + p.ne + e) * C(h[c1]);
        if constexpr (VK) he = he + ld(p.vkh + e) * C(p.hv[v0]) + ld(p.vkh + p.ne + e) * C(p.hv[v1]);
        double s = double(C(a.Fb[e]) * he) +
                   double(ld(a.ecurl + e) * C(a.zb[v0]) + ld(a.ecurl + p.ne + e) * C(a.zb[v1])) +
                   2.0 * double(ue) * (double(ld(a.ekin + e)) * a.Bb[c0] + double(ld(a.ekin + p.ne + e)) * a.Bb[c1]);
        if constexpr (VK)
            s += 2.0 * double(ue) * double(ld(a.evkw + e) * C(a.Kvb[v0]) + ld(a.evkw + p.ne + e) * C(a.Kvb[v1]));
        adj_store(r, e, s, rin);
    } else {
        const int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const AdjIn rin = adj_load(r, p.ne + i);
        const uint32_t bits = ld(a.cbits + i);
        double s = p.g * a.Bb[i];
        for (int k = 0; k < p.kc; ++k) {
            if (!((bits >> (16 + k)) & 1)) continue;
            const int32_t e = ld(p.ce + k * p.nc + i);
            s += double(C(a.Fb[e]) * ld(p.eh + ((bits >> k) & 1) * p.ne + e) * C(u[e]));
        }
        for (int j = 0; j < a.CW; ++j) s += double(ld(a.cvh + j * p.nc + i) * C(a.hvb[ld(a.corner + j * p.nc + i)]));
        adj_store(r, p.ne + i, s, rin);
    }
}

// the arguments a captured adjoint sweep bakes in
struct SweepKey {
    const double* y0 = nullptr;
    const double* ybarT = nullptr;
    double* ybar0 = nullptr;
    double* yT = nullptr;
    double dt = 0;
    int64_t N = -1, ck = -1;
    hipStream_t stream = nullptr;
    bool operator==(const SweepKey& o) const {
        return y0 == o.y0 && ybarT == o.ybarT && ybar0 == o.ybar0 && yT == o.yT && dt == o.dt && N == o.N &&
               ck == o.ck && stream == o.stream;
    }
};

// per-handle adjoint state
struct AdjState {
    std::vector<DevBuf> owned;
    AdjParams<Fp64> a64{};
    AdjParams<Fp32Bulk> a32{};
    TabAdj<Fp64> t64{};
    TabAdj<Fp32Bulk> t32{};
    // work of trisk_rk4_adjoint, kept (and grown) across calls so a captured sweep stays valid
    std::vector<DevBuf> work;
    int64_t work_ck_states = 0, work_seg_states = 0;
    double *ckp = nullptr, *seg = nullptr, *lam = nullptr, *acc = nullptr, *kbar = nullptr, *run = nullptr;
    // per step of a segment: the stage states s2, s3, s4 (su, sh: 3 per step) and, on the
    // cell-pair path, q_v and G at the step's input and at s2, s3, s4 (sqv, sG: 4 per step), kept
    // from the segment's recomputation for the VJPs at those states
    std::vector<void*> su, sqv, sG;
    std::vector<double*> sh;
    double* scratch = nullptr;  // the discarded output of a segment's last step
    // the sweep of the last graph-replayed call
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SweepKey key;
};

static void drop_sweep(AdjState* st) {
    if (st->exec) (void)hipGraphExecDestroy(st->exec);
    if (st->graph) (void)hipGraphDestroy(st->graph);
    st->exec = nullptr;
    st->graph = nullptr;
    st->key = SweepKey{};
}

static void adj_release(trisk_handle* H) {
    auto* st = static_cast<AdjState*>(H->adj);
    if (!st) return;
    drop_sweep(st);
    for (auto& b : st->work) (void)hipFree(b.p);
    for (auto& b : st->owned) (void)hipFree(b.p);
    delete st;
    H->adj = nullptr;
}

template <class P>
static AdjParams<P>& adj_params(AdjState* st) {
    if constexpr (P::id == TRISK_FP64)
        return st->a64;
    else
        return st->a32;
}

template <class P>
static TabAdj<P>& tab_params(AdjState* st) {
    if constexpr (P::id == TRISK_FP64)
        return st->t64;
    else
        return st->t32;
}

template <class P>
static const Params<P>& fwd_params(const trisk_handle* H) {
    if constexpr (P::id == TRISK_FP64)
        return H->p64;
    else
        return H->p32;
}

template <class P>
static Params<P>& params_of(trisk_handle* H) {
    if constexpr (P::id == TRISK_FP64)
        return H->p64;
    else
        return H->p32;
}

// The tables of the table-path adjoint (TabAdj), from the layout in the device numbering.
template <class P>
static int table_adj_setup(trisk_handle* H, AdjState* st) {
    using C = typename P::C;
    using I = typename P::I;
    const Layout& L = H->L;
    const Params<P>& p = fwd_params<P>(H);
    TabAdj<P>& a = tab_params<P>(st);
    const int64_t nc = L.nc, ne = L.ne, nv = L.nv;
    const int KC = L.kc;
    if (L.kv != 3) return set_err(TRISK_EINVAL, "the adjoint of the table kernels needs three edges per vertex");
    const bool gm = L.gassmann, vk = L.vertex_ke, pvform = L.gassmann && !L.gassmann_mean;
    const int K = gm ? L.kg : L.kt;
    const std::vector<int32_t>& NB = gm ? L.ge : L.te;
    const std::vector<double>& Wt = gm ? L.gw : L.tw;
    auto slot_of = [](const std::vector<int32_t>& t, int k_max, int64_t n, int64_t j, int64_t x) {
        for (int k = 0; k < k_max; ++k)
            if (t[size_t(k) * n + j] == x) return k;
        return -1;
    };
    // the PV edges of the Gassmann form (the slot order's PV slots when no table was kept)
    std::vector<int32_t> PV;
    if (pvform) {
        PV.resize(size_t(K) * ne);
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < K; ++k)
                PV[size_t(k) * ne + e] =
                    L.gp.empty() ? L.ge[size_t(gassmann_pv_slot(k)) * ne + e] : L.gp[size_t(k) * ne + e];
    }
    // (e, k) pairs with a nonzero weight whose target to[k, e] is f, by e then k
    auto transpose = [&](const std::vector<int32_t>& to, std::vector<int32_t>& te, std::vector<uint8_t>& tk) {
        std::vector<int32_t> cnt(ne, 0);
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < K; ++k)
                if (Wt[size_t(k) * ne + e] != 0.0) ++cnt[to[size_t(k) * ne + e]];
        const int width = ne ? *std::max_element(cnt.begin(), cnt.end()) : 0;
        te.assign(size_t(std::max(width, 1)) * ne, 0);
        tk.assign(size_t(std::max(width, 1)) * ne, uint8_t(255));
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < K; ++k)
                if (Wt[size_t(k) * ne + e] != 0.0) {
                    const int64_t f = to[size_t(k) * ne + e];
                    const int j = cnt[f]++;
                    te[size_t(j) * ne + f] = int32_t(e);
                    tk[size_t(j) * ne + f] = uint8_t(k);
                }
        return width;
    };
    std::vector<int32_t> tfe, tqe;
    std::vector<uint8_t> tfk, tqk;
    a.K = K;
    a.TW = transpose(NB, tfe, tfk);
    a.QW = pvform ? transpose(PV, tqe, tqk) : 0;
    // cells: side and slot bits; the continuity and kinetic-energy weights of every edge side
    const bool flux = P::id == TRISK_FP32_BULK && L.gflux;
    std::vector<uint32_t> cbits(nc, 0);
    for (int64_t i = 0; i < nc; ++i)
        for (int k = 0; k < KC; ++k) {
            if (L.div[size_t(k) * nc + i] == 0.0) continue;
            const int64_t e = L.ce[size_t(k) * nc + i];
            if (L.ec[e] != i && L.ec[ne + e] != i) return set_err(TRISK_EINVAL, "cell_edges and edge_cells disagree");
            cbits[i] |= (1u << (16 + k)) | (L.ec[ne + e] == i ? 1u << k : 0u);
        }
    std::vector<double> ecurl(2 * ne), ekin(2 * ne), evkw(vk ? 2 * ne : 0), edc(2 * ne);
    for (int64_t e = 0; e < ne; ++e)
        for (int sd = 0; sd < 2; ++sd) {
            const int64_t c = L.ec[sd * ne + e], v = L.ev[sd * ne + e];
            const int kc = slot_of(L.ce, KC, nc, c, e), kv = slot_of(L.ve, 3, nv, v, e);
            if (kc < 0 || kv < 0) return set_err(TRISK_EINVAL, "edge tables and cell or vertex tables disagree");
            const double dv = L.div[size_t(kc) * nc + c];
            ekin[sd * ne + e] = L.kin[size_t(kc) * nc + c];
            edc[sd * ne + e] = flux ? (dv > 0 ? 1.0 : -1.0) / L.area_c[c] : dv;
            ecurl[sd * ne + e] = L.curl[size_t(kv) * nv + v];
            if (vk) evkw[sd * ne + e] = L.vkw[size_t(kv) * nv + v];
        }
    // vertices: the side of v in each edge, its slot among the corners of each cell; the
    // corners of every cell with their vertex_height weights
    std::vector<uint32_t> vbits(nv, 0);
    std::vector<int32_t> ncorner(nc, 0);
    for (int64_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            const int64_t e = L.ve[size_t(k) * nv + v], i = L.vc[size_t(k) * nv + v];
            if (L.ev[e] != v && L.ev[ne + e] != v) return set_err(TRISK_EINVAL, "vertex_edges and edge_vertices disagree");
            if (L.ev[ne + e] == v) vbits[v] |= 1u << k;
            if (vk) {
                const int sl = slot_of(L.cv, KC, nc, i, v);
                if (sl < 0 || sl > 15) return set_err(TRISK_EINVAL, "vertex_cells and cell_vertices disagree");
                vbits[v] |= uint32_t(sl) << (8 + 4 * k);
            }
            ++ncorner[i];
        }
    a.CW = nc ? *std::max_element(ncorner.begin(), ncorner.end()) : 0;
    std::vector<int32_t> corner(size_t(std::max(a.CW, 1)) * nc, 0);
    std::vector<double> cvh(size_t(std::max(a.CW, 1)) * nc, 0.0);
    std::fill(ncorner.begin(), ncorner.end(), 0);
    for (int64_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            const int64_t i = L.vc[size_t(k) * nv + v];
            const int j = ncorner[i]++;
            corner[size_t(j) * nc + i] = int32_t(v);
            cvh[size_t(j) * nc + i] = L.vh[size_t(k) * nv + v];
        }
    auto coef = [&](const std::vector<double>& v, const C** out) -> int {
        if constexpr (sizeof(C) == sizeof(double))
            return upload(st->owned, v, reinterpret_cast<const double**>(out));
        else
            return upload(st->owned, to_float(v), reinterpret_cast<const float**>(out));
    };
    int rc = 0;
    a.nb = gm ? p.ge : p.te;
    a.w = gm ? p.gw : p.tw;
    if (pvform) rc |= upload(st->owned, PV, &a.pv);
    rc |= upload(st->owned, tfe, &a.tf_e);
    rc |= upload(st->owned, tfk, &a.tf_k);
    if (pvform) {
        rc |= upload(st->owned, tqe, &a.tq_e);
        rc |= upload(st->owned, tqk, &a.tq_k);
    }
    rc |= upload(st->owned, cbits, &a.cbits);
    rc |= upload(st->owned, vbits, &a.vbits);
    rc |= coef(ecurl, &a.ecurl);
    rc |= coef(ekin, &a.ekin);
    if (vk) rc |= coef(evkw, &a.evkw);
    rc |= coef(edc, &a.edc);
    rc |= upload(st->owned, corner, &a.corner);
    rc |= coef(cvh, &a.cvh);
    if (rc) return rc;
    void* q;
    auto get = [&](size_t bytes) -> void* {
        rc = alloc(st->owned, bytes, &q);
        return rc ? nullptr : q;
    };
    if (!(a.cF = static_cast<I*>(get(size_t(K) * ne * sizeof(I)))) ||
        !(a.cQ = static_cast<I*>(get(size_t(K) * ne * sizeof(I)))) || !(a.qown = static_cast<I*>(get(ne * sizeof(I)))) ||
        !(a.Fb = static_cast<I*>(get(ne * sizeof(I)))) || !(a.Zb = static_cast<I*>(get(ne * sizeof(I)))) ||
        !(a.Eb = static_cast<I*>(get(ne * sizeof(I)))) || !(a.Bb = static_cast<double*>(get(nc * sizeof(double)))) ||
        !(a.zb = static_cast<I*>(get(nv * sizeof(I)))) || !(a.hvb = static_cast<I*>(get(nv * sizeof(I)))) ||
        !(a.Kvb = static_cast<I*>(get(nv * sizeof(I)))))
        return rc;
    return 0;
}

template <class P>
static int adj_setup(trisk_handle* H) {
    using C = typename P::C;
    using I = typename P::I;
    if (H->adj) return 0;
    auto* st = new AdjState;
    H->adj = st;
    H->adj_release = adj_release;
    if (!H->L.pair) {
        const int rc = table_adj_setup<P>(H, st);
        if (rc) adj_release(H);  // a later call builds the tables again and reports the same error
        return rc;
    }
    const Layout& L = H->L;
    AdjParams<P>& w = adj_params<P>(st);
    auto coef = [&](const std::vector<double>& v, const C** out) -> int {
        if constexpr (sizeof(C) == sizeof(double))
            return upload(st->owned, v, reinterpret_cast<const double**>(out));
        else
            return upload(st->owned, to_float(v), reinterpret_cast<const float**>(out));
    };
    int rc = 0;
    rc |= coef(L.adj_ecurl, &w.ecurl);
    rc |= coef(L.adj_ekin, &w.ekin);
    rc |= upload(st->owned, L.adj_cv3, &w.cv3);
    rc |= upload(st->owned, L.adj_vslot, &w.vslot);
    if (rc) return rc;
    void* q;
    if ((rc = alloc(st->owned, L.nc * sizeof(double), &q))) return rc;
    w.Bbar = static_cast<double*>(q);
    if ((rc = alloc(st->owned, 6 * L.nc * sizeof(I), &q))) return rc;
    w.Tbar = static_cast<I*>(q);
    if ((rc = alloc(st->owned, 6 * L.nc * sizeof(I), &q))) return rc;
    w.QC = static_cast<I*>(q);
    if ((rc = alloc(st->owned, L.nv * sizeof(I), &q))) return rc;
    w.zbar = static_cast<I*>(q);
    if ((rc = alloc(st->owned, 3 * L.nv * sizeof(I), &q))) return rc;
    w.VH = static_cast<I*>(q);
    if ((rc = alloc(st->owned, 2 * L.ne * sizeof(I), &q))) return rc;
    w.GL = static_cast<I*>(q);
    return 0;
}

// One vector-Jacobian product of the table kernels (T1-T4 after P1 and P2 of the stage).
template <class P, class UIn>
static int table_vjp(trisk_handle* H, const UIn* u, const double* h, const double* cot, double c, const AdjRk& r,
                     hipStream_t s) {
    const Layout& L = H->L;
    const Params<P>& p = fwd_params<P>(H);
    const TabAdj<P>& a = tab_params<P>(static_cast<AdjState*>(H->adj));
    int rc = table_state_pass(H, u, std::is_same<UIn, double>::value, h, s);
    if (rc) return rc;
    const bool gm = L.gassmann && !L.gassmann_mean, rh = L.gassmann, vk = L.vertex_ke;
    const int64_t te = tiles_per_xcd(L.ne), tv = tiles_per_xcd(L.nv);
    const dim3 ge(unsigned(kXcd * te)), gv(unsigned(kXcd * tv)), b(kBlock);
    const int64_t t0 = tiles_per_xcd(L.ne), t1 = tiles_per_xcd(L.nc);
    const dim3 g2(unsigned(kXcd * (t0 + t1)));
    if (gm) {
        hipLaunchKernelGGL((tadj_slot_kernel<P, true>), ge, b, 0, s, p, a, cot, c, te);
        hipLaunchKernelGGL((tadj_flux_kernel<P, true, true>), g2, b, 0, s, p, a, cot, c, t0, t1);
    } else {
        hipLaunchKernelGGL((tadj_slot_kernel<P, false>), ge, b, 0, s, p, a, cot, c, te);
        if (rh)
            hipLaunchKernelGGL((tadj_flux_kernel<P, false, true>), g2, b, 0, s, p, a, cot, c, t0, t1);
        else
            hipLaunchKernelGGL((tadj_flux_kernel<P, false, false>), g2, b, 0, s, p, a, cot, c, t0, t1);
    }
    if (rh && vk)
        hipLaunchKernelGGL((tadj_vertex_kernel<P, UIn, true, true>), gv, b, 0, s, p, a, u, tv);
    else if (rh)
        hipLaunchKernelGGL((tadj_vertex_kernel<P, UIn, true, false>), gv, b, 0, s, p, a, u, tv);
    else if (vk)
        hipLaunchKernelGGL((tadj_vertex_kernel<P, UIn, false, true>), gv, b, 0, s, p, a, u, tv);
    else
        hipLaunchKernelGGL((tadj_vertex_kernel<P, UIn, false, false>), gv, b, 0, s, p, a, u, tv);
    if (vk)
        hipLaunchKernelGGL((tadj_final_kernel<P, UIn, true>), g2, b, 0, s, p, a, u, h, r, t0, t1);
    else
        hipLaunchKernelGGL((tadj_final_kernel<P, UIn, false>), g2, b, 0, s, p, a, u, h, r, t0, t1);
    HIP_TRY(hipGetLastError());
    return 0;
}

// One vector-Jacobian product at the stage (u, h) for the cotangent c * cot, ending in the
// epilogue r.  Cell-pair path: with kqv and kG given, the forward q_v and G of (u, h) (all the
// VJP reads of W1) are those kept there by the segment's recomputation, and W1 is not run again.
template <class P, class UIn>
static int vjp(trisk_handle* H, const UIn* u, const double* h, const double* cot, double c, const AdjRk& r,
               hipStream_t s, void* kqv = nullptr, void* kG = nullptr) {
    if (!H->L.pair) return table_vjp<P, UIn>(H, u, h, cot, c, r, s);
    const Layout& L = H->L;
    auto* st = static_cast<AdjState*>(H->adj);
    Params<P> p = fwd_params<P>(H);
    const AdjParams<P>& w = adj_params<P>(st);
    if (kqv && kG) {
        p.qv = static_cast<typename P::I*>(kqv);
        p.G = static_cast<typename P::I*>(kG);
    } else {
        const int rc = pair_state_pass(H, u, std::is_same<UIn, double>::value, h, s);
        if (rc) return rc;
    }
    // XCD x takes the x-th eighth of each range (tile_of), so the gathers of its tiles stay in
    // its L2
    const int64_t tc = tiles_per_xcd(L.nc), tv = tiles_per_xcd(L.nv), te = tiles_per_xcd(L.ne);
    const dim3 gc(unsigned(kXcd * tc)), gv(unsigned(kXcd * tv)), ge(unsigned(kXcd * te)), b(kBlock);
    hipLaunchKernelGGL((adj_cell_kernel<P>), gc, b, 0, s, p, w, cot, c, tc);
    hipLaunchKernelGGL((adj_vertex_kernel<P, UIn>), gv, b, 0, s, p, w, u, h, tv);
    hipLaunchKernelGGL((adj_edge_kernel<P, UIn>), ge, b, 0, s, p, w, u, h, r, te);
    hipLaunchKernelGGL((adj_height_kernel<P, UIn>), gc, b, 0, s, p, w, u, r, tc);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class P>
static int tendency_vjp(trisk_handle* H, const double* y, const double* lam, double* ybar, hipStream_t s) {
    int rc = adj_setup<P>(H);
    if (rc) return rc;
    AdjRk r{};
    r.mode = 0;
    r.out = ybar;
    return vjp<P, double>(H, y, y + H->L.ne, lam, 1.0, r, s);
}

// Work of a sweep with nck checkpoints and segments of up to seg states.
template <class P>
static int sweep_work(trisk_handle* H, AdjState* st, int64_t nck, int64_t seg) {
    using U = typename P::U;
    if (nck <= st->work_ck_states && seg <= st->work_seg_states) return 0;
    nck = std::max(nck, st->work_ck_states);
    seg = std::max(seg, st->work_seg_states);
    drop_sweep(st);
    HIP_TRY(hipDeviceSynchronize());  // a replay in flight may still use the old work
    for (auto& b : st->work) (void)hipFree(b.p);
    st->work.clear();
    st->work_ck_states = st->work_seg_states = 0;
    const Layout& L = H->L;
    const size_t sb = (L.ne + L.nc) * sizeof(double);
    void* q;
    int rc;
    auto get = [&](size_t bytes) -> void* {
        rc = alloc(st->work, bytes, &q);
        return rc ? nullptr : q;
    };
    if (!(st->ckp = static_cast<double*>(get(nck * sb))) || !(st->seg = static_cast<double*>(get(seg * sb))) ||
        !(st->lam = static_cast<double*>(get(sb))) || !(st->acc = static_cast<double*>(get(sb))) ||
        !(st->kbar = static_cast<double*>(get(sb))) || !(st->run = static_cast<double*>(get(sb))) ||
        !(st->scratch = static_cast<double*>(get(sb))))
        return rc;
    st->su.assign(size_t(3 * seg), nullptr);
    st->sh.assign(size_t(3 * seg), nullptr);
    for (int64_t k = 0; k < 3 * seg; ++k)
        if (!(st->su[k] = get(L.ne * sizeof(U))) || !(st->sh[k] = static_cast<double*>(get(L.nc * sizeof(double)))))
            return rc;
    st->sqv.assign(L.pair ? size_t(4 * seg) : 0, nullptr);
    st->sG.assign(L.pair ? size_t(4 * seg) : 0, nullptr);
    for (size_t k = 0; k < st->sqv.size(); ++k)
        if (!(st->sqv[k] = get(L.nv * sizeof(typename P::I))) || !(st->sG[k] = get(L.ne * sizeof(typename P::I))))
            return rc;
    st->work_ck_states = nck;
    st->work_seg_states = seg;
    return 0;
}

// Issues the checkpointed adjoint of N RK4 steps on s: the forward run keeps the state
// every ck steps; each segment is recomputed from its checkpoint, and each step's stage
// states again from its input, before the step is transposed.  Nothing synchronises.
template <class P>
static int enqueue_sweep(trisk_handle* H, AdjState* st, const double* y0, double dt, int64_t N, int64_t ck,
                         const double* ybarT, double* ybar0, double* yT, hipStream_t s) {
    using U = typename P::U;
    const Layout& L = H->L;
    const int64_t n = L.ne + L.nc;
    const size_t sb = n * sizeof(double);
    const int64_t nck = N > 0 ? (N + ck - 1) / ck : 0;
    int rc;
    double* C0 = st->ckp;
    double* S0 = st->seg;
    double* y = yT ? yT : st->run;
    HIP_TRY(hipMemcpyAsync(y, y0, sb, hipMemcpyDeviceToDevice, s));
    for (int64_t k = 0; k < N; ++k) {
        if (k % ck == 0) HIP_TRY(hipMemcpyAsync(C0 + (k / ck) * n, y, sb, hipMemcpyDeviceToDevice, s));
        if ((rc = rk4_steps_direct(H, y, dt, 1, s))) return rc;
    }
    double* la = st->lam;
    HIP_TRY(hipMemcpyAsync(la, ybarT, sb, hipMemcpyDeviceToDevice, s));
    AdjRk a{};
    a.lam = la;
    a.acc = st->acc;
    a.kbar = st->kbar;
    a.dt = dt;
    Rk r{};
    r.acc = H->acc;
    r.dt = dt;
    r.dt2 = dt / 2;
    r.dt6 = dt / 6;
    Params<P>& fp = params_of<P>(H);
    const auto fwd_qv = fp.qv;
    const auto fwd_G = fp.G;
    for (int64_t g = nck - 1; g >= 0; --g) {
        const int64_t first = g * ck, len = std::min(ck, N - first);
        // the segment again from its checkpoint, stage by stage with the forward's own
        // arithmetic: step k's stage states stay in su/sh[3k + i], and on the cell-pair path the
        // q_v and G at its tendency inputs (y_k, s2, s3, s4) in sqv/sG[4k + i].  A step's last
        // tendency updates a copy of y_k in place into y_{k+1}; the segment's last step leaves
        // it out (its output is the next checkpoint) and its VJP at s4 runs W1 instead.
        HIP_TRY(hipMemcpyAsync(S0, C0 + g * n, sb, hipMemcpyDeviceToDevice, s));
        for (int64_t k = 0; k < len; ++k) {
            double* yn = S0 + k * n;
            double* ynext = k + 1 < len ? S0 + (k + 1) * n : st->scratch;
            HIP_TRY(hipMemcpyAsync(ynext, yn, sb, hipMemcpyDeviceToDevice, s));
            r.y = ynext;
            const int stages = k + 1 < len ? 4 : 3;
            for (int i = 0; i < stages; ++i) {
                r.mode = i + 1;
                if (i < 3) {
                    r.su = st->su[3 * k + i];
                    r.sh = st->sh[3 * k + i];
                }
                const void* uin = i == 0 ? static_cast<const void*>(yn) : st->su[3 * k + i - 1];
                const double* hin = i == 0 ? yn + L.ne : st->sh[3 * k + i - 1];
                if (L.pair) {
                    fp.qv = static_cast<typename P::I*>(st->sqv[4 * k + i]);
                    fp.G = static_cast<typename P::I*>(st->sG[4 * k + i]);
                }
                rc = tendency_stage(H, uin, i == 0, hin, r, s);
                fp.qv = fwd_qv;
                fp.G = fwd_G;
                if (rc) return rc;
            }
        }
        for (int64_t k = len - 1; k >= 0; --k) {
            double* yn = S0 + k * n;
            auto kqv = [&](int i) { return L.pair && (i < 3 || k + 1 < len) ? st->sqv[4 * k + i] : nullptr; };
            auto kG = [&](int i) { return L.pair && (i < 3 || k + 1 < len) ? st->sG[4 * k + i] : nullptr; };
            const U* su2 = static_cast<const U*>(st->su[3 * k + 2]);
            const U* su1 = static_cast<const U*>(st->su[3 * k + 1]);
            const U* su0 = static_cast<const U*>(st->su[3 * k]);
            a.mode = 4;
            if ((rc = vjp<P, U>(H, su2, st->sh[3 * k + 2], la, dt / 6, a, s, kqv(3), kG(3)))) return rc;
            a.mode = 3;
            if ((rc = vjp<P, U>(H, su1, st->sh[3 * k + 1], a.kbar, 1.0, a, s, kqv(2), kG(2)))) return rc;
            a.mode = 2;
            if ((rc = vjp<P, U>(H, su0, st->sh[3 * k], a.kbar, 1.0, a, s, kqv(1), kG(1)))) return rc;
            a.mode = 1;
            if ((rc = vjp<P, double>(H, yn, yn + L.ne, a.kbar, 1.0, a, s, kqv(0), kG(0)))) return rc;
        }
    }
    HIP_TRY(hipMemcpyAsync(ybar0, la, sb, hipMemcpyDeviceToDevice, s));
    return 0;
}

// The adjoint sweep, replayed through one captured HIP graph when the handle has graphs on
// (trisk_set_graphs) and the arguments match the last capture.
template <class P>
static int rk4_adjoint(trisk_handle* H, const double* y0, double dt, int64_t N, int64_t ck, const double* ybarT,
                       double* ybar0, double* yT, hipStream_t s) {
    int rc = adj_setup<P>(H);
    if (rc) return rc;
    auto* st = static_cast<AdjState*>(H->adj);
    const int64_t nck = N > 0 ? (N + ck - 1) / ck : 0;
    if ((rc = sweep_work<P>(H, st, std::max<int64_t>(nck, 1), std::min(ck, std::max<int64_t>(N, 1))))) return rc;
    const bool graph = H->graphs && !H->timing && s != nullptr;
    if (!graph) {
        if ((rc = enqueue_sweep<P>(H, st, y0, dt, N, ck, ybarT, ybar0, yT, s))) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        return 0;
    }
    SweepKey key;
    key.y0 = y0;
    key.ybarT = ybarT;
    key.ybar0 = ybar0;
    key.yT = yT;
    key.dt = dt;
    key.N = N;
    key.ck = ck;
    key.stream = s;
    if (!st->exec || !(st->key == key)) {
        drop_sweep(st);
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = enqueue_sweep<P>(H, st, y0, dt, N, ck, ybarT, ybar0, yT, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(TRISK_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        st->graph = g;
        HIP_TRY(hipGraphInstantiate(&st->exec, g, nullptr, nullptr, 0));
        st->key = key;
    }
    HIP_TRY(hipGraphLaunch(st->exec, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

static int adj_check(trisk_handle* h) {
    if (!h) return set_err(TRISK_EINVAL, "null handle");
    return 0;
}

}  // namespace trisk

using namespace trisk;

extern "C" {

int trisk_tendency_vjp(trisk_handle* h, const double* y_dev, const double* dybar_dev, double* ybar_dev,
                       void* stream) {
    if (h && h->L.icon) return set_err(TRISK_EINVAL, "no adjoint for the ICON triangle model");
    int rc = adj_check(h);
    if (rc) return rc;
    if (!y_dev || !dybar_dev || !ybar_dev) return set_err(TRISK_EINVAL, "null state pointer");
    if (ybar_dev == y_dev || ybar_dev == dybar_dev) return set_err(TRISK_EINVAL, "ybar must not alias y or dybar");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    return h->precision == TRISK_FP64 ? tendency_vjp<Fp64>(h, y_dev, dybar_dev, ybar_dev, s)
                                      : tendency_vjp<Fp32Bulk>(h, y_dev, dybar_dev, ybar_dev, s);
}

int trisk_rk4_adjoint(trisk_handle* h, const double* y0_dev, double dt, int64_t nsteps, int64_t checkpoint_every,
                      const double* ybarT_dev, double* ybar0_dev, double* yT_dev, void* stream) {
    if (h && h->L.icon) return set_err(TRISK_EINVAL, "no adjoint for the ICON triangle model");
    int rc = adj_check(h);
    if (rc) return rc;
    if (!y0_dev || !ybarT_dev || !ybar0_dev) return set_err(TRISK_EINVAL, "null state pointer");
    if (nsteps < 0 || checkpoint_every < 0) return set_err(TRISK_EINVAL, "nsteps and checkpoint_every must be >= 0");
    if (yT_dev && (yT_dev == y0_dev || yT_dev == ybarT_dev || yT_dev == ybar0_dev))
        return set_err(TRISK_EINVAL, "yT must not alias the other vectors");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t ck = checkpoint_every > 0 ? checkpoint_every : 8;
    return h->precision == TRISK_FP64
               ? rk4_adjoint<Fp64>(h, y0_dev, dt, nsteps, ck, ybarT_dev, ybar0_dev, yT_dev, s)
               : rk4_adjoint<Fp32Bulk>(h, y0_dev, dt, nsteps, ck, ybarT_dev, ybar0_dev, yT_dev, s);
}

}  // extern "C"
