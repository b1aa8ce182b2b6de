// gfx950 kernels and C ABI of the TRiSK shallow-water step (include/trisk_hip.h).
//
// One tendency (TRiSKShallowWater.jl:379-394) is three launches, each covering two
// entity ranges (cells | vertices, edges | cells, edges | cells):
//   P1 cell_vertex : K_i, B_i = K_i + g h_i  (kinetic_energy_kernel!, :202-210)
//                    q_v = (zeta_v + f_v)/h_v, and zeta_v, h_v, K_v when needed
//                    (vertex_pv_kernel! :246-259, vertex_kinetic_energy_kernel! :212-219)
//   P2 edge        : F_e = h_e u_e and the edge PV q_e, stored as one (F, q) pair
//                    (mass_flux_kernel! :232-244, edge_pv_kernel! :261-264,
//                    rhombus_pv_kernel! :266-273); with vertex kinetic energy also the
//                    cells' vertex part of K and B (cell_from_vertex_kernel! :222-230)
//   P3 update      : du_e (momentum_kernel! :275-284 or gassmann_momentum_kernel!
//                    :286-296) and dh_i (continuity_kernel! :298-305), each followed by
//                    the RK4 stage update of rk4_step! (RungeKutta.jl:18-28) or a plain
//                    store of the tendency.
// Every gather sums its slots in the reference's table order; FP contraction is off, so
// the FP64 path rounds exactly as the reference formulas (and the oracle) do.  Each
// thread owns one entity (gathers only, no atomics); block -> tile mapping gives each
// XCD contiguous ranges of the Hilbert-ordered entities so neighbour data is reused in
// its L2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "trisk_device.hpp"

namespace trisk {

static thread_local std::string g_error;
// The message slot of the handle whose entry point is running on this thread (ErrorScope).
static thread_local std::string* g_handle_error = nullptr;

int set_err(int code, const std::string& msg) {
    g_error = msg;
    if (g_handle_error) *g_handle_error = msg;
    return code;
}

// While an entry point on handle h runs, its errors are also kept in h (for
// trisk_handle_last_error); nested calls on other handles restore the outer slot.
struct ErrorScope {
    std::string* prev;
    explicit ErrorScope(const trisk_handle* h) : prev(g_handle_error) { g_handle_error = h ? &h->last_error : nullptr; }
    ~ErrorScope() { g_handle_error = prev; }
    ErrorScope(const ErrorScope&) = delete;
    ErrorScope& operator=(const ErrorScope&) = delete;
};

int lloyd_set_err(int code, const std::string& msg) { return set_err(code, msg); }

// ---------------------------------------------------------------------------------
// Stencil widths of the table kernels: compile-time for hexagon meshes (6 cell edges, 3
// vertex edges, 10 TRiSK and 12 Gassmann terms), so that every gather loop unrolls and its
// table loads issue together, or the handle's runtime widths.  The summation order is the
// same either way.

template <int KC, int KV, int KT, int KG>
struct Widths {
    static constexpr int cw = KC, vw = KV, tw = KT, gw = KG;
    template <class P>
    __device__ static int kc(const Params<P>& p) { return KC ? KC : p.kc; }
    template <class P>
    __device__ static int kv(const Params<P>& p) { return KV ? KV : p.kv; }
    template <class P>
    __device__ static int kt(const Params<P>& p) { return KT ? KT : p.kt; }
    template <class P>
    __device__ static int kg(const Params<P>& p) { return KG ? KG : p.kg; }
};
using AnyWidths = Widths<0, 0, 0, 0>;
using HexWidths = Widths<6, 3, 10, 12>;

// ---------------------------------------------------------------------------------
// P1: cells (K, B) | vertices (q_v [, zeta_v, h_v] [, K_v])

// QV false: q_v is not stored (the Gassmann flux forms the edge PV from zeta_v and h_v).
// REC (FP32-bulk, hexagon widths): the rows come from the entity records r1c / r1v staged
// through LDS instead of the slot-major arrays; the values and the order of the sums are
// the same.
template <class P, class UIn, class HIn, bool VK, bool ZH, class W = AnyWidths, bool QV = true, bool REC = false>
__global__ __launch_bounds__(kBlock) void cell_vertex_kernel(Params<P> p, const UIn* __restrict__ u,
                                                             const HIn* __restrict__ h, int64_t t0, int64_t t1) {
    using C = typename P::C;
    using I = typename P::I;
    static_assert(!REC || (P::id == TRISK_FP32_BULK && W::cw == 6 && W::vw == 3), "records: FP32-bulk hexagons");
    // vertex record: indices 4 words, curl 3, vh 3, [f when q_v is formed], [vkw 3]
    constexpr int RF = QV ? 1 : 0, RV = 10 + RF + (VK ? 3 : 0);
    __shared__ __attribute__((aligned(16))) int32_t lds[REC ? kBlock * RV : 1];
    Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        int64_t i = t.first + threadIdx.x;
        C kin[W::cw ? W::cw : 1];
        IndexWords<W::cw> ew;
        if constexpr (REC) {
            if (!stage_records<9>(p.r1c, t.first, p.nc, lds) || i >= p.nc) return;
            const int32_t* r = lds + 9 * threadIdx.x;
#pragma unroll
            for (int q = 0; q < 3; ++q) ew.w[q] = r[q];
#pragma unroll
            for (int k = 0; k < 6; ++k) kin[k] = fbits(r[3 + k]);
        } else {
            if (i >= p.nc) return;
            ew = index_words<W::cw>(p.ce16p, p.nc, i);
        }
        const Index<W::cw> E(ew, p.ce, p.nc, i, 3 * i);
        C s = 0;
        for (int k = 0; k < W::kc(p); ++k) {
            C ue = C(u[E[k]]);
            if constexpr (REC)
                s += kin[k] * (ue * ue);
            else
                s += ld(p.kin + k * p.nc + i) * (ue * ue);
        }
        double hi = double(h[i]);
        double K = double(s);
        if (VK) {
            // completed with the vertex part in P2; a tendency reads only B, the diagnostics K
            if (p.K_out) p.K[i] = K;
            p.B[i] = K + p.g * hi;
        } else {
            p.B[i] = K + p.g * hi;
            if (p.K_out) p.K_out[i] = K;
        }
    } else {
        int64_t v = t.first + threadIdx.x;
        IndexWords<W::vw> vew, vcw;
        C cu[W::vw ? W::vw : 1], ch[W::vw ? W::vw : 1], kw[W::vw ? W::vw : 1], f = 0;
        if constexpr (REC) {
            if (!stage_records<RV>(p.r1v, t.first, p.nv, lds) || v >= p.nv) return;
            const int32_t* r = lds + RV * threadIdx.x;
            vew.w[0] = r[0];
            vew.w[1] = r[1];
            vcw.w[0] = r[2];
            vcw.w[1] = r[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                cu[k] = fbits(r[4 + k]);
                ch[k] = fbits(r[7 + k]);
                if (VK) kw[k] = fbits(r[10 + RF + k]);
            }
            if (QV) f = fbits(r[10]);
        } else {
            if (v >= p.nv) return;
            vew = index_words<W::vw>(p.ve16p, p.nv, v);
            vcw = index_words<W::vw>(p.vc16p, p.nv, v);
        }
        const Index<W::vw> VE(vew, p.ve, p.nv, v, (3 * v) / 2), VC(vcw, p.vc, p.nv, v, v / 2);
        C z = 0, eta = 0;
        if constexpr (REC) {
            for (int k = 0; k < 3; ++k) z += cu[k] * C(u[VE[k]]);
            for (int k = 0; k < 3; ++k) eta += ch[k] * C(h[VC[k]]);
        } else {
            for (int k = 0; k < W::kv(p); ++k) z += ld(p.curl + k * p.nv + v) * C(u[VE[k]]);
            for (int k = 0; k < W::kv(p); ++k) eta += ld(p.vh + k * p.nv + v) * C(h[VC[k]]);
        }
        if (QV || p.qv_out) {
            if constexpr (!REC) f = ld(p.fv + v);
            C q = (z + f) / eta;
            if (QV) p.qv[v] = I(q);
            if (p.qv_out) {
                p.qv_out[v] = double(q);
                p.zv_out[v] = double(z);
                p.hv_out[v] = double(eta);
            }
        }
        if (ZH) {
            p.zv[v] = I(z);
            p.hv[v] = I(eta);
        }
        if (VK) {
            C s = 0;
            for (int k = 0; k < W::kv(p); ++k) {
                C ue = C(u[VE[k]]);
                if constexpr (REC)
                    s += kw[k] * (ue * ue);
                else
                    s += ld(p.vkw + k * p.nv + v) * (ue * ue);
            }
            p.Kv[v] = I(s);
        }
    }
}

// ---------------------------------------------------------------------------------
// P2: edges (F, q_e) | cells (vertex part of K, B; only with vertex kinetic energy)

// REC (FP32-bulk, hexagon widths): the rows from the records r2e / r2c staged through LDS.
// BH (vertex kinetic energy, every cell kinetic weight 0): P1 skipped the cells, and the cell
// part forms B = g h + the vertex part itself; P1 would have stored 0 + g h, which is g h, so B
// is bit for bit the same.
// With the records P1 stores A_v zeta_v and A_v h_v (its record coefficients carry A_v), so
// the rhombus PV (r_0 zeta_0 + r_1 zeta_1 + f_e)/(r_0 h_0 + r_1 h_1), r_s = A_s/(A_0 + A_1),
// is (A_0 zeta_0 + A_1 zeta_1 + (A_0 + A_1) f_e)/(A_0 h_0 + A_1 h_1): one coefficient per
// edge, and the vertex part of the mass flux takes vkh_s / A_s.
template <class P, class UIn, class HIn, bool VK, bool GM, class W = AnyWidths, bool REC = false, bool BH = false>
__global__ __launch_bounds__(kBlock) void edge_kernel(Params<P> p, const UIn* __restrict__ u,
                                                      const HIn* __restrict__ h, int64_t t0, int64_t t1) {
    using C = typename P::C;
    using I = typename P::I;
    static_assert(!REC || (P::id == TRISK_FP32_BULK && W::cw == 6), "records: FP32-bulk hexagons");
    constexpr int RE = 4 + (VK ? 2 : 0) + (GM ? 1 : 0);
    __shared__ __attribute__((aligned(16))) int32_t lds[REC ? kBlock * 9 : 1];
    Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        int64_t e = t.first + threadIdx.x;
        IndexWords<W::cw ? 2 : 0> ecw, evw;
        C eh0, eh1, vk0 = 0, vk1 = 0, r0 = 0, r1 = 0, fe = 0;
        if constexpr (REC) {
            if (!stage_records<RE>(p.r2e, t.first, p.ne, lds) || e >= p.ne) return;
            const int32_t* r = lds + RE * threadIdx.x;
            ecw.w[0] = r[0];
            evw.w[0] = r[1];
            eh0 = fbits(r[2]);
            eh1 = fbits(r[3]);
            if (VK) {
                vk0 = fbits(r[4]);
                vk1 = fbits(r[5]);
            }
            if (GM) fe = fbits(r[RE - 1]);  // (A_0 + A_1) f_e
        } else {
            if (e >= p.ne) return;
            ecw = index_words<W::cw ? 2 : 0>(p.ec16p, p.ne, e);
            evw = index_words<W::cw ? 2 : 0>(p.ev16p, p.ne, e);
            eh0 = ld(p.eh + e);
            eh1 = ld(p.eh + p.ne + e);
            if (VK) {
                vk0 = ld(p.vkh + e);
                vk1 = ld(p.vkh + p.ne + e);
            }
            if (GM) {
                r0 = ld(p.grh + e);
                r1 = ld(p.grh + p.ne + e);
                fe = ld(p.gf + e);
            }
        }
        const Index<W::cw ? 2 : 0> EC(ecw, p.ec, p.ne, e, e / 3), EV(evw, p.ev, p.ne, e, (2 * e) / 3);
        const int32_t c0 = EC[0], c1 = EC[1], v0 = EV[0], v1 = EV[1];
        C he = eh0 * C(h[c0]) + eh1 * C(h[c1]);
        if (VK) he = he + vk0 * C(p.hv[v0]) + vk1 * C(p.hv[v1]);
        C F = he * C(u[e]);
        C q;
        if (GM) {
            if constexpr (REC) {
                q = (C(p.zv[v0]) + C(p.zv[v1]) + fe) / (C(p.hv[v0]) + C(p.hv[v1]));
            } else {
                C z = r0 * C(p.zv[v0]) + r1 * C(p.zv[v1]);
                C eta = r0 * C(p.hv[v0]) + r1 * C(p.hv[v1]);
                q = (z + fe) / eta;
            }
        } else {
            q = (C(p.qv[v0]) + C(p.qv[v1])) / 2;
        }
        p.Fq[e] = Pair<I>{I(F), I(q)};
    } else {
        if (!VK) return;
        int64_t i = t.first + threadIdx.x;
        IndexWords<W::cw> cvw;
        C kk[W::cw ? W::cw : 1];
        if constexpr (REC) {
            if (!stage_records<9>(p.r2c, t.first, p.nc, lds) || i >= p.nc) return;
            const int32_t* r = lds + 9 * threadIdx.x;
#pragma unroll
            for (int q = 0; q < 3; ++q) cvw.w[q] = r[q];
#pragma unroll
            for (int k = 0; k < 6; ++k) kk[k] = fbits(r[3 + k]);
        } else {
            if (i >= p.nc) return;
            cvw = index_words<W::cw>(p.cv16p, p.nc, i);
        }
        const Index<W::cw> CV(cvw, p.cv, p.nc, i, 2 * i);
        C s = 0;
        for (int k = 0; k < W::kc(p); ++k) {
            if constexpr (REC)
                s += kk[k] * C(p.Kv[CV[k]]);
            else
                s += ld(p.vkk + k * p.nc + i) * C(p.Kv[CV[k]]);
        }
        if constexpr (BH)
            p.B[i] = p.g * double(h[i]) + double(s);
        else
            p.B[i] = p.B[i] + double(s);
        if (p.K_out) {
            double K = p.K[i] + double(s);
            p.K[i] = K;
            p.K_out[i] = K;
        }
    }
}

// ---------------------------------------------------------------------------------
// P3: edges (du) | cells (dh), then the RK4 stage update

// GSLOT (12 Gassmann slots in the order of gassmann_coefficients): the PV of slot k is the q
// of the pair already gathered for slot gassmann_pv_slot(k), so neither the PV-edge table nor
// a second gather is needed; the products and their order are those of the table form.
// GF (flux form, Layout::gflux): the pairs hold G = l F and the continuity is
// -(sum_k s_k G_k)/A_i, with the signs from csign.  GREC (FP32-bulk slot form): each edge
// reads its slot edges, cells, 1/d_e and weights as one record (Params::rg) with five 16-byte
// loads instead of 20 slot-major loads; the products and their order are unchanged.
template <class P, bool GM, bool GMEAN, class W = AnyWidths, bool GSLOT = false, bool GF = false, bool GREC = false>
__global__ __launch_bounds__(kBlock) void update_kernel(Params<P> p, Rk r, int64_t t0, int64_t t1) {
    using C = typename P::C;
    using I = typename P::I;
    using U = typename P::U;
    Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        int64_t e = t.first + threadIdx.x;
        if constexpr (GREC) {
            static_assert(GSLOT && P::id == TRISK_FP32_BULK, "records: FP32-bulk slot form");
            // the block's records staged through LDS with coalesced 16-byte loads; at a stride
            // of 80 bytes the 16-byte LDS reads of 16 lanes fall on distinct banks
            __shared__ __attribute__((aligned(16))) int32_t srec[kBlock * 20];
            if (t.first >= p.ne) return;  // the whole block
            // the RK4 loads first: they do not wait for the records
            const RkIn rin = e < p.ne ? rk_load(r, e) : RkIn{0.0, 0.0};
            const int64_t n = p.ne - t.first < kBlock ? p.ne - t.first : kBlock;
            const f32x4* src = reinterpret_cast<const f32x4*>(p.rg + 20 * t.first);
            for (int w = threadIdx.x; w < int(5 * n); w += kBlock)
                reinterpret_cast<f32x4*>(srec)[w] = __builtin_nontemporal_load(src + w);
            __syncthreads();
            if (e >= p.ne) return;
            double du;
            int32_t rec[20];
            __builtin_memcpy(rec, &srec[20 * threadIdx.x], 80);
            IndexWords<12> gw16;
#pragma unroll
            for (int q = 0; q < 6; ++q) gw16.w[q] = rec[q];
            IndexWords<2> ec16;
            ec16.w[0] = rec[6];
            const Index<2> EC(ec16, p.ec, p.ne, e, e / 3);
            const Index<12> GE(gw16, p.ge, p.ne, e, e);
            float invd, wk[12];
            __builtin_memcpy(&invd, &rec[7], 4);
            __builtin_memcpy(wk, &rec[8], 48);
            const double grad = (p.B[EC[1]] - p.B[EC[0]]) * double(invd);
            Pair<I> f[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) f[k] = p.Fq[GE[k]];
            C s = 0;
#pragma unroll
            for (int k = 0; k < 12; ++k) s += wk[k] * C(f[k].F) * C(f[gassmann_pv_slot(k)].q);
            du = double(s) - grad;
            rk_store<U, kStreamAll>(r, e, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<U*>(r.su) + e : nullptr);
            return;
        }
        if (e >= p.ne) return;
        const RkIn rin = rk_load(r, e);
        double du;
        const Index<W::cw ? 2 : 0> EC(p.ec16p, p.ec, p.ne, e, e / 3);
        double grad = (p.B[EC[1]] - p.B[EC[0]]) * double(ld(p.invd + e));
        if (!GM) {
            const Index<W::tw> TE(p.te16p, p.te, p.ne, e, e);
            C qe = C(p.Fq[e].q);
            C s = 0;
            for (int k = 0; k < W::kt(p); ++k) {
                Pair<I> f = p.Fq[TE[k]];
                s += ld(p.tw + k * p.ne + e) * C(f.F) * (qe + C(f.q));
            }
            du = double(s / 2) - grad;
        } else if constexpr (GSLOT) {
            static_assert(W::gw == 12 && !GMEAN, "the slot form needs the 12 hexagon-mesh slots");
            const Index<12> GE(p.ge16p, p.ge, p.ne, e, e);
            Pair<I> f[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) f[k] = p.Fq[GE[k]];
            C s = 0;
#pragma unroll
            for (int k = 0; k < 12; ++k) s += ld(p.gw + k * p.ne + e) * C(f[k].F) * C(f[gassmann_pv_slot(k)].q);
            du = double(s) - grad;
        } else {
            const Index<W::gw> GE(p.ge16p, p.ge, p.ne, e, e);
            const Index<GMEAN ? 0 : W::gw> GP(p.gp16p, p.gp, p.ne, e, e);
            C qe = C(p.Fq[e].q);
            C s = 0;
            for (int k = 0; k < W::kg(p); ++k) {
                Pair<I> f = p.Fq[GE[k]];
                C q = GMEAN ? (qe + C(f.q)) / 2 : C(p.Fq[GP[k]].q);
                s += ld(p.gw + k * p.ne + e) * C(f.F) * q;
            }
            du = double(s) - grad;
        }
        rk_store<U, kStreamAll>(r, e, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<U*>(r.su) + e : nullptr);
    } else {
        int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const RkIn rin = rk_load(r, p.ne + i);
        const Index<W::cw> CE(p.ce16p, p.ce, p.nc, i, 3 * i);
        double s = 0.0;
        if constexpr (GF) {
            const uint32_t m = ld(p.csign + i);
            const double area = ld(p.area_c + i);
            for (int k = 0; k < W::kc(p); ++k) {
                const double G = double(p.Fq[CE[k]].F);
                const uint32_t b = (m >> (2 * k)) & 3u;
                s += b == 1u ? G : b == 2u ? -G : 0.0;
            }
            s = s / area;
        } else {
            for (int k = 0; k < W::kc(p); ++k) s += ld(p.div + k * p.nc + i) * double(p.Fq[CE[k]].F);
        }
        rk_store<double, kStreamAll>(r, p.ne + i, -s, rin, r.mode >= 1 && r.mode <= 3 ? r.sh + i : nullptr);
    }
}

// ---------------------------------------------------------------------------------
// Cell-pair path (TRiSK PV flux, cell kinetic energy): three launches per tendency.
// Inside a cell the TRiSK weights are antisymmetric, w_{k',k} = -w_{k,k'} (the kite
// fractions of a cell add up to one), so a cell needs only the 15 weights of its edge
// pairs k < k' (TRiSKGrid.jl:140-155) instead of every edge storing ten coefficients and
// indices.  With G_e = l_e F_e the PV flux of edge e is
//   (1/2 d_e) sum over its two cells c of H_c(e),  H_c(e_k) = sum_k' w_{k,k'} G_k' (q_k + q_k'),
// and the continuity is dh_i = -(sum_k s_k G_k)/A_i (TRiSKShallowWater.jl:275-305).

// W1: vertices (q_v) | edges (G = l_e h_e u_e).  REC: the tables as entity records read with
// vector loads (fewer instructions; faster for the FP32 tables), else slot-major arrays (faster
// for the FP64 ones, whose records span several cache lines).  B is formed by the cell kernel,
// which gathers u at the same edges as G.
// PLANE (rk4_steps' (G, u) plane, one pair per edge): kPlaneNone keeps G and the stage u in
// separate arrays; kPlaneWrite (the first stage of a call) reads u and writes the edges' (G,
// C(u)) pairs; kPlaneRead (later stages, no edge range) gathers the vertices' u from the plane,
// which holds C(u) of the stage input, the value the vertices would convert u to.
enum PairPlane { kPlaneNone = 0, kPlaneWrite = 1, kPlaneRead = 2 };

template <class P, class UIn, bool REC, int PLANE = kPlaneNone>
__global__ __launch_bounds__(kBlock) void pair_state_kernel(Params<P> p, const UIn* __restrict__ u,
                                                            const double* __restrict__ h, int64_t t0, int64_t t1) {
    using C = typename P::C;
    using I = typename P::I;
    Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        int64_t v = t.first + threadIdx.x;
        if (v >= p.nv) return;
        int32_t idx[6];
        C cf[8];
        load_vertex_record<REC>(p, v, idx, cf);
        C uz[3], hz[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (PLANE == kPlaneRead)
                uz[k] = C(p.gu[idx[k]].q);
            else
                uz[k] = C(u[idx[k]]);
            hz[k] = C(h[idx[3 + k]]);
        }
        p.qv[v] = I(pair_vertex_pv(cf, uz, hz));
    } else {
        int64_t e = t.first + threadIdx.x;
        if (e >= p.ne) return;
        int32_t c[2];
        C lh[2];
        if constexpr (REC) {
            float d;
            load_edge_record(p, e, c, d, lh);
        } else {
            load_edge_cells<REC>(p, e, c);
            ldrec(p.re_lh + 2 * e, lh);
        }
        const C ue = C(u[e]);
        const I G = I(pair_mass_flux(lh, C(h[c[0]]), C(h[c[1]]), ue));
        if constexpr (PLANE == kPlaneWrite)
            p.gu[e] = Pair<I>{G, I(ue)};
        else
            p.G[e] = G;
    }
}

// W2: cells: K_i and B_i = K_i + g h_i at the stage (u, h), the PV-flux half sums of their
// edges, and the continuity with its RK4 update (XCD x takes the x-th eighth of the cells, so
// the G, u and q_v its tiles share stay in its L2).  The FP32-bulk kernel stages its block's
// cell records through LDS.  PLANE: G and u gathered as one (G, u) pair per edge.
template <class P, class UIn, bool PLANE = false>
__global__ __launch_bounds__(kBlock) void pair_cell_kernel(Params<P> p, const UIn* __restrict__ u,
                                                           const double* __restrict__ h, Rk r, int64_t t0) {
    using C = typename P::C;
    const int64_t nc = p.nc;
    const int64_t first = tile_of(blockIdx.x, t0, 0).first;
    const int64_t i = first + threadIdx.x;
    // the streaming loads first: they do not wait for the gathers
    const RkIn rin = i < nc ? rk_load(r, p.ne + i) : RkIn{0.0, 0.0};
    const double area = i < nc ? ld(p.area_c + i) : 1.0;
    const double hi = i < nc ? ld(h + i) : 0.0;
    int32_t meta, E[6], V[6];
    C W[15], kin[6];
    if constexpr (P::id == TRISK_FP32_BULK) {
        // the block's cell records (76 B each), staged through LDS with 16-byte coalesced
        // loads: a quarter of the load instructions of the slot-major arrays
        constexpr int RW = kPairRecord;
        __shared__ __attribute__((aligned(16))) int32_t rec[kBlock * RW];
        if (first >= nc) return;  // the whole block
        const int64_t ncell = nc - first < kBlock ? nc - first : kBlock;
        const int nw = int(ncell * RW), nw4 = nw / 4;
        const int32_t* src = p.rc_pair + first * RW;
        for (int w = threadIdx.x; w < nw4; w += kBlock)
            reinterpret_cast<f32x4*>(rec)[w] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src) + w);
        for (int w = 4 * nw4 + threadIdx.x; w < nw; w += kBlock) rec[w] = ld(src + w);
        __syncthreads();
        if (i >= nc) return;
        const int32_t* rr = rec + RW * threadIdx.x;
        meta = rr[0];
        const bool wide = int16_t(rr[1] & 0xffff) == kWide;
        const int32_t be = int32_t(3 * i), bv = int32_t(2 * i);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int32_t pk = rr[1 + k];
            E[k] = wide ? ld(p.ce6 + k * nc + i) : be + int16_t(pk & 0xffff);
            V[k] = wide ? ld(p.ccv + k * nc + i) : bv + int16_t(uint32_t(pk) >> 16);
        }
        C R[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            __builtin_memcpy(&R[k], &rr[7 + k], 4);
            __builtin_memcpy(&kin[k], &rr[13 + k], 4);
        }
        pair_weights(meta, R, W);
    } else {
        if (i >= nc) return;
        load_cell_pairs<false>(p, i, meta, E, V, W);
#pragma unroll
        for (int k = 0; k < 6; ++k) kin[k] = k < p.kc ? ld(p.kin + k * nc + i) : C(0);
    }
    const int n = meta & 15;
    C G[6], Q[6], ue[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if constexpr (PLANE) {
            const Pair<typename P::I> gk = p.gu[E[k]];
            G[k] = C(gk.F);
            ue[k] = C(gk.q);
        } else {
            G[k] = C(p.G[E[k]]);
            ue[k] = C(u[E[k]]);
        }
        Q[k] = C(p.qv[V[k]]);
    }
    p.B[i] = double(pair_kinetic(kin, ue)) + p.g * hi;
    C H[6];
    pair_half_sums(n, W, G, Q, H);
    // the half sums go out cell-major, 6 per cell in one contiguous run (slots past n are
    // never read); the edge kernel gathers each edge's two from its cells
    using I = typename P::I;
    I* out = p.T + 6 * i;
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = k < n ? I(H[k]) : I(0);
    double div = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) {
            const int side = (meta >> (4 + k)) & 1;
            div += side ? -double(G[k]) : double(G[k]);
        }
    }
    rk_store<double, kStreamAcc>(r, p.ne + i, -div / area, rin, r.mode >= 1 && r.mode <= 3 ? r.sh + i : nullptr);
}

// W3: edges: du = (H_left + H_right) / (2 d_e) - (B_r - B_l) / d_e, then the RK4 update.
// GNEXT (RK4 stages): also G_e = l_e h_e u_e of the state this stage writes, from the heights
// the cell kernel has just written and the edge's own new value, so that the next stage's
// state pass skips its edge range.  The operands and operations are those of the state pass.
// PLANE (with GNEXT): the new G and C(u) go to the (G, u) plane as one pair, in place of G and
// the stage vector (modes 1-3; mode 4 still writes y).
template <class P, bool GNEXT, bool PLANE = false>
__global__ __launch_bounds__(kBlock) void pair_edge_kernel(Params<P> p, Rk r, int64_t t0) {
    using C = typename P::C;
    using I = typename P::I;
    using U = typename P::U;
    const int64_t e = tile_of(blockIdx.x, t0, 0).first + threadIdx.x;
    if (e >= p.ne) return;
    // the edge's own streamed values first, so that they do not wait for the cell record
    const RkIn rin = rk_load(r, e);
    const uint32_t sl = ld(p.eslot + e);
    C d;
    C lh[2] = {C(0), C(0)};
    int32_t c[2];
    if constexpr (P::id == TRISK_FP32_BULK) {
        load_edge_record(p, e, c, d, lh);  // one 16-byte load: cells, d, l h
    } else {
        d = ld(p.dual + e);
        if constexpr (GNEXT) ldrec(p.re_lh + 2 * e, lh);
        load_edge_cells<false>(p, e, c);
    }
    // then every gather from the two cells: the half sums (their sum is the same in either
    // order), B and, with GNEXT, the heights of the state this stage writes (the stage
    // heights, or y after the last stage)
    const I hs0 = p.T[6 * int64_t(c[0]) + (sl & 15)], hs1 = p.T[6 * int64_t(c[1]) + (sl >> 4)];
    const double b0 = p.B[c[0]], b1 = p.B[c[1]];
    double hn0 = 0.0, hn1 = 0.0;
    if constexpr (GNEXT) {
        const double* hn = r.mode == 4 ? r.y + p.ne : r.sh;
        hn0 = hn[c[0]];
        hn1 = hn[c[1]];
    }
    const double du = pair_edge_tendency(C(hs0), C(hs1), d, b0, b1);
    if constexpr (PLANE) {
        static_assert(GNEXT, "the (G, u) plane is formed with the next stage's G");
        rk_store<U, kStreamAcc, false>(r, e, du, rin, static_cast<U*>(nullptr));
        const C un = C(rk_next(r, du, rin));
        p.gu[e] = Pair<I>{I(pair_mass_flux(lh, C(hn0), C(hn1), un)), I(un)};
    } else {
        rk_store<U, kStreamAcc>(r, e, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<U*>(r.su) + e : nullptr);
        if constexpr (GNEXT) p.G[e] = I(pair_mass_flux(lh, C(hn0), C(hn1), C(rk_next(r, du, rin))));
    }
}

// ---------------------------------------------------------------------------------
// generic gather y_j = sum_k w[k, j] x[idx[k, j]] (edge_sum_kernel!, :307-314)

__global__ __launch_bounds__(kBlock) void edge_sum_kernel(int64_t n, int K, const int32_t* __restrict__ idx,
                                                          const double* __restrict__ w, const double* __restrict__ x,
                                                          double* __restrict__ out) {
    int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += w[k * n + j] * x[idx[k * n + j]];
    out[j] = s;
}

// ---------------------------------------------------------------------------------
// handle


}  // namespace trisk

using namespace trisk;

namespace trisk {

template <class P>
static Params<P>& params(trisk_handle* h);
template <>
Params<Fp64>& params<Fp64>(trisk_handle* h) {
    return h->p64;
}
template <>
Params<Fp32Bulk>& params<Fp32Bulk>(trisk_handle* h) {
    return h->p32;
}

// The FP32-bulk table kernels run in flux form when the layout allows it; FP64 keeps the
// reference's divergence weights (bit-exact).
template <class P>
static bool flux_form(const trisk_handle* H) {
    return P::id == TRISK_FP32_BULK && H->L.gflux;
}

template <class P>
static int setup_params(trisk_handle* H) {
    using C = typename P::C;
    using I = typename P::I;
    Layout& L = H->L;
    Params<P>& p = params<P>(H);
    std::vector<DevBuf>& o = H->owned;
    p.nc = L.nc;
    p.ne = L.ne;
    p.nv = L.nv;
    p.kc = L.kc;
    p.kv = L.kv;
    p.kt = L.kt;
    p.kg = L.kg;
    p.g = L.gravity;
    auto coef = [&](const std::vector<double>& v, const C** out) -> int {
        if constexpr (sizeof(C) == sizeof(double)) {
            return upload(o, v, reinterpret_cast<const double**>(out));
        } else {
            return upload(o, to_float(v), reinterpret_cast<const float**>(out));
        }
    };
    // flux form (FP32-bulk): the edge heights times l_e, so that the edge kernel forms G = l F,
    // and the momentum weights over l of the edge they multiply
    const bool gf = flux_form<P>(H);
    auto flux = [&](const std::vector<double>& v, const std::vector<int32_t>* idx, const C** out) -> int {
        if constexpr (sizeof(C) == sizeof(float)) {
            std::vector<float> w(v.size());
            const int64_t ne = L.ne;
#pragma omp parallel for schedule(static)
            for (int64_t j = 0; j < int64_t(v.size()); ++j)
                w[j] = float(idx ? v[j] / L.len[(*idx)[j]] : v[j] * L.len[j % ne]);
            return upload(o, w, reinterpret_cast<const float**>(out));
        } else {
            return coef(v, out);
        }
    };
    int rc = 0;
    rc |= upload(o, L.ce, &p.ce);
    if (gf)
        rc |= upload(o, L.csign, &p.csign);
    else
        rc |= upload(o, L.div, &p.div);
    rc |= coef(L.kin, &p.kin);
    rc |= upload(o, L.cv, &p.cv);
    rc |= coef(L.vkk, &p.vkk);
    rc |= upload(o, L.area_c, &p.area_c);
    rc |= upload(o, L.ec, &p.ec);
    rc |= upload(o, L.ev, &p.ev);
    rc |= gf ? flux(L.eh, nullptr, &p.eh) : coef(L.eh, &p.eh);
    rc |= coef(L.invd, &p.invd);
    rc |= upload(o, L.te, &p.te);
    rc |= gf ? flux(L.tw, &L.te, &p.tw) : coef(L.tw, &p.tw);
    rc |= gf ? flux(L.vkh, nullptr, &p.vkh) : coef(L.vkh, &p.vkh);
    rc |= upload(o, L.ge, &p.ge);
    rc |= upload(o, L.gp, &p.gp);
    rc |= gf ? flux(L.gw, &L.ge, &p.gw) : coef(L.gw, &p.gw);
    p.r1c = p.r1v = p.r2e = p.r2c = nullptr;
    if (P::id == TRISK_FP32_BULK && !L.pair && L.kc == 6 && L.kv == 3 && L.kt == 10 && (L.kg == 0 || L.kg == 12)) {
        // the P1 and P2 records of the FP32-bulk table kernels (Params::r1c, r1v, r2e, r2c)
        const int64_t nc = L.nc, ne = L.ne, nv = L.nv;
        const bool vk = L.vertex_ke, gm = L.gassmann;
        auto bits = [](double x) {
            const float f = float(x);
            int32_t w;
            std::memcpy(&w, &f, 4);
            return w;
        };
        std::vector<int32_t> rc1(size_t(9) * nc), rc2(vk ? size_t(9) * nc : 0);
        for (int64_t i = 0; i < nc; ++i) {
            for (int q = 0; q < 3; ++q) rc1[9 * i + q] = L.ce16p[size_t(q) * nc + i];
            for (int k = 0; k < 6; ++k) rc1[9 * i + 3 + k] = bits(L.kin[size_t(k) * nc + i]);
            if (!vk) continue;
            for (int q = 0; q < 3; ++q) rc2[9 * i + q] = L.cv16p[size_t(q) * nc + i];
            for (int k = 0; k < 6; ++k) rc2[9 * i + 3 + k] = bits(L.vkk[size_t(k) * nc + i]);
        }
        // the vertex records of P1 (QV: with f, which the Gassmann flux does not read)
        const int RF = gm ? 0 : 1, RV = 10 + RF + (vk ? 3 : 0);
        std::vector<int32_t> rv(size_t(RV) * nv);
        for (int64_t v = 0; v < nv; ++v) {
            int32_t* r = &rv[size_t(RV) * v];
            for (int q = 0; q < 2; ++q) {
                r[q] = L.ve16p[size_t(q) * nv + v];
                r[2 + q] = L.vc16p[size_t(q) * nv + v];
            }
            // A_v zeta_v and A_v h_v (the edge kernel's records take the areas out again)
            const double av = L.area_v[v];
            for (int k = 0; k < 3; ++k) {
                r[4 + k] = bits(av * L.curl[size_t(k) * nv + v]);
                r[7 + k] = bits(av * L.vh[size_t(k) * nv + v]);
                if (vk) r[10 + RF + k] = bits(L.vkw[size_t(k) * nv + v]);
            }
            if (RF) r[10] = bits(av * L.fv[v]);
        }
        const int RE = 4 + (vk ? 2 : 0) + (gm ? 1 : 0);
        std::vector<int32_t> re(size_t(RE) * ne);
        for (int64_t e = 0; e < ne; ++e) {
            int32_t* r = &re[size_t(RE) * e];
            const double l = gf ? L.len[e] : 1.0;  // flux form: G = l F
            const double a0 = L.area_v[L.ev[e]], a1 = L.area_v[L.ev[ne + e]];
            r[0] = L.ec16p[e];
            r[1] = L.ev16p[e];
            r[2] = bits(L.eh[e] * l);
            r[3] = bits(L.eh[ne + e] * l);
            if (vk) {
                r[4] = bits(L.vkh[e] * l / a0);
                r[5] = bits(L.vkh[ne + e] * l / a1);
            }
            if (gm) r[RE - 1] = bits((a0 + a1) * L.gf[e]);
        }
        rc |= upload(o, rc1, &p.r1c);
        rc |= upload(o, rv, &p.r1v);
        rc |= upload(o, re, &p.r2e);
        if (vk) rc |= upload(o, rc2, &p.r2c);
    }
    p.rg = nullptr;
    if (P::id == TRISK_FP32_BULK && L.gassmann_slots && L.kc == 6 && L.kv == 3 && L.kt == 10 && L.kg == 12) {
        // the Gassmann edge records of the slot form (Params::rg)
        const int64_t ne = L.ne;
        std::vector<int32_t> rec(size_t(20) * ne);
        for (int64_t e = 0; e < ne; ++e) {
            int32_t* r = &rec[size_t(20) * e];
            for (int q = 0; q < 6; ++q) r[q] = L.ge16p[size_t(q) * ne + e];
            r[6] = L.ec16p[e];
            const float invd = float(L.invd[e]);
            std::memcpy(&r[7], &invd, 4);
            for (int k = 0; k < 12; ++k) {
                const size_t j = size_t(k) * ne + e;
                const float w = float(gf ? L.gw[j] / L.len[L.ge[j]] : L.gw[j]);
                std::memcpy(&r[8 + k], &w, 4);
            }
        }
        rc |= upload(o, rec, &p.rg);
    }
    rc |= coef(L.grh, &p.grh);
    rc |= coef(L.gf, &p.gf);
    rc |= upload(o, L.ve, &p.ve);
    rc |= coef(L.curl, &p.curl);
    rc |= upload(o, L.vc, &p.vc);
    rc |= coef(L.vh, &p.vh);
    rc |= coef(L.fv, &p.fv);
    rc |= coef(L.vkw, &p.vkw);
    rc |= upload(o, L.area_v, &p.area_v);
    rc |= upload(o, L.ce16p, &p.ce16p);
    rc |= upload(o, L.cv16p, &p.cv16p);
    rc |= upload(o, L.ec16p, &p.ec16p);
    rc |= upload(o, L.ev16p, &p.ev16p);
    rc |= upload(o, L.te16p, &p.te16p);
    rc |= upload(o, L.ge16p, &p.ge16p);
    rc |= upload(o, L.gp16p, &p.gp16p);
    rc |= upload(o, L.ve16p, &p.ve16p);
    rc |= upload(o, L.vc16p, &p.vc16p);
    if (L.pair) {
        rc |= upload(o, L.cmeta, &p.cmeta);
        rc |= coef(L.dual, &p.dual);
        if (H->precision == TRISK_FP32_BULK) {
            rc |= upload(o, L.rc_state, &p.rc_state);
            rc |= coef(L.rc_kin, &p.rc_kin);
            rc |= upload(o, L.rv_idx, &p.rv_idx);
            rc |= coef(L.rv_coef, &p.rv_coef);
            rc |= upload(o, L.rc_state16, &p.rc_state16);
            rc |= upload(o, L.rv_idx16, &p.rv_idx16);
            rc |= upload(o, L.re_cells16, &p.re_cells16);
            {
                std::vector<int32_t> rec(size_t(L.ne) * 4);
                for (int64_t e = 0; e < L.ne; ++e) {
                    const float w[3] = {float(L.dual[e]), float(L.re_lh[2 * e]), float(L.re_lh[2 * e + 1])};
                    rec[4 * e] = L.re_cells16[e];
                    std::memcpy(&rec[4 * e + 1], w, sizeof(w));
                }
                rc |= upload(o, rec, &p.re_rec);
            }
            rc |= upload(o, L.pc_ev16, &p.pc_ev16);
            rc |= upload(o, L.rc_pair, &p.rc_pair);
        }
        rc |= upload(o, L.re_cells, &p.re_cells);
        rc |= upload(o, L.eslot, &p.eslot);
        rc |= coef(L.re_lh, &p.re_lh);
        rc |= upload(o, L.rc_pce, &p.ce6);
        rc |= upload(o, L.ccv, &p.ccv);
        rc |= coef(L.pr, &p.pr);
    }
    if (rc) return rc;
    void* q;
    if ((rc = alloc(o, L.nc * sizeof(double), &q))) return rc;
    p.B = static_cast<double*>(q);
    if ((rc = alloc(o, L.nc * sizeof(double), &q))) return rc;
    p.K = static_cast<double*>(q);
    if ((rc = alloc(o, L.ne * sizeof(Pair<I>), &q))) return rc;
    p.Fq = static_cast<Pair<I>*>(q);
    if ((rc = alloc(o, L.ne * sizeof(I), &q))) return rc;
    p.G = static_cast<I*>(q);
    p.gu = nullptr;
    if (L.pair) {
        if ((rc = alloc(o, L.ne * sizeof(Pair<I>), &q))) return rc;
        p.gu = static_cast<Pair<I>*>(q);
    }
    if ((rc = alloc(o, 6 * L.nc * sizeof(I), &q))) return rc;
    p.T = static_cast<I*>(q);
    for (I** w : {&p.qv, &p.zv, &p.hv, &p.Kv}) {
        if ((rc = alloc(o, L.nv * sizeof(I), &q))) return rc;
        *w = static_cast<I*>(q);
    }
    p.K_out = p.qv_out = p.zv_out = p.hv_out = nullptr;
    // RK4 work
    int64_t n = L.ne + L.nc;
    if ((rc = alloc(o, n * sizeof(double), &q))) return rc;
    H->acc = static_cast<double*>(q);
    if ((rc = alloc(o, L.ne * sizeof(typename P::U), &q))) return rc;
    H->su = q;
    if ((rc = alloc(o, L.nc * sizeof(double), &q))) return rc;
    H->sh = static_cast<double*>(q);
    if ((rc = alloc(o, 2 * H->reduce_blocks * sizeof(double), &q))) return rc;
    H->partial = static_cast<double*>(q);
    for (double** w : {&H->diag_K, &H->diag_qv, &H->diag_zv, &H->diag_hv}) {
        size_t cnt = w == &H->diag_K ? L.nc : L.nv;
        if ((rc = alloc(o, cnt * sizeof(double), &q))) return rc;
        *w = static_cast<double*>(q);
    }
    return 0;
}


// One tendency: P1, P2, P3 with the given stage input and update mode.  plane (RK4 stages with
// gnext, the stage input written by the previous edge kernel or, with state_edges, the state):
// G and u travel in the (G, u) plane.
template <class P, class UIn>
static int run_pair(trisk_handle* H, const UIn* u, const double* h, const Rk& r, hipStream_t s, bool state_edges,
                    bool gnext, bool plane) {
    const Params<P>& p = params<P>(H);
    const Layout& L = H->L;
    hipEvent_t stop;
    int rc;
    {
        // state_edges false: G is already there, formed by the previous stage's edge kernel
        int64_t t0 = tiles_per_xcd(L.nv), t1 = state_edges ? tiles_per_xcd(L.ne) : 0;
        if ((rc = timed_begin(H, K_PAIR_STATE, s, &stop))) return rc;
        constexpr bool rec = P::id == TRISK_FP32_BULK;
        const dim3 grid(unsigned(kXcd * (t0 + t1)));
        if (!plane)
            hipLaunchKernelGGL((pair_state_kernel<P, UIn, rec>), grid, dim3(kBlock), 0, s, p, u, h, t0, t1);
        else if (state_edges)
            hipLaunchKernelGGL((pair_state_kernel<P, UIn, rec, kPlaneWrite>), grid, dim3(kBlock), 0, s, p, u, h, t0,
                               t1);
        else
            hipLaunchKernelGGL((pair_state_kernel<P, UIn, rec, kPlaneRead>), grid, dim3(kBlock), 0, s, p, u, h, t0,
                               t1);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    {
        if ((rc = timed_begin(H, K_PAIR_CELL, s, &stop))) return rc;
        const int64_t t0 = tiles_per_xcd(L.nc);
        if (plane)
            hipLaunchKernelGGL((pair_cell_kernel<P, UIn, true>), dim3(unsigned(kXcd * t0)), dim3(kBlock), 0, s, p, u, h,
                               r, t0);
        else
            hipLaunchKernelGGL((pair_cell_kernel<P, UIn>), dim3(unsigned(kXcd * t0)), dim3(kBlock), 0, s, p, u, h, r,
                               t0);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    {
        if ((rc = timed_begin(H, K_PAIR_EDGE, s, &stop))) return rc;
        const int64_t t0 = tiles_per_xcd(L.ne);
        if (gnext && plane)
            hipLaunchKernelGGL((pair_edge_kernel<P, true, true>), dim3(unsigned(kXcd * t0)), dim3(kBlock), 0, s, p, r,
                               t0);
        else if (gnext)
            hipLaunchKernelGGL((pair_edge_kernel<P, true>), dim3(unsigned(kXcd * t0)), dim3(kBlock), 0, s, p, r, t0);
        else
            hipLaunchKernelGGL((pair_edge_kernel<P, false>), dim3(unsigned(kXcd * t0)), dim3(kBlock), 0, s, p, r, t0);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    return 0;
}

template <class P, class W, bool GF>
static void launch_update(trisk_handle* H, dim3 grid, hipStream_t s, const Rk& r, int64_t t0, int64_t t1) {
    const Params<P>& p = params<P>(H);
    const Layout& L = H->L;
    if (!L.gassmann)
        hipLaunchKernelGGL((update_kernel<P, false, false, W, false, GF>), grid, dim3(kBlock), 0, s, p, r, t0, t1);
    else if (L.gassmann_mean)
        hipLaunchKernelGGL((update_kernel<P, true, true, W, false, GF>), grid, dim3(kBlock), 0, s, p, r, t0, t1);
    else if constexpr (W::gw == 12) {
        if (p.rg)
            hipLaunchKernelGGL((update_kernel<P, true, false, W, true, GF, P::id == TRISK_FP32_BULK>), grid, dim3(kBlock),
                               0, s, p, r, t0, t1);
        else if (L.gassmann_slots)
            hipLaunchKernelGGL((update_kernel<P, true, false, W, true, GF>), grid, dim3(kBlock), 0, s, p, r, t0, t1);
        else
            hipLaunchKernelGGL((update_kernel<P, true, false, W, false, GF>), grid, dim3(kBlock), 0, s, p, r, t0, t1);
    } else
        hipLaunchKernelGGL((update_kernel<P, true, false, W, false, GF>), grid, dim3(kBlock), 0, s, p, r, t0, t1);
}

template <class P, class UIn, class W, bool REC>
static void launch_p1(const Params<P>& p, dim3 grid, hipStream_t s, const UIn* u, const double* h, int64_t t0,
                      int64_t t1, bool vk, bool gm, bool zh) {
    if (vk && gm)
        hipLaunchKernelGGL((cell_vertex_kernel<P, UIn, double, true, true, W, false, REC>), grid, dim3(kBlock), 0, s, p,
                           u, h, t0, t1);
    else if (vk)
        hipLaunchKernelGGL((cell_vertex_kernel<P, UIn, double, true, true, W, true, REC>), grid, dim3(kBlock), 0, s, p, u,
                           h, t0, t1);
    else if (gm)
        hipLaunchKernelGGL((cell_vertex_kernel<P, UIn, double, false, true, W, false, REC>), grid, dim3(kBlock), 0, s, p,
                           u, h, t0, t1);
    else if (zh)
        hipLaunchKernelGGL((cell_vertex_kernel<P, UIn, double, false, true, W, true, REC>), grid, dim3(kBlock), 0, s, p,
                           u, h, t0, t1);
    else
        hipLaunchKernelGGL((cell_vertex_kernel<P, UIn, double, false, false, W, true, REC>), grid, dim3(kBlock), 0, s, p,
                           u, h, t0, t1);
}

template <class P, class UIn, class W, bool REC>
static void launch_p2(const Params<P>& p, dim3 grid, hipStream_t s, const UIn* u, const double* h, int64_t t0,
                      int64_t t1, bool vk, bool gm, bool bh = false) {
    if (vk && gm && bh)
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, true, true, W, REC, true>), grid, dim3(kBlock), 0, s, p, u, h,
                           t0, t1);
    else if (vk && gm)
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, true, true, W, REC>), grid, dim3(kBlock), 0, s, p, u, h, t0, t1);
    else if (vk && bh)
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, true, false, W, REC, true>), grid, dim3(kBlock), 0, s, p, u, h,
                           t0, t1);
    else if (vk)
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, true, false, W, REC>), grid, dim3(kBlock), 0, s, p, u, h, t0, t1);
    else if (gm)
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, false, true, W, REC>), grid, dim3(kBlock), 0, s, p, u, h, t0, t1);
    else
        hipLaunchKernelGGL((edge_kernel<P, UIn, double, false, false, W, REC>), grid, dim3(kBlock), 0, s, p, u, h, t0,
                           t1);
}

template <class P, class UIn, class W>
static int run_widths(trisk_handle* H, const UIn* u, const double* h, const Rk& r, hipStream_t s) {
    const Params<P>& p = params<P>(H);
    const Layout& L = H->L;
    const bool vk = L.vertex_ke, gm = L.gassmann;
    const bool zh = gm || vk || p.qv_out != nullptr;
    hipEvent_t stop;
    int rc;
    // records (FP32-bulk, hexagon widths) when the handle holds them
    constexpr bool can_rec = P::id == TRISK_FP32_BULK && W::cw == 6 && W::vw == 3;
    const bool rec = can_rec && p.r1c != nullptr;
    // vertex kinetic energy with every cell weight 0: P1 covers the vertices only and P2 forms B
    // (not when the diagnostics ask for K)
    const bool bh = vk && L.cell_ke_zero && p.K_out == nullptr;
    // P1
    {
        int64_t t0 = bh ? 0 : tiles_per_xcd(L.nc), t1 = tiles_per_xcd(L.nv);
        dim3 grid(unsigned(kXcd * (t0 + t1)));
        if ((rc = timed_begin(H, K_CELL_VERTEX, s, &stop))) return rc;
        if (rec)
            launch_p1<P, UIn, W, can_rec>(p, grid, s, u, h, t0, t1, vk, gm, zh);
        else
            launch_p1<P, UIn, W, false>(p, grid, s, u, h, t0, t1, vk, gm, zh);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    // P2
    {
        int64_t t0 = tiles_per_xcd(L.ne), t1 = vk ? tiles_per_xcd(L.nc) : 0;
        dim3 grid(unsigned(kXcd * (t0 + t1)));
        if ((rc = timed_begin(H, K_EDGE, s, &stop))) return rc;
        if (rec)
            launch_p2<P, UIn, W, can_rec>(p, grid, s, u, h, t0, t1, vk, gm, bh);
        else
            launch_p2<P, UIn, W, false>(p, grid, s, u, h, t0, t1, vk, gm, bh);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    // P3
    {
        int64_t t0 = tiles_per_xcd(L.ne), t1 = tiles_per_xcd(L.nc);
        dim3 grid(unsigned(kXcd * (t0 + t1)));
        if ((rc = timed_begin(H, K_UPDATE, s, &stop))) return rc;
        if (flux_form<P>(H))
            launch_update<P, W, true>(H, grid, s, r, t0, t1);
        else
            launch_update<P, W, false>(H, grid, s, r, t0, t1);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }
    return 0;
}

template <class P>
static bool pair_run(trisk_handle* H) {
    return H->L.pair && params<P>(H).qv_out == nullptr;
}

// state_edges, gnext: the cell-pair path's fusion of G into the previous stage's edge kernel
template <class P, class UIn>
static int run_tendency(trisk_handle* H, const UIn* u, const double* h, const Rk& r, hipStream_t s,
                        bool state_edges = true, bool gnext = false, bool plane = false) {
    const Layout& L = H->L;
    if (pair_run<P>(H)) return run_pair<P, UIn>(H, u, h, r, s, state_edges, gnext, plane);
    const bool hex = L.kc == 6 && L.kv == 3 && L.kt == 10 && (L.kg == 0 || L.kg == 12);
    return hex ? run_widths<P, UIn, HexWidths>(H, u, h, r, s) : run_widths<P, UIn, AnyWidths>(H, u, h, r, s);
}

template <class P>
static int rk4_steps(trisk_handle* H, double* y, double dt, int64_t nsteps, hipStream_t s) {
    using U = typename P::U;
    const Layout& L = H->L;
    Rk r{};
    r.y = y;
    r.acc = H->acc;
    r.su = H->su;
    r.sh = H->sh;
    r.dt = dt;
    r.dt2 = dt / 2;
    r.dt6 = dt / 6;
    int rc;
    // cell-pair path: each stage's edge kernel also forms G of the state it writes, so only the
    // first stage of the call runs the state pass over the edges; with the (G, u) plane that
    // edge kernel writes G and the stage u as one pair (no stage vector), which the next
    // stage's state and cell kernels gather
    const bool fuse = pair_run<P>(H);
    const bool plane = fuse && H->pair_plane;
    for (int64_t step = 0; step < nsteps; ++step) {
        r.mode = 1;
        if ((rc = run_tendency<P, double>(H, y, y + L.ne, r, s, !fuse || step == 0, fuse, plane))) return rc;
        for (int mode = 2; mode <= 4; ++mode) {
            r.mode = mode;
            const bool last = mode == 4 && step + 1 == nsteps;
            if ((rc = run_tendency<P, U>(H, static_cast<const U*>(H->su), H->sh, r, s, !fuse, fuse && !last, plane)))
                return rc;
        }
    }
    return 0;
}

template <class P>
static int tendency(trisk_handle* H, const double* y, double* dy, hipStream_t s) {
    Rk r{};
    r.mode = 0;
    r.dy = dy;
    return run_tendency<P, double>(H, y, y + H->L.ne, r, s);
}

// P1 (and with vertex kinetic energy the cell part of P2) with the FP64 diagnostic outputs
template <class P>
static int diagnostic_pass(trisk_handle* H, const double* y, double* K_out, double* qv_out, double* zv_out,
                           double* hv_out, hipStream_t s) {
    Params<P>& p = params<P>(H);
    const Layout& L = H->L;
    p.K_out = K_out;
    p.qv_out = qv_out;
    p.zv_out = zv_out;
    p.hv_out = hv_out;
    const double* u = y;
    const double* h = y + L.ne;
    int64_t t0 = tiles_per_xcd(L.nc), t1 = tiles_per_xcd(L.nv);
    if (L.vertex_ke)
        hipLaunchKernelGGL((cell_vertex_kernel<P, double, double, true, true>), dim3(unsigned(kXcd * (t0 + t1))),
                           dim3(kBlock), 0, s, p, u, h, t0, t1);
    else
        hipLaunchKernelGGL((cell_vertex_kernel<P, double, double, false, true>), dim3(unsigned(kXcd * (t0 + t1))),
                           dim3(kBlock), 0, s, p, u, h, t0, t1);
    hipError_t e1 = hipGetLastError();
    if (e1 == hipSuccess && L.vertex_ke) {
        int64_t e0 = tiles_per_xcd(L.ne), c1 = tiles_per_xcd(L.nc);
        if (L.gassmann)
            hipLaunchKernelGGL((edge_kernel<P, double, double, true, true>), dim3(unsigned(kXcd * (e0 + c1))),
                               dim3(kBlock), 0, s, p, u, h, e0, c1);
        else
            hipLaunchKernelGGL((edge_kernel<P, double, double, true, false>), dim3(unsigned(kXcd * (e0 + c1))),
                               dim3(kBlock), 0, s, p, u, h, e0, c1);
        e1 = hipGetLastError();
    }
    p.K_out = p.qv_out = p.zv_out = p.hv_out = nullptr;
    if (e1 != hipSuccess) return set_err(TRISK_EHIP, std::string("diagnostic pass: ") + hipGetErrorString(e1));
    return 0;
}


// ---------------------------------------------------------------------------------
// Static scratch of trisk_diagnostics_noalloc: the block partials (sum, compensation) of
// the three diagnostic sums.  Every GPU the library runs on holds one copy until the
// process ends; calls that overlap on one GPU share it and race.

constexpr int kNoallocTerms = 3;
constexpr int kNoallocMaxBlocks = 1024;
constexpr int kNoallocDoubles = kNoallocTerms * 2 * kNoallocMaxBlocks;  // 48 KiB
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static bool g_scratch_done[kMaxDevices];
static double* g_scratch_ptr[kMaxDevices];     // the device address of `scratch` on each GPU
static hipError_t g_scratch_err[kMaxDevices];  // the result of its lookup

// Looks up the address of `scratch` on `device`, which must be current: once per device,
// at handle creation.  A failed lookup is kept and reported by trisk_diagnostics_noalloc.
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_done[device]) return;
    void* ptr = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? static_cast<double*>(ptr) : nullptr;
    g_scratch_done[device] = true;
}

static double* noalloc_scratch(int device, hipError_t* err) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kMaxDevices || !g_scratch_done[device]) {
        *err = hipErrorNotInitialized;
        return nullptr;
    }
    *err = g_scratch_err[device];
    return g_scratch_ptr[device];
}

// The host combination of reduce() on the device: a fixed-order Neumaier sum of the block
// partials of each term, bitwise the same as trisk_diagnostics' (FP contraction is off).
static __global__ __launch_bounds__(64) void noalloc_combine_kernel(int blocks, const double* __restrict__ part,
                                                                  double* __restrict__ out) {
    const int term = threadIdx.x;
    if (term >= kNoallocTerms) return;
    const double* p = part + size_t(term) * 2 * blocks;
    double S = 0.0, Cc = 0.0;
    for (int b = 0; b < blocks; ++b) {
        double x = p[2 * b], t = S + x;
        if (fabs(S) >= fabs(x))
            Cc += (S - t) + x;
        else
            Cc += (x - t) + S;
        S = t;
        Cc += p[2 * b + 1];
    }
    out[term] = term == 2 ? (S + Cc) / 2 : S + Cc;
}

// Non-finite entries of u and of h in the state y, and vertices whose height h_v (of the
// last diagnostic pass, in diag_hv) is not > 0 or NaN: the divisor of q_v.  Synchronous.
int state_counts(trisk_handle* H, const double* y, int64_t counts[3], hipStream_t s) {
    const Layout& L = H->L;
    double c[3];
    int rc;
    if ((rc = reduce(H, 3, L.ne, nullptr, y, nullptr, &c[0], s))) return rc;
    if ((rc = reduce(H, 3, L.nc, nullptr, y + L.ne, nullptr, &c[1], s))) return rc;
    if ((rc = reduce(H, 4, L.nv, nullptr, H->diag_hv, nullptr, &c[2], s))) return rc;
    for (int k = 0; k < 3; ++k) counts[k] = int64_t(c[k]);
    return 0;
}

// TRISK_ENONFINITE with the counts in the message when the state y (whose diagnostic pass
// has just run) is not finite or has a vertex height that is not positive.
int finite_status(trisk_handle* H, const double* y, hipStream_t s) {
    int64_t c[3];
    int rc = state_counts(H, y, c, s);
    if (rc) return rc;
    if (c[0] || c[1] || c[2])
        return set_err(TRISK_ENONFINITE, "state check failed: " + std::to_string(c[0]) + " non-finite u, " +
                                             std::to_string(c[1]) + " non-finite h, " + std::to_string(c[2]) +
                                             " vertices with h_v <= 0 or NaN");
    return 0;
}

}  // namespace trisk

#define CHECK_HANDLE(h)             \
    ErrorScope error_scope_((h)); \
    if (!(h)) return set_err(TRISK_EINVAL, "null handle")
#define CHECK_TRISK(h) \
    if ((h)->L.icon) return set_err(TRISK_EINVAL, "not available for the ICON triangle model")

extern "C" {

const char* trisk_last_error(void) { return g_error.c_str(); }

const char* trisk_handle_last_error(const trisk_handle* h) { return h ? h->last_error.c_str() : "null handle"; }

static void drop_graph(trisk_handle* h);

int trisk_check_tables(const trisk_model_desc* desc, int32_t* recomputed) {
    Layout L;
    std::string err;
    if (!build_layout(desc, L, err)) return set_err(TRISK_EINVAL, err);
    build_pairs(desc, L);
    if (recomputed) *recomputed = L.pair ? 1 : 0;
    g_error = L.pair_reason;
    return 0;
}

int trisk_create(const trisk_model_desc* desc, const trisk_options* opt, trisk_handle** out) {
    if (!out) return set_err(TRISK_EINVAL, "null output pointer");
    *out = nullptr;
    trisk_options o{};
    if (opt) o = *opt;
    if (o.precision != TRISK_FP64 && o.precision != TRISK_FP32_BULK)
        return set_err(TRISK_EINVAL, "precision must be TRISK_FP64 or TRISK_FP32_BULK");
    if (o.block_size != 0 && o.block_size != kBlock)
        return set_err(TRISK_EINVAL, "block_size must be 0 or the library's workgroup size");
    auto* H = new trisk_handle;
    H->precision = o.precision;
    // the (G, u) plane: 8-byte FP32 pairs measured faster, 16-byte FP64 pairs slower (they
    // double the lines the vertex kernel gathers u from)
    H->pair_plane = o.precision == TRISK_FP32_BULK;
    H->device = o.device;
    std::string err;
    if (!build_layout(desc, H->L, err)) {
        delete H;
        return set_err(TRISK_EINVAL, err);
    }
    if (o.path < 0 || o.path > 2) {
        delete H;
        return set_err(TRISK_EINVAL, "path must be TRISK_PATH_AUTO, TRISK_PATH_TABLES or TRISK_PATH_PAIR");
    }
    const bool want_pair = o.path == TRISK_PATH_PAIR || (o.path == TRISK_PATH_AUTO && o.precision == TRISK_FP32_BULK);
    if (want_pair) {
        build_pairs(desc, H->L, o.narrow_limit);
        if (!H->L.pair && o.path == TRISK_PATH_PAIR) {
            std::string why = H->L.pair_reason;
            delete H;
            return set_err(TRISK_EINVAL, "cell-pair path unavailable: " + why);
        }
    } else {
        H->L.pair_reason = o.precision == TRISK_FP64 ? "FP64: coefficient tables, bit-exact with the reference"
                                                     : "coefficient tables requested";
    }
    if (o.precision == TRISK_FP32_BULK && !H->L.pair) build_flux_form(desc, H->L);
    int ndev = 0;
    hipError_t he = hipGetDeviceCount(&ndev);
    if (he != hipSuccess || o.device < 0 || o.device >= ndev || hipSetDevice(o.device) != hipSuccess) {
        delete H;
        return set_err(TRISK_EHIP, "no such HIP device");
    }
    noalloc_setup(o.device);
    int rc = o.precision == TRISK_FP64 ? setup_params<Fp64>(H) : setup_params<Fp32Bulk>(H);
    if (rc) {
        for (auto& b : H->owned) (void)hipFree(b.p);
        delete H;
        return rc;
    }
    *out = H;
    return 0;
}

int trisk_destroy(trisk_handle* h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    drop_graph(h);
    if (h->adj_release) h->adj_release(h);
    if (h->icon_release) h->icon_release(h);
    for (auto& e : h->events) (void)hipEventDestroy(e);
    for (auto& b : h->owned) (void)hipFree(b.p);
    delete h;
    return 0;
}

int trisk_state_length(const trisk_handle* h, int64_t* n) {
    CHECK_HANDLE(h);
    if (!n) return set_err(TRISK_EINVAL, "null output pointer");
    *n = h->L.ne + h->L.nc;
    return 0;
}

int trisk_permutation(const trisk_handle* h, int64_t* cells, int64_t* edges, int64_t* vertices) {
    CHECK_HANDLE(h);
    if (cells) std::memcpy(cells, h->L.cperm.data(), h->L.nc * sizeof(int64_t));
    if (edges) std::memcpy(edges, h->L.eperm.data(), h->L.ne * sizeof(int64_t));
    if (vertices) std::memcpy(vertices, h->L.vperm.data(), h->L.nv * sizeof(int64_t));
    return 0;
}

static const std::vector<int64_t>* perm_of(const trisk_handle* h, int kind) {
    return kind == 0 ? &h->L.cperm : kind == 1 ? &h->L.eperm : kind == 2 ? &h->L.vperm : nullptr;
}

int trisk_upload_field(trisk_handle* h, int32_t kind, const double* x_host, double* x_dev, void* stream) {
    CHECK_HANDLE(h);
    const std::vector<int64_t>* perm = perm_of(h, kind);
    if (!perm || !x_host || !x_dev) return set_err(TRISK_EINVAL, "bad field kind or pointer");
    HIP_TRY(hipSetDevice(h->device));
    std::vector<double> tmp(perm->size());
    for (size_t j = 0; j < tmp.size(); ++j) tmp[j] = x_host[(*perm)[j]];
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(x_dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int trisk_download_field(trisk_handle* h, int32_t kind, const double* x_dev, double* x_host, void* stream) {
    CHECK_HANDLE(h);
    const std::vector<int64_t>* perm = perm_of(h, kind);
    if (!perm || !x_host || !x_dev) return set_err(TRISK_EINVAL, "bad field kind or pointer");
    HIP_TRY(hipSetDevice(h->device));
    std::vector<double> tmp(perm->size());
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(tmp.data(), x_dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t j = 0; j < tmp.size(); ++j) x_host[(*perm)[j]] = tmp[j];
    return 0;
}

int trisk_upload_state(trisk_handle* h, const double* y_host, double* y_dev, void* stream) {
    CHECK_HANDLE(h);
    int rc = trisk_upload_field(h, 1, y_host, y_dev, stream);
    if (rc) return rc;
    return trisk_upload_field(h, 0, y_host + h->L.ne, y_dev + h->L.ne, stream);
}

int trisk_download_state(trisk_handle* h, const double* y_dev, double* y_host, void* stream) {
    CHECK_HANDLE(h);
    int rc = trisk_download_field(h, 1, y_dev, y_host, stream);
    if (rc) return rc;
    return trisk_download_field(h, 0, y_dev + h->L.ne, y_host + h->L.ne, stream);
}

int trisk_tendency(trisk_handle* h, const double* y_dev, double* dy_dev, void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev || !dy_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->L.icon) return icon_tendency(h, y_dev, dy_dev, s);
    return h->precision == TRISK_FP64 ? tendency<Fp64>(h, y_dev, dy_dev, s) : tendency<Fp32Bulk>(h, y_dev, dy_dev, s);
}

static int run_steps(trisk_handle* h, double* y, double dt, int64_t n, hipStream_t s) {
    if (h->L.icon) return icon_run_steps(h, y, dt, n, s);
    return h->precision == TRISK_FP64 ? rk4_steps<Fp64>(h, y, dt, n, s) : rk4_steps<Fp32Bulk>(h, y, dt, n, s);
}

}  // extern "C"

namespace trisk {

template <class P, class UIn>
static int state_pass(trisk_handle* H, const UIn* u, const double* h, hipStream_t s) {
    const Layout& L = H->L;
    const int64_t t0 = tiles_per_xcd(L.nv), t1 = tiles_per_xcd(L.ne);
    constexpr bool rec = P::id == TRISK_FP32_BULK;
    hipLaunchKernelGGL((pair_state_kernel<P, UIn, rec>), dim3(unsigned(kXcd * (t0 + t1))), dim3(kBlock), 0, s,
                       params<P>(H), u, h, t0, t1);
    HIP_TRY(hipGetLastError());
    return 0;
}

int pair_state_pass(trisk_handle* H, const void* u, bool u_double, const double* h, hipStream_t s) {
    if (H->precision == TRISK_FP64) return state_pass<Fp64, double>(H, static_cast<const double*>(u), h, s);
    if (u_double) return state_pass<Fp32Bulk, double>(H, static_cast<const double*>(u), h, s);
    return state_pass<Fp32Bulk, float>(H, static_cast<const float*>(u), h, s);
}

int tendency_stage(trisk_handle* H, const void* u, bool u_double, const double* h, const Rk& r, hipStream_t s) {
    if (H->precision == TRISK_FP64)
        return run_tendency<Fp64, double>(H, static_cast<const double*>(u), h, r, s, true, false);
    if (u_double) return run_tendency<Fp32Bulk, double>(H, static_cast<const double*>(u), h, r, s, true, false);
    return run_tendency<Fp32Bulk, float>(H, static_cast<const float*>(u), h, r, s, true, false);
}

// P1 and P2 of the table kernels at (u, h): B, the vertex quantities (zeta_v, h_v, q_v, K_v
// as the options need them) and the (F, q_e) pairs
template <class P, class UIn>
static int table_state(trisk_handle* H, const UIn* u, const double* h, hipStream_t s) {
    const Params<P>& p = params<P>(H);
    const Layout& L = H->L;
    const bool vk = L.vertex_ke, gm = L.gassmann, zh = true;
    const bool hex = L.kc == 6 && L.kv == 3 && L.kt == 10 && (L.kg == 0 || L.kg == 12);
    const int64_t c0 = tiles_per_xcd(L.nc), v1 = tiles_per_xcd(L.nv), e0 = tiles_per_xcd(L.ne),
                  c1 = vk ? tiles_per_xcd(L.nc) : 0;
    const dim3 g1(unsigned(kXcd * (c0 + v1))), g2(unsigned(kXcd * (e0 + c1)));
    if (hex) {
        launch_p1<P, UIn, HexWidths, false>(p, g1, s, u, h, c0, v1, vk, gm, zh);
        launch_p2<P, UIn, HexWidths, false>(p, g2, s, u, h, e0, c1, vk, gm);
    } else {
        launch_p1<P, UIn, AnyWidths, false>(p, g1, s, u, h, c0, v1, vk, gm, zh);
        launch_p2<P, UIn, AnyWidths, false>(p, g2, s, u, h, e0, c1, vk, gm);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int table_state_pass(trisk_handle* H, const void* u, bool u_double, const double* h, hipStream_t s) {
    if (H->precision == TRISK_FP64) return table_state<Fp64, double>(H, static_cast<const double*>(u), h, s);
    if (u_double) return table_state<Fp32Bulk, double>(H, static_cast<const double*>(u), h, s);
    return table_state<Fp32Bulk, float>(H, static_cast<const float*>(u), h, s);
}

int rk4_steps_direct(trisk_handle* H, double* y, double dt, int64_t n, hipStream_t s) {
    return run_steps(H, y, dt, n, s);
}

}  // namespace trisk

extern "C" {


static void drop_graph(trisk_handle* h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_y = nullptr;
}

constexpr int64_t kGraphSteps = 8;

int trisk_rk4_steps(trisk_handle* h, double* y_dev, double dt, int64_t nsteps, void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev) return set_err(TRISK_EINVAL, "null state pointer");
    if (nsteps < 0) return set_err(TRISK_EINVAL, "nsteps must be non-negative");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->timing) {
        h->events_used = 0;
        return run_steps(h, y_dev, dt, nsteps, s);
    }
    if (!h->graphs || s == nullptr || nsteps < kGraphSteps) return run_steps(h, y_dev, dt, nsteps, s);
    // replay a captured graph of kGraphSteps steps; the kernels' arguments (y, dt and the
    // handle's buffers) are baked in, so the graph is rebuilt when y or dt change
    if (!h->graph_exec || h->graph_y != y_dev || h->graph_dt != dt) {
        drop_graph(h);
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = run_steps(h, y_dev, dt, kGraphSteps, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(TRISK_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        h->graph = g;
        HIP_TRY(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
        h->graph_y = y_dev;
        h->graph_dt = dt;
    }
    int64_t reps = nsteps / kGraphSteps;
    for (int64_t r = 0; r < reps; ++r) HIP_TRY(hipGraphLaunch(h->graph_exec, s));
    return run_steps(h, y_dev, dt, nsteps - reps * kGraphSteps, s);
}

int trisk_set_pair_plane(trisk_handle* h, int32_t on) {
    CHECK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    h->pair_plane = on != 0;
    drop_graph(h);  // a captured graph holds the kernels of the other layout
    return 0;
}

int trisk_set_graphs(trisk_handle* h, int32_t on) {
    CHECK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    h->graphs = on != 0;
    if (!h->graphs) drop_graph(h);
    return 0;
}

int trisk_kinetic_energy(trisk_handle* h, const double* y_dev, double* K_dev, double* B_dev, void* stream) {
    CHECK_HANDLE(h);
    CHECK_TRISK(h);
    if (!y_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = h->precision == TRISK_FP64
                 ? diagnostic_pass<Fp64>(h, y_dev, K_dev ? K_dev : h->diag_K, h->diag_qv, h->diag_zv, h->diag_hv, s)
                 : diagnostic_pass<Fp32Bulk>(h, y_dev, K_dev ? K_dev : h->diag_K, h->diag_qv, h->diag_zv,
                                             h->diag_hv, s);
    if (rc) return rc;
    if (B_dev) {
        const double* B = h->precision == TRISK_FP64 ? h->p64.B : h->p32.B;
        HIP_TRY(hipMemcpyAsync(B_dev, B, h->L.nc * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return 0;
}

int trisk_vertex_pv(trisk_handle* h, const double* y_dev, double* qv_dev, double* zv_dev, double* hv_dev,
                    void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->L.icon) return icon_vertex_pv(h, y_dev, qv_dev, zv_dev, hv_dev, s);
    double* q = qv_dev ? qv_dev : h->diag_qv;
    double* z = zv_dev ? zv_dev : h->diag_zv;
    double* e = hv_dev ? hv_dev : h->diag_hv;
    return h->precision == TRISK_FP64 ? diagnostic_pass<Fp64>(h, y_dev, h->diag_K, q, z, e, s)
                                      : diagnostic_pass<Fp32Bulk>(h, y_dev, h->diag_K, q, z, e, s);
}

int trisk_diagnostics(trisk_handle* h, const double* y_dev, double out[3], void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev || !out) return set_err(TRISK_EINVAL, "null pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->L.icon) {
        int rc = icon_diagnostics(h, y_dev, out, s);
        return rc ? rc : h->finite_check ? finite_status(h, y_dev, s) : 0;
    }
    const Layout& L = h->L;
    int rc = trisk_vertex_pv(h, y_dev, nullptr, nullptr, nullptr, stream);  // also fills diag_K
    if (rc) return rc;
    const double* area_c = h->precision == TRISK_FP64 ? h->p64.area_c : h->p32.area_c;
    const double* area_v = h->precision == TRISK_FP64 ? h->p64.area_v : h->p32.area_v;
    if ((rc = reduce(h, 0, L.nc, area_c, y_dev + L.ne, nullptr, &out[0], s))) return rc;
    if ((rc = reduce(h, 1, L.nc, area_c, y_dev + L.ne, h->diag_K, &out[1], s))) return rc;
    if ((rc = reduce(h, 2, L.nv, area_v, h->diag_hv, h->diag_qv, &out[2], s))) return rc;
    out[2] /= 2;
    return h->finite_check ? finite_status(h, y_dev, s) : 0;
}

int trisk_set_finite_check(trisk_handle* h, int32_t on) {
    CHECK_HANDLE(h);
    h->finite_check = on != 0;
    return 0;
}

int trisk_check_state(trisk_handle* h, const double* y_dev, int64_t counts[3], void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev || !counts) return set_err(TRISK_EINVAL, "null pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the vertex heights h_v of the state (the diagnostic pass of either model)
    int rc = h->L.icon ? icon_vertex_pv(h, y_dev, nullptr, nullptr, nullptr, s)
                       : trisk_vertex_pv(h, y_dev, nullptr, nullptr, nullptr, stream);
    if (rc) return rc;
    return state_counts(h, y_dev, counts, s);
}

int trisk_diagnostics_noalloc(trisk_handle* h, const double* y_dev, double* out_dev, void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev || !out_dev) return set_err(TRISK_EINVAL, "null pointer");
    const int blocks = h->reduce_blocks;
    if (blocks < 1 || blocks > kNoallocMaxBlocks)
        return set_err(TRISK_EINVAL, "trisk_diagnostics_noalloc: " + std::to_string(kNoallocTerms * 2 * blocks) +
                                         " doubles of partial sums exceed the static scratch of " +
                                         std::to_string(kNoallocDoubles));
    hipError_t serr;
    double* sp = noalloc_scratch(h->device, &serr);
    if (!sp)
        return set_err(TRISK_EHIP, std::string("trisk_diagnostics_noalloc: no scratch address on device ") +
                                       std::to_string(h->device) + ": " + hipGetErrorString(serr));
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Layout& L = h->L;
    // the diagnostic pass (K, q_v, h_v into the handle's work vectors), as trisk_diagnostics
    const double* K = h->diag_K;
    int rc = h->L.icon ? icon_diagnostic_pass(h, y_dev, &K, s)
                       : trisk_vertex_pv(h, y_dev, nullptr, nullptr, nullptr, stream);
    if (rc) return rc;
    const double* area_c = h->precision == TRISK_FP64 || L.icon ? h->p64.area_c : h->p32.area_c;
    const double* area_v = h->precision == TRISK_FP64 || L.icon ? h->p64.area_v : h->p32.area_v;
    const dim3 grid{unsigned(blocks)}, block{unsigned(kBlock)};
    hipLaunchKernelGGL(reduce_kernel, grid, block, 0, s, 0, L.nc, area_c, y_dev + L.ne, nullptr, L.gravity, sp);
    hipLaunchKernelGGL(reduce_kernel, grid, block, 0, s, 1, L.nc, area_c, y_dev + L.ne, K, L.gravity,
                       sp + 2 * blocks);
    hipLaunchKernelGGL(reduce_kernel, grid, block, 0, s, 2, L.nv, area_v, h->diag_hv, h->diag_qv, L.gravity,
                       sp + 4 * blocks);
    hipLaunchKernelGGL(noalloc_combine_kernel, dim3(1), dim3(64), 0, s, blocks, sp, out_dev);
    HIP_TRY(hipGetLastError());
    return 0;
}

int trisk_relative_vorticity(trisk_handle* h, const double* y_dev, double* zeta_dev, void* stream) {
    CHECK_HANDLE(h);
    if (!y_dev || !zeta_dev) return set_err(TRISK_EINVAL, "null pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->L.icon) return icon_relative_vorticity(h, y_dev, zeta_dev, s);
    const Layout& L = h->L;
    // FP64: the curl of u from the FP64 curl table; FP32-bulk: the vertex pass's FP32 curl
    if (h->precision == TRISK_FP64) {
        hipLaunchKernelGGL(edge_sum_kernel, dim3(unsigned((L.nv + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, L.nv,
                           L.kv, h->p64.ve, h->p64.curl, y_dev, zeta_dev);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    return trisk_vertex_pv(h, y_dev, nullptr, zeta_dev, nullptr, stream);
}

int trisk_tangential_velocity(trisk_handle* h, const double* u_dev, double* vt_dev, void* stream) {
    CHECK_HANDLE(h);
    if (!u_dev || !vt_dev) return set_err(TRISK_EINVAL, "null pointer");
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->L.icon) return icon_tangential_velocity(h, u_dev, vt_dev, s);
    const Layout& L = h->L;
    if (h->precision != TRISK_FP64)
        return set_err(TRISK_EINVAL, "tangential_velocity is available on FP64 handles");
    hipLaunchKernelGGL(edge_sum_kernel, dim3(unsigned((L.ne + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, L.ne, L.kt,
                       h->p64.te, h->p64.tw, u_dev, vt_dev);
    HIP_TRY(hipGetLastError());
    return 0;
}

int trisk_set_timing(trisk_handle* h, int32_t on) {
    CHECK_HANDLE(h);
    h->timing = on != 0;  // recorded events stay readable until the next timed call
    return 0;
}

int trisk_kernel_times(trisk_handle* h, int32_t max, const char** names, double* avg_ms, int64_t* launches,
                       int32_t* count) {
    CHECK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    double ms[K_COUNT] = {0};
    int64_t n[K_COUNT] = {0};
    for (size_t i = 0; i + 1 < h->events_used; i += 2) {
        HIP_TRY(hipEventSynchronize(h->events[i + 1]));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, h->events[i], h->events[i + 1]));
        int k = h->event_kernel[i / 2];
        ms[k] += t;
        n[k] += 1;
    }
    int c = 0;
    for (int k = 0; k < K_COUNT && c < max; ++k) {
        if (n[k] == 0) continue;  // kernels this handle's path does not launch
        if (names) names[c] = kKernelNames[k];
        if (avg_ms) avg_ms[c] = ms[k] / double(n[k]);
        if (launches) launches[c] = n[k];
        ++c;
    }
    if (count) *count = c;
    return 0;
}

int trisk_kernel_bytes(trisk_handle* h, int32_t max, const char** names, double* bytes, int32_t* count) {
    CHECK_HANDLE(h);
    if (!bytes && max > 0) return set_err(TRISK_EINVAL, "null bytes array");
    const Layout& L = h->L;
    // Algorithmic HBM bytes of one launch: every element of every array the launch must
    // read or write counted once (gathered arrays by their unique size), in a stage 2/3
    // launch (state read as the stage vector, RK4 accumulator read and written).
    const bool f64 = h->precision == TRISK_FP64;
    const double c = f64 ? 8.0 : 4.0;   // coefficient and intermediate bytes
    const double uw = f64 ? 8.0 : 4.0;  // stage velocity
    const double iw = 4.0;              // int32 index
    const double nc = double(L.nc), ne = double(L.ne), nv = double(L.nv);
    const double KC = L.kc, KV = L.kv, KT = L.kt;
    const double rk_e = 8.0 + 16.0 + uw, rk_c = 8.0 + 16.0 + 8.0;  // y, acc r/w, stage
    std::vector<std::pair<const char*, double>> v;
    if (L.icon) {
        v = icon_kernel_bytes(h);
    } else if (L.pair) {
        const double vc = f64 ? 7 * c : 6 * c;                  /* FP32: f rides in the index record */
        const double rec8 = f64 ? 6 * iw : 16.0;                 /* FP32: 16-bit deltas, 8-wide records */
        const double ew = f64 ? 2 * iw : 4.0;                    /* edge cells */
        const double cw = f64 ? 12 * iw : 24.0;                  /* cell edges and corners */
        // rk4_steps forms G in the previous stage's edge kernel: the state pass of stages after
        // the first covers the vertices, and the edge kernel reads l h and the new heights
        // with the (G, u) plane the vertices gather u from it (both halves of its lines)
        double w1 = nv * (rec8 + vc + c)                         /* indices, coefficients, q_v */
                    + ne * (h->pair_plane ? 2 * c : uw) + nc * 8.0;  /* u, h, gathered */
        double w2 = nc * (cw + 6 * c + iw + 8.0 + rk_c)         /* edges, corners, kite fractions, meta, A, RK4 */
                    + nc * (6 * c + 8.0 + 8.0)                   /* KE weights, h, B */
                    + ne * c + nv * c + ne * uw                  /* G, q_v, u */
                    + 6 * nc * c;                                /* half sums, 6 per cell */
        double w3 = ne * (2 * c + 1.0 + ew + c + rk_e)          /* half sums, slots, cells, d, RK4 */
                    + nc * 8.0                                   /* B */
                    + ne * (2 * c + c) + nc * 8.0;               /* l h weights, G; the new heights */
        v = {{kKernelNames[K_PAIR_STATE], w1}, {kKernelNames[K_PAIR_CELL], w2}, {kKernelNames[K_PAIR_EDGE], w3}};
    } else {
        // Gassmann: zeta_v, h_v written in P1 and gathered in P2 with the rhombus weights and
        // f at the edges; 12-wide edge, PV-edge (not for the mean, nor in the slot form) and
        // weight tables in P3.
        // Vertex kinetic energy: K_v and its weights in P1, the vertex part of h_e in P2, and
        // the cells' vertex part of K and B (kites, corners, K_v gathers) in P2.
        // On hexagon meshes the index tables are read as 16-bit deltas (2 B per index).
        const bool vk = L.vertex_ke, gm = L.gassmann;
        const double KG = L.kg;
        const bool hex = L.kc == 6 && L.kv == 3 && L.kt == 10 && (L.kg == 0 || L.kg == 12);
        const double ix = hex ? 2.0 : iw;
        // the FP32-bulk records (P1, P2 rows as entity records)
        const bool rec = !f64 && h->p32.r1c != nullptr;
        // q_v is stored only for the TRiSK flux (the Gassmann edge PV comes from zeta_v, h_v);
        // the Gassmann records then leave out f_v too
        // with every cell kinetic weight 0 (vertex kinetic energy) P1 skips the cells, whose h
        // its vertices still gather, and P2 reads h instead of B
        const bool bh = vk && L.cell_ke_zero;
        double p1 = (bh ? nc * 8.0 : nc * (KC * (ix + c) + 8.0 + 8.0)) + ne * uw +
                    nv * (KV * 2 * (ix + c) + (gm && rec ? 0.0 : c) + (gm ? 0.0 : c) + (gm || vk ? 2 * c : 0.0) +
                          (vk ? KV * c + c : 0.0));
        // with vertex kinetic energy P1 writes B and P2 adds the vertex part (K only for the
        // diagnostics)
        // the FP32-bulk records fold the rhombus weights into one coefficient per edge
        double p2 = ne * (2 * ix + 2 * c + 2 * ix + uw + 2 * c + (vk ? 2 * c : 0.0) + (gm ? (rec ? 1 : 3) * c : 0.0)) +
                    nc * 8.0 + nv * c * (gm || vk ? 2.0 : 1.0) + (vk ? nc * (KC * (c + ix) + 16.0) + nv * c : 0.0);
        const bool gp_table = !L.gassmann_mean && !(hex && L.gassmann_slots);
        const double pv = gm ? KG * (ix + c + (gp_table ? ix : 0.0)) : KT * (ix + c);
        // flux form (FP32-bulk): the continuity reads the 2-bit signs and A_i, not the divergence
        const double cont = flux_form<Fp32Bulk>(h) && !f64 ? KC * ix + 4.0 + 8.0 : KC * (ix + 8.0);
        double p3 = ne * (pv + 2 * ix + c + 2 * c + rk_e) + nc * (cont + rk_c);
        v = {{kKernelNames[K_CELL_VERTEX], p1}, {kKernelNames[K_EDGE], p2}, {kKernelNames[K_UPDATE], p3}};
    }
    int k = 0;
    for (; k < int(v.size()) && k < max; ++k) {
        if (names) names[k] = v[k].first;
        bytes[k] = v[k].second;
    }
    if (count) *count = k;
    return 0;
}

int trisk_wide_counts(const trisk_handle* h, int64_t counts[3]) {
    CHECK_HANDLE(h);
    if (!counts) return set_err(TRISK_EINVAL, "null counts array");
    counts[0] = h->L.wide_cells;
    counts[1] = h->L.wide_vertices;
    counts[2] = h->L.wide_edges;
    return 0;
}

int trisk_info(const trisk_handle* h, int32_t* recomputed, const char** note) {
    CHECK_HANDLE(h);
    if (recomputed) *recomputed = h->L.pair ? 1 : 0;
    if (note) *note = h->L.pair_reason.c_str();
    return 0;
}

}  // extern "C"
