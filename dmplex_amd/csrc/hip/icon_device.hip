// The ICON triangle shallow-water model on gfx950 (ICONShallowWater, src/FiniteVolumes/
// ICONShallowWater.jl): normal velocities at the edges, heights at the triangle
// circumcenters, vorticity at the vertices.
//
// One tendency is three launches (five with divergence damping), each a gather over one
// kind of entity in the device numbering (cells along the Hilbert curve, edges and
// vertices by first appearance), tiled per XCD like the TRiSK kernels:
//   E  icon_edge_kernel        v_e = sum_k c_k u_k (RBF), (K_e, F_e)        edges
//   C  icon_cell_vertex_kernel K_i = sum_e b u-K_e, dh_i = -sum_e n l F / A,
//                              [D_i = -div u], RK4 stage of h               cells
//                              zeta_v = sum_e t d u / A_v                   vertices
//   [G icon_damp_edge_kernel   G_e = (D_r - D_l) / d_e                      edges
//    H icon_damp_cell_kernel   D_i = -div G                                 cells]
//   M icon_momentum_kernel     du_e (+ nu G), RK4 stage of u                edges
// The RK4 stage update of rk4_step! is fused into C (heights) and M (velocities).  M
// reads the stage's input heights while C writes the next stage's, so the height stage
// alternates between two buffers.  Every FP64 formula keeps the reference's order of
// operations (no contraction), so the FP64 results equal the reference algorithm bit for bit.
//
// FP32-bulk (TRISK_FP32_BULK): the same kernels with FP32 coefficients, intermediates and
// stage velocity; the state, the stage heights, the RK4 accumulator, the pressure gradient
// g (h_r - h_l) / d and the continuity stay FP64.  The edge kernel then stores G_e = l_e F_e
// (its height weights carry l_e) and a cell's height tendency is -(sum of its edges' +-G_e)
// / A_i, summed in FP64: every G_e enters its two cells with opposite signs, so the global
// mass changes only by the rounding of the division.  Diagnostics run the FP64 kernels.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "trisk_device.hpp"

namespace trisk {

// (K_e, F_e) of an edge, one aligned pair (FP32-bulk: (K_e, l_e F_e))
template <class T>
struct alignas(2 * sizeof(T)) KeFT {
    T Ke, F;
};
using KeF = KeFT<double>;

// Device tables and work of one precision: C = double (FP64) or float (FP32-bulk
// coefficients and intermediates).
template <class C>
struct IconParamsT {
    int64_t nc, ne, nv;
    int kv, kr;
    double g;
    // cells
    const int32_t* ce;
    const C* div;
    const C* bln;
    // edges
    const int32_t* ec;
    const int32_t* ev;
    const C* eh;  // FP32-bulk: l_e times the height weights
    const C* grad;
    const C* invd;
    const int32_t* re;
    const C* rw;
    const C* fe;
    const C* nu;
    // vertices
    const int32_t* ve;
    const C* curl;
    const int32_t* vc;
    const C* vh;
    const C* fv;
    // the index tables as 16-bit deltas (two per word; kWide: read the int32 row): cell edges
    // from 3i/2, edge cells from 2e/3, edge vertices from e/3, RBF edges from e, vertex edges
    // from 3v
    const int32_t *ce16p, *ec16p, *ev16p, *re16p, *ve16p;
    // FP32-bulk continuity: bit k set when the normal of cell edge k points inwards, and A_i
    const uint8_t* cneg;
    const double* area;
    // FP32-bulk entity records (8 words each, NULL for FP64), stored as two 16-byte chunks in two
    // arrays (words 0-3 of entity j at 4 j, words 4-7 at 4 (n + j)), so that each of the two
    // 16-byte loads is contiguous across a wave; the values and words of the arrays above: per
    // edge for icon_edge_kernel [re16 2, ec16, l h 2, rw 0..2] (rw 3 stays in rw); per cell
    // [ce16 2, bln 3, inward-normal bits, A_i as FP64]; per vertex [ve16 3, curl 0..4] (curl 5
    // stays in curl).  (A record for the momentum kernel measured 9 % slower than its separate
    // loads: 100 against 92 us at triangle level 9.)
    const int32_t *rec_e, *rec_c, *rec_v;
    // work
    C* v;           // ne: RBF tangential velocity
    KeFT<C>* KF;    // ne: edge kinetic energy and mass flux
    C* K;           // nc: cell kinetic energy
    C* zeta;        // nv: relative vorticity
    C* D;           // nc: divergence-damping work
    C* G;           // ne: divergence-damping work
};
using IconParams = IconParamsT<double>;

struct IconState {
    IconParams p{};                  // FP64 tables (the FP64 model and every diagnostic)
    IconParamsT<float> p32{};        // FP32-bulk tables (stepping of an FP32-bulk handle)
    double* sh2 = nullptr;           // second height stage buffer
};

// the 8 words of entity j of a chunk-major record array (two contiguous 16-byte loads)
__device__ __forceinline__ void ld_record8(const int32_t* __restrict__ rec, int64_t n, int64_t j, int32_t (&a)[8]) {
    int32_t lo[4], hi[4];
    ldrec(rec + 4 * j, lo);
    ldrec(rec + 4 * (n + j), hi);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        a[w] = lo[w];
        a[4 + w] = hi[w];
    }
}

// ---------------------------------------------------------------------------------
// E: icon_edge_kernel! (ICONShallowWater.jl:149-159)

template <class C, class UIn, int KR>
__global__ __launch_bounds__(kBlock) void icon_edge_kernel(IconParamsT<C> p, const UIn* __restrict__ u,
                                                           const double* __restrict__ h, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    C s = 0, F;
    const C ue = C(u[e]);
    if constexpr (KR > 0) {
        // both index records and the edge's streamed values first, then every gather
        IndexWords<KR> wre{};
        IndexWords<2> wec{};
        C w[KR] = {}, eh0 = C(0), eh1 = C(0);
        bool rec = false;
        if constexpr (KR == 4 && std::is_same_v<C, float>) {
            if (p.rec_e) {  // the FP32 edge record: two 16-byte loads and the last RBF weight
                int32_t a[8];
                ld_record8(p.rec_e, p.ne, e, a);
                wre.w[0] = a[0];
                wre.w[1] = a[1];
                wec.w[0] = a[2];
                __builtin_memcpy(&eh0, &a[3], 4);
                __builtin_memcpy(&eh1, &a[4], 4);
#pragma unroll
                for (int k = 0; k < 3; ++k) __builtin_memcpy(&w[k], &a[5 + k], 4);
                w[3] = ld(p.rw + 3 * p.ne + e);
                rec = true;
            }
        }
        if (!rec) {
            wre = index_words<KR>(p.re16p, p.ne, e);
            wec = index_words<2>(p.ec16p, p.ne, e);
#pragma unroll
            for (int k = 0; k < KR; ++k) w[k] = ld(p.rw + k * p.ne + e);
            eh0 = ld(p.eh + e);
            eh1 = ld(p.eh + p.ne + e);
        }
        const Index<KR> RE(wre, p.re, p.ne, e, e);
        const Index<2> EC(wec, p.ec, p.ne, e, (2 * e) / 3);
        C ur[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) ur[k] = C(u[RE[k]]);
        const C hl = C(h[EC[0]]), hr = C(h[EC[1]]);
#pragma unroll
        for (int k = 0; k < KR; ++k) s += w[k] * ur[k];
        F = (eh0 * hl + eh1 * hr) * ue;
    } else {
        for (int k = 0; k < p.kr; ++k) s += ld(p.rw + k * p.ne + e) * C(u[ld(p.re + k * p.ne + e)]);
        const Index<2> EC(p.ec16p, p.ec, p.ne, e, (2 * e) / 3);
        F = (ld(p.eh + e) * C(h[EC[0]]) + ld(p.eh + p.ne + e) * C(h[EC[1]])) * ue;
    }
    p.v[e] = s;
    p.KF[e] = KeFT<C>{(ue * ue + s * s) / 2, F};
}

// ---------------------------------------------------------------------------------
// C: cells: K (edge_sum_kernel! with the bilinear weights), the continuity
// (continuity_kernel!, TRiSKShallowWater.jl:298-305) with the RK4 stage of h, and with
// damping D = -div u; vertices: zeta (edge_sum_kernel! with the curl weights)

template <class C, class UIn, int KV, bool DAMP>
__global__ __launch_bounds__(kBlock) void icon_cell_vertex_kernel(IconParamsT<C> p, const UIn* __restrict__ u, Rk r,
                                                                  int64_t t0, int64_t t1) {
    const Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        const int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const RkIn rin = rk_load(r, p.ne + i);
        IndexWords<3> cw{};
        C wb[3] = {}, wd[3] = {};
        int neg = 0;
        double area = 1.0;
        bool rec = false;
        if constexpr (std::is_same_v<C, float>) {
            if (p.rec_c) {  // the FP32 cell record: edges, bilinear weights, normal bits, A_i
                int32_t a[8];
                ld_record8(p.rec_c, p.nc, i, a);
                cw.w[0] = a[0];
                cw.w[1] = a[1];
#pragma unroll
                for (int k = 0; k < 3; ++k) __builtin_memcpy(&wb[k], &a[2 + k], 4);
                neg = a[5];
                __builtin_memcpy(&area, &a[6], 8);
                rec = true;
            }
        }
        if (!rec) {
            cw = index_words<3>(p.ce16p, p.nc, i);
#pragma unroll
            for (int k = 0; k < 3; ++k) wb[k] = ld(p.bln + k * p.nc + i);
            if constexpr (std::is_same_v<C, float>) {
                neg = ld(p.cneg + i);
                area = ld(p.area + i);
            }
        }
        if constexpr (std::is_same_v<C, double> || DAMP) {
#pragma unroll
            for (int k = 0; k < 3; ++k) wd[k] = ld(p.div + k * p.nc + i);
        }
        const Index<3> E(cw, p.ce, p.nc, i, (3 * i) / 2);
        KeFT<C> kf[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) kf[k] = p.KF[E[k]];
        C K = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) K += wb[k] * kf[k].Ke;
        double dh;
        if constexpr (std::is_same_v<C, double>) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += wd[k] * kf[k].F;
            dh = -s;
        } else {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += (neg >> k) & 1 ? -double(kf[k].F) : double(kf[k].F);
            dh = -s / area;
        }
        p.K[i] = K;
        if constexpr (DAMP) {
            C d = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) d += wd[k] * C(u[E[k]]);
            p.D[i] = -d;
        }
        rk_store<double, kStreamAll>(r, p.ne + i, dh, rin, r.mode >= 1 && r.mode <= 3 ? r.sh + i : nullptr);
    } else {
        const int64_t v = t.first + threadIdx.x;
        if (v >= p.nv) return;
        C z = 0;
        if constexpr (KV > 0) {
            IndexWords<KV> vw{};
            C w[KV] = {};
            bool rec = false;
            if constexpr (KV == 6 && std::is_same_v<C, float>) {
                if (p.rec_v) {  // the FP32 vertex record: edges, curl 0..4; curl 5 from its row
                    int32_t a[8];
                    ld_record8(p.rec_v, p.nv, v, a);
#pragma unroll
                    for (int q = 0; q < 3; ++q) vw.w[q] = a[q];
#pragma unroll
                    for (int k = 0; k < 5; ++k) __builtin_memcpy(&w[k], &a[3 + k], 4);
                    w[5] = ld(p.curl + 5 * p.nv + v);
                    rec = true;
                }
            }
            if (!rec) {
                vw = index_words<KV>(p.ve16p, p.nv, v);
#pragma unroll
                for (int k = 0; k < KV; ++k) w[k] = ld(p.curl + k * p.nv + v);
            }
            const Index<KV> VE(vw, p.ve, p.nv, v, 3 * v);
#pragma unroll
            for (int k = 0; k < KV; ++k) z += w[k] * C(u[VE[k]]);
        } else {
            for (int k = 0; k < p.kv; ++k) z += ld(p.curl + k * p.nv + v) * C(u[ld(p.ve + k * p.nv + v)]);
        }
        p.zeta[v] = z;
    }
}

// ---------------------------------------------------------------------------------
// divergence damping (divergence_damping!, ICONShallowWater.jl:209-217):
// G = grad D (edge_gradient_kernel!, :173-176, scale 1), D = -div G

template <class C>
__global__ __launch_bounds__(kBlock) void icon_damp_edge_kernel(IconParamsT<C> p, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    const Index<2> EC(p.ec16p, p.ec, p.ne, e, (2 * e) / 3);
    const int32_t l = EC[0], r = EC[1];
    p.G[e] = (C(1) * (p.D[r] - p.D[l])) * ld(p.invd + e);
}

template <class C>
__global__ __launch_bounds__(kBlock) void icon_damp_cell_kernel(IconParamsT<C> p, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const Index<3> E(p.ce16p, p.ce, p.nc, i, (3 * i) / 2);
    C s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += ld(p.div + k * p.nc + i) * p.G[E[k]];
    p.D[i] = -s;
}

// ---------------------------------------------------------------------------------
// M: icon_momentum_kernel! (:161-171), with damping the last edge_gradient_kernel!
// (scale -1) and du += nu G, then the RK4 stage of u

template <class C, bool DAMP>
__global__ __launch_bounds__(kBlock) void icon_momentum_kernel(IconParamsT<C> p, const double* __restrict__ h, Rk r,
                                                               int64_t t0) {
    using U = C;  // stage velocity
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    // the edge's own values and both index records first, then every gather
    const RkIn rin = rk_load(r, e);
    const IndexWords<2> wec = index_words<2>(p.ec16p, p.ne, e), wev = index_words<2>(p.ev16p, p.ne, e);
    const C g1 = ld(p.grad + e), g2 = ld(p.grad + p.ne + e);
    const C invd = ld(p.invd + e);
    const C fe = ld(p.fe + e), ve = ld(p.v + e), Ke = p.KF[e].Ke;
    const C nu = DAMP ? ld(p.nu + e) : C(0);
    const Index<2> EC(wec, p.ec, p.ne, e, (2 * e) / 3), EV(wev, p.ev, p.ne, e, e / 3);
    const int32_t l = EC[0], rr = EC[1];
    const int32_t v0 = EV[0], v1 = EV[1];
    const C z0 = p.zeta[v0], z1 = p.zeta[v1], Kr = p.K[rr], Kl = p.K[l];
    const double hr = h[rr], hl = h[l];
    C Dr = C(0), Dl = C(0);
    if constexpr (DAMP) {
        Dr = p.D[rr];
        Dl = p.D[l];
    }
    const C zeta = (z0 + z1) / 2;
    const C grad_K = Ke * (g1 - g2) + g2 * Kr - g1 * Kl;
    // the pressure gradient in FP64 in both precisions (for FP64 this is the reference's
    // left-to-right order of (f + zeta) v - grad K - g (h_r - h_l) / d)
    double du = double((fe + zeta) * ve - grad_K) - p.g * (hr - hl) * double(invd);
    if constexpr (DAMP) {
        const C G = (C(-1) * (Dr - Dl)) * invd;
        du = du + double(nu * G);
    }
    rk_store<U, kStreamAll>(r, e, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<U*>(r.su) + e : nullptr);
}

// ---------------------------------------------------------------------------------
// diagnostics: vertex_pv_kernel! (TRiSKShallowWater.jl:246-259) and the plain gathers

__global__ __launch_bounds__(kBlock) void icon_vertex_pv_kernel(IconParams p, const double* __restrict__ u,
                                                                const double* __restrict__ h, double* __restrict__ qv,
                                                                double* __restrict__ zv, double* __restrict__ hv) {
    const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (v >= p.nv) return;
    double z = 0.0, eta = 0.0;
    for (int k = 0; k < p.kv; ++k) z += p.curl[k * p.nv + v] * u[p.ve[k * p.nv + v]];
    for (int k = 0; k < p.kv; ++k) eta += p.vh[k * p.nv + v] * h[p.vc[k * p.nv + v]];
    zv[v] = z;
    hv[v] = eta;
    qv[v] = (z + p.fv[v]) / eta;
}

__global__ __launch_bounds__(kBlock) void icon_gather_kernel(int64_t n, int K, const int32_t* __restrict__ idx,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ x, double* __restrict__ out) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += w[k * n + j] * x[idx[k * n + j]];
    out[j] = s;
}

// ---------------------------------------------------------------------------------
// host side

static IconParams& iparams(trisk_handle* H) { return H->icon->p; }

static unsigned grid1(int64_t n) { return unsigned(kXcd * tiles_per_xcd(n)); }

// One tendency with the stage epilogue r: u, h are the stage's input; r.sh receives the
// next stage's heights (modes 1-3).
template <class C, class UIn>
static int icon_stage(trisk_handle* H, const IconParamsT<C>& p, const UIn* u, const double* h, const Rk& r,
                      hipStream_t s) {
    const bool damp = p.nu != nullptr;
    const int64_t te = tiles_per_xcd(p.ne), tc = tiles_per_xcd(p.nc), tv = tiles_per_xcd(p.nv);
    const dim3 b(kBlock);
    hipEvent_t stop;
    int rc;
    if ((rc = timed_begin(H, K_ICON_EDGE, s, &stop))) return rc;
    if (p.kr == 4)
        hipLaunchKernelGGL((icon_edge_kernel<C, UIn, 4>), dim3(grid1(p.ne)), b, 0, s, p, u, h, te);
    else
        hipLaunchKernelGGL((icon_edge_kernel<C, UIn, 0>), dim3(grid1(p.ne)), b, 0, s, p, u, h, te);
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;

    if ((rc = timed_begin(H, K_ICON_CELL_VERTEX, s, &stop))) return rc;
    const dim3 gcv(unsigned(kXcd * (tc + tv)));
    if (p.kv == 6) {
        if (damp)
            hipLaunchKernelGGL((icon_cell_vertex_kernel<C, UIn, 6, true>), gcv, b, 0, s, p, u, r, tc, tv);
        else
            hipLaunchKernelGGL((icon_cell_vertex_kernel<C, UIn, 6, false>), gcv, b, 0, s, p, u, r, tc, tv);
    } else {
        if (damp)
            hipLaunchKernelGGL((icon_cell_vertex_kernel<C, UIn, 0, true>), gcv, b, 0, s, p, u, r, tc, tv);
        else
            hipLaunchKernelGGL((icon_cell_vertex_kernel<C, UIn, 0, false>), gcv, b, 0, s, p, u, r, tc, tv);
    }
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;

    if (damp) {
        if ((rc = timed_begin(H, K_ICON_DAMP_EDGE, s, &stop))) return rc;
        hipLaunchKernelGGL((icon_damp_edge_kernel<C>), dim3(grid1(p.ne)), b, 0, s, p, te);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
        if ((rc = timed_begin(H, K_ICON_DAMP_CELL, s, &stop))) return rc;
        hipLaunchKernelGGL((icon_damp_cell_kernel<C>), dim3(grid1(p.nc)), b, 0, s, p, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }

    if ((rc = timed_begin(H, K_ICON_MOMENTUM, s, &stop))) return rc;
    if (damp)
        hipLaunchKernelGGL((icon_momentum_kernel<C, true>), dim3(grid1(p.ne)), b, 0, s, p, h, r, te);
    else
        hipLaunchKernelGGL((icon_momentum_kernel<C, false>), dim3(grid1(p.ne)), b, 0, s, p, h, r, te);
    HIP_TRY(hipGetLastError());
    return timed_end(stop, s);
}

int icon_tendency(trisk_handle* H, const double* y, double* dy, hipStream_t s) {
    Rk r{};
    r.mode = 0;
    r.dy = dy;
    if (H->precision == TRISK_FP32_BULK) return icon_stage(H, H->icon->p32, y, y + H->L.ne, r, s);
    return icon_stage(H, H->icon->p, y, y + H->L.ne, r, s);
}

template <class C>
static int icon_steps(trisk_handle* H, const IconParamsT<C>& p, double* y, double dt, int64_t n, hipStream_t s) {
    const int64_t ne = H->L.ne;
    C* su = static_cast<C*>(H->su);
    double* sh[2] = {H->sh, H->icon->sh2};
    Rk r{};
    r.y = y;
    r.acc = H->acc;
    r.su = su;
    r.dt = dt;
    r.dt2 = dt / 2;
    r.dt6 = dt / 6;
    int rc;
    for (int64_t step = 0; step < n; ++step) {
        // stage k reads (u, h) of stage k-1 and writes stage k: heights sh[0], sh[1], sh[0]
        for (int mode = 1; mode <= 4; ++mode) {
            r.mode = mode;
            r.sh = mode <= 3 ? sh[(mode + 1) % 2] : nullptr;
            if (mode == 1) {
                rc = icon_stage(H, p, static_cast<const double*>(y), y + ne, r, s);
            } else {
                rc = icon_stage(H, p, static_cast<const C*>(su), sh[mode % 2], r, s);
            }
            if (rc) return rc;
        }
    }
    return 0;
}

int icon_run_steps(trisk_handle* H, double* y, double dt, int64_t n, hipStream_t s) {
    if (H->precision == TRISK_FP32_BULK) return icon_steps(H, H->icon->p32, y, dt, n, s);
    return icon_steps(H, H->icon->p, y, dt, n, s);
}

// K at the cells and zeta at the vertices of the state y (FP64 tables in both precisions):
// the stage kernels with the tendency epilogue into the RK4 accumulator, which holds nothing
// between steps
static int icon_kinetic_pass(trisk_handle* H, const double* y, hipStream_t s) {
    Rk r{};
    r.mode = 0;
    r.dy = H->acc;
    IconParams& p = iparams(H);
    const int64_t te = tiles_per_xcd(p.ne), tc = tiles_per_xcd(p.nc), tv = tiles_per_xcd(p.nv);
    const double* u = y;
    const double* h = y + p.ne;
    hipLaunchKernelGGL((icon_edge_kernel<double, double, 0>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, u, h, te);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((icon_cell_vertex_kernel<double, double, 0, false>), dim3(unsigned(kXcd * (tc + tv))),
                       dim3(kBlock), 0, s, p, u, r, tc, tv);
    HIP_TRY(hipGetLastError());
    return 0;
}

int icon_vertex_pv(trisk_handle* H, const double* y, double* qv, double* zv, double* hv, hipStream_t s) {
    if (!y) return set_err(TRISK_EINVAL, "null state pointer");
    IconParams& p = iparams(H);
    hipLaunchKernelGGL(icon_vertex_pv_kernel, dim3(unsigned((p.nv + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, p, y,
                       y + p.ne, qv ? qv : H->diag_qv, zv ? zv : H->diag_zv, hv ? hv : H->diag_hv);
    HIP_TRY(hipGetLastError());
    return 0;
}

int icon_diagnostic_pass(trisk_handle* H, const double* y, const double** K, hipStream_t s) {
    int rc;
    if ((rc = icon_kinetic_pass(H, y, s))) return rc;
    if ((rc = icon_vertex_pv(H, y, nullptr, nullptr, nullptr, s))) return rc;
    *K = iparams(H).K;
    return 0;
}

int icon_diagnostics(trisk_handle* H, const double* y, double out[3], hipStream_t s) {
    const Layout& L = H->L;
    IconParams& p = iparams(H);
    int rc;
    if ((rc = icon_kinetic_pass(H, y, s))) return rc;
    if ((rc = icon_vertex_pv(H, y, nullptr, nullptr, nullptr, s))) return rc;
    if ((rc = reduce(H, 0, L.nc, H->p64.area_c, y + L.ne, nullptr, &out[0], s))) return rc;
    if ((rc = reduce(H, 1, L.nc, H->p64.area_c, y + L.ne, p.K, &out[1], s))) return rc;
    if ((rc = reduce(H, 2, L.nv, H->p64.area_v, H->diag_hv, H->diag_qv, &out[2], s))) return rc;
    out[2] /= 2;
    return 0;
}

int icon_relative_vorticity(trisk_handle* H, const double* y, double* zeta, hipStream_t s) {
    if (!y || !zeta) return set_err(TRISK_EINVAL, "null pointer");
    IconParams& p = iparams(H);
    hipLaunchKernelGGL(icon_gather_kernel, dim3(unsigned((p.nv + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, p.nv,
                       p.kv, p.ve, p.curl, y, zeta);
    HIP_TRY(hipGetLastError());
    return 0;
}

int icon_tangential_velocity(trisk_handle* H, const double* u, double* vt, hipStream_t s) {
    if (!u || !vt) return set_err(TRISK_EINVAL, "null pointer");
    IconParams& p = iparams(H);
    hipLaunchKernelGGL(icon_gather_kernel, dim3(unsigned((p.ne + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, p.ne,
                       p.kr, p.re, p.rw, u, vt);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Algorithmic HBM bytes of one launch in a stage-2/3 step (state read as the stage
// vectors, RK4 accumulator read and written, stage written); gathered arrays are counted
// once at their full size.
std::vector<std::pair<const char*, double>> icon_kernel_bytes(const trisk_handle* H) {
    const Layout& L = H->L;
    const bool damp = !L.damp.empty();
    const double nc = double(L.nc), ne = double(L.ne), nv = double(L.nv), KR = L.kr, KV = L.kv;
    // index tables as 16-bit deltas (2 B per index); the RBF and vertex-edge tables of other
    // widths than the triangle mesh's 4 and 6 are read as int32
    const bool f32 = H->precision == TRISK_FP32_BULK;
    const double ix = 2.0, ixr = L.kr == 4 ? 2.0 : 4.0, ixv = L.kv == 6 ? 2.0 : 4.0;
    // coefficients, intermediates and stage velocity (FP32-bulk: 4 B); heights, y and the
    // accumulator are FP64.  u: stage velocity; rk: y, accumulator read and written, stage
    const double c = f32 ? 4.0 : 8.0, u = c, rk_h = 8.0 + 16.0 + 8.0, rk_u = 8.0 + 16.0 + u;
    std::vector<std::pair<const char*, double>> v;
    // E: RBF table, u, cells, c_lin_e, h, (K_e, F_e), v
    v.push_back({kKernelNames[K_ICON_EDGE], ne * (KR * (ixr + c) + u + 2 * ix + 2 * c + 2 * c + c) + nc * 8.0});
    // C: cell edges, bilinear and divergence weights (FP32-bulk: bilinear weights, inward
    //    bits and A_i), (K_e, F_e), K, RK4 of h [u, D, the divergence weights];
    //    vertex edges, curl, u, zeta
    const double cw = f32 ? 3 * c + 1.0 + 8.0 + (damp ? 3 * c : 0.0) : 6 * c;
    v.push_back({kKernelNames[K_ICON_CELL_VERTEX], nc * (3 * ix + cw + c + rk_h + (damp ? c : 0.0)) +
                                                       ne * (2 * c + u) + nv * (KV * (ixv + c) + c)});
    if (damp) {
        v.push_back({kKernelNames[K_ICON_DAMP_EDGE], ne * (2 * ix + c + c) + nc * c});
        v.push_back({kKernelNames[K_ICON_DAMP_CELL], nc * (3 * ix + 3 * c + c) + ne * c});
    }
    // M: cells, vertices, gradient, 1/d, f_e, v, K_e, K, h, zeta, RK4 of u [nu, D]
    v.push_back({kKernelNames[K_ICON_MOMENTUM], ne * (4 * ix + 2 * c + c + c + c + c + rk_u + (damp ? c : 0.0)) +
                                                    nc * (c + 8.0 + (damp ? c : 0.0)) + nv * c});
    return v;
}

static void icon_release(trisk_handle* H) {
    delete H->icon;
    H->icon = nullptr;
}

static int icon_setup(trisk_handle* H) {
    const Layout& L = H->L;
    auto& o = H->owned;
    H->icon = new IconState;
    H->icon_release = icon_release;
    IconParams& p = H->icon->p;
    p.nc = L.nc;
    p.ne = L.ne;
    p.nv = L.nv;
    p.kv = L.kv;
    p.kr = L.kr;
    p.g = L.gravity;
    int rc = 0;
    rc |= upload(o, L.ce, &p.ce);
    rc |= upload(o, L.div, &p.div);
    rc |= upload(o, L.bln, &p.bln);
    rc |= upload(o, L.ec, &p.ec);
    rc |= upload(o, L.ev, &p.ev);
    rc |= upload(o, L.eh, &p.eh);
    rc |= upload(o, L.grad, &p.grad);
    rc |= upload(o, L.invd, &p.invd);
    rc |= upload(o, L.re, &p.re);
    rc |= upload(o, L.rw, &p.rw);
    rc |= upload(o, L.fe, &p.fe);
    if (!L.damp.empty()) rc |= upload(o, L.damp, &p.nu);
    rc |= upload(o, L.ve, &p.ve);
    rc |= upload(o, L.curl, &p.curl);
    rc |= upload(o, L.vc, &p.vc);
    rc |= upload(o, L.vh, &p.vh);
    rc |= upload(o, L.fv, &p.fv);
    rc |= upload(o, L.ce16p, &p.ce16p);
    rc |= upload(o, L.ec16p, &p.ec16p);
    rc |= upload(o, L.ev16p, &p.ev16p);
    rc |= upload(o, L.re16p, &p.re16p);
    rc |= upload(o, L.ve16p, &p.ve16p);
    rc |= upload(o, L.area_c, &H->p64.area_c);
    rc |= upload(o, L.area_v, &H->p64.area_v);
    if (rc) return rc;
    const bool f32 = H->precision == TRISK_FP32_BULK;
    if (f32) {
        // FP32-bulk: FP32 coefficients; the height weights times l_e = |div| A of the edge's
        // first cell slot; per cell the inward-normal bits of its edges
        IconParamsT<float>& q = H->icon->p32;
        q.nc = p.nc;
        q.ne = p.ne;
        q.nv = p.nv;
        q.kv = p.kv;
        q.kr = p.kr;
        q.g = p.g;
        q.ce = p.ce;
        q.ec = p.ec;
        q.ev = p.ev;
        q.re = p.re;
        q.ve = p.ve;
        q.vc = p.vc;
        q.ce16p = p.ce16p;
        q.ec16p = p.ec16p;
        q.ev16p = p.ev16p;
        q.re16p = p.re16p;
        q.ve16p = p.ve16p;
        q.area = H->p64.area_c;
        std::vector<double> len(L.ne, 0.0);
        std::vector<char> seen(L.ne, 0);
        std::vector<uint8_t> neg(L.nc, 0);
        for (int64_t i = 0; i < L.nc; ++i)
            for (int k = 0; k < 3; ++k) {
                const double d = L.div[size_t(k) * L.nc + i];
                const int32_t e = L.ce[size_t(k) * L.nc + i];
                if (d < 0) neg[i] |= uint8_t(1u << k);
                if (!seen[e]) {
                    seen[e] = 1;
                    len[e] = std::fabs(d) * L.area_c[i];
                }
            }
        std::vector<double> lh(L.eh.size());
        for (int64_t e = 0; e < L.ne; ++e) {
            lh[e] = len[e] * L.eh[e];
            lh[L.ne + e] = len[e] * L.eh[L.ne + e];
        }
        rc |= upload(o, to_float(L.div), &q.div);
        rc |= upload(o, to_float(L.bln), &q.bln);
        rc |= upload(o, to_float(lh), &q.eh);
        rc |= upload(o, to_float(L.grad), &q.grad);
        rc |= upload(o, to_float(L.invd), &q.invd);
        rc |= upload(o, to_float(L.rw), &q.rw);
        rc |= upload(o, to_float(L.fe), &q.fe);
        if (!L.damp.empty()) rc |= upload(o, to_float(L.damp), &q.nu);
        rc |= upload(o, to_float(L.curl), &q.curl);
        rc |= upload(o, neg, &q.cneg);
        if (rc) return rc;
        // entity records (triangle meshes: 4 RBF edges, 6 vertex edges)
        if (L.kr == 4 && L.kv == 6) {
            const int64_t ne = L.ne, nc = L.nc, nv = L.nv;
            auto fw = [](double x) {
                const float f = float(x);
                int32_t w;
                std::memcpy(&w, &f, 4);
                return w;
            };
            std::vector<int32_t> re(8 * size_t(ne)), rcl(8 * size_t(nc)), rv(8 * size_t(nv));
            for (int64_t e = 0; e < ne; ++e) {
                int32_t* a = &re[8 * e];
                a[0] = L.re16p[e];
                a[1] = L.re16p[ne + e];
                a[2] = L.ec16p[e];
                a[3] = fw(lh[e]);
                a[4] = fw(lh[ne + e]);
                for (int k = 0; k < 3; ++k) a[5 + k] = fw(L.rw[size_t(k) * ne + e]);
            }
            for (int64_t i = 0; i < nc; ++i) {
                int32_t* c = &rcl[8 * i];
                c[0] = L.ce16p[i];
                c[1] = L.ce16p[nc + i];
                for (int k = 0; k < 3; ++k) c[2 + k] = fw(L.bln[size_t(k) * nc + i]);
                c[5] = neg[i];
                std::memcpy(&c[6], &L.area_c[i], 8);
            }
            for (int64_t v = 0; v < nv; ++v) {
                int32_t* a = &rv[8 * v];
                for (int q2 = 0; q2 < 3; ++q2) a[q2] = L.ve16p[size_t(q2) * nv + v];
                for (int k = 0; k < 5; ++k) a[3 + k] = fw(L.curl[size_t(k) * nv + v]);
            }
            // chunk-major: words 0-3 of every entity, then words 4-7
            auto chunked = [](const std::vector<int32_t>& aos) {
                const size_t n = aos.size() / 8;
                std::vector<int32_t> out(aos.size());
                for (size_t j = 0; j < n; ++j)
                    for (int w = 0; w < 8; ++w) out[(w < 4 ? 4 * j : 4 * (n + j)) + w % 4] = aos[8 * j + w];
                return out;
            };
            rc |= upload(o, chunked(re), &q.rec_e);
            rc |= upload(o, chunked(rcl), &q.rec_c);
            rc |= upload(o, chunked(rv), &q.rec_v);
            if (rc) return rc;
        }
    }
    void* q;
    struct Buf {
        void** out;
        size_t bytes;
    } bufs[] = {{reinterpret_cast<void**>(&p.v), L.ne * sizeof(double)},
                {reinterpret_cast<void**>(&p.KF), L.ne * sizeof(KeF)},
                {reinterpret_cast<void**>(&p.K), L.nc * sizeof(double)},
                {reinterpret_cast<void**>(&p.zeta), L.nv * sizeof(double)},
                {reinterpret_cast<void**>(&p.D), L.nc * sizeof(double)},
                {reinterpret_cast<void**>(&p.G), L.ne * sizeof(double)},
                {reinterpret_cast<void**>(&H->acc), (L.ne + L.nc) * sizeof(double)},
                {&H->su, L.ne * (f32 ? sizeof(float) : sizeof(double))},
                {reinterpret_cast<void**>(&H->sh), L.nc * sizeof(double)},
                {reinterpret_cast<void**>(&H->icon->sh2), L.nc * sizeof(double)},
                {reinterpret_cast<void**>(&H->partial), 2 * H->reduce_blocks * sizeof(double)},
                {reinterpret_cast<void**>(&H->diag_K), L.nc * sizeof(double)},
                {reinterpret_cast<void**>(&H->diag_qv), L.nv * sizeof(double)},
                {reinterpret_cast<void**>(&H->diag_zv), L.nv * sizeof(double)},
                {reinterpret_cast<void**>(&H->diag_hv), L.nv * sizeof(double)}};
    for (auto& b : bufs) {
        if ((rc = alloc(o, b.bytes, &q))) return rc;
        *b.out = q;
    }
    if (f32) {
        IconParamsT<float>& w = H->icon->p32;
        Buf wb[] = {{reinterpret_cast<void**>(&w.v), L.ne * sizeof(float)},
                    {reinterpret_cast<void**>(&w.KF), L.ne * sizeof(KeFT<float>)},
                    {reinterpret_cast<void**>(&w.K), L.nc * sizeof(float)},
                    {reinterpret_cast<void**>(&w.zeta), L.nv * sizeof(float)},
                    {reinterpret_cast<void**>(&w.D), L.nc * sizeof(float)},
                    {reinterpret_cast<void**>(&w.G), L.ne * sizeof(float)}};
        for (auto& b : wb) {
            if ((rc = alloc(o, b.bytes, &q))) return rc;
            *b.out = q;
        }
    }
    return 0;
}

}  // namespace trisk

using namespace trisk;

extern "C" {

int icon_check_tables(const icon_model_desc* desc) {
    Layout L;
    std::string err;
    if (!build_icon_layout(desc, L, err)) return set_err(TRISK_EINVAL, err);
    return 0;
}

int icon_create(const icon_model_desc* desc, const trisk_options* opt, trisk_handle** out) {
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
    H->device = o.device;
    std::string err;
    if (!build_icon_layout(desc, H->L, err, o.narrow_limit)) {
        delete H;
        return set_err(TRISK_EINVAL, err);
    }
    int ndev = 0;
    hipError_t he = hipGetDeviceCount(&ndev);
    if (he != hipSuccess || o.device < 0 || o.device >= ndev || hipSetDevice(o.device) != hipSuccess) {
        delete H;
        return set_err(TRISK_EHIP, "no such HIP device");
    }
    noalloc_setup(o.device);
    int rc = icon_setup(H);
    if (rc) {
        for (auto& b : H->owned) (void)hipFree(b.p);
        if (H->icon_release) H->icon_release(H);
        delete H;
        return rc;
    }
    *out = H;
    return 0;
}

}  // extern "C"
