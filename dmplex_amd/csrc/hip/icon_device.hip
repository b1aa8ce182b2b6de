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
#include <cstring>

#include "trisk_device.hpp"

namespace trisk {

struct alignas(16) KeF {
    double Ke, F;
};

struct IconParams {
    int64_t nc, ne, nv;
    int kv, kr;
    double g;
    // cells
    const int32_t* ce;
    const double* div;
    const double* bln;
    // edges
    const int32_t* ec;
    const int32_t* ev;
    const double* eh;
    const double* grad;
    const double* invd;
    const int32_t* re;
    const double* rw;
    const double* fe;
    const double* nu;
    // vertices
    const int32_t* ve;
    const double* curl;
    const int32_t* vc;
    const double* vh;
    const double* fv;
    // the index tables as 16-bit deltas (two per word; kWide: read the int32 row): cell edges
    // from 3i/2, edge cells from 2e/3, edge vertices from e/3, RBF edges from e, vertex edges
    // from 3v
    const int32_t *ce16p, *ec16p, *ev16p, *re16p, *ve16p;
    // work
    double* v;    // ne: RBF tangential velocity
    KeF* KF;      // ne: edge kinetic energy and mass flux
    double* K;    // nc: cell kinetic energy
    double* zeta; // nv: relative vorticity
    double* D;    // nc: divergence-damping work
    double* G;    // ne: divergence-damping work
};

struct IconState {
    IconParams p{};
    double* sh2 = nullptr;  // second height stage buffer
};

// ---------------------------------------------------------------------------------
// E: icon_edge_kernel! (ICONShallowWater.jl:149-159)

template <int KR>
__global__ __launch_bounds__(kBlock) void icon_edge_kernel(IconParams p, const double* __restrict__ u,
                                                           const double* __restrict__ h, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    double s = 0.0;
    if constexpr (KR > 0) {
        const Index<KR> RE(p.re16p, p.re, p.ne, e, e);
        double w[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) w[k] = ld(p.rw + k * p.ne + e);
#pragma unroll
        for (int k = 0; k < KR; ++k) s += w[k] * u[RE[k]];
    } else {
        for (int k = 0; k < p.kr; ++k) s += ld(p.rw + k * p.ne + e) * u[ld(p.re + k * p.ne + e)];
    }
    const double ue = u[e];
    const Index<2> EC(p.ec16p, p.ec, p.ne, e, (2 * e) / 3);
    const int32_t l = EC[0], r = EC[1];
    const double F = (ld(p.eh + e) * h[l] + ld(p.eh + p.ne + e) * h[r]) * ue;
    p.v[e] = s;
    p.KF[e] = KeF{(ue * ue + s * s) / 2, F};
}

// ---------------------------------------------------------------------------------
// C: cells: K (edge_sum_kernel! with the bilinear weights), the continuity
// (continuity_kernel!, TRiSKShallowWater.jl:298-305) with the RK4 stage of h, and with
// damping D = -div u; vertices: zeta (edge_sum_kernel! with the curl weights)

template <int KV, bool DAMP>
__global__ __launch_bounds__(kBlock) void icon_cell_vertex_kernel(IconParams p, const double* __restrict__ u, Rk r,
                                                                  int64_t t0, int64_t t1) {
    const Tile t = tile_of(blockIdx.x, t0, t1);
    if (t.range == 0) {
        const int64_t i = t.first + threadIdx.x;
        if (i >= p.nc) return;
        const RkIn rin = rk_load(r, p.ne + i);
        const Index<3> E(p.ce16p, p.ce, p.nc, i, (3 * i) / 2);
        double wb[3], wd[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            wb[k] = ld(p.bln + k * p.nc + i);
            wd[k] = ld(p.div + k * p.nc + i);
        }
        KeF kf[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) kf[k] = p.KF[E[k]];
        double K = 0.0, s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) K += wb[k] * kf[k].Ke;
#pragma unroll
        for (int k = 0; k < 3; ++k) s += wd[k] * kf[k].F;
        p.K[i] = K;
        if constexpr (DAMP) {
            double d = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) d += wd[k] * u[E[k]];
            p.D[i] = -d;
        }
        rk_store<double>(r, p.ne + i, -s, rin, r.mode >= 1 && r.mode <= 3 ? r.sh + i : nullptr);
    } else {
        const int64_t v = t.first + threadIdx.x;
        if (v >= p.nv) return;
        double z = 0.0;
        if constexpr (KV > 0) {
            const Index<KV> VE(p.ve16p, p.ve, p.nv, v, 3 * v);
            double w[KV];
#pragma unroll
            for (int k = 0; k < KV; ++k) w[k] = ld(p.curl + k * p.nv + v);
#pragma unroll
            for (int k = 0; k < KV; ++k) z += w[k] * u[VE[k]];
        } else {
            for (int k = 0; k < p.kv; ++k) z += ld(p.curl + k * p.nv + v) * u[ld(p.ve + k * p.nv + v)];
        }
        p.zeta[v] = z;
    }
}

// ---------------------------------------------------------------------------------
// divergence damping (divergence_damping!, ICONShallowWater.jl:209-217):
// G = grad D (edge_gradient_kernel!, :173-176, scale 1), D = -div G

__global__ __launch_bounds__(kBlock) void icon_damp_edge_kernel(IconParams p, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    const Index<2> EC(p.ec16p, p.ec, p.ne, e, (2 * e) / 3);
    const int32_t l = EC[0], r = EC[1];
    p.G[e] = (1.0 * (p.D[r] - p.D[l])) * ld(p.invd + e);
}

__global__ __launch_bounds__(kBlock) void icon_damp_cell_kernel(IconParams p, int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t i = t.first + threadIdx.x;
    if (i >= p.nc) return;
    const Index<3> E(p.ce16p, p.ce, p.nc, i, (3 * i) / 2);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) s += ld(p.div + k * p.nc + i) * p.G[E[k]];
    p.D[i] = -s;
}

// ---------------------------------------------------------------------------------
// M: icon_momentum_kernel! (:161-171), with damping the last edge_gradient_kernel!
// (scale -1) and du += nu G, then the RK4 stage of u

template <bool DAMP>
__global__ __launch_bounds__(kBlock) void icon_momentum_kernel(IconParams p, const double* __restrict__ h, Rk r,
                                                               int64_t t0) {
    const Tile t = tile_of(blockIdx.x, t0, 0);
    const int64_t e = t.first + threadIdx.x;
    if (e >= p.ne) return;
    const RkIn rin = rk_load(r, e);
    const Index<2> EC(p.ec16p, p.ec, p.ne, e, (2 * e) / 3), EV(p.ev16p, p.ev, p.ne, e, e / 3);
    const int32_t l = EC[0], rr = EC[1];
    const int32_t v0 = EV[0], v1 = EV[1];
    const double g1 = ld(p.grad + e), g2 = ld(p.grad + p.ne + e);
    const double invd = ld(p.invd + e);
    const double zeta = (p.zeta[v0] + p.zeta[v1]) / 2;
    const double grad_K = p.KF[e].Ke * (g1 - g2) + g2 * p.K[rr] - g1 * p.K[l];
    double du = (ld(p.fe + e) + zeta) * ld(p.v + e) - grad_K - p.g * (h[rr] - h[l]) * invd;
    if constexpr (DAMP) {
        const double G = (-1.0 * (p.D[rr] - p.D[l])) * invd;
        du = du + ld(p.nu + e) * G;
    }
    rk_store<double>(r, e, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<double*>(r.su) + e : nullptr);
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
static int icon_stage(trisk_handle* H, const double* u, const double* h, const Rk& r, hipStream_t s) {
    IconParams& p = iparams(H);
    const bool damp = p.nu != nullptr;
    const int64_t te = tiles_per_xcd(p.ne), tc = tiles_per_xcd(p.nc), tv = tiles_per_xcd(p.nv);
    hipEvent_t stop;
    int rc;
    if ((rc = timed_begin(H, K_ICON_EDGE, s, &stop))) return rc;
    if (p.kr == 4)
        hipLaunchKernelGGL((icon_edge_kernel<4>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, u, h, te);
    else
        hipLaunchKernelGGL((icon_edge_kernel<0>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, u, h, te);
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;

    if ((rc = timed_begin(H, K_ICON_CELL_VERTEX, s, &stop))) return rc;
    const dim3 gcv(unsigned(kXcd * (tc + tv)));
    if (p.kv == 6) {
        if (damp)
            hipLaunchKernelGGL((icon_cell_vertex_kernel<6, true>), gcv, dim3(kBlock), 0, s, p, u, r, tc, tv);
        else
            hipLaunchKernelGGL((icon_cell_vertex_kernel<6, false>), gcv, dim3(kBlock), 0, s, p, u, r, tc, tv);
    } else {
        if (damp)
            hipLaunchKernelGGL((icon_cell_vertex_kernel<0, true>), gcv, dim3(kBlock), 0, s, p, u, r, tc, tv);
        else
            hipLaunchKernelGGL((icon_cell_vertex_kernel<0, false>), gcv, dim3(kBlock), 0, s, p, u, r, tc, tv);
    }
    HIP_TRY(hipGetLastError());
    if ((rc = timed_end(stop, s))) return rc;

    if (damp) {
        if ((rc = timed_begin(H, K_ICON_DAMP_EDGE, s, &stop))) return rc;
        hipLaunchKernelGGL(icon_damp_edge_kernel, dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, te);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
        if ((rc = timed_begin(H, K_ICON_DAMP_CELL, s, &stop))) return rc;
        hipLaunchKernelGGL(icon_damp_cell_kernel, dim3(grid1(p.nc)), dim3(kBlock), 0, s, p, tc);
        HIP_TRY(hipGetLastError());
        if ((rc = timed_end(stop, s))) return rc;
    }

    if ((rc = timed_begin(H, K_ICON_MOMENTUM, s, &stop))) return rc;
    if (damp)
        hipLaunchKernelGGL((icon_momentum_kernel<true>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, h, r, te);
    else
        hipLaunchKernelGGL((icon_momentum_kernel<false>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, h, r, te);
    HIP_TRY(hipGetLastError());
    return timed_end(stop, s);
}

int icon_tendency(trisk_handle* H, const double* y, double* dy, hipStream_t s) {
    Rk r{};
    r.mode = 0;
    r.dy = dy;
    return icon_stage(H, y, y + H->L.ne, r, s);
}

int icon_run_steps(trisk_handle* H, double* y, double dt, int64_t n, hipStream_t s) {
    const int64_t ne = H->L.ne;
    double* su = static_cast<double*>(H->su);
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
            const double* u = mode == 1 ? y : su;
            const double* h = mode == 1 ? y + ne : sh[mode % 2];
            if ((rc = icon_stage(H, u, h, r, s))) return rc;
        }
    }
    return 0;
}

// K at the cells and zeta at the vertices of the state y: the stage kernels with the
// tendency epilogue into the RK4 accumulator, which holds nothing between steps
static int icon_kinetic_pass(trisk_handle* H, const double* y, hipStream_t s) {
    Rk r{};
    r.mode = 0;
    r.dy = H->acc;
    IconParams& p = iparams(H);
    const int64_t te = tiles_per_xcd(p.ne), tc = tiles_per_xcd(p.nc), tv = tiles_per_xcd(p.nv);
    const double* u = y;
    const double* h = y + p.ne;
    hipLaunchKernelGGL((icon_edge_kernel<0>), dim3(grid1(p.ne)), dim3(kBlock), 0, s, p, u, h, te);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((icon_cell_vertex_kernel<0, false>), dim3(unsigned(kXcd * (tc + tv))), dim3(kBlock), 0, s, p,
                       u, r, tc, tv);
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
    const double ix = 2.0, ixr = L.kr == 4 ? 2.0 : 4.0, ixv = L.kv == 6 ? 2.0 : 4.0, c = 8.0,
                 rk = 8.0 + 16.0 + 8.0;
    std::vector<std::pair<const char*, double>> v;
    // E: RBF table, u, cells, c_lin_e, h, (K_e, F_e), v
    v.push_back({kKernelNames[K_ICON_EDGE], ne * (KR * (ixr + c) + 8.0 + 2 * ix + 2 * c + 16.0 + 8.0) + nc * 8.0});
    // C: cell edges, bilinear and divergence weights, (K_e, F_e), K, RK4 of h [u, D];
    //    vertex edges, curl, u, zeta
    v.push_back({kKernelNames[K_ICON_CELL_VERTEX], nc * (3 * ix + 6 * c + 8.0 + rk + (damp ? 8.0 : 0.0)) +
                                                       ne * (16.0 + 8.0) + nv * (KV * (ixv + c) + 8.0)});
    if (damp) {
        v.push_back({kKernelNames[K_ICON_DAMP_EDGE], ne * (2 * ix + c + 8.0) + nc * 8.0});
        v.push_back({kKernelNames[K_ICON_DAMP_CELL], nc * (3 * ix + 3 * c + 8.0) + ne * 8.0});
    }
    // M: cells, vertices, gradient, 1/d, f_e, v, K_e, K, h, zeta, RK4 of u [nu, D]
    v.push_back({kKernelNames[K_ICON_MOMENTUM], ne * (4 * ix + 2 * c + c + c + 8.0 + 8.0 + rk + (damp ? c : 0.0)) +
                                                    nc * (16.0 + (damp ? 8.0 : 0.0)) + nv * 8.0});
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
                {&H->su, L.ne * sizeof(double)},
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
    if (o.precision != TRISK_FP64) return set_err(TRISK_EINVAL, "the ICON triangle model runs in TRISK_FP64");
    if (o.block_size != 0 && o.block_size != kBlock)
        return set_err(TRISK_EINVAL, "block_size must be 0 or the library's workgroup size");
    auto* H = new trisk_handle;
    H->precision = TRISK_FP64;
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
