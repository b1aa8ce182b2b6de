// Adjoint of the TRiSK shallow-water tendency and of the RK4 step on the cell-pair path
// (trisk_tendency_vjp, trisk_rk4_adjoint in include/trisk_hip.h).
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
//   A3   edges: Gbar_e from its two sides; ubar_e = Gbar_e h_e + sum curl zbar + 2 u_e sum kin Bbar
//   A4   cells: hbar_i = g Bbar_i + sum_k Gbar_k lh_k u_k + sum_k vh_k hbar_vk
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
    const C* cvh;          // 6 x nc
    const int32_t* vslot;  // nv
    double* Bbar;          // nc
    I* Tbar;               // 6 x nc, cell-major: cell i's partial of Gbar at its edge k in Tbar[6 i + k]
    I* QC;                 // nc x 6  (cell, corner) partials of qbar
    I* zbar;               // nv
    I* ebar;               // nv      cotangent of h_v
    I* Gbar;               // ne
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
                                                          double c) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
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

// A2: vertices (the stage state recomputes zeta_v, h_v and q_v as the forward does)
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_vertex_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                            const double* __restrict__ h) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
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
    w.ebar[v] = I(-qb * qv / eta);
}

// A3: edges
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_edge_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                          const double* __restrict__ h, AdjRk r) {
    using C = typename P::C;
    using I = typename P::I;
    const int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x;
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
    w.Gbar[e] = I(Gb);
    adj_store(r, e, s, rin);
}

// A4: cells
template <class P, class UIn>
__global__ __launch_bounds__(kBlock) void adj_height_kernel(Params<P> p, AdjParams<P> w, const UIn* __restrict__ u,
                                                            AdjRk r) {
    using C = typename P::C;
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= p.nc) return;
    const AdjIn rin = adj_load(r, p.ne + i);
    const int32_t meta = ld(p.cmeta + i);
    const int n = meta & 15;
    double s = p.g * w.Bbar[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) {
            const int64_t e = ld(p.ce6 + k * p.nc + i);
            const int side = (meta >> (4 + k)) & 1;
            s += double(C(w.Gbar[e]) * ld(p.re_lh + 2 * e + side) * C(u[e]));
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k < n) s += double(ld(w.cvh + k * p.nc + i) * C(w.ebar[ld(p.ccv + k * p.nc + i)]));
    }
    adj_store(r, p.ne + i, s, rin);
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
    // work of trisk_rk4_adjoint, kept (and grown) across calls so a captured sweep stays valid
    std::vector<DevBuf> work;
    int64_t work_ck_states = 0, work_seg_states = 0;
    double *ckp = nullptr, *seg = nullptr, *lam = nullptr, *acc = nullptr, *kbar = nullptr, *run = nullptr;
    void* su[3] = {nullptr, nullptr, nullptr};
    double* sh[3] = {nullptr, nullptr, nullptr};
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
static int adj_setup(trisk_handle* H) {
    using C = typename P::C;
    using I = typename P::I;
    if (H->adj) return 0;
    auto* st = new AdjState;
    H->adj = st;
    H->adj_release = adj_release;
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
    rc |= coef(L.adj_cvh, &w.cvh);
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
    if ((rc = alloc(st->owned, L.nv * sizeof(I), &q))) return rc;
    w.ebar = static_cast<I*>(q);
    if ((rc = alloc(st->owned, L.ne * sizeof(I), &q))) return rc;
    w.Gbar = static_cast<I*>(q);
    return 0;
}

template <class P>
static const Params<P>& fwd_params(const trisk_handle* H) {
    if constexpr (P::id == TRISK_FP64)
        return H->p64;
    else
        return H->p32;
}

// One vector-Jacobian product at the stage (u, h) for the cotangent c * cot, ending in the
// epilogue r.
template <class P, class UIn>
static int vjp(trisk_handle* H, const UIn* u, const double* h, const double* cot, double c, const AdjRk& r,
               hipStream_t s) {
    const Layout& L = H->L;
    const Params<P>& p = fwd_params<P>(H);
    const AdjParams<P>& w = adj_params<P>(static_cast<AdjState*>(H->adj));
    int rc = pair_state_pass(H, u, std::is_same<UIn, double>::value, h, s);
    if (rc) return rc;
    const dim3 gc(unsigned((L.nc + kBlock - 1) / kBlock)), gv(unsigned((L.nv + kBlock - 1) / kBlock)),
        ge(unsigned((L.ne + kBlock - 1) / kBlock)), b(kBlock);
    hipLaunchKernelGGL((adj_cell_kernel<P>), gc, b, 0, s, p, w, cot, c);
    hipLaunchKernelGGL((adj_vertex_kernel<P, UIn>), gv, b, 0, s, p, w, u, h);
    hipLaunchKernelGGL((adj_edge_kernel<P, UIn>), ge, b, 0, s, p, w, u, h, r);
    hipLaunchKernelGGL((adj_height_kernel<P, UIn>), gc, b, 0, s, p, w, u, r);
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
        !(st->kbar = static_cast<double*>(get(sb))) || !(st->run = static_cast<double*>(get(sb))))
        return rc;
    for (int k = 0; k < 3; ++k)
        if (!(st->su[k] = get(L.ne * sizeof(U))) || !(st->sh[k] = static_cast<double*>(get(L.nc * sizeof(double)))))
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
    for (int64_t g = nck - 1; g >= 0; --g) {
        const int64_t first = g * ck, len = std::min(ck, N - first);
        HIP_TRY(hipMemcpyAsync(S0, C0 + g * n, sb, hipMemcpyDeviceToDevice, s));
        for (int64_t k = 1; k < len; ++k) {
            HIP_TRY(hipMemcpyAsync(S0 + k * n, S0 + (k - 1) * n, sb, hipMemcpyDeviceToDevice, s));
            if ((rc = rk4_steps_direct(H, S0 + k * n, dt, 1, s))) return rc;
        }
        for (int64_t k = len - 1; k >= 0; --k) {
            double* yn = S0 + k * n;
            // stage states s2, s3, s4 of the step from yn (the forward's own arithmetic)
            r.y = yn;
            for (int i = 0; i < 3; ++i) {
                r.mode = i + 1;
                r.su = st->su[i];
                r.sh = st->sh[i];
                const void* uin = i == 0 ? static_cast<const void*>(yn) : st->su[i - 1];
                const double* hin = i == 0 ? yn + L.ne : st->sh[i - 1];
                if ((rc = pair_tendency_stage(H, uin, i == 0, hin, r, s))) return rc;
            }
            a.mode = 4;
            if ((rc = vjp<P, U>(H, static_cast<const U*>(st->su[2]), st->sh[2], la, dt / 6, a, s))) return rc;
            a.mode = 3;
            if ((rc = vjp<P, U>(H, static_cast<const U*>(st->su[1]), st->sh[1], a.kbar, 1.0, a, s))) return rc;
            a.mode = 2;
            if ((rc = vjp<P, U>(H, static_cast<const U*>(st->su[0]), st->sh[0], a.kbar, 1.0, a, s))) return rc;
            a.mode = 1;
            if ((rc = vjp<P, double>(H, yn, yn + L.ne, a.kbar, 1.0, a, s))) return rc;
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
    if (!h->L.pair)
        return set_err(TRISK_EINVAL,
                       "the adjoint runs on the cell-pair path (TRiSK PV flux, cell kinetic energy; create the "
                       "model with TRISK_PATH_PAIR): " + h->L.pair_reason);
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
