// Batched ensembles of TRiSK shallow-water models on meshes of several resolutions
// (trisk_ensemble_* in include/trisk_hip.h).
//
// The reference advances one model per rk4_step! loop (RungeKutta.jl:18-28,
// examples/trisk_shallow_water.jl:92-100).  Here the members of a level share that level's
// tables, and every work vector is member-interleaved: member m of entity j sits at
// j * M + m.  A thread owns one (entity, member) item, so the M lanes of one entity read its
// table record once (a broadcast) and every neighbour gather reads M consecutive values.
// Each kernel of the cell-pair tendency (state, cell, edge: see trisk_device.hip) is one
// launch over all members of all levels; its blocks walk a per-XCD list of tiles built on
// the host (the schedule), optionally as a persistent grid.  The arithmetic is the
// single-model kernels' (trisk_device.hpp), so every member rounds exactly as a model of
// its own.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "trisk_device.hpp"

namespace trisk {

constexpr int kMaxLevels = 16;
constexpr int kEnsKinds = 3;  // state, cell, edge kernels
// tile lists: one per kernel, and the state kernel's without the edges (RK4 stages whose G the
// previous stage's edge kernel formed)
constexpr int kEnsLists = 4, kStateNoEdges = 3;

// n / d for n < 2^31 with one 32x32->64-bit product: with l = ceil(log2 d) and
// mul = ceil(2^(31+l) / d) (< 2^32), n * mul / 2^(31+l) = n/d + n e / (d 2^(31+l)) with
// 0 <= e < d <= 2^l, so the excess over n/d is below 1/d and the floor is exact.
struct FastDiv {
    uint32_t mul = 1u << 31;
    uint32_t shift = 31;
    __host__ __device__ uint32_t operator()(uint32_t n) const { return uint32_t((uint64_t(n) * mul) >> shift); }
};

inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    FastDiv f;
    f.shift = 31 + l;
    f.mul = uint32_t(((uint64_t(1) << f.shift) + d - 1) / d);
    return f;
}

// One level: the tables of its model, the member-interleaved work vectors and its dt.
template <class P>
struct EnsLevel {
    Params<P> p;       // tables; B, qv, G, T point to the level's (entity x member) vectors
    FastDiv md[4];     // division by M / V for V = 1, 2, 4, 8 members per lane
    int32_t M;         // members
    int64_t yoff;      // first element of the level's block [u (ne M); h (nc M)] of the state
    double dt, dt2, dt6;
    double* acc;       // (ne + nc) M  RK4 accumulator
    void* su;          // ne M         stage velocity (type P::U)
    double* sh;        // nc M         stage height
};

// A run of tiles of one range (cells, vertices or edges of one level) on one XCD.  The state
// kernel takes a level's vertex run backward, so that it starts where the previous launch (the
// edge kernel, forward through the same XCD's share) ended, and, in the first stage of a call,
// the edge run that follows starts where the vertex run ended.
constexpr int32_t kEnsBackward = 8;
struct EnsSeg {
    int32_t level, kind;  // kind: 0 cells, 1 vertices, 2 edges, 3 vertices and edges interleaved; | kEnsBackward
                          // for a run taken last tile first
    int32_t tile0;        // first tile of the run within its range
    int32_t ntiles;
    int32_t cum;          // tiles before the run in its XCD's list
    int32_t nitems;       // items (entity x member) of the range
};

struct EnsArgs {
    const void* levels;          // EnsLevel<P>[nlevels] in device memory
    const EnsSeg* seg;           // runs of all XCDs
    int32_t first[kXcd + 1];     // XCD x's runs: seg[first[x] .. first[x + 1])
    double* y;                   // ensemble state
    double* dy;                  // tendency output (mode 0)
    int32_t mode;                // Rk::mode
    int32_t from_y;              // stage input: the state itself (1) or the stage registers (0)
};

template <class P>
__device__ __forceinline__ Rk level_rk(const EnsLevel<P>& L, const EnsArgs& a) {
    Rk r;
    r.mode = a.mode;
    r.dy = a.dy ? a.dy + L.yoff : nullptr;
    r.y = a.y + L.yoff;
    r.acc = L.acc;
    r.su = L.su;
    r.sh = L.sh;
    r.dt = L.dt;
    r.dt2 = L.dt2;
    r.dt6 = L.dt6;
    return r;
}

// body(level, kind, item) for every item of the tiles of this block's XCD (blockIdx mod 8:
// the hardware deals workgroups to the XCDs round-robin); the XCD's blocks stride over its
// tile list, so a grid smaller than the list is a persistent launch.  A run of kind 3 holds
// the vertices and edges of one level interleaved along the Hilbert curve, in groups of five
// tiles per cell tile k: tiles 5k, 5k+1 are vertex tiles 2k, 2k+1 and tiles 5k+2..5k+4 edge
// tiles 3k..3k+2 (a level has fewer than 2 vertices and 3 edges per cell), so the blocks
// running together touch one stretch of the mesh; count(level, kind) gives the items of
// the level's cells (0), vertices (1) or edges (2).
// The level records and tile runs are written by the host before a launch and only read by
// the kernels: they are loaded through the constant address space, which lets the compiler
// use scalar loads for these block-uniform reads.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;

template <class P>
__device__ __forceinline__ EnsLevel<P> load_level(const void* levels, int level) {
    static_assert(sizeof(EnsLevel<P>) % 8 == 0 && alignof(EnsLevel<P>) >= 8, "levels are arrays of 8-byte words");
    constexpr int W = int(sizeof(EnsLevel<P>) / 8);
    cptr<uint64_t> src = (cptr<uint64_t>)levels + int64_t(level) * W;
    uint64_t w[W];
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] = src[k];
    EnsLevel<P> L;
    __builtin_memcpy(&L, w, sizeof(L));
    return L;
}

template <int BLK, class N, class F>
__device__ __forceinline__ void for_each_item(const EnsArgs& a, N&& count, F&& body) {
    const int x = int(blockIdx.x % kXcd);
    const int nbx = int(gridDim.x / kXcd);
    const int s0 = a.first[x], s1 = a.first[x + 1];
    if (s1 <= s0) return;
    const cptr<EnsSeg> seg = (cptr<EnsSeg>)a.seg;
    const int total = seg[s1 - 1].cum + seg[s1 - 1].ntiles;
    int s = s0;
    for (int t = int(blockIdx.x / kXcd); t < total; t += nbx) {
        while (s + 1 < s1 && t >= seg[s + 1].cum) ++s;  // t only grows
        EnsSeg g;
        g.level = seg[s].level;
        g.kind = seg[s].kind;
        g.tile0 = seg[s].tile0;
        g.ntiles = seg[s].ntiles;
        g.cum = seg[s].cum;
        g.nitems = seg[s].nitems;
        const bool back = (g.kind & kEnsBackward) != 0;
        int64_t tile = g.tile0 + (back ? g.ntiles - 1 - (t - g.cum) : t - g.cum);
        int kind = g.kind & ~kEnsBackward;
        int64_t nitems = g.nitems;
        if (kind == 3) {
            const int64_t sup = tile / 5, r = tile - 5 * sup;
            kind = r <= 1 ? 1 : 2;
            tile = kind == 1 ? 2 * sup + r : 3 * sup + (r - 2);
            nitems = count(g.level, kind);
        }
        const int64_t j = tile * BLK + threadIdx.x;
        if (j < nitems) body(g.level, kind, uint32_t(j));
    }
}

// V consecutive members at an address aligned to V * sizeof(T) (cached vector accesses)
template <class T, int V>
struct alignas(sizeof(T) * V) Lanes {
    T v[V];
};
// p points to global memory; element by element through a global pointer, which the compiler
// merges into global (not flat) vector accesses
template <class T, int V>
__device__ __forceinline__ Lanes<T, V> vload(const T* p) {
    const gptr<T> q = (gptr<T>)p;
    Lanes<T, V> x;
#pragma unroll
    for (int v = 0; v < V; ++v) x.v[v] = q[v];
    return x;
}
template <class T, int V>
__device__ __forceinline__ void vstore_nt(T* p, const Lanes<T, V>& x) {
    __attribute__((address_space(1))) T* q = (__attribute__((address_space(1))) T*)p;
#pragma unroll
    for (int v = 0; v < V; ++v) __builtin_nontemporal_store(x.v[v], q + v);
}
template <class T, int V>
__device__ __forceinline__ void vstore(T* p, const Lanes<T, V>& x) {
    __attribute__((address_space(1))) T* q = (__attribute__((address_space(1))) T*)p;
#pragma unroll
    for (int v = 0; v < V; ++v) q[v] = x.v[v];
}

// rk_load / rk_store (trisk_device.hpp) for V consecutive elements, with the same arithmetic
// per element: y and acc are loaded before the gathers; the accumulator, state and tendency
// are written with nontemporal stores, the stage vectors (gathered next) with plain ones
template <int V>
struct RkInLanes {
    Lanes<double, V> y, acc;
};

template <int V>
__device__ __forceinline__ RkInLanes<V> rk_load_lanes(const Rk& r, int64_t j) {
    RkInLanes<V> x;
#pragma unroll
    for (int v = 0; v < V; ++v) x.y.v[v] = x.acc.v[v] = 0.0;
    if (r.mode >= 1) x.y = vload<double, V>(r.y + j);
    if (r.mode >= 2) x.acc = vload<double, V>(r.acc + j);
    return x;
}

template <class U, int V>
__device__ __forceinline__ void rk_store_lanes(const Rk& r, int64_t j, const Lanes<double, V>& k,
                                               const RkInLanes<V>& x, U* stage) {
    Lanes<U, V> st;
    if (r.mode == 0) {
        vstore_nt(r.dy + j, k);
    } else if (r.mode <= 3) {
        Lanes<double, V> acc = k;
        if (r.mode > 1) {
#pragma unroll
            for (int v = 0; v < V; ++v) acc.v[v] = x.acc.v[v] + 2 * k.v[v];
        }
        vstore_nt(r.acc + j, acc);
        const double c = r.mode == 3 ? r.dt : r.dt2;
#pragma unroll
        for (int v = 0; v < V; ++v) st.v[v] = U(x.y.v[v] + c * k.v[v]);
        vstore(stage, st);
    } else {
        Lanes<double, V> y;
#pragma unroll
        for (int v = 0; v < V; ++v) y.v[v] = x.y.v[v] + r.dt6 * (x.acc.v[v] + k.v[v]);
        vstore_nt(r.y + j, y);
    }
}

template <int V>
constexpr int lane_log() {
    return V == 1 ? 0 : V == 2 ? 1 : V == 4 ? 2 : 3;
}

// Item j of a range owns members [m0, m0 + V) of entity ent, elements [j V, j V + V) of the
// range's member-interleaved vectors (j V = ent M + m0).
struct Item {
    uint32_t ent, m0, o;
};
template <int V, class P>
__device__ __forceinline__ Item item_of(const EnsLevel<P>& L, uint32_t j) {
    const uint32_t ent = L.md[lane_log<V>()](j);
    const uint32_t o = j * V;
    return {ent, o - ent * uint32_t(L.M), o};
}

// W1 of all levels: vertices (q_v), edges (G, in the first stage of a call); cf.
// pair_state_kernel.  B is formed by the cell kernel, which gathers u at the edges where it
// gathers G.
template <class P, class UIn, int BLK, int V>
__device__ __forceinline__ void ens_state_body(const EnsArgs& a) {
    using C = typename P::C;
    using I = typename P::I;
    constexpr bool NARROW = P::id == TRISK_FP32_BULK;
    auto count = [&](int level, int kind) -> int64_t {
        const EnsLevel<P> L = load_level<P>(a.levels, level);
        return (kind == 0 ? L.p.nc : kind == 1 ? L.p.nv : L.p.ne) * (L.M / V);
    };
    for_each_item<BLK>(a, count, [&](int level, int kind, uint32_t j) {
        const EnsLevel<P> L = load_level<P>(a.levels, level);
        const Params<P>& p = L.p;
        const uint32_t M = uint32_t(L.M);
        const Item it = item_of<V>(L, j);
        const UIn* u = a.from_y ? reinterpret_cast<const UIn*>(a.y + L.yoff) : static_cast<const UIn*>(L.su);
        const double* h = a.from_y ? a.y + L.yoff + p.ne * M : L.sh;
        if (kind == 1) {
            int32_t idx[6];
            C cf[8];
            load_vertex_record<NARROW>(p, it.ent, idx, cf);
            Lanes<UIn, V> uz[3];
            Lanes<double, V> hz[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                uz[k] = vload<UIn, V>(u + uint32_t(idx[k]) * M + it.m0);
                hz[k] = vload<double, V>(h + uint32_t(idx[3 + k]) * M + it.m0);
            }
            Lanes<I, V> q;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                C uu[3], hh[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    uu[k] = C(uz[k].v[v]);
                    hh[k] = C(hz[k].v[v]);
                }
                q.v[v] = I(pair_vertex_pv(cf, uu, hh));
            }
            vstore(p.qv + it.o, q);
        } else {
            int32_t c[2];
            C lh[2];
            if constexpr (NARROW) {
                float d;
                load_edge_record(p, it.ent, c, d, lh);
            } else {
                load_edge_cells<NARROW>(p, it.ent, c);
                ldrec(p.re_lh + 2 * int64_t(it.ent), lh);
            }
            const Lanes<double, V> h0 = vload<double, V>(h + uint32_t(c[0]) * M + it.m0);
            const Lanes<double, V> h1 = vload<double, V>(h + uint32_t(c[1]) * M + it.m0);
            const Lanes<UIn, V> ue = vload<UIn, V>(u + it.o);
            Lanes<I, V> G;
#pragma unroll
            for (int v = 0; v < V; ++v) G.v[v] = I(pair_mass_flux(lh, C(h0.v[v]), C(h1.v[v]), C(ue.v[v])));
            vstore(p.G + it.o, G);
        }
    });
}

// W2 of all levels: K and B = K + g h at the stage (u, h), the PV-flux half sums and the
// continuity with its RK4 update; cf. pair_cell_kernel.  Half sums are stored cell-major:
// member m of cell i's slot k at T[(6 i + k) M + m].
// One item of the cell kernel, from its cell's decoded table row (meta, edges, corners, pair
// weights, kinetic-energy weights).
template <class P, class UIn, int V>
__device__ __forceinline__ void ens_cell_item(const EnsArgs& a, const EnsLevel<P>& L, uint32_t j, int32_t meta,
                                              const int32_t (&E)[6], const int32_t (&V6)[6],
                                              const typename P::C (&W)[15], const typename P::C (&kin)[6],
                                              const RkInLanes<V>& rin) {
    using C = typename P::C;
    using I = typename P::I;
    const Params<P>& p = L.p;
    const uint32_t M = uint32_t(L.M);
    const Item it = item_of<V>(L, j);
    const Rk r = level_rk(L, a);
    // the stage heights: those of the state itself when the input is y (already loaded by
    // rk_load_lanes in RK4 modes; a tendency call, mode 0, loads them here)
    const double* hs = a.from_y ? a.y + L.yoff + p.ne * M : L.sh;
    const Lanes<double, V> hh = a.from_y && r.mode >= 1 ? rin.y : vload<double, V>(hs + it.o);
    const UIn* u = a.from_y ? reinterpret_cast<const UIn*>(a.y + L.yoff) : static_cast<const UIn*>(L.su);
    const int n = meta & 15;
    Lanes<I, V> Gk[6], Qk[6], Hk[6];
    Lanes<UIn, V> uk[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        Gk[k] = vload<I, V>(p.G + uint32_t(E[k]) * M + it.m0);
        Qk[k] = vload<I, V>(p.qv + uint32_t(V6[k]) * M + it.m0);
        uk[k] = vload<UIn, V>(u + uint32_t(E[k]) * M + it.m0);
    }
    const double area = ld(p.area_c + it.ent);
    Lanes<double, V> dh, B;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        C G[6], Q[6], H[6], ue[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            G[k] = C(Gk[k].v[v]);
            Q[k] = C(Qk[k].v[v]);
            ue[k] = C(uk[k].v[v]);
        }
        B.v[v] = double(pair_kinetic(kin, ue)) + p.g * hh.v[v];
        pair_half_sums(n, W, G, Q, H);
        double div = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            Hk[k].v[v] = I(H[k]);
            if (k < n) div += (meta >> (4 + k)) & 1 ? -double(G[k]) : double(G[k]);
        }
        dh.v[v] = -div / area;
    }
    vstore(p.B + it.o, B);
#pragma unroll
    for (int k = 0; k < 6; ++k) vstore(p.T + (6 * it.ent + uint32_t(k)) * M + it.m0, Hk[k]);
    rk_store_lanes<double, V>(r, p.ne * M + it.o, dh, rin, r.mode >= 1 && r.mode <= 3 ? L.sh + it.o : nullptr);
}

// W2 of all levels: K and B = K + g h at the stage (u, h), the PV-flux half sums and the
// continuity with its RK4 update; cf. pair_cell_kernel.  Half sums are stored cell-major:
// member m of cell i's slot k at T[(6 i + k) M + m].  FP32-bulk: the tiles walk as in
// for_each_item, and each stages the cell records of its items (76 B, the single model's
// rc_pair) through LDS with 16-byte loads, from a start rounded down to a multiple of 4 cells
// (16-byte aligned); its items then read their cell's record there.  FP64 reads the slot-major
// tables.
template <class P, class UIn, int BLK, int V>
__device__ __forceinline__ void ens_cell_body(const EnsArgs& a) {
    using C = typename P::C;
    constexpr bool NARROW = P::id == TRISK_FP32_BULK;
    if constexpr (!NARROW) {
        auto count = [](int, int) -> int64_t { return 0; };  // runs of one kind only
        for_each_item<BLK>(a, count, [&](int level, int, uint32_t j) {
            const EnsLevel<P> L = load_level<P>(a.levels, level);
            const Params<P>& p = L.p;
            const Item it = item_of<V>(L, j);
            const RkInLanes<V> rin = rk_load_lanes<V>(level_rk(L, a), p.ne * uint32_t(L.M) + it.o);
            int32_t meta, E[6], V6[6];
            C W[15], kin[6];
            load_cell_pairs<false>(p, it.ent, meta, E, V6, W);
#pragma unroll
            for (int k = 0; k < 6; ++k) kin[k] = k < p.kc ? ld(p.kin + k * p.nc + it.ent) : C(0);
            ens_cell_item<P, UIn, V>(a, L, j, meta, E, V6, W, kin, rin);
        });
    } else {
        constexpr int RW = kPairRecord;
        __shared__ __attribute__((aligned(16))) int32_t srec[(BLK + 4) * RW];
        const int x = int(blockIdx.x % kXcd);
        const int nbx = int(gridDim.x / kXcd);
        const int s0 = a.first[x], s1 = a.first[x + 1];
        if (s1 <= s0) return;  // the whole block
        const cptr<EnsSeg> seg = (cptr<EnsSeg>)a.seg;
        const int total = seg[s1 - 1].cum + seg[s1 - 1].ntiles;
        int s = s0;
        // t is the same for every thread of the block, so the barriers are reached by all
        for (int t = int(blockIdx.x / kXcd); t < total; t += nbx) {
            while (s + 1 < s1 && t >= seg[s + 1].cum) ++s;  // t only grows
            const int32_t kind = seg[s].kind;
            const int32_t ntiles = seg[s].ntiles, cum = seg[s].cum;
            const int64_t tile = seg[s].tile0 + ((kind & kEnsBackward) ? ntiles - 1 - (t - cum) : t - cum);
            const int64_t nitems = seg[s].nitems;
            const EnsLevel<P> L = load_level<P>(a.levels, seg[s].level);
            const Params<P>& p = L.p;
            const int64_t j0 = tile * BLK, jl = (j0 + BLK < nitems ? j0 + BLK : nitems) - 1;
            const int64_t e0 = int64_t(L.md[lane_log<V>()](uint32_t(j0))) & ~int64_t(3);
            const int64_t e1 = int64_t(L.md[lane_log<V>()](uint32_t(jl))) + 1;
            const int64_t j = j0 + threadIdx.x;
            const bool on = j <= jl;
            // the RK4 loads first: they do not wait for the records
            RkInLanes<V> rin{};
            if (on) rin = rk_load_lanes<V>(level_rk(L, a), p.ne * uint32_t(L.M) + uint32_t(j) * V);
            __syncthreads();  // the previous tile's records have been read
            const int nw = int((e1 - e0) * RW), nw4 = nw / 4;
            const int32_t* src = p.rc_pair + e0 * RW;
            for (int w = threadIdx.x; w < nw4; w += BLK)
                reinterpret_cast<f32x4*>(srec)[w] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src) + w);
            for (int w = 4 * nw4 + threadIdx.x; w < nw; w += BLK) srec[w] = ld(src + w);
            __syncthreads();
            if (!on) continue;
            const Item it = item_of<V>(L, uint32_t(j));
            const int32_t* rr = srec + RW * (int64_t(it.ent) - e0);
            const int32_t meta = rr[0];
            const bool wide = int16_t(rr[1] & 0xffff) == kWide;
            const int32_t be = int32_t(3 * it.ent), bv = int32_t(2 * it.ent);
            int32_t E[6], V6[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int32_t pk = rr[1 + k];
                E[k] = wide ? ld(p.ce6 + k * p.nc + it.ent) : be + int16_t(pk & 0xffff);
                V6[k] = wide ? ld(p.ccv + k * p.nc + it.ent) : bv + int16_t(uint32_t(pk) >> 16);
            }
            C R[6], kin[6], W[15];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                __builtin_memcpy(&R[k], &rr[7 + k], 4);
                __builtin_memcpy(&kin[k], &rr[13 + k], 4);
            }
            pair_weights(meta, R, W);
            ens_cell_item<P, UIn, V>(a, L, uint32_t(j), meta, E, V6, W, kin, rin);
        }
    }
}

// W3 of all levels: du and its RK4 update; cf. pair_edge_kernel.  GNEXT (RK4 stages): also G of
// the state the stage writes, so that the next stage's state kernel skips the edges.
template <class P, int BLK, int V, bool GNEXT>
__device__ __forceinline__ void ens_edge_body(const EnsArgs& a) {
    using C = typename P::C;
    using I = typename P::I;
    using U = typename P::U;
    constexpr bool NARROW = P::id == TRISK_FP32_BULK;
    auto count = [](int, int) -> int64_t { return 0; };  // runs of one kind only
    for_each_item<BLK>(a, count, [&](int level, int, uint32_t j) {
        const EnsLevel<P> L = load_level<P>(a.levels, level);
        const Params<P>& p = L.p;
        const uint32_t M = uint32_t(L.M);
        const Item it = item_of<V>(L, j);
        const Rk r = level_rk(L, a);
        const RkInLanes<V> rin = rk_load_lanes<V>(r, it.o);
        // the edge's own streamed values first, so that they do not wait for the cell record
        const uint32_t sl = ld(p.eslot + it.ent);
        C d;
        C lh[2] = {C(0), C(0)};
        int32_t c[2];
        if constexpr (NARROW) {
            load_edge_record(p, it.ent, c, d, lh);  // one 16-byte load: cells, d, l h
        } else {
            d = ld(p.dual + it.ent);
            if constexpr (GNEXT) ldrec(p.re_lh + 2 * int64_t(it.ent), lh);
            load_edge_cells<NARROW>(p, it.ent, c);
        }
        const Lanes<I, V> t0 = vload<I, V>(p.T + (6 * uint32_t(c[0]) + (sl & 15)) * M + it.m0);
        const Lanes<I, V> t1 = vload<I, V>(p.T + (6 * uint32_t(c[1]) + (sl >> 4)) * M + it.m0);
        const Lanes<double, V> bl = vload<double, V>(p.B + uint32_t(c[0]) * M + it.m0);
        const Lanes<double, V> br = vload<double, V>(p.B + uint32_t(c[1]) * M + it.m0);
        Lanes<double, V> hn0, hn1;
        if constexpr (GNEXT) {
            // the heights of the state this stage writes: the stage heights, or y after the last stage
            const double* hn = r.mode == 4 ? r.y + p.ne * M : L.sh;
            hn0 = vload<double, V>(hn + uint32_t(c[0]) * M + it.m0);
            hn1 = vload<double, V>(hn + uint32_t(c[1]) * M + it.m0);
        }
        Lanes<double, V> du;
#pragma unroll
        for (int v = 0; v < V; ++v) du.v[v] = pair_edge_tendency(C(t0.v[v]), C(t1.v[v]), d, bl.v[v], br.v[v]);
        rk_store_lanes<U, V>(r, it.o, du, rin, r.mode >= 1 && r.mode <= 3 ? static_cast<U*>(L.su) + it.o : nullptr);
        if constexpr (GNEXT) {
            // the new values exactly as rk_store_lanes forms them
            Lanes<I, V> G;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double un = r.mode <= 3 ? rin.y.v[v] + (r.mode == 3 ? r.dt : r.dt2) * du.v[v]
                                              : rin.y.v[v] + r.dt6 * (rin.acc.v[v] + du.v[v]);
                G.v[v] = I(pair_mass_flux(lh, C(hn0.v[v]), C(hn1.v[v]), C(un)));
            }
            vstore(p.G + it.o, G);
        }
    });
}

template <class P, class UIn, int BLK, int V>
__global__ __launch_bounds__(BLK) void ens_state_kernel(EnsArgs a) {
    ens_state_body<P, UIn, BLK, V>(a);
}
template <class P, class UIn, int BLK, int V>
__global__ __launch_bounds__(BLK) void ens_cell_kernel(EnsArgs a) {
    ens_cell_body<P, UIn, BLK, V>(a);
}
template <class P, int BLK, int V, bool GNEXT>
__global__ __launch_bounds__(BLK) void ens_edge_kernel(EnsArgs a) {
    ens_edge_body<P, BLK, V, GNEXT>(a);
}

// One RK stage of every member in one cooperative launch: the state, cell and edge phases of
// the kernels above, with a grid-wide barrier between them, since each phase gathers what the
// previous one wrote anywhere on the mesh.  The launch is cooperative (hipLaunchCooperativeKernel),
// so every workgroup of the grid is resident at once and the barrier is reached by all of
// them; the grid (a multiple of the XCD count) is sized by the occupancy of this kernel.  The
// phases walk their own tile lists as persistent loops (kFused: block size and members per
// work-item of each phase).
// The grid-wide barrier costs in proportion to the workgroups that reach it (about 31, 75 and
// 144 us per barrier for 256, 1024 and 2048 workgroups of 256 on MI355X), so the stage is
// slower than three launches replayed from a graph: 0.36 against 0.19 ms per stage for the
// 64-member ensemble.  Four members per work-item in the state and cell phases hold the grid
// at two workgroups per CU (the cell phase's registers), which measured faster than two
// members per work-item at four per CU (0.47 ms).
constexpr int kFusedBlock = 256, kFusedV0 = 4, kFusedV1 = 4, kFusedV2 = 2;
template <class P, class UIn, bool GNEXT>
__global__ __launch_bounds__(kFusedBlock) void ens_stage_kernel(EnsArgs as, EnsArgs ac, EnsArgs ae) {
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    ens_state_body<P, UIn, kFusedBlock, kFusedV0>(as);
    g.sync();
    ens_cell_body<P, UIn, kFusedBlock, kFusedV1>(ac);
    g.sync();
    ens_edge_body<P, kFusedBlock, kFusedV2, GNEXT>(ae);
}

// member m of a level block <-> a single-member state (dir 0: get, 1: set)
__global__ __launch_bounds__(256) void ens_member_kernel(double* __restrict__ block, double* __restrict__ one,
                                                         int64_t ne, int64_t n, int32_t M, int32_t m, int32_t dir) {
    const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (k >= n) return;
    const int64_t at = k < ne ? k * M + m : ne * M + (k - ne) * M + m;
    if (dir == 0)
        one[k] = block[at];
    else
        block[at] = one[k];
}

// timing kinds: the three kernels, and the fused stage
constexpr int kEnsFused = kEnsKinds, kEnsTimed = kEnsKinds + 1;
static const char* kEnsNames[kEnsTimed] = {"ens_state_kernel", "ens_cell_kernel", "ens_edge_kernel",
                                           "ens_stage_kernel"};

}  // namespace trisk

using namespace trisk;

struct trisk_ensemble {
    int precision = TRISK_FP64;
    int device = 0;
    int cus = 256;
    int nlevels = 0;
    std::vector<trisk_handle*> models;  // one per level: its tables (device) and numbering
    std::vector<int32_t> M;
    std::vector<int64_t> yoff;
    int64_t n_state = 0;
    std::vector<DevBuf> owned;
    std::vector<EnsLevel<Fp64>> l64;
    std::vector<EnsLevel<Fp32Bulk>> l32;
    void* levels_dev = nullptr;
    std::vector<double> dts;
    // schedules of the three kernels
    trisk_schedule sched[kEnsKinds]{};
    EnsSeg* seg_dev[kEnsLists]{};
    size_t seg_cap[kEnsLists]{};
    int32_t first[kEnsLists][kXcd + 1]{};
    int64_t grid[kEnsLists]{};
    double* scratch = nullptr;  // one member of the largest level
    // the fused stage (trisk_ensemble_set_fused): one cooperative launch per RK stage, with its
    // own tile lists and grid
    bool fused = false;
    EnsSeg* fseg_dev[kEnsLists]{};
    size_t fseg_cap[kEnsLists]{};
    int32_t ffirst[kEnsLists][kXcd + 1]{};
    int64_t fgrid = 0;
    // graph replay
    bool graphs = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    double* graph_y = nullptr;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> events;
    std::vector<int> event_kernel;
    size_t events_used = 0;
};

namespace trisk {

static int lanes_of(const trisk_schedule& s) { return s.members_per_lane > 0 ? s.members_per_lane : 1; }

static bool valid_schedule(const trisk_ensemble* E, const trisk_schedule& s) {
    const int V = lanes_of(s);
    bool ok = (s.block_size == 128 || s.block_size == 256 || s.block_size == 512) && s.blocks_per_cu >= 0 &&
              s.blocks_per_cu <= 64 &&
              (s.split == TRISK_SPLIT_CHUNKED || s.split == TRISK_SPLIT_BLOCKED || s.split == TRISK_SPLIT_INTERLEAVED) &&
              (V == 1 || V == 2 || V == 4 || V == 8);
    for (int l = 0; ok && l < E->nlevels; ++l) ok = E->M[l] % V == 0;
    return ok;
}

// State vectors must be aligned to the widest member group any kernel loads.
static int check_alignment(const trisk_ensemble* E, const void* p) {
    int V = E->fused ? std::max(kFusedV0, std::max(kFusedV1, kFusedV2)) : 1;
    for (const auto& s : E->sched) V = std::max(V, lanes_of(s));
    if (reinterpret_cast<uintptr_t>(p) % (8 * V) != 0)
        return set_err(TRISK_EINVAL, "ensemble state vectors must be aligned to " + std::to_string(8 * V) +
                                         " bytes for this schedule");
    return 0;
}

// Where build_list_to writes a tile list: its device copy (grown as needed), the first run of
// every XCD and the grid that walks it.
struct ListOut {
    EnsSeg** dev;
    size_t* cap;
    int32_t* first;  // kXcd + 1
    int64_t* grid;
};

// The runs of tile list `list` per XCD under schedule `sc`.
static int build_list_to(trisk_ensemble* E, int list, const trisk_schedule& sc, const ListOut& out) {
    const int kind = list == kStateNoEdges ? TRISK_ENS_STATE : list;
    const int64_t blk = sc.block_size;
    const int64_t V = lanes_of(sc);
    struct Range {
        int32_t level, kind;
        int64_t items, tiles;
    };
    std::vector<Range> ranges;
    for (int l = 0; l < E->nlevels; ++l) {
        const Layout& L = E->models[l]->L;
        const int64_t M = E->M[l] / V;  // items per entity
        if (kind == TRISK_ENS_STATE) {
            ranges.push_back({l, 1, L.nv * M, 0});
            if (list != kStateNoEdges) ranges.push_back({l, 2, L.ne * M, 0});
        } else if (kind == TRISK_ENS_CELL) {
            ranges.push_back({l, 0, L.nc * M, 0});
        } else {
            ranges.push_back({l, 2, L.ne * M, 0});
        }
    }
    int64_t all = 0;
    for (auto& r : ranges) {
        r.tiles = (r.items + blk - 1) / blk;
        all += r.tiles;
    }
    std::vector<std::vector<EnsSeg>> per(kXcd);
    auto add = [&](int x, const Range& r, int64_t lo, int64_t hi) {
        if (lo >= hi) return;
        per[x].push_back({r.level, r.kind, int32_t(lo), int32_t(hi - lo), 0, int32_t(r.items)});
    };
    // (the interleaved groups hold edge tiles: the list without edges takes the chunked split)
    if (sc.split == TRISK_SPLIT_INTERLEAVED && list == TRISK_ENS_STATE) {
        // XCD x: the x-th contiguous chunk of every level's interleaved vertices and edges, in
        // whole groups of 5 tiles (the vertex and edge tiles of one stretch of cell tile size)
        ranges.clear();
        for (int l = 0; l < E->nlevels; ++l) {
            const int64_t items = E->models[l]->L.nc * (E->M[l] / V);
            ranges.push_back({l, 3, items, 5 * ((items + blk - 1) / blk)});
        }
        for (int x = 0; x < kXcd; ++x)
            for (const auto& r : ranges) {
                const int64_t groups = r.tiles / 5, c = (groups + kXcd - 1) / kXcd;
                add(x, r, 5 * std::min(groups, x * c), 5 * std::min(groups, (x + 1) * c));
            }
    } else if (sc.split != TRISK_SPLIT_BLOCKED) {
        // XCD x: the x-th contiguous chunk of every range (a level's vertices backward)
        for (int x = 0; x < kXcd; ++x)
            for (const auto& r : ranges) {
                const int64_t c = (r.tiles + kXcd - 1) / kXcd;
                add(x, r, std::min(r.tiles, x * c), std::min(r.tiles, (x + 1) * c));
                if (r.kind == 1 && !per[x].empty() && per[x].back().kind == 1) per[x].back().kind |= kEnsBackward;
            }
    } else {
        // XCD x: the x-th contiguous chunk of the ranges laid end to end
        const int64_t c = (all + kXcd - 1) / kXcd;
        for (int x = 0; x < kXcd; ++x) {
            const int64_t lo = x * c, hi = std::min(all, (x + 1) * c);
            int64_t base = 0;
            for (const auto& r : ranges) {
                add(x, r, std::max<int64_t>(lo - base, 0), std::min(r.tiles, hi - base));
                base += r.tiles;
            }
        }
    }
    std::vector<EnsSeg> segs;
    int64_t most = 0;
    for (int x = 0; x < kXcd; ++x) {
        out.first[x] = int32_t(segs.size());
        int64_t cum = 0;
        for (auto g : per[x]) {
            g.cum = int32_t(cum);
            cum += g.ntiles;
            segs.push_back(g);
        }
        most = std::max(most, cum);
    }
    out.first[kXcd] = int32_t(segs.size());
    int64_t nbx = most;
    if (sc.blocks_per_cu > 0) nbx = std::min<int64_t>(most, int64_t(sc.blocks_per_cu) * ((E->cus + kXcd - 1) / kXcd));
    *out.grid = kXcd * std::max<int64_t>(nbx, 1);
    if (segs.empty()) segs.push_back({0, 0, 0, 0, 0, 0});
    if (segs.size() > *out.cap) {
        void* q;
        int rc = alloc(E->owned, segs.size() * sizeof(EnsSeg), &q);
        if (rc) return rc;
        *out.dev = static_cast<EnsSeg*>(q);
        *out.cap = segs.size();
    }
    HIP_TRY(hipMemcpy(*out.dev, segs.data(), segs.size() * sizeof(EnsSeg), hipMemcpyHostToDevice));
    return 0;
}

static int build_list(trisk_ensemble* E, int list) {
    const int kind = list == kStateNoEdges ? TRISK_ENS_STATE : list;
    return build_list_to(E, list, E->sched[kind],
                         ListOut{&E->seg_dev[list], &E->seg_cap[list], E->first[list], &E->grid[list]});
}

// The tile lists of kernel `kind` (the state kernel has two).
static int build_schedule(trisk_ensemble* E, int kind) {
    int rc = build_list(E, kind);
    if (!rc && kind == TRISK_ENS_STATE) rc = build_list(E, kStateNoEdges);
    return rc;
}

template <class P>
static std::vector<EnsLevel<P>>& ens_levels(trisk_ensemble* E);
template <>
std::vector<EnsLevel<Fp64>>& ens_levels<Fp64>(trisk_ensemble* E) {
    return E->l64;
}
template <>
std::vector<EnsLevel<Fp32Bulk>>& ens_levels<Fp32Bulk>(trisk_ensemble* E) {
    return E->l32;
}

template <class P>
static Params<P>& model_params(trisk_handle* h) {
    if constexpr (P::id == TRISK_FP64)
        return h->p64;
    else
        return h->p32;
}

template <class P>
static int setup_levels(trisk_ensemble* E) {
    using I = typename P::I;
    using U = typename P::U;
    auto& lv = ens_levels<P>(E);
    lv.resize(E->nlevels);
    int64_t biggest = 0;
    for (int l = 0; l < E->nlevels; ++l) {
        const Layout& L = E->models[l]->L;
        const int64_t M = E->M[l];
        EnsLevel<P>& e = lv[l];
        e.p = model_params<P>(E->models[l]);
        for (int k = 0; k < 4; ++k) e.md[k] = make_fastdiv(uint32_t(M % (1 << k) == 0 ? M >> k : 1));
        e.M = int32_t(M);
        e.yoff = E->yoff[l];
        e.dt = e.dt2 = e.dt6 = 0;
        void* q;
        int rc;
        auto get = [&](size_t bytes) -> void* {
            rc = alloc(E->owned, bytes, &q);
            return rc ? nullptr : q;
        };
        e.p.B = static_cast<double*>(get(L.nc * M * sizeof(double)));
        if (rc) return rc;
        e.p.qv = static_cast<I*>(get(L.nv * M * sizeof(I)));
        if (rc) return rc;
        e.p.G = static_cast<I*>(get(L.ne * M * sizeof(I)));
        if (rc) return rc;
        e.p.T = static_cast<I*>(get(6 * L.nc * M * sizeof(I)));
        if (rc) return rc;
        e.acc = static_cast<double*>(get((L.ne + L.nc) * M * sizeof(double)));
        if (rc) return rc;
        e.su = get(L.ne * M * sizeof(U));
        if (rc) return rc;
        e.sh = static_cast<double*>(get(L.nc * M * sizeof(double)));
        if (rc) return rc;
        e.p.K = nullptr;
        e.p.Fq = nullptr;
        e.p.zv = e.p.hv = e.p.Kv = nullptr;
        e.p.K_out = e.p.qv_out = e.p.zv_out = e.p.hv_out = nullptr;
        biggest = std::max(biggest, L.ne + L.nc);
    }
    void* q;
    int rc = alloc(E->owned, E->nlevels * sizeof(EnsLevel<P>), &q);
    if (rc) return rc;
    E->levels_dev = q;
    if ((rc = alloc(E->owned, biggest * sizeof(double), &q))) return rc;
    E->scratch = static_cast<double*>(q);
    HIP_TRY(hipMemcpy(E->levels_dev, lv.data(), lv.size() * sizeof(EnsLevel<P>), hipMemcpyHostToDevice));
    return 0;
}

// Per-level dt into the device copy of the levels (only when they change; the stream is
// drained first because captured graphs read them).
template <class P>
static int set_dts(trisk_ensemble* E, const double* dt, hipStream_t s) {
    bool same = true;
    for (int l = 0; l < E->nlevels; ++l) same = same && E->dts[l] == dt[l];
    if (same) return 0;
    auto& lv = ens_levels<P>(E);
    for (int l = 0; l < E->nlevels; ++l) {
        E->dts[l] = dt[l];
        lv[l].dt = dt[l];
        lv[l].dt2 = dt[l] / 2;
        lv[l].dt6 = dt[l] / 6;
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(E->levels_dev, lv.data(), lv.size() * sizeof(EnsLevel<P>), hipMemcpyHostToDevice));
    return 0;
}

static int ens_timed_begin(trisk_ensemble* E, int kind, hipStream_t s, hipEvent_t* stop) {
    *stop = nullptr;
    if (!E->timing) return 0;
    if (E->events_used + 2 > E->events.size()) {
        for (int k = 0; k < 256; ++k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            E->events.push_back(e);
        }
        E->event_kernel.resize(E->events.size() / 2);
    }
    size_t i = E->events_used;
    E->events_used += 2;
    E->event_kernel[i / 2] = kind;
    HIP_TRY(hipEventRecord(E->events[i], s));
    *stop = E->events[i + 1];
    return 0;
}

template <class P, class UIn, int BLK, int V>
static void launch_kind(int kind, bool gnext, dim3 grid, const EnsArgs& a, hipStream_t s) {
    if (kind == TRISK_ENS_STATE)
        hipLaunchKernelGGL((ens_state_kernel<P, UIn, BLK, V>), grid, dim3(BLK), 0, s, a);
    else if (kind == TRISK_ENS_CELL)
        hipLaunchKernelGGL((ens_cell_kernel<P, UIn, BLK, V>), grid, dim3(BLK), 0, s, a);
    else if (gnext)
        hipLaunchKernelGGL((ens_edge_kernel<P, BLK, V, true>), grid, dim3(BLK), 0, s, a);
    else
        hipLaunchKernelGGL((ens_edge_kernel<P, BLK, V, false>), grid, dim3(BLK), 0, s, a);
}

template <class P, class UIn, int BLK>
static void launch_lanes(int kind, bool gnext, int V, dim3 grid, const EnsArgs& a, hipStream_t s) {
    if (V == 8)
        launch_kind<P, UIn, BLK, 8>(kind, gnext, grid, a, s);
    else if (V == 4)
        launch_kind<P, UIn, BLK, 4>(kind, gnext, grid, a, s);
    else if (V == 2)
        launch_kind<P, UIn, BLK, 2>(kind, gnext, grid, a, s);
    else
        launch_kind<P, UIn, BLK, 1>(kind, gnext, grid, a, s);
}

// One tendency of every member: the three kernels with the given stage input and RK4 mode.
// state_edges false: G is there already, formed by the previous stage's edge kernel; gnext: the
// edge kernel forms G of the state it writes.
template <class P, class UIn>
static int ens_tendency(trisk_ensemble* E, double* y, double* dy, int mode, bool from_y, hipStream_t s,
                        bool state_edges = true, bool gnext = false) {
    EnsArgs a{};
    a.levels = E->levels_dev;
    a.y = y;
    a.dy = dy;
    a.mode = mode;
    a.from_y = from_y ? 1 : 0;
    if (E->fused) {
        EnsArgs as = a, ac = a, ae = a;
        const int ls = state_edges ? int(TRISK_ENS_STATE) : kStateNoEdges;
        as.seg = E->fseg_dev[ls];
        std::memcpy(as.first, E->ffirst[ls], sizeof(as.first));
        ac.seg = E->fseg_dev[TRISK_ENS_CELL];
        std::memcpy(ac.first, E->ffirst[TRISK_ENS_CELL], sizeof(ac.first));
        ae.seg = E->fseg_dev[TRISK_ENS_EDGE];
        std::memcpy(ae.first, E->ffirst[TRISK_ENS_EDGE], sizeof(ae.first));
        hipEvent_t stop;
        int rc = ens_timed_begin(E, kEnsFused, s, &stop);
        if (rc) return rc;
        void* args[] = {&as, &ac, &ae};
        const void* f = gnext ? reinterpret_cast<const void*>(&ens_stage_kernel<P, UIn, true>)
                              : reinterpret_cast<const void*>(&ens_stage_kernel<P, UIn, false>);
        HIP_TRY(hipLaunchCooperativeKernel(f, dim3(unsigned(E->fgrid)), dim3(kFusedBlock), args, 0, s));
        if (stop) HIP_TRY(hipEventRecord(stop, s));
        return 0;
    }
    for (int kind = 0; kind < kEnsKinds; ++kind) {
        const int list = kind == TRISK_ENS_STATE && !state_edges ? kStateNoEdges : kind;
        a.seg = E->seg_dev[list];
        std::memcpy(a.first, E->first[list], sizeof(a.first));
        hipEvent_t stop;
        int rc = ens_timed_begin(E, kind, s, &stop);
        if (rc) return rc;
        const dim3 grid(unsigned(E->grid[list]));
        const int V = lanes_of(E->sched[kind]);
        switch (E->sched[kind].block_size) {
            case 128:
                launch_lanes<P, UIn, 128>(kind, gnext, V, grid, a, s);
                break;
            case 512:
                launch_lanes<P, UIn, 512>(kind, gnext, V, grid, a, s);
                break;
            default:
                launch_lanes<P, UIn, 256>(kind, gnext, V, grid, a, s);
                break;
        }
        HIP_TRY(hipGetLastError());
        if (stop) HIP_TRY(hipEventRecord(stop, s));
    }
    return 0;
}

template <class P>
static int ens_steps(trisk_ensemble* E, double* y, int64_t nsteps, hipStream_t s) {
    using U = typename P::U;
    int rc;
    // each stage's edge kernel forms G of the state it writes: only the first stage of the
    // call runs the state kernel over the edges
    for (int64_t k = 0; k < nsteps; ++k) {
        if ((rc = ens_tendency<P, double>(E, y, nullptr, 1, true, s, k == 0, true))) return rc;
        for (int mode = 2; mode <= 4; ++mode)
            if ((rc = ens_tendency<P, U>(E, y, nullptr, mode, false, s, false, !(mode == 4 && k + 1 == nsteps))))
                return rc;
    }
    return 0;
}

static int run_ens_steps(trisk_ensemble* E, double* y, int64_t n, hipStream_t s) {
    return E->precision == TRISK_FP64 ? ens_steps<Fp64>(E, y, n, s) : ens_steps<Fp32Bulk>(E, y, n, s);
}

static void ens_drop_graph(trisk_ensemble* E) {
    if (E->graph_exec) (void)hipGraphExecDestroy(E->graph_exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->graph_exec = nullptr;
    E->graph = nullptr;
    E->graph_y = nullptr;
}

constexpr int64_t kEnsGraphSteps = 8;

static int ens_member_check(trisk_ensemble* E, int32_t level, int32_t member) {
    if (level < 0 || level >= E->nlevels) return set_err(TRISK_EINVAL, "level out of range");
    if (member < 0 || member >= E->M[level]) return set_err(TRISK_EINVAL, "member out of range");
    return 0;
}

static int ens_member_copy(trisk_ensemble* E, int32_t level, int32_t member, double* y, double* one, int dir,
                           hipStream_t s) {
    const Layout& L = E->models[level]->L;
    const int64_t n = L.ne + L.nc;
    hipLaunchKernelGGL(ens_member_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, y + E->yoff[level], one,
                       L.ne, n, E->M[level], member, dir);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace trisk

#define CHECK_ENS(e) \
    if (!(e)) return set_err(TRISK_EINVAL, "null ensemble")

extern "C" {

int trisk_ensemble_destroy(trisk_ensemble* e) {
    if (!e) return 0;
    (void)hipSetDevice(e->device);
    ens_drop_graph(e);
    for (auto& ev : e->events) (void)hipEventDestroy(ev);
    for (auto& b : e->owned) (void)hipFree(b.p);
    for (auto* h : e->models) trisk_destroy(h);
    delete e;
    return 0;
}

int trisk_ensemble_create(const trisk_model_desc* const* levels, const int32_t* members, int32_t nlevels,
                          const trisk_options* opt, trisk_ensemble** out) {
    if (!out) return set_err(TRISK_EINVAL, "null output pointer");
    *out = nullptr;
    if (!levels || !members || nlevels < 1 || nlevels > kMaxLevels)
        return set_err(TRISK_EINVAL, "need 1 to 16 levels with their member counts");
    trisk_options o{};
    if (opt) o = *opt;
    if (o.path != TRISK_PATH_AUTO && o.path != TRISK_PATH_PAIR)
        return set_err(TRISK_EINVAL, "ensembles run the cell-pair path (TRISK_PATH_AUTO or TRISK_PATH_PAIR)");
    o.path = TRISK_PATH_PAIR;
    auto* E = new trisk_ensemble;
    E->precision = o.precision;
    E->device = o.device;
    E->nlevels = nlevels;
    for (int l = 0; l < nlevels; ++l) {
        if (members[l] < 1) {
            trisk_ensemble_destroy(E);
            return set_err(TRISK_EINVAL, "every level needs at least one member");
        }
        trisk_handle* h = nullptr;
        int rc = trisk_create(levels[l], &o, &h);
        if (rc) {
            std::string why = trisk_last_error();
            trisk_ensemble_destroy(E);
            return set_err(rc, "level " + std::to_string(l) + ": " + why);
        }
        E->models.push_back(h);
        const int64_t items = (h->L.ne + h->L.nc) * int64_t(members[l]);
        if (2 * h->L.ne * int64_t(members[l]) >= (int64_t(1) << 31) ||
            6 * h->L.nc * int64_t(members[l]) >= (int64_t(1) << 31) || items >= (int64_t(1) << 31)) {
            trisk_ensemble_destroy(E);
            return set_err(TRISK_EINVAL, "level " + std::to_string(l) + ": entities x members exceed 2^30");
        }
        E->M.push_back(members[l]);
        E->yoff.push_back(E->n_state);
        E->n_state += items;
    }
    E->dts.assign(nlevels, 0.0);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, E->device) == hipSuccess && cus > 0)
        E->cus = cus;
    int rc = E->precision == TRISK_FP64 ? setup_levels<Fp64>(E) : setup_levels<Fp32Bulk>(E);
    for (int k = 0; k < kEnsKinds && !rc; ++k) {
        E->sched[k] = trisk_schedule{256, 0, TRISK_SPLIT_CHUNKED, 1};
        rc = build_schedule(E, k);
    }
    if (rc) {
        std::string why = trisk_last_error();
        trisk_ensemble_destroy(E);
        return set_err(rc, why);
    }
    *out = E;
    return 0;
}

int trisk_ensemble_state_length(const trisk_ensemble* e, int64_t* n) {
    CHECK_ENS(e);
    if (!n) return set_err(TRISK_EINVAL, "null pointer");
    *n = e->n_state;
    return 0;
}

int trisk_ensemble_level(trisk_ensemble* e, int32_t level, trisk_handle** model, int64_t* offset,
                         int32_t* members) {
    CHECK_ENS(e);
    if (level < 0 || level >= e->nlevels) return set_err(TRISK_EINVAL, "level out of range");
    if (model) *model = e->models[level];
    if (offset) *offset = e->yoff[level];
    if (members) *members = e->M[level];
    return 0;
}

int trisk_ensemble_get_member(trisk_ensemble* e, int32_t level, int32_t member, const double* y_dev,
                              double* y_member_dev, void* stream) {
    CHECK_ENS(e);
    int rc = ens_member_check(e, level, member);
    if (rc) return rc;
    if (!y_dev || !y_member_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(e->device));
    return ens_member_copy(e, level, member, const_cast<double*>(y_dev), y_member_dev, 0,
                           static_cast<hipStream_t>(stream));
}

int trisk_ensemble_set_member(trisk_ensemble* e, int32_t level, int32_t member, const double* y_member_dev,
                              double* y_dev, void* stream) {
    CHECK_ENS(e);
    int rc = ens_member_check(e, level, member);
    if (rc) return rc;
    if (!y_dev || !y_member_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(e->device));
    return ens_member_copy(e, level, member, y_dev, const_cast<double*>(y_member_dev), 1,
                           static_cast<hipStream_t>(stream));
}

int trisk_ensemble_tendency(trisk_ensemble* e, const double* y_dev, double* dy_dev, void* stream) {
    CHECK_ENS(e);
    if (!y_dev || !dy_dev) return set_err(TRISK_EINVAL, "null state pointer");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    e->events_used = 0;
    int rc = check_alignment(e, y_dev);
    if (!rc) rc = check_alignment(e, dy_dev);
    if (rc) return rc;
    double* y = const_cast<double*>(y_dev);  // mode 0 only reads the state
    return e->precision == TRISK_FP64 ? ens_tendency<Fp64, double>(e, y, dy_dev, 0, true, s)
                                      : ens_tendency<Fp32Bulk, double>(e, y, dy_dev, 0, true, s);
}

int trisk_ensemble_rk4_steps(trisk_ensemble* e, double* y_dev, const double* dt, int64_t nsteps, void* stream) {
    CHECK_ENS(e);
    if (!y_dev || !dt) return set_err(TRISK_EINVAL, "null pointer");
    if (nsteps < 0) return set_err(TRISK_EINVAL, "nsteps must be non-negative");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = check_alignment(e, y_dev);
    if (rc) return rc;
    rc = e->precision == TRISK_FP64 ? set_dts<Fp64>(e, dt, s) : set_dts<Fp32Bulk>(e, dt, s);
    if (rc) return rc;
    if (e->timing) {
        e->events_used = 0;
        return run_ens_steps(e, y_dev, nsteps, s);
    }
    // the fused stage runs by direct cooperative launches (four per step)
    if (!e->graphs || e->fused || s == nullptr || nsteps < kEnsGraphSteps) return run_ens_steps(e, y_dev, nsteps, s);
    // the graph bakes in y and the schedules; the per-level dt is read from device memory
    if (!e->graph_exec || e->graph_y != y_dev) {
        ens_drop_graph(e);
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = run_ens_steps(e, y_dev, kEnsGraphSteps, s);
        hipGraph_t g = nullptr;
        hipError_t he = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (he != hipSuccess) return set_err(TRISK_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
        e->graph = g;
        HIP_TRY(hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0));
        e->graph_y = y_dev;
    }
    const int64_t reps = nsteps / kEnsGraphSteps;
    for (int64_t r = 0; r < reps; ++r) HIP_TRY(hipGraphLaunch(e->graph_exec, s));
    return run_ens_steps(e, y_dev, nsteps - reps * kEnsGraphSteps, s);
}

int trisk_ensemble_set_graphs(trisk_ensemble* e, int32_t on) {
    CHECK_ENS(e);
    HIP_TRY(hipSetDevice(e->device));
    e->graphs = on != 0;
    if (!e->graphs) ens_drop_graph(e);
    return 0;
}

int trisk_ensemble_set_fused(trisk_ensemble* e, int32_t on) {
    CHECK_ENS(e);
    HIP_TRY(hipSetDevice(e->device));
    ens_drop_graph(e);
    if (!on) {
        e->fused = false;
        return 0;
    }
    int coop = 0;
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, e->device));
    if (!coop) return set_err(TRISK_EINVAL, "the device does not support cooperative launches");
    for (int l = 0; l < e->nlevels; ++l)
        if (e->M[l] % kFusedV0 || e->M[l] % kFusedV1 || e->M[l] % kFusedV2)
            return set_err(TRISK_EINVAL, "the fused stage needs a member count divisible by " +
                                             std::to_string(std::max(kFusedV0, kFusedV1)) + " at every level");
    // workgroups of the fused kernel that one CU holds at once (the least over its variants)
    int per_cu = 1 << 30;
    auto occupancy = [&](const void* f) -> int {
        int n = 0;
        hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, kFusedBlock, 0);
        if (err != hipSuccess) return set_err(TRISK_EHIP, std::string("hipOccupancyMaxActiveBlocksPerMultiprocessor: ") +
                                                              hipGetErrorString(err));
        per_cu = std::min(per_cu, n);
        return 0;
    };
    int rc;
    if (e->precision == TRISK_FP64) {
        if ((rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp64, double, true>))) ||
            (rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp64, double, false>))))
            return rc;
    } else {
        if ((rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp32Bulk, double, true>))) ||
            (rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp32Bulk, double, false>))) ||
            (rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp32Bulk, float, true>))) ||
            (rc = occupancy(reinterpret_cast<const void*>(&ens_stage_kernel<Fp32Bulk, float, false>))))
            return rc;
    }
    if (per_cu < 1) return set_err(TRISK_EINVAL, "the fused stage kernel does not fit a CU");
    // a multiple of the XCD count, so that every XCD walks its share with the same stride
    e->fgrid = int64_t(per_cu) * ((e->cus / kXcd) * kXcd);
    if (e->fgrid < kXcd) return set_err(TRISK_EINVAL, "too few compute units for the fused stage");
    const int lanes[kEnsLists] = {kFusedV0, kFusedV1, kFusedV2, kFusedV0};
    for (int list = 0; list < kEnsLists; ++list) {
        int64_t grid = 0;
        const trisk_schedule sc{kFusedBlock, per_cu, TRISK_SPLIT_CHUNKED, lanes[list]};
        if ((rc = build_list_to(e, list, sc, ListOut{&e->fseg_dev[list], &e->fseg_cap[list], e->ffirst[list], &grid})))
            return rc;
    }
    e->fused = true;
    return 0;
}

int trisk_ensemble_diagnostics(trisk_ensemble* e, const double* y_dev, double* out, void* stream) {
    CHECK_ENS(e);
    if (!y_dev || !out) return set_err(TRISK_EINVAL, "null pointer");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t k = 0;
    for (int l = 0; l < e->nlevels; ++l)
        for (int m = 0; m < e->M[l]; ++m, ++k) {
            int rc = ens_member_copy(e, l, m, const_cast<double*>(y_dev), e->scratch, 0, s);
            if (rc) return rc;
            if ((rc = trisk_diagnostics(e->models[l], e->scratch, out + 3 * k, stream))) return rc;
        }
    return 0;
}

int trisk_ensemble_set_schedule(trisk_ensemble* e, const trisk_schedule sched[3]) {
    CHECK_ENS(e);
    if (!sched) return set_err(TRISK_EINVAL, "null schedule");
    for (int k = 0; k < kEnsKinds; ++k)
        if (!valid_schedule(e, sched[k]))
            return set_err(TRISK_EINVAL, "schedule: block_size 128/256/512, blocks_per_cu 0..64, split CHUNKED/BLOCKED/INTERLEAVED, "
                                         "members_per_lane 1/2/4/8 dividing every level's member count");
    HIP_TRY(hipSetDevice(e->device));
    ens_drop_graph(e);
    HIP_TRY(hipDeviceSynchronize());  // runs in flight may still read the old tile lists
    for (int k = 0; k < kEnsKinds; ++k) {
        e->sched[k] = sched[k];
        int rc = build_schedule(e, k);
        if (rc) return rc;
    }
    return 0;
}

int trisk_ensemble_get_schedule(const trisk_ensemble* e, trisk_schedule sched[3]) {
    CHECK_ENS(e);
    if (!sched) return set_err(TRISK_EINVAL, "null schedule");
    for (int k = 0; k < kEnsKinds; ++k) sched[k] = e->sched[k];
    return 0;
}

int trisk_ensemble_set_timing(trisk_ensemble* e, int32_t on) {
    CHECK_ENS(e);
    e->timing = on != 0;
    return 0;
}

int trisk_ensemble_kernel_times(trisk_ensemble* e, int32_t max, const char** names, double* avg_ms,
                                int64_t* launches, int32_t* count) {
    CHECK_ENS(e);
    HIP_TRY(hipSetDevice(e->device));
    double ms[kEnsTimed] = {0};
    int64_t n[kEnsTimed] = {0};
    for (size_t i = 0; i + 1 < e->events_used; i += 2) {
        HIP_TRY(hipEventSynchronize(e->events[i + 1]));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, e->events[i], e->events[i + 1]));
        ms[e->event_kernel[i / 2]] += t;
        n[e->event_kernel[i / 2]] += 1;
    }
    int c = 0;
    for (int k = 0; k < kEnsTimed && c < max; ++k) {
        if (n[k] == 0) continue;
        if (names) names[c] = kEnsNames[k];
        if (avg_ms) avg_ms[c] = ms[k] / double(n[k]);
        if (launches) launches[c] = n[k];
        ++c;
    }
    if (count) *count = c;
    return 0;
}

int trisk_ensemble_kernel_bytes(trisk_ensemble* e, int32_t max, const char** names, double* bytes,
                                int32_t* count) {
    CHECK_ENS(e);
    // Algorithmic HBM bytes of one stage-2/3 launch: each level's tables once (shared by its
    // members), each member's state, intermediates and RK4 registers once.
    const bool f64 = e->precision == TRISK_FP64;
    const double c = f64 ? 8.0 : 4.0, uw = c, iw = 4.0;
    const double rk_e = 8.0 + 16.0 + uw, rk_c = 8.0 + 16.0 + 8.0;
    const double vc = f64 ? 7 * c : 6 * c, rec8 = f64 ? 6 * iw : 16.0, ew = f64 ? 2 * iw : 4.0,
                 cw = f64 ? 12 * iw : 24.0;
    double b[kEnsKinds] = {0, 0, 0};
    for (int l = 0; l < e->nlevels; ++l) {
        const Layout& L = e->models[l]->L;
        const double nc = double(L.nc), ne = double(L.ne), nv = double(L.nv), M = e->M[l];
        // RK4 stages after the first: G comes from the previous stage's edge kernel, which
        // reads l h and the new heights for it
        b[0] += nv * (rec8 + vc) + M * (nv * c + ne * uw + nc * 8.0);
        // the cell kernel forms B: KE weights, u gathered, the stage heights, B written
        b[1] += nc * (cw + 6 * c + iw + 8.0 + 6 * c) + M * (nc * rk_c + ne * c + nv * c + 6 * nc * c + ne * uw +
                                                          nc * (8.0 + 8.0));
        b[2] += ne * (ew + c + 1.0 + 2 * c) + M * (ne * (2 * c + rk_e + c) + nc * (8.0 + 8.0));
    }
    // the fused stage moves what its three phases move
    const double all[kEnsTimed] = {b[0], b[1], b[2], b[0] + b[1] + b[2]};
    int k = 0;
    for (; k < kEnsTimed && k < max; ++k) {
        if (names) names[k] = kEnsNames[k];
        bytes[k] = all[k];
    }
    if (count) *count = k;
    return 0;
}

// Times every candidate schedule on a copy of the state and keeps, per kernel, the one
// with the shortest mean launch.
int trisk_ensemble_autotune(trisk_ensemble* e, const double* y_dev, const double* dt, int32_t steps, void* stream,
                            trisk_schedule best[3], double best_ms[3]) {
    CHECK_ENS(e);
    if (!y_dev || !dt) return set_err(TRISK_EINVAL, "null pointer");
    if (steps < 1) return set_err(TRISK_EINVAL, "steps must be positive");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = e->precision == TRISK_FP64 ? set_dts<Fp64>(e, dt, s) : set_dts<Fp32Bulk>(e, dt, s);
    if (rc) return rc;
    double* work = nullptr;
    HIP_TRY(hipMalloc(&work, e->n_state * sizeof(double)));
    struct Free {
        double* p;
        ~Free() { (void)hipFree(p); }
    } guard{work};
    HIP_TRY(hipMemcpyAsync(work, y_dev, e->n_state * sizeof(double), hipMemcpyDeviceToDevice, s));
    const bool timing = e->timing, fused = e->fused;
    e->fused = false;  // the candidates are the three-launch schedules
    trisk_schedule keep[kEnsKinds];
    for (int k = 0; k < kEnsKinds; ++k) keep[k] = e->sched[k];
    trisk_schedule win[kEnsKinds] = {keep[0], keep[1], keep[2]};
    double win_ms[kEnsKinds] = {1e300, 1e300, 1e300};
    ens_drop_graph(e);
    e->timing = true;
    for (int lanes : {1, 2, 4, 8})
        for (int blk : {128, 256, 512})
            for (int bpc : {0, 2, 4, 8})
                for (int split : {TRISK_SPLIT_CHUNKED, TRISK_SPLIT_BLOCKED, TRISK_SPLIT_INTERLEAVED}) {
                    const trisk_schedule cand{blk, bpc, split, lanes};
                    if (rc || !valid_schedule(e, cand)) continue;
                    HIP_TRY(hipStreamSynchronize(s));
                    for (int k = 0; k < kEnsKinds && !rc; ++k) {
                        e->sched[k] = cand;
                        rc = build_schedule(e, k);
                    }
                    e->events_used = 0;
                    if (!rc) rc = run_ens_steps(e, work, 1, s);  // warm-up
                    e->events_used = 0;
                    if (!rc) rc = run_ens_steps(e, work, steps, s);
                    double ms[kEnsKinds];
                    int32_t cnt = 0;
                    if (!rc) rc = trisk_ensemble_kernel_times(e, kEnsKinds, nullptr, ms, nullptr, &cnt);
                    if (rc) continue;
                    for (int k = 0; k < kEnsKinds && k < cnt; ++k)
                        if (ms[k] < win_ms[k]) {
                            win_ms[k] = ms[k];
                            win[k] = e->sched[k];
                        }
                }
    e->timing = timing;
    e->fused = fused;
    e->events_used = 0;
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < kEnsKinds; ++k) {
        e->sched[k] = rc ? keep[k] : win[k];
        int rc2 = build_schedule(e, k);
        if (!rc) rc = rc2;
    }
    if (rc) return rc;
    for (int k = 0; k < kEnsKinds; ++k) {
        if (best) best[k] = win[k];
        if (best_ms) best_ms[k] = win_ms[k];
    }
    return 0;
}

}  // extern "C"
