// Device-side definitions shared by the TRiSK kernels (trisk_device.hip) and the ensemble
// kernels (trisk_ensemble.hip): precision policies, kernel parameters, XCD tile mapping,
// streamed loads, the RK4 stage update and the handle of one model.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "trisk_layout.hpp"

namespace trisk {

// Records `msg` as the thread's last error (trisk_last_error) and returns `code`.
int set_err(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return set_err(TRISK_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

// workgroup size of the single-model, ICON and MPAS kernels (4 waves of 64; 128 and 512
// measured equal or slower on every path)
constexpr int kBlock = 256;
constexpr int kXcd = 8;

// ---------------------------------------------------------------------------------
// precision policies

struct Fp64 {
    using C = double;  // coefficients of the stencils
    using I = double;  // intermediates F, q_e, q_v, zeta_v, h_v, K_v
    using U = double;  // stage velocity
    static constexpr int id = TRISK_FP64;
};
struct Fp32Bulk {
    using C = float;
    using I = float;
    using U = float;
    static constexpr int id = TRISK_FP32_BULK;
};

template <class I>
struct Pair {
    I F, q;
};
template <>
struct alignas(16) Pair<double> {
    double F, q;
};
template <>
struct alignas(8) Pair<float> {
    float F, q;
};

// ---------------------------------------------------------------------------------
// kernel parameters (device pointers, entity-major tables)

template <class P>
struct Params {
    using C = typename P::C;
    using I = typename P::I;
    int64_t nc, ne, nv;
    int kc, kv, kt, kg;
    double g;
    // cells
    const int32_t* ce;
    const double* div;
    const C* kin;
    const int32_t* cv;
    const C* vkk;
    const double* area_c;
    // edges
    const int32_t* ec;
    const int32_t* ev;
    const C* eh;
    const C* invd;
    const int32_t* te;
    const C* tw;
    const C* vkh;
    const int32_t* ge;
    const int32_t* gp;
    const C* gw;
    const C* grh;
    const C* gf;
    const uint32_t* csign;  // nc: flux form (Layout::gflux), 2 sign bits per cell slot
    // vertices
    const int32_t* ve;
    const C* curl;
    const int32_t* vc;
    const C* vh;
    const C* fv;
    const C* vkw;
    const double* area_v;
    // work
    double* B;          // nc
    double* K;          // nc (vertex KE or diagnostics)
    Pair<I>* Fq;        // ne
    I* qv;              // nv
    I* zv;              // nv (Gassmann, diagnostics)
    I* hv;              // nv (Gassmann, vertex KE, diagnostics)
    I* Kv;              // nv (vertex KE)
    // cell-pair path
    const int32_t* cmeta;  // nc: size | negative-sign bits << 4
    const int32_t* ccv;    // 6 x nc corners
    const C* pr;           // 6 x nc kite fractions (the pair weights are rebuilt from them)
    const C* len;          // ne: l_e
    const C* dual;         // ne: d_e
    I* G;                  // ne: l_e F_e
    Pair<I>* gu;           // ne: (G_e, u_e) of the stage input, one pair per edge (the (G, u) plane)
    I* T;                  // 6 x nc, cell-major: PV-flux half sum of cell i at its edge k in T[6 i + k]
    const uint8_t* eslot;  // ne: slot of e in its left cell | slot in its right cell << 4
    // entity records of the cell-pair path
    const int32_t* rc_state;  // nc x 6
    const C* rc_kin;          // nc x 6
    const int32_t* rv_idx;    // nv x 6
    const C* rv_coef;         // nv x 4, then nv x 2 (FP32-bulk)
    const int32_t* re_cells;  // ne x 2
    const C* re_lh;           // ne x 2
    const int32_t* ce6;       // 6 x nc cell edges (pads: the first edge)
    // 16-bit deltas of the slot-major index tables (table kernels; kWide: the int32 row)
    const int32_t *ce16p, *cv16p, *ec16p, *ev16p, *te16p, *ge16p, *gp16p, *ve16p, *vc16p;
    // 16-bit deltas (sentinel kWide: read the int32 tables)
    const int16_t* rc_state16;  // nc x 8
    const int16_t* rv_idx16;    // nv x 8
    const int32_t* re_cells16;  // ne
    const int32_t* re_rec;      // ne x 4 (FP32-bulk): the cells (as re_cells16), the FP32 bits of d_e, l h_0, l h_1
    const int32_t* pc_ev16;     // 6 x nc
    const int32_t* rc_pair;     // nc x kPairRecord (FP32-bulk cell kernel records)
    // FP32-bulk Gassmann slot form: one 80-byte record per edge, the 12 slot edges as 16-bit
    // deltas (6 words, kWide: the int32 table ge), the two cells (one word, kWide: ec), the
    // FP32 bits of 1/d_e and of the 12 weights
    const int32_t* rg;          // ne x 20
    // FP32-bulk table kernels on hexagon meshes: the slot-major tables of P1 and P2 as entity
    // records (16-bit index words first, then the FP32 bits of the coefficients), staged
    // through LDS: P1 cells [ce 3 words, kin 6], vertices [ve 2, vc 2, A_v curl 3, A_v vh 3,
    // A_v f, vkw 3 with vertex KE]; P2 edges [ec, ev, eh 2, vkh_s / A_s 2 with vertex KE,
    // (A_0 + A_1) f_e with the Gassmann flux], cells [cv 3, vkk 6] (vertex KE)
    const int32_t *r1c, *r1v, *r2e, *r2c;
    // optional FP64 outputs of the diagnostics (NULL: not written)
    double* K_out;
    double* qv_out;
    double* zv_out;
    double* hv_out;
};

// Block b of a grid of kXcd * (t0 + t1) blocks -> (range, tile).  The hardware deals
// blocks round-robin to the XCDs; XCD x receives tiles [x*t0, (x+1)*t0) of range 0
// and [x*t1, (x+1)*t1) of range 1, contiguous in the Hilbert order.  Range 1 runs backward
// through its share, so it starts on the region range 0 ended on, whose gathered data is
// still in the XCD's L2 (level-10 Gassmann step: 7.25 -> 7.07 ms FP32-bulk, 12.42 -> 12.22
// ms FP64; no change at level 9, in ICON or in MPAS).
struct Tile {
    int range;
    int64_t first;  // first entity of the tile (may be past the range's end)
};
__device__ __forceinline__ Tile tile_of(int64_t b, int64_t t0, int64_t t1) {
    int64_t x = b % kXcd, j = b / kXcd;
    if (j < t0) return {0, (x * t0 + j) * kBlock};
    return {1, (x * t1 + (t1 - 1 - (j - t0))) * kBlock};
}

// The second range runs backward through each XCD's share, so it starts where the first
// range ended and finds that region's gathered data still in the XCD's L2 (the cell-pair state
// kernel: 3.4 % fewer HBM reads at level 9).
__device__ __forceinline__ Tile tile_of3(int64_t b, int64_t t0, int64_t t1, int64_t t2) {
    int64_t x = b % kXcd, j = b / kXcd;
    if (j < t0) return {0, (x * t0 + j) * kBlock};
    if (j < t0 + t1) return {1, (x * t1 + (t1 - 1 - (j - t0))) * kBlock};
    return {2, (x * t2 + (j - t0 - t1)) * kBlock};
}

inline int64_t tiles_per_xcd(int64_t n) { return ((n + kBlock - 1) / kBlock + kXcd - 1) / kXcd; }

// A pointer into global memory (device allocations, never LDS or scratch): loads through it
// are global_load instructions.  Kernels that read their pointers from a table in memory
// (the ensemble's level records) rather than from their arguments would otherwise issue
// flat loads, whose completion is tracked together with the LDS and scalar counters.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;

// Streamed tables are read once per launch: nontemporal loads keep them from evicting
// the gathered state and intermediates from the caches.  p points to global memory.
template <class T>
__device__ __forceinline__ T ld(const T* p) {
    return __builtin_nontemporal_load((gptr<T>)p);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// N consecutive elements starting at p (one entity record, global memory) with 16-byte
// (or, for records of a size not divisible by 16, 8-byte) nontemporal vector loads; p must
// be aligned to that width.
template <class T, int N>
__device__ __forceinline__ void ldrec(const T* __restrict__ p, T (&out)[N]) {
    constexpr int bytes = N * int(sizeof(T));
    if constexpr (bytes % 16 == 0) {
        gptr<f32x4> q = (gptr<f32x4>)p;
#pragma unroll
        for (int c = 0; c < bytes / 16; ++c) {
            f32x4 v = __builtin_nontemporal_load(q + c);
            __builtin_memcpy(reinterpret_cast<char*>(out) + 16 * c, &v, 16);
        }
    } else {
        static_assert(bytes % 8 == 0, "records are multiples of 8 bytes");
        gptr<f32x2> q = (gptr<f32x2>)p;
#pragma unroll
        for (int c = 0; c < bytes / 8; ++c) {
            f32x2 v = __builtin_nontemporal_load(q + c);
            __builtin_memcpy(reinterpret_cast<char*>(out) + 8 * c, &v, 8);
        }
    }
}


// Copies the records of entities [first, min(first + kBlock, n)), R int32 words each, into
// LDS with coalesced 16-byte nontemporal loads (first is a multiple of kBlock, so the span is
// 16-byte aligned) and waits for the block; false (for the whole block) past the range's end.
template <int R>
__device__ __forceinline__ bool stage_records(const int32_t* __restrict__ rec, int64_t first, int64_t n,
                                              int32_t* lds) {
    if (first >= n) return false;
    const int64_t cnt = n - first < kBlock ? n - first : kBlock;
    const int nw = int(cnt * R), nw4 = nw / 4;
    const int32_t* src = rec + first * R;
    for (int w = threadIdx.x; w < nw4; w += kBlock)
        reinterpret_cast<f32x4*>(lds)[w] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src) + w);
    for (int w = 4 * nw4 + threadIdx.x; w < nw; w += kBlock) lds[w] = ld(src + w);
    __syncthreads();
    return true;
}

// The FP32 value whose bits are word w
__device__ __forceinline__ float fbits(int32_t w) {
    float f;
    __builtin_memcpy(&f, &w, 4);
    return f;
}

// ---------------------------------------------------------------------------------
// RK4 stage update (rk4_step!, RungeKutta.jl:18-28)

struct Rk {
    int mode;            // 0: store the tendency in dy; 1..4: RK4 stage
    double* dy;          // mode 0
    double* y;           // state
    double* acc;         // k1 + 2k2 + 2k3 accumulator
    void* su;            // stage velocity (ne, type U)
    double* sh;          // stage height (nc)
    double dt, dt2, dt6;
};

// Slot k of entity j of a slot-major index table.  With a compile-time width K the row is
// decoded up front from the 16-bit deltas packed two per word (base + delta; the int32 row
// when word 0 holds kWide), so all its loads issue together; with K = 0 each slot is read
// from the int32 table when used.
// The packed words of row j, loaded ahead of their decoding (index_words), let a kernel issue
// the record loads of several tables together before any of them is decoded.
template <int K>
struct IndexWords {
    int32_t w[K > 0 ? (K + 1) / 2 : 1];
};
template <int K>
__device__ __forceinline__ IndexWords<K> index_words(const int32_t* p16, int64_t n, int64_t j) {
    IndexWords<K> r{};
    if constexpr (K > 0) {
#pragma unroll
        for (int q = 0; q < (K + 1) / 2; ++q) r.w[q] = ld(p16 + q * n + j);
    }
    return r;
}

template <int K>
struct Index {
    int32_t v[K > 0 ? K : 1];
    const int32_t* t32;
    int64_t n, j;
    __device__ __forceinline__ Index(const int32_t* p16, const int32_t* t32_, int64_t n_, int64_t j_, int64_t base)
        : Index(index_words<K>(p16, n_, j_), t32_, n_, j_, base) {}
    __device__ __forceinline__ Index(const IndexWords<K>& words, const int32_t* t32_, int64_t n_, int64_t j_,
                                     int64_t base)
        : t32(t32_), n(n_), j(j_) {
        if constexpr (K > 0) {
            const int32_t* w = words.w;
            if (int16_t(w[0] & 0xffff) == kWide) {
#pragma unroll
                for (int k = 0; k < K; ++k) v[k] = ld(t32 + k * n + j);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    v[k] = int32_t(base) + (k % 2 ? int16_t(uint32_t(w[k / 2]) >> 16) : int16_t(w[k / 2] & 0xffff));
            }
        }
    }
    __device__ __forceinline__ int32_t operator[](int k) const {
        if constexpr (K > 0)
            return v[k];
        else
            return ld(t32 + k * n + j);
    }
};

// The RK4 stage update of entity j with tendency k: y + c dt k and the accumulator, in the
// order of rk4_step!'s broadcasts (RungeKutta.jl:18-28).  rk_load reads y_j and acc_j, which
// a kernel issues first so that these streaming loads do not wait for its gathers; rk_store
// writes the accumulator and the stage (modes 1-3), the new state (mode 4) or the tendency
// (mode 0).
struct RkIn {
    double y, acc;
};

__device__ __forceinline__ RkIn rk_load(const Rk& r, int64_t j) {
    RkIn v{0.0, 0.0};
    if (r.mode >= 1) v.y = r.y[j];
    if (r.mode >= 2) v.acc = r.acc[j];
    return v;
}

// Which of rk_store's writes are nontemporal (streaming) stores: none, the accumulator, state
// and tendency (read next by the same entity's next stage), or also the stage vector.
// Measured: kStreamAll is faster in the table path's update kernel and the ICON kernels;
// the cell-pair kernels take kStreamAcc, since their stage vectors are gathered by the next
// kernel, which ran slower when they were streamed.
enum RkStores { kPlain = 0, kStreamAcc = 1, kStreamAll = 2 };

// The value rk_store leaves as the next stage's input (RK4 modes): the stage value, or y after
// the last stage.  A stage vector of type U holds U(x), which reads back as the same C(x).
__device__ __forceinline__ double rk_next(const Rk& r, double k, const RkIn& v) {
    switch (r.mode) {
        case 1:
        case 2:
            return v.y + r.dt2 * k;
        case 3:
            return v.y + r.dt * k;
        default:
            return v.y + r.dt6 * (v.acc + k);
    }
}

// STAGE false: modes 1-3 leave the stage vector to the caller (the (G, u) plane holds it).
template <class U, int STORES = kPlain, bool STAGE = true>
__device__ __forceinline__ void rk_store(const Rk& r, int64_t j, double k, const RkIn& v, U* stage) {
    auto acc = [](auto x, auto* p) {
        if constexpr (STORES != kPlain)
            __builtin_nontemporal_store(x, p);
        else
            *p = x;
    };
    auto stg = [](auto x, auto* p) {
        if constexpr (STORES == kStreamAll)
            __builtin_nontemporal_store(x, p);
        else
            *p = x;
    };
    switch (r.mode) {
        case 0:
            acc(k, r.dy + j);
            break;
        case 1:
            acc(k, r.acc + j);
            if constexpr (STAGE) stg(U(v.y + r.dt2 * k), stage);
            break;
        case 2:
            acc(v.acc + 2 * k, r.acc + j);
            if constexpr (STAGE) stg(U(v.y + r.dt2 * k), stage);
            break;
        case 3:
            acc(v.acc + 2 * k, r.acc + j);
            if constexpr (STAGE) stg(U(v.y + r.dt * k), stage);
            break;
        default:
            acc(v.y + r.dt6 * (v.acc + k), r.y + j);
            break;
    }
}

// The two cells of edge e from the packed 16-bit deltas (base e/3), or the int32 pair.
template <class P>
__device__ __forceinline__ void edge_cells_narrow(const Params<P>& p, int64_t e, int32_t (&c)[2]) {
    const int32_t pk = ld(p.re_cells16 + e);
    const int16_t d0 = int16_t(pk & 0xffff), d1 = int16_t(uint32_t(pk) >> 16);
    if (d0 == kWide) {
        ldrec(p.re_cells + 2 * e, c);
    } else {
        const int32_t b = int32_t(e / 3);
        c[0] = b + d0;
        c[1] = b + d1;
    }
}

// The FP32-bulk record of edge e (one 16-byte load): its two cells (16-bit deltas from e/3, or
// the int32 pair for an edge flagged kWide), d_e and the two weights l_e h_e,s of the mass flux.
template <class P>
__device__ __forceinline__ void load_edge_record(const Params<P>& p, int64_t e, int32_t (&c)[2], float& d,
                                                 float (&lh)[2]) {
    int32_t w[4];
    ldrec(p.re_rec + 4 * e, w);
    const int16_t d0 = int16_t(w[0] & 0xffff), d1 = int16_t(uint32_t(w[0]) >> 16);
    if (d0 == kWide) {
        ldrec(p.re_cells + 2 * e, c);
    } else {
        const int32_t b = int32_t(e / 3);
        c[0] = b + d0;
        c[1] = b + d1;
    }
    __builtin_memcpy(&d, &w[1], 4);
    __builtin_memcpy(&lh[0], &w[2], 4);
    __builtin_memcpy(&lh[1], &w[3], 4);
}

// Table records of the cell-pair path.  NARROW: the FP32-bulk layout (16-bit index
// deltas from a linear base, entity records read with vector loads, int32 fallback for
// entities flagged kWide); otherwise the slot-major int32 tables.

// cell i: its six edge slots and kinetic-energy weights (pads: a valid index, weight 0)
template <bool NARROW, class P>
__device__ __forceinline__ void load_cell_state(const Params<P>& p, int64_t i, int32_t (&E)[6],
                                                typename P::C (&kin)[6]) {
    using C = typename P::C;
    if (NARROW) {
        int16_t d[8];
        ldrec(p.rc_state16 + 8 * i, d);
        if (d[0] == kWide) {
            ldrec(p.rc_state + 6 * i, E);
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) E[k] = int32_t(3 * i) + d[k];
        }
        ldrec(p.rc_kin + 6 * i, kin);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            E[k] = k < p.kc ? ld(p.ce + k * p.nc + i) : 0;
            kin[k] = k < p.kc ? ld(p.kin + k * p.nc + i) : C(0);
        }
    }
}

// vertex v: edges (3), cells (3); coefficients curl (3), vertex heights (3), f
template <bool NARROW, class P>
__device__ __forceinline__ void load_vertex_record(const Params<P>& p, int64_t v, int32_t (&idx)[6],
                                                   typename P::C (&cf)[8]) {
    if constexpr (NARROW) {  // FP32 records: indices and f in 16 bytes, curl and vertex heights in 16 + 8
        int16_t d[8];
        ldrec(p.rv_idx16 + 8 * v, d);
        if (d[0] == kWide) {
            ldrec(p.rv_idx + 6 * v, idx);
        } else {
            const int32_t b0 = int32_t((3 * v) >> 1), b1 = int32_t(v >> 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                idx[k] = b0 + d[k];
                idx[3 + k] = b1 + d[3 + k];
            }
        }
        // the coefficients as a 16-byte record (curl 3, the first vertex height) and an 8-byte
        // one (the other two) in the array after it: every load of a wave is contiguous (one
        // 24-byte record read as three 8-byte loads touched each cache line three times)
        typename P::C c4[4], c2[2];
        ldrec(p.rv_coef + 4 * v, c4);
        ldrec(p.rv_coef + 4 * p.nv + 2 * v, c2);
#pragma unroll
        for (int k = 0; k < 4; ++k) cf[k] = c4[k];
        cf[4] = c2[0];
        cf[5] = c2[1];
        static_assert(sizeof(typename P::C) == 4, "the records hold FP32 coefficients");
        __builtin_memcpy(&cf[6], &d[6], 4);
        cf[7] = 0;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            idx[k] = ld(p.ve + k * p.nv + v);
            idx[3 + k] = ld(p.vc + k * p.nv + v);
            cf[k] = ld(p.curl + k * p.nv + v);
            cf[3 + k] = ld(p.vh + k * p.nv + v);
        }
        cf[6] = ld(p.fv + v);
        cf[7] = 0;
    }
}

// edge e: left and right cell
template <bool NARROW, class P>
__device__ __forceinline__ void load_edge_cells(const Params<P>& p, int64_t e, int32_t (&c)[2]) {
    if (NARROW)
        edge_cells_narrow(p, e, c);
    else
        ldrec(p.re_cells + 2 * e, c);
}

// w_{k,k'} = s_k s_k' (1/2 - sum_{j=k+1..k'} R_j) (TRiSKGrid.jl:143-155) from the kite fractions R
// and the sign bits of meta, in the order and association of the host construction, so FP64
// weights are bitwise the host's
template <class C>
__device__ __forceinline__ void pair_weights(int32_t meta, const C (&R)[6], C (&W)[15]) {
    int q = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const C sk = (meta >> (4 + k)) & 1 ? C(-1) : C(1);
        C S = 0;
#pragma unroll
        for (int k2 = k + 1; k2 < 6; ++k2, ++q) {
            S += R[k2];
            const C s2 = (meta >> (4 + k2)) & 1 ? C(-1) : C(1);
            W[q] = sk * s2 * (C(0.5) - S);
        }
    }
}

// cell i: size and side bits, edges and corners (6 slots, pads valid), pair weights (those of
// slots beyond the size are not used)
template <bool NARROW, class P>
__device__ __forceinline__ void load_cell_pairs(const Params<P>& p, int64_t i, int32_t& meta, int32_t (&E)[6],
                                                int32_t (&V)[6], typename P::C (&W)[15]) {
    const int64_t nc = p.nc;
    meta = ld(p.cmeta + i);
    bool wide = !NARROW;
    if (NARROW) {
        int32_t pk[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) pk[k] = ld(p.pc_ev16 + k * nc + i);
        wide = int16_t(pk[0] & 0xffff) == kWide;
        if (!wide) {
            const int32_t be = int32_t(3 * i), bv = int32_t(2 * i);
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                E[k] = be + int16_t(pk[k] & 0xffff);
                V[k] = bv + int16_t(uint32_t(pk[k]) >> 16);
            }
        }
    }
    if (wide) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            E[k] = ld(p.ce6 + k * nc + i);
            V[k] = ld(p.ccv + k * nc + i);
        }
    }
    typename P::C R[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) R[k] = ld(p.pr + k * nc + i);
    pair_weights(meta, R, W);
}

// ---------------------------------------------------------------------------------
// per-entity arithmetic of the cell-pair path, shared by the single-model kernels and the
// ensemble kernels so that every member rounds exactly as a model of its own

// K_i = sum_k kin_k u_k^2 over the six slots (pads have a zero weight)
template <class C>
__device__ __forceinline__ C pair_kinetic(const C (&kin)[6], const C (&ue)[6]) {
    C s = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) s += kin[k] * (ue[k] * ue[k]);
    return s;
}

// q_v = (zeta_v + f_v) / h_v from the vertex record (curl 3, vertex heights 3, f)
template <class C>
__device__ __forceinline__ C pair_vertex_pv(const C (&cf)[8], const C (&uz)[3], const C (&hz)[3]) {
    C z = 0, eta = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) z += cf[k] * uz[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) eta += cf[3 + k] * hz[k];
    return (z + cf[6]) / eta;
}

// G_e = l_e h_e u_e
template <class C>
__device__ __forceinline__ C pair_mass_flux(const C (&lh)[2], C h0, C h1, C u) {
    return (lh[0] * h0 + lh[1] * h1) * u;
}

// the PV-flux half sums H_k = sum_k' w_{k,k'} G_k' (q_k + q_k') of a cell with n edges,
// q_k the mean of the PV at the two corners of edge k
template <class C>
__device__ __forceinline__ void pair_half_sums(int n, const C (&W)[15], const C (&G)[6], const C (&Q)[6],
                                               C (&H)[6]) {
    C qe[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) qe[k] = (Q[k] + (k + 1 < n ? Q[k + 1 < 6 ? k + 1 : 0] : Q[0])) / 2;
#pragma unroll
    for (int k = 0; k < 6; ++k) H[k] = 0;
    int q = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int k2 = k + 1; k2 < 6; ++k2, ++q) {
            if (k2 < n) {
                const C tq = W[q] * (qe[k] + qe[k2]);
                H[k] += tq * G[k2];
                H[k2] -= tq * G[k];
            }
        }
}

// du_e = (H_left + H_right) / (2 d_e) - (B_right - B_left) / d_e
template <class C>
__device__ __forceinline__ double pair_edge_tendency(C hs0, C hs1, C d, double Bl, double Br) {
    const C flux = (hs0 + hs1) / (2 * d);
    const double grad = (Br - Bl) / double(d);
    return double(flux) - grad;
}

// ---------------------------------------------------------------------------------
// compensated sums

__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// ---------------------------------------------------------------------------------
// compensated sums: Neumaier per thread over a fixed chunk, then a wavefront butterfly and
// the waves of the block in order, all on (sum, compensation) pairs combined with Knuth's
// exact two-sum.  Both partners of a butterfly step compute the same pair (the two-sum
// error of a + b is that of b + a), so the result does not depend on scheduling.

// s + e = a + b exactly
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (s, c) <- (s, c) + (s2, c2)
__device__ __forceinline__ void pair_add(double& s, double& c, double s2, double c2) {
    double t, e;
    two_sum(s, s2, t, e);
    s = t;
    c = (c + c2) + e;
}

// term 0: A_i h_i; 1: A_i h_i (K_i + g h_i / 2); 2: A_v h_v q_v^2; 3: 1 for every entry of a that
// is NaN or +-Inf (an exponent of all ones), else 0; 4: 1 for every entry of a that is not > 0
// (NaN included), else 0.  Terms 3 and 4 are exact counts (area and b unused).
static __global__ __launch_bounds__(kBlock) void reduce_kernel(int term, int64_t n, const double* __restrict__ area,
                                                        const double* __restrict__ a, const double* __restrict__ b,
                                                        double g, double* __restrict__ partial) {
    constexpr int kWave = 64, kWaves = kBlock / kWave;
    __shared__ double ws[kWaves], wc[kWaves];
    int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    int64_t lo = int64_t(blockIdx.x) * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0, c = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        double x;
        if (term == 0)
            x = area[i] * a[i];
        else if (term == 1)
            x = area[i] * a[i] * (b[i] + g * a[i] / 2);
        else if (term == 2)
            x = area[i] * a[i] * (b[i] * b[i]);
        else if (term == 3)
            x = (__double_as_longlong(a[i]) & 0x7ff0000000000000ll) == 0x7ff0000000000000ll ? 1.0 : 0.0;
        else
            x = a[i] > 0.0 ? 0.0 : 1.0;
        neumaier(s, c, x);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) pair_add(s, c, __shfl_xor(s, off), __shfl_xor(c, off));
    const int wave = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        ws[wave] = s;
        wc[wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ws[0], Cc = wc[0];
        for (int k = 1; k < kWaves; ++k) pair_add(S, Cc, ws[k], wc[k]);
        partial[2 * blockIdx.x] = S;
        partial[2 * blockIdx.x + 1] = Cc;
    }
}

// ---------------------------------------------------------------------------------
// device memory owned by a handle

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

template <class T>
inline int upload(std::vector<DevBuf>& owned, const std::vector<T>& v, const T** out) {
    if (v.empty()) {
        *out = nullptr;
        return 0;
    }
    DevBuf b;
    b.bytes = v.size() * sizeof(T);
    if (hipMalloc(&b.p, b.bytes) != hipSuccess) return set_err(TRISK_ENOMEM, "hipMalloc of a table failed");
    owned.push_back(b);
    HIP_TRY(hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice));
    *out = static_cast<const T*>(b.p);
    return 0;
}

inline int alloc(std::vector<DevBuf>& owned, size_t bytes, void** out) {
    DevBuf b;
    b.bytes = bytes ? bytes : 16;
    if (hipMalloc(&b.p, b.bytes) != hipSuccess) return set_err(TRISK_ENOMEM, "hipMalloc of a work vector failed");
    owned.push_back(b);
    *out = b.p;
    return 0;
}

inline std::vector<float> to_float(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

}  // namespace trisk

// One model on one device: its tables, work vectors, RK4 registers, captured graph and timing events.
namespace trisk {

// Hot-path kernels, for the timing events and the byte model.
enum KernelId {
    K_CELL_VERTEX = 0, K_EDGE = 1, K_UPDATE = 2, K_PAIR_STATE = 3, K_PAIR_CELL = 4, K_PAIR_EDGE = 5,
    K_ICON_EDGE = 6, K_ICON_CELL_VERTEX = 7, K_ICON_DAMP_EDGE = 8, K_ICON_DAMP_CELL = 9, K_ICON_MOMENTUM = 10,
    K_COUNT = 11
};
inline constexpr const char* kKernelNames[K_COUNT] = {
    "cell_vertex_kernel",      "edge_kernel",           "update_kernel",          "pair_state_kernel",
    "pair_cell_kernel",        "pair_edge_kernel",      "icon_edge_kernel",       "icon_cell_vertex_kernel",
    "icon_damp_edge_kernel",   "icon_damp_cell_kernel", "icon_momentum_kernel"};

struct IconState;  // device tables and work of an ICON handle (icon_device.hip)

}  // namespace trisk

struct trisk_handle {
    trisk::Layout L;
    int precision = TRISK_FP64;
    int device = 0;
    std::vector<trisk::DevBuf> owned;
    trisk::Params<trisk::Fp64> p64{};
    trisk::Params<trisk::Fp32Bulk> p32{};
    // RK4 work (acc, stage) and reduction scratch
    double* acc = nullptr;
    void* su = nullptr;    // stage velocity (ne, FP64 or FP32)
    double* sh = nullptr;  // stage height (nc)
    double* partial = nullptr;
    int reduce_blocks = 1024;
    double* diag_K = nullptr;
    double* diag_qv = nullptr;
    double* diag_zv = nullptr;
    double* diag_hv = nullptr;
    bool finite_check = false;         // trisk_set_finite_check
    bool pair_plane = false;           // trisk_set_pair_plane: rk4_steps keeps (G, u) in one plane
    mutable std::string last_error;    // trisk_handle_last_error
    std::vector<double> host_stage;
    // graph replay of trisk_rk4_steps: one captured graph of kGraphSteps steps per (y, dt)
    bool graphs = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    double* graph_y = nullptr;
    double graph_dt = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> events;     // pairs (start, stop)
    std::vector<int> event_kernel;      // kernel id per pair
    size_t events_used = 0;
    // adjoint work (trisk_adjoint.hip), allocated on first use and released with the handle
    void* adj = nullptr;
    void (*adj_release)(trisk_handle*) = nullptr;
    // the ICON triangle model (L.icon): its device tables and work, released with the handle
    trisk::IconState* icon = nullptr;
    void (*icon_release)(trisk_handle*) = nullptr;
};

namespace trisk {

// Forward passes for the adjoint (trisk_device.hip).  The stage input u is double (the state
// y) when u_double, else the handle's stage type (P::U).
// Cell-pair path, W1 only: B, q_v and G of the stage (u, h).
int pair_state_pass(trisk_handle* H, const void* u, bool u_double, const double* h, hipStream_t s);
// Table kernels, P1 and P2 only: B, zeta_v, h_v, q_v, K_v and the (F, q_e) pairs of the stage.
int table_state_pass(trisk_handle* H, const void* u, bool u_double, const double* h, hipStream_t s);
// One tendency of the handle's path with the RK4 stage epilogue r.
int tendency_stage(trisk_handle* H, const void* u, bool u_double, const double* h, const Rk& r, hipStream_t s);
// n fused RK4 steps of y with direct launches.
int rk4_steps_direct(trisk_handle* H, double* y, double dt, int64_t n, hipStream_t s);

// Timing events around one launch (when the handle's timing is on).
inline int timed_begin(trisk_handle* H, int kernel, hipStream_t s, hipEvent_t* stop) {
    *stop = nullptr;
    if (!H->timing) return 0;
    if (H->events_used + 2 > H->events.size()) {
        for (int k = 0; k < 256; ++k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            H->events.push_back(e);
        }
        H->event_kernel.resize(H->events.size() / 2);
    }
    size_t i = H->events_used;
    H->events_used += 2;
    H->event_kernel[i / 2] = kernel;
    HIP_TRY(hipEventRecord(H->events[i], s));
    *stop = H->events[i + 1];
    return 0;
}

inline int timed_end(hipEvent_t stop, hipStream_t s) {
    if (stop) HIP_TRY(hipEventRecord(stop, s));
    return 0;
}

// Compensated sum of a diagnostic term (reduce_kernel) over n entities: a fixed-order
// Neumaier combination of the block partials on the host.  Synchronous.
inline int reduce(trisk_handle* H, int term, int64_t n, const double* area, const double* a, const double* b,
                  double* out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3(unsigned(H->reduce_blocks)), dim3(kBlock), 0, s, term, n, area, a, b,
                       H->L.gravity, H->partial);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(2 * H->reduce_blocks);
    HIP_TRY(hipMemcpyAsync(part.data(), H->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double S = 0.0, Cc = 0.0;
    for (int b2 = 0; b2 < H->reduce_blocks; ++b2) {
        double x = part[2 * b2], t = S + x;
        if (std::fabs(S) >= std::fabs(x))
            Cc += (S - t) + x;
        else
            Cc += (x - t) + S;
        S = t;
        Cc += part[2 * b2 + 1];
    }
    *out = S + Cc;
    return 0;
}

// The generic calls on an ICON handle (icon_device.hip).
int icon_tendency(trisk_handle* H, const double* y, double* dy, hipStream_t s);
int icon_run_steps(trisk_handle* H, double* y, double dt, int64_t n, hipStream_t s);
int icon_diagnostics(trisk_handle* H, const double* y, double out[3], hipStream_t s);
// K, q_v and h_v of the state y (the diagnostic pass); *K receives the device cell K.
int icon_diagnostic_pass(trisk_handle* H, const double* y, const double** K, hipStream_t s);
// Per-device address of the library's static scratch (trisk_diagnostics_noalloc), looked
// up once per device with that device current (trisk_device.hip).
void noalloc_setup(int device);
int icon_vertex_pv(trisk_handle* H, const double* y, double* qv, double* zv, double* hv, hipStream_t s);
int icon_relative_vorticity(trisk_handle* H, const double* y, double* zeta, hipStream_t s);
int icon_tangential_velocity(trisk_handle* H, const double* u, double* vt, hipStream_t s);
std::vector<std::pair<const char*, double>> icon_kernel_bytes(const trisk_handle* H);

}  // namespace trisk

