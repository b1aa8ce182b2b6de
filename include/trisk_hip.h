/*
 * trisk_hip.h -- C ABI of libtrisk_hip.so, the MI355X (gfx950) implementation of the
 * nonlinear TRiSK C-grid shallow-water step of dmplex.jl.
 *
 * The library replaces the device side of
 *   TRiSKShallowWater(dm; ...)   src/FiniteVolumes/TRiSKShallowWater.jl:94-170  -> trisk_create
 *   tendency!(dy, y, m)          src/FiniteVolumes/TRiSKShallowWater.jl:379-394 -> trisk_tendency
 *   rk4_step!(y, dt, tendency!, RK4Cache(y), m)
 *                                src/TimeSteppers/RungeKutta.jl:2-28            -> trisk_rk4_steps
 *   total_mass / total_energy / potential_enstrophy
 *                                src/FiniteVolumes/TRiSKShallowWater.jl:397-419 -> trisk_diagnostics
 *   relative_vorticity / tangential_velocity
 *                                src/FiniteVolumes/TRiSKShallowWater.jl:422-433 -> trisk_relative_vorticity,
 *                                                                                  trisk_tangential_velocity
 *   kinetic_energy! / vertex_pv! src/FiniteVolumes/TRiSKShallowWater.jl:332-360 -> trisk_kinetic_energy,
 *                                                                                  trisk_vertex_pv
 *   to_backend(backend, x)       src/FEM/DeviceSolvers.jl:7-11                  -> trisk_upload_state
 *   gather_state(m, y)           src/FiniteVolumes/TRiSKShallowWater.jl:191-195 -> trisk_download_state
 *
 * The caller hands over the model's tables exactly as the reference constructor builds
 * them (host pointers, borrowed for the duration of trisk_create): k x n column-major
 * tables (entry (k, j) at k + j*K, the memory of a Julia k x n Matrix) with 0-based
 * int64 indices (the reference's Int tables minus one); padded slots have weight 0.
 * The library validates every index, renumbers cells along the Hilbert curve of
 * `cell_points` (edges and vertices by first appearance), stores every table
 * entity-major (struct of arrays, int32 indices) in device memory, and owns that memory.
 *
 * Device state vectors `y` hold [u (ne); h (nc)] in the library's internal (Hilbert)
 * order; trisk_upload_state / trisk_download_state convert from and to the reference's
 * global order.  Device vectors are allocated by the caller (any hipMalloc'ed memory)
 * and must hold trisk_state_length() doubles.  Every call is stream-ordered on the
 * given hipStream_t (passed as void*, NULL = the null stream) and returns without a
 * device synchronisation unless stated.  A handle is not re-entrant: calls on one
 * handle must not overlap.  All calls return 0 or a negative TRISK_E* code; the
 * message is in trisk_last_error() (thread-local).
 */
#ifndef TRISK_HIP_H
#define TRISK_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TRISK_OK = 0,
    TRISK_EINVAL = -1,  /* bad argument or table (the reference's ArgumentError) */
    TRISK_EHIP = -2,    /* a HIP runtime error */
    TRISK_ENOMEM = -3,  /* device allocation failed */
    TRISK_ENONFINITE = -4  /* the state check found NaN/Inf in y or a vertex height h_v <= 0 */
};

/* pv_flux keyword (TRiSKShallowWater.jl:94-98) */
enum { TRISK_PV_TRISK = 0, TRISK_PV_GASSMANN = 1, TRISK_PV_GASSMANN_MEAN = 2 };

/* arithmetic of the stencils */
enum {
    TRISK_FP64 = 0,      /* everything in FP64: matches the FP64 oracle */
    TRISK_FP32_BULK = 1  /* FP32 coefficients, fluxes and stencil arithmetic; FP64 state,
                            heights, Bernoulli function, RK4 accumulation and sums */
};

/* The fields of the reference's TRiSKShallowWater struct (TRiSKShallowWater.jl:67-92). */
typedef struct {
    int64_t nc, ne, nv;          /* cells, edges, vertices */
    int64_t kc, kv, kt, kg;      /* widths: cell edges, vertex edges, TRiSK stencil, Gassmann (12 or 0) */
    const int64_t* cell_edges;   /* kc x nc */
    const double* divergence;    /* kc x nc   n_{e,i} l_e / A_i */
    const double* kinetic;       /* kc x nc   alpha l_e d_{e,i} / (2 A_i) */
    const int64_t* edge_cells;   /* 2 x ne    (left, right) */
    const int64_t* edge_vertices;/* 2 x ne */
    const double* edge_height;   /* 2 x ne    alpha d_{e,i} / d_e */
    const double* inverse_dual;  /* ne        1 / d_e */
    const int64_t* vertex_edges; /* kv x nv */
    const double* curl;          /* kv x nv   t_{e,v} d_e / A_v */
    const int64_t* vertex_cells; /* kv x nv */
    const double* vertex_height; /* kv x nv   K_{i,v} / A_v */
    const double* coriolis;      /* nv        f at the vertices */
    const int64_t* trisk_edges;  /* kt x ne */
    const double* trisk;         /* kt x ne   w_{e,e'} l_{e'} / d_e */
    const double* cell_areas;    /* nc */
    const double* vertex_areas;  /* nv */
    double gravity;
    /* vertex_kinetic (alpha != 1), all NULL otherwise */
    const int64_t* cell_vertices;/* kc x nc */
    const double* vk_kites;      /* kc x nc   (1 - alpha) K_{i,v} / A_i */
    const double* vk_weights;    /* kv x nv   A_{e,v} / A_v */
    const double* vk_height;     /* 2 x ne    2 (1 - alpha) A_{e,v} / (l_e d_e) */
    /* gassmann (pv_flux != TRISK_PV_TRISK), all NULL otherwise */
    const int64_t* g_edges;      /* kg x ne */
    const int64_t* g_pv_edges;   /* kg x ne */
    const double* g_weights;     /* kg x ne */
    const double* g_rhombus;     /* 2 x ne    A_v / (A_v1 + A_v2) */
    const double* g_coriolis;    /* ne        f at the edge points */
    int32_t pv_flux;             /* TRISK_PV_* */
    int32_t reserved;
    const double* cell_points;   /* nc x 3 (NULL: keep the given order) for the Hilbert layout */
    /* TRiSKGrid fields the coefficients were folded from (TRiSKGrid.jl:37-60), optional;
     * they enable TRISK_PATH_PAIR, which is only used when `trisk`, `divergence` and
     * `inverse_dual` agree with their recomputation from these fields. */
    const double* primal_lengths; /* ne */
    const double* dual_lengths;   /* ne */
    const double* cell_kites;     /* kc x nc */
} trisk_model_desc;

/* Kernel paths.  TABLES streams the reference's coefficient tables and evaluates every
 * formula in the reference's order (FP64: bit-exact with the reference algorithm).  PAIR
 * rebuilds the 15 antisymmetric TRiSK weights of each cell's edge pairs from its kite
 * fractions (grid fields) instead of streaming ten coefficients and indices per edge (fewer
 * bytes; equal to the tables up to rounding).  AUTO: TABLES for TRISK_FP64, PAIR (when the grid fields
 * are given and consistent) for TRISK_FP32_BULK. */
enum { TRISK_PATH_AUTO = 0, TRISK_PATH_TABLES = 1, TRISK_PATH_PAIR = 2 };

typedef struct {
    int32_t precision;           /* TRISK_FP64 or TRISK_FP32_BULK */
    int32_t device;              /* HIP device ordinal */
    int32_t block_size;          /* threads per workgroup (0: library default) */
    int32_t path;                /* TRISK_PATH_*: which kernels run the TRiSK/cell-KE step */
    int32_t narrow_limit;        /* FP32-bulk pair path and ICON model: largest 16-bit index
                                    delta used (0: 32767); entities beyond it read int32
                                    indices */
    int32_t reserved[3];
} trisk_options;

typedef struct trisk_handle trisk_handle;

/* Thread-local text of the last error. */
const char* trisk_last_error(void);
/* Text of the last error of a call on handle h (kept in the handle until the next error
 * on it; "" if none). */
const char* trisk_handle_last_error(const trisk_handle* h);

/* Host-only validation of a descriptor (every index in range, table widths), and
 * whether TRISK_PATH_PAIR is available (*recomputed); the reason is in
 * trisk_last_error().  Touches no device. */
int trisk_check_tables(const trisk_model_desc* desc, int32_t* recomputed);

/* TRiSKShallowWater(grid; ...; backend) (TRiSKShallowWater.jl:94-170): validates and
 * re-lays out the tables and uploads them.  Synchronous. */
int trisk_create(const trisk_model_desc* desc, const trisk_options* opt, trisk_handle** out);
int trisk_destroy(trisk_handle* h);

/* ne + nc (velocity_dofs, height_dofs; TRiSKShallowWater.jl:173-174). */
int trisk_state_length(const trisk_handle* h, int64_t* n);

/* new -> reference index maps of cells (nc), edges (ne) and vertices (nv); any may be NULL. */
int trisk_permutation(const trisk_handle* h, int64_t* cells, int64_t* edges, int64_t* vertices);

/* initial_state / to_backend: host [u; h] in reference order -> device state.  Synchronous. */
int trisk_upload_state(trisk_handle* h, const double* y_host, double* y_dev, void* stream);
/* gather_state: device state -> host [u; h] in reference order.  Synchronous. */
int trisk_download_state(trisk_handle* h, const double* y_dev, double* y_host, void* stream);
/* A device field on cells (kind 0), edges (1) or vertices (2) -> host, reference order.  Synchronous. */
int trisk_download_field(trisk_handle* h, int32_t kind, const double* x_dev, double* x_host, void* stream);
/* Host field in reference order -> device, internal order.  Synchronous. */
int trisk_upload_field(trisk_handle* h, int32_t kind, const double* x_host, double* x_dev, void* stream);

/* tendency!(dy, y, m): dy = [du; dh] of the state y (both device, internal order). */
int trisk_tendency(trisk_handle* h, const double* y_dev, double* dy_dev, void* stream);

/* nsteps times rk4_step!(y, dt, tendency!, RK4Cache(y), m) in place, with the RK4 stage
 * update fused into the tendency kernels (same arithmetic as the reference broadcasts). */
int trisk_rk4_steps(trisk_handle* h, double* y_dev, double dt, int64_t nsteps, void* stream);

/* Cell-pair path: trisk_rk4_steps keeps each stage's mass flux G_e and velocity u_e as one
 * (G, u) pair per edge (on = 1, the default on FP32-bulk handles), which the next stage's
 * vertex and cell kernels gather, or as two arrays (on = 0, the default on FP64 handles, whose
 * 16-byte pairs would double the lines the vertex kernel gathers u from).  The same values
 * either way (bitwise); the pairs halve the cell kernel's gathers at the edges.  No counterpart in the reference (its kernels read
 * u and F from separate arrays, TRiSKShallowWater.jl:202-305): a layout switch for tests and
 * measurement. */
int trisk_set_pair_plane(trisk_handle* h, int32_t on);

/* Replay trisk_rk4_steps through a captured HIP graph of 8 RK4 steps (on = 1; default 0).
 * The graph is captured on the call's stream (which must not be NULL) the first time and
 * again whenever y_dev or dt change; runs shorter than 8 steps and timed runs launch the
 * kernels directly.  Same results as direct launches. */
int trisk_set_graphs(trisk_handle* h, int32_t on);

/* out[0] = total_mass, out[1] = total_energy, out[2] = potential_enstrophy, summed with
 * Neumaier compensation in a fixed order (bitwise reproducible).  Synchronous. */
int trisk_diagnostics(trisk_handle* h, const double* y_dev, double out[3], void* stream);

/* trisk_diagnostics without allocation or synchronisation.  No counterpart in the
 * reference: it is for callers that cannot allocate or wait (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  out_dev (device memory, 3
 * doubles) receives the values trisk_diagnostics returns, bitwise, stream-ordered on
 * `stream`.  The block partials live in a static device array of the library (48 KiB per
 * GPU, held until the process ends): the handle's reduction may use at most 1024 blocks
 * (TRISK_EINVAL otherwise).  All calls on one GPU share that array, so calls that overlap
 * on one GPU (other streams, other threads) race.  No hipMalloc, hipFree or host
 * allocation. */
int trisk_diagnostics_noalloc(trisk_handle* h, const double* y_dev, double* out_dev, void* stream);

/* Failure detection (no counterpart in the reference, which has no NaN checks).
 * trisk_check_state: counts[0] = entries of u that are NaN or Inf, counts[1] = the same for
 * h, counts[2] = vertices whose height h_v (the divisor of q_v = (zeta_v + f_v) / h_v,
 * TRiSKShallowWater.jl:258) is not > 0 or is NaN.  Synchronous.
 * trisk_set_finite_check(h, 1): trisk_diagnostics also runs this check and returns
 * TRISK_ENONFINITE (out[] still filled) when any count is nonzero; default 0, the
 * reference's behaviour of returning NaN sums. */
int trisk_check_state(trisk_handle* h, const double* y_dev, int64_t counts[3], void* stream);
int trisk_set_finite_check(trisk_handle* h, int32_t on);

/* kinetic_energy!(m, u, h): K and B at the cells (device, internal order). */
int trisk_kinetic_energy(trisk_handle* h, const double* y_dev, double* K_dev, double* B_dev, void* stream);
/* vertex_pv!(m, u, h): q_v, zeta_v, h_v at the vertices (device, internal order). */
int trisk_vertex_pv(trisk_handle* h, const double* y_dev, double* qv_dev, double* zv_dev, double* hv_dev,
                    void* stream);
/* relative_vorticity(m, y): zeta at the vertices. */
int trisk_relative_vorticity(trisk_handle* h, const double* y_dev, double* zeta_dev, void* stream);
/* tangential_velocity(m, u): (1/d_e) sum_e' w l_e' u_e' at the edges; u_dev has ne entries. */
int trisk_tangential_velocity(trisk_handle* h, const double* u_dev, double* vt_dev, void* stream);

/* Average duration in ms of each hot-path kernel over the last trisk_rk4_steps call made
 * with timing enabled (HIP events on the call's stream); names[i] receive static strings.
 * trisk_set_timing(h, 1) enables the events (they add no synchronisation inside the call). */
int trisk_set_timing(trisk_handle* h, int32_t on);
int trisk_kernel_times(trisk_handle* h, int32_t max, const char** names, double* avg_ms, int64_t* launches,
                       int32_t* count);

/* Algorithmic HBM bytes of one launch of each hot-path kernel of this handle's path
 * (each array element the kernel must touch counted once), with the kernel names, for an
 * RK4 stage of trisk_rk4_steps after the first of the call.  On the cell-pair path the edge
 * kernel of each such stage also forms the mass flux G of the state it writes, so the next
 * stage's state kernel covers the vertices only (the cell kernel forms K and B).  With vertex
 * kinetic energy whose cell weights are all 0, the table path's first kernel covers the
 * vertices only and the second forms B. */
int trisk_kernel_bytes(trisk_handle* h, int32_t max, const char** names, double* bytes, int32_t* count);

/* Cells, vertices and edges whose indices do not fit the 16-bit deltas of the FP32-bulk
 * pair path or of the ICON model's tables (they are read from int32 tables; for ICON the
 * edge count adds up its cell, vertex and RBF tables). */
int trisk_wide_counts(const trisk_handle* h, int64_t counts[3]);

/* Whether this handle runs TRISK_PATH_PAIR (*recomputed = 1) or the table path, and why
 * (text owned by the handle). */
int trisk_info(const trisk_handle* h, int32_t* recomputed, const char** note);

/* --- ICON triangle model ----------------------------------------------------------
 * ICONShallowWater(grid; gravity, coriolis, rbf_scale, divergence_damping, kinetic_gradient)
 *   src/FiniteVolumes/ICONShallowWater.jl:96-143                            -> icon_create
 * The triangular C grid of ICON's dynamical core: RBF tangential velocity, bilinear
 * edge-to-cell kinetic energy, ICON's kinetic-energy gradient, vertex vorticity averaged to
 * the edges and optional fourth-order divergence damping.  icon_create returns a
 * trisk_handle, and the generic calls take it as they take a TRiSK handle (both models
 * are the reference's CGridShallowWater, with the same state layout [u; h]):
 * trisk_destroy, trisk_state_length, trisk_permutation, trisk_upload/download_state and
 * _field, trisk_tendency (tendency!, ICONShallowWater.jl:191-207), trisk_rk4_steps,
 * trisk_set_graphs, trisk_diagnostics, trisk_vertex_pv, trisk_relative_vorticity,
 * trisk_tangential_velocity (the RBF tangential velocity, :219-224), trisk_set_timing,
 * trisk_kernel_times, trisk_kernel_bytes, trisk_wide_counts and trisk_info.  The TRiSK-only
 * calls (trisk_kinetic_energy, trisk_tendency_vjp, trisk_rk4_adjoint) return
 * TRISK_EINVAL on an ICON handle.  TRISK_FP64 follows the reference's arithmetic bit for
 * bit; TRISK_FP32_BULK runs FP32 stencils with the FP64 state, heights, pressure gradient,
 * continuity (mass conserved exactly) and RK4 accumulation, and FP64 diagnostics. */
typedef struct {
    int64_t nc, ne, nv;
    int64_t kv;                        /* vertex table width (6 on icosahedral meshes) */
    int64_t kr;                        /* RBF stencil width (4 on triangles) */
    const int64_t* cell_edges;         /* 3 x nc */
    const double* divergence;          /* 3 x nc: n l / A */
    const double* bilinear;            /* 3 x nc: e_bln_c_s (bilinear_weights, :82-94) */
    const int64_t* edge_cells;         /* 2 x ne */
    const int64_t* edge_vertices;      /* 2 x ne */
    const double* edge_height;         /* 2 x ne: c_lin_e */
    const double* gradient;            /* 2 x ne: coeff_gradekin */
    const double* inverse_dual;        /* ne */
    const int64_t* rbf_edges;          /* kr x ne (rbf_coefficients, TRiSKGrid.jl:254-281) */
    const double* rbf;                 /* kr x ne */
    const double* edge_coriolis;       /* ne */
    const int64_t* vertex_edges;       /* kv x nv */
    const double* curl;                /* kv x nv: t d / A_v */
    const int64_t* vertex_cells;       /* kv x nv */
    const double* vertex_height;       /* kv x nv: K_iv / A_v (diagnostics) */
    const double* coriolis;            /* nv (diagnostics) */
    const double* cell_areas;          /* nc */
    const double* vertex_areas;        /* nv */
    double gravity;
    const double* divergence_damping;  /* ne (nu_4 per edge), or NULL without damping */
    const double* cell_points;         /* nc x 3: Hilbert order of the cells, or NULL */
} icon_model_desc;

/* Host-only validation of an ICON descriptor (every index in range).  Touches no device. */
int icon_check_tables(const icon_model_desc* desc);
/* ICONShallowWater(grid; ...; backend): validates, re-lays out and uploads the tables.
 * Synchronous. */
int icon_create(const icon_model_desc* desc, const trisk_options* opt, trisk_handle** out);

/* --- MPAS-Gassmann TRiSK operators with levels ---------------------------------------
 * The explicit tendencies of one stage of MPASGassmannNonhydrostatic,
 *   slow_tendencies!(m, ys, pressure)  src/Nonhydrostatic/MPAS/Gassmann/Stepping.jl:22-56
 * with the TRiSK operators of src/Nonhydrostatic/MPAS/Operators.jl:6-154 (diagnose, vertex
 * and cell-from-vertex kinetic energy, w kinetic energy, eta, edge density and mass flux,
 * vertex and edge PV, the w tendencies) and vn_tendency_kernel! with edge_w
 * (Gassmann/Tendencies.jl:4-67).  Over orography (terrain-following levels) also the
 * contravariant correction w_c (concorr_cell_kernel!, concorr_interface_kernel!,
 * contravariant_w_kernel!, w_concorr_tendency_kernel!; Stepping.jl:58-76), the deviation of
 * pi from the reference atmosphere (exner_deviation_centred_kernel!) and the pressure
 * gradient along the levels or at constant height (pressure_gradient, Tendencies.jl:13-30;
 * height_difference, Common/CGridOperators.jl:193-200).  Fields are (points x levels) with the
 * point index fastest, as the reference's (n, nlev) arrays: vn (ne x nlev), w (nc x nlev+1),
 * rho and rhotheta (nc x nlev).  The library renumbers the points (cells along the Hilbert
 * curve, edges and vertices by first appearance); mpas_ops_upload_field / _download_field
 * convert. */
typedef struct mpas_ops mpas_ops;

typedef struct {
    int64_t nc, ne, nv, nlev;
    int64_t kc, kv, kt;             /* cell, vertex and TRiSK table widths */
    /* fields of MPASGassmannNonhydrostatic (MPASGassmannNonhydrostatic.jl:133-176) */
    const int64_t* cell_edges;      /* kc x nc */
    const double* kinetic;          /* kc x nc: alpha l d_ei / 2A */
    const double* inner;            /* kc x nc: l d_ei / A */
    const int64_t* edge_cells;      /* 2 x ne */
    const int64_t* edge_vertices;   /* 2 x ne */
    const double* edge_weight;      /* 2 x ne: d_ei / d_e */
    const double* inverse_dual;     /* ne: 1 / (a d_e) */
    const int64_t* vertex_edges;    /* kv x nv */
    const double* curl;             /* kv x nv: t d / (a A_v) */
    const int64_t* vertex_cells;    /* kv x nv */
    const double* vertex_weight;    /* kv x nv: K_iv / A_v */
    const double* coriolis;         /* nv */
    const int32_t* trisk_index;     /* ne x kt x 3, edge first: e', and the edges a, b of (q_a + q_b)/2 */
    const double* trisk;            /* ne x kt, edge first: w l' / 2d */
    const double* rhombus;          /* 2 x ne */
    const double* edge_coriolis;    /* ne */
    /* column metrics: cells dz (nc x nlev), dzh and wlow (nc x nlev+1);
     * edges dz (ne x nlev), dzh (ne x nlev+1) (metric.dz, metric.dzh) */
    const double* dz;
    const double* dzh;
    const double* wlow;
    const double* edge_dz;
    const double* edge_dzh;
    /* vertex kinetic energy (kinetic_energy != :cell), else NULL */
    const int64_t* cell_vertices;   /* kc x nc */
    const double* vk_kites;         /* kc x nc: (1 - alpha) K_iv / A_i */
    const double* vk_weights;       /* kv x nv: A_ev / A_v */
    int32_t iap_pv;                 /* 1: rhombus PV (pv_flux = :iap), 0: mean of the vertex PV */
    int32_t vertical;               /* 0 advective, 1 vector invariant, 2 ICON-IAP mass fluxes */
    int32_t five_band;              /* vertical_solver = :five_band */
    int32_t reserved;
    const double* cell_points;      /* nc x 3: Hilbert order of the cells, or NULL */
    /* orography (MPASGassmannNonhydrostatic.jl:220-240): 0 flat ground (the fields below
     * unused), 1 terrain-following levels with the pressure gradient along the levels,
     * 2 with the pressure gradient at constant height (pressure_gradient = :height) */
    int32_t orography;
    int32_t reserved2;
    const double* z_full;           /* nc x nlev: full-level heights of the cells (metric.z_full) */
    const double* pi_ref;           /* nc x nlev: reference-atmosphere Exner pressure there (metric.πref) */
    const double* ground_q;         /* nc x 3: quadratic extrapolation weights to the ground (metric.q) */
    const double* ddxn;             /* ne x nlev: slope of the levels along the edge normal (metric.ddxn) */
    const int32_t* zlev;            /* 2 x ne x nlev: 0-based level below the edge height in each edge cell
                                       (metric.zlev - 1); orography 2 only, nlev >= 2 */
    const double* zwt;              /* 2 x ne x nlev: the interpolation weights (metric.zwt); orography 2 only */
} mpas_ops_desc;

/* Work fields of a slow-tendency call (mpas_ops_field): kind, points x levels. */
enum {
    MPAS_THETA = 0,  /* theta, cells x nlev */
    MPAS_EXNER = 1,  /* pi, cells x nlev */
    MPAS_K = 2,      /* K (horizontal, vertex and w parts as the options add them), cells x nlev */
    MPAS_QV = 3,     /* q_v, vertices x nlev */
    MPAS_ZETA = 4,   /* zeta_v, vertices x nlev */
    MPAS_RHOV = 5,   /* rho_v, vertices x nlev */
    MPAS_QE = 6,     /* q_e, edges x nlev */
    MPAS_RHOE = 7,   /* rho_e, edges x nlev */
    MPAS_FS = 8,     /* rho_e vn, edges x nlev */
    MPAS_ETA = 9,    /* eta, edges x nlev+1 */
    MPAS_TW = 10,    /* w tendency, cells x nlev+1 */
    MPAS_TVN = 11,   /* vn tendency, edges x nlev */
    MPAS_KV = 12,    /* vertex kinetic energy, vertices x nlev */
    MPAS_WCF = 13,   /* contravariant correction w_c at the full levels, cells x nlev (orography) */
    MPAS_WCH = 14,   /* w_c at the interfaces, cells x nlev+1 (orography) */
    MPAS_WHAT = 15,  /* w - w_c, cells x nlev+1 (orography) */
    MPAS_PIP = 16,   /* pi - pi_ref, cells x nlev (orography, pressure > 0) */
    MPAS_DPIDZ = 17, /* d(pi - pi_ref)/dz, centred, cells x nlev (orography, pressure > 0) */
    MPAS_FIELDS = 18
};

/* Host-only validation of a descriptor (index ranges, widths).  Touches no device. */
int mpas_ops_check(const mpas_ops_desc* desc);
/* Validates, renumbers and uploads the tables and allocates the work fields.  Synchronous. */
int mpas_ops_create(const mpas_ops_desc* desc, int32_t device, mpas_ops** out);
int mpas_ops_destroy(mpas_ops* h);
/* new -> reference index maps of cells, edges and vertices; any may be NULL. */
int mpas_ops_permutation(const mpas_ops* h, int64_t* cells, int64_t* edges, int64_t* vertices);
/* A host field on cells (kind 0), edges (1) or vertices (2) with `levels` levels (point
 * fastest, reference numbering) -> device (internal numbering), and back.  Synchronous. */
int mpas_ops_upload_field(mpas_ops* h, int32_t kind, int64_t levels, const double* host, double* dev, void* stream);
int mpas_ops_download_field(mpas_ops* h, int32_t kind, int64_t levels, const double* dev, double* host,
                            void* stream);
/* slow_tendencies!(m, ys, pressure): theta, pi, K, the PV, eta, tw and tvn of the stage
 * state (vn, w, rho, rhotheta), device fields in the internal numbering, into the work
 * fields. */
int mpas_slow_tendencies(mpas_ops* h, const double* vn, const double* w, const double* rho, const double* rhotheta,
                         double pressure, void* stream);
/* Device pointer and shape (points, levels) of a work field (MPAS_*).  The kernels keep the
 * pairs MPAS_FS/MPAS_QE, MPAS_THETA/MPAS_EXNER, MPAS_ZETA/MPAS_RHOV and MPAS_RHOE/MPAS_ETA
 * interleaved in one plane each (one 16-byte gather for both); for these fields the call waits
 * for the device, writes the two fields of the pair from the plane and returns the asked one. */
int mpas_ops_field(mpas_ops* h, int32_t which, double** dev, int64_t* points, int64_t* levels);
/* HIP-event timing of the kernels of mpas_slow_tendencies (as trisk_set_timing and
 * trisk_kernel_times), and their algorithmic HBM bytes per launch. */
int mpas_ops_set_timing(mpas_ops* h, int32_t on);
int mpas_ops_kernel_times(mpas_ops* h, int32_t max, const char** names, double* avg_ms, int64_t* launches,
                          int32_t* count);
int mpas_ops_kernel_bytes(mpas_ops* h, int32_t max, const char** names, double* bytes, int32_t* count);

/* --- adjoint ------------------------------------------------------------------------
 * No counterpart in the reference (it has no adjoint).  These are the pullbacks of
 * tendency! (src/FiniteVolumes/TRiSKShallowWater.jl:379-394) and of nsteps rk4_step!
 * (src/TimeSteppers/RungeKutta.jl:18-28) that a ChainRulesCore rrule calls: the transposed
 * linearisation of the handle's own forward arithmetic, on either path: the cell-pair path
 * (TRiSK PV flux, cell kinetic energy) or the table kernels (every PV flux and
 * kinetic-energy form; meshes with three edges per vertex).  Gather-only kernels, so the
 * result does not depend on scheduling.  ICON handles return TRISK_EINVAL. */

/* ybar = (d tendency / dy)^T dybar at the state y (device vectors, internal order; ybar
 * must not alias y or dybar). */
int trisk_tendency_vjp(trisk_handle* h, const double* y_dev, const double* dybar_dev, double* ybar_dev,
                       void* stream);

/* The cotangent ybar0 of y0 for the cotangent ybarT of y_T, y_T being y0 after nsteps RK4
 * steps of dt (as trisk_rk4_steps).  The forward run keeps the state every
 * checkpoint_every steps (0: 8) and each segment is recomputed from its checkpoint during
 * the backward sweep, keeping every step's stage states for its transposition.  yT_dev
 * (may be NULL) receives y_T.  ybar0 may alias ybarT.  The handle keeps
 * (ceil(nsteps / checkpoint_every) + checkpoint_every + 5) state vectors of work memory,
 * plus per step of a segment its three stage states and, on the cell-pair path, four copies
 * of q_v and G (grown as needed), until trisk_destroy.  With trisk_set_graphs(h, 1) and a
 * stream, the whole sweep is captured as one HIP graph and replayed while the pointers,
 * dt, nsteps and checkpoint_every stay those of the capture.  Synchronous. */
int trisk_rk4_adjoint(trisk_handle* h, const double* y0_dev, double dt, int64_t nsteps, int64_t checkpoint_every,
                      const double* ybarT_dev, double* ybar0_dev, double* yT_dev, void* stream);

/* --- ensembles ----------------------------------------------------------------------
 * No counterpart in the reference, which advances each model with its own rk4_step! loop
 * (src/TimeSteppers/RungeKutta.jl:18-28, examples/trisk_shallow_water.jl:92-100).  An
 * ensemble holds `nlevels` models (meshes of any resolution; TRiSK PV flux and cell
 * kinetic energy, i.e. TRISK_PATH_PAIR) with members[l] states each, and advances them
 * together: each of the three kernels of a tendency is one launch over every member of
 * every level.
 *
 * The ensemble state is one device vector of trisk_ensemble_state_length() doubles: the
 * blocks of the levels in order, level l's block [u (ne x M); h (nc x M)] holding member m
 * of entity j at j*M + m, entities in the level model's internal order.  Fill and read it
 * member by member with trisk_ensemble_set_member / get_member and a single-member device
 * state of the level's model (trisk_ensemble_level, trisk_upload_state,
 * trisk_download_state).  The level models belong to the ensemble.  Calls on one ensemble
 * must not overlap. */
typedef struct trisk_ensemble trisk_ensemble;

/* How one kernel's tiles (block_size items each) are laid on the device. */
enum { TRISK_SPLIT_CHUNKED = 0,     /* XCD x takes the x-th eighth of every level's range */
       TRISK_SPLIT_BLOCKED = 1,     /* XCD x takes the x-th eighth of all ranges laid end to end */
       TRISK_SPLIT_INTERLEAVED = 2  /* as CHUNKED; the state kernel's first-stage launch takes the
                                       vertices and edges of each stretch of the curve together */ };
enum { TRISK_ENS_STATE = 0, TRISK_ENS_CELL = 1, TRISK_ENS_EDGE = 2 };  /* the three kernels */

typedef struct {
    int32_t block_size;          /* 128, 256 or 512 work-items per workgroup */
    int32_t blocks_per_cu;       /* persistent grid of blocks_per_cu workgroups per CU; 0: one per tile */
    int32_t split;               /* TRISK_SPLIT_* */
    int32_t members_per_lane;    /* 1, 2, 4 or 8 consecutive members per work-item (0: 1); must divide
                                    every level's member count, and the state vectors passed to the
                                    ensemble must then be aligned to 8 x members_per_lane bytes */
} trisk_schedule;

/* Creates the level models (trisk_create with TRISK_PATH_PAIR; opt->path must be AUTO or
 * PAIR) and the member-interleaved work vectors.  Synchronous. */
int trisk_ensemble_create(const trisk_model_desc* const* levels, const int32_t* members, int32_t nlevels,
                          const trisk_options* opt, trisk_ensemble** out);
int trisk_ensemble_destroy(trisk_ensemble* e);
int trisk_ensemble_state_length(const trisk_ensemble* e, int64_t* n);
/* Level l's model (owned by the ensemble), the offset of its block and its member count;
 * any output may be NULL. */
int trisk_ensemble_level(trisk_ensemble* e, int32_t level, trisk_handle** model, int64_t* offset,
                         int32_t* members);
/* Member `member` of level `level`: ensemble state <-> single-model device state. */
int trisk_ensemble_get_member(trisk_ensemble* e, int32_t level, int32_t member, const double* y_dev,
                              double* y_member_dev, void* stream);
int trisk_ensemble_set_member(trisk_ensemble* e, int32_t level, int32_t member, const double* y_member_dev,
                              double* y_dev, void* stream);
/* tendency! of every member, dy in the layout of y. */
int trisk_ensemble_tendency(trisk_ensemble* e, const double* y_dev, double* dy_dev, void* stream);
/* nsteps RK4 steps of every member, level l with time step dt[l]; bitwise equal to
 * trisk_rk4_steps of each member on its level's model. */
int trisk_ensemble_rk4_steps(trisk_ensemble* e, double* y_dev, const double* dt, int64_t nsteps, void* stream);
/* Replay trisk_ensemble_rk4_steps through a captured graph of 8 steps (as trisk_set_graphs). */
int trisk_ensemble_set_graphs(trisk_ensemble* e, int32_t on);
/* on != 0: every RK stage (and tendency) of the ensemble is one cooperative launch
 * (hipLaunchCooperativeKernel) whose workgroups run the state, cell and edge phases with a
 * grid-wide barrier between them, instead of three launches; its grid holds as many workgroups
 * as the device keeps resident at once, and each phase walks its tiles XCD by XCD.  Results
 * are bitwise those of the three launches.  Needs every level's member count divisible by 4
 * and a device with cooperative launches (TRISK_EINVAL otherwise); graph replay is not used
 * while it is on, and trisk_ensemble_autotune tunes the three-launch schedules.  The grid-wide
 * barriers make it slower than the three launches of a replayed graph (the default).  No counterpart
 * in the reference, which steps one model at a time (RungeKutta.jl:18-28). */
int trisk_ensemble_set_fused(trisk_ensemble* e, int32_t on);
/* out[3k .. 3k+2] = trisk_diagnostics of member k, members numbered level by level.  Synchronous. */
int trisk_ensemble_diagnostics(trisk_ensemble* e, const double* y_dev, double* out, void* stream);
/* The schedules of the three kernels (TRISK_ENS_*); the default is {256, 0, CHUNKED, 1}. */
int trisk_ensemble_set_schedule(trisk_ensemble* e, const trisk_schedule sched[3]);
int trisk_ensemble_get_schedule(const trisk_ensemble* e, trisk_schedule sched[3]);
/* Times `steps` RK4 steps under each candidate schedule (block size x persistence x
 * split x members per lane) on a device copy of y_dev (y_dev is not changed), keeps per kernel the one with
 * the shortest mean launch, and returns it with that time (either output may be NULL).
 * Synchronous. */
int trisk_ensemble_autotune(trisk_ensemble* e, const double* y_dev, const double* dt, int32_t steps, void* stream,
                            trisk_schedule best[3], double best_ms[3]);
/* Kernel timing and algorithmic bytes, as trisk_set_timing / trisk_kernel_times / trisk_kernel_bytes
 * (the three kernels, then the fused stage of trisk_ensemble_set_fused: four entries). */
int trisk_ensemble_set_timing(trisk_ensemble* e, int32_t on);
int trisk_ensemble_kernel_times(trisk_ensemble* e, int32_t max, const char** names, double* avg_ms,
                                int64_t* launches, int32_t* count);
int trisk_ensemble_kernel_bytes(trisk_ensemble* e, int32_t max, const char** names, double* bytes,
                                int32_t* count);

/* --- mesh setup on the device ------------------------------------------------------
 * lloyd!(x, triangles; fixed=1:nfixed, tolerance, history, max_iterations)
 * (src/Grids/Icosahedral.jl:126-184) on HIP device `device`: x (host, nv x 3) is moved in
 * place to the centroids of the spherical Voronoi cells of the counter-clockwise
 * triangulation tri (host, nt x 3); star_offsets/stars are the triangle stars of the
 * vertices (dmx_triangle_stars of libdmplex_host.so).  Returns 0 when converged, 1 when
 * max_iterations was reached (the reference's warning); *residual (may be NULL) receives
 * the last residual in units of the mean edge length.  The Delaunay check of the
 * result is left to the caller (dmx_delaunay).  Synchronous. */
int trisk_lloyd(int32_t device, double* x, int64_t nv, const int64_t* tri, int64_t nt, const int64_t* star_offsets,
                const int64_t* stars, int64_t nfixed, double tolerance, int32_t history, int64_t max_iterations,
                int64_t* iterations, double* residual);

#ifdef __cplusplus
}
#endif
#endif
