/*
 * dmplex_host.h -- C ABI of libdmplex_host.so, the host-side (CPU, OpenMP) setup
 * code of the TRiSK shallow-water path: mesh generation, the Plex topology, the
 * TRiSK geometry tables, Hilbert keys and the seeded synthetic initial state.
 *
 * In the Julia drop-in these steps stay in the reference package (they run once,
 * before the hot loop); this library restates them natively so that the Python
 * host mirror and the tests can build level-9/10 meshes in seconds.
 *
 * Conventions (all 0-based):
 *   - Plex points are numbered cells [0, nc), edges [nc, nc+ne), vertices
 *     [nc+ne, nc+ne+nv) (src/Topology/Plex.jl:5-10, shifted to 0-based).
 *   - Tables are "k x n column-major": entry (k, j) of an n-entity table of
 *     width K lives at k + j*K, which is the memory layout of a Julia k x n
 *     Matrix (src/FiniteVolumes/TRiSKGrid.jl:28-31).  Padded slots hold index 0
 *     and weight 0 (the reference's index 1 / weight 0).
 *   - Points (coordinates, centers, edge points, normals) are n x 3, one
 *     contiguous xyz triple per point.
 * Every function returns 0 on success, a negative code on error; the message is
 * available from dmx_last_error().
 */
#ifndef DMPLEX_HOST_H
#define DMPLEX_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Thread-local text of the last error. */
const char* dmx_last_error(void);

/* --- icosahedral mesh, Lloyd SCVT ------------------------------------------------ */

/* Replaces create_spherical_mesh (src/Grids/Icosahedral.jl:22-61) up to the final
 * plex_from_polygons call: bisection of the icosahedron `subdivisions` times, with
 * Lloyd's algorithm after every level when centroidal != 0.  Outputs
 * nv = 10*4^n + 2 vertices (coords, nv x 3) and nt = 20*4^n triangles (nt x 3).
 * iterations (may be NULL) receives the Lloyd iteration count of each level
 * (array of length subdivisions). */
int dmx_icosahedral(int subdivisions, int centroidal, double* coords, int64_t* triangles,
                    int64_t* iterations);

/* Replaces lloyd! (src/Grids/Icosahedral.jl:126-184), in place on x (nv x 3) for the
 * counter-clockwise triangulation tri (nt x 3).  Vertices [0, nfixed) stay fixed.
 * history = 0 is plain Lloyd.  Returns the iteration count in *iterations;
 * error -3 if the result is not Delaunay (the reference's ArgumentError). */
int dmx_lloyd(double* x, int64_t nv, const int64_t* tri, int64_t nt, int64_t nfixed,
              double tolerance, int history, int64_t max_iterations, int64_t* iterations);

/* The 12 vertices and 20 triangles of the icosahedron (Icosahedral.jl:24-35). */
int dmx_icosahedron(double* coords, int64_t* triangles);

/* One bisection level (Icosahedral.jl:38-48): nv + ne new vertices in x_out
 * ((nv + 3*nt/2) x 3) and 4*nt triangles in tri_out. */
int dmx_bisect(int64_t nv, const double* x, int64_t nt, const int64_t* tri, double* x_out, int64_t* tri_out);

/* Counter-clockwise triangle star of every vertex (Icosahedral.jl:130-147), CSR:
 * star_offsets (nv + 1), stars (3 * nt). */
int dmx_triangle_stars(int64_t nv, const int64_t* tri, int64_t nt, int64_t* star_offsets, int64_t* stars);

/* is_delaunay for every triangle (Icosahedral.jl:210-216): 0, or -3 with a message. */
int dmx_delaunay(const double* x, int64_t nv, const int64_t* tri, int64_t nt);

/* Mean length of the first edge of every triangle, the scale of Lloyd's tolerance
 * (Icosahedral.jl:148), summed in a thread-count independent order. */
int dmx_mean_edge_length(const double* x, const int64_t* tri, int64_t nt, double* h);

/* Lloyd step residual only: largest |centroid - x| over free vertices (lloyd_step!,
 * src/Grids/Icosahedral.jl:186-207), without moving x. */
int dmx_lloyd_residual(const double* x, int64_t nv, const int64_t* tri, int64_t nt,
                       int64_t nfixed, double* residual);

/* --- Plex -------------------------------------------------------------------------- */

typedef struct dmx_plex_build dmx_plex_build;

/* Replaces plex_from_polygons (src/Topology/Plex.jl:49-97) on the unit sphere:
 * polygon c has vertices poly_vertices[poly_offsets[c] .. poly_offsets[c+1]).
 * Cells given clockwise are reversed; edges are numbered by first appearance.
 * dims[0..2] receive nc, ne, nv and dims[3] the cone array length. */
int dmx_plex_from_polygons(int64_t npoly, const int64_t* poly_offsets, const int64_t* poly_vertices,
                           int64_t nv, const double* coords, dmx_plex_build** out, int64_t dims[4]);
/* Copies the CSR arrays (npoints+1 offsets, dims[3] cones/orientations/supports). */
int dmx_plex_export(const dmx_plex_build* b, int64_t* cone_offsets, int64_t* cones,
                    int8_t* orientations, int64_t* support_offsets, int64_t* supports);
void dmx_plex_free(dmx_plex_build* b);

/* Read-only view of a Plex held by the caller. */
typedef struct {
    int64_t nc, ne, nv;
    const int64_t* cone_offsets;     /* npoints + 1 */
    const int64_t* cones;
    const int8_t* orientations;      /* 0 or -1 per cone entry */
    const int64_t* support_offsets;  /* npoints + 1 */
    const int64_t* supports;
    const double* coordinates;       /* nv x 3 */
} dmx_plex;

/* Polygons of the Voronoi dual of a triangle Plex (create_hexagon_mesh,
 * src/Grids/Hexagonal.jl:11-17 with dual_polygons/sort_ring,
 * src/Grids/FileGrids.jl:28-42): poly_offsets (nv+1), poly_vertices (3*nt) and
 * centers (nt x 3, the circumcenters, cell_centers(...; method=:circumcenter),
 * src/Grids/CellCenters.jl:37-42). */
int dmx_dual_polygons(const dmx_plex* primal, int64_t* poly_offsets, int64_t* poly_vertices,
                      double* centers);

/* Largest |cos| between primal edges and the great circles of their cell centers
 * (orthogonality_defect, src/Grids/CellCenters.jl:90-101). */
int dmx_orthogonality_defect(const dmx_plex* dm, const double* centers, double* defect);

/* Checks the mesh invariants of check_plex (src/Topology/Plex.jl:329-360);
 * returns 0 if they hold, -4 with a message otherwise. */
int dmx_check_plex(const dmx_plex* dm);

/* --- TRiSK geometry -------------------------------------------------------------- */

/* Output arrays of TRiSKGrid (src/FiniteVolumes/TRiSKGrid.jl:37-60), caller-allocated.
 * kc = max cell size, kv = max vertex degree, kt = max TRiSK stencil. */
typedef struct {
    int64_t kc, kv, kt;
    double* edge_points;      /* ne x 3 */
    double* edge_normals;     /* ne x 3 */
    double* primal_lengths;   /* ne */
    double* dual_lengths;     /* ne */
    int64_t* edge_cells;      /* 2 x ne (left, right) */
    int64_t* edge_vertices;   /* 2 x ne */
    double* cell_distances;   /* 2 x ne */
    double* cell_areas;       /* nc */
    double* vertex_areas;     /* nv */
    int64_t* cell_edges;      /* kc x nc */
    double* cell_signs;       /* kc x nc */
    int64_t* cell_vertices;   /* kc x nc */
    double* cell_kites;       /* kc x nc */
    int64_t* vertex_edges;    /* kv x nv */
    double* vertex_signs;     /* kv x nv */
    int64_t* vertex_cells;    /* kv x nv */
    double* vertex_kites;     /* kv x nv */
    int64_t* trisk_edges;     /* kt x ne */
    double* trisk_weights;    /* kt x ne */
} dmx_grid_tables;

/* Table widths of TRiSKGrid for this mesh (kc, kv, kt). */
int dmx_trisk_grid_widths(const dmx_plex* dm, int64_t widths[3]);
/* TRiSKGrid(dm; centers) (src/FiniteVolumes/TRiSKGrid.jl:76-162). Error -2 for an
 * open mesh (the reference's ArgumentError at :92). */
int dmx_trisk_grid(const dmx_plex* dm, const double* centers, dmx_grid_tables* g);

/* vertex_edge_areas (src/FiniteVolumes/TRiSKGrid.jl:179-191): by_vertex (kv x nv),
 * by_edge (2 x ne). */
int dmx_vertex_edge_areas(int64_t nc, int64_t ne, int64_t nv, const dmx_grid_tables* g,
                          const double* centers, const double* vertices, double* by_vertex,
                          double* by_edge);

/* gassmann_coefficients (src/FiniteVolumes/TRiSKGrid.jl:214-251): edges, pv_edges,
 * weights (12 x ne).  Error -2 when a vertex does not have three edges. */
int dmx_gassmann_coefficients(int64_t nc, int64_t ne, int64_t nv, const dmx_grid_tables* g,
                              const double* vertices, int64_t* edges, int64_t* pv_edges,
                              double* weights);

/* --- ICON triangle model tables -------------------------------------------------- */

/* rbf_coefficients(g; scale) (src/FiniteVolumes/TRiSKGrid.jl:254-281): for every edge the
 * other edges of its two cells, the inverse-multiquadric vector-RBF weights of the
 * tangential velocity, normalized along x_e x n_e.  edges, weights: kr x ne (pads index
 * 0, weight 0).  With edges == NULL only *width (the longest stencil) is returned.
 * Error -2 when an RBF matrix is not positive definite. */
int dmx_rbf_coefficients(int64_t nc, int64_t ne, const dmx_grid_tables* g, double scale, int64_t kr,
                         int64_t* edges, double* weights, int64_t* width);
/* bilinear_weights(g) (src/FiniteVolumes/ICONShallowWater.jl:79-96): 3 x nc weights with
 * sum 1 and no first moment in the tangent plane of the cell center.  Error -3 unless
 * every cell is a triangle. */
int dmx_bilinear_weights(int64_t nc, const dmx_grid_tables* g, const double* centers, double* weights);

/* --- ICON-IAP TRiSK tables of the MPAS-Gassmann operators ------------------------ */

/* trisk_pv_edges(g) (src/FiniteVolumes/TRiSKGrid.jl:318-362): kt x ne, for TRiSK slot k of
 * edge e the edge whose rhombus PV replaces the mean (q_e + q_e')/2 (i_cori_method = 4),
 * -1 for the mean.  Error -3 if the TRiSK table is not in the order of the cell walk. */
int dmx_trisk_pv_edges(int64_t nc, int64_t ne, const dmx_grid_tables* g, int64_t* pv_edges);
/* iap_trisk_weights(g) (TRiSKGrid.jl:364-403): kt x ne TRiSK weights from the kite fractions
 * of planar right triangles; vertices nv x 3. */
int dmx_iap_trisk_weights(int64_t nc, int64_t ne, const dmx_grid_tables* g, const double* vertices,
                          double* weights);

/* --- ordering -------------------------------------------------------------------- */

/* Hilbert key of integer points (hilbert_key, src/Topology/Ordering.jl:10-41):
 * n points x (n x 3), each coordinate in [0, 2^bits). */
int dmx_hilbert_keys(int64_t n, const uint64_t* x, int bits, uint64_t* keys);
/* hilbert_permutation (src/Topology/Ordering.jl:85-92) of n points (n x 3):
 * perm[c] = old index of new entity c (stable sort of the keys). */
int dmx_hilbert_permutation(int64_t n, const double* points, int bits, int64_t* perm);

/* --- grid files (NetCDF) --------------------------------------------------------- */

/* Read-only NetCDF file: the storage under read_variable / haskey / ds.attrib of the
 * grid-file readers (src/Grids/FileGrids.jl:8-11, 67, 101-102, which use NCDatasets.jl).
 * Classic (CDF-1), 64-bit offset (CDF-2) and 64-bit data (CDF-5) files are parsed
 * natively; NetCDF-4 files go through the HDF5 C library, loaded on first use
 * (DMPLEX_HDF5_LIBRARY or the system's libhdf5.so).  Error -5: the file cannot be
 * opened; -2: not a NetCDF file, malformed, or NetCDF-4 without HDF5; -3: no such
 * variable, a wrong count, or an integer read of a floating-point variable. */
typedef struct NcFile dmx_netcdf;
int dmx_nc_open(const char* path, dmx_netcdf** out);
void dmx_nc_close(dmx_netcdf* f);
/* 1 classic, 2 64-bit offset, 5 64-bit data, 4 NetCDF-4 (HDF5). */
int dmx_nc_format(const dmx_netcdf* f);
int64_t dmx_nc_num_vars(const dmx_netcdf* f);
int dmx_nc_var_name(const dmx_netcdf* f, int64_t i, char* buf, int64_t cap);
/* 1 if the file has the variable, else 0 (haskey(ds, name)). */
int dmx_nc_has_var(const dmx_netcdf* f, const char* name);
/* NetCDF type code (1 byte, 2 char, 3 short, 4 int, 5 float, 6 double, 7 ubyte,
 * 8 ushort, 9 uint, 10 int64, 11 uint64) and shape in ncdump (C) order. */
int dmx_nc_var_info(const dmx_netcdf* f, const char* name, int* type, int* ndims, int64_t* shape,
                    int max_dims);
/* The whole variable, raw (no fill-value or scale conversion), in C order, converted to
 * double / int64; n must equal its number of values. */
int dmx_nc_read_f64(const dmx_netcdf* f, const char* name, double* out, int64_t n);
int dmx_nc_read_i64(const dmx_netcdf* f, const char* name, int64_t* out, int64_t n);
/* Text of a character global attribute; returns 1 when the file has none of that name. */
int dmx_nc_global_text(const dmx_netcdf* f, const char* name, char* buf, int64_t cap, int64_t* len);

/* sort_ring over many rings (src/Grids/FileGrids.jl:22-41): ring r, the slice
 * ids[offsets[r], offsets[r+1]) of indices into points (npoints x 3), is sorted in place
 * counter-clockwise around normalize(axes[r]) (nrings x 3) as seen from outside. */
int dmx_sort_rings(int64_t nrings, const int64_t* offsets, int64_t* ids, int64_t npoints,
                   const double* points, const double* axes);

/* --- synthetic initial state ---------------------------------------------------- */

/* psi(x) = sum_{l=1..lmax} sum_{m=-l..l} coef[l*l - 1 + (m + l)] Y_lm(x) with real
 * orthonormal spherical harmonics, at n points (n x 3) on the unit sphere. */
int dmx_streamfunction(int lmax, const double* coef, int64_t n, const double* points, double* psi);

#ifdef __cplusplus
}
#endif
#endif
