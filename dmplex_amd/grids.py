"""Mesh generators of the TRiSK path: icosahedral, SCVT (Lloyd) and hexagon meshes.

Mirrors ``src/Grids/Icosahedral.jl``, ``src/Grids/Hexagonal.jl`` and
``src/Grids/CellCenters.jl``; the loops run in ``libdmplex_host.so``.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native as N
from .topology import Plex, plex_from_polygons


def _lloyd_device(x: np.ndarray, tri: np.ndarray, device: int, fixed: int = 12, tolerance: float = 1e-8,
                  history: int = 16, max_iterations: int = 10_000, info: dict | None = None) -> int:
    """lloyd! on a HIP device (trisk_lloyd of libtrisk_hip.so), then the host Delaunay check."""
    from . import _device as D

    nv, nt = len(x), len(tri)
    off = np.empty(nv + 1, dtype=np.int64)
    star = np.empty(3 * nt, dtype=np.int64)
    N.check(N.host().dmx_triangle_stars(nv, N.ptr(tri, N.i64p), nt, N.ptr(off, N.i64p), N.ptr(star, N.i64p)))
    its = np.zeros(1, dtype=np.int64)
    res = np.zeros(1)
    rc = D.lib().trisk_lloyd(device, N.ptr(x, N.f64p), nv, N.ptr(tri, N.i64p), nt, N.ptr(off, N.i64p),
                             N.ptr(star, N.i64p), fixed, tolerance, history, max_iterations, N.ptr(its, N.i64p),
                             N.ptr(res, N.f64p))
    if rc < 0:
        D.check(rc)
    if rc > 0:  # the reference's @warn (Icosahedral.jl:183)
        warnings.warn(f"Lloyd's algorithm did not converge in {int(its[0])} iterations (residual "
                      f"{float(res[0]):.3e} mean edge lengths, tolerance {tolerance:.1e})", RuntimeWarning, stacklevel=3)
    if info is not None:
        info.update(iterations=int(its[0]), residual=float(res[0]), converged=rc == 0)
    N.check(N.host().dmx_delaunay(N.ptr(x, N.f64p), nv, N.ptr(tri, N.i64p), nt))
    return int(its[0])


def _icosahedral(subdivisions: int, centroidal: bool, device: int | None = None, log=None):
    if subdivisions < 0:
        raise ValueError("subdivisions must be non-negative")
    if device is None:
        nv, nt = 10 * 4 ** subdivisions + 2, 20 * 4 ** subdivisions
        x = np.empty((nv, 3))
        tri = np.empty((nt, 3), dtype=np.int64)
        its = np.zeros(max(subdivisions, 1), dtype=np.int64)
        N.check(N.host().dmx_icosahedral(subdivisions, int(centroidal), N.ptr(x, N.f64p), N.ptr(tri, N.i64p),
                                         N.ptr(its, N.i64p)))
        its = its[:subdivisions]
        if centroidal:
            for level, n in enumerate(its, start=1):
                if n >= 10_000:  # the reference's @warn (Icosahedral.jl:183)
                    warnings.warn(f"Lloyd's algorithm did not converge in {n} iterations at level {level}",
                                  RuntimeWarning, stacklevel=3)
        return x, tri, its
    x = np.empty((12, 3))
    tri = np.empty((20, 3), dtype=np.int64)
    N.check(N.host().dmx_icosahedron(N.ptr(x, N.f64p), N.ptr(tri, N.i64p)))
    its = []
    for level in range(1, subdivisions + 1):
        nv, nt = len(x), len(tri)
        x2 = np.empty((nv + 3 * nt // 2, 3))
        tri2 = np.empty((4 * nt, 3), dtype=np.int64)
        N.check(N.host().dmx_bisect(nv, N.ptr(x, N.f64p), nt, N.ptr(tri, N.i64p), N.ptr(x2, N.f64p),
                                    N.ptr(tri2, N.i64p)))
        x, tri = x2, tri2
        if centroidal:
            info = {}
            its.append(_lloyd_device(x, tri, device, info=info))
            if log:
                state = "converged" if info["converged"] else "stopped (not converged)"
                log(f"level {level}: {len(x)} vertices, Lloyd {state} after {its[-1]} iterations, "
                    f"residual {info['residual']:.3e} mean edge lengths")
    return x, tri, np.array(its, dtype=np.int64)


def create_spherical_mesh(subdivisions: int, spring_beta=None, centroidal: bool = False,
                          device: int | None = None, log=None) -> Plex:
    """Icosahedral triangle mesh on the unit sphere (Icosahedral.jl:22-61).

    ``centroidal=True`` runs Lloyd's algorithm after every bisection (SCVT), on the host
    or, with ``device=<HIP ordinal>``, on the GPU (the way to build levels 8-10 in
    seconds).  The spring dynamics option of the reference (``spring_beta``) is not part
    of this path.
    """
    if spring_beta is not None:
        raise NotImplementedError("spring dynamics is outside the TRiSK SCVT path")
    x, tri, _ = _icosahedral(subdivisions, centroidal, device, log)
    return plex_from_polygons((np.arange(0, 3 * len(tri) + 1, 3, dtype=np.int64), tri.reshape(-1)), x)


def create_hexagon_mesh(subdivisions: int, spring_beta=None, centroidal: bool = False,
                        primal: Plex | None = None, device: int | None = None, log=None) -> Plex:
    """Voronoi dual of the icosahedral mesh (Hexagonal.jl:11-17).

    Triangle circumcenters become vertices and every triangle vertex a polygon; the
    result is an exact spherical Voronoi diagram whose generators are the icosahedral
    vertices (kept as ``dm.generators``).
    """
    if primal is None:
        primal = create_spherical_mesh(subdivisions, spring_beta, centroidal, device, log)
    pv = primal.view()
    off = np.empty(primal.nv + 1, dtype=np.int64)
    verts = np.empty(3 * primal.nc, dtype=np.int64)
    centers = np.empty((primal.nc, 3))
    N.check(N.host().dmx_dual_polygons(C.byref(pv), N.ptr(off, N.i64p), N.ptr(verts, N.i64p),
                                       N.ptr(centers, N.f64p)))
    return plex_from_polygons((off, verts), centers, generators=primal.coordinates.copy())


def lloyd_(x: np.ndarray, triangles: np.ndarray, fixed: int = 0, tolerance: float = 1e-8,
           history: int = 16, max_iterations: int = 10_000, device: int | None = None) -> int:
    """Lloyd's algorithm with Anderson mixing, in place (Icosahedral.jl:126-184).

    ``fixed`` is the number of leading vertices that do not move (the reference's
    ``fixed=1:12`` is ``fixed=12``).  Returns the number of iterations; raises
    ``ValueError`` if the triangulation does not stay Delaunay.
    """
    if x.dtype != np.float64 or not x.flags.c_contiguous:
        raise ValueError("x must be a C-contiguous float64 (n, 3) array")
    tri = np.ascontiguousarray(triangles, dtype=np.int64)
    if device is not None:
        return _lloyd_device(x, tri, device, fixed, tolerance, history, max_iterations)
    its = np.zeros(1, dtype=np.int64)
    rc = N.check(N.host().dmx_lloyd(N.ptr(x, N.f64p), len(x), N.ptr(tri, N.i64p), len(tri), fixed, tolerance,
                                    history, max_iterations, N.ptr(its, N.i64p)), warn_ok=True)
    if rc > 0:  # the reference's @warn (Icosahedral.jl:183)
        warnings.warn(f"Lloyd's algorithm did not converge in {int(its[0])} iterations", RuntimeWarning,
                      stacklevel=2)
    return int(its[0])


def lloyd_residual(x: np.ndarray, triangles: np.ndarray, fixed: int = 0) -> float:
    """Largest distance of a free vertex from its Voronoi centroid (lloyd_step!)."""
    tri = np.ascontiguousarray(triangles, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = C.c_double()
    N.check(N.host().dmx_lloyd_residual(N.ptr(x, N.f64p), len(x), N.ptr(tri, N.i64p), len(tri), fixed,
                                        C.byref(r)))
    return r.value


def orthogonality_defect(dm: Plex, centers) -> float:
    """Largest |cos| between a primal edge and the dual great circle (CellCenters.jl:90-101)."""
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    out = C.c_double()
    N.check(N.host().dmx_orthogonality_defect(C.byref(dm.view()), N.ptr(c, N.f64p), C.byref(out)))
    return out.value


def _normalize(a: np.ndarray) -> np.ndarray:
    return a / np.linalg.norm(a, axis=-1, keepdims=True)


def _edge_cells(dm: Plex) -> tuple[np.ndarray, np.ndarray]:
    """(left, right) cell of every edge, -1 when missing (get_edge_cells, Plex.jl:210-218)."""
    nconn = dm.cone_offsets[dm.nc]
    cone, orient = dm.cones[:nconn], dm.cone_orientations[:nconn]
    owner = np.repeat(np.arange(dm.nc), np.diff(dm.cone_offsets[:dm.nc + 1]))
    left = np.full(dm.ne, -1, dtype=np.int64)
    right = np.full(dm.ne, -1, dtype=np.int64)
    left[cone[orient == 0] - dm.nc] = owner[orient == 0]
    right[cone[orient != 0] - dm.nc] = owner[orient != 0]
    return left, right


def voronoi_generators(dm: Plex, iterations: int = 4) -> np.ndarray:
    """Generators of the best-fitting Voronoi diagram by inverse iteration on the
    homogeneous bisector conditions (CellCenters.jl:49-81), with a sparse LU solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    left, right = _edge_cells(dm)
    inner = np.flatnonzero((left >= 0) & (right >= 0))
    i, j = left[inner], right[inner]
    first = dm.cones[dm.cone_offsets[dm.nc + inner]]
    second = dm.cones[dm.cone_offsets[dm.nc + inner] + 1]
    a = dm.coordinates[first - dm.nc - dm.ne]
    b = dm.coordinates[second - dm.nc - dm.ne]
    n = _normalize(np.cross(a, b))
    pick = np.where((np.abs(n[:, 0]) < 0.9)[:, None], np.array([1.0, 0, 0]), np.array([0, 1.0, 0]))
    t1 = _normalize(np.cross(n, pick))
    t2 = np.cross(n, t1)
    # three rows per edge, (t1, -1), (t2, -1), (n, +1); six entries per row, w_k on g_i
    # and s w_k on g_j for k = x, y, z
    m = len(inner)
    w = np.stack([t1, t2, n], axis=1)                                   # m x 3 rows x 3
    sgn = np.array([-1.0, -1.0, 1.0])[None, :, None]
    row = np.broadcast_to(np.arange(3 * m).reshape(m, 3, 1), (m, 3, 3))
    k = np.arange(3)[None, None, :]
    rows = np.concatenate([row.ravel(), row.ravel()])
    cols = np.concatenate([np.broadcast_to(3 * i[:, None, None] + k, (m, 3, 3)).ravel(),
                           np.broadcast_to(3 * j[:, None, None] + k, (m, 3, 3)).ravel()])
    vals = np.concatenate([w.ravel(), (sgn * w).ravel()])
    nrow = 3 * m
    A = sp.csr_matrix((vals, (rows, cols)), shape=(nrow, 3 * dm.nc))
    B = (A.T @ A).tocsc()
    B = B + 1e-10 * abs(B.diagonal()).max() * sp.identity(B.shape[0], format="csc")
    lu = spl.splu(B)
    centroids = cell_centers(dm, method="centroid")
    x = centroids.reshape(-1).copy()
    for _ in range(iterations):
        x = lu.solve(x)
        x /= np.linalg.norm(x)
    gens = _normalize(x.reshape(-1, 3))
    if np.sum(np.einsum("ij,ij->i", gens, centroids)) < 0:
        gens = -gens
    return gens


def cell_centers(dm: Plex, method: str = "auto", tolerance: float = 1e-8) -> np.ndarray:
    """Cell centers that coincide with the orthogonal dual's vertices (CellCenters.jl:25-47)."""
    if method == "auto":
        sizes = np.diff(dm.cone_offsets[:dm.nc + 1])
        if np.all(sizes == 3):
            return cell_centers(dm, method="circumcenter")
        gens = dm.generators if dm.generators is not None else cell_centers(dm, method="voronoi")
        if orthogonality_defect(dm, gens) < tolerance:
            return np.array(gens, dtype=np.float64)
        return cell_centers(dm, method="centroid")
    if method == "centroid":
        sizes = np.diff(dm.cone_offsets[:dm.nc + 1])
        nconn = dm.cone_offsets[dm.nc]
        cone, orient = dm.cones[:nconn], dm.cone_orientations[:nconn]
        start = np.where(orient == 0, dm.cones[dm.cone_offsets[cone]], dm.cones[dm.cone_offsets[cone] + 1])
        out = np.zeros((dm.nc, 3))
        np.add.at(out, np.repeat(np.arange(dm.nc), sizes), dm.coordinates[start - dm.nc - dm.ne])
        return _normalize(out)
    if method == "circumcenter":
        tri = dm.cell_vertex_table()
        if tri.shape[1] != 3:
            raise ValueError("circumcenters need a triangle mesh")
        a, b, d = (dm.coordinates[tri[:, k]] for k in range(3))
        return _normalize(np.cross(b - a, d - a))
    if method == "voronoi":
        return voronoi_generators(dm)
    raise ValueError(f"unknown center method {method}")
