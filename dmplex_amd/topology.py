"""The ``Plex`` mesh topology (a DMPlex-like DAG of a surface mesh).

Mirrors ``src/Topology/Plex.jl`` and ``src/Topology/Ordering.jl`` with 0-based
numbering: points are cells ``[0, nc)``, edges ``[nc, nc+ne)``, vertices
``[nc+ne, nc+ne+nv)`` (Plex.jl:5-10).  Every cone entry carries an orientation,
``0`` if the cell traverses the edge from its first to its second vertex and ``-1``
otherwise; edges run from their lower to their higher vertex, cells are
counter-clockwise seen from outside, so the cell with orientation ``0`` lies left
of the edge (Plex.jl:12-16).  The heavy lifting is native (``libdmplex_host.so``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class Plex:
    """Interpolated surface mesh in CSR form (Plex.jl:23-34), on the unit sphere."""

    def __init__(self, nc, ne, nv, cone_offsets, cones, orientations, support_offsets, supports,
                 coordinates, generators=None):
        self.nc, self.ne, self.nv = int(nc), int(ne), int(nv)
        self.cone_offsets = np.ascontiguousarray(cone_offsets, dtype=np.int64)
        self.cones = np.ascontiguousarray(cones, dtype=np.int64)
        self.cone_orientations = np.ascontiguousarray(orientations, dtype=np.int8)
        self.support_offsets = np.ascontiguousarray(support_offsets, dtype=np.int64)
        self.supports = np.ascontiguousarray(supports, dtype=np.int64)
        self.coordinates = np.ascontiguousarray(coordinates, dtype=np.float64).reshape(-1, 3)
        self.surface = "sphere"
        # Voronoi generators of a hexagon mesh built here (the primal vertices); the
        # reference recovers them by inverse iteration (CellCenters.jl:49-81) to 1e-12.
        self.generators = generators
        self._view = None

    # --- strata ------------------------------------------------------------------
    def num_cells(self) -> int:
        return self.nc

    def num_edges(self) -> int:
        return self.ne

    def num_vertices(self) -> int:
        return self.nv

    def cell_range(self) -> range:
        return range(0, self.nc)

    def edge_range(self) -> range:
        return range(self.nc, self.nc + self.ne)

    def vertex_range(self) -> range:
        return range(self.nc + self.ne, self.nc + self.ne + self.nv)

    def cell_index(self, c: int) -> int:
        return c

    def edge_index(self, e: int) -> int:
        return e - self.nc

    def vertex_index(self, v: int) -> int:
        return v - self.nc - self.ne

    def get_coordinates(self, v: int) -> np.ndarray:
        return self.coordinates[self.vertex_index(v)]

    # --- cones and supports ----------------------------------------------------------
    def get_cone(self, p: int) -> np.ndarray:
        return self.cones[self.cone_offsets[p]:self.cone_offsets[p + 1]]

    def get_cone_orientation(self, p: int) -> np.ndarray:
        return self.cone_orientations[self.cone_offsets[p]:self.cone_offsets[p + 1]]

    def get_cone_size(self, p: int) -> int:
        return int(self.cone_offsets[p + 1] - self.cone_offsets[p])

    def get_support(self, p: int) -> np.ndarray:
        return self.supports[self.support_offsets[p]:self.support_offsets[p + 1]]

    def edge_sign(self, cell: int, i: int) -> int:
        """``+1`` if cone entry ``i`` (0-based) of ``cell`` runs along the edge (Plex.jl:201)."""
        return 1 if self.cone_orientations[self.cone_offsets[cell] + i] == 0 else -1

    def get_edge_cells(self, e: int) -> tuple[int, int]:
        """(left, right) cell points of edge ``e``, ``-1`` when missing (Plex.jl:210-218)."""
        left = right = -1
        for c in self.get_support(e):
            k = self.cone_offsets[c] + int(np.flatnonzero(self.get_cone(c) == e)[0])
            if self.cone_orientations[k] == 0:
                left = int(c)
            else:
                right = int(c)
        return left, right

    def get_cell_vertices(self, cell: int) -> np.ndarray:
        """Vertex points of ``cell`` counter-clockwise; vertex i starts cone edge i (Plex.jl:226-230)."""
        cone, orient = self.get_cone(cell), self.get_cone_orientation(cell)
        return np.array([self.cones[self.cone_offsets[e] + (0 if o == 0 else 1)] for e, o in zip(cone, orient)],
                        dtype=np.int64)

    def cell_coordinates(self, cell: int) -> np.ndarray:
        return self.coordinates[self.get_cell_vertices(cell) - self.nc - self.ne]

    # --- native view -------------------------------------------------------------
    def view(self) -> N.DmxPlex:
        if self._view is None:
            self._view = N.DmxPlex(self.nc, self.ne, self.nv, N.ptr(self.cone_offsets, N.i64p),
                                   N.ptr(self.cones, N.i64p), N.ptr(self.cone_orientations, N.i8p),
                                   N.ptr(self.support_offsets, N.i64p), N.ptr(self.supports, N.i64p),
                                   N.ptr(self.coordinates, N.f64p))
        return self._view

    def cell_vertex_table(self) -> np.ndarray:
        """(nc, k) vertex indices of every cell of constant size k (e.g. triangles)."""
        sizes = np.diff(self.cone_offsets[:self.nc + 1])
        k = int(sizes[0])
        if not np.all(sizes == k):
            raise ValueError("cells of different sizes")
        cone = self.cones[:self.nc * k].reshape(self.nc, k)
        orient = self.cone_orientations[:self.nc * k].reshape(self.nc, k)
        first = self.cones[self.cone_offsets[cone]]
        second = self.cones[self.cone_offsets[cone] + 1]
        return np.where(orient == 0, first, second) - self.nc - self.ne


def plex_from_polygons(polygons, coordinates, generators=None) -> Plex:
    """Interpolated Plex from a cell-vertex list (Plex.jl:49-97), on the unit sphere.

    ``polygons`` is a list of vertex-index sequences or a tuple ``(offsets, vertices)``
    in CSR form.  Clockwise cells are reversed; edges are numbered by first appearance.
    """
    coords = np.ascontiguousarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if isinstance(polygons, tuple):
        offsets, verts = (np.ascontiguousarray(a, dtype=np.int64) for a in polygons)
    else:
        polys = [np.asarray(p, dtype=np.int64) for p in polygons]
        offsets = np.zeros(len(polys) + 1, dtype=np.int64)
        offsets[1:] = np.cumsum([len(p) for p in polys])
        verts = np.concatenate(polys) if polys else np.zeros(0, dtype=np.int64)
    if np.all(coords[:, 2] == 0):
        raise NotImplementedError("planar meshes are outside the TRiSK shallow-water path")
    lib = N.host()
    handle = C.c_void_p()
    dims = np.zeros(4, dtype=np.int64)
    N.check(lib.dmx_plex_from_polygons(len(offsets) - 1, N.ptr(offsets, N.i64p), N.ptr(verts, N.i64p),
                                       len(coords), N.ptr(coords, N.f64p), C.byref(handle),
                                       N.ptr(dims, N.i64p)))
    nc, ne, nv, ncone = (int(d) for d in dims)
    npts = nc + ne + nv
    co, cones = np.empty(npts + 1, np.int64), np.empty(ncone, np.int64)
    orient, so, sup = np.empty(ncone, np.int8), np.empty(npts + 1, np.int64), np.empty(ncone, np.int64)
    try:
        N.check(lib.dmx_plex_export(handle, N.ptr(co, N.i64p), N.ptr(cones, N.i64p), N.ptr(orient, N.i8p),
                                    N.ptr(so, N.i64p), N.ptr(sup, N.i64p)))
    finally:
        lib.dmx_plex_free(handle)
    return Plex(nc, ne, nv, co, cones, orient, so, sup, coords, generators)


def check_plex(dm: Plex) -> bool:
    """Mesh invariants of ``check_plex`` (Plex.jl:329-360); raises ``RuntimeError``."""
    lib = N.host()
    if lib.dmx_check_plex(C.byref(dm.view())) != 0:
        raise RuntimeError(lib.dmx_last_error().decode())
    return True


def hilbert_key(x, bits: int) -> int:
    """Index of the integer point ``x`` along the 3-D Hilbert curve (Ordering.jl:10-41)."""
    return int(hilbert_keys(np.asarray([x], dtype=np.uint64), bits)[0])


def hilbert_keys(x: np.ndarray, bits: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 3)
    keys = np.empty(len(x), dtype=np.uint64)
    N.check(N.host().dmx_hilbert_keys(len(x), N.ptr(x, N.u64p), bits, N.ptr(keys, N.u64p)))
    return keys


def cell_centroids(dm: Plex) -> np.ndarray:
    """Mean of the vertex coordinates of every cell (as in hilbert_permutation)."""
    sizes = np.diff(dm.cone_offsets[:dm.nc + 1])
    cone = dm.cones[:dm.cone_offsets[dm.nc]]
    orient = dm.cone_orientations[:dm.cone_offsets[dm.nc]]
    start = np.where(orient == 0, dm.cones[dm.cone_offsets[cone]], dm.cones[dm.cone_offsets[cone] + 1])
    pts = dm.coordinates[start - dm.nc - dm.ne]
    owner = np.repeat(np.arange(dm.nc), sizes)
    out = np.zeros((dm.nc, 3))
    np.add.at(out, owner, pts)
    return out / sizes[:, None]


def hilbert_permutation(dm: Plex, bits: int = 16) -> np.ndarray:
    """Cells in the order of the Hilbert curve of their centroids (Ordering.jl:85-92)."""
    pts = np.ascontiguousarray(cell_centroids(dm))
    perm = np.empty(dm.nc, dtype=np.int64)
    N.check(N.host().dmx_hilbert_permutation(dm.nc, N.ptr(pts, N.f64p), bits, N.ptr(perm, N.i64p)))
    return perm


def hilbert_order(dm: Plex, bits: int = 16) -> Plex:
    """The same mesh with cells along the Hilbert curve, vertices and edges by first
    appearance (Ordering.jl:54-83; labels are not carried)."""
    perm = hilbert_permutation(dm, bits)
    polys = [dm.get_cell_vertices(int(c)) - dm.nc - dm.ne for c in perm]
    vnew = np.full(dm.nv, -1, dtype=np.int64)
    n = 0
    for poly in polys:
        for v in poly:
            if vnew[v] < 0:
                vnew[v] = n
                n += 1
    if n != dm.nv:
        raise ValueError("vertices without cells")
    coords = np.empty_like(dm.coordinates)
    coords[vnew] = dm.coordinates
    gens = None if dm.generators is None else dm.generators[perm]
    return plex_from_polygons([vnew[p] for p in polys], coords, gens)
