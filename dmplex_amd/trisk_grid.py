"""``TRiSKGrid``: geometry and connectivity of the TRiSK C grid on a spherical Voronoi mesh.

Mirrors ``src/FiniteVolumes/TRiSKGrid.jl`` (struct :37-60, constructor :76-162,
``vertex_edge_areas`` :179-191, ``gassmann_coefficients`` :214-251).  Tables are numpy
arrays of shape ``(k, n)`` in Fortran order, i.e. the memory of the reference's Julia
``k x n`` matrices, indexed ``table[k, j]`` with 0-based indices; padded slots hold index
0 and weight 0.  Point lists (``centers``, ``vertices``, ``edge_points``,
``edge_normals``) are ``(n, 3)`` arrays.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native as N
from .grids import cell_centers, orthogonality_defect
from .topology import Plex


def _table(k: int, n: int, dtype) -> np.ndarray:
    return np.zeros((k, n), dtype=dtype, order="F")


def arc_length(a, b):
    """Great-circle distance on the unit sphere (TRiSKGrid.jl:2)."""
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.sum(a * b, axis=-1))


def signed_triangle_area(a, b, c):
    """Signed spherical triangle area, negative when clockwise (TRiSKGrid.jl:5)."""
    return 2 * np.arctan2(np.sum(a * np.cross(b, c), axis=-1),
                          1 + np.sum(a * b, axis=-1) + np.sum(b * c, axis=-1) + np.sum(c * a, axis=-1))


class TRiSKGrid:
    """Voronoi-Delaunay geometry of the TRiSK C grid (TRiSKGrid.jl:7-60)."""

    def __init__(self, dm: Plex, centers=None, tolerance: float = 1e-8):
        if dm.surface != "sphere":
            raise ValueError("TRiSKGrid needs a spherical mesh")
        g = cell_centers(dm) if centers is None else np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        if g.shape[0] != dm.nc:
            raise ValueError("one center per cell is needed")
        defect = orthogonality_defect(dm, g)
        if not defect < tolerance:
            warnings.warn(f"the mesh has no orthogonal dual (defect {defect}): TRiSK operators lose consistency")
        self.dm = dm
        self.centers = np.ascontiguousarray(g)
        self.vertices = dm.coordinates
        nc, ne, nv = dm.nc, dm.ne, dm.nv
        widths = np.zeros(3, dtype=np.int64)
        view = dm.view()
        N.check(N.host().dmx_trisk_grid_widths(C.byref(view), N.ptr(widths, N.i64p)))
        kc, kv, kt = (int(w) for w in widths)
        self.edge_points = np.empty((ne, 3))
        self.edge_normals = np.empty((ne, 3))
        self.primal_lengths = np.empty(ne)
        self.dual_lengths = np.empty(ne)
        self.edge_cells = _table(2, ne, np.int64)
        self.edge_vertices = _table(2, ne, np.int64)
        self.cell_distances = _table(2, ne, np.float64)
        self.cell_areas = np.empty(nc)
        self.vertex_areas = np.empty(nv)
        self.cell_edges = _table(kc, nc, np.int64)
        self.cell_signs = _table(kc, nc, np.float64)
        self.cell_vertices = _table(kc, nc, np.int64)
        self.cell_kites = _table(kc, nc, np.float64)
        self.vertex_edges = _table(kv, nv, np.int64)
        self.vertex_signs = _table(kv, nv, np.float64)
        self.vertex_cells = _table(kv, nv, np.int64)
        self.vertex_kites = _table(kv, nv, np.float64)
        self.trisk_edges = _table(kt, ne, np.int64)
        self.trisk_weights = _table(kt, ne, np.float64)
        N.check(N.host().dmx_trisk_grid(C.byref(view), N.ptr(self.centers, N.f64p), C.byref(self._tables())))

    def _tables(self) -> N.DmxGridTables:
        f, i = N.f64p, N.i64p
        return N.DmxGridTables(
            self.cell_edges.shape[0], self.vertex_edges.shape[0], self.trisk_edges.shape[0],
            N.ptr(self.edge_points, f), N.ptr(self.edge_normals, f), N.ptr(self.primal_lengths, f),
            N.ptr(self.dual_lengths, f), N.ptr(self.edge_cells, i), N.ptr(self.edge_vertices, i),
            N.ptr(self.cell_distances, f), N.ptr(self.cell_areas, f), N.ptr(self.vertex_areas, f),
            N.ptr(self.cell_edges, i), N.ptr(self.cell_signs, f), N.ptr(self.cell_vertices, i),
            N.ptr(self.cell_kites, f), N.ptr(self.vertex_edges, i), N.ptr(self.vertex_signs, f),
            N.ptr(self.vertex_cells, i), N.ptr(self.vertex_kites, f), N.ptr(self.trisk_edges, i),
            N.ptr(self.trisk_weights, f))

    def num_cells(self) -> int:
        return len(self.centers)

    def num_edges(self) -> int:
        return len(self.edge_points)

    def num_vertices(self) -> int:
        return len(self.vertices)


def vertex_edge_areas(g: TRiSKGrid):
    """Signed areas A_{e,v} of (left generator, right generator, vertex) per vertex
    (``by_vertex``, kv x nv) and per edge (``by_edge``, 2 x ne) (TRiSKGrid.jl:179-191)."""
    nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
    by_vertex = _table(g.vertex_edges.shape[0], nv, np.float64)
    by_edge = _table(2, ne, np.float64)
    N.check(N.host().dmx_vertex_edge_areas(nc, ne, nv, C.byref(g._tables()), N.ptr(g.centers, N.f64p),
                                           N.ptr(np.ascontiguousarray(g.vertices), N.f64p),
                                           N.ptr(by_vertex, N.f64p), N.ptr(by_edge, N.f64p)))
    return by_vertex, by_edge


def gassmann_coefficients(g: TRiSKGrid):
    """(edges, pv_edges, weights), 12 x ne, of Gassmann's PV flux reconstruction
    (TRiSKGrid.jl:214-251); raises ``ValueError`` unless every vertex has three edges."""
    nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
    E = _table(12, ne, np.int64)
    P = _table(12, ne, np.int64)
    W = _table(12, ne, np.float64)
    N.check(N.host().dmx_gassmann_coefficients(nc, ne, nv, C.byref(g._tables()),
                                               N.ptr(np.ascontiguousarray(g.vertices), N.f64p),
                                               N.ptr(E, N.i64p), N.ptr(P, N.i64p), N.ptr(W, N.f64p)))
    return E, P, W
