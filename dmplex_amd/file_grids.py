"""Readers of ICON and MPAS grid files (``src/Grids/FileGrids.jl``).

Cells and vertices keep the numbering of the file; edges are numbered by
``plex_from_polygons``.  The file access (NetCDF classic, 64-bit offset, 64-bit data
and NetCDF-4/HDF5) is native, in ``libdmplex_host.so``; so is the ring sort of the
dual meshes.  The reference raises ``ArgumentError`` and ``DimensionMismatch`` where
these functions raise ``ValueError``.

Beyond the reference, a mesh read here keeps the Voronoi generators the file holds
as ``dm.generators``: the MPAS cell generators (``x/y/zCell``) of a primal MPAS mesh
and the ICON vertices of an ICON dual.  ``cell_centers`` takes them when they fit the
mesh (orthogonality defect below its tolerance), which is the point the reference's
inverse iteration (``CellCenters.jl:49-81``) converges to, without its global solve.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .topology import Plex, plex_from_polygons

NC_TYPES = {1: "byte", 2: "char", 3: "short", 4: "int", 5: "float", 6: "double", 7: "ubyte", 8: "ushort",
            9: "uint", 10: "int64", 11: "uint64"}
_FLOAT_TYPES = (5, 6)


class NetCDFFile:
    """A NetCDF file open for reading (the part of ``NCDatasets.NCDataset`` the readers use).

    ``name in ds`` is ``haskey(ds, name)``, ``ds.variable(name)`` the raw values in
    ``ncdump`` (C) order, float64 for floating-point and int64 for integer variables,
    ``ds.attribute(name)`` the text of a character global attribute or ``None``.
    """

    def __init__(self, path):
        self.path = os.fspath(path)
        self._h = C.c_void_p()
        N.check(N.host().dmx_nc_open(self.path.encode(), C.byref(self._h)))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            N.host().dmx_nc_close(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _handle(self) -> C.c_void_p:
        if self._h is None:
            raise ValueError(f"{self.path} is closed")
        return self._h

    @property
    def format(self) -> str:
        return {1: "classic", 2: "64bit_offset", 5: "64bit_data", 4: "netcdf4"}[
            N.host().dmx_nc_format(self._handle())]

    def keys(self) -> list[str]:
        lib, h = N.host(), self._handle()
        names = []
        buf = C.create_string_buffer(4096)
        for i in range(lib.dmx_nc_num_vars(h)):
            N.check(lib.dmx_nc_var_name(h, i, buf, len(buf)))
            names.append(buf.value.decode())
        return names

    def __contains__(self, name: str) -> bool:
        return bool(N.host().dmx_nc_has_var(self._handle(), name.encode()))

    def info(self, name: str) -> tuple[str, tuple[int, ...]]:
        """(NetCDF type name, shape in C order) of a variable."""
        t, nd = C.c_int(), C.c_int()
        shape = np.zeros(16, dtype=np.int64)
        N.check(N.host().dmx_nc_var_info(self._handle(), name.encode(), C.byref(t), C.byref(nd),
                                         N.ptr(shape, N.i64p), len(shape)))
        return NC_TYPES[t.value], tuple(int(s) for s in shape[:nd.value])

    def variable(self, name: str) -> np.ndarray:
        lib, h = N.host(), self._handle()
        t, shape = self.info(name)
        if t == "char":
            raise ValueError(f"variable {name} is a character variable")
        floating = t in ("float", "double")
        out = np.empty(shape, dtype=np.float64 if floating else np.int64)
        if floating:
            N.check(lib.dmx_nc_read_f64(h, name.encode(), N.ptr(out, N.f64p), out.size))
        else:
            N.check(lib.dmx_nc_read_i64(h, name.encode(), N.ptr(out, N.i64p), out.size))
        return out

    def attribute(self, name: str) -> str | None:
        buf = C.create_string_buffer(1 << 16)
        n = np.zeros(1, dtype=np.int64)
        rc = N.host().dmx_nc_global_text(self._handle(), name.encode(), buf, len(buf), N.ptr(n, N.i64p))
        if rc == 1:
            return None
        N.check(rc)
        return buf.raw[:int(n[0])].decode(errors="replace")


def read_variable(ds: NetCDFFile, name: str) -> np.ndarray:
    """A whole variable, raw, without fill-value or scaling conversion (FileGrids.jl:4-11).

    The array is in ``ncdump`` (C) order, i.e. with the dimensions reversed with respect
    to the reference's Julia array.
    """
    return ds.variable(name)


def lonlat_to_xyz(lon, lat) -> np.ndarray:
    """Points on the unit sphere from longitude and latitude in radians (FileGrids.jl:13)."""
    lon, lat = np.asarray(lon, dtype=np.float64), np.asarray(lat, dtype=np.float64)
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=-1)


def rows_per_entity(table: np.ndarray, n: int) -> np.ndarray:
    """A connectivity table as ``n x k``, one row per entity (FileGrids.jl:15-20).

    ``table`` is in C order, so the reference's first test (Julia ``size(table, 1)``)
    is the last axis here.
    """
    table = np.asarray(table)
    if table.ndim != 2:
        raise ValueError(f"connectivity table of shape {table.shape} is not a matrix")
    if table.shape[1] == n:
        return np.ascontiguousarray(table.T)
    if table.shape[0] == n:
        return table
    raise ValueError(f"connectivity table of size {table.shape} does not match {n} entities")


def _sort_rings(offsets: np.ndarray, ids: np.ndarray, coordinates: np.ndarray, axes: np.ndarray) -> np.ndarray:
    ids = np.ascontiguousarray(ids, dtype=np.int64).copy()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    coords = np.ascontiguousarray(coordinates, dtype=np.float64).reshape(-1, 3)
    axes = np.ascontiguousarray(axes, dtype=np.float64).reshape(-1, 3)
    if len(axes) != len(offsets) - 1:
        raise ValueError("one axis per ring is needed")
    N.check(N.host().dmx_sort_rings(len(axes), N.ptr(offsets, N.i64p), N.ptr(ids, N.i64p), len(coords),
                                    N.ptr(coords, N.f64p), N.ptr(axes, N.f64p)))
    return ids


def sort_ring(ids, coordinates, axis) -> np.ndarray:
    """The points ``coordinates[ids]`` sorted counter-clockwise around ``axis``, seen
    from outside the sphere (FileGrids.jl:22-33)."""
    ids = np.asarray(ids, dtype=np.int64)
    return _sort_rings(np.array([0, len(ids)]), ids, coordinates, np.asarray(axis, dtype=np.float64)[None])


def dual_polygons(rings, coordinates, centers):
    """Polygons of a dual mesh (FileGrids.jl:35-41): ring ``v`` lists the primal cells
    around primal vertex ``v``, the corners of dual cell ``v``, sorted around
    ``normalize(centers[v])``.  ``rings`` is a list of index sequences or a CSR tuple
    ``(offsets, ids)``; the result is a CSR tuple for ``plex_from_polygons``.
    """
    if isinstance(rings, tuple):
        offsets, ids = rings
    else:
        offsets = np.zeros(len(rings) + 1, dtype=np.int64)
        offsets[1:] = np.cumsum([len(r) for r in rings])
        ids = np.concatenate([np.asarray(r, dtype=np.int64) for r in rings]) if len(rings) else np.zeros(0, np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)
    return offsets, _sort_rings(offsets, ids, coordinates, centers)


def _csr_rows(table: np.ndarray, keep: np.ndarray):
    """CSR polygons of the kept entries of every row of ``table`` (row order preserved)."""
    counts = keep.sum(axis=1)
    offsets = np.zeros(len(table) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    return offsets, np.ascontiguousarray(table[keep], dtype=np.int64)


def _normalized_xyz(ds: NetCDFFile, suffix: str) -> np.ndarray:
    x, y, z = (read_variable(ds, axis + suffix) for axis in "xyz")
    r = np.sqrt(x * x + y * y + z * z)
    return np.stack([x / r, y / r, z / r], axis=-1)


def read_icon_grid(filename, dual: bool = False) -> Plex:
    """ICON triangular grid from a grid file (``vlon``, ``vlat``, ``vertex_of_cell``),
    or with ``dual=True`` its hexagonal/pentagonal dual: vertices at the triangle
    centers (``clon``, ``clat``), cells from ``cells_of_vertex``; global grids only
    (FileGrids.jl:43-68)."""
    with NetCDFFile(filename) as ds:
        vertices = lonlat_to_xyz(read_variable(ds, "vlon"), read_variable(ds, "vlat"))
        centers = lonlat_to_xyz(read_variable(ds, "clon"), read_variable(ds, "clat"))
        if not dual:
            table = rows_per_entity(read_variable(ds, "vertex_of_cell"), len(centers))
            return plex_from_polygons(_csr_rows(table - 1, np.ones(table.shape, dtype=bool)), vertices)
        table = rows_per_entity(read_variable(ds, "cells_of_vertex"), len(vertices))
    keep = (table >= 1) & (table <= len(centers))
    if not np.all(keep.sum(axis=1) >= 5):
        raise ValueError("the dual of a regional ICON grid is not supported")
    polys = dual_polygons(_csr_rows(table - 1, keep), centers, vertices)
    return plex_from_polygons(polys, centers, generators=vertices)


def read_mpas_grid(filename, dual: bool = False) -> Plex:
    """MPAS Voronoi grid from a grid file (``x/y/zVertex``, ``verticesOnCell``,
    ``nEdgesOnCell``), projected onto the unit sphere; with ``dual=True`` the Delaunay
    triangulation: vertices at the generators (``x/y/zCell``), triangles from
    ``cellsOnVertex``, without those touching a missing cell (FileGrids.jl:70-96)."""
    with NetCDFFile(filename) as ds:
        flag = ds.attribute("on_a_sphere")
        if flag is not None and flag.strip() != "YES":
            raise ValueError(f"{os.fspath(filename)} is a planar MPAS mesh")
        vertices, centers = _normalized_xyz(ds, "Vertex"), _normalized_xyz(ds, "Cell")
        if not dual:
            table = rows_per_entity(read_variable(ds, "verticesOnCell"), len(centers))
            nedges = read_variable(ds, "nEdgesOnCell").reshape(-1)
            if len(nedges) != len(centers) or np.any(nedges < 0) or np.any(nedges > table.shape[1]):
                raise ValueError("nEdgesOnCell does not match verticesOnCell")
            keep = np.arange(table.shape[1])[None, :] < nedges[:, None]
            return plex_from_polygons(_csr_rows(table - 1, keep), vertices, generators=centers)
        table = rows_per_entity(read_variable(ds, "cellsOnVertex"), len(vertices))
    whole = np.all((table >= 1) & (table <= len(centers)), axis=1)
    tri = np.ascontiguousarray(table[whole] - 1, dtype=np.int64)
    return plex_from_polygons((np.arange(0, tri.size + 1, tri.shape[1], dtype=np.int64), tri.reshape(-1)), centers)


def read_grid(filename, dual: bool = False) -> Plex:
    """An ICON or MPAS grid file, told apart by its variables (FileGrids.jl:98-107)."""
    with NetCDFFile(filename) as ds:
        kind = "icon" if "vertex_of_cell" in ds else "mpas" if "verticesOnCell" in ds else None
    if kind == "icon":
        return read_icon_grid(filename, dual)
    if kind == "mpas":
        return read_mpas_grid(filename, dual)
    raise ValueError(f"{os.fspath(filename)} is neither an ICON nor an MPAS grid file")
