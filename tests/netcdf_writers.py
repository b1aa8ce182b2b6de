"""Writers of NetCDF test files (test infrastructure, not collected by pytest).

* classic and 64-bit offset files: ``scipy.io.netcdf_file`` (an independent writer);
* 64-bit data (CDF-5) files: written here from the published format specification;
* NetCDF-4 files: HDF5 datasets in the root group, written through the HDF5 C library
  (ctypes), with the dimension scales netCDF-4 adds for dimensions without variables.

``write_icon_file`` and ``write_mpas_file`` restate the reference's test writers
(test/grids.jl:3-50): the same variables, layouts and padding.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

FORMATS = ["classic", "64bit_offset", "64bit_data", "netcdf4"]

_CDF_TYPES = {np.dtype("int8"): (1, "b"), np.dtype("S1"): (2, "c"), np.dtype("int16"): (3, "h"),
              np.dtype("int32"): (4, "i"), np.dtype("float32"): (5, "f"), np.dtype("float64"): (6, "d"),
              np.dtype("uint8"): (7, "B"), np.dtype("uint16"): (8, "H"), np.dtype("uint32"): (9, "I"),
              np.dtype("int64"): (10, "q"), np.dtype("uint64"): (11, "Q")}


# --- HDF5 ---------------------------------------------------------------------------------
_H5 = None


def hdf5_library():
    """The HDF5 C library (ctypes), or None when the system has none."""
    global _H5
    if _H5 is not None:
        return _H5 or None
    names = [os.environ.get("DMPLEX_HDF5_LIBRARY"), "libhdf5.so", "libhdf5_serial.so", "libhdf5.so.103",
             "/opt/conda/lib/libhdf5.so"]
    for n in names:
        if not n:
            continue
        try:
            lib = C.CDLL(n)
            break
        except OSError:
            continue
    else:
        _H5 = False
        return None
    hid, herr = C.c_int64, C.c_int
    sig = {
        "H5open": ([], herr), "H5Fcreate": ([C.c_char_p, C.c_uint, hid, hid], hid), "H5Fclose": ([hid], herr),
        "H5Screate_simple": ([C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p], hid), "H5Screate": ([C.c_int], hid),
        "H5Sclose": ([hid], herr), "H5Tcopy": ([hid], hid), "H5Tset_size": ([hid, C.c_size_t], herr),
        "H5Tclose": ([hid], herr), "H5Pcreate": ([hid], hid), "H5Pclose": ([hid], herr),
        "H5Pset_chunk": ([hid, C.c_int, C.POINTER(C.c_ulonglong)], herr), "H5Pset_deflate": ([hid, C.c_uint], herr),
        "H5Pset_shuffle": ([hid], herr), "H5Pset_libver_bounds": ([hid, C.c_int, C.c_int], herr),
        "H5Dcreate2": ([hid, C.c_char_p, hid, hid, hid, hid, hid], hid),
        "H5Dwrite": ([hid, hid, hid, hid, hid, C.c_void_p], herr), "H5Dclose": ([hid], herr),
        "H5Acreate2": ([hid, C.c_char_p, hid, hid, hid, hid], hid), "H5Awrite": ([hid, hid, C.c_void_p], herr),
        "H5Aclose": ([hid], herr),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res
    if lib.H5open() < 0:
        _H5 = False
        return None
    _H5 = lib
    return lib


def _h5_global(lib, name: str) -> int:
    return C.c_int64.in_dll(lib, name).value


def _h5_string_attr(lib, obj: int, name: str, text: str, variable: bool = False) -> None:
    t = lib.H5Tcopy(_h5_global(lib, "H5T_C_S1_g"))
    space = lib.H5Screate(0)  # H5S_SCALAR
    if variable:
        lib.H5Tset_size(t, C.c_size_t(-1).value)  # H5T_VARIABLE
        buf = C.c_char_p(text.encode())
        a = lib.H5Acreate2(obj, name.encode(), t, space, 0, 0)
        assert a >= 0 and lib.H5Awrite(a, t, C.byref(buf)) >= 0
    else:
        raw = text.encode()
        lib.H5Tset_size(t, max(len(raw), 1))
        a = lib.H5Acreate2(obj, name.encode(), t, space, 0, 0)
        assert a >= 0 and lib.H5Awrite(a, t, C.create_string_buffer(raw, max(len(raw), 1))) >= 0
    lib.H5Aclose(a)
    lib.H5Sclose(space)
    lib.H5Tclose(t)


def write_hdf5(path, dims: dict, variables: list, attributes: dict, libver_latest: bool = False,
               chunk: int = 0, deflate: int = 0, variable_strings: bool = False) -> None:
    """A NetCDF-4-style file: one dataset per variable, a dimension scale (with netCDF-4's
    NAME attribute) for every dimension no variable is named after, string attributes."""
    lib = hdf5_library()
    if lib is None:
        raise RuntimeError("no HDF5 library")
    native = {np.dtype("int8"): "H5T_NATIVE_SCHAR_g", np.dtype("uint8"): "H5T_NATIVE_UCHAR_g",
              np.dtype("int16"): "H5T_NATIVE_SHORT_g", np.dtype("uint16"): "H5T_NATIVE_USHORT_g",
              np.dtype("int32"): "H5T_NATIVE_INT_g", np.dtype("uint32"): "H5T_NATIVE_UINT_g",
              np.dtype("int64"): "H5T_NATIVE_LLONG_g", np.dtype("uint64"): "H5T_NATIVE_ULLONG_g",
              np.dtype("float32"): "H5T_NATIVE_FLOAT_g", np.dtype("float64"): "H5T_NATIVE_DOUBLE_g"}
    fapl = 0
    if libver_latest:
        fapl = lib.H5Pcreate(_h5_global(lib, "H5P_CLS_FILE_ACCESS_ID_g"))
        lib.H5Pset_libver_bounds(fapl, 2, 2)  # H5F_LIBVER_LATEST (1.10: V110)
    f = lib.H5Fcreate(os.fspath(path).encode(), 2, 0, fapl)  # H5F_ACC_TRUNC
    assert f >= 0, path
    if fapl:
        lib.H5Pclose(fapl)
    named = {v[0] for v in variables}
    for dim, n in dims.items():
        if dim in named:
            continue
        d = np.zeros(n, dtype=np.float32)
        space = lib.H5Screate_simple(1, (C.c_ulonglong * 1)(n), None)
        ds = lib.H5Dcreate2(f, dim.encode(), _h5_global(lib, "H5T_NATIVE_FLOAT_g"), space, 0, 0, 0)
        lib.H5Dwrite(ds, _h5_global(lib, "H5T_NATIVE_FLOAT_g"), 0, 0, 0, d.ctypes.data)
        _h5_string_attr(lib, ds, "CLASS", "DIMENSION_SCALE")
        _h5_string_attr(lib, ds, "NAME", f"This is a netCDF dimension but not a netCDF variable.{n:10d}")
        lib.H5Dclose(ds)
        lib.H5Sclose(space)
    for name, _, data in variables:
        data = np.ascontiguousarray(data)
        shape = (C.c_ulonglong * max(data.ndim, 1))(*data.shape)
        space = lib.H5Screate_simple(data.ndim, shape, None)
        dcpl = 0
        if chunk and data.ndim:
            dcpl = lib.H5Pcreate(_h5_global(lib, "H5P_CLS_DATASET_CREATE_ID_g"))
            cshape = (C.c_ulonglong * data.ndim)(*[max(1, min(chunk, s)) for s in data.shape])
            lib.H5Pset_chunk(dcpl, data.ndim, cshape)
            if deflate:
                lib.H5Pset_shuffle(dcpl)
                lib.H5Pset_deflate(dcpl, deflate)
        t = _h5_global(lib, native[data.dtype])
        ds = lib.H5Dcreate2(f, name.encode(), t, space, 0, dcpl, 0)
        assert ds >= 0, name
        assert lib.H5Dwrite(ds, t, 0, 0, 0, data.ctypes.data) >= 0
        lib.H5Dclose(ds)
        if dcpl:
            lib.H5Pclose(dcpl)
        lib.H5Sclose(space)
    for name, text in attributes.items():
        _h5_string_attr(lib, f, name, text, variable_strings)
    lib.H5Fclose(f)


# --- CDF-5 --------------------------------------------------------------------------------
def write_cdf5(path, dims: dict, variables: list, attributes: dict, record_dim: str | None = None,
               numrecs: int = 0) -> None:
    """A 64-bit data (CDF-5) file.  ``dims[record_dim]`` is ignored: the record
    dimension's length is ``numrecs``; record variables have it first."""
    def name(s):
        b = s.encode()
        return struct.pack(">q", len(b)) + b + b"\0" * (-len(b) % 4)

    def values(arr):
        code, fmt = _CDF_TYPES[arr.dtype]
        raw = arr.astype(arr.dtype.newbyteorder(">")).tobytes() if arr.dtype.kind != "S" else arr.tobytes()
        return code, raw + b"\0" * (-len(raw) % 4)

    dim_names = list(dims)
    head = b"CDF\x05" + struct.pack(">q", numrecs)
    head += struct.pack(">iq", 10, len(dims)) if dims else struct.pack(">iq", 0, 0)
    for d in dim_names:
        head += name(d) + struct.pack(">q", 0 if d == record_dim else dims[d])
    if attributes:
        head += struct.pack(">iq", 12, len(attributes))
        for k, text in attributes.items():
            raw = text.encode()
            head += name(k) + struct.pack(">iq", 2, len(raw)) + raw + b"\0" * (-len(raw) % 4)
    else:
        head += struct.pack(">iq", 0, 0)
    recvars = [v for v in variables if record_dim is not None and v[1] and v[1][0] == record_dim]
    fixed = [v for v in variables if v not in recvars]
    # header size first (offsets are fixed-width), then the offsets
    def var_entries(offsets):
        out = struct.pack(">iq", 11, len(variables)) if variables else struct.pack(">iq", 0, 0)
        for vname, vdims, data in variables:
            code = _CDF_TYPES[np.asarray(data).dtype][0]
            out += name(vname) + struct.pack(">q", len(vdims))
            out += b"".join(struct.pack(">q", dim_names.index(d)) for d in vdims)
            out += struct.pack(">iq", 0, 0) + struct.pack(">i", code)
            n = np.asarray(data).dtype.itemsize
            for s in (np.asarray(data).shape[1:] if vdims[:1] == (record_dim,) else np.asarray(data).shape):
                n *= s
            out += struct.pack(">qq", n + (-n % 4) if len(recvars) != 1 or vdims[:1] != (record_dim,) else n,
                               offsets.get(vname, 0))
        return out

    hsize = len(head) + len(var_entries({}))
    offsets, pos, body = {}, hsize, b""
    for vname, _, data in fixed:
        offsets[vname] = pos
        raw = values(np.ascontiguousarray(data))[1]
        body += raw
        pos += len(raw)
    recsize = 0
    for vname, _, data in recvars:
        offsets[vname] = pos + recsize
        n = np.asarray(data).dtype.itemsize * int(np.prod(np.asarray(data).shape[1:]))
        recsize += n if len(recvars) == 1 else n + (-n % 4)
    for r in range(numrecs):
        for vname, _, data in recvars:
            a = np.ascontiguousarray(np.asarray(data)[r])
            raw = a.astype(a.dtype.newbyteorder(">")).tobytes()
            body += raw + (b"" if len(recvars) == 1 else b"\0" * (-len(raw) % 4))
    with open(path, "wb") as f:
        f.write(head + var_entries(offsets) + body)


# --- dispatch -----------------------------------------------------------------------------
def write_netcdf(path, fmt: str, dims: dict, variables: list, attributes: dict | None = None, **hdf5_options):
    """``variables``: (name, dimension names in C order, array)."""
    attributes = attributes or {}
    if fmt in ("classic", "64bit_offset"):
        from scipy.io import netcdf_file

        with netcdf_file(os.fspath(path), "w", version=1 if fmt == "classic" else 2) as f:
            for d, n in dims.items():
                f.createDimension(d, n)
            for k, text in attributes.items():
                setattr(f, k, text)
            for name, vdims, data in variables:
                data = np.asarray(data)
                v = f.createVariable(name, data.dtype if data.dtype.kind != "S" else "c", vdims)
                v[:] = data
    elif fmt == "64bit_data":
        write_cdf5(path, dims, variables, attributes)
    elif fmt == "netcdf4":
        write_hdf5(path, dims, variables, attributes, **hdf5_options)
    else:
        raise ValueError(fmt)


# --- the reference's grid-file writers (test/grids.jl:3-50) -------------------------------
def xyz_to_lonlat(x: np.ndarray):
    x = np.asarray(x)
    return np.arctan2(x[:, 1], x[:, 0]), np.arcsin(np.clip(x[:, 2], -1, 1))


def cell_vertices(dm):
    """CSR (offsets, 0-based vertex indices) of every cell's counter-clockwise ring."""
    sizes = np.diff(dm.cone_offsets[:dm.nc + 1])
    nconn = dm.cone_offsets[dm.nc]
    cone, orient = dm.cones[:nconn], dm.cone_orientations[:nconn]
    start = np.where(orient == 0, dm.cones[dm.cone_offsets[cone]], dm.cones[dm.cone_offsets[cone] + 1])
    offsets = np.zeros(dm.nc + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    return offsets, start - dm.nc - dm.ne


def vertex_stars(dm, width: int, drop_cells=()) -> np.ndarray:
    """nv x width table of the 1-based cells around every vertex in increasing order,
    0-padded; cells in ``drop_cells`` are left out."""
    offsets, verts = cell_vertices(dm)
    owner = np.repeat(np.arange(dm.nc), np.diff(offsets))
    keep = ~np.isin(owner, np.fromiter(drop_cells, dtype=np.int64, count=len(drop_cells)))
    order = np.argsort(verts, kind="stable")
    order = order[keep[order]]
    v, c = verts[order], owner[order]
    first = np.searchsorted(v, np.arange(dm.nv))
    slot = np.arange(len(v)) - first[v]
    table = np.zeros((dm.nv, width), dtype=np.int32)
    table[v, slot] = c + 1
    return table


def write_icon_file(path, dm, fmt: str = "netcdf4", drop_cells=(), **options) -> None:
    """ICON-like file of a triangle mesh: vlon/vlat, clon/clat (circumcenters),
    vertex_of_cell(nv, cell) and cells_of_vertex(ne, vertex), 1-based, 0-padded.
    Cells in ``drop_cells`` are left out of cells_of_vertex (a regional grid)."""
    from dmplex_amd import cell_centers

    centers = cell_centers(dm)
    offsets, verts = cell_vertices(dm)
    if not np.all(np.diff(offsets) == 3):
        raise ValueError("an ICON grid has triangles only")
    voc = (verts.reshape(-1, 3) + 1).astype(np.int32)
    cov = vertex_stars(dm, 6, drop_cells)
    vlon, vlat = xyz_to_lonlat(dm.coordinates)
    clon, clat = xyz_to_lonlat(centers)
    write_netcdf(path, fmt, {"vertex": dm.nv, "cell": dm.nc, "nv": 3, "ne": 6},
                 [("vlon", ("vertex",), vlon), ("vlat", ("vertex",), vlat), ("clon", ("cell",), clon),
                  ("clat", ("cell",), clat), ("vertex_of_cell", ("nv", "cell"), np.ascontiguousarray(voc.T)),
                  ("cells_of_vertex", ("ne", "vertex"), np.ascontiguousarray(cov.T))], {}, **options)


def write_mpas_file(path, dm, fmt: str = "netcdf4", radius: float = 6.371229e6, centers=None,
                    on_a_sphere: str = "YES", drop_cells=(), **options) -> None:
    """MPAS-like file of a Voronoi mesh: x/y/z Vertex and Cell scaled by ``radius``,
    verticesOnCell(nCells, maxEdges) padded with nv + 1, nEdgesOnCell, and
    cellsOnVertex(nVertices, vertexDegree) with 0 for the cells in ``drop_cells``."""
    from dmplex_amd import cell_centers

    if centers is None:
        centers = cell_centers(dm)
    offsets, verts = cell_vertices(dm)
    noc = np.diff(offsets).astype(np.int32)
    voc = np.full((dm.nc, 6), dm.nv + 1, dtype=np.int32)
    owner = np.repeat(np.arange(dm.nc), noc)
    voc[owner, np.arange(len(verts)) - offsets[owner]] = verts + 1
    cov = vertex_stars(dm, 3)
    if drop_cells:
        cov[np.isin(cov - 1, list(drop_cells))] = 0
    variables = []
    for name, pts, dim in (("Vertex", dm.coordinates, "nVertices"), ("Cell", centers, "nCells")):
        for k, axis in enumerate("xyz"):
            variables.append((axis + name, (dim,), radius * np.asarray(pts)[:, k]))
    variables += [("verticesOnCell", ("nCells", "maxEdges"), voc), ("nEdgesOnCell", ("nCells",), noc),
                  ("cellsOnVertex", ("nVertices", "vertexDegree"), cov)]
    write_netcdf(path, fmt, {"nCells": dm.nc, "nVertices": dm.nv, "maxEdges": 6, "vertexDegree": 3}, variables,
                 {"on_a_sphere": on_a_sphere}, **options)
