"""Grid-file input: restates the reference's "ICON and MPAS readers" test set
(test/grids.jl:52-85) over every NetCDF format, and checks the native NetCDF reader
against independent writers (scipy for classic and 64-bit offset files, the HDF5 C
library for NetCDF-4, a CDF-5 writer from the format specification)."""
import glob
import os

import numpy as np
import pytest
from netcdf_writers import (FORMATS, hdf5_library, write_cdf5, write_hdf5, write_icon_file, write_mpas_file,
                            write_netcdf)

from dmplex_amd import _native as N
from dmplex_amd import TRiSKGrid, cell_centers, check_plex, create_hexagon_mesh, create_spherical_mesh
from dmplex_amd.file_grids import (NetCDFFile, dual_polygons, lonlat_to_xyz, read_grid, read_icon_grid,
                                   read_mpas_grid, rows_per_entity, sort_ring)

needs_hdf5 = pytest.mark.skipif(hdf5_library() is None, reason="no HDF5 C library on this system")


def _formats():
    return [pytest.param(f, marks=needs_hdf5) if f == "netcdf4" else f for f in FORMATS]


def total_area(dm) -> float:
    return float(np.sum(TRiSKGrid(dm).cell_areas))


# --- the NetCDF reader ------------------------------------------------------------------------
def _scipy_files():
    import scipy.io

    return sorted(glob.glob(os.path.join(os.path.dirname(scipy.io.__file__), "tests", "data", "*.nc")))


def _compare_with_scipy(path):
    from scipy.io import netcdf_file

    with netcdf_file(path, "r", mmap=False) as ref, NetCDFFile(path) as ds:
        assert sorted(ds.keys()) == sorted(ref.variables)
        for name, v in ref.variables.items():
            t, shape = ds.info(name)
            assert shape == tuple(v.shape), name
            if v.typecode() == "c":
                assert t == "char"
                continue
            np.testing.assert_array_equal(ds.variable(name), np.asarray(v[:]).astype(ds.variable(name).dtype))
        for name, value in ref._attributes.items():
            if isinstance(value, bytes):
                assert ds.attribute(name) == value.decode()


@pytest.mark.parametrize("path", _scipy_files(), ids=os.path.basename)
def test_reader_matches_scipy_on_its_sample_files(path):
    """scipy's own NetCDF test files (record variables, fill values, char variables)."""
    _compare_with_scipy(path)


@pytest.mark.parametrize("fmt", ["classic", "64bit_offset"])
@pytest.mark.parametrize("nrecvars", [0, 1, 3])
def test_reader_matches_scipy(tmp_path, fmt, nrecvars):
    """Every classic type, multi-dimensional variables and 0/1/several record variables
    (a single record variable is stored without padding)."""
    from scipy.io import netcdf_file

    rng = np.random.default_rng(7)
    path = str(tmp_path / "t.nc")
    with netcdf_file(path, "w", version=1 if fmt == "classic" else 2) as f:
        f.createDimension("time", None)
        f.createDimension("x", 5)
        f.createDimension("y", 3)
        f.title = "reader test"
        f.on_a_sphere = "YES      "
        for code, data in (("b", rng.integers(-128, 127, (5, 3))), ("h", rng.integers(-999, 999, (5,))),
                           ("i", rng.integers(-2 ** 31, 2 ** 31 - 1, (3, 5))), ("f", rng.standard_normal((5, 3))),
                           ("d", rng.standard_normal((3,)))):
            dims = ("x", "y") if data.shape == (5, 3) else ("y", "x") if data.shape == (3, 5) else \
                ("x",) if data.shape == (5,) else ("y",)
            v = f.createVariable("v_" + code, code, dims)
            v[:] = data
        for k in range(nrecvars):
            code = "bhd"[k]  # byte and short slabs are padded inside each record
            v = f.createVariable(f"r{k}", code, ("time", "y") if k != 1 else ("time",))
            v[:] = rng.integers(-100, 100, (4, 3) if k != 1 else (4,))
        c = f.createVariable("name", "c", ("x",))
        c[:] = np.array(list(b"abcde"), dtype="S1")
    _compare_with_scipy(path)
    with NetCDFFile(path) as ds:
        assert ds.format == fmt
        assert ds.attribute("on_a_sphere").strip() == "YES" and ds.attribute("missing") is None
        with pytest.raises(ValueError, match="not an integer"):
            out = np.empty(3, dtype=np.int64)
            N.check(N.host().dmx_nc_read_i64(ds._h, b"v_d", N.ptr(out, N.i64p), 3))
        with pytest.raises(ValueError, match="character"):
            ds.variable("name")
        with pytest.raises(ValueError, match="no variable"):
            ds.variable("absent")


@pytest.mark.parametrize("numrecs", [0, 2])
def test_cdf5_reader(tmp_path, numrecs):
    """64-bit data files: 64-bit counts and offsets and the unsigned/int64 types."""
    rng = np.random.default_rng(3)
    arrays = {
        "i8": rng.integers(-2 ** 62, 2 ** 62, (4, 3)).astype(np.int64),
        "u8": rng.integers(0, 2 ** 63, (3,)).astype(np.uint64),
        "u4": rng.integers(0, 2 ** 32, (4,)).astype(np.uint32),
        "u2": rng.integers(0, 2 ** 16, (3,)).astype(np.uint16),
        "u1": rng.integers(0, 256, (4, 3)).astype(np.uint8),
        "f4": rng.standard_normal((3, 4)).astype(np.float32),
        "f8": rng.standard_normal((4,)),
    }
    dims = {"a": 4, "b": 3, "t": 0}
    dimmap = {(4, 3): ("a", "b"), (3,): ("b",), (4,): ("a",), (3, 4): ("b", "a")}
    variables = [(k, dimmap[v.shape], v) for k, v in arrays.items()]
    rec = rng.standard_normal((numrecs, 3))
    rec2 = rng.integers(-5, 5, (numrecs, 3)).astype(np.int16)
    variables += [("rec", ("t", "b"), rec), ("rec2", ("t", "b"), rec2)]
    path = str(tmp_path / "c5.nc")
    write_cdf5(path, dims, variables, {"on_a_sphere": "YES"}, record_dim="t", numrecs=numrecs)
    with NetCDFFile(path) as ds:
        assert ds.format == "64bit_data"
        assert ds.info("rec") == ("double", (numrecs, 3))
        for name, _, data in variables:
            got = ds.variable(name)
            want = data.astype(np.float64) if data.dtype.kind == "f" else data.astype(np.int64)
            np.testing.assert_array_equal(got, want)


@needs_hdf5
@pytest.mark.parametrize("layout", ["contiguous", "chunked-deflate", "latest-format", "vlen-strings"])
def test_netcdf4_reader(tmp_path, layout):
    """NetCDF-4 files: datasets in the root group, the pure dimension scales netCDF-4
    adds are not variables; chunked and compressed storage, the newest file format with
    dense link storage (more than 8 links), fixed and variable-length string attributes."""
    rng = np.random.default_rng(11)
    variables = [(f"v{k}", ("n", "m"), rng.standard_normal((6, 4))) for k in range(10)]
    variables += [("idx", ("n",), rng.integers(-50, 50, 6).astype(np.int32)),
                  ("big", ("m",), rng.integers(-2 ** 40, 2 ** 40, 4).astype(np.int64)),
                  ("small", ("n",), rng.standard_normal(6).astype(np.float32))]
    opts = {"contiguous": {}, "chunked-deflate": {"chunk": 3, "deflate": 4},
            "latest-format": {"libver_latest": True, "chunk": 2}, "vlen-strings": {"variable_strings": True}}[layout]
    path = str(tmp_path / "n4.nc")
    write_hdf5(path, {"n": 6, "m": 4, "unused": 2}, variables, {"on_a_sphere": "YES", "title": "x" * 300}, **opts)
    with NetCDFFile(path) as ds:
        assert ds.format == "netcdf4"
        assert sorted(ds.keys()) == sorted(v[0] for v in variables)
        assert "unused" not in ds and "v3" in ds
        assert ds.attribute("on_a_sphere") == "YES" and ds.attribute("title") == "x" * 300
        assert ds.attribute("absent") is None
        for name, _, data in variables:
            got = ds.variable(name)
            np.testing.assert_array_equal(got, data.astype(got.dtype))
        assert ds.info("small") == ("float", (6,)) and ds.info("big") == ("int64", (4,))


def test_reader_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        NetCDFFile(tmp_path / "missing.nc")
    junk = tmp_path / "junk.nc"
    junk.write_bytes(b"this is not a netcdf file at all")
    with pytest.raises(ValueError, match="not a NetCDF file"):
        NetCDFFile(junk)
    good = str(tmp_path / "good.nc")
    write_netcdf(good, "64bit_offset", {"x": 1000}, [("a", ("x",), np.arange(1000.0))])
    raw = open(good, "rb").read()
    for cut in (6, 30, len(raw) - 8):  # inside the header, before and inside the data
        bad = tmp_path / f"cut{cut}.nc"
        bad.write_bytes(raw[:cut])
        with pytest.raises(ValueError):
            NetCDFFile(bad)


@pytest.mark.parametrize("fmt", _formats())
def test_reader_survives_mutated_files(tmp_path, fmt):
    """Random byte changes and truncations of a valid file: the reader either reads every
    variable or raises ValueError/OSError; it never crashes or reads outside the file.
    ``make -C oracle sanitize-test`` runs this test (with the other host-library tests)
    against AddressSanitizer/UBSan builds of the host library and the oracle."""
    import random

    rng = random.Random(7)
    path = str(tmp_path / "base.nc")
    dims = {"x": 50, "y": 3} if fmt in ("netcdf4", "64bit_data") else {"t": None, "x": 50, "y": 3}
    write_netcdf(path, fmt, dims, [("a", ("x",), np.arange(50.0)),
                                   ("b", ("y", "x"), np.arange(150, dtype=np.int32).reshape(3, 50))])
    raw = open(path, "rb").read()
    outcomes = set()
    for trial in range(60):
        b = bytearray(raw)
        for _ in range(rng.randint(1, 4)):
            b[rng.randrange(0, min(len(b), 600))] = rng.randrange(256)
        if trial % 5 == 0:
            b = b[: rng.randrange(1, len(b))]
        bad = tmp_path / "bad.nc"
        bad.write_bytes(bytes(b))
        try:
            f = NetCDFFile(bad)
        except (ValueError, OSError):
            outcomes.add("rejected")
            continue
        for k in f.keys():
            try:
                f.variable(k)
                outcomes.add("read")
            except (ValueError, OSError):
                outcomes.add("variable rejected")
    assert "rejected" in outcomes


# --- ring sort --------------------------------------------------------------------------------
def test_sort_ring_brute_force():
    """sort_ring against a direct restatement: counter-clockwise seen from outside,
    starting at the smallest atan2 angle measured from the first point (ties stable)."""
    import math

    rng = np.random.default_rng(5)
    for _ in range(50):
        axis = rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        pts = axis + 0.1 * rng.standard_normal((7, 3))
        pts /= np.linalg.norm(pts, axis=1)[:, None]
        ids = rng.permutation(7)[:6]
        got = sort_ring(ids, pts, axis)
        e1 = pts[ids[0]] - axis

        def angle(i):
            v = pts[i] - axis
            return math.atan2(float(np.dot(np.cross(e1, v), axis)), float(np.dot(e1, v)))

        assert list(got) == sorted(ids, key=angle)
        # counter-clockwise: consecutive corners turn left around the axis
        ring = pts[got]
        turn = [np.dot(np.cross(ring[k] - axis, ring[(k + 1) % len(ring)] - axis), axis) for k in range(len(ring) - 1)]
        assert all(t > 0 for t in turn)


def test_dual_polygons_and_rows_per_entity():
    dm = create_spherical_mesh(1)
    centers = cell_centers(dm)
    stars = [[] for _ in range(dm.nv)]
    for c in dm.cell_range():
        for v in dm.get_cell_vertices(c):
            stars[dm.vertex_index(int(v))].append(c)
    offsets, ids = dual_polygons(stars, centers, dm.coordinates)
    assert list(np.diff(offsets)) == [len(s) for s in stars]
    for v in range(dm.nv):
        ring = ids[offsets[v]:offsets[v + 1]]
        assert sorted(ring) == sorted(stars[v])
    t = np.arange(12).reshape(3, 4)
    assert rows_per_entity(t, 4).shape == (4, 3) and np.array_equal(rows_per_entity(t, 4), t.T)
    assert rows_per_entity(t, 3) is t
    with pytest.raises(ValueError, match="does not match"):
        rows_per_entity(t, 5)
    np.testing.assert_allclose(np.linalg.norm(lonlat_to_xyz([0.3, -2.0], [0.1, 1.2]), axis=1), 1.0)
    np.testing.assert_allclose(lonlat_to_xyz(np.pi / 2, 0.0), [0, 1, 0], atol=1e-16)


# --- ICON and MPAS readers (test/grids.jl:52-85) ----------------------------------------------
@pytest.mark.parametrize("fmt", _formats())
def test_icon_and_mpas_readers(tmp_path, fmt):
    icosahedral = create_spherical_mesh(2)
    icon = tmp_path / "icon.nc"
    write_icon_file(icon, icosahedral, fmt)
    dm = read_grid(icon)
    assert check_plex(dm)
    assert (dm.nc, dm.ne, dm.nv) == (icosahedral.nc, icosahedral.ne, icosahedral.nv)
    np.testing.assert_allclose(dm.coordinates, icosahedral.coordinates, atol=1e-15)
    dual = read_icon_grid(icon, dual=True)
    assert check_plex(dual)
    assert dual.nc == icosahedral.nv
    assert np.sum(np.diff(dual.cone_offsets[:dual.nc + 1]) == 5) == 12
    assert total_area(dual) == pytest.approx(4 * np.pi, rel=1e-12)

    hexagonal = create_hexagon_mesh(2)
    mpas = tmp_path / "mpas.nc"
    write_mpas_file(mpas, hexagonal, fmt)
    dm = read_grid(mpas)
    assert check_plex(dm)
    assert dm.nc == hexagonal.nc
    np.testing.assert_allclose(np.linalg.norm(dm.coordinates, axis=1), 1.0, rtol=1e-15)
    delaunay = read_mpas_grid(mpas, dual=True)
    assert check_plex(delaunay)
    assert delaunay.nc == hexagonal.nv
    assert np.all(np.diff(delaunay.cone_offsets[:delaunay.nc + 1]) == 3)
    assert total_area(delaunay) == pytest.approx(4 * np.pi, rel=1e-12)

    other = tmp_path / "other.nc"
    write_netcdf(other, fmt, {"x": 3}, [("temperature", ("x",), np.zeros(3))])
    with pytest.raises(ValueError, match="neither an ICON nor an MPAS"):
        read_grid(other)


@pytest.mark.parametrize("fmt", ["64bit_offset", pytest.param("netcdf4", marks=needs_hdf5)])
def test_round_trip_keeps_the_mesh(tmp_path, fmt):
    """Cells and vertices keep the file's numbering and the rings their order, so a mesh
    written and read back is the same Plex (edges numbered alike), and its Voronoi
    generators come back from x/y/zCell."""
    for src, write in ((create_hexagon_mesh(3, centroidal=True), write_mpas_file),
                       (create_spherical_mesh(3, centroidal=True), write_icon_file)):
        path = tmp_path / "m.nc"
        write(path, src, fmt)
        dm = read_grid(path)
        for a in ("cone_offsets", "cones", "cone_orientations", "support_offsets", "supports"):
            assert np.array_equal(getattr(dm, a), getattr(src, a)), a
        # x/y/z scaled and normalized (MPAS), or through longitude and latitude (ICON)
        tol = 4e-16 if write is write_mpas_file else 1e-14
        np.testing.assert_allclose(dm.coordinates, src.coordinates, rtol=0, atol=tol)
        if write is write_mpas_file:
            np.testing.assert_allclose(dm.generators, src.generators, rtol=0, atol=4e-16)
            np.testing.assert_allclose(cell_centers(dm), src.generators, rtol=0, atol=4e-16)
            # the reference's inverse iteration recovers the same generators
            from dmplex_amd import voronoi_generators
            np.testing.assert_allclose(voronoi_generators(dm), src.generators, atol=1e-12)


def test_icon_dual_is_the_voronoi_mesh(tmp_path):
    """The dual of an ICON (SCVT) grid is the hexagon mesh of the same generators."""
    tri = create_spherical_mesh(3, centroidal=True)
    path = tmp_path / "icon.nc"
    write_icon_file(path, tri, "64bit_offset")
    dual = read_icon_grid(path, dual=True)
    hexm = create_hexagon_mesh(3, primal=tri)
    assert (dual.nc, dual.ne, dual.nv) == (hexm.nc, hexm.ne, hexm.nv)
    # longitude/latitude round trip: a few ulps
    np.testing.assert_allclose(dual.generators, tri.coordinates, rtol=0, atol=1e-14)
    np.testing.assert_allclose(cell_centers(dual), hexm.generators, rtol=0, atol=1e-14)
    g1, g2 = TRiSKGrid(dual), TRiSKGrid(hexm)
    np.testing.assert_allclose(np.sort(g1.cell_areas), np.sort(g2.cell_areas), rtol=1e-12)


def test_regional_and_planar_files(tmp_path):
    tri = create_spherical_mesh(2)
    icon = tmp_path / "regional.nc"
    write_icon_file(icon, tri, "classic", drop_cells={0, 1, 2})
    assert check_plex(read_grid(icon))  # the triangles themselves are whole
    with pytest.raises(ValueError, match="regional ICON"):
        read_icon_grid(icon, dual=True)
    hexm = create_hexagon_mesh(2)
    mpas = tmp_path / "regional_mpas.nc"
    write_mpas_file(mpas, hexm, "classic", drop_cells={5})
    delaunay = read_mpas_grid(mpas, dual=True)  # triangles touching cell 5 are dropped
    assert delaunay.nc == hexm.nv - hexm.get_cone_size(5)
    planar = tmp_path / "planar.nc"
    write_mpas_file(planar, hexm, "classic", on_a_sphere="NO")
    with pytest.raises(ValueError, match="planar"):
        read_mpas_grid(planar)
