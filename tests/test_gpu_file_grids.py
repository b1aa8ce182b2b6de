"""The TRiSK step on meshes read from grid files (read_grid, src/Grids/FileGrids.jl:98-107),
against the FP64 oracle: an MPAS Voronoi file and the hexagonal dual of an ICON file,
written at test time from meshes generated here."""
import sys

import numpy as np
import pytest
from netcdf_writers import hdf5_library, write_icon_file, write_mpas_file

from dmplex_amd import (HIPBackend, TRiSKShallowWater, create_hexagon_mesh, create_spherical_mesh, gather_state,
                        rk4_steps_, tendency_)
from dmplex_amd.file_grids import read_grid, read_icon_grid
from dmplex_amd.synthetic import earth_parameters, seeded_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
FMT = "netcdf4" if hdf5_library() is not None else "64bit_offset"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_mpas_file_model_matches_oracle_and_source_mesh(tmp_path):
    say(f"MPAS file ({FMT}) level 4: FP64 RK4 vs oracle")
    src = create_hexagon_mesh(4, centroidal=True)
    path = tmp_path / "x1.2562.nc"
    write_mpas_file(path, src, FMT, centers=src.generators)
    dm = read_grid(path)
    m = TRiSKShallowWater(dm, **earth_parameters())  # centers: the file's generators
    y = seeded_state(m.grid)
    ref = Oracle(m).rk4_steps(y.copy(), 600.0, 3)
    yd = m.upload_state(y)
    rk4_steps_(yd, 600.0, 3, m)
    got = gather_state(m, yd)
    assert np.array_equal(got, ref), rel(got, ref)
    # the same mesh up to the rounding of x/y/z: the same trajectory to ~1e-13
    m0 = TRiSKShallowWater(src, centers=src.generators, **earth_parameters())
    y0 = m0.upload_state(seeded_state(m0.grid))
    rk4_steps_(y0, 600.0, 3, m0)
    assert rel(got, gather_state(m0, y0)) < 1e-11


def test_icon_dual_model_fp32_bulk(tmp_path):
    say(f"ICON file ({FMT}) level 4 dual: FP32-bulk cell-pair tendency vs oracle")
    tri = create_spherical_mesh(4, centroidal=True)
    path = tmp_path / "icon_R2B2.nc"
    write_icon_file(path, tri, FMT)
    dual = read_icon_grid(path, dual=True)
    m = TRiSKShallowWater(dual, path="pair", backend=HIPBackend(precision="fp32_bulk"), **earth_parameters())
    assert m.path_info()[0]
    y = seeded_state(m.grid)
    ref = Oracle(m).tendency(y)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    got = gather_state(m, dy)
    ne = dual.ne
    assert rel(got[:ne], ref[:ne]) < 1e-3
    assert rel(got[ne:], ref[ne:]) < 2e-4
    A = m.cell_areas
    assert abs(np.dot(A, got[ne:])) < 1e-12 * np.dot(A, np.abs(got[ne:]))
