"""Meshes and topology: restates test/grids.jl:105-131,209-222 and test/topology.jl:42-60
of the reference for the host library (icosahedral bisection, Lloyd SCVT, hexagon dual,
Plex conventions, Hilbert order)."""
import itertools

import numpy as np
import pytest

from dmplex_amd import (TRiSKGrid, cell_centers, check_plex, create_hexagon_mesh, create_spherical_mesh,
                        hilbert_key, hilbert_keys, hilbert_order, hilbert_permutation, lloyd_, lloyd_residual,
                        orthogonality_defect, plex_from_polygons, voronoi_generators)
from dmplex_amd.grids import _icosahedral
from dmplex_amd.topology import cell_centroids


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4])
def test_counts_and_plex_invariants(n):
    tri = create_spherical_mesh(n)
    hexm = create_hexagon_mesh(n)
    assert (tri.nc, tri.ne, tri.nv) == (20 * 4 ** n, 30 * 4 ** n, 10 * 4 ** n + 2)
    assert (hexm.nc, hexm.ne, hexm.nv) == (10 * 4 ** n + 2, 30 * 4 ** n, 20 * 4 ** n)
    assert check_plex(tri) and check_plex(hexm)
    sizes = np.diff(hexm.cone_offsets[:hexm.nc + 1])
    assert np.sum(sizes == 5) == 12 and np.all((sizes == 5) | (sizes == 6))


def test_plex_conventions():
    """Edges run low -> high vertex, orientation 0 = left cell, edges by first appearance."""
    dm = create_hexagon_mesh(2)
    for e in dm.edge_range():
        a, b = dm.get_cone(e)
        assert a < b
        left, right = dm.get_edge_cells(e)
        assert left >= 0 and right >= 0 and left != right
        k = list(dm.get_cone(left)).index(e)
        assert dm.edge_sign(left, k) == 1
    # first appearance: the first cell's edges are 0..n-1 in order
    first = dm.get_cone(0) - dm.nc
    assert list(first) == list(range(len(first)))
    # a clockwise polygon is reversed
    x = np.array([[1, 0, 0.1], [0, 1, 0.1], [-1, 0, 0.1], [0, -1, 0.1], [0, 0, 1.0]])
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    x = np.vstack([x, [[0, 0, -1.0]]])
    polys = [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4], [1, 0, 5], [2, 1, 5], [3, 2, 5], [0, 3, 5]]
    cw = [p[::-1] for p in polys]
    assert check_plex(plex_from_polygons(polys, x)) and check_plex(plex_from_polygons(cw, x))


def test_circumcenters_equidistant():
    ico = create_spherical_mesh(2)
    c = cell_centers(ico)
    tri = ico.cell_vertex_table()
    d = np.stack([np.sum(c * ico.coordinates[tri[:, k]], axis=1) for k in range(3)], axis=1)
    assert np.max(d.max(axis=1) - d.min(axis=1)) < 1e-14


def test_voronoi_generators_recover_primal_vertices():
    ico, hexm = create_spherical_mesh(2), create_hexagon_mesh(2)
    gens = voronoi_generators(hexm)
    assert orthogonality_defect(hexm, gens) < 1e-12
    assert np.max(np.linalg.norm(gens - ico.coordinates, axis=1)) < 1e-12
    assert np.max(np.linalg.norm(hexm.generators - ico.coordinates, axis=1)) == 0


def test_scvt_fixed_point_and_fixed_vertices():
    x, tri, its = _icosahedral(4, True)
    x0, _, _ = _icosahedral(0, False)
    assert np.array_equal(x[:12], x0)
    assert np.max(np.abs(np.linalg.norm(x, axis=1) - 1)) < 1e-14
    y = x.copy()
    assert lloyd_(y, tri, fixed=12) < 3
    assert np.max(np.linalg.norm(y - x, axis=1)) < 1e-9
    assert lloyd_residual(x, tri, fixed=12) < 1e-7


def test_anderson_matches_plain_lloyd():
    _, tri, _ = _icosahedral(4, False)
    x, _, _ = _icosahedral(4, False)
    y, z = x.copy(), x.copy()
    na = lloyd_(y, tri, fixed=12)
    npl = lloyd_(z, tri, fixed=12, history=0)
    assert na < npl
    assert np.max(np.linalg.norm(y - z, axis=1)) < 1e-7


def test_lloyd_warns_when_it_stops_short():
    """The reference warns when lloyd! reaches max_iterations (Icosahedral.jl:183)."""
    x, tri, _ = _icosahedral(3, False)
    with pytest.warns(RuntimeWarning, match="did not converge in 5 iterations"):
        assert lloyd_(x, tri, fixed=12, max_iterations=5) == 5
    y, _, _ = _icosahedral(3, False)
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert lloyd_(y, tri, fixed=12) < 10_000  # converged: no warning


def test_scvt_improves_trisk_tangential_reconstruction():
    def tangential_error(centroidal):
        tri = create_spherical_mesh(4, centroidal=centroidal)
        g = TRiSKGrid(create_hexagon_mesh(4, primal=tri), centers=tri.coordinates)
        assert orthogonality_defect(g.dm, g.centers) < 1e-10
        w = np.array([0.3, -0.2, 1.0])
        u = np.sum(np.cross(w, g.edge_points) * g.edge_normals, axis=1)
        vt = np.sum(g.trisk_weights * g.primal_lengths[g.trisk_edges] * u[g.trisk_edges], axis=0) / g.dual_lengths
        exact = np.sum(np.cross(w, g.edge_points) * np.cross(g.edge_points, g.edge_normals), axis=1)
        return np.max(np.abs(vt - exact))

    assert tangential_error(True) < 0.1 * tangential_error(False)


def test_hilbert_key_is_a_hilbert_curve():
    """bits = 2 and 3: a bijection onto 0..8^bits-1 whose consecutive points are lattice neighbours."""
    for bits in (2, 3):
        side = 2 ** bits
        pts = np.array(list(itertools.product(range(side), repeat=3)), dtype=np.uint64)
        keys = hilbert_keys(pts, bits)
        assert np.array_equal(np.sort(keys), np.arange(side ** 3, dtype=np.uint64))
        order = pts[np.argsort(keys)].astype(np.int64)
        steps = np.abs(np.diff(order, axis=0)).sum(axis=1)
        assert np.all(steps == 1)
    assert hilbert_key((0, 0, 0), 16) == 0


def test_hilbert_order():
    dm = create_hexagon_mesh(2)
    new = hilbert_order(dm)
    assert check_plex(new)
    assert (new.nc, new.ne, new.nv) == (dm.nc, dm.ne, dm.nv)
    perm = hilbert_permutation(dm)
    assert np.allclose(cell_centroids(new), cell_centroids(dm)[perm])
    # locality: consecutive cells are close
    c = cell_centroids(new)
    step = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert np.median(step) < 2.5 * np.sqrt(4 * np.pi / dm.nc)


def test_scvt_levels_converge():
    _, _, its = _icosahedral(5, True)
    assert len(its) == 5 and np.all(its >= 1) and np.all(its < 10000)
