"""The MPAS-Gassmann stage operators on the host: tables and oracle.

Restates the reference's operator tests of the extruded C grid (test/extruded.jl:55-105):
the PV usage of i_cori_method = 4 is symmetric, the PV flux does no work, and the vorticity
term exchanges horizontal and vertical kinetic energy without loss (both in the energy
sum_e l d rho_e vn^2/2 + sum_i A rho w^2/2), for the ICON-IAP and the vector-invariant forms.
These pin the signs, the weights and the level indices of the edge-first TRiSK tables,
eta, edge_w and the w tendency.  Closed forms pin the rest: a resting atmosphere has no
tendency, the advective w tendency of a horizontally uniform w is -w dw/dz.
"""
import numpy as np
import pytest

from dmplex_amd import TRiSKGrid, create_hexagon_mesh
from dmplex_amd.mpas_ops import (Cvd, P0, Rd, MPASOperatorTables, iap_trisk_weights, trisk_pv_edges,
                                 vertical_grid, vertical_grid_from_half)
from oracle.oracle import MpasOracle


@pytest.fixture(scope="module")
def grid():
    return TRiSKGrid(create_hexagon_mesh(3))


def test_vertical_grid():
    v = vertical_grid(4, 10000.0)
    np.testing.assert_allclose(v.z_half, [0, 2500, 5000, 7500, 10000])
    np.testing.assert_allclose(v.dzh[1:4], 2500)
    np.testing.assert_allclose(v.wlow[1:4], 0.5)
    assert v.dzh[0] == v.dzh[4] == v.wlow[0] == v.wlow[4] == 0
    s = vertical_grid(10, 30000.0, stretch=3.0)
    assert s.dz[-1] / s.dz[0] == pytest.approx(3.0) and s.z_half[-1] == pytest.approx(30000.0)
    with pytest.raises(ValueError):
        vertical_grid_from_half([0.0, 1.0, 1.0])


def test_pv_edges_symmetric_and_iap_weights(grid):
    g = grid
    P = trisk_pv_edges(g)
    for e in range(g.num_edges()):
        for k in range(g.trisk_edges.shape[0]):
            if g.trisk_weights[k, e] == 0:
                continue
            e2 = g.trisk_edges[k, e]
            back = np.flatnonzero(g.trisk_edges[:, e2] == e)
            assert P[k, e] in P[back, e2]
    # a pentagon/hexagon walk: s = n/2 (hexagon opposite edge) is the mean, the rest name an edge
    assert np.any(P == -1) and np.any(P >= 0)
    # planar kites: close to the exact spherical weights, on the same stencil
    W = iap_trisk_weights(g)
    assert np.max(np.abs(W - g.trisk_weights)) < 1e-2
    assert not np.any(W[g.trisk_weights == 0][g.trisk_edges[g.trisk_weights == 0] == 0])


def _state(g, v, rng):
    ne, nc, nl = g.num_edges(), g.num_cells(), len(v.dz)
    vn = rng.standard_normal((nl, ne))
    w = rng.standard_normal((nl + 1, nc))
    w[0] = 0
    w[-1] = 0
    rho = 1 + 0.3 * rng.random((nl, nc))
    return vn, w, rho


@pytest.mark.parametrize("pv_flux,vertical_advection", [("iap", "iap"), ("trisk", "vector_invariant")])
def test_pv_flux_and_vorticity_energy(grid, pv_flux, vertical_advection):
    g = grid
    v = vertical_grid(4, 10000.0)
    rng = np.random.default_rng(1)
    vn, w, rho = _state(g, v, rng)
    ne, nc, nl = g.num_edges(), g.num_cells(), len(v.dz)
    Ae = g.primal_lengths * g.dual_lengths
    l, r = g.edge_cells
    rhoe = (g.cell_distances[0][None, :] * rho[:, l] + g.cell_distances[1][None, :] * rho[:, r]) / g.dual_lengths
    rhoh = np.zeros((nl + 1, nc))
    for j in range(1, nl):
        rhoh[j] = v.wlow[j] * rho[j - 1] + (1 - v.wlow[j]) * rho[j]
    m = MPASOperatorTables(g, v, radius=1.0, pv_flux=pv_flux, vertical_advection=vertical_advection,
                           trisk_weights="exact")
    mw = vertical_advection == "iap"
    o = MpasOracle(m)
    W = o.work
    W["K"][:] = 0
    W["exner"][:] = 0
    W["theta"][:] = 1
    o.kernel("vertex_pv", vn=vn, rho=rho)
    o.kernel("edge_pv")
    o.kernel("edge_density", vn=vn, rho=rho)
    np.testing.assert_allclose(W["rhoe"], rhoe, rtol=1e-14)

    def tendency(qe, w_, eta):
        W["qe"][:], W["eta"][:] = qe, eta
        o.kernel("vn_tendency", vn=vn, w=w_, rho=rho, pressure=1.0)
        return W["tvn"].copy()

    # the PV flux does no work
    F = rhoe * vn
    t = tendency(W["qe"].copy(), np.zeros_like(w), np.zeros((nl + 1, ne)))
    work = (Ae[None, :] * F * t).T @ v.dz
    assert abs(work.sum()) < 1e-12 * ((Ae[None, :] * np.abs(F * t)).T @ v.dz).sum()
    # the vorticity term exchanges horizontal and vertical kinetic energy without loss
    o.kernel("eta", vn=vn, w=w)
    o.kernel("w_tendency_vi", vn=vn)
    t = tendency(np.zeros((nl, ne)), w, W["eta"].copy())
    a = ((Ae[None, :] * ((rhoe if mw else 1) * vn * t)).T @ v.dz).sum()
    b = ((g.cell_areas[None, :] * ((rhoh if mw else 1) * w * W["tw"])).T @ v.dzh).sum()
    assert abs(a) > 0 and abs(a + b) < 1e-12 * abs(a)


@pytest.mark.parametrize("opts", [dict(), dict(pv_flux="trisk", vertical_advection="vector_invariant",
                                                  vertical_solver="tridiagonal"),
                                  dict(vertical_advection="advective", vertical_solver="tridiagonal"),
                                  dict(kinetic_energy="mixed")])
def test_rest_state(grid, opts):
    """A horizontally uniform atmosphere at rest has no tendency: tvn = 0 and tw = 0 exactly."""
    g = grid
    v = vertical_grid(6, 20000.0, stretch=2.0)
    m = MPASOperatorTables(g, v, **opts)
    nl, nc, ne = len(v.dz), g.num_cells(), g.num_edges()
    rho = np.repeat(np.exp(-v.z_full / 8000.0)[:, None], nc, axis=1)
    rt = rho * 300.0 * np.exp(v.z_full / 20000.0)[:, None]
    out = MpasOracle(m).slow_tendencies(np.zeros((nl, ne)), np.zeros((nl + 1, nc)), rho, rt, 1.0)
    assert not np.any(out["tvn"]) and not np.any(out["tw"]) and not np.any(out["K"])
    np.testing.assert_array_equal(out["theta"], rt / rho)
    # pow differs in the last bit between libms (numpy's SIMD pow, glibc, the device's)
    np.testing.assert_allclose(out["exner"], (Rd * rt / P0) ** (Rd / Cvd), rtol=5e-16, atol=0)


def test_advective_w_tendency(grid):
    """Horizontally uniform w: no horizontal advection, tw = -w (w_up - w_down)/(dz_below + dz_above)."""
    g = grid
    v = vertical_grid(5, 10000.0, stretch=1.5)
    m = MPASOperatorTables(g, v, vertical_advection="advective", vertical_solver="tridiagonal")
    nl, nc, ne = len(v.dz), g.num_cells(), g.num_edges()
    rng = np.random.default_rng(3)
    wcol = np.concatenate([[0.0], rng.standard_normal(nl - 1), [0.0]])
    w = np.repeat(wcol[:, None], nc, axis=1)
    vn = rng.standard_normal((nl, ne))
    rho = np.ones((nl, nc))
    out = MpasOracle(m).slow_tendencies(vn, w, rho, 300 * rho, 0.0)
    for j in range(1, nl):
        want = -(wcol[j] * (wcol[j + 1] - wcol[j - 1]) / (v.dz[j - 1] + v.dz[j]))
        np.testing.assert_array_equal(out["tw"][j], np.full(nc, want))
    assert not np.any(out["tw"][0]) and not np.any(out["tw"][nl])


def test_table_errors(grid):
    v = vertical_grid(3, 1000.0)
    with pytest.raises(ValueError, match="five_band"):
        MPASOperatorTables(grid, v, vertical_advection="advective")
    with pytest.raises(ValueError, match="pv_flux"):
        MPASOperatorTables(grid, v, pv_flux="gassmann")
    m = MPASOperatorTables(grid, v)
    m.check()
    bad = m.trisk_index.copy(order="F")
    bad[3, 2, 1] = grid.num_edges()
    m.trisk_index = bad
    with pytest.raises(ValueError, match="trisk_index"):
        m.check()
