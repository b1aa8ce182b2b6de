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
from dmplex_amd.mpas_ops import (ICON_CPD, ICON_GRAV, ICON_RD, Cpd, Cvd, P0, Rd, MPASOperatorTables,
                                 column_metrics, dcmip_mountain_height, dcmip_rest, ground_extrapolation,
                                 height_difference, height_interpolation, iap_trisk_weights, lonlat,
                                 reference_atmosphere, terrain_following, trisk_pv_edges, vertical_grid,
                                 vertical_grid_from_half)
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


# --- orography: terrain-following levels (test/extruded.jl:440-517) ---------------------

def test_terrain_following_and_column_metrics():
    """test/extruded.jl:441-449: the levels follow the ground, flat above flat_height, the basic
    coordinate, crossing levels rejected; the column thicknesses add up to the column height."""
    v = vertical_grid(10, 12000.0)
    h = np.array([0.0, 1000.0, -200.0])
    z = terrain_following(v, h, flat_height=8000.0)
    assert np.array_equal(z[0], v.z_half) and z[1, 0] == 1000.0 and z[2, 0] == -200.0
    above = v.z_half >= 8000
    assert np.array_equal(z[1, above], v.z_half[above])
    np.testing.assert_allclose(terrain_following(v, h, coordinate="basic")[1], v.z_half + 1000 * (1 - v.z_half / 12000))
    with pytest.raises(ValueError, match="cross"):
        terrain_following(v, [20000.0])
    c = column_metrics(z)
    np.testing.assert_allclose(c.dz.sum(axis=1), 12000 - h)
    np.testing.assert_allclose(c.z_full[1], (z[1, :-1] + z[1, 1:]) / 2)


def test_height_interpolation_known_values():
    """test/extruded.jl:481-487: the layers and weights of the edge heights, and height_difference."""
    edge_cells = np.array([[0], [1]])
    zc = np.array([[0.0, 100.0, 300.0], [500.0, 700.0, 900.0]])
    lev, wt = height_interpolation(edge_cells, zc, np.array([[600.0, 200.0, 1000.0]]))
    assert np.array_equal(lev[:, 0, :], [[1, 1, 1], [0, 0, 1]])
    np.testing.assert_allclose(wt[:, 0, 0], [2.5, 0.5])
    np.testing.assert_allclose(wt[:, 0, 1], [0.5, -1.5])
    pip = np.array([[1.0, 2.0, 4.0], [3.0, 5.0, 7.0]])
    assert height_difference(pip, edge_cells, lev, wt, 0, 0) == pytest.approx((3 + 0.5 * 2) - (2 + 2.5 * 2))


def test_reference_atmosphere_is_hydrostatic():
    """ICON's reference atmosphere: T(0) = 288.15 K, p(0) = 101325 Pa and c_p dpi/dz = -g pi/T
    (hydrostatic balance), by central differences."""
    T0, pi0 = reference_atmosphere(0.0)
    assert T0 == pytest.approx(288.15) and pi0 == pytest.approx((101325.0 / P0) ** (ICON_RD / ICON_CPD), rel=1e-15)
    for z in (0.0, 3000.0, 11000.0, 25000.0):
        dz = 0.5
        dpi = (reference_atmosphere(z + dz)[1] - reference_atmosphere(z - dz)[1]) / (2 * dz)
        T, pi = reference_atmosphere(z)
        assert abs(dpi + ICON_GRAV * pi / (ICON_CPD * T)) < 1e-8 * abs(dpi)


def test_ground_extrapolation_exact_for_quadratics():
    v = vertical_grid(8, 12000.0, stretch=3.0)
    z = terrain_following(v, np.array([0.0, 700.0, 1800.0]), coordinate="basic")
    c = column_metrics(z)
    q = ground_extrapolation(c)
    for i in range(3):
        f = lambda x: 2.0 - 3e-3 * x + 4e-7 * x * x  # noqa: E731
        assert np.dot(q[i], f(c.z_full[i, :3])) == pytest.approx(f(z[i, 0]), rel=1e-12)


def _mountain(grid, levels=10, top=12000.0, **kw):
    v = vertical_grid(levels, top)
    return MPASOperatorTables(grid, v, surface_height=dcmip_mountain_height, coordinate="basic", flat_height=top, **kw)


def _rest_state(m):
    """The DCMIP 2012 test 2-0 atmosphere at rest (T = T0 - Gamma z) at the levels of every column."""
    g = m.grid
    T, p = dcmip_rest(m.z_full)
    rho = p / (Rd * T)
    theta = T * (P0 / p) ** (Rd / Cpd)
    nl, ne, nc = m.nlev, g.num_edges(), g.num_cells()
    return np.zeros((nl, ne)), np.zeros((nl + 1, nc)), rho, rho * theta


@pytest.mark.parametrize("pressure_gradient", ["levels", "height"])
def test_pressure_gradient_at_rest_over_the_mountain(grid, pressure_gradient):
    """At rest over the DCMIP mountain the horizontal pressure gradient at constant height
    vanishes; along the sloping levels, with the metric correction (or interpolated to the edge
    height), the discrete one is a truncation error, a small fraction of the uncorrected
    gradient along the levels.  Flat ground gives exactly zero."""
    m = _mountain(grid, pressure_gradient=pressure_gradient)
    assert m.orography and np.max(m.cell_metrics.z_half[:, 0]) > 1500
    state = _rest_state(m)
    tvn = MpasOracle(m).slow_tendencies(*state, 1.0)["tvn"]
    planted = _mountain(grid, pressure_gradient="levels")
    planted.ddxn = np.zeros_like(planted.ddxn)   # no slope term: the gradient along the levels
    raw = MpasOracle(planted).slow_tendencies(*state, 1.0)["tvn"]
    # 7.5e-4 and 2.5e-5 m/s^2 at hexagon 3 x 10 levels; the reference's 24 steps of 300 s over this
    # mountain stay below 0.5 m/s (test/extruded.jl:517), 0.18 m/s at this rate
    assert np.max(np.abs(raw)) > 5e-4
    assert np.max(np.abs(tvn)) < 0.06 * np.max(np.abs(raw))
    flat = MPASOperatorTables(grid, vertical_grid(10, 12000.0))
    assert not flat.orography
    out = MpasOracle(flat).slow_tendencies(*_rest_state(flat), 1.0)
    assert not np.any(out["tvn"])


def _zonal_wind_correction(n):
    """(wcf, exact, w, out) of the zonal wind a omega cos(lat) over the mountain at hexagon level n."""
    g = TRiSKGrid(create_hexagon_mesh(n))
    m = _mountain(g, levels=4)
    a, omega = m.radius, 1e-5
    vel = a * omega * np.stack([-g.edge_points[:, 1], g.edge_points[:, 0], np.zeros(g.num_edges())], axis=1)
    vn = np.repeat(np.sum(vel * g.edge_normals, axis=1)[None, :], m.nlev, axis=0)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((m.nlev + 1, g.num_cells()))
    rho = np.ones((m.nlev, g.num_cells()))
    out = MpasOracle(m).slow_tendencies(vn, w, rho, 300 * rho, 0.0)
    lon, lat = lonlat(g.centers)
    dl = 1e-6
    dhdlon = (dcmip_mountain_height(lon + dl, lat) - dcmip_mountain_height(lon - dl, lat)) / (2 * dl)
    zeta = np.asarray(m.vgrid.z_half)
    b = 1 - zeta / zeta[-1]
    exact = omega * dhdlon[None, :] * ((b[:-1] + b[1:]) / 2)[:, None]
    return out["wcf"], exact, w, out


def test_contravariant_correction_of_a_zonal_wind():
    """w_c = u . grad z of the full levels: for the zonal wind u = a omega cos(lat) (solid-body
    rotation about the axis) it is omega dz/dlon, which the cell inner product of vn dz/dn
    approximates to second order (relative L2 error 0.175 at hexagon 5, 0.046 at hexagon 6; with
    the sign of the slopes flipped it would be about 2); w_c vanishes at the lid and the vertical
    coupling takes w - w_c."""
    errs = []
    for n in (5, 6):
        wcf, exact, w, out = _zonal_wind_correction(n)
        errs.append(np.linalg.norm(wcf - exact) / np.linalg.norm(exact))
    assert errs[1] < 0.08 and errs[1] < errs[0] / 3, errs
    assert not np.any(out["wch"][-1])
    np.testing.assert_array_equal(out["what"], w - out["wch"])
