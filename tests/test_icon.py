"""The ICON triangle model on the host: its tables and the oracle.

Restates the reference's "ICON triangle shallow water" test set (test/finite_volumes.jl:
112-164) on the oracle, and pins the oracle to the TC2 height errors the reference
publishes for this model (docs/STATUS.md:40: L2 error 8.5e-5 on 1280 cells and 3.9e-6 on
82k cells, plain bisection meshes), to closed forms (the pressure gradient of a fluid at
rest, mass conservation) and to the discrete identity that makes the divergence damping
dissipative.
"""
import numpy as np
import pytest

from dmplex_amd import (ICONTables, TRiSKGrid, bilinear_weights, create_hexagon_mesh, create_spherical_mesh,
                        orthogonality_defect, rbf_coefficients)
from oracle.oracle import IconOracle

U0 = 0.2


def tc2_state(g):
    """Solid-body rotation in geostrophic balance (test/finite_volumes.jl:126-128)."""
    x = g.edge_points
    vel = U0 * np.stack([-x[:, 1], x[:, 0], np.zeros(len(x))], axis=1)
    u = np.sum(vel * g.edge_normals, axis=1)
    h = 1 - (U0 + U0 ** 2 / 2) * g.centers[:, 2] ** 2
    return np.concatenate([u, h])


def tc2_error(m, check_invariants=True):
    """L2 height error after t = 1 with dt = 0.2 min d_e (test/finite_volumes.jl:129-142)."""
    g = m.grid
    o = IconOracle(m)
    y = tc2_state(g)
    y0 = y.copy()
    mass0, energy0 = o.total_mass(y), o.total_energy(y)
    dt = 0.2 * g.dual_lengths.min()
    o.rk4_steps(y, dt, int(np.ceil(1 / dt)))
    if check_invariants:
        assert abs(o.total_mass(y) - mass0) < 1e-13
        assert abs(o.total_energy(y) - energy0) / energy0 < 1e-5
    dh = (y - y0)[g.num_edges():]
    return float(np.sqrt(np.sum(g.cell_areas * dh ** 2) / (4 * np.pi)))


def test_rbf_and_bilinear_weights():
    g = TRiSKGrid(create_spherical_mesh(3))  # circumcenters
    assert orthogonality_defect(g.dm, g.centers) < 1e-12
    E, W = rbf_coefficients(g)
    assert E.shape == (4, g.num_edges())
    omega = np.array([0.3, -0.2, 1.0])
    solid = np.einsum("ij,ij->i", np.cross(omega, g.edge_points), g.edge_normals)
    tangential = np.einsum("ij,ij->i", np.cross(omega, g.edge_points), np.cross(g.edge_points, g.edge_normals))
    assert np.max(np.abs(np.sum(W * solid[E], axis=0) - tangential)) < 0.01
    # exact for a constant vector along r x n_e at the edge itself, by the normalization
    t = np.cross(g.edge_points, g.edge_normals)
    exact = np.sum(W * np.einsum("kej,ej->ke", g.edge_normals[E], t), axis=0)
    np.testing.assert_allclose(exact, 1.0, rtol=1e-12)
    # the stencil is the other edges of the two cells
    for e in range(0, g.num_edges(), 97):
        cells = g.edge_cells[:, e]
        others = sorted(set(g.cell_edges[:, cells].ravel()) - {e})
        assert sorted(E[:, e]) == others
    B = bilinear_weights(g)
    np.testing.assert_allclose(B.sum(axis=0), 1.0, rtol=0, atol=1e-14)
    moment = np.einsum("ki,kij->ij", B, g.edge_points[g.cell_edges])
    assert np.max(np.linalg.norm(np.cross(g.centers, moment), axis=1)) < 1e-14


def test_tc2_published_errors():
    """The oracle reproduces the reference's published TC2 errors of this model."""
    e3 = tc2_error(ICONTables(create_spherical_mesh(3)))
    assert round(e3, 6) == pytest.approx(8.5e-5, abs=0.05e-5)
    e6 = tc2_error(ICONTables(create_spherical_mesh(6)), check_invariants=False)
    assert e6 == pytest.approx(3.9e-6, abs=0.05e-6)


def test_tc2_convergence_and_centered_gradient():
    e3 = tc2_error(ICONTables(create_spherical_mesh(3)))
    e4 = tc2_error(ICONTables(create_spherical_mesh(4)))
    assert e3 < 1e-4 and e4 < e3 / 3
    scvt = create_spherical_mesh(4, centroidal=True)
    e4i = tc2_error(ICONTables(scvt))
    e4c = tc2_error(ICONTables(scvt, kinetic_gradient="centered"))
    assert e4c < 2 * e4i
    e5 = tc2_error(ICONTables(create_spherical_mesh(5, centroidal=True)))
    assert e5 < e4i / 3.5  # second order on optimized meshes


def test_fluid_at_rest_and_mass():
    """u = 0: du_e = -g (h_r - h_l) / d_e exactly and dh = 0; any state conserves mass."""
    m = ICONTables(create_spherical_mesh(3), gravity=9.80616)
    g = m.grid
    rng = np.random.default_rng(2)
    h = 1 + 0.1 * rng.standard_normal(g.num_cells())
    y = np.concatenate([np.zeros(g.num_edges()), h])
    dy = IconOracle(m).tendency(y)
    l, r = g.edge_cells
    np.testing.assert_array_equal(dy[:g.num_edges()], -m.gravity * (h[r] - h[l]) * (1 / g.dual_lengths))
    assert not np.any(dy[g.num_edges():])
    y = np.concatenate([rng.standard_normal(g.num_edges()), h])
    dh = IconOracle(m).tendency(y)[g.num_edges():]
    assert abs(np.dot(g.cell_areas, dh)) < 1e-14 * np.dot(g.cell_areas, np.abs(dh))


def test_divergence_damping():
    """A uniform nu and the same nu per edge agree (test/finite_volumes.jl:160-163), and
    the damping term du = -nu (grad div)^2 u satisfies <u, du> = -nu |grad div u|^2 in
    the edge inner product sum l d u v, the discrete statement that it dissipates."""
    dm = create_spherical_mesh(3)
    g = TRiSKGrid(dm)

    def damped(nu):
        m = ICONTables(g, divergence_damping=nu)
        y = tc2_state(g)
        y[g.num_edges():] = 1 + 0.1 * g.centers[:, 0]
        return IconOracle(m).tendency(y), m, y

    a, m, y = damped(1e-4)
    b, _, _ = damped(np.full(g.num_edges(), 1e-4))
    c, _, _ = damped(0.0)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(g.num_edges())
    y = np.concatenate([u, np.ones(g.num_cells())])
    du = IconOracle(m).tendency(y)[:g.num_edges()] - IconOracle(ICONTables(g)).tendency(y)[:g.num_edges()]
    ce, l, r = g.cell_edges, g.edge_cells[0], g.edge_cells[1]
    div = np.sum(m.divergence * u[ce], axis=0)
    grad_div = (div[r] - div[l]) / g.dual_lengths
    w = g.primal_lengths * g.dual_lengths
    lhs, rhs = np.dot(w * u, du), -1e-4 * np.dot(w, grad_div ** 2)
    assert rhs < 0 and lhs == pytest.approx(rhs, rel=1e-9)


def test_table_errors():
    with pytest.raises(ValueError, match="triangle mesh"):
        ICONTables(create_hexagon_mesh(2))
    with pytest.raises(ValueError, match="kinetic_gradient"):
        ICONTables(create_spherical_mesh(1), kinetic_gradient="upwind")
    with pytest.raises(ValueError, match="one value per edge"):
        ICONTables(create_spherical_mesh(1), divergence_damping=np.ones(3))
    m = ICONTables(create_spherical_mesh(2))
    m.check_tables()
    bad = m.rbf_edges.copy(order="F")
    bad[2, 5] = m.grid.num_edges()
    m.rbf_edges = bad
    with pytest.raises(ValueError, match="out of range in rbf_edges"):
        m.check_tables()
