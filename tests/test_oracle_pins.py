"""Independent pins of the oracle's diagnostics and RK4 driver.

* ``oracle_potential_enstrophy``, ``oracle_edge_sum`` (relative vorticity, TRiSK
  tangential velocity) and ``oracle_total_energy`` against dense operators assembled
  from the geometry (TRiSKShallowWater.jl:397-433), and against the reference's own
  solid-body tolerances (test/finite_volumes.jl:60-64);
* ``oracle_rk4_steps`` against the generic RK4 of the package (RungeKutta.jl:18-28, itself
  pinned by the harmonic-oscillator known answer) driven by the oracle's tendency, bit for
  bit, and by its temporal order (Richardson ratio 2^4 at a fixed mesh);
* the ICON oracle's enstrophy, vorticity and RBF tangential velocity likewise
  (ICONShallowWater.jl:209-224; test/finite_volumes.jl:115-120).

Each check is shown to be sharp: a planted mistake of the kind it guards against (a
missing 1/2, a transposed table stride, a swapped stage coefficient) fails it.
"""
import numpy as np
import pytest

from dense_operators import operators, potential_enstrophy, total_energy, vertex_pv
from dmplex_amd import ICONTables, ModelTables, RK4Cache, create_hexagon_mesh, create_spherical_mesh, rk4_step_
from dmplex_amd.synthetic import earth_parameters, seeded_state
from oracle.oracle import IconOracle, Oracle


def _model(n, **kw):
    dm = create_hexagon_mesh(n, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **earth_parameters(), **kw)
    return t, seeded_state(t.grid)


def _transposed_stride(table):
    """The k x n column-major table as a reader with a transposed stride sees it: entry
    (q, j) taken from memory position q n + j instead of q + j k."""
    k, n = table.shape
    return np.asfortranarray(table.ravel(order="F").reshape(k, n))


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("n", [1, 2])
def test_diagnostics_match_dense_operators(n):
    t, y = _model(n)
    g = t.grid
    ne = g.num_edges()
    o = Oracle(t)
    ops = operators(g)
    # relative vorticity = Curl u; tangential velocity = (1/d_e) sum_e' w l_e' u_e'
    assert _rel(o.relative_vorticity(y), ops["Curl"] @ y[:ne]) < 1e-12
    Wl = ops["W"] * g.primal_lengths[None, :] / g.dual_lengths[:, None]
    assert _rel(o.tangential_velocity(y[:ne]), Wl @ y[:ne]) < 1e-12
    qv, zeta, hv = o.vertex_pv(y)
    q_ref, z_ref, h_ref = vertex_pv(ops, y, ne, t.coriolis)
    assert _rel(qv, q_ref) < 1e-12 and _rel(zeta, z_ref) < 1e-12 and _rel(hv, h_ref) < 1e-12
    # compensated sums against plain dense sums of the same terms
    ens = potential_enstrophy(g, ops, y, t.coriolis)
    assert abs(o.potential_enstrophy(y) - ens) < 1e-12 * abs(ens)
    en = total_energy(g, ops, y, t.gravity)
    assert abs(o.total_energy(y) - en) < 1e-12 * abs(en)
    mass = float(np.sum(g.cell_areas * y[ne:]))
    assert abs(o.total_mass(y) - mass) < 1e-14 * abs(mass)


def test_enstrophy_closed_form_uniform_pv():
    """For a state whose PV is the same constant q0 at every vertex, the enstrophy is
    q0^2/2 sum_v A_v h_v exactly (sum_v A_v h_v = sum_i A_i h_i, the kites partition the
    cells).  Choose u = 0 and f_v = q0 h_v: then q_v = q0."""
    dm = create_hexagon_mesh(2, centroidal=True)
    q0 = 0.37
    h = 1.0 + 0.1 * dm.generators[:, 0] + 0.05 * dm.generators[:, 2] ** 2
    t = ModelTables(dm, centers=dm.generators)
    g = t.grid
    t.coriolis = q0 * (operators(g)["Hv"] @ h)
    y = np.concatenate([np.zeros(g.num_edges()), h])
    o = Oracle(t)
    qv, _, _ = o.vertex_pv(y)
    assert np.max(np.abs(qv - q0)) < 1e-14
    expect = q0 ** 2 / 2 * np.sum(g.cell_areas * h)
    assert abs(o.potential_enstrophy(y) - expect) < 1e-13 * expect


def test_solid_body_vorticity_and_tangential_velocity():
    """test/finite_volumes.jl:60-64 through the oracle's edge sums: solid-body rotation
    about z has vorticity 2 z and tangential velocity r x n_e . (z x r)."""
    dm = create_hexagon_mesh(3, centroidal=True)
    t = ModelTables(dm, centers=dm.generators)
    g = t.grid
    vel = np.stack([-g.edge_points[:, 1], g.edge_points[:, 0], np.zeros(g.num_edges())], axis=1)
    u = np.sum(vel * g.edge_normals, axis=1)
    y = np.concatenate([u, np.ones(g.num_cells())])
    o = Oracle(t)
    zeta = o.relative_vorticity(y)
    assert np.max(np.abs(zeta - 2 * g.vertices[:, 2])) < 0.03
    tang = np.sum(vel * np.cross(g.edge_points, g.edge_normals), axis=1)
    assert np.max(np.abs(o.tangential_velocity(u) - tang)) < 0.15


def test_planted_diagnostic_mistakes_fail_the_dense_checks():
    t, y = _model(2)
    g = t.grid
    ne = g.num_edges()
    o = Oracle(t)
    ops = operators(g)
    ens = potential_enstrophy(g, ops, y, t.coriolis)
    # a missing 1/2
    assert abs(2 * o.potential_enstrophy(y) - ens) > 1e-3 * abs(ens)
    # a transposed stride in edge_sum: the k x n table read as if it were stored n x k
    bad = o.edge_sum(_transposed_stride(t.vertex_edges), _transposed_stride(t.curl), y[:ne])
    assert _rel(bad, ops["Curl"] @ y[:ne]) > 1e-3
    # the sign of the tangent (vertex pair swapped)
    assert _rel(o.edge_sum(t.vertex_edges, -t.curl, y[:ne]), ops["Curl"] @ y[:ne]) > 1.0


def _generic_rk4(o, y, dt, nsteps, weights=(1, 2, 2, 1), third=1.0):
    """rk4_step_ of the package with the oracle's tendency; ``weights``/``third`` plant
    a wrong stage coefficient when they differ from the classical ones."""
    def tend(dy, x):
        dy[...] = o.tendency(x)

    if weights == (1, 2, 2, 1) and third == 1.0:
        cache = RK4Cache(y)
        for _ in range(nsteps):
            rk4_step_(y, dt, tend, cache)
        return y
    w1, w2, w3, w4 = weights
    for _ in range(nsteps):
        k1 = o.tendency(y)
        k2 = o.tendency(y + dt / 2 * k1)
        k3 = o.tendency(y + dt / 2 * k2)
        k4 = o.tendency(y + third * dt * k3)
        y[...] = y + dt / 6 * (w1 * k1 + w2 * k2 + w3 * k3 + w4 * k4)
    return y


@pytest.mark.parametrize("kw", [dict(), dict(pv_flux="gassmann", kinetic_energy="mixed")])
def test_oracle_rk4_equals_generic_rk4(kw):
    t, y0 = _model(3, **kw)
    o = Oracle(t)
    a = o.rk4_steps(y0.copy(), 900.0, 3)
    b = _generic_rk4(o, y0.copy(), 900.0, 3)
    assert np.array_equal(a, b)


def _richardson(step, y0, T, n):
    """(y_n - y_2n) / (y_2n - y_4n) of the integrator ``step(y, dt, nsteps)`` to time T."""
    ys = [step(y0.copy(), T / m, m) for m in (n, 2 * n, 4 * n)]
    return np.linalg.norm(ys[0] - ys[1]) / np.linalg.norm(ys[1] - ys[2])


def test_oracle_rk4_is_fourth_order_and_planted_coefficients_are_not():
    t, y0 = _model(2)
    o = Oracle(t)
    T = 3 * 3600.0  # 8, 16 and 32 steps: in the asymptotic range, above the rounding floor
    ratio = _richardson(lambda y, dt, m: o.rk4_steps(y, dt, m), y0, T, 8)
    assert 13.0 < ratio < 19.0
    for planted in (dict(weights=(1, 2, 1, 2)), dict(weights=(2, 1, 2, 1)), dict(third=0.5)):
        bad = _richardson(lambda y, dt, m: _generic_rk4(o, y, dt, m, **planted), y0, T, 8)
        assert bad < 4.0, planted


# --- ICON ---------------------------------------------------------------------------

def _icon(n=2):
    m = ICONTables(create_spherical_mesh(n), **earth_parameters())
    g = m.grid
    rng = np.random.default_rng(5)
    x = g.edge_points
    omega = np.array([0.3, -0.2, 1.0]) * 1e-5
    u = np.einsum("ij,ij->i", np.cross(omega, x), g.edge_normals) + 1e-7 * rng.standard_normal(len(x))
    h = 8000 + 20 * g.centers[:, 0] + rng.standard_normal(g.num_cells())
    return m, np.concatenate([u, h])


def test_icon_diagnostics_match_dense_operators():
    m, y = _icon()
    g = m.grid
    ne = g.num_edges()
    o = IconOracle(m)
    ops = operators(g)
    assert _rel(o.relative_vorticity(y), ops["Curl"] @ y[:ne]) < 1e-12
    R = np.zeros((ne, ne))
    for k in range(m.rbf_edges.shape[0]):
        np.add.at(R, (np.arange(ne), m.rbf_edges[k]), m.rbf[k])
    assert _rel(o.tangential_velocity(y[:ne]), R @ y[:ne]) < 1e-12
    qv, zeta, hv = o.vertex_pv(y)
    q_ref, z_ref, h_ref = vertex_pv(ops, y, ne, m.coriolis)
    assert _rel(qv, q_ref) < 1e-12 and _rel(hv, h_ref) < 1e-12
    ens = potential_enstrophy(g, ops, y, m.coriolis)
    assert abs(o.potential_enstrophy(y) - ens) < 1e-12 * abs(ens)
    assert abs(2 * o.potential_enstrophy(y) - ens) > 1e-3 * abs(ens)


def test_icon_rbf_tangential_velocity_solid_body():
    """test/finite_volumes.jl:115-120 through the ICON oracle's edge sum."""
    m = ICONTables(create_spherical_mesh(3))
    g = m.grid
    omega = np.array([0.3, -0.2, 1.0])
    solid = np.einsum("ij,ij->i", np.cross(omega, g.edge_points), g.edge_normals)
    tang = np.einsum("ij,ij->i", np.cross(omega, g.edge_points), np.cross(g.edge_points, g.edge_normals))
    o = IconOracle(m)
    assert np.max(np.abs(o.tangential_velocity(solid) - tang)) < 0.01
    # a transposed stride in the 4 x ne table fails it
    bad = Oracle.edge_sum(o, _transposed_stride(m.rbf_edges), _transposed_stride(m.rbf), solid)
    assert np.max(np.abs(bad - tang)) > 0.1
