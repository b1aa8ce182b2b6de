"""The oracle (FP64 CPU restatement of the reference path), pinned by the reference's
own tests (test/finite_volumes.jl:79-110,184-229; test/shallow_water.jl:34-40) and by
dense operators assembled independently of the gather tables (brute force on small
meshes), so that a dropped term, a wrong sign or a transposed index fails."""
import numpy as np
import pytest

from dense_operators import operators
from dmplex_amd import ModelTables, RK4Cache, TRiSKGrid, create_hexagon_mesh, create_spherical_mesh, rk4_step_
from dmplex_amd.synthetic import earth_parameters, harmonic_coefficients, seeded_state, streamfunction, williamson2
from oracle.oracle import Oracle


def tc2_run(n, pv_flux):
    dm = create_hexagon_mesh(n)
    t = ModelTables(dm, pv_flux=pv_flux)
    o = Oracle(t)
    y = williamson2(t.grid)
    y0 = y.copy()
    rng = np.random.default_rng(3)
    z = y + 0.01 * rng.standard_normal(len(y))
    dz = o.tendency(z)
    eps = 1e-6
    dEdt = (o.total_energy(z + eps * dz) - o.total_energy(z - eps * dz)) / (2 * eps)
    mass0, energy0 = o.total_mass(y), o.total_energy(y)
    dt = 0.2 * np.sqrt(4 * np.pi / dm.nc)
    o.rk4_steps(y, dt, int(np.ceil(1 / dt)))
    dh = (y - y0)[dm.ne:]
    return dict(dEdt=abs(dEdt) / np.linalg.norm(dz), mass=abs(o.total_mass(y) - mass0),
                energy=abs(o.total_energy(y) - energy0) / energy0,
                l2=np.sqrt(np.sum(t.cell_areas * dh ** 2) / (4 * np.pi)))


@pytest.mark.parametrize("pv_flux", ["trisk", "gassmann", "gassmann_mean"])
def test_williamson2_conservation_and_convergence(pv_flux):
    r3, r4 = tc2_run(3, pv_flux), tc2_run(4, pv_flux)
    for r in (r3, r4):
        assert r["dEdt"] < 1e-8
        assert r["mass"] < 1e-13
        assert r["energy"] < 1e-8
    assert r3["l2"] < 2e-3
    assert r4["l2"] < r3["l2"] / 2.5


def test_vertex_kinetic_energy_orderings():
    """Cell / mixed / vertex kinetic energy (finite_volumes.jl:184-229) on the SCVT
    hexagon mesh of level 4 with the primal vertices as generators.  The reference runs
    this on its spring-dynamics mesh (not part of this path) and asserts a TC2 error ratio
    of 2.5; on the SCVT mesh the ratio is 2.37, asserted here as > 2."""
    tri = create_spherical_mesh(4, centroidal=True)
    g = TRiSKGrid(create_hexagon_mesh(4, primal=tri), centers=tri.coordinates)
    u0 = 0.2
    errors = []
    for ke in ("cell", "mixed", "vertex"):
        t = ModelTables(g, kinetic_energy=ke)
        o = Oracle(t)
        y = williamson2(g)
        y0 = y.copy()
        rng = np.random.default_rng(3)
        z = y + 0.01 * rng.standard_normal(len(y))
        dz = o.tendency(z)
        eps = 1e-6
        dEdt = (o.total_energy(z + eps * dz) - o.total_energy(z - eps * dz)) / (2 * eps)
        assert abs(dEdt) < 1e-8 * np.linalg.norm(dz)
        K, _ = o.kinetic_energy(y)
        vel2 = u0 ** 2 * (g.centers[:, 0] ** 2 + g.centers[:, 1] ** 2)
        kerror = np.max(np.abs(K - vel2 / 2)) / (u0 ** 2 / 2)
        dt = 0.2 * np.sqrt(4 * np.pi / g.num_cells())
        o.rk4_steps(y, dt, int(np.ceil(1 / dt)))
        errors.append((kerror, np.max(np.abs((y - y0)[g.num_edges():]))))
    assert errors[2][0] < errors[0][0] / 2
    assert errors[2][1] < errors[0][1] / 2
    assert errors[2][1] < errors[1][1] < errors[0][1]


def test_rk4_known_answer():
    for dt in (0.1, 0.05):
        y = np.array([1.0, 0.0])
        rk4_step_(y, dt, lambda dy, y: dy.__setitem__(slice(None), (y[1], -y[0])), RK4Cache(y))
        assert abs(y[0] - np.cos(dt)) < dt ** 5 / 100


def dense_tendency(t, y):
    """The TRiSK tendency from dense operators assembled from the geometry alone."""
    g = t.grid
    ne = g.num_edges()
    l, d = g.primal_lengths, g.dual_lengths
    left, right = g.edge_cells
    a, b = g.edge_vertices
    ops = operators(g)
    D, Kmat, Curl, Hv, W = (ops[k] for k in ("D", "K", "Curl", "Hv", "W"))
    u, h = y[:ne], y[ne:]
    K = Kmat @ (u * u)
    B = K + t.gravity * h
    zeta = Curl @ u
    hv = Hv @ h
    qv = (zeta + t.coriolis) / hv
    qe = (qv[a] + qv[b]) / 2
    F = (g.cell_distances[0] * h[left] + g.cell_distances[1] * h[right]) / d * u
    Wl = W * l[None, :] / d[:, None]
    du = 0.5 * (Wl @ F * qe + Wl @ (F * qe)) - (B[right] - B[left]) / d
    dh = -D @ F
    return np.concatenate([du, dh])


@pytest.mark.parametrize("n", [1, 2])
def test_oracle_matches_dense_operators(n):
    dm = create_hexagon_mesh(n, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **earth_parameters())
    y = seeded_state(t.grid)
    ref = dense_tendency(t, y)
    got = Oracle(t).tendency(y)
    ne = dm.ne
    for sl in (slice(0, ne), slice(ne, None)):
        assert np.max(np.abs(got[sl] - ref[sl])) < 1e-11 * np.max(np.abs(ref[sl]))


def test_dense_check_catches_a_planted_sign_error():
    """The dense check is sharp: flipping one gradient sign in a copy of the tables fails it."""
    dm = create_hexagon_mesh(2, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **earth_parameters())
    y = seeded_state(t.grid)
    ref = dense_tendency(t, y)
    bad = ModelTables(dm, centers=dm.generators, **earth_parameters())
    bad.edge_cells = np.asfortranarray(bad.edge_cells[::-1])  # left/right swapped: gradient sign flips
    got = Oracle(bad).tendency(y)
    assert np.max(np.abs(got[:dm.ne] - ref[:dm.ne])) > 1e-3 * np.max(np.abs(ref[:dm.ne]))


def test_oracle_independent_of_thread_count():
    dm = create_hexagon_mesh(4, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **earth_parameters())
    y = seeded_state(t.grid)
    o = Oracle(t)
    n0 = o.threads()
    try:
        o.set_threads(1)
        a = o.rk4_steps(y.copy(), 600.0, 3)
        ea = o.total_energy(a)
        o.set_threads(max(n0, 4))
        b = o.rk4_steps(y.copy(), 600.0, 3)
        eb = o.total_energy(b)
    finally:
        o.set_threads(n0)
    assert np.array_equal(a, b) and ea == eb


def test_seeded_state_recipe():
    a, omega, g = 6.371e6, 7.292e-5, 9.80616
    dm = create_hexagon_mesh(4, centroidal=True)
    grid = TRiSKGrid(dm, centers=dm.generators)
    y = seeded_state(grid, noise=0.0)
    ne = dm.ne
    u, h = y[:ne], y[ne:]
    assert np.isclose(np.max(np.abs(u)) * a, 30.0, rtol=1e-12)
    # discretely non-divergent
    D = np.sum(grid.cell_signs * grid.primal_lengths[grid.cell_edges] * u[grid.cell_edges], axis=0) / grid.cell_areas
    assert np.max(np.abs(D)) < 1e-12 * np.max(np.abs(u))
    assert np.all(h > 7000) and np.all(h < 9000)
    # the edge velocity is the normal component of r x grad(psi): compare with psi = z
    coef = np.zeros(80)
    coef[1] = 1.0  # Y_10, proportional to z
    psi_v = streamfunction(grid.vertices, coef)
    un = (psi_v[grid.edge_vertices[0]] - psi_v[grid.edge_vertices[1]]) / grid.primal_lengths
    c10 = np.sqrt(3 / (4 * np.pi))
    exact = c10 * np.sum(np.cross(grid.edge_points, [0, 0, 1.0]) * grid.edge_normals, axis=1)
    assert np.max(np.abs(un - exact)) < 1e-2 * np.max(np.abs(exact))
    # same seeds, same state; other seeds differ
    assert np.array_equal(seeded_state(grid), seeded_state(grid))
    assert not np.array_equal(seeded_state(grid), seeded_state(grid, seed_coef=3))
    assert harmonic_coefficients(1).shape == (80,)


def test_spherical_harmonics_orthonormal():
    dm = create_hexagon_mesh(5)
    grid = TRiSKGrid(dm)
    A = grid.cell_areas
    ys = []
    for idx in (0, 3, 7, 20, 79):
        coef = np.zeros(80)
        coef[idx] = 1.0
        ys.append(streamfunction(grid.centers, coef))
    G = np.array([[np.sum(A * p * q) for q in ys] for p in ys])
    assert np.max(np.abs(G - np.eye(len(ys)))) < 5e-3


# --- the oracle's adjoint (no reference counterpart: pinned by finite differences) -------------

ADJOINT_OPTIONS = [dict(), dict(kinetic_energy="mixed"), dict(kinetic_energy="vertex"),
                   dict(pv_flux="gassmann"), dict(pv_flux="gassmann", kinetic_energy="vertex"),
                   dict(pv_flux="gassmann_mean", kinetic_energy="mixed")]
ADJOINT_IDS = ["trisk-cell", "trisk-mixed", "trisk-vertex", "gassmann-cell", "gassmann-vertex", "gassmann_mean-mixed"]


def _adjoint_case(n=3, **opts):
    dm = create_hexagon_mesh(n, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **opts, **earth_parameters())
    y = seeded_state(t.grid)
    ne = dm.ne
    scale = np.concatenate([np.full(ne, np.std(y[:ne])), np.full(dm.nc, np.std(y[ne:]))])
    return t, y, ne, scale


@pytest.mark.parametrize("opts", ADJOINT_OPTIONS, ids=ADJOINT_IDS)
def test_oracle_tendency_vjp_dot_product(opts):
    """<w, J v> by central differences of the tendency equals <J^T w, v>, for every PV flux
    and kinetic-energy form."""
    t, y, ne, scale = _adjoint_case(**opts)
    o = Oracle(t)
    rng = np.random.default_rng(11)
    for _ in range(4):
        v = rng.standard_normal(len(y)) * scale
        w = rng.standard_normal(len(y)) / scale
        eps = 1e-4
        jv = (o.tendency(y + eps * v) - o.tendency(y - eps * v)) / (2 * eps)
        a, b = w @ jv, o.tendency_vjp(y, w) @ v
        assert abs(a - b) < 1e-8 * abs(b)


@pytest.mark.parametrize("opts", [ADJOINT_OPTIONS[0], ADJOINT_OPTIONS[4]], ids=[ADJOINT_IDS[0], ADJOINT_IDS[4]])
def test_oracle_rk4_adjoint_matches_finite_differences(opts):
    """Gradient of J = sum A (h_T - h*)^2 after 6 RK4 steps against central differences."""
    t, y, ne, scale = _adjoint_case(**opts)
    o = Oracle(t)
    dt, N = 1200.0, 6
    hstar = seeded_state(t.grid, seed_coef=3, seed_noise=4)[ne:]
    A = t.cell_areas

    def J(y0):
        yT = o.rk4_steps(y0.copy(), dt, N)
        return np.sum(A * (yT[ne:] - hstar) ** 2)

    yT = o.rk4_steps(y.copy(), dt, N)
    ybar = np.zeros_like(y)
    ybar[ne:] = 2 * A * (yT[ne:] - hstar)
    g = o.rk4_adjoint(y, dt, N, ybar)
    rng = np.random.default_rng(12)
    for _ in range(4):
        v = rng.standard_normal(len(y)) * scale
        eps = 1e-5
        fd = (J(y + eps * v) - J(y - eps * v)) / (2 * eps)
        assert abs(fd - g @ v) < 1e-7 * abs(g @ v)


def test_adjoint_check_is_sharp():
    """The dot-product check fails for the adjoint of a slightly different model (kinetic
    weights scaled by 1 + 1e-3), so a dropped or mis-weighted term cannot pass it."""
    t, y, ne, scale = _adjoint_case()
    o = Oracle(t)
    bad = ModelTables(t.grid, **earth_parameters())
    bad.kinetic = np.asfortranarray(bad.kinetic * (1 + 1e-3))
    ob = Oracle(bad)
    rng = np.random.default_rng(13)
    v = rng.standard_normal(len(y)) * scale
    w = rng.standard_normal(len(y)) / scale
    eps = 1e-4
    a = w @ ((o.tendency(y + eps * v) - o.tendency(y - eps * v)) / (2 * eps))
    assert abs(a - ob.tendency_vjp(y, w) @ v) > 1e-6 * abs(a)
    assert abs(a - o.tendency_vjp(y, w) @ v) < 1e-8 * abs(a)


@pytest.mark.parametrize("table", ["kites", "height", "rhombus", "weights"])
def test_adjoint_check_is_sharp_for_gassmann_and_vertex_ke(table):
    """The same check fails when one of the tables only the Gassmann flux or the vertex
    kinetic energy uses is perturbed by 1e-3 in the adjoint's model alone."""
    t, y, ne, scale = _adjoint_case(pv_flux="gassmann", kinetic_energy="vertex")
    o = Oracle(t)
    bad = ModelTables(t.grid, pv_flux="gassmann", kinetic_energy="vertex", **earth_parameters())
    tables = bad.gassmann if table in ("rhombus", "weights") else bad.vertex_kinetic
    tables[table] = np.asfortranarray(tables[table] * (1 + 1e-3))
    ob = Oracle(bad)
    rng = np.random.default_rng(14)
    v = rng.standard_normal(len(y)) * scale
    w = rng.standard_normal(len(y)) / scale
    eps = 1e-4
    a = w @ ((o.tendency(y + eps * v) - o.tendency(y - eps * v)) / (2 * eps))
    assert abs(a - ob.tendency_vjp(y, w) @ v) > 1e-6 * abs(a)
    assert abs(a - o.tendency_vjp(y, w) @ v) < 1e-8 * abs(a)
