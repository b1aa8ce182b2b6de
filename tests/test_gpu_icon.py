"""Parity of the ICON triangle model (icon_create and the generic C ABI) with the FP64
oracle.  The FP64 kernels keep every formula of ICONShallowWater.jl in the reference's
order without contraction, so tendencies and RK4 states must agree bit for bit; the
diagnostics are compensated sums and agree to a few ulps.  FP32-bulk agrees to FP32
rounding and conserves mass exactly."""
import sys

import numpy as np
import pytest

from dmplex_amd import (HIPBackend, ICONShallowWater, create_spherical_mesh, diagnostics, gather_state,
                        initial_state, relative_vorticity, rk4_step_, rk4_steps_, RK4Cache, tangential_velocity,
                        tendency_, tendency_vjp, vertex_pv_)
from oracle.oracle import IconOracle

pytestmark = pytest.mark.gpu
_MESHES = {}


def mesh(n, centroidal=True):
    if (n, centroidal) not in _MESHES:
        _MESHES[n, centroidal] = create_spherical_mesh(n, centroidal=centroidal)
    return _MESHES[n, centroidal]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def icon_damping(m, dt):
    """ICON's fourth-order damping coefficient nu = divdamp_fac A^2 / dt (divdamp_fac 0.0025)."""
    return 0.0025 * (4 * np.pi / m.grid.num_cells()) ** 2 / dt


def seeded(m, seed=0):
    """Solid-body flow with a random perturbation of u and h (every term of the tendency active)."""
    g = m.grid
    rng = np.random.default_rng(seed)
    x = g.edge_points
    vel = 0.2 * np.stack([-x[:, 1], x[:, 0], np.zeros(len(x))], axis=1)
    u = np.sum(vel * g.edge_normals, axis=1) + 0.01 * rng.standard_normal(g.num_edges())
    h = 1 - 0.22 * g.centers[:, 2] ** 2 + 0.01 * rng.standard_normal(g.num_cells())
    return np.concatenate([u, h])


@pytest.mark.parametrize("level,centroidal", [(3, False), (4, True)])
@pytest.mark.parametrize("gradient", ["icon", "centered"])
@pytest.mark.parametrize("damping", [0.0, 1e-4])
def test_tendency_bitwise(level, centroidal, gradient, damping):
    say(f"icon tendency level {level} centroidal={centroidal} {gradient} damping={damping}")
    m = ICONShallowWater(mesh(level, centroidal), gravity=9.80616, kinetic_gradient=gradient,
                         divergence_damping=damping)
    y = seeded(m)
    ref = IconOracle(m).tendency(y)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    got = gather_state(m, dy)
    assert np.array_equal(got, ref), rel(got, ref)


@pytest.mark.parametrize("damping", [False, True])
def test_rk4_bitwise(damping):
    say(f"icon rk4 level 4 damping={damping}")
    plain = ICONShallowWater(mesh(4))
    dt = 0.2 * plain.grid.dual_lengths.min()
    m = ICONShallowWater(plain.grid, divergence_damping=icon_damping(plain, dt)) if damping else plain
    y = seeded(m, 1)
    ref = IconOracle(m).rk4_steps(y.copy(), dt, 5)
    yd = m.upload_state(y)
    rk4_steps_(yd, dt, 5, m)
    got = gather_state(m, yd)
    assert np.all(np.isfinite(ref)) and np.array_equal(got, ref), rel(got, ref)
    # rk4_step! with tendency! dispatches to the fused step
    yd = m.upload_state(y)
    for _ in range(5):
        rk4_step_(yd, dt, tendency_, RK4Cache(yd), m)
    assert np.array_equal(gather_state(m, yd), ref)


def test_graph_replay_and_diagnostics():
    say("icon graphs + diagnostics level 5")
    plain = ICONShallowWater(mesh(5))
    dt = 0.2 * plain.grid.dual_lengths.min()
    m = ICONShallowWater(plain.grid, divergence_damping=icon_damping(plain, dt))
    del plain
    y = seeded(m, 2)
    direct = m.upload_state(y)
    rk4_steps_(direct, dt, 19, m)
    m.use_graphs(True)
    import torch

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        graphed = m.upload_state(y)
        rk4_steps_(graphed, dt, 11, m)
        rk4_steps_(graphed, dt, 8, m)  # a replay of the graph captured by the first call
    s.synchronize()
    m.use_graphs(False)
    assert torch.isfinite(direct).all()
    assert torch.equal(direct, graphed)
    o = IconOracle(m)
    yh = gather_state(m, direct)
    mass, energy, enstrophy = diagnostics(m, direct)
    assert mass == pytest.approx(o.total_mass(yh), rel=1e-14)
    assert energy == pytest.approx(o.total_energy(yh), rel=1e-13)
    assert enstrophy == pytest.approx(o.potential_enstrophy(yh), rel=1e-13)
    qv, zv, hv = (m.download_field(2, a) for a in vertex_pv_(m, direct))
    rq, rz, rh = o.vertex_pv(yh)
    assert np.array_equal(qv, rq) and np.array_equal(zv, rz) and np.array_equal(hv, rh)
    assert np.array_equal(m.download_field(2, relative_vorticity(m, direct)), o.relative_vorticity(yh))
    ne = m.grid.num_edges()
    assert np.array_equal(tangential_velocity(m, yh[:ne]), o.tangential_velocity(yh[:ne]))


def test_tc2_on_the_gpu():
    """The reference's steady-state test on the device: errors as the oracle's (8.5e-5 at
    1280 cells), mass to round-off, energy to 1e-5."""
    say("icon tc2 level 3 on the device")
    m = ICONShallowWater(create_spherical_mesh(3))
    g = m.grid
    x = g.edge_points
    u = np.sum(0.2 * np.stack([-x[:, 1], x[:, 0], np.zeros(len(x))], axis=1) * g.edge_normals, axis=1)
    h = 1 - (0.2 + 0.02) * g.centers[:, 2] ** 2
    y = initial_state(m, lambda p: 0.2 * np.stack([-p[:, 1], p[:, 0], np.zeros(len(p))], axis=1),
                      lambda p: 1 - (0.2 + 0.02) * p[:, 2] ** 2)
    np.testing.assert_array_equal(gather_state(m, y), np.concatenate([u, h]))
    mass0, energy0, _ = diagnostics(m, y)
    dt = 0.2 * g.dual_lengths.min()
    rk4_steps_(y, dt, int(np.ceil(1 / dt)), m)
    mass, energy, _ = diagnostics(m, y)
    assert abs(mass - mass0) < 1e-13 and abs(energy - energy0) / energy0 < 1e-5
    dh = gather_state(m, y)[g.num_edges():] - h
    err = float(np.sqrt(np.sum(g.cell_areas * dh ** 2) / (4 * np.pi)))
    assert err == pytest.approx(8.5e-5, abs=0.05e-5)


def test_trisk_only_calls_and_options():
    m = ICONShallowWater(mesh(2))
    y = m.upload_state(seeded(m))
    with pytest.raises(ValueError, match="ICON"):
        tendency_vjp(m, y, y)
    assert m.wide_counts() == (0, 0, 0)


def test_narrow_and_wide_indices_agree():
    """The 16-bit index deltas and the int32 rows they fall back to give the same bits: a
    model whose deltas are limited to 40 reads many rows from the int32 tables."""
    say("icon narrow and wide indices, level 4")
    kw = dict(gravity=9.80616, divergence_damping=1e-4)
    narrow = ICONShallowWater(mesh(4), **kw)
    wide = ICONShallowWater(mesh(4), narrow_limit=40, **kw)
    assert narrow.wide_counts() == (0, 0, 0)
    assert all(n > 0 for n in wide.wide_counts())
    y = seeded(narrow)
    ref = IconOracle(narrow).tendency(y)
    for m in (narrow, wide):
        dy = m.zeros()
        tendency_(dy, m.upload_state(y), m)
        assert np.array_equal(gather_state(m, dy), ref)


@pytest.mark.parametrize("damping", [False, True])
def test_fp32_bulk_close_to_the_oracle(damping):
    """FP32-bulk (FP32 stencils; FP64 state, heights, pressure gradient, continuity and RK4
    accumulation) against the FP64 oracle on the seeded Earth state, level 4: the tendency
    (the height tendency is a small residual of the flux divergence), 20 RK4 steps, exact
    mass conservation, and graph replay bitwise equal to direct launches."""
    say(f"icon fp32_bulk level 4 damping={damping}")
    from dmplex_amd.synthetic import earth_parameters, seeded_state

    dt = 640.0
    nu = 0.0025 * (4 * np.pi / mesh(4).nc) ** 2 / dt if damping else 0.0
    m64 = ICONShallowWater(mesh(4), divergence_damping=nu, **earth_parameters())
    m = ICONShallowWater(m64.grid, divergence_damping=nu, backend=HIPBackend(0, "fp32_bulk"), **earth_parameters())
    y = seeded_state(m.grid)
    o = IconOracle(m64)
    ne = m.grid.num_edges()
    ref = o.tendency(y)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    got = gather_state(m, dy)
    # measured: u 3e-7, h 5e-5
    assert rel(got[:ne], ref[:ne]) < 2e-6 and rel(got[ne:], ref[ne:]) < 5e-4
    ref = o.rk4_steps(y.copy(), dt, 20)
    yd = m.upload_state(y)
    mass0 = diagnostics(m, yd)[0]
    rk4_steps_(yd, dt, 20, m)
    got = gather_state(m, yd)
    # measured after 50 steps: u 2e-7, h 2e-8
    assert rel(got[:ne], ref[:ne]) < 2e-6 and rel(got[ne:], ref[ne:]) < 2e-7
    assert abs(diagnostics(m, yd)[0] - mass0) <= 1e-15 * mass0
    m.use_graphs(True)
    yg = m.upload_state(y)
    rk4_steps_(yg, dt, 20, m)
    assert bool((yg == yd).all())


def test_fp32_bulk_tc2():
    """The reference's steady-state test in FP32-bulk: the same 8.5e-5 error as FP64 at
    1280 cells, mass to round-off."""
    say("icon tc2 level 3 fp32_bulk")
    m = ICONShallowWater(create_spherical_mesh(3), backend=HIPBackend(0, "fp32_bulk"))
    g = m.grid
    y = initial_state(m, lambda p: 0.2 * np.stack([-p[:, 1], p[:, 0], np.zeros(len(p))], axis=1),
                      lambda p: 1 - (0.2 + 0.02) * p[:, 2] ** 2)
    h = 1 - (0.2 + 0.02) * g.centers[:, 2] ** 2
    mass0, energy0, _ = diagnostics(m, y)
    dt = 0.2 * g.dual_lengths.min()
    rk4_steps_(y, dt, int(np.ceil(1 / dt)), m)
    mass, energy, _ = diagnostics(m, y)
    assert abs(mass - mass0) < 1e-13 and abs(energy - energy0) / energy0 < 1e-5
    dh = gather_state(m, y)[g.num_edges():] - h
    err = float(np.sqrt(np.sum(g.cell_areas * dh ** 2) / (4 * np.pi)))
    assert err == pytest.approx(8.5e-5, abs=0.05e-5)
