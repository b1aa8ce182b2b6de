"""The step at the benchmark's full sizes: SCVT level 9 (2,621,442 cells, TRiSK PV flux, cell
kinetic energy, dt = 45 s) and level 10 (10,485,762 cells, Gassmann PV flux, vertex kinetic
energy); the ICON model on triangle level 9, the MPAS-Gassmann stage
operators on hexagon 7 x 30 levels and the 64-member ensemble on levels 4-7, the sizes they are
measured at.  The oracle finishes one tendency in about a second at these
sizes, so the FP64 path is checked bit for bit against it; longer runs are checked through
size-independent properties: mass conservation, agreement between the FP64 table and cell-pair
paths, the FP32-bulk error, and bitwise determinism of direct launches against graph replay."""
import gc
import sys

import numpy as np
import pytest

from dmplex_amd import (HIPBackend, TRiSKGrid, TRiSKShallowWater, create_hexagon_mesh, diagnostics, gather_state,
                        rk4_steps_, tendency_)
from dmplex_amd.synthetic import earth_parameters, seeded_state
from oracle.oracle import Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def level9():
    say("level 9 SCVT mesh (Lloyd on the GPU)")
    dm = create_hexagon_mesh(9, centroidal=True, device=0)
    grid = TRiSKGrid(dm, centers=dm.generators)
    return grid, seeded_state(grid)


def test_level9_fp64_bitwise_with_oracle(level9):
    grid, y = level9
    say("level 9 FP64: tendency and 2 RK4 steps against the oracle")
    m = TRiSKShallowWater(grid, **earth_parameters())
    o = Oracle(m)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    assert np.array_equal(gather_state(m, dy), o.tendency(y))
    yd = m.upload_state(y)
    rk4_steps_(yd, 45.0, 2, m)
    assert np.array_equal(gather_state(m, yd), o.rk4_steps(y.copy(), 45.0, 2))


def test_level9_paths_conservation_and_determinism(level9):
    grid, y = level9
    steps = 20
    say(f"level 9: {steps} steps on the FP64 table and pair paths and in FP32-bulk")
    out = {}
    for key, precision, path in (("fp64", "fp64", "tables"), ("fp64_pair", "fp64", "pair"),
                                 ("fp32_bulk", "fp32_bulk", "auto")):
        m = TRiSKShallowWater(grid, backend=HIPBackend(0, precision), path=path, **earth_parameters())
        yd = m.upload_state(y)
        mass0 = diagnostics(m, yd)[0]
        rk4_steps_(yd, 45.0, steps, m)
        mass = diagnostics(m, yd)[0]
        assert abs(mass - mass0) <= 1e-14 * mass0, key
        out[key] = m.download_state(yd)
        if key == "fp32_bulk":
            # the same steps replayed from captured graphs give the same bits
            m.use_graphs(True)
            yg = m.upload_state(y)
            rk4_steps_(yg, 45.0, steps, m)
            assert np.array_equal(m.download_state(yg), out[key])
        del m
    ne = grid.num_edges()
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(out["fp64_pair"][sl], out["fp64"][sl]) < 1e-12
    assert rel(out["fp32_bulk"][ne:], out["fp64"][ne:]) < 1e-6
    assert rel(out["fp32_bulk"][:ne], out["fp64"][:ne]) < 1e-4


def test_level10_gassmann_vertex_ke():
    """Config 2's mesh and method: create_hexagon_mesh(10; centroidal=true) (10,485,762
    cells; Lloyd on the GPU), Gassmann PV flux and vertex kinetic energy.  FP64: one tendency
    and two RK4 steps of 22.5 s bit for bit with the oracle.  FP32-bulk: the same two steps to
    FP32 rounding, with the mass kept exactly."""
    say("level 10 SCVT mesh (Lloyd on the GPU)")
    dm = create_hexagon_mesh(10, centroidal=True, device=0)
    grid = TRiSKGrid(dm, centers=dm.generators)
    del dm
    y = seeded_state(grid)
    say("level 10: Gassmann PV flux and vertex kinetic energy, FP64 tendency and 2 RK4 steps")
    kw = dict(pv_flux="gassmann", kinetic_energy="vertex", **earth_parameters())
    m = TRiSKShallowWater(grid, **kw)
    o = Oracle(m)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    assert np.array_equal(gather_state(m, dy), o.tendency(y))
    del dy
    yd = m.upload_state(y)
    rk4_steps_(yd, 22.5, 2, m)
    ref = o.rk4_steps(y.copy(), 22.5, 2)
    assert np.array_equal(gather_state(m, yd), ref)
    del m, yd, o
    gc.collect()
    say("level 10: FP32-bulk, 2 RK4 steps")
    m32 = TRiSKShallowWater(grid, backend=HIPBackend(0, "fp32_bulk"), **kw)
    yd = m32.upload_state(y)
    mass0 = diagnostics(m32, yd)[0]
    rk4_steps_(yd, 22.5, 2, m32)
    got = gather_state(m32, yd)
    ne = grid.num_edges()
    eu, eh = rel(got[:ne], ref[:ne]), rel(got[ne:], ref[ne:])
    say(f"  FP32-bulk after 2 steps: u {eu:.2e}, h {eh:.2e}")
    assert eu < 2e-6 and eh < 2e-7
    assert abs(diagnostics(m32, yd)[0] - mass0) <= 1e-15 * mass0
    del m32, yd, ref, grid
    gc.collect()


def test_icon_level9_bitwise_with_oracle():
    """ICONShallowWater on the SCVT triangle mesh of level 9 (5,242,880 cells), with ICON's
    divergence damping: one tendency and two RK4 steps of 20 s bit for bit in FP64, and the
    two steps to FP32 rounding with exact mass in FP32-bulk."""
    from dmplex_amd import ICONShallowWater, create_spherical_mesh
    from oracle.oracle import IconOracle

    say("ICON level 9 triangles: tendency and 2 RK4 steps against the oracle")
    dm = create_spherical_mesh(9, centroidal=True, device=0)
    dt = 20.0
    nu = 0.0025 * (4 * np.pi / dm.nc) ** 2 / dt
    m = ICONShallowWater(dm, divergence_damping=nu, **earth_parameters())
    y = seeded_state(m.grid)
    o = IconOracle(m)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    assert np.array_equal(gather_state(m, dy), o.tendency(y))
    yd = m.upload_state(y)
    rk4_steps_(yd, dt, 2, m)
    ref2 = o.rk4_steps(y.copy(), dt, 2)
    assert np.array_equal(gather_state(m, yd), ref2)
    # FP32-bulk: two steps to FP32 rounding, mass exactly
    m32 = ICONShallowWater(m.grid, divergence_damping=nu, backend=HIPBackend(0, "fp32_bulk"), **earth_parameters())
    yd = m32.upload_state(y)
    mass0 = diagnostics(m32, yd)[0]
    rk4_steps_(yd, dt, 2, m32)
    got = gather_state(m32, yd)
    ne = m.grid.num_edges()
    assert rel(got[:ne], ref2[:ne]) < 2e-6 and rel(got[ne:], ref2[ne:]) < 2e-7
    assert abs(diagnostics(m32, yd)[0] - mass0) <= 1e-15 * mass0


def test_mpas_hexagon7_30_levels_with_oracle():
    """The MPAS-Gassmann stage tendencies on SCVT hexagon 7 x 30 levels (the measured
    workload): every field bit for bit, the Exner pressure to 1 ulp."""
    from dmplex_amd.mpas_ops import FIELDS, OROGRAPHY_FIELDS, MPASOperators, vertical_grid
    from oracle.oracle import MpasOracle

    say("MPAS hexagon 7 x 30 levels: one slow_tendencies call against the oracle")
    dm = create_hexagon_mesh(7, centroidal=True, device=0)
    g = TRiSKGrid(dm, centers=dm.generators)
    v = vertical_grid(30, 30000.0)
    m = MPASOperators(g, v)
    rng = np.random.default_rng(5)
    nl, nc, ne = 30, g.num_cells(), g.num_edges()
    rho = np.exp(-v.z_full / 8000.0)[:, None] * (1 + 0.05 * rng.random((nl, nc)))
    rt = rho * 300.0 * np.exp(v.z_full / 20000.0)[:, None]
    w = 0.5 * rng.standard_normal((nl + 1, nc))
    w[0] = w[-1] = 0
    vn = 20 * rng.standard_normal((nl, ne))
    ref = MpasOracle(m).slow_tendencies(vn, w, rho, rt, 0.0)
    m.slow_tendencies(*m.upload_state(vn, w, rho, rt), pressure=0.0)
    for f in FIELDS:
        if (f == "Kv" and m.vertex_kinetic is None) or (f == "eta" and m.vertical == 0) or f in OROGRAPHY_FIELDS:
            continue
        got = m.field(f)
        if f == "exner":
            np.testing.assert_allclose(got, ref[f], rtol=5e-16, atol=0, err_msg=f)
        else:
            assert np.array_equal(got, ref[f]), f


def test_ensemble_64_members_levels_4_to_7():
    """The measured ensemble: 16 members on each of SCVT levels 4-7, FP32-bulk, per-level dt
    1200/600/300/150 s, under the autotuned schedule with graph replay.  After 8 steps every
    member equals a single model of its level bit for bit and has kept its mass."""
    from dmplex_amd import ModelTables, TRiSKEnsemble

    say("ensemble: 64 members on levels 4-7, autotuned, 8 steps against single models")
    levels, dts, M = (4, 5, 6, 7), [1200.0, 600.0, 300.0, 150.0], 16
    tabs = []
    for n in levels:
        dm = create_hexagon_mesh(n, centroidal=True, device=0)
        tabs.append(ModelTables(dm, centers=dm.generators, **earth_parameters()))
    ens = TRiSKEnsemble(tabs, [M] * 4, backend=HIPBackend(0, "fp32_bulk"))
    y = ens.zeros()

    def state(li, m):
        return seeded_state(tabs[li].grid, seed_coef=1000 + M * li + m, seed_noise=2000 + M * li + m)

    for li in range(4):
        for m in range(M):
            ens.set_member(y, li, m, state(li, m))
    ens.autotune(y, dts)
    ens.use_graphs(True)
    mass0 = ens.diagnostics(y)[:, 0]
    ens.rk4_steps_(y, dts, 8)
    mass = ens.diagnostics(y)[:, 0]
    assert np.all(np.abs(mass - mass0) <= 1e-14 * mass0)
    for li in range(4):
        s = TRiSKShallowWater(tabs[li].grid, path="pair", backend=HIPBackend(0, "fp32_bulk"), **earth_parameters())
        for m in range(M):
            yd = s.upload_state(state(li, m))
            rk4_steps_(yd, dts[li], 8, s)
            assert np.array_equal(ens.get_member(y, li, m), s.download_state(yd)), (levels[li], m)
