"""Parity of the HIP path (through the C ABI of libtrisk_hip.so) with the FP64 oracle.

The FP64 kernels evaluate every formula in the reference's order without FMA
contraction, so tendencies, RK4 states and the gathered diagnostics must agree with
the oracle bit for bit; the compensated sums agree to a few ulps.  Sizes are the small
meshes the oracle finishes in seconds; test_gpu_full_size.py covers levels 9 and 10.
"""
import sys

import numpy as np
import pytest

from dmplex_amd import (HIPBackend, ModelTables, TRiSKShallowWater, create_hexagon_mesh, diagnostics,
                        gather_state, kinetic_energy_, relative_vorticity, rk4_steps_, tangential_velocity,
                        tendency_, vertex_pv_)
from dmplex_amd.synthetic import earth_parameters, seeded_state, williamson2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

_MESHES = {}


def mesh(n):
    if n not in _MESHES:
        _MESHES[n] = create_hexagon_mesh(n, centroidal=True)
    return _MESHES[n]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("pv_flux", ["trisk", "gassmann", "gassmann_mean"])
@pytest.mark.parametrize("kinetic_energy", ["cell", "mixed", "vertex"])
def test_tendency_bitwise(pv_flux, kinetic_energy):
    say(f"tendency parity level 4 {pv_flux} {kinetic_energy}")
    dm = mesh(4)
    m = TRiSKShallowWater(dm, centers=dm.generators, pv_flux=pv_flux, kinetic_energy=kinetic_energy,
                          **earth_parameters())
    y = seeded_state(m.grid)
    ref = Oracle(m).tendency(y)
    yd = m.upload_state(y)
    dy = m.zeros()
    tendency_(dy, yd, m)
    got = gather_state(m, dy)
    assert np.array_equal(got, ref), rel(got, ref)


@pytest.mark.parametrize("pv_flux", ["trisk", "gassmann", "gassmann_mean"])
def test_rk4_bitwise(pv_flux):
    say(f"rk4 parity level 5 {pv_flux}")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, pv_flux=pv_flux, **earth_parameters())
    y = seeded_state(m.grid)
    o = Oracle(m)
    ref = o.rk4_steps(y.copy(), 600.0, 12)
    yd = m.upload_state(y)
    rk4_steps_(yd, 600.0, 12, m)
    got = gather_state(m, yd)
    assert np.array_equal(got, ref), rel(got, ref)


def _update_bytes(m):
    import ctypes as C

    from dmplex_amd import _device as D
    names = (C.c_char_p * 8)()
    by = np.zeros(8)
    n = C.c_int32()
    D.check(D.lib().trisk_kernel_bytes(m.handle, 8, names, by.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
    return dict(zip([names[k].decode() for k in range(n.value)], by))["update_kernel"]


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_gassmann_slot_and_table_forms(precision):
    """Gassmann tables in the slot order of gassmann_coefficients run without the PV-edge
    table (each slot's PV comes from the pair gathered for its inner edge); the same tables
    with the slots reversed fall back to reading it.  Both agree with the oracle on the
    tables they were given: bit for bit in FP64, to FP32 rounding in FP32-bulk."""
    say(f"gassmann slot and table forms level 4 {precision}")
    from dmplex_amd import _device as D
    dm = mesh(4)
    m = TRiSKShallowWater(dm, centers=dm.generators, pv_flux="gassmann", backend=HIPBackend(0, precision),
                          **earth_parameters())
    y = seeded_state(m.grid)
    results = []
    for form in ("slot", "table"):
        if form == "table":
            slot_bytes = _update_bytes(m)
            for key in ("edges", "pv_edges", "weights"):
                m.gassmann[key] = np.asfortranarray(m.gassmann[key][::-1])  # k x n column-major
            D.check(D.lib().trisk_destroy(m.handle))
            m._handle = None
            m._create()
            # the PV-edge table is read again: 12 indices of 2 bytes per edge and tendency
            assert _update_bytes(m) == slot_bytes + 24 * m.grid.num_edges()
        ref = Oracle(m).tendency(y)
        dy = m.zeros()
        tendency_(dy, m.upload_state(y), m)
        got = gather_state(m, dy)
        if precision == "fp64":
            assert np.array_equal(got, ref), (form, rel(got, ref))
        else:
            ne = m.grid.num_edges()
            # FP32 stencils: u to a few 1e-7, h (a divergence with cancellation) to a few 1e-5
            assert rel(got[:ne], ref[:ne]) < 2e-6 and rel(got[ne:], ref[ne:]) < 1e-4, form
        results.append(got)
    if precision == "fp32_bulk":
        ne = m.grid.num_edges()
        assert not np.array_equal(results[0][:ne], results[1][:ne])  # another summation order
        assert rel(results[1][:ne], results[0][:ne]) < 2e-6
        assert np.array_equal(results[0][ne:], results[1][ne:])  # dh does not use the PV flux


@pytest.mark.parametrize("pv_flux", ["gassmann", "gassmann_mean"])
@pytest.mark.parametrize("kinetic_energy", ["mixed", "vertex"])
def test_fp32_bulk_gassmann_vertex_ke(pv_flux, kinetic_energy):
    """FP32-bulk stencils (FP64 state, heights, RK4 accumulation) with the Gassmann PV flux
    and mixed or vertex kinetic energy (TRiSKShallowWater.jl:212-244,266-296) against the
    FP64 oracle on the SCVT mesh of level 5: the tendency to FP32 rounding with
    sum A dh = 0, then 12 RK4 steps of 600 s with the mass kept exactly."""
    say(f"fp32-bulk {pv_flux} {kinetic_energy} level 5")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, pv_flux=pv_flux, kinetic_energy=kinetic_energy,
                          backend=HIPBackend(0, "fp32_bulk"), **earth_parameters())
    y = seeded_state(m.grid)
    o = Oracle(m)
    ref = o.tendency(y)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    got = gather_state(m, dy)
    ne = dm.ne
    eu, eh = rel(got[:ne], ref[:ne]), rel(got[ne:], ref[ne:])
    say(f"  tendency: u {eu:.2e}, h {eh:.2e}")
    # FP32 stencils: du to a few 1e-7; dh, a divergence that cancels to ~1e-2 of its terms, to ~1e-5
    assert eu < 2e-6 and eh < 1e-4
    A = m.cell_areas
    assert abs(np.dot(A, got[ne:])) < 1e-12 * np.dot(A, np.abs(got[ne:]))
    yd = m.upload_state(y)
    mass0 = diagnostics(m, yd)[0]
    rk4_steps_(yd, 600.0, 12, m)
    got = gather_state(m, yd)
    ref = o.rk4_steps(y.copy(), 600.0, 12)
    eu, eh = rel(got[:ne], ref[:ne]), rel(got[ne:], ref[ne:])
    say(f"  12 steps: u {eu:.2e}, h {eh:.2e}")
    assert eu < 1e-6 and eh < 1e-7
    assert abs(diagnostics(m, yd)[0] - mass0) <= 1e-14 * mass0


def test_level5_one_day_144_steps():
    """Level 5, TRiSK, cell KE, dt = 600 s, 144 steps (one day)."""
    say("level 5, 144 steps of 600 s")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, **earth_parameters())
    y0 = seeded_state(m.grid)
    o = Oracle(m)
    ref = o.rk4_steps(y0.copy(), 600.0, 144)
    yd = m.upload_state(y0)
    mass0, energy0, _ = diagnostics(m, yd)
    rk4_steps_(yd, 600.0, 144, m)
    got = gather_state(m, yd)
    ne = dm.ne
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(got[sl], ref[sl]) < 1e-12
    assert np.array_equal(got, ref)
    mass, energy, _ = diagnostics(m, yd)
    assert abs(mass - mass0) / mass0 < 1e-14
    assert abs(mass - o.total_mass(ref)) / mass0 < 1e-15
    assert abs(energy - o.total_energy(ref)) / energy0 < 1e-14
    # the energy drift equals the oracle's
    drift_gpu = (energy - energy0) / energy0
    drift_ref = (o.total_energy(ref) - o.total_energy(y0)) / o.total_energy(y0)
    assert abs(drift_gpu - drift_ref) < 1e-14


def test_diagnostics_match_oracle():
    say("diagnostics level 4")
    dm = mesh(4)
    m = TRiSKShallowWater(dm, centers=dm.generators, kinetic_energy="mixed", **earth_parameters())
    y = seeded_state(m.grid)
    o = Oracle(m)
    mass, energy, ens = diagnostics(m, y)
    assert abs(mass - o.total_mass(y)) <= 4e-16 * abs(mass)
    assert abs(energy - o.total_energy(y)) <= 4e-16 * abs(energy)
    assert abs(ens - o.potential_enstrophy(y)) <= 4e-16 * abs(ens)
    K, B = kinetic_energy_(m, y).K, kinetic_energy_(m, y).B
    Kr, Br = o.kinetic_energy(y)
    assert np.array_equal(K, Kr) and np.array_equal(B, Br)
    q, z, h = vertex_pv_(m, y)
    qr, zr, hr = o.vertex_pv(y)
    assert np.array_equal(q, qr) and np.array_equal(z, zr) and np.array_equal(h, hr)


def test_vorticity_and_tangential_velocity():
    say("relative vorticity / tangential velocity level 3")
    dm = mesh(3)
    m = TRiSKShallowWater(dm, centers=dm.generators)
    y = williamson2(m.grid)
    o = Oracle(m)
    assert np.array_equal(relative_vorticity(m, y), o.relative_vorticity(y))
    u = y[:dm.ne]
    assert np.array_equal(tangential_velocity(m, u), o.tangential_velocity(u))


def test_hilbert_layout_is_a_permutation_and_invisible():
    say("layout permutation level 4")
    dm = mesh(4)
    m = TRiSKShallowWater(dm, centers=dm.generators)
    c, e, v = m.permutation()
    for p, n in ((c, dm.nc), (e, dm.ne), (v, dm.nv)):
        assert np.array_equal(np.sort(p), np.arange(n))
    plain = TRiSKShallowWater(dm, centers=dm.generators, hilbert=False)
    y = williamson2(m.grid)
    a, b = m.zeros(), plain.zeros()
    tendency_(a, m.upload_state(y), m)
    tendency_(b, plain.upload_state(y), plain)
    assert np.array_equal(gather_state(m, a), gather_state(plain, b))


def test_fp32_bulk_close_to_fp64():
    say("fp32-bulk tendency level 5")
    dm = mesh(5)
    kw = dict(centers=dm.generators, **earth_parameters())
    m32 = TRiSKShallowWater(dm, backend=HIPBackend(precision="fp32_bulk"), **kw)
    y = seeded_state(m32.grid)
    ref = Oracle(m32).tendency(y)
    dy = m32.zeros()
    tendency_(dy, m32.upload_state(y), m32)
    got = gather_state(m32, dy)
    ne = dm.ne
    # FP32 stencils: the divergence and the PV flux cancel to ~1e-2 of their terms
    assert rel(got[:ne], ref[:ne]) < 1e-3
    assert rel(got[ne:], ref[ne:]) < 2e-4
    # mass: the continuity stays FP64, so sum A dh telescopes
    A = m32.cell_areas
    assert abs(np.dot(A, got[ne:])) < 1e-12 * np.dot(A, np.abs(got[ne:]))


@pytest.mark.parametrize("level", [3, 5])
def test_pair_path_fp64_matches_oracle(level):
    """The cell-pair formulation (15 antisymmetric weights per cell) equals the reference
    tendency up to rounding: 1e-12 relative per field, norm-wise."""
    say(f"pair path fp64 level {level}")
    dm = mesh(level)
    m = TRiSKShallowWater(dm, centers=dm.generators, path="pair", **earth_parameters())
    assert m.path_info()[0]
    y = seeded_state(m.grid)
    o = Oracle(m)
    ref = o.tendency(y)
    dy = m.zeros()
    tendency_(dy, m.upload_state(y), m)
    got = gather_state(m, dy)
    ne = dm.ne
    assert rel(got[:ne], ref[:ne]) < 1e-12 and rel(got[ne:], ref[ne:]) < 1e-12
    yd = m.upload_state(y)
    rk4_steps_(yd, 600.0, 20, m)
    ref = o.rk4_steps(y.copy(), 600.0, 20)
    got = gather_state(m, yd)
    assert rel(got[:ne], ref[:ne]) < 1e-12 and rel(got[ne:], ref[ne:]) < 1e-12
    mass0 = o.total_mass(y)
    assert abs(diagnostics(m, yd)[0] - mass0) / mass0 < 1e-14


def test_fp32_narrow_and_wide_indices_agree():
    """16-bit index deltas and the int32 fallback give the same FP32-bulk tendency; a low
    delta limit forces a mix of narrow and wide entities on a small mesh."""
    say("narrow vs wide indices level 5")
    dm = mesh(5)
    kw = dict(centers=dm.generators, backend=HIPBackend(precision="fp32_bulk"), **earth_parameters())
    narrow = TRiSKShallowWater(dm, **kw)
    mixed = TRiSKShallowWater(dm, narrow_limit=60, **kw)
    wc, wv, we = mixed.wide_counts()
    assert 0 < wc < dm.nc and 0 < wv < dm.nv and 0 < we < dm.ne
    assert narrow.wide_counts() == (0, 0, 0)
    y = seeded_state(narrow.grid)
    a, b = narrow.zeros(), mixed.zeros()
    tendency_(a, narrow.upload_state(y), narrow)
    tendency_(b, mixed.upload_state(y), mixed)
    assert np.array_equal(gather_state(narrow, a), gather_state(mixed, b))
    ya, yb = narrow.upload_state(y), mixed.upload_state(y)
    rk4_steps_(ya, 600.0, 4, narrow)
    rk4_steps_(yb, 600.0, 4, mixed)
    assert np.array_equal(gather_state(narrow, ya), gather_state(mixed, yb))


def test_pair_path_fp32_close_to_tables():
    say("pair vs tables fp32 level 5")
    dm = mesh(5)
    kw = dict(centers=dm.generators, backend=HIPBackend(precision="fp32_bulk"), **earth_parameters())
    fast = TRiSKShallowWater(dm, **kw)
    slow = TRiSKShallowWater(dm, path="tables", **kw)
    assert fast.path_info()[0] and not slow.path_info()[0]
    y = seeded_state(fast.grid)
    a, b = fast.zeros(), slow.zeros()
    tendency_(a, fast.upload_state(y), fast)
    tendency_(b, slow.upload_state(y), slow)
    ga, gb = gather_state(fast, a), gather_state(slow, b)
    ref = Oracle(fast).tendency(y)
    ne = dm.ne
    for g in (ga, gb):
        assert rel(g[:ne], ref[:ne]) < 1e-3 and rel(g[ne:], ref[ne:]) < 2e-4
    A = fast.cell_areas
    assert abs(np.dot(A, ga[ne:])) < 1e-12 * np.dot(A, np.abs(ga[ne:]))


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_pair_rk4_call_boundaries_bitwise(precision):
    """On the cell-pair path each RK4 stage's edge kernel forms the mass flux G of the state it
    writes, and the first stage of a call forms it in the state pass instead.  Twelve one-step
    calls (G always from the state pass at stage 1) and one twelve-step call (G from the
    previous step's last stage) give the same bits."""
    say(f"pair rk4 call boundaries {precision}")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, path="pair", backend=HIPBackend(precision=precision),
                          **earth_parameters())
    y0 = seeded_state(m.grid)
    a, b = m.upload_state(y0), m.upload_state(y0)
    for _ in range(12):
        rk4_steps_(a, 600.0, 1, m)
    rk4_steps_(b, 600.0, 12, m)
    assert np.array_equal(gather_state(m, a), gather_state(m, b))


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_pair_plane_equals_separate_arrays(precision):
    """RK4 stages that keep G_e and the stage velocity u_e as one (G, u) pair per edge give the
    bits of the same stages with G and the stage vector in two arrays, over one call of 13
    steps and 13 calls of one step (each call's first stage forms the pairs from the state)."""
    say(f"pair plane {precision}")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, path="pair", backend=HIPBackend(precision=precision),
                          **earth_parameters())
    y0 = seeded_state(m.grid)
    out = {}
    for on in (False, True):
        m.use_pair_plane(on)
        a, b = m.upload_state(y0), m.upload_state(y0)
        rk4_steps_(a, 600.0, 13, m)
        for _ in range(13):
            rk4_steps_(b, 600.0, 1, m)
        out[on] = gather_state(m, a), gather_state(m, b)
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][0])
    assert np.array_equal(out[False][1], out[False][0])


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_graph_replay_equals_direct_launches(precision):
    """Graph replay (captured once, replayed with eager work between replays) gives the
    same bits as direct launches."""
    import torch

    say(f"graph replay {precision}")
    dm = mesh(5)
    m = TRiSKShallowWater(dm, centers=dm.generators, backend=HIPBackend(precision=precision), **earth_parameters())
    y0 = seeded_state(m.grid)
    a = m.upload_state(y0)
    rk4_steps_(a, 600.0, 24, m)
    ref = gather_state(m, a)
    m.use_graphs(True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b = m.upload_state(y0)
        for _ in range(3):
            rk4_steps_(b, 600.0, 8, m)
            torch.cuda.current_stream().synchronize()
            junk = torch.ones(1000, device="cuda").sum()  # eager work between replays
            assert junk.item() == 1000.0
        torch.cuda.current_stream().synchronize()
    assert np.array_equal(gather_state(m, b), ref)


@pytest.mark.parametrize("model", ["trisk", "icon"])
def test_state_check_reports_nan_and_zero_vertex_height(model):
    """Failure detection: a NaN planted in u, an Inf in h and a zero-height cell (whose
    vertex heights h_v, the divisor of q_v at TRiSKShallowWater.jl:258, then vanish) are
    counted by check_state; with the finite check on, diagnostics raises; a clean state
    passes.  The conditions are set from the host; every kernel runs to its end."""
    from dmplex_amd import ICONShallowWater, create_spherical_mesh
    from dmplex_amd._device import NonFiniteStateError

    say(f"state check {model}")
    if model == "trisk":
        dm = mesh(3)
        m = TRiSKShallowWater(dm, centers=dm.generators, **earth_parameters())
    else:
        m = ICONShallowWater(create_spherical_mesh(3), **earth_parameters())
    y = seeded_state(m.grid)
    ne, nc = m.grid.num_edges(), m.grid.num_cells()
    assert m.check_state(y) == (0, 0, 0)
    m.set_finite_check(True)
    diagnostics(m, m.upload_state(y))  # a finite state passes
    bad = y.copy()
    bad[5] = np.nan
    bad[ne + 7] = np.inf
    bad[ne + 8] = -np.inf
    nu, nh, nhv = m.check_state(bad)
    assert (nu, nh) == (1, 2) and nhv > 0  # h_v of the vertices around the infinite cells
    with pytest.raises(NonFiniteStateError, match="1 non-finite u, 2 non-finite h"):
        diagnostics(m, m.upload_state(bad))
    assert "non-finite" in m.last_error()
    dry = y.copy()
    dry[ne:] = 0.0  # no fluid: every h_v is zero
    assert m.check_state(dry) == (0, 0, m.grid.num_vertices())
    with pytest.raises(NonFiniteStateError, match="vertices with h_v"):
        diagnostics(m, m.upload_state(dry))
    m.set_finite_check(False)
    mass, _, _ = diagnostics(m, m.upload_state(bad))  # the reference's behaviour: NaN sums
    assert not np.isfinite(mass)
    assert nc > 0


@pytest.mark.parametrize("pv_flux,kinetic_energy", [("trisk", "cell"), ("gassmann", "vertex")])
def test_julia_extension_descriptor_tendency(pv_flux, kinetic_energy):
    """A handle created from the descriptor the Julia extension builds (1-based struct
    tables, 0-based copies, Gassmann and vertex-KE tables passed) gives the oracle's FP64
    tendency bit for bit."""
    import ctypes as C

    import torch

    from dmplex_amd import _device as D
    from julia_descriptor import descriptor, julia_struct

    say(f"julia descriptor {pv_flux} {kinetic_energy}")
    dm = mesh(4)
    t = ModelTables(dm, centers=dm.generators, pv_flux=pv_flux, kinetic_energy=kinetic_energy, **earth_parameters())
    desc, keep = descriptor(julia_struct(t))
    h = C.c_void_p()
    D.check(D.lib().trisk_create(C.byref(desc), C.byref(D.TriskOptions(D.TRISK_FP64, 0, 0, 0, 0)), C.byref(h)))
    try:
        y = seeded_state(t.grid)
        n = len(y)
        yd = torch.zeros(n, dtype=torch.float64, device="cuda")
        dy = torch.zeros_like(yd)
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        f64p = C.POINTER(C.c_double)
        D.check(D.lib().trisk_upload_state(h, y.ctypes.data_as(f64p), C.c_void_p(yd.data_ptr()), s))
        D.check(D.lib().trisk_tendency(h, C.c_void_p(yd.data_ptr()), C.c_void_p(dy.data_ptr()), s))
        got = np.empty(n)
        D.check(D.lib().trisk_download_state(h, C.c_void_p(dy.data_ptr()), got.ctypes.data_as(f64p), s))
    finally:
        D.check(D.lib().trisk_destroy(h))
    assert np.array_equal(got, Oracle(t).tendency(y))


def test_invalid_tables_are_rejected():
    say("invalid tables")
    dm = mesh(2)
    t = ModelTables(dm, centers=dm.generators)
    bad = t.cell_edges.copy(order="F")
    bad[0, 0] = dm.ne + 5
    t.cell_edges = bad
    m = TRiSKShallowWater.__new__(TRiSKShallowWater)
    m.__dict__.update(t.__dict__)
    m.backend, m._hilbert, m._handle = HIPBackend(), True, None
    with pytest.raises(ValueError, match="out of range"):
        m._create()


def test_device_lloyd_matches_host_lloyd():
    """Lloyd on the GPU converges to the host SCVT (same fixed point) at level 5."""
    from dmplex_amd import lloyd_residual
    from dmplex_amd.grids import _icosahedral

    say("device Lloyd level 5")
    xh, tri, its_h = _icosahedral(5, True)
    xd, tri_d, its_d = _icosahedral(5, True, device=0)
    assert np.array_equal(tri, tri_d)
    assert np.array_equal(xh[:12], xd[:12])
    assert np.max(np.linalg.norm(xh - xd, axis=1)) < 1e-9
    assert lloyd_residual(xd, tri, fixed=12) < 1e-8 * 0.06
    assert np.all(np.abs(its_h - its_d) <= 3)
