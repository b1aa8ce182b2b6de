"""The hand-written adjoint (trisk_tendency_vjp, trisk_rk4_adjoint) through the C ABI.

The reference has no adjoint and Julia AD is not available here, so the gradient is pinned
by (1) the oracle's independent scatter-add adjoint (itself pinned by finite differences on
the CPU), (2) central finite differences of the device forward model in random directions,
and (3) the dot-product test <w, J v> = <J^T w, v>.  The level-7 workload (163,842 cells,
40 RK4 steps of 150 s, checkpoints every 8 steps) is checked against 32 finite-difference
directions with eps = 1e-6.  The table path (Gassmann fluxes, mixed and vertex kinetic
energy) is checked against the oracle for six option sets in both precisions and, with the
Gassmann flux and vertex kinetic energy at level 6, against 16 finite-difference directions.
"""
import sys

import numpy as np
import pytest

from dmplex_amd import (HIPBackend, TRiSKShallowWater, cell_areas_device, create_hexagon_mesh, height_misfit,
                        rk4_adjoint, rk4_steps, rk4_steps_, rrule_rk4_steps, tendency_, tendency_vjp)
from dmplex_amd.synthetic import earth_parameters, seeded_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

_MESHES = {}


def mesh(n):
    if n not in _MESHES:
        _MESHES[n] = create_hexagon_mesh(n, centroidal=True, device=0)
    return _MESHES[n]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def model(n, precision="fp64"):
    dm = mesh(n)
    return TRiSKShallowWater(dm, centers=dm.generators, path="pair", backend=HIPBackend(0, precision),
                             **earth_parameters())


def scales(m, y):
    ne = m.grid.num_edges()
    return np.concatenate([np.full(ne, np.std(y[:ne])), np.full(m.grid.num_cells(), np.std(y[ne:]))])


def test_tendency_vjp_matches_oracle():
    say("tendency vjp vs oracle, level 4")
    m = model(4)
    y = seeded_state(m.grid)
    rng = np.random.default_rng(1)
    w = rng.standard_normal(len(y)) / scales(m, y)
    ref = Oracle(m).tendency_vjp(y, w)
    got = m.download_state(tendency_vjp(m, m.upload_state(y), m.upload_state(w)))
    ne = m.grid.num_edges()
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(got[sl], ref[sl]) < 1e-11


def test_tendency_vjp_dot_product_on_device():
    say("tendency vjp dot-product test, level 5")
    m = model(5)
    y = seeded_state(m.grid)
    s = scales(m, y)
    rng = np.random.default_rng(2)
    yd = m.upload_state(y)
    for _ in range(3):
        v = rng.standard_normal(len(y)) * s
        w = rng.standard_normal(len(y)) / s
        eps = 1e-4
        fp, fm = m.zeros(), m.zeros()
        tendency_(fp, m.upload_state(y + eps * v), m)
        tendency_(fm, m.upload_state(y - eps * v), m)
        a = float(((fp - fm) / (2 * eps)) @ m.upload_state(w))
        b = float(tendency_vjp(m, yd, m.upload_state(w)) @ m.upload_state(v))
        assert abs(a - b) < 1e-8 * abs(b)


def test_rk4_adjoint_matches_oracle():
    say("rk4 adjoint vs oracle, level 4, 10 steps")
    m = model(4)
    y = seeded_state(m.grid)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(len(y)) / scales(m, y)
    ref = Oracle(m).rk4_adjoint(y, 600.0, 10, lam)
    g, yT = rk4_adjoint(m, m.upload_state(y), 600.0, 10, m.upload_state(lam), checkpoint_every=4)
    got = m.download_state(g)
    ne = m.grid.num_edges()
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(got[sl], ref[sl]) < 1e-10
    yd = m.upload_state(y)
    rk4_steps_(yd, 600.0, 10, m)
    assert bool((yT == yd).all())


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_checkpointing_does_not_change_the_gradient(precision):
    say(f"checkpoint invariance {precision}")
    m = model(4, precision)
    y = m.upload_state(seeded_state(m.grid))
    lam = m.upload_state(np.random.default_rng(4).standard_normal(m.state_length) * 1e-3)
    gs = [rk4_adjoint(m, y, 600.0, 13, lam, checkpoint_every=c)[0] for c in (1, 3, 8, 13, 40)]
    for g in gs[1:]:
        assert bool((g == gs[0]).all())


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_graph_replayed_sweep_equals_direct(precision):
    """With graphs on, the captured sweep is replayed; new contents of the same buffers give
    the new gradient, bitwise as the direct launches."""
    import torch

    say(f"adjoint sweep graph replay {precision}")
    m = model(4, precision)
    y = m.upload_state(seeded_state(m.grid))
    rng = np.random.default_rng(6)
    lams = [m.upload_state(rng.standard_normal(m.state_length) * 1e-3) for _ in range(3)]
    ref = [rk4_adjoint(m, y, 600.0, 11, lam, checkpoint_every=4)[0] for lam in lams]
    m.use_graphs(True)
    lam, out, yT = m.zeros(), m.zeros(), m.zeros()
    from dmplex_amd import _device as D
    import ctypes as C

    for k in (0, 1, 2, 1):
        lam.copy_(lams[k])
        torch.cuda.synchronize()
        D.check(D.lib().trisk_rk4_adjoint(m.handle, C.c_void_p(y.data_ptr()), 600.0, 11, 4,
                                          C.c_void_p(lam.data_ptr()), C.c_void_p(out.data_ptr()),
                                          C.c_void_p(yT.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        assert bool((out == ref[k]).all())


def level7_misfit(precision):
    m = model(7, precision)
    y0 = seeded_state(m.grid)
    hstar = m.upload_state(seeded_state(m.grid, seed_coef=3, seed_noise=4))[m.grid.num_edges():].clone()
    return m, y0, hstar


def test_level7_gradient_matches_finite_differences():
    """Level 7, 40 RK4 steps of 150 s, FP64: dJ/dy0 against 32 central differences."""
    say("adjoint: level 7, 40 steps, 32 FD directions")
    m, y0, hstar = level7_misfit("fp64")
    dt, N = 150.0, 40
    yd = m.upload_state(y0)
    yT, pullback = rrule_rk4_steps(m, yd, dt, N, checkpoint_every=8)
    J0, ybar = height_misfit(m, yT, hstar)
    g = pullback(ybar)
    s = m.upload_state(scales(m, y0))
    import torch

    gen = torch.Generator(device="cuda").manual_seed(5)
    worst = 0.0
    for k in range(32):
        v = torch.randn(m.state_length, dtype=torch.float64, device="cuda", generator=gen) * s
        eps = 1e-6
        Jp = height_misfit(m, rk4_steps(m, yd + eps * v, dt, N), hstar)[0]
        Jm = height_misfit(m, rk4_steps(m, yd - eps * v, dt, N), hstar)[0]
        fd = (Jp - Jm) / (2 * eps)
        ad = float(g @ v)
        worst = max(worst, abs(fd - ad) / abs(ad))
    say(f"worst relative FD difference {worst:.2e}")
    # central differences at eps = 1e-6 carry a rounding error of about 1e-16 |J| / (eps |dJ.v|)
    # (measured worst case over the 32 directions: 8e-7)
    assert worst < 5e-6


def test_fp32_bulk_gradient_close_to_fp64():
    say("adjoint: level 7 FP32-bulk gradient vs FP64")
    grads = {}
    for p in ("fp64", "fp32_bulk"):
        m, y0, hstar = level7_misfit(p)
        yT, pullback = rrule_rk4_steps(m, m.upload_state(y0), 150.0, 40)
        grads[p] = m.download_state(pullback(height_misfit(m, yT, hstar)[1]))
    ne = m.grid.num_edges()
    for sl in (slice(0, ne), slice(ne, None)):
        a, b = grads["fp32_bulk"][sl], grads["fp64"][sl]
        err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        say(f"relative L2 difference {err:.2e}")
        assert err < 1e-5  # measured: 6e-7 (u), 1e-7 (h)


def test_torch_autograd_equals_pullback():
    import torch

    say("torch autograd through rk4_steps")
    m = model(4)
    hstar = m.upload_state(seeded_state(m.grid, seed_coef=3, seed_noise=4))[m.grid.num_edges():].clone()
    y0 = m.upload_state(seeded_state(m.grid)).requires_grad_(True)
    yT = rk4_steps(m, y0, 600.0, 5)
    A = cell_areas_device(m)
    J = torch.sum(A * (yT[m.grid.num_edges():] - hstar) ** 2)
    J.backward()
    _, ybar = height_misfit(m, yT.detach(), hstar)
    ref, _ = rk4_adjoint(m, y0.detach(), 600.0, 5, ybar)
    assert torch.allclose(y0.grad, ref, rtol=1e-12, atol=0)


def test_adjoint_error_paths():
    say("adjoint error paths")
    dm = mesh(2)
    m = TRiSKShallowWater(dm, centers=dm.generators, **earth_parameters())
    y = m.upload_state(seeded_state(m.grid))
    from dmplex_amd import _device as D
    import ctypes as C

    with pytest.raises(ValueError, match="alias"):
        D.check(D.lib().trisk_tendency_vjp(m.handle, C.c_void_p(y.data_ptr()), C.c_void_p(y.data_ptr()),
                                           C.c_void_p(y.data_ptr()), None))
    with pytest.raises(ValueError, match="nsteps"):
        rk4_adjoint(m, y, 600.0, -1, y.clone())


# --- the table path: every PV flux and kinetic-energy form ---------------------------------

TABLE_OPTIONS = [("trisk", "cell"), ("trisk", "vertex"), ("gassmann", "cell"), ("gassmann", "vertex"),
                 ("gassmann", "mixed"), ("gassmann_mean", "mixed")]


def table_model(n, pv_flux, kinetic_energy, precision="fp64"):
    dm = mesh(n)
    return TRiSKShallowWater(dm, centers=dm.generators, pv_flux=pv_flux, kinetic_energy=kinetic_energy,
                             path="tables", backend=HIPBackend(0, precision), **earth_parameters())


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("pv_flux,kinetic_energy", TABLE_OPTIONS)
def test_table_tendency_vjp_matches_oracle(pv_flux, kinetic_energy, precision):
    """The table kernels' pullback (TRiSKShallowWater.jl:202-305) against the oracle's
    scatter-add adjoint, itself pinned by finite differences on the CPU: FP64 to rounding,
    FP32-bulk (FP32 stencils and cotangents) to FP32 rounding."""
    say(f"table-path tendency vjp vs oracle, level 4, {pv_flux} {kinetic_energy} {precision}")
    m = table_model(4, pv_flux, kinetic_energy, precision)
    assert not m.path_info()[0]
    y = seeded_state(m.grid)
    rng = np.random.default_rng(21)
    w = rng.standard_normal(len(y)) / scales(m, y)
    ref = Oracle(m).tendency_vjp(y, w)
    got = m.download_state(tendency_vjp(m, m.upload_state(y), m.upload_state(w)))
    ne = m.grid.num_edges()
    tol = 1e-11 if precision == "fp64" else 2e-5
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(got[sl], ref[sl]) < tol, (sl, rel(got[sl], ref[sl]))


@pytest.mark.parametrize("pv_flux,kinetic_energy", [("gassmann", "vertex"), ("trisk", "mixed")])
def test_table_rk4_adjoint_matches_oracle(pv_flux, kinetic_energy):
    say(f"table-path rk4 adjoint vs oracle, level 4, 10 steps, {pv_flux} {kinetic_energy}")
    m = table_model(4, pv_flux, kinetic_energy)
    y = seeded_state(m.grid)
    rng = np.random.default_rng(22)
    lam = rng.standard_normal(len(y)) / scales(m, y)
    ref = Oracle(m).rk4_adjoint(y, 600.0, 10, lam)
    g, yT = rk4_adjoint(m, m.upload_state(y), 600.0, 10, m.upload_state(lam), checkpoint_every=4)
    got = m.download_state(g)
    ne = m.grid.num_edges()
    for sl in (slice(0, ne), slice(ne, None)):
        assert rel(got[sl], ref[sl]) < 1e-10
    yd = m.upload_state(y)
    rk4_steps_(yd, 600.0, 10, m)
    assert bool((yT == yd).all())


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_table_checkpointing_does_not_change_the_gradient(precision):
    say(f"table-path checkpoint invariance {precision}")
    m = table_model(4, "gassmann", "vertex", precision)
    y = m.upload_state(seeded_state(m.grid))
    lam = m.upload_state(np.random.default_rng(23).standard_normal(m.state_length) * 1e-3)
    gs = [rk4_adjoint(m, y, 600.0, 9, lam, checkpoint_every=c)[0] for c in (1, 4, 9)]
    for g in gs[1:]:
        assert bool((g == gs[0]).all())


def test_gassmann_vertex_ke_gradient_matches_finite_differences():
    """Config 2's method (Gassmann PV flux, vertex kinetic energy) at level 6, 20 RK4 steps of
    300 s, FP64: dJ/dy0 of the height misfit against 16 central differences, and the
    FP32-bulk gradient against the FP64 one."""
    import torch

    say("table-path adjoint: level 6 gassmann vertex, 20 steps, 16 FD directions")
    grads = {}
    for precision in ("fp64", "fp32_bulk"):
        m = table_model(6, "gassmann", "vertex", precision)
        y0 = seeded_state(m.grid)
        hstar = m.upload_state(seeded_state(m.grid, seed_coef=3, seed_noise=4))[m.grid.num_edges():].clone()
        yd = m.upload_state(y0)
        yT, pullback = rrule_rk4_steps(m, yd, 300.0, 20, checkpoint_every=5)
        grads[precision] = g = pullback(height_misfit(m, yT, hstar)[1])
        if precision != "fp64":
            continue
        s = m.upload_state(scales(m, y0))
        gen = torch.Generator(device="cuda").manual_seed(24)
        worst = 0.0
        for _ in range(16):
            v = torch.randn(m.state_length, dtype=torch.float64, device="cuda", generator=gen) * s
            eps = 1e-6
            Jp = height_misfit(m, rk4_steps(m, yd + eps * v, 300.0, 20), hstar)[0]
            Jm = height_misfit(m, rk4_steps(m, yd - eps * v, 300.0, 20), hstar)[0]
            ad = float(g @ v)
            worst = max(worst, abs((Jp - Jm) / (2 * eps) - ad) / abs(ad))
        say(f"  worst relative FD difference {worst:.2e}")
        assert worst < 5e-6
    a, b = grads["fp32_bulk"].cpu().numpy(), grads["fp64"].cpu().numpy()
    err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    say(f"  FP32-bulk vs FP64 gradient: relative L2 {err:.2e}")
    assert err < 1e-4
