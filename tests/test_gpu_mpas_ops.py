"""Parity of the MPAS-Gassmann stage operators (mpas_ops_* of libtrisk_hip.so) with the FP64
oracle, on flat ground and over orography.  Every field equals the oracle's bit for bit except the Exner pressure, where the
device's pow and the host's may differ in the last bit, and the vn tendency with the
pressure gradient, which inherits that difference through (pi_r - pi_l)."""
import sys

import numpy as np
import pytest

from dmplex_amd import TRiSKGrid, create_hexagon_mesh
from dmplex_amd.mpas_ops import FIELDS, OROGRAPHY_FIELDS, MPASOperators, vertical_grid
from oracle.oracle import MpasOracle

pytestmark = pytest.mark.gpu
_GRIDS = {}


def grid(n):
    if n not in _GRIDS:
        _GRIDS[n] = TRiSKGrid(create_hexagon_mesh(n, centroidal=True))
    return _GRIDS[n]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state(g, v, seed):
    rng = np.random.default_rng(seed)
    nl, nc, ne = len(v.dz), g.num_cells(), g.num_edges()
    rho = np.exp(-v.z_full / 8000.0)[:, None] * (1 + 0.05 * rng.random((nl, nc)))
    rt = rho * 300.0 * np.exp(v.z_full / 20000.0)[:, None] * (1 + 0.01 * rng.random((nl, nc)))
    vn = 20 * rng.standard_normal((nl, ne))
    w = 0.5 * rng.standard_normal((nl + 1, nc))
    w[0] = 0
    w[-1] = 0
    return vn, w, rho, rt


OPTIONS = {
    "iap": dict(),
    "trisk-vi-tridiagonal": dict(pv_flux="trisk", vertical_advection="vector_invariant", vertical_solver="tridiagonal",
                                 trisk_weights="exact"),
    "advective": dict(vertical_advection="advective", vertical_solver="tridiagonal"),
    "mixed-ke": dict(kinetic_energy="mixed", vertical_solver="tridiagonal"),
    "mixed-ke-five-band": dict(kinetic_energy="mixed"),
}


@pytest.mark.parametrize("level,nlev", [(3, 7), (4, 30)])
@pytest.mark.parametrize("name", list(OPTIONS))
@pytest.mark.parametrize("pressure", [0.0, 1.0])
def test_slow_tendencies_match_oracle(level, nlev, name, pressure):
    say(f"mpas slow tendencies hexagon {level} x {nlev} levels, {name}, pressure {pressure}")
    g = grid(level)
    v = vertical_grid(nlev, 30000.0, stretch=3.0)
    m = MPASOperators(g, v, **OPTIONS[name])
    vn, w, rho, rt = state(g, v, level * 100 + nlev)
    ref = MpasOracle(m).slow_tendencies(vn, w, rho, rt, pressure)
    m.slow_tendencies(*m.upload_state(vn, w, rho, rt), pressure=pressure)
    for f in FIELDS:
        if f == "Kv" and m.vertex_kinetic is None:
            continue
        if f == "eta" and m.vertical == 0:
            continue
        if f in OROGRAPHY_FIELDS:  # flat ground: not allocated
            continue
        got = m.field(f)
        if f == "exner":
            np.testing.assert_allclose(got, ref[f], rtol=5e-16, atol=0, err_msg=f)
        elif f == "tvn" and pressure:
            scale = np.max(np.abs(ref[f]))
            np.testing.assert_allclose(got, ref[f], rtol=0, atol=1e-12 * scale, err_msg=f)
        else:
            assert np.array_equal(got, ref[f]), (f, float(np.max(np.abs(got - ref[f]))))


def test_timing_and_errors():
    g = grid(2)
    v = vertical_grid(9, 10000.0)
    m = MPASOperators(g, v)
    dev = m.upload_state(*state(g, v, 1))
    m.set_timing(True)
    m.slow_tendencies(*dev, pressure=0.0)
    stats = m.kernel_stats()
    names = {s["name"] for s in stats}
    assert {"mpas_cell_vertex_kernel", "mpas_edge_kernel", "mpas_w_tendency_kernel", "mpas_w_kinetic_kernel",
            "mpas_vn_tendency_kernel"} <= names
    assert all(s["bytes"] > 0 and s["avg_ms"] > 0 for s in stats)
    with pytest.raises(ValueError, match="field must have shape"):
        m.upload(0, np.zeros((3, 5)))


@pytest.mark.parametrize("pressure_gradient", ["levels", "height"])
@pytest.mark.parametrize("name", ["iap", "advective", "mixed-ke-five-band"])
@pytest.mark.parametrize("pressure", [0.0, 1.0])
def test_slow_tendencies_over_orography_match_oracle(pressure_gradient, name, pressure):
    """Terrain-following levels over the DCMIP mountain (SLEVE coordinate): the contravariant
    correction, w - w_c, pi', its vertical derivative and every tendency equal the oracle's, bit
    for bit except where the last bit of pow enters (pi, pi' and the pressure gradient)."""
    from dmplex_amd.mpas_ops import dcmip_mountain_height, dcmip_rest

    say(f"mpas slow tendencies over orography, {pressure_gradient}, {name}, pressure {pressure}")
    g = grid(4)
    v = vertical_grid(12, 20000.0, stretch=2.0)
    m = MPASOperators(g, v, surface_height=dcmip_mountain_height, flat_height=16000.0,
                      pressure_gradient=pressure_gradient, **OPTIONS[name])
    assert m.orography
    rng = np.random.default_rng(7)
    T, p = dcmip_rest(m.z_full)
    rho = p / (287.0 * T) * (1 + 0.01 * rng.random(T.shape))
    rt = rho * T * (1e5 / p) ** (287.0 / 1004.0) * (1 + 0.002 * rng.random(T.shape))
    vn = 20 * rng.standard_normal((m.nlev, g.num_edges()))
    w = 0.5 * rng.standard_normal((m.nlev + 1, g.num_cells()))
    w[-1] = 0
    ref = MpasOracle(m).slow_tendencies(vn, w, rho, rt, pressure)
    m.slow_tendencies(*m.upload_state(vn, w, rho, rt), pressure=pressure)
    for f in FIELDS:
        if f == "Kv" and m.vertex_kinetic is None or f == "eta" and m.vertical == 0:
            continue
        if f in ("pip", "dpidz") and not pressure:
            continue
        got = m.field(f)
        if f == "exner":
            np.testing.assert_allclose(got, ref[f], rtol=5e-16, atol=0, err_msg=f)
        elif f in ("pip", "dpidz") or f == "tvn" and pressure:
            scale = np.max(np.abs(ref[f if f != "pip" else "exner"]))
            np.testing.assert_allclose(got, ref[f], rtol=0, atol=1e-12 * scale, err_msg=f)
        else:
            assert np.array_equal(got, ref[f]), (f, float(np.max(np.abs(got - ref[f]))))
    assert set(OROGRAPHY_FIELDS) <= set(FIELDS)
