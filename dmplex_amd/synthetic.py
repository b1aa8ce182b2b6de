"""Seeded synthetic initial states of the benchmark configurations.

The model runs on the unit sphere, as the reference does (src/Grids/Grids.jl:1): with
radius ``a``, rotation ``Omega`` and gravity ``g`` the dimensional problem maps onto
``gravity = g / a^2`` and ``coriolis(x) = 2 Omega x_3``, with velocities stored as
``u / a`` (rad/s) and heights in metres.  Every TRiSK coefficient is dimensionless or
scales as 1/a, so this is exact (TRiSKShallowWater.jl:106-123).

State recipe: streamfunction ``psi = sum_{l=1..8} sum_{m=-l..l} a_lm Y_lm`` with real
orthonormal spherical harmonics and ``a_lm ~ N(0, 1) l^-1.5``; edge velocities from the
discrete skew gradient at the vertices, ``u_e = (psi_{v1} - psi_{v2}) / l_e`` (the
normal component of ``r x grad psi``, discretely non-divergent), rescaled so that
``max |u| = 30 m/s``; heights in linear balance ``h = h0 + f psi / g`` at the
generators plus ``N(0, 5 m)`` white noise.
"""
from __future__ import annotations

import numpy as np

from . import _native as N

EARTH_RADIUS = 6.371e6
EARTH_OMEGA = 7.292e-5
EARTH_GRAVITY = 9.80616


def earth_parameters(a: float = EARTH_RADIUS, omega: float = EARTH_OMEGA, g: float = EARTH_GRAVITY) -> dict:
    """``gravity`` and ``coriolis`` keywords of the dimensional problem on the unit sphere."""
    return dict(gravity=g / a ** 2, coriolis=lambda x: 2 * omega * x[..., 2])


def harmonic_coefficients(seed: int, lmax: int = 8) -> np.ndarray:
    """a_lm ~ N(0, 1) l^-1.5 for l = 1..lmax, m = -l..l (index l^2 - 1 + m + l)."""
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal(lmax * (lmax + 2))
    ls = np.concatenate([np.full(2 * l + 1, l) for l in range(1, lmax + 1)])
    return coef * ls.astype(np.float64) ** -1.5


def streamfunction(points: np.ndarray, coef: np.ndarray, lmax: int = 8) -> np.ndarray:
    """sum of a_lm Y_lm at unit vectors ``points`` (n, 3)."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    out = np.empty(len(pts))
    N.check(N.host().dmx_streamfunction(lmax, N.ptr(coef, N.f64p), len(pts), N.ptr(pts, N.f64p),
                                        N.ptr(out, N.f64p)))
    return out


def seeded_state(grid, seed_coef: int = 1, seed_noise: int = 2, umax: float = 30.0, h0: float = 8000.0,
                 noise: float = 5.0, a: float = EARTH_RADIUS, omega: float = EARTH_OMEGA,
                 g: float = EARTH_GRAVITY, lmax: int = 8) -> np.ndarray:
    """[u; h] in the reference's global order, unit-sphere units (u in rad/s, h in m)."""
    coef = harmonic_coefficients(seed_coef, lmax)
    psi_v = streamfunction(grid.vertices, coef, lmax)
    psi_c = streamfunction(grid.centers, coef, lmax)
    v1, v2 = grid.edge_vertices[0], grid.edge_vertices[1]
    u = (psi_v[v1] - psi_v[v2]) / grid.primal_lengths
    scale = (umax / a) / np.max(np.abs(u))
    u = u * scale
    psi_c = psi_c * scale
    f = 2 * omega * grid.centers[:, 2]
    h = h0 + f * psi_c / (g / a ** 2)
    if noise:
        h = h + np.random.default_rng(seed_noise).normal(0.0, noise, len(h))
    return np.concatenate([u, h])


def williamson2(grid, u0: float = 0.2):
    """Williamson test case 2 in the reference's units a = Omega = g = 1
    (test/finite_volumes.jl:80-82): solid-body rotation in geostrophic balance."""
    vel = u0 * np.stack([-grid.edge_points[:, 1], grid.edge_points[:, 0], np.zeros(len(grid.edge_points))], axis=1)
    u = np.sum(vel * grid.edge_normals, axis=1)
    h = 1 - (u0 + u0 ** 2 / 2) * grid.centers[:, 2] ** 2
    return np.concatenate([u, h])
