"""``ICONShallowWater`` on MI355X: the triangular C grid of ICON's dynamical core.

Mirrors ``src/FiniteVolumes/ICONShallowWater.jl``: the constructor folds the
coefficients exactly as the reference does (:96-143), with the vector-RBF tangential
velocity (``rbf_coefficients``, ``TRiSKGrid.jl:254-281``) and the bilinear edge-to-cell
weights (``bilinear_weights``, :82-94) computed in ``libdmplex_host.so``, and hands them to
``icon_create`` of ``libtrisk_hip.so``.  The model is a ``CGridShallowWater`` like
``TRiSKShallowWater``: ``initial_state``, ``tendency_``, ``rk4_step_``/``rk4_steps_``,
``gather_state``, the diagnostics, ``relative_vorticity`` and ``tangential_velocity``
(here the RBF one) take it unchanged.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _device as D
from . import _native as N
from ._native import f64p, i64p, ptr
from .trisk_grid import TRiSKGrid
from .trisk_model import CGridShallowWater, HIPBackend, _F, _torch


def rbf_coefficients(g: TRiSKGrid, scale: float = 0.5):
    """(edges, weights), both kr x ne: the vector-RBF stencil (the other edges of the two
    cells, four on triangles) and weights of the tangential velocity at every edge,
    ``sum_k weights[k, e] u[edges[k, e]] ~ u . (r x n_e)`` (TRiSKGrid.jl:254-281)."""
    nc, ne = g.num_cells(), g.num_edges()
    tables = g._tables()
    width = np.zeros(1, dtype=np.int64)
    N.check(N.host().dmx_rbf_coefficients(nc, ne, C.byref(tables), float(scale), 0, None, None, ptr(width, i64p)))
    kr = int(width[0])
    edges = np.zeros((kr, ne), dtype=np.int64, order="F")
    weights = np.zeros((kr, ne), order="F")
    N.check(N.host().dmx_rbf_coefficients(nc, ne, C.byref(tables), float(scale), kr, ptr(edges, i64p),
                                          ptr(weights, f64p), None))
    return edges, weights


def bilinear_weights(g: TRiSKGrid) -> np.ndarray:
    """3 x nc bilinear edge-to-cell weights of ICON (``e_bln_c_s``): partition of unity
    without first moment in the tangent plane of the cell center (:82-94)."""
    w = np.zeros((3, g.num_cells()), order="F")
    tables = g._tables()
    N.check(N.host().dmx_bilinear_weights(g.num_cells(), C.byref(tables), ptr(g.centers, f64p), ptr(w, f64p)))
    return w


class ICONTables:
    """Host-side tables of an ICON triangle model, folded as the reference constructor
    does (ICONShallowWater.jl:96-143), in the reference's layout; needs no GPU."""

    def __init__(self, dm_or_grid, centers=None, gravity: float = 1.0, coriolis=None, rbf_scale: float = 0.5,
                 divergence_damping=0.0, kinetic_gradient: str = "icon"):
        grid = dm_or_grid if isinstance(dm_or_grid, TRiSKGrid) else TRiSKGrid(dm_or_grid, centers)
        if kinetic_gradient not in ("icon", "centered"):
            raise ValueError(f"unknown kinetic_gradient {kinetic_gradient}")
        if coriolis is None:
            coriolis = lambda x: 2 * x[..., 2]  # noqa: E731  (the reference default, Omega = 1)
        g = grid
        nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
        if g.cell_edges.shape[0] != 3 or np.any(g.cell_signs == 0):
            raise ValueError("the ICON model needs a triangle mesh")
        self.grid, self.gravity = grid, float(gravity)
        self.kinetic_gradient = kinetic_gradient
        l, d, dc = g.primal_lengths, g.dual_lengths, g.cell_distances
        self.cell_edges = g.cell_edges
        self.divergence = _F(g.cell_signs * l[g.cell_edges] / g.cell_areas[None, :])
        self.bilinear = bilinear_weights(g)
        self.edge_cells, self.edge_vertices = g.edge_cells, g.edge_vertices
        self.edge_height = _F(np.stack([dc[1], dc[0]]) / d[None, :])
        if kinetic_gradient == "icon":
            self.gradient = _F(np.stack([dc[1] / (dc[0] * d), dc[0] / (dc[1] * d)]))
        else:
            self.gradient = _F(np.stack([1 / d, 1 / d]))
        self.inverse_dual = 1 / d
        self.curl = _F(g.vertex_signs * d[g.vertex_edges] / g.vertex_areas[None, :])
        self.rbf_edges, self.rbf = rbf_coefficients(g, rbf_scale)
        self.edge_coriolis = np.ascontiguousarray(coriolis(g.edge_points), dtype=np.float64).reshape(ne)
        self.vertex_edges, self.vertex_cells = g.vertex_edges, g.vertex_cells
        self.vertex_height = _F(g.vertex_kites / g.vertex_areas[None, :])
        self.coriolis = np.ascontiguousarray(coriolis(g.vertices), dtype=np.float64).reshape(nv)
        self.cell_areas, self.vertex_areas = g.cell_areas, g.vertex_areas
        if np.ndim(divergence_damping) == 0:
            nu = float(divergence_damping)
            self.divergence_damping = None if nu == 0 else np.full(ne, nu)
        else:
            nu = np.ascontiguousarray(divergence_damping, dtype=np.float64).reshape(-1)
            if len(nu) != ne:
                raise ValueError(f"divergence_damping needs one value per edge ({ne})")
            self.divergence_damping = nu

    def descriptor(self, hilbert: bool = True) -> D.IconModelDesc:
        """The C-ABI view of the host tables (valid while the model lives)."""
        g = self.grid
        i, f = i64p, f64p
        return D.IconModelDesc(
            g.num_cells(), g.num_edges(), g.num_vertices(), self.vertex_edges.shape[0], self.rbf_edges.shape[0],
            ptr(self.cell_edges, i), ptr(self.divergence, f), ptr(self.bilinear, f), ptr(self.edge_cells, i),
            ptr(self.edge_vertices, i), ptr(self.edge_height, f), ptr(self.gradient, f), ptr(self.inverse_dual, f),
            ptr(self.rbf_edges, i), ptr(self.rbf, f), ptr(self.edge_coriolis, f), ptr(self.vertex_edges, i),
            ptr(self.curl, f), ptr(self.vertex_cells, i), ptr(self.vertex_height, f), ptr(self.coriolis, f),
            ptr(self.cell_areas, f), ptr(self.vertex_areas, f), self.gravity,
            ptr(self.divergence_damping, f) if self.divergence_damping is not None else None,
            ptr(g.centers, f) if hilbert else None)

    def check_tables(self) -> None:
        """Host-side validation through the C ABI (no device)."""
        desc = self.descriptor()
        D.check(D.lib().icon_check_tables(C.byref(desc)))


class ICONShallowWater(ICONTables, CGridShallowWater):
    """Nonlinear rotating shallow water on the triangular C grid of ICON
    (ICONShallowWater.jl:1-48), on a HIP device: FP64 (the reference's arithmetic, bit for
    bit) or FP32-bulk (``HIPBackend(precision="fp32_bulk")``: FP32 stencils, FP64 state,
    heights, pressure gradient, continuity and RK4 accumulation)."""

    def __init__(self, dm_or_grid, centers=None, gravity: float = 1.0, coriolis=None, rbf_scale: float = 0.5,
                 divergence_damping=0.0, kinetic_gradient: str = "icon", comm=None, part=None, backend=None,
                 hilbert: bool = True, narrow_limit: int = 0):
        if comm is not None or part is not None:
            raise NotImplementedError("the distributed model (comm/part) is not part of this one-GPU build")
        if backend is None:
            backend = HIPBackend()
        if not isinstance(backend, HIPBackend):
            raise NotImplementedError("only HIPBackend() is available: the product has no CPU path")
        super().__init__(dm_or_grid, centers, gravity, coriolis, rbf_scale, divergence_damping, kinetic_gradient)
        self.backend = backend
        torch = _torch()
        torch.cuda.set_device(backend.device)
        opts = D.TriskOptions(D.TRISK_FP64 if backend.precision == "fp64" else D.TRISK_FP32_BULK, backend.device, 0,
                              0, int(narrow_limit))
        h = C.c_void_p()
        desc = self.descriptor(hilbert)
        D.check(D.lib().icon_create(C.byref(desc), C.byref(opts), C.byref(h)))
        self._finish_create(h)
