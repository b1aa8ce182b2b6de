"""``TRiSKShallowWater`` on MI355X: the reference's model API over ``libtrisk_hip.so``.

Mirrors ``src/FiniteVolumes/TRiSKShallowWater.jl`` and ``src/TimeSteppers/RungeKutta.jl``:
the constructor folds the coefficients of the TRiSK operators exactly as the reference
does (:94-168) and hands them to the device library (``trisk_create``), which re-lays
them out in Hilbert order and owns the device memory.  ``tendency_`` (``tendency!``),
``rk4_step_`` (``rk4_step!``, fused on the device for this model) and the diagnostics
call the C ABI.  Julia's ``!`` functions end in ``_`` here.

Device state vectors are ``torch`` float64 CUDA tensors of length ``ne + nc`` in the
library's internal order (velocity block first, as ``velocity_dofs`` /
``height_dofs``); ``gather_state`` returns the reference's global order.  Functions
that take a state also accept a host numpy vector in the reference order and then
return host results in the reference order.  There is no CPU backend: without the HIP
library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _device as D
from ._native import f64p, i64p, ptr
from .trisk_grid import TRiSKGrid, gassmann_coefficients, vertex_edge_areas
from .topology import Plex

KINETIC = {"cell": 1.0, "mixed": 3 / 8, "vertex": 0.0}


class HIPBackend:
    """Backend of the model: HIP device ``device`` with stencil ``precision``
    ``"fp64"`` (default) or ``"fp32_bulk"`` (FP32 stencils, FP64 state and sums)."""

    def __init__(self, device: int = 0, precision: str = "fp64"):
        if precision not in ("fp64", "fp32_bulk"):
            raise ValueError("precision must be 'fp64' or 'fp32_bulk'")
        self.device = int(device)
        self.precision = precision

    def __repr__(self):
        return f"HIPBackend(device={self.device}, precision='{self.precision}')"


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the TRiSK model runs on MI355X only")
    return torch


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _F(a) -> np.ndarray:
    return np.asfortranarray(a)


class ModelTables:
    """Host-side coefficient tables of a TRiSK model, folded exactly as the reference
    constructor does (TRiSKShallowWater.jl:94-168), in the reference's layout.  This is
    what ``trisk_create`` receives; it needs no GPU."""

    def __init__(self, dm_or_grid, centers=None, gravity: float = 1.0, coriolis=None, pv_flux: str = "trisk",
                 kinetic_energy="cell"):
        if pv_flux not in ("trisk", "gassmann", "gassmann_mean"):
            raise ValueError("pv_flux must be :trisk, :gassmann or :gassmann_mean")
        if isinstance(kinetic_energy, str):
            if kinetic_energy not in KINETIC:
                raise ValueError("kinetic_energy must be :cell, :mixed, :vertex or a weight")
            alpha = KINETIC[kinetic_energy]
        else:
            alpha = float(kinetic_energy)
        grid = dm_or_grid if isinstance(dm_or_grid, TRiSKGrid) else TRiSKGrid(dm_or_grid, centers)
        if coriolis is None:
            coriolis = lambda x: 2 * x[..., 2]  # noqa: E731  (the reference default, Omega = 1)
        self.grid, self.gravity, self.alpha = grid, float(gravity), alpha
        self.pv_flux = pv_flux
        g = grid
        nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
        l, d = g.primal_lengths, g.dual_lengths
        ce, cs = g.cell_edges, g.cell_signs
        cells = np.arange(nc)[None, :]
        side = np.where(g.edge_cells[0, ce] == cells, 0, 1)
        real = cs != 0
        A = g.cell_areas[None, :]
        # coefficients of the global grid (TRiSKShallowWater.jl:105-123)
        self.cell_edges = g.cell_edges
        self.divergence = _F(np.where(real, cs * l[ce] / A, 0.0))
        self.kinetic = _F(np.where(real, alpha * l[ce] * g.cell_distances[side, ce] / (2 * A), 0.0))
        self.curl = _F(g.vertex_signs * d[g.vertex_edges] / g.vertex_areas[None, :])
        self.vertex_height = _F(g.vertex_kites / g.vertex_areas[None, :])
        self.trisk = _F(g.trisk_weights * l[g.trisk_edges] / d[None, :])
        self.edge_cells, self.edge_vertices = g.edge_cells, g.edge_vertices
        self.edge_height = _F(alpha * g.cell_distances / d[None, :])
        self.inverse_dual = 1 / d
        self.vertex_edges, self.vertex_cells, self.trisk_edges = g.vertex_edges, g.vertex_cells, g.trisk_edges
        self.coriolis = np.ascontiguousarray(coriolis(g.vertices), dtype=np.float64).reshape(nv)
        self.cell_areas, self.vertex_areas = g.cell_areas, g.vertex_areas
        self.gassmann = None
        if pv_flux != "trisk":
            E, P, W = gassmann_coefficients(g)
            rh = g.vertex_areas[g.edge_vertices]
            rh = _F(rh / rh.sum(axis=0, keepdims=True))
            self.gassmann = dict(edges=E, pv_edges=P, weights=W, rhombus=rh,
                                 coriolis=np.ascontiguousarray(coriolis(g.edge_points), dtype=np.float64).reshape(ne),
                                 mean=pv_flux == "gassmann_mean")
        self.vertex_kinetic = None
        if alpha != 1:
            by_vertex, by_edge = vertex_edge_areas(g)
            self.vertex_kinetic = dict(
                cell_vertices=g.cell_vertices, kites=_F((1 - alpha) * g.cell_kites / A),
                weights=_F(by_vertex / g.vertex_areas[None, :]),
                height=_F((1 - alpha) * 2 * by_edge / (l * d)[None, :]))


    def descriptor(self, hilbert: bool = True) -> D.TriskModelDesc:
        """The C-ABI view of the host tables (valid while the model lives)."""
        g = self.grid
        vk, gm = self.vertex_kinetic, self.gassmann
        i, f = i64p, f64p
        desc = D.TriskModelDesc(
            g.num_cells(), g.num_edges(), g.num_vertices(), self.cell_edges.shape[0],
            self.vertex_edges.shape[0], self.trisk_edges.shape[0], 12 if gm else 0,
            ptr(self.cell_edges, i), ptr(self.divergence, f), ptr(self.kinetic, f), ptr(self.edge_cells, i),
            ptr(self.edge_vertices, i), ptr(self.edge_height, f), ptr(self.inverse_dual, f),
            ptr(self.vertex_edges, i), ptr(self.curl, f), ptr(self.vertex_cells, i), ptr(self.vertex_height, f),
            ptr(self.coriolis, f), ptr(self.trisk_edges, i), ptr(self.trisk, f), ptr(self.cell_areas, f),
            ptr(self.vertex_areas, f), self.gravity,
            ptr(vk["cell_vertices"], i) if vk else None, ptr(vk["kites"], f) if vk else None,
            ptr(vk["weights"], f) if vk else None, ptr(vk["height"], f) if vk else None,
            ptr(gm["edges"], i) if gm else None, ptr(gm["pv_edges"], i) if gm else None,
            ptr(gm["weights"], f) if gm else None, ptr(gm["rhombus"], f) if gm else None,
            ptr(gm["coriolis"], f) if gm else None, D.PV_FLUX[self.pv_flux], 0,
            ptr(g.centers, f) if hilbert else None,
            ptr(g.primal_lengths, f), ptr(g.dual_lengths, f), ptr(g.cell_kites, f))
        return desc

    def check_tables(self) -> tuple[bool, str]:
        """Host-side validation through the C ABI (no device): (cell-pair path available?, reason)."""
        flag = C.c_int32()
        desc = self.descriptor()
        D.check(D.lib().trisk_check_tables(C.byref(desc), C.byref(flag)))
        return bool(flag.value), D.lib().trisk_last_error().decode()


class CGridShallowWater:
    """Device state and the generic calls of the C-grid shallow-water models
    (``CGridShallowWater``, TRiSKShallowWater.jl:172-198): the handle of the device library,
    the state layout [u; h] and the transfers between the reference's global order and the
    library's internal order.  ``TRiSKShallowWater`` and ``ICONShallowWater`` derive from it."""

    _handle = None

    def _finish_create(self, h: C.c_void_p) -> None:
        self._handle = h
        n = np.zeros(1, dtype=np.int64)
        D.check(D.lib().trisk_state_length(h, ptr(n, i64p)))
        self.state_length = int(n[0])

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                D.lib().trisk_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def handle(self) -> C.c_void_p:
        return self._handle

    def wide_counts(self) -> tuple[int, int, int]:
        """Cells, vertices, edges read with int32 instead of 16-bit-delta indices."""
        n = np.zeros(3, dtype=np.int64)
        D.check(D.lib().trisk_wide_counts(self._handle, ptr(n, i64p)))
        return tuple(int(x) for x in n)

    def set_finite_check(self, on: bool = True) -> None:
        """With the check on, ``diagnostics`` raises ``NonFiniteStateError`` for a state with
        NaN/Inf entries or a vertex height h_v <= 0 (trisk_set_finite_check)."""
        D.check(D.lib().trisk_set_finite_check(self._handle, int(on)))

    def check_state(self, y) -> tuple[int, int, int]:
        """(non-finite u entries, non-finite h entries, vertices with h_v <= 0 or NaN) of a
        state (trisk_check_state)."""
        torch = _torch()
        yd, _ = self._device_state(y)
        n = np.zeros(3, dtype=np.int64)
        D.check(D.lib().trisk_check_state(self._handle, C.c_void_p(yd.data_ptr()), ptr(n, i64p), _stream(torch)))
        return int(n[0]), int(n[1]), int(n[2])

    def last_error(self) -> str:
        """Message of the last failed call on this model's handle."""
        return D.lib().trisk_handle_last_error(self._handle).decode()

    def use_graphs(self, on: bool = True) -> None:
        """Replay the RK4 steps through a captured HIP graph (trisk_set_graphs)."""
        D.check(D.lib().trisk_set_graphs(self._handle, int(on)))

    def use_pair_plane(self, on: bool = True) -> None:
        """Cell-pair path: keep each RK4 stage's G_e and u_e as one pair per edge (the
        FP32-bulk default) or in two arrays (the FP64 default); the results are bitwise the
        same (trisk_set_pair_plane)."""
        D.check(D.lib().trisk_set_pair_plane(self._handle, int(on)))

    def permutation(self):
        """(cells, edges, vertices): reference index of every device entity."""
        g = self.grid
        c, e, v = (np.empty(n, dtype=np.int64) for n in (g.num_cells(), g.num_edges(), g.num_vertices()))
        D.check(D.lib().trisk_permutation(self._handle, ptr(c, i64p), ptr(e, i64p), ptr(v, i64p)))
        return c, e, v

    # --- state transfer -----------------------------------------------------------

    def zeros(self):
        torch = _torch()
        return torch.zeros(self.state_length, dtype=torch.float64, device=f"cuda:{self.backend.device}")

    def upload_state(self, y_host: np.ndarray):
        """Host [u; h] in reference order -> device state (internal order)."""
        torch = _torch()
        y_host = np.ascontiguousarray(y_host, dtype=np.float64)
        if y_host.shape != (self.state_length,):
            raise ValueError(f"state must have {self.state_length} entries")
        y = self.zeros()
        D.check(D.lib().trisk_upload_state(self._handle, ptr(y_host, f64p), C.c_void_p(y.data_ptr()), _stream(torch)))
        return y

    def download_state(self, y) -> np.ndarray:
        torch = _torch()
        out = np.empty(self.state_length)
        D.check(D.lib().trisk_download_state(self._handle, C.c_void_p(y.data_ptr()), ptr(out, f64p),
                                             _stream(torch)))
        return out

    def download_field(self, kind: int, x) -> np.ndarray:
        """Device field on cells (0), edges (1) or vertices (2) -> host, reference order."""
        torch = _torch()
        n = (self.grid.num_cells(), self.grid.num_edges(), self.grid.num_vertices())[kind]
        out = np.empty(n)
        D.check(D.lib().trisk_download_field(self._handle, kind, C.c_void_p(x.data_ptr()), ptr(out, f64p),
                                             _stream(torch)))
        return out

    def upload_field(self, kind: int, x_host: np.ndarray):
        torch = _torch()
        x_host = np.ascontiguousarray(x_host, dtype=np.float64)
        x = torch.empty(len(x_host), dtype=torch.float64, device=f"cuda:{self.backend.device}")
        D.check(D.lib().trisk_upload_field(self._handle, kind, ptr(x_host, f64p), C.c_void_p(x.data_ptr()),
                                           _stream(torch)))
        return x

    def _device_state(self, y):
        """(device tensor, was_host)."""
        if isinstance(y, np.ndarray):
            return self.upload_state(y), True
        return y, False


class TRiSKShallowWater(ModelTables, CGridShallowWater):
    """Nonlinear rotating shallow water in vector-invariant form on the Voronoi C grid
    with the TRiSK operators (TRiSKShallowWater.jl:1-66), on a HIP device."""

    def __init__(self, dm_or_grid, centers=None, gravity: float = 1.0, coriolis=None, pv_flux: str = "trisk",
                 kinetic_energy="cell", comm=None, part=None, backend=None, hilbert: bool = True,
                 path: str = "auto", narrow_limit: int = 0):
        if comm is not None or part is not None:
            raise NotImplementedError("the distributed model (comm/part) is not part of this one-GPU build")
        if backend is None:
            backend = HIPBackend()
        if not isinstance(backend, HIPBackend):
            raise NotImplementedError("only HIPBackend() is available: the product has no CPU path")
        super().__init__(dm_or_grid, centers, gravity, coriolis, pv_flux, kinetic_energy)
        self.backend = backend
        self._hilbert = hilbert
        if path not in D.PATHS:
            raise ValueError("path must be 'auto', 'tables' or 'pair'")
        self._path = path
        self._narrow_limit = int(narrow_limit)
        self._handle = None
        self._create()

    def _create(self) -> None:
        torch = _torch()
        torch.cuda.set_device(self.backend.device)
        opts = D.TriskOptions(D.TRISK_FP64 if self.backend.precision == "fp64" else D.TRISK_FP32_BULK,
                              self.backend.device, 0, D.PATHS[getattr(self, "_path", "auto")],
                              getattr(self, "_narrow_limit", 0))
        h = C.c_void_p()
        desc = self.descriptor(self._hilbert)
        D.check(D.lib().trisk_create(C.byref(desc), C.byref(opts), C.byref(h)))
        self._finish_create(h)

    def path_info(self) -> tuple[bool, str]:
        """(cell-pair path?, reason) of the device kernels."""
        flag = C.c_int32()
        note = C.c_char_p()
        D.check(D.lib().trisk_info(self._handle, C.byref(flag), C.byref(note)))
        return bool(flag.value), note.value.decode()


# --- state layout (TRiSKShallowWater.jl:172-195) ---------------------------------------

def velocity_dofs(m: CGridShallowWater) -> slice:
    return slice(0, m.grid.num_edges())


def height_dofs(m: CGridShallowWater) -> slice:
    return slice(m.grid.num_edges(), m.grid.num_edges() + m.grid.num_cells())


global_velocity_dofs, global_height_dofs = velocity_dofs, height_dofs


def initial_state_values(m: CGridShallowWater, u: np.ndarray, h: np.ndarray):
    """Device state from host edge-normal velocities and cell heights (reference order)."""
    return m.upload_state(np.concatenate([np.asarray(u, dtype=np.float64), np.asarray(h, dtype=np.float64)]))


def initial_state(m: CGridShallowWater, velocity, height):
    """Normal components of ``velocity(x)`` at the edge points and ``height(x)`` at the
    generators (TRiSKShallowWater.jl:184-188); both callables take an (n, 3) array."""
    g = m.grid
    vel = np.asarray(velocity(g.edge_points), dtype=np.float64).reshape(-1, 3)
    u = np.sum(vel * g.edge_normals, axis=1)
    h = np.asarray(height(g.centers), dtype=np.float64).reshape(-1)
    return initial_state_values(m, u, h)


def gather_state(m: CGridShallowWater, y) -> np.ndarray:
    """Global state [u; h] on the host (TRiSKShallowWater.jl:191-195)."""
    if isinstance(y, np.ndarray):
        return y.copy()
    return m.download_state(y)


# --- tendency and time stepping -----------------------------------------------------

def tendency_(dy, y, m: CGridShallowWater):
    """``tendency!(dy, y, m)`` (TRiSKShallowWater.jl:379-394): writes and returns ``dy``."""
    torch = _torch()
    if isinstance(y, np.ndarray):
        yd = m.upload_state(y)
        out = m.zeros()
        D.check(D.lib().trisk_tendency(m.handle, C.c_void_p(yd.data_ptr()), C.c_void_p(out.data_ptr()),
                                       _stream(torch)))
        dy[...] = m.download_state(out)
        return dy
    if dy.dtype != torch.float64 or y.dtype != torch.float64 or not dy.is_contiguous() or not y.is_contiguous():
        raise ValueError("device state vectors must be contiguous float64 tensors")
    if dy.numel() != m.state_length or y.numel() != m.state_length:
        raise ValueError(f"state vectors must have {m.state_length} entries")
    D.check(D.lib().trisk_tendency(m.handle, C.c_void_p(y.data_ptr()), C.c_void_p(dy.data_ptr()), _stream(torch)))
    return dy


class RK4Cache:
    """Work arrays of ``rk4_step!`` (RungeKutta.jl:2-10).  For a ``TRiSKShallowWater``
    the fused device step keeps its stage and accumulator inside the library, so the
    five vectors are only allocated when another tendency is stepped."""

    def __init__(self, y):
        self._like = y
        self._vectors = None

    def vectors(self):
        if self._vectors is None:
            y = self._like
            if isinstance(y, np.ndarray):
                self._vectors = [np.empty_like(y) for _ in range(5)]
            else:
                self._vectors = [y.new_empty(y.shape) for _ in range(5)]
        return self._vectors


def rk4_step_(y, dt: float, tendency, cache: RK4Cache, *args):
    """Classical RK4 step in place (RungeKutta.jl:18-28).

    With ``tendency_`` and a ``TRiSKShallowWater`` or ``ICONShallowWater`` the whole step
    runs on the device as fused launches (``trisk_rk4_steps``); otherwise the reference's
    broadcasts are evaluated on whatever array type ``y`` is.
    """
    if tendency is tendency_ and len(args) == 1 and isinstance(args[0], CGridShallowWater) \
            and not isinstance(y, np.ndarray):
        return rk4_steps_(y, dt, 1, args[0])
    k1, k2, k3, k4, stage = cache.vectors()
    tendency(k1, y, *args)
    stage[...] = y + dt / 2 * k1
    tendency(k2, stage, *args)
    stage[...] = y + dt / 2 * k2
    tendency(k3, stage, *args)
    stage[...] = y + dt * k3
    tendency(k4, stage, *args)
    y[...] = y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return y


def rk4_steps_(y, dt: float, nsteps: int, m: CGridShallowWater):
    """``nsteps`` fused RK4 steps of the model on the device state ``y``, in place."""
    torch = _torch()
    if y.dtype != torch.float64 or not y.is_contiguous() or y.numel() != m.state_length:
        raise ValueError(f"the state must be a contiguous float64 tensor of {m.state_length} entries")
    D.check(D.lib().trisk_rk4_steps(m.handle, C.c_void_p(y.data_ptr()), float(dt), int(nsteps), _stream(torch)))
    return y


# --- diagnostics (TRiSKShallowWater.jl:396-433) -----------------------------------------

def diagnostics(m: CGridShallowWater, y) -> tuple[float, float, float]:
    """(total mass, total energy, potential enstrophy), compensated sums."""
    torch = _torch()
    yd, _ = m._device_state(y)
    out = np.zeros(3)
    D.check(D.lib().trisk_diagnostics(m.handle, C.c_void_p(yd.data_ptr()), ptr(out, f64p), _stream(torch)))
    return float(out[0]), float(out[1]), float(out[2])


def total_mass(m: CGridShallowWater, y) -> float:
    """Sum_i A_i h_i (TRiSKShallowWater.jl:397-399)."""
    return diagnostics(m, y)[0]


def total_energy(m: CGridShallowWater, y) -> float:
    """Sum_i A_i (h_i K_i + g h_i^2 / 2) (TRiSKShallowWater.jl:406-411)."""
    return diagnostics(m, y)[1]


def potential_enstrophy(m: CGridShallowWater, y) -> float:
    """Sum_v A_v h_v q_v^2 / 2 (TRiSKShallowWater.jl:414-419)."""
    return diagnostics(m, y)[2]


class KineticWork:
    def __init__(self, K, B):
        self.K, self.B = K, B


def kinetic_energy_(m: TRiSKShallowWater, y) -> KineticWork:
    """K and B at the cells of the state ``y`` (kinetic_energy!, TRiSKShallowWater.jl:332-340)."""
    torch = _torch()
    yd, host = m._device_state(y)
    K = torch.empty(m.grid.num_cells(), dtype=torch.float64, device=yd.device)
    B = torch.empty_like(K)
    D.check(D.lib().trisk_kinetic_energy(m.handle, C.c_void_p(yd.data_ptr()), C.c_void_p(K.data_ptr()),
                                         C.c_void_p(B.data_ptr()), _stream(torch)))
    if host:
        return KineticWork(m.download_field(0, K), m.download_field(0, B))
    return KineticWork(K, B)


def vertex_pv_(m: CGridShallowWater, y):
    """(q_v, zeta_v, h_v) at the vertices (vertex_pv!, TRiSKShallowWater.jl:355-360)."""
    torch = _torch()
    yd, host = m._device_state(y)
    out = [torch.empty(m.grid.num_vertices(), dtype=torch.float64, device=yd.device) for _ in range(3)]
    D.check(D.lib().trisk_vertex_pv(m.handle, C.c_void_p(yd.data_ptr()), *(C.c_void_p(o.data_ptr()) for o in out),
                                    _stream(torch)))
    if host:
        return tuple(m.download_field(2, o) for o in out)
    return tuple(out)


def relative_vorticity(m: CGridShallowWater, y):
    """zeta_v at the vertices (TRiSKShallowWater.jl:422-426)."""
    torch = _torch()
    yd, host = m._device_state(y)
    z = torch.empty(m.grid.num_vertices(), dtype=torch.float64, device=yd.device)
    D.check(D.lib().trisk_relative_vorticity(m.handle, C.c_void_p(yd.data_ptr()), C.c_void_p(z.data_ptr()),
                                             _stream(torch)))
    return m.download_field(2, z) if host else z


def tangential_velocity(m: CGridShallowWater, u):
    """(1/d_e) Sum w_{e,e'} l_e' u_e' at the edges (TRiSKShallowWater.jl:429-433)."""
    torch = _torch()
    host = isinstance(u, np.ndarray)
    ud = m.upload_field(1, u) if host else u
    v = torch.empty(m.grid.num_edges(), dtype=torch.float64, device=ud.device)
    D.check(D.lib().trisk_tangential_velocity(m.handle, C.c_void_p(ud.data_ptr()), C.c_void_p(v.data_ptr()),
                                              _stream(torch)))
    return m.download_field(1, v) if host else v


# --- adjoint (the pullbacks behind a ChainRulesCore rrule of tendency! / rk4_step!) ----------

def _check_state(m: TRiSKShallowWater, *ys) -> None:
    torch = _torch()
    for y in ys:
        if y.dtype != torch.float64 or not y.is_contiguous() or y.numel() != m.state_length:
            raise ValueError(f"device states must be contiguous float64 tensors of {m.state_length} entries")


def tendency_vjp(m: TRiSKShallowWater, y, dybar):
    """(d tendency / dy)^T dybar at the device state ``y`` (the pullback of ``tendency!``,
    TRiSKShallowWater.jl:379-394, for every PV flux and kinetic-energy form)."""
    torch = _torch()
    _check_state(m, y, dybar)
    out = m.zeros()
    D.check(D.lib().trisk_tendency_vjp(m.handle, C.c_void_p(y.data_ptr()), C.c_void_p(dybar.data_ptr()),
                                       C.c_void_p(out.data_ptr()), _stream(torch)))
    return out


def rk4_adjoint(m: TRiSKShallowWater, y0, dt: float, nsteps: int, ybar, checkpoint_every: int = 8):
    """(ybar0, yT): the cotangent of ``y0`` for the cotangent ``ybar`` of the state after
    ``nsteps`` RK4 steps (RungeKutta.jl:18-28), with checkpoints every ``checkpoint_every``
    steps and recomputation in between; ``yT`` is that state."""
    torch = _torch()
    _check_state(m, y0, ybar)
    out, yT = m.zeros(), m.zeros()
    D.check(D.lib().trisk_rk4_adjoint(m.handle, C.c_void_p(y0.data_ptr()), float(dt), int(nsteps),
                                      int(checkpoint_every), C.c_void_p(ybar.data_ptr()), C.c_void_p(out.data_ptr()),
                                      C.c_void_p(yT.data_ptr()), _stream(torch)))
    return out, yT


def rrule_rk4_steps(m: TRiSKShallowWater, y0, dt: float, nsteps: int, checkpoint_every: int = 8):
    """The rrule of ``nsteps`` fused ``rk4_step!``: ``(yT, pullback)`` with
    ``pullback(ybarT) -> ybar0`` (ChainRulesCore's convention)."""
    yT = y0.clone()
    rk4_steps_(yT, dt, nsteps, m)

    def pullback(ybar):
        return rk4_adjoint(m, y0, dt, nsteps, ybar.contiguous(), checkpoint_every)[0]

    return yT, pullback


_AUTOGRAD = None


def rk4_steps(m: TRiSKShallowWater, y0, dt: float, nsteps: int, checkpoint_every: int = 8):
    """``nsteps`` RK4 steps from ``y0`` (not in place) as a differentiable torch operation:
    the backward pass runs ``trisk_rk4_adjoint``."""
    global _AUTOGRAD
    torch = _torch()
    if _AUTOGRAD is None:
        class RK4Steps(torch.autograd.Function):
            @staticmethod
            def forward(ctx, y, model, dt_, n_, ck_):
                ctx.model, ctx.dt, ctx.n, ctx.ck = model, dt_, n_, ck_
                ctx.save_for_backward(y)
                yT = y.detach().clone()
                rk4_steps_(yT, dt_, n_, model)
                return yT

            @staticmethod
            def backward(ctx, ybar):
                (y,) = ctx.saved_tensors
                g, _ = rk4_adjoint(ctx.model, y.detach(), ctx.dt, ctx.n, ybar.contiguous(), ctx.ck)
                return g, None, None, None, None

        _AUTOGRAD = RK4Steps
    return _AUTOGRAD.apply(y0, m, float(dt), int(nsteps), int(checkpoint_every))


def cell_areas_device(m: TRiSKShallowWater):
    """Cell areas as a device tensor in the library's internal order."""
    torch = _torch()
    if getattr(m, "_areas_dev", None) is None:
        perm = m.permutation()[0]
        m._areas_dev = torch.tensor(m.cell_areas[perm], dtype=torch.float64, device=f"cuda:{m.backend.device}")
    return m._areas_dev


def height_misfit(m: TRiSKShallowWater, y, hstar):
    """J = sum_i A_i (h_i - h*_i)^2 and its cotangent [0; 2 A (h - h*)] (device order;
    ``hstar`` a device vector of the cells)."""
    torch = _torch()
    A = cell_areas_device(m)
    ne = m.grid.num_edges()
    d = y[ne:] - hstar
    ybar = torch.zeros_like(y)
    ybar[ne:] = 2 * A * d
    return float(torch.sum(A * d * d).item()), ybar
