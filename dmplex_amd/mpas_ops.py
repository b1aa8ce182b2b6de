"""The TRiSK operators of the MPAS-Gassmann nonhydrostatic dycore with a level dimension.

Mirrors the parts of ``src/Nonhydrostatic`` that produce the explicit tendencies of one
stage on flat ground or over orography, ``slow_tendencies!(m, ys, pressure)`` (``MPAS/Gassmann/Stepping.jl:
22-56``): the vertical grid (``Common/VerticalGrid.jl:16-42``), the ICON-IAP TRiSK tables
(``trisk_pv_edges``, ``iap_trisk_weights``, ``FiniteVolumes/TRiSKGrid.jl:318-403``), the
horizontal tables of ``MPASGassmannNonhydrostatic`` (``MPASGassmannNonhydrostatic.jl:
196-256``), the terrain-following levels and metrics of orography (``Common/VerticalGrid.jl:
45-125``, ``Common/CGridOperators.jl:171-191``) and the device operators (``mpas_ops_*`` of
``libtrisk_hip.so``).  The time stepping, the vertical solvers and the diffusion are not
part of it.

Fields are host arrays of shape ``(nlev, n)`` (C order), the memory of the reference's
``(n, nlev)`` arrays, or torch tensors in the library's point numbering.
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple

import numpy as np

from . import _device as D
from . import _native as N
from ._native import f64p, i64p, ptr
from .trisk_grid import TRiSKGrid
from .trisk_model import KINETIC, ModelTables, _torch, _stream

Rd, Cpd, P0 = 287.0, 1004.0, 1.0e5
Cvd = Cpd - Rd
GRAV = 9.80616
# ICON's constants of its reference atmosphere (Common/Constants.jl:20-23)
ICON_RD, ICON_CPD, ICON_GRAV = 287.04, 1004.64, 9.80665

VerticalGrid = namedtuple("VerticalGrid", "z_half z_full dz dzh wlow")

FIELDS = {"theta": 0, "exner": 1, "K": 2, "qv": 3, "zeta": 4, "rhov": 5, "qe": 6, "rhoe": 7, "Fs": 8, "eta": 9,
          "tw": 10, "tvn": 11, "Kv": 12, "wcf": 13, "wch": 14, "what": 15, "pip": 16, "dpidz": 17}
_KIND = {"theta": 0, "exner": 0, "K": 0, "qv": 2, "zeta": 2, "rhov": 2, "qe": 1, "rhoe": 1, "Fs": 1, "eta": 1,
         "tw": 0, "tvn": 1, "Kv": 2, "wcf": 0, "wch": 0, "what": 0, "pip": 0, "dpidz": 0}
OROGRAPHY_FIELDS = ("wcf", "wch", "what", "pip", "dpidz")


def vertical_grid_from_half(z_half) -> VerticalGrid:
    """VerticalGrid(z_half) (VerticalGrid.jl:24-36): full levels, thicknesses, distances of
    the layer centres at the interfaces and the weight of the layer below."""
    z = np.asarray(z_half, dtype=np.float64)
    nlev = len(z) - 1
    if nlev < 1 or not np.all(np.diff(z) > 0):
        raise ValueError("z_half must increase")
    zf = (z[:nlev] + z[1:]) / 2
    dz = np.diff(z)
    dzh, wlow = np.zeros(nlev + 1), np.zeros(nlev + 1)
    dzh[1:nlev] = zf[1:] - zf[:-1]
    wlow[1:nlev] = (zf[1:] - z[1:nlev]) / dzh[1:nlev]
    return VerticalGrid(z, zf, dz, dzh, wlow)


def vertical_grid(nlev: int, z_top: float, stretch: float = 1.0) -> VerticalGrid:
    """VerticalGrid(nlev, z_top; stretch) (VerticalGrid.jl:38-42): layers growing
    geometrically to ``stretch`` times the lowest."""
    r = float(stretch) ** (1 / max(nlev - 1, 1))
    dz = np.array([r ** (k - 1) for k in range(1, nlev + 1)])
    return vertical_grid_from_half(np.concatenate([[0.0], np.cumsum(dz) * (z_top / np.sum(dz))]))


ColumnMetrics = namedtuple("ColumnMetrics", "z_half z_full dz dzh wlow")


def column_metrics(z_half) -> ColumnMetrics:
    """column_metrics(z_half) (VerticalGrid.jl:117-125): the metrics of vertical_grid_from_half
    for every column of the interfaces ``z_half`` (columns x nlev+1), as columns x layers or
    interfaces arrays."""
    z = np.asarray(z_half, dtype=np.float64)
    nlev = z.shape[1] - 1
    if nlev < 1 or not np.all(np.diff(z, axis=1) > 0):
        raise ValueError("z_half must increase in every column")
    zf = (z[:, :nlev] + z[:, 1:]) / 2
    dz = np.diff(z, axis=1)
    dzh, wlow = np.zeros((len(z), nlev + 1)), np.zeros((len(z), nlev + 1))
    dzh[:, 1:nlev] = zf[:, 1:] - zf[:, :-1]
    wlow[:, 1:nlev] = (zf[:, 1:] - z[:, 1:nlev]) / dzh[:, 1:nlev]
    return ColumnMetrics(z, zf, dz, dzh, wlow)


def terrain_following(vgrid: VerticalGrid, h, coordinate: str = "sleve", flat_height: float = 16000.0,
                      decay_scale: float = 4000.0, decay_exponent: float = 1.35) -> np.ndarray:
    """terrain_following (VerticalGrid.jl:57-69): interfaces z = zeta + h_i b(zeta) over the
    reference levels zeta of ``vgrid`` (columns x nlev+1); ``coordinate`` ``"sleve"`` (one
    decay scale up to ``flat_height``, flat above) or ``"basic"`` (b = 1 - zeta/z_top)."""
    zeta = np.asarray(vgrid.z_half, dtype=np.float64)
    if zeta[0] != 0:
        raise ValueError("the reference levels must start at 0")
    F, sc, n = min(float(flat_height), zeta[-1]), float(decay_scale), float(decay_exponent)
    if coordinate == "basic":
        b = 1 - zeta / zeta[-1]
    elif coordinate == "sleve":
        b = np.array([np.sinh((F / sc) ** n - (z / sc) ** n) / np.sinh((F / sc) ** n) if z < F else 0.0
                      for z in zeta])
    else:
        raise ValueError("coordinate must be :sleve or :basic")
    z = zeta[None, :] + np.asarray(h, dtype=np.float64)[:, None] * b[None, :]
    if not np.all(np.diff(z, axis=1) > 0):
        raise ValueError("the levels cross over the orography")
    return z


def lonlat(x):
    """(longitude, latitude) of points on the sphere (TestCases.jl:266)."""
    x = np.asarray(x, dtype=np.float64)
    return np.arctan2(x[..., 1], x[..., 0]), np.arcsin(np.clip(x[..., 2] / np.linalg.norm(x, axis=-1), -1, 1))


def dcmip_mountain_height(lon, lat, h0: float = 2000.0, R: float = 3 * np.pi / 4, zeta: float = np.pi / 16,
                          lon_c: float = 3 * np.pi / 2, lat_c: float = 0.0):
    """The mountain of DCMIP 2012 test 2-0 (TestCases.jl:57-60)."""
    r = np.arccos(np.clip(np.sin(lat_c) * np.sin(lat) + np.cos(lat_c) * np.cos(lat) * np.cos(lon - lon_c), -1, 1))
    return np.where(r < R, h0 / 2 * (1 + np.cos(np.pi * r / R)) * np.cos(np.pi * r / zeta) ** 2, 0.0)


def dcmip_rest(z, T0: float = 300.0, gamma: float = 0.0065):
    """The atmosphere at rest of DCMIP 2012 test 2-0 (TestCases.jl:68): (T, p) at height z."""
    z = np.asarray(z, dtype=np.float64)
    return T0 - gamma * z, P0 * (1 - gamma * z / T0) ** (GRAV / (Rd * gamma))


def reference_atmosphere(z):
    """ICON's reference atmosphere (VerticalGrid.jl:97-100): (T, pi) at height z."""
    z = np.asarray(z, dtype=np.float64)
    p = 101325.0 * np.exp(-ICON_GRAV / ICON_RD * 10000.0 / 213.15 * np.log((np.exp(z / 10000.0) * 213.15 + 75.0) / 288.15))
    return 213.15 + 75.0 * np.exp(-z / 10000.0), (p / P0) ** (ICON_RD / ICON_CPD)


def ground_extrapolation(c: ColumnMetrics) -> np.ndarray:
    """Weights (q1, q2, q3) of the quadratic extrapolation of the lowest three layers of every
    column to the ground (ICON wgtfacq; VerticalGrid.jl:103-110): columns x 3."""
    n = c.z_full.shape[0]
    if c.z_full.shape[1] < 3:
        return np.tile([1.0, 0.0, 0.0], (n, 1))
    z1, z2, z3 = (c.z_full[:, k] - c.z_half[:, 0] for k in range(3))
    q3 = z1 * z2 / (z2 - z3) / (z1 - z3)
    q2 = (z1 - q3 * (z1 - z3)) / (z1 - z2)
    return np.stack([1 - q2 - q3, q2, q3], axis=1)


def height_interpolation(edge_cells, zc, ze):
    """height_interpolation (CGridOperators.jl:178-191): for side s of every edge e and level
    k, the 0-based layer j of the cell below the edge height ze[e, k] (clamped to 0..nlev-2) and
    the weight (ze - zc[j]) / (zc[j+1] - zc[j]); arrays 2 x ne x nlev."""
    ne, nlev = ze.shape
    lev, wt = np.zeros((2, ne, nlev), dtype=np.int32), np.zeros((2, ne, nlev))
    if nlev < 2:
        return lev, wt
    for s in range(2):
        col = zc[edge_cells[s]]                                          # ne x nlev
        j = np.array([np.searchsorted(col[e], ze[e], side="right") for e in range(ne)]) - 1
        j = np.clip(j, 0, nlev - 2)
        lo = np.take_along_axis(col, j, axis=1)
        hi = np.take_along_axis(col, j + 1, axis=1)
        lev[s], wt[s] = j, (ze - lo) / (hi - lo)
    return lev, wt


def height_difference(pip, edge_cells, lev, wt, e: int, k: int) -> float:
    """height_difference (CGridOperators.jl:194-200): pi' of the right cell minus the left one,
    both at the edge height (pip cells x levels, lev/wt from height_interpolation)."""
    l, r = edge_cells[0, e], edge_cells[1, e]
    jl, tl, jr, tr = lev[0, e, k], wt[0, e, k], lev[1, e, k], wt[1, e, k]
    return pip[r, jr] + tr * (pip[r, jr + 1] - pip[r, jr]) - pip[l, jl] - tl * (pip[l, jl + 1] - pip[l, jl])


def trisk_pv_edges(g: TRiSKGrid) -> np.ndarray:
    """kt x ne: the edge whose rhombus PV replaces the mean for each TRiSK pair
    (i_cori_method = 4, TRiSKGrid.jl:318-362), -1 for the mean."""
    out = np.empty(g.trisk_edges.shape, dtype=np.int64, order="F")
    tables = g._tables()
    N.check(N.host().dmx_trisk_pv_edges(g.num_cells(), g.num_edges(), C.byref(tables), ptr(out, i64p)))
    return out


def iap_trisk_weights(g: TRiSKGrid) -> np.ndarray:
    """kt x ne TRiSK weights of ICON-IAP from planar kite fractions (TRiSKGrid.jl:364-403)."""
    out = np.empty(g.trisk_weights.shape, order="F")
    tables = g._tables()
    N.check(N.host().dmx_iap_trisk_weights(g.num_cells(), g.num_edges(), C.byref(tables),
                                           ptr(np.ascontiguousarray(g.vertices), f64p), ptr(out, f64p)))
    return out


class MPASOperatorTables:
    """Host tables of the explicit MPAS-Gassmann operators on flat ground, folded as the
    reference constructor does (MPASGassmannNonhydrostatic.jl:196-256); needs no GPU.

    ``pv_flux`` ``"iap"`` (rhombus PV, i_cori_method = 4) or ``"trisk"``;
    ``vertical_advection`` ``"iap"``, ``"vector_invariant"`` or ``"advective"``;
    ``trisk_weights`` ``"iap"`` (planar kites) or ``"exact"``; ``vertical_solver``
    ``"five_band"`` or ``"tridiagonal"`` (where K_w enters K); ``kinetic_energy`` as for
    ``TRiSKShallowWater``.
    """

    def __init__(self, grid: TRiSKGrid, vgrid: VerticalGrid, radius: float = 6.37122e6, omega: float = 7.292e-5,
                 coriolis=None, kinetic_energy="cell", pv_flux: str = "iap", vertical_advection: str = "iap",
                 trisk_weights: str = "iap", vertical_solver: str = "five_band", surface_height=None, z_half=None,
                 coordinate: str = "sleve", flat_height: float = 16000.0, decay_scale: float = 4000.0,
                 decay_exponent: float = 1.35, pressure_gradient: str = "levels"):
        if pv_flux not in ("iap", "trisk"):
            raise ValueError("pv_flux must be :iap or :trisk")
        vertical = {"advective": 0, "vector_invariant": 1, "iap": 2}.get(vertical_advection)
        if vertical is None:
            raise ValueError("vertical_advection must be :iap, :vector_invariant or :advective")
        if trisk_weights not in ("iap", "exact"):
            raise ValueError("trisk_weights must be :iap or :exact")
        if vertical_solver not in ("five_band", "tridiagonal"):
            raise ValueError("vertical_solver must be :five_band or :tridiagonal")
        if vertical_solver == "five_band" and vertical == 0:
            raise ValueError("vertical_solver = :five_band needs the vector-invariant vertical advection")
        if pressure_gradient not in ("levels", "height"):
            raise ValueError("pressure_gradient must be :levels or :height")
        if coriolis is None:
            coriolis = lambda x: 2 * omega * x[..., 2]  # noqa: E731
        a = float(radius)
        g = grid
        sw = ModelTables(g, coriolis=coriolis, kinetic_energy=kinetic_energy)
        self.grid, self.vgrid, self.radius = g, vgrid, a
        self.iap_pv, self.vertical, self.five_band = pv_flux == "iap", vertical, vertical_solver == "five_band"
        nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
        self.nlev = len(vgrid.dz)
        self.cell_edges, self.kinetic = sw.cell_edges, sw.kinetic
        self.edge_cells, self.edge_vertices = g.edge_cells, g.edge_vertices
        self.vertex_edges, self.vertex_cells = g.vertex_edges, g.vertex_cells
        self.edge_weight = np.asfortranarray(g.cell_distances / g.dual_lengths[None, :])
        self.inverse_dual = sw.inverse_dual / a
        self.curl = np.asfortranarray(sw.curl / a)
        self.vertex_weight = sw.vertex_height
        self.coriolis = sw.coriolis
        ce, cs = g.cell_edges, g.cell_signs
        side = np.where(g.edge_cells[0, ce] == np.arange(nc)[None, :], 0, 1)
        self.inner = np.asfortranarray(np.where(cs != 0, g.primal_lengths[ce] * g.cell_distances[side, ce]
                                                / g.cell_areas[None, :], 0.0))
        rh = g.vertex_areas[g.edge_vertices]
        self.rhombus = np.asfortranarray(rh / rh.sum(axis=0, keepdims=True))
        self.edge_coriolis = np.ascontiguousarray(coriolis(g.edge_points), dtype=np.float64).reshape(ne)
        # TRiSK tables with the edge first, the PV of each neighbour by two edges, 1/2 in the weight
        pv = trisk_pv_edges(g) if self.iap_pv else np.full(g.trisk_edges.shape, -1, dtype=np.int64)
        W = iap_trisk_weights(g) if trisk_weights == "iap" else g.trisk_weights
        e = np.broadcast_to(np.arange(ne)[None, :], g.trisk_edges.shape)
        t = W * g.primal_lengths[g.trisk_edges] / (2 * g.dual_lengths[None, :])
        e2 = g.trisk_edges
        zero = t == 0
        a_, b_, c_ = (np.where(zero, e, e2), np.where(zero, e, np.where(pv < 0, e, pv)),
                      np.where(zero, e, np.where(pv < 0, e2, pv)))
        self.trisk = np.asfortranarray(t.T)                                           # ne x kt
        self.trisk_index = np.asfortranarray(np.stack([a_.T, b_.T, c_.T], axis=2).astype(np.int32))  # ne x kt x 3
        self.vertex_kinetic = sw.vertex_kinetic
        # the interfaces of the cells (model_levels, VerticalGrid.jl:77-89): z_half, or
        # terrain-following over surface_height(lon, lat) (or the heights of the cells), or flat
        nl = self.nlev
        if z_half is not None:
            zi = np.asarray(z_half, dtype=np.float64)
            if zi.shape != (nc, nl + 1):
                raise ValueError("z_half must be cells x nlev+1")
        elif surface_height is not None:
            hs = surface_height if not callable(surface_height) else surface_height(*lonlat(g.centers))
            zi = terrain_following(vgrid, np.asarray(hs, dtype=np.float64).reshape(nc), coordinate, flat_height,
                                   decay_scale, decay_exponent)
        else:
            zi = np.broadcast_to(np.asarray(vgrid.z_half)[None, :], (nc, nl + 1))
        self.orography = bool(np.any(zi != np.asarray(vgrid.z_half)[None, :]))
        cm = column_metrics(zi)
        # the edges weight the interfaces of their cells with d_ei/d_e (MPASGassmannNonhydrostatic.jl:226-227)
        l_, r_ = g.edge_cells
        em = column_metrics(self.edge_weight[0][:, None] * zi[l_] + self.edge_weight[1][:, None] * zi[r_])
        self.cell_metrics, self.edge_metrics = cm, em
        self.dz = np.ascontiguousarray(cm.dz.T)
        self.dzh = np.ascontiguousarray(cm.dzh.T)
        self.wlow = np.ascontiguousarray(cm.wlow.T)
        self.edge_dz = np.ascontiguousarray(em.dz.T)
        self.edge_dzh = np.ascontiguousarray(em.dzh.T)
        # orography metrics (:228-233): slopes of the levels along the edge normals, the height
        # interpolation of pressure_gradient = :height, pi of the reference atmosphere, wgtfacq
        self.pressure_gradient = pressure_gradient
        self.z_full = np.ascontiguousarray(cm.z_full.T)                                   # nlev x nc
        self.ddxn = np.ascontiguousarray(((cm.z_full[r_] - cm.z_full[l_]) / (a * g.dual_lengths[:, None])).T)
        zlev, zwt = height_interpolation(g.edge_cells, cm.z_full, em.z_full)
        self.zlev = np.ascontiguousarray(zlev.transpose(2, 1, 0))                          # nlev x ne x 2
        self.zwt = np.ascontiguousarray(zwt.transpose(2, 1, 0))
        self.pi_ref = np.ascontiguousarray(reference_atmosphere(cm.z_full)[1].T)
        self.ground_q = np.ascontiguousarray(ground_extrapolation(cm).T)                   # 3 x nc

    def descriptor(self, hilbert: bool = True) -> D.MpasOpsDesc:
        g = self.grid
        vk = self.vertex_kinetic
        i, f = i64p, f64p
        return D.MpasOpsDesc(
            g.num_cells(), g.num_edges(), g.num_vertices(), self.nlev, self.cell_edges.shape[0],
            self.vertex_edges.shape[0], self.trisk.shape[1],
            ptr(self.cell_edges, i), ptr(self.kinetic, f), ptr(self.inner, f), ptr(self.edge_cells, i),
            ptr(self.edge_vertices, i), ptr(self.edge_weight, f), ptr(self.inverse_dual, f), ptr(self.vertex_edges, i),
            ptr(self.curl, f), ptr(self.vertex_cells, i), ptr(self.vertex_weight, f), ptr(self.coriolis, f),
            ptr(self.trisk_index, C.POINTER(C.c_int32)), ptr(self.trisk, f), ptr(self.rhombus, f),
            ptr(self.edge_coriolis, f), ptr(self.dz, f), ptr(self.dzh, f), ptr(self.wlow, f), ptr(self.edge_dz, f),
            ptr(self.edge_dzh, f),
            ptr(vk["cell_vertices"], i) if vk else None, ptr(vk["kites"], f) if vk else None,
            ptr(vk["weights"], f) if vk else None,
            int(self.iap_pv), self.vertical, int(self.five_band), 0, ptr(g.centers, f) if hilbert else None,
            self.orography_mode(), 0,
            *((ptr(self.z_full, f), ptr(self.pi_ref, f), ptr(self.ground_q, f), ptr(self.ddxn, f))
              if self.orography else (None,) * 4),
            *((ptr(self.zlev, C.POINTER(C.c_int32)), ptr(self.zwt, f)) if self.orography_mode() == 2 else (None, None)))

    def orography_mode(self) -> int:
        """0 flat ground, 1 the pressure gradient along the levels, 2 at constant height."""
        return 0 if not self.orography else 1 if self.pressure_gradient == "levels" else 2

    def check(self) -> None:
        desc = self.descriptor()
        D.check(D.lib().mpas_ops_check(C.byref(desc)))


class MPASOperators(MPASOperatorTables):
    """The device operators (``mpas_ops_create``): ``slow_tendencies`` of a stage state on
    a HIP device, in FP64."""

    def __init__(self, grid: TRiSKGrid, vgrid: VerticalGrid, device: int = 0, hilbert: bool = True, **kwargs):
        super().__init__(grid, vgrid, **kwargs)
        torch = _torch()
        torch.cuda.set_device(device)
        self.device = device
        h = C.c_void_p()
        desc = self.descriptor(hilbert)
        D.check(D.lib().mpas_ops_create(C.byref(desc), device, C.byref(h)))
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                D.lib().mpas_ops_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def handle(self):
        return self._handle

    def _points(self, kind: int) -> int:
        g = self.grid
        return (g.num_cells(), g.num_edges(), g.num_vertices())[kind]

    def upload(self, kind: int, host: np.ndarray):
        """Host field (levels, points) of cells (0), edges (1) or vertices (2) -> device tensor."""
        torch = _torch()
        host = np.ascontiguousarray(host, dtype=np.float64)
        if host.ndim != 2 or host.shape[1] != self._points(kind):
            raise ValueError(f"field must have shape (levels, {self._points(kind)})")
        x = torch.empty(host.size, dtype=torch.float64, device=f"cuda:{self.device}")
        D.check(D.lib().mpas_ops_upload_field(self._handle, kind, host.shape[0], ptr(host, f64p),
                                              C.c_void_p(x.data_ptr()), _stream(torch)))
        return x

    def download(self, kind: int, levels: int, dev) -> np.ndarray:
        torch = _torch()
        out = np.empty((levels, self._points(kind)))
        dptr = dev if isinstance(dev, int) else dev.data_ptr()
        D.check(D.lib().mpas_ops_download_field(self._handle, kind, levels, C.c_void_p(dptr), ptr(out, f64p),
                                                _stream(torch)))
        return out

    def upload_state(self, vn, w, rho, rhotheta):
        """Host (vn, w, rho, rhotheta) -> device tensors in the library's numbering."""
        return self.upload(1, vn), self.upload(0, w), self.upload(0, rho), self.upload(0, rhotheta)

    def slow_tendencies(self, vn, w, rho, rhotheta, pressure: float = 1.0) -> None:
        """slow_tendencies!(m, ys, pressure) of device fields (library numbering)."""
        torch = _torch()
        D.check(D.lib().mpas_slow_tendencies(self._handle, *(C.c_void_p(x.data_ptr()) for x in (vn, w, rho, rhotheta)),
                                             float(pressure), _stream(torch)))

    def field(self, name: str) -> np.ndarray:
        """A work field on the host, (levels, points) in the reference numbering."""
        p = C.c_void_p()
        n, lev = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
        D.check(D.lib().mpas_ops_field(self._handle, FIELDS[name], C.byref(p), ptr(n, i64p), ptr(lev, i64p)))
        return self.download(_KIND[name], int(lev[0]), p.value)

    def set_timing(self, on: bool = True) -> None:
        D.check(D.lib().mpas_ops_set_timing(self._handle, int(on)))

    def kernel_stats(self):
        names = (C.c_char_p * 8)()
        avg, nl = np.zeros(8), np.zeros(8, dtype=np.int64)
        cnt = C.c_int32()
        D.check(D.lib().mpas_ops_kernel_times(self._handle, 8, names, ptr(avg, f64p), ptr(nl, i64p), C.byref(cnt)))
        bn = (C.c_char_p * 8)()
        by = np.zeros(8)
        c2 = C.c_int32()
        D.check(D.lib().mpas_ops_kernel_bytes(self._handle, 8, bn, ptr(by, f64p), C.byref(c2)))
        bytes_of = {bn[i].decode(): float(by[i]) for i in range(c2.value)}
        return [dict(name=names[i].decode(), avg_ms=float(avg[i]), launches=int(nl[i]),
                     bytes=bytes_of.get(names[i].decode(), 0.0)) for i in range(cnt.value)]
