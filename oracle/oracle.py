"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of ``liboracle.so``, the FP64 CPU
restatement of the reference TRiSK path (trisk_oracle.c), of the ICON triangle model
(icon_oracle.c) and of the MPAS-Gassmann stage tendencies (mpas_oracle.c).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it,
as the checker; the product (``dmplex_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
i64, f64 = C.c_int64, C.c_double
i64p, f64p = C.POINTER(C.c_int64), C.POINTER(C.c_double)


class OracleModel(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("kc", i64), ("kv", i64), ("kt", i64), ("kg", i64),
                ("cell_edges", i64p), ("divergence", f64p), ("kinetic", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_height", f64p), ("inverse_dual", f64p),
                ("vertex_edges", i64p), ("curl", f64p), ("vertex_cells", i64p), ("vertex_height", f64p),
                ("coriolis", f64p), ("trisk_edges", i64p), ("trisk", f64p), ("cell_areas", f64p),
                ("vertex_areas", f64p), ("gravity", f64),
                ("cell_vertices", i64p), ("vk_kites", f64p), ("vk_weights", f64p), ("vk_height", f64p),
                ("g_edges", i64p), ("g_pv_edges", i64p), ("g_weights", f64p), ("g_rhombus", f64p),
                ("g_coriolis", f64p), ("g_mean", C.c_int),
                ("K", f64p), ("B", f64p), ("F", f64p), ("qe", f64p), ("zv", f64p), ("hv", f64p), ("qv", f64p),
                ("Kv", f64p)]


class IconOracleModel(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("kv", i64), ("kr", i64),
                ("cell_edges", i64p), ("divergence", f64p), ("bilinear", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_height", f64p), ("gradient", f64p), ("inverse_dual", f64p),
                ("rbf_edges", i64p), ("rbf", f64p), ("edge_coriolis", f64p), ("vertex_edges", i64p),
                ("curl", f64p), ("vertex_cells", i64p), ("vertex_height", f64p), ("coriolis", f64p),
                ("cell_areas", f64p), ("vertex_areas", f64p), ("gravity", f64), ("damping", f64p),
                ("v", f64p), ("Ke", f64p), ("K", f64p), ("F", f64p), ("zv", f64p), ("hv", f64p), ("qv", f64p),
                ("D", f64p), ("G", f64p)]


_lib = None


def build() -> Path:
    """Compiles liboracle.so (make)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "liboracle.so"


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # DMPLEX_ORACLE_LIBRARY selects another build of the oracle (the sanitizer build)
        path = Path(os.environ.get("DMPLEX_ORACLE_LIBRARY") or HERE / "liboracle.so")
        if not path.exists():
            build()
        d = C.CDLL(str(path))
        vp = C.c_void_p
        M = C.POINTER(OracleModel)
        for name, args, res in [
            ("oracle_tendency", [M, vp, vp], None),
            ("oracle_rk4_step", [M, vp, f64, vp, vp, vp, vp, vp], None),
            ("oracle_rk4_steps", [M, vp, f64, i64, vp, vp, vp, vp, vp], None),
            ("oracle_total_mass", [M, vp], f64),
            ("oracle_total_energy", [M, vp], f64),
            ("oracle_potential_enstrophy", [M, vp], f64),
            ("oracle_kinetic_energy", [M, vp, vp], None),
            ("oracle_vertex_pv", [M, vp, vp], None),
            ("oracle_mass_flux", [M, vp, vp], None),
            ("oracle_edge_sum", [i64, i64, vp, vp, vp, vp], None),
            ("oracle_tendency_vjp", [M, vp, vp, vp], C.c_int),
            ("oracle_rk4_adjoint", [M, vp, f64, i64, vp], C.c_int),
            ("icon_oracle_tendency", [C.POINTER(IconOracleModel), vp, vp], None),
            ("icon_oracle_rk4_steps", [C.POINTER(IconOracleModel), vp, f64, i64, vp, vp, vp, vp, vp], None),
            ("icon_oracle_total_mass", [C.POINTER(IconOracleModel), vp], f64),
            ("icon_oracle_total_energy", [C.POINTER(IconOracleModel), vp], f64),
            ("icon_oracle_potential_enstrophy", [C.POINTER(IconOracleModel), vp], f64),
            ("icon_oracle_kinetic_energy", [C.POINTER(IconOracleModel), vp, vp], None),
            ("icon_oracle_vertex_pv", [C.POINTER(IconOracleModel), vp, vp], None),
            ("mpas_oracle_slow_tendencies", [C.c_void_p, vp, vp, vp, vp, f64], None),
            ("mpas_oracle_kernel", [C.c_void_p, C.c_int, vp, vp, vp, vp, f64], C.c_int),
            ("oracle_threads", [], C.c_int),
            ("oracle_set_threads", [C.c_int], None),
        ]:
            fn = getattr(d, name)
            fn.argtypes = args
            fn.restype = res
        _lib = d
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Oracle:
    """The reference algorithm on the host tables of a model (any object with the
    TRiSKShallowWater table attributes, in the reference's layout)."""

    def __init__(self, m):
        self.m = m
        g = m.grid
        nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
        self.nc, self.ne, self.nv = nc, ne, nv
        self.work = {k: np.zeros(n) for k, n in
                     dict(K=nc, B=nc, F=ne, qe=ne, zv=nv, hv=nv, qv=nv, Kv=nv).items()}
        vk, gm = m.vertex_kinetic, m.gassmann
        w = self.work
        self.s = OracleModel(
            nc, ne, nv, m.cell_edges.shape[0], m.vertex_edges.shape[0], m.trisk_edges.shape[0], 12 if gm else 0,
            _p(m.cell_edges, i64p), _p(m.divergence, f64p), _p(m.kinetic, f64p), _p(m.edge_cells, i64p),
            _p(m.edge_vertices, i64p), _p(m.edge_height, f64p), _p(m.inverse_dual, f64p),
            _p(m.vertex_edges, i64p), _p(m.curl, f64p), _p(m.vertex_cells, i64p), _p(m.vertex_height, f64p),
            _p(m.coriolis, f64p), _p(m.trisk_edges, i64p), _p(m.trisk, f64p), _p(m.cell_areas, f64p),
            _p(m.vertex_areas, f64p), m.gravity,
            _p(vk["cell_vertices"], i64p) if vk else None, _p(vk["kites"], f64p) if vk else None,
            _p(vk["weights"], f64p) if vk else None, _p(vk["height"], f64p) if vk else None,
            _p(gm["edges"], i64p) if gm else None, _p(gm["pv_edges"], i64p) if gm else None,
            _p(gm["weights"], f64p) if gm else None, _p(gm["rhombus"], f64p) if gm else None,
            _p(gm["coriolis"], f64p) if gm else None, int(gm["mean"]) if gm else 0,
            *(_p(w[k], f64p) for k in ("K", "B", "F", "qe", "zv", "hv", "qv", "Kv")))
        self._cache = None

    def _ptr(self, a):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(C.c_void_p)

    def tendency(self, y: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.empty_like(y)
        lib().oracle_tendency(C.byref(self.s), self._ptr(y), self._ptr(dy))
        return dy

    def rk4_steps(self, y: np.ndarray, dt: float, nsteps: int) -> np.ndarray:
        """nsteps RK4 steps in place on the host vector y."""
        if self._cache is None or len(self._cache[0]) != len(y):
            self._cache = [np.empty_like(y) for _ in range(5)]
        k = self._cache
        lib().oracle_rk4_steps(C.byref(self.s), self._ptr(y), dt, nsteps, *(self._ptr(a) for a in k))
        return y

    def tendency_vjp(self, y: np.ndarray, lam: np.ndarray) -> np.ndarray:
        """(d tendency / dy)^T lam at y (every PV flux and kinetic-energy form)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.empty_like(y)
        if lib().oracle_tendency_vjp(C.byref(self.s), self._ptr(y), self._ptr(lam), self._ptr(out)) != 0:
            raise ValueError("oracle_tendency_vjp failed")
        return out

    def rk4_adjoint(self, y0: np.ndarray, dt: float, nsteps: int, ybar: np.ndarray) -> np.ndarray:
        """The cotangent of y0 for the cotangent ``ybar`` of the state after nsteps RK4 steps."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        out = np.array(ybar, dtype=np.float64)
        if lib().oracle_rk4_adjoint(C.byref(self.s), self._ptr(y0), dt, nsteps, self._ptr(out)) != 0:
            raise ValueError("oracle_rk4_adjoint failed")
        return out

    def total_mass(self, y):
        return lib().oracle_total_mass(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def total_energy(self, y):
        return lib().oracle_total_energy(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def potential_enstrophy(self, y):
        return lib().oracle_potential_enstrophy(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def kinetic_energy(self, y):
        y = np.ascontiguousarray(y)
        lib().oracle_kinetic_energy(C.byref(self.s), self._ptr(y), self._ptr(y[self.ne:]))
        return self.work["K"].copy(), self.work["B"].copy()

    def vertex_pv(self, y):
        y = np.ascontiguousarray(y)
        lib().oracle_vertex_pv(C.byref(self.s), self._ptr(y), self._ptr(y[self.ne:]))
        return self.work["qv"].copy(), self.work["zv"].copy(), self.work["hv"].copy()

    def edge_sum(self, edges, weights, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        k, n = edges.shape
        out = np.empty(n)
        lib().oracle_edge_sum(n, k, edges.ctypes.data_as(C.c_void_p), weights.ctypes.data_as(C.c_void_p),
                              self._ptr(x), self._ptr(out))
        return out

    def relative_vorticity(self, y):
        return self.edge_sum(self.m.vertex_edges, self.m.curl, y[:self.ne])

    def tangential_velocity(self, u):
        return self.edge_sum(self.m.trisk_edges, self.m.trisk, u)

    @staticmethod
    def threads() -> int:
        return lib().oracle_threads()

    @staticmethod
    def set_threads(n: int) -> None:
        lib().oracle_set_threads(n)


class IconOracle:
    """The reference ICON triangle model (ICONShallowWater.jl) on the host tables of an
    ``ICONTables``/``ICONShallowWater`` (reference layout)."""

    def __init__(self, m):
        self.m = m
        g = m.grid
        nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
        self.nc, self.ne, self.nv = nc, ne, nv
        self.work = {k: np.zeros(n) for k, n in dict(v=ne, Ke=ne, K=nc, F=ne, zv=nv, hv=nv, qv=nv, D=nc, G=ne).items()}
        w = self.work
        self.s = IconOracleModel(
            nc, ne, nv, m.vertex_edges.shape[0], m.rbf_edges.shape[0],
            _p(m.cell_edges, i64p), _p(m.divergence, f64p), _p(m.bilinear, f64p), _p(m.edge_cells, i64p),
            _p(m.edge_vertices, i64p), _p(m.edge_height, f64p), _p(m.gradient, f64p), _p(m.inverse_dual, f64p),
            _p(m.rbf_edges, i64p), _p(m.rbf, f64p), _p(m.edge_coriolis, f64p), _p(m.vertex_edges, i64p),
            _p(m.curl, f64p), _p(m.vertex_cells, i64p), _p(m.vertex_height, f64p), _p(m.coriolis, f64p),
            _p(m.cell_areas, f64p), _p(m.vertex_areas, f64p), m.gravity, _p(m.divergence_damping, f64p),
            *(_p(w[k], f64p) for k in ("v", "Ke", "K", "F", "zv", "hv", "qv", "D", "G")))
        self._cache = None

    def _ptr(self, a):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data_as(C.c_void_p)

    def tendency(self, y: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.empty_like(y)
        lib().icon_oracle_tendency(C.byref(self.s), self._ptr(y), self._ptr(dy))
        return dy

    def rk4_steps(self, y: np.ndarray, dt: float, nsteps: int) -> np.ndarray:
        """nsteps RK4 steps in place on the host vector y."""
        if self._cache is None or len(self._cache[0]) != len(y):
            self._cache = [np.empty_like(y) for _ in range(5)]
        lib().icon_oracle_rk4_steps(C.byref(self.s), self._ptr(y), dt, nsteps, *(self._ptr(a) for a in self._cache))
        return y

    def total_mass(self, y):
        return lib().icon_oracle_total_mass(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def total_energy(self, y):
        return lib().icon_oracle_total_energy(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def potential_enstrophy(self, y):
        return lib().icon_oracle_potential_enstrophy(C.byref(self.s), self._ptr(np.ascontiguousarray(y)))

    def kinetic_energy(self, y):
        """(v, K_e, F at the edges; K at the cells) of the state y."""
        y = np.ascontiguousarray(y)
        lib().icon_oracle_kinetic_energy(C.byref(self.s), self._ptr(y), self._ptr(y[self.ne:]))
        w = self.work
        return w["v"].copy(), w["Ke"].copy(), w["F"].copy(), w["K"].copy()

    def vertex_pv(self, y):
        y = np.ascontiguousarray(y)
        lib().icon_oracle_vertex_pv(C.byref(self.s), self._ptr(y), self._ptr(y[self.ne:]))
        return self.work["qv"].copy(), self.work["zv"].copy(), self.work["hv"].copy()

    def relative_vorticity(self, y):
        return Oracle.edge_sum(self, self.m.vertex_edges, self.m.curl, y[:self.ne])

    def tangential_velocity(self, u):
        return Oracle.edge_sum(self, self.m.rbf_edges, self.m.rbf, u)


class MpasOracleModel(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("nlev", i64), ("kc", i64), ("kv", i64), ("kt", i64),
                ("cell_edges", i64p), ("kinetic", f64p), ("inner", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_weight", f64p), ("inverse_dual", f64p), ("vertex_edges", i64p),
                ("curl", f64p), ("vertex_cells", i64p), ("vertex_weight", f64p), ("coriolis", f64p),
                ("trisk_index", C.POINTER(C.c_int32)), ("trisk", f64p), ("rhombus", f64p), ("edge_coriolis", f64p),
                ("dz", f64p), ("dzh", f64p), ("wlow", f64p), ("edge_dz", f64p), ("edge_dzh", f64p),
                ("cell_vertices", i64p), ("vk_kites", f64p), ("vk_weights", f64p),
                ("iap_pv", C.c_int32), ("vertical", C.c_int32), ("five_band", C.c_int32), ("reserved", C.c_int32)] + \
               [(n, f64p) for n in ("theta", "exner", "K", "qv", "zeta", "rhov", "qe", "rhoe", "Fs", "eta", "tw",
                                    "tvn", "Kv")] + \
               [("orography", C.c_int32), ("reserved2", C.c_int32), ("z_full", f64p), ("pi_ref", f64p),
                ("ground_q", f64p), ("ddxn", f64p), ("zlev", C.POINTER(C.c_int32)), ("zwt", f64p)] + \
               [(n, f64p) for n in ("wcf", "wch", "what", "pip", "dpidz")]


class MpasOracle:
    """The reference's explicit MPAS-Gassmann stage tendencies on flat ground or over orography
    (slow_tendencies!) on the host tables of an ``MPASOperatorTables``.  Fields are
    (levels, points) arrays in the reference numbering."""

    SHAPES = dict(theta="c", exner="c", K="c", qv="v", zeta="v", rhov="v", qe="e", rhoe="e", Fs="e", eta="e+",
                  tw="c+", tvn="e", Kv="v", wcf="c", wch="c+", what="c+", pip="c", dpidz="c")

    def __init__(self, m):
        self.m = m
        g = m.grid
        n = dict(c=g.num_cells(), e=g.num_edges(), v=g.num_vertices())
        L = m.nlev
        self.work = {k: np.zeros((L + 1 if s.endswith("+") else L, n[s[0]])) for k, s in self.SHAPES.items()}
        vk = m.vertex_kinetic
        w = self.work
        self.s = MpasOracleModel(
            n["c"], n["e"], n["v"], L, m.cell_edges.shape[0], m.vertex_edges.shape[0], m.trisk.shape[1],
            _p(m.cell_edges, i64p), _p(m.kinetic, f64p), _p(m.inner, f64p), _p(m.edge_cells, i64p),
            _p(m.edge_vertices, i64p), _p(m.edge_weight, f64p), _p(m.inverse_dual, f64p), _p(m.vertex_edges, i64p),
            _p(m.curl, f64p), _p(m.vertex_cells, i64p), _p(m.vertex_weight, f64p), _p(m.coriolis, f64p),
            _p(m.trisk_index, C.POINTER(C.c_int32)), _p(m.trisk, f64p), _p(m.rhombus, f64p),
            _p(m.edge_coriolis, f64p), _p(m.dz, f64p), _p(m.dzh, f64p), _p(m.wlow, f64p), _p(m.edge_dz, f64p),
            _p(m.edge_dzh, f64p),
            _p(vk["cell_vertices"], i64p) if vk else None, _p(vk["kites"], f64p) if vk else None,
            _p(vk["weights"], f64p) if vk else None,
            int(m.iap_pv), m.vertical, int(m.five_band), 0,
            *(_p(w[k], f64p) for k in ("theta", "exner", "K", "qv", "zeta", "rhov", "qe", "rhoe", "Fs", "eta", "tw",
                                       "tvn", "Kv")),
            m.orography_mode(), 0,
            *((_p(m.z_full, f64p), _p(m.pi_ref, f64p), _p(m.ground_q, f64p), _p(m.ddxn, f64p))
              if m.orography else (None,) * 4),
            *((_p(m.zlev, C.POINTER(C.c_int32)), _p(m.zwt, f64p)) if m.orography_mode() == 2 else (None, None)),
            *(_p(w[k], f64p) for k in ("wcf", "wch", "what", "pip", "dpidz")))

    def slow_tendencies(self, vn, w, rho, rhotheta, pressure: float = 1.0) -> dict:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (vn, w, rho, rhotheta)]
        lib().mpas_oracle_slow_tendencies(C.byref(self.s), *(a.ctypes.data_as(C.c_void_p) for a in arrs),
                                          float(pressure))
        return {k: v.copy() for k, v in self.work.items()}

    KERNELS = dict(diagnose=0, vertex_kinetic=1, cell_from_vertex=2, w_kinetic=3, vertex_pv=4, edge_pv=5,
                   edge_density=6, eta=7, w_tendency_vi=8, w_tendency=9, vn_tendency=10, concorr_cell=11,
                   concorr_interface=12, contravariant_w=13, exner_deviation=14, w_concorr_tendency=15)

    def kernel(self, name: str, vn=None, w=None, rho=None, rhotheta=None, pressure: float = 1.0) -> None:
        """One reference kernel on the current work fields (``self.work``, edited in place)."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (vn, w, rho, rhotheta)]
        ptrs = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in keep]
        if lib().mpas_oracle_kernel(C.byref(self.s), self.KERNELS[name], *ptrs, float(pressure)) != 0:
            raise ValueError(name)
