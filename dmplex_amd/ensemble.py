"""Batched ensembles: many TRiSK shallow-water states on meshes of several resolutions,
advanced together on one device (``trisk_ensemble_*`` of ``include/trisk_hip.h``).

The reference steps each model with its own ``rk4_step!`` loop (RungeKutta.jl:18-28,
examples/trisk_shallow_water.jl:92-100).  ``TRiSKEnsemble`` takes one ``ModelTables`` per
level (TRiSK PV flux, cell kinetic energy) and a member count per level.  Each kernel of the
tendency is one launch over every member of every level, with the tables of a level shared
by its members.  Every member evolves bitwise as it would on a ``TRiSKShallowWater`` of its
own level with ``path="pair"``.

The ensemble state is one float64 CUDA tensor, level blocks in order, members interleaved
per entity; ``set_member`` / ``get_member`` move one member in the reference's global
order ``[u; h]``.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _device as D
from ._native import f64p, i64p, ptr
from .trisk_model import HIPBackend, ModelTables, _stream, _torch

SPLITS = {"chunked": 0, "blocked": 1, "interleaved": 2}
KERNELS = ("ens_state_kernel", "ens_cell_kernel", "ens_edge_kernel")


class TriskSchedule(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("blocks_per_cu", C.c_int32), ("split", C.c_int32),
                ("members_per_lane", C.c_int32)]

    def as_dict(self) -> dict:
        return dict(block_size=self.block_size, blocks_per_cu=self.blocks_per_cu,
                    split={v: k for k, v in SPLITS.items()}[self.split], members_per_lane=self.members_per_lane)


Schedule3 = TriskSchedule * 3


def _vp(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


class TRiSKEnsemble:
    """Members ``members[l]`` on the model tables ``levels[l]``, on a HIP device."""

    def __init__(self, levels: Sequence[ModelTables], members: Sequence[int], backend: HIPBackend | None = None):
        if len(levels) != len(members) or not levels:
            raise ValueError("one member count per level is needed")
        backend = backend or HIPBackend()
        torch = _torch()
        torch.cuda.set_device(backend.device)
        self.backend = backend
        self.levels = list(levels)          # the descriptors point into these tables
        self.members = [int(m) for m in members]
        descs = [t.descriptor() for t in self.levels]
        self._descs = descs
        arr = (C.POINTER(D.TriskModelDesc) * len(descs))(*[C.pointer(d) for d in descs])
        mem = np.asarray(self.members, dtype=np.int32)
        opts = D.TriskOptions(D.TRISK_FP64 if backend.precision == "fp64" else D.TRISK_FP32_BULK,
                              backend.device, 0, D.PATHS["pair"], 0)
        h = C.c_void_p()
        D.check(D.lib().trisk_ensemble_create(arr, mem.ctypes.data_as(C.POINTER(C.c_int32)), len(descs),
                                              C.byref(opts), C.byref(h)))
        self._handle = h
        n = np.zeros(1, dtype=np.int64)
        D.check(D.lib().trisk_ensemble_state_length(h, ptr(n, i64p)))
        self.state_length = int(n[0])
        self._level_handles = []
        self.offsets = []
        for lev in range(len(descs)):
            lh, off = C.c_void_p(), C.c_int64()
            D.check(D.lib().trisk_ensemble_level(h, lev, C.byref(lh), C.byref(off), None))
            self._level_handles.append(lh)
            self.offsets.append(int(off.value))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                D.lib().trisk_ensemble_destroy(h)
            except Exception:
                pass
            self._handle = None

    @property
    def handle(self) -> C.c_void_p:
        return self._handle

    @property
    def num_members(self) -> int:
        return sum(self.members)

    def cells(self) -> int:
        """Cell-member count: sum over levels of cells x members."""
        return sum(t.grid.num_cells() * m for t, m in zip(self.levels, self.members))

    def _level_length(self, level: int) -> int:
        g = self.levels[level].grid
        return g.num_edges() + g.num_cells()

    def _check(self, y) -> None:
        torch = _torch()
        if y.dtype != torch.float64 or not y.is_contiguous() or y.numel() != self.state_length:
            raise ValueError(f"ensemble states are contiguous float64 tensors of {self.state_length} entries")

    # --- members ----------------------------------------------------------------------
    def zeros(self):
        torch = _torch()
        return torch.zeros(self.state_length, dtype=torch.float64, device=f"cuda:{self.backend.device}")

    def set_member(self, y, level: int, member: int, y_host: np.ndarray) -> None:
        """Member ``member`` of ``level`` from a host [u; h] in the reference's order."""
        torch = _torch()
        self._check(y)
        y_host = np.ascontiguousarray(y_host, dtype=np.float64)
        n = self._level_length(level)
        if y_host.shape != (n,):
            raise ValueError(f"a level-{level} state has {n} entries")
        one = torch.empty(n, dtype=torch.float64, device=y.device)
        lib, s = D.lib(), _stream(torch)
        D.check(lib.trisk_upload_state(self._level_handles[level], ptr(y_host, f64p), _vp(one), s))
        D.check(lib.trisk_ensemble_set_member(self._handle, level, member, _vp(one), _vp(y), s))
        torch.cuda.current_stream().synchronize()

    def member_device(self, y, level: int, member: int):
        """Member ``member`` of ``level`` as a single-model device state (internal order)."""
        torch = _torch()
        self._check(y)
        one = torch.empty(self._level_length(level), dtype=torch.float64, device=y.device)
        D.check(D.lib().trisk_ensemble_get_member(self._handle, level, member, _vp(y), _vp(one), _stream(torch)))
        return one

    def get_member(self, y, level: int, member: int) -> np.ndarray:
        """Member ``member`` of ``level`` as a host [u; h] in the reference's order."""
        torch = _torch()
        one = self.member_device(y, level, member)
        out = np.empty(one.numel())
        D.check(D.lib().trisk_download_state(self._level_handles[level], _vp(one), ptr(out, f64p), _stream(torch)))
        return out

    # --- stepping ---------------------------------------------------------------------
    def tendency_(self, dy, y):
        """tendency! of every member; ``dy`` has the layout of ``y``."""
        torch = _torch()
        self._check(y)
        self._check(dy)
        D.check(D.lib().trisk_ensemble_tendency(self._handle, _vp(y), _vp(dy), _stream(torch)))
        return dy

    def rk4_steps_(self, y, dts: Sequence[float], nsteps: int):
        """``nsteps`` RK4 steps of every member in place, level ``l`` with time step ``dts[l]``."""
        torch = _torch()
        self._check(y)
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        if dt.shape != (len(self.levels),):
            raise ValueError("one time step per level is needed")
        D.check(D.lib().trisk_ensemble_rk4_steps(self._handle, _vp(y), ptr(dt, f64p), int(nsteps), _stream(torch)))
        return y

    def use_graphs(self, on: bool = True) -> None:
        D.check(D.lib().trisk_ensemble_set_graphs(self._handle, int(on)))

    def use_fused(self, on: bool = True) -> None:
        """One cooperative launch per RK stage for all members (state, cell and edge phases with
        a grid-wide barrier between them) instead of three launches; every member rounds the
        same either way.  Needs member counts divisible by 4; graphs are not used meanwhile."""
        D.check(D.lib().trisk_ensemble_set_fused(self._handle, int(on)))

    def diagnostics(self, y) -> np.ndarray:
        """(members, 3): total mass, total energy, potential enstrophy of every member."""
        torch = _torch()
        self._check(y)
        out = np.zeros(3 * self.num_members)
        D.check(D.lib().trisk_ensemble_diagnostics(self._handle, _vp(y), ptr(out, f64p), _stream(torch)))
        return out.reshape(-1, 3)

    # --- schedule -----------------------------------------------------------------------
    def schedule(self) -> list[dict]:
        s = Schedule3()
        D.check(D.lib().trisk_ensemble_get_schedule(self._handle, s))
        return [x.as_dict() for x in s]

    def set_schedule(self, sched: Sequence[dict]) -> None:
        s = Schedule3(*[TriskSchedule(int(d["block_size"]), int(d["blocks_per_cu"]), SPLITS[d["split"]],
                                      int(d.get("members_per_lane", 1))) for d in sched])
        D.check(D.lib().trisk_ensemble_set_schedule(self._handle, s))

    def autotune(self, y, dts: Sequence[float], steps: int = 4) -> tuple[list[dict], list[float]]:
        """Times the candidate schedules on a copy of ``y`` and keeps the fastest per kernel;
        returns (schedules, mean launch ms)."""
        torch = _torch()
        self._check(y)
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        best = Schedule3()
        ms = np.zeros(3)
        D.check(D.lib().trisk_ensemble_autotune(self._handle, _vp(y), ptr(dt, f64p), int(steps), _stream(torch),
                                                best, ptr(ms, f64p)))
        return [x.as_dict() for x in best], [float(v) for v in ms]

    # --- timing -------------------------------------------------------------------------
    def set_timing(self, on: bool) -> None:
        D.check(D.lib().trisk_ensemble_set_timing(self._handle, int(on)))

    def kernel_stats(self) -> list[dict]:
        """Mean duration, launches and algorithmic bytes of each kernel over the last timed call."""
        # the three kernels, and the fused stage (use_fused)
        names = (C.c_char_p * 4)()
        avg = np.zeros(4)
        nl = np.zeros(4, dtype=np.int64)
        cnt = C.c_int32()
        D.check(D.lib().trisk_ensemble_kernel_times(self._handle, 4, names, ptr(avg, f64p), ptr(nl, i64p),
                                                    C.byref(cnt)))
        by = np.zeros(4)
        bn = (C.c_char_p * 4)()
        cnt2 = C.c_int32()
        D.check(D.lib().trisk_ensemble_kernel_bytes(self._handle, 4, bn, ptr(by, f64p), C.byref(cnt2)))
        bytes_of = {bn[i].decode(): float(by[i]) for i in range(cnt2.value)}
        return [dict(name=names[i].decode(), avg_ms=float(avg[i]), launches=int(nl[i]),
                     bytes=bytes_of.get(names[i].decode(), 0.0)) for i in range(cnt.value)]
