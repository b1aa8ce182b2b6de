"""ctypes binding of ``libtrisk_hip.so`` (include/trisk_hip.h).

The library is the only compute path of the package: if it is missing, or no HIP
device is visible, these calls raise.
"""
from __future__ import annotations

import ctypes as C

from ._native import LIB_DIR, NativeError, f64p, i64, i64p

vp = C.c_void_p


class TriskModelDesc(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("kc", i64), ("kv", i64), ("kt", i64), ("kg", i64),
                ("cell_edges", i64p), ("divergence", f64p), ("kinetic", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_height", f64p), ("inverse_dual", f64p),
                ("vertex_edges", i64p), ("curl", f64p), ("vertex_cells", i64p), ("vertex_height", f64p),
                ("coriolis", f64p), ("trisk_edges", i64p), ("trisk", f64p), ("cell_areas", f64p),
                ("vertex_areas", f64p), ("gravity", C.c_double),
                ("cell_vertices", i64p), ("vk_kites", f64p), ("vk_weights", f64p), ("vk_height", f64p),
                ("g_edges", i64p), ("g_pv_edges", i64p), ("g_weights", f64p), ("g_rhombus", f64p),
                ("g_coriolis", f64p), ("pv_flux", C.c_int32), ("reserved", C.c_int32),
                ("cell_points", f64p), ("primal_lengths", f64p), ("dual_lengths", f64p),
                ("cell_kites", f64p)]


class IconModelDesc(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("kv", i64), ("kr", i64),
                ("cell_edges", i64p), ("divergence", f64p), ("bilinear", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_height", f64p), ("gradient", f64p), ("inverse_dual", f64p),
                ("rbf_edges", i64p), ("rbf", f64p), ("edge_coriolis", f64p), ("vertex_edges", i64p),
                ("curl", f64p), ("vertex_cells", i64p), ("vertex_height", f64p), ("coriolis", f64p),
                ("cell_areas", f64p), ("vertex_areas", f64p), ("gravity", C.c_double),
                ("divergence_damping", f64p), ("cell_points", f64p)]


class MpasOpsDesc(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("nlev", i64), ("kc", i64), ("kv", i64), ("kt", i64),
                ("cell_edges", i64p), ("kinetic", f64p), ("inner", f64p), ("edge_cells", i64p),
                ("edge_vertices", i64p), ("edge_weight", f64p), ("inverse_dual", f64p), ("vertex_edges", i64p),
                ("curl", f64p), ("vertex_cells", i64p), ("vertex_weight", f64p), ("coriolis", f64p),
                ("trisk_index", C.POINTER(C.c_int32)), ("trisk", f64p), ("rhombus", f64p), ("edge_coriolis", f64p),
                ("dz", f64p), ("dzh", f64p), ("wlow", f64p), ("edge_dz", f64p), ("edge_dzh", f64p),
                ("cell_vertices", i64p), ("vk_kites", f64p), ("vk_weights", f64p),
                ("iap_pv", C.c_int32), ("vertical", C.c_int32), ("five_band", C.c_int32), ("reserved", C.c_int32),
                ("cell_points", f64p), ("orography", C.c_int32), ("reserved2", C.c_int32), ("z_full", f64p),
                ("pi_ref", f64p), ("ground_q", f64p), ("ddxn", f64p), ("zlev", C.POINTER(C.c_int32)), ("zwt", f64p)]


class TriskOptions(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32), ("block_size", C.c_int32),
                ("path", C.c_int32), ("narrow_limit", C.c_int32), ("reserved", C.c_int32 * 3)]


TRISK_FP64, TRISK_FP32_BULK = 0, 1
PV_FLUX = {"trisk": 0, "gassmann": 1, "gassmann_mean": 2}
PATHS = {"auto": 0, "tables": 1, "pair": 2}

# Every entry point declared in include/trisk_hip.h, with its argument types.
SIGNATURES = {
    "trisk_check_tables": [C.POINTER(TriskModelDesc), C.POINTER(C.c_int32)],
    "trisk_create": [C.POINTER(TriskModelDesc), C.POINTER(TriskOptions), C.POINTER(vp)],
    "trisk_destroy": [vp],
    "trisk_state_length": [vp, i64p],
    "trisk_permutation": [vp, i64p, i64p, i64p],
    "trisk_upload_state": [vp, f64p, vp, vp],
    "trisk_download_state": [vp, vp, f64p, vp],
    "trisk_download_field": [vp, C.c_int32, vp, f64p, vp],
    "trisk_upload_field": [vp, C.c_int32, f64p, vp, vp],
    "trisk_tendency": [vp, vp, vp, vp],
    "trisk_rk4_steps": [vp, vp, C.c_double, i64, vp],
    "trisk_diagnostics": [vp, vp, f64p, vp],
    "trisk_check_state": [vp, vp, i64p, vp],
    "trisk_diagnostics_noalloc": [vp, vp, vp, vp],
    "trisk_set_finite_check": [vp, C.c_int32],
    "trisk_kinetic_energy": [vp, vp, vp, vp, vp],
    "trisk_vertex_pv": [vp, vp, vp, vp, vp, vp],
    "trisk_relative_vorticity": [vp, vp, vp, vp],
    "trisk_tangential_velocity": [vp, vp, vp, vp],
    "trisk_set_timing": [vp, C.c_int32],
    "trisk_set_graphs": [vp, C.c_int32],
    "trisk_set_pair_plane": [vp, C.c_int32],
    "trisk_kernel_times": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, i64p, C.POINTER(C.c_int32)],
    "trisk_kernel_bytes": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, C.POINTER(C.c_int32)],
    "trisk_info": [vp, C.POINTER(C.c_int32), C.POINTER(C.c_char_p)],
    "trisk_wide_counts": [vp, i64p],
    "trisk_tendency_vjp": [vp, vp, vp, vp, vp],
    "trisk_rk4_adjoint": [vp, vp, C.c_double, i64, i64, vp, vp, vp, vp],
    "trisk_ensemble_create": [C.POINTER(C.POINTER(TriskModelDesc)), C.POINTER(C.c_int32), C.c_int32,
                              C.POINTER(TriskOptions), C.POINTER(vp)],
    "trisk_ensemble_destroy": [vp],
    "trisk_ensemble_state_length": [vp, i64p],
    "trisk_ensemble_level": [vp, C.c_int32, C.POINTER(vp), C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
    "trisk_ensemble_get_member": [vp, C.c_int32, C.c_int32, vp, vp, vp],
    "trisk_ensemble_set_member": [vp, C.c_int32, C.c_int32, vp, vp, vp],
    "trisk_ensemble_tendency": [vp, vp, vp, vp],
    "trisk_ensemble_rk4_steps": [vp, vp, f64p, i64, vp],
    "trisk_ensemble_set_graphs": [vp, C.c_int32],
    "trisk_ensemble_set_fused": [vp, C.c_int32],
    "trisk_ensemble_diagnostics": [vp, vp, f64p, vp],
    "trisk_ensemble_set_schedule": [vp, vp],
    "trisk_ensemble_get_schedule": [vp, vp],
    "trisk_ensemble_autotune": [vp, vp, f64p, C.c_int32, vp, vp, f64p],
    "trisk_ensemble_set_timing": [vp, C.c_int32],
    "trisk_ensemble_kernel_times": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, i64p, C.POINTER(C.c_int32)],
    "trisk_ensemble_kernel_bytes": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, C.POINTER(C.c_int32)],
    "icon_check_tables": [C.POINTER(IconModelDesc)],
    "icon_create": [C.POINTER(IconModelDesc), C.POINTER(TriskOptions), C.POINTER(vp)],
    "mpas_ops_check": [C.POINTER(MpasOpsDesc)],
    "mpas_ops_create": [C.POINTER(MpasOpsDesc), C.c_int32, C.POINTER(vp)],
    "mpas_ops_destroy": [vp],
    "mpas_ops_permutation": [vp, i64p, i64p, i64p],
    "mpas_ops_upload_field": [vp, C.c_int32, i64, f64p, vp, vp],
    "mpas_ops_download_field": [vp, C.c_int32, i64, vp, f64p, vp],
    "mpas_slow_tendencies": [vp, vp, vp, vp, vp, C.c_double, vp],
    "mpas_ops_field": [vp, C.c_int32, C.POINTER(vp), i64p, i64p],
    "mpas_ops_set_timing": [vp, C.c_int32],
    "mpas_ops_kernel_times": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, i64p, C.POINTER(C.c_int32)],
    "mpas_ops_kernel_bytes": [vp, C.c_int32, C.POINTER(C.c_char_p), f64p, C.POINTER(C.c_int32)],
    "trisk_lloyd": [C.c_int32, f64p, i64, i64p, i64, i64p, i64p, i64, C.c_double, C.c_int32, i64, i64p, f64p],
}

_lib = None


def lib() -> C.CDLL:
    """The device library (loaded once); raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64 under the same soname: loaded first, it is the one
        # HIP runtime of the process (loaded after ours, it would be a second runtime and
        # torch would see no device), so the state tensors and the kernels share it
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = LIB_DIR / "libtrisk_hip.so"
        if not path.exists():
            raise NativeError(f"{path} is missing: run `python -m dmplex_amd.build`")
        dll = C.CDLL(str(path))
        dll.trisk_last_error.restype = C.c_char_p
        dll.trisk_last_error.argtypes = []
        dll.trisk_handle_last_error.restype = C.c_char_p
        dll.trisk_handle_last_error.argtypes = [vp]
        for name, args in SIGNATURES.items():
            fn = getattr(dll, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = dll
    return _lib


TRISK_ENONFINITE = -4


class NonFiniteStateError(FloatingPointError):
    """The state check of a model found NaN/Inf in the state or a vertex height <= 0."""


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().trisk_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        if rc == TRISK_ENONFINITE:
            raise NonFiniteStateError(msg)
        raise NativeError(f"libtrisk_hip call failed ({rc}): {msg}")
