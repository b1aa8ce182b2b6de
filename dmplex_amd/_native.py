"""ctypes bindings of the two native libraries of the package.

``libdmplex_host.so`` (host setup code) and ``libtrisk_hip.so`` (device path) are
built in-tree by :mod:`dmplex_amd.build`.  A missing library is an error: there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

i64 = C.c_int64
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
i8p = C.POINTER(C.c_int8)
u64p = C.POINTER(C.c_uint64)


class NativeError(RuntimeError):
    """Error reported by a native library call."""


def ptr(a: np.ndarray | None, ctype=None):
    """Raw pointer of a contiguous numpy array (or NULL)."""
    if a is None:
        return None
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError("array must be contiguous")
    p = a.ctypes.data_as(C.c_void_p)
    return C.cast(p, ctype) if ctype is not None else p


class DmxPlex(C.Structure):
    _fields_ = [("nc", i64), ("ne", i64), ("nv", i64), ("cone_offsets", i64p), ("cones", i64p),
                ("orientations", i8p), ("support_offsets", i64p), ("supports", i64p),
                ("coordinates", f64p)]


class DmxGridTables(C.Structure):
    _fields_ = [("kc", i64), ("kv", i64), ("kt", i64),
                ("edge_points", f64p), ("edge_normals", f64p), ("primal_lengths", f64p),
                ("dual_lengths", f64p), ("edge_cells", i64p), ("edge_vertices", i64p),
                ("cell_distances", f64p), ("cell_areas", f64p), ("vertex_areas", f64p),
                ("cell_edges", i64p), ("cell_signs", f64p), ("cell_vertices", i64p),
                ("cell_kites", f64p), ("vertex_edges", i64p), ("vertex_signs", f64p),
                ("vertex_cells", i64p), ("vertex_kites", f64p), ("trisk_edges", i64p),
                ("trisk_weights", f64p)]


_host = None


def host() -> C.CDLL:
    """The host setup library (loaded once)."""
    global _host
    if _host is None:
        # DMPLEX_HOST_LIBRARY selects another build of the same library (the sanitizer build)
        path = Path(os.environ.get("DMPLEX_HOST_LIBRARY") or LIB_DIR / "libdmplex_host.so")
        if not path.exists():
            raise NativeError(f"{path} is missing: run `python -m dmplex_amd.build`")
        lib = C.CDLL(str(path))
        lib.dmx_last_error.restype = C.c_char_p
        sig = {
            "dmx_icosahedral": [C.c_int, C.c_int, f64p, i64p, i64p],
            "dmx_lloyd": [f64p, i64, i64p, i64, i64, C.c_double, C.c_int, i64, i64p],
            "dmx_lloyd_residual": [f64p, i64, i64p, i64, i64, f64p],
            "dmx_icosahedron": [f64p, i64p],
            "dmx_bisect": [i64, f64p, i64, i64p, f64p, i64p],
            "dmx_triangle_stars": [i64, i64p, i64, i64p, i64p],
            "dmx_delaunay": [f64p, i64, i64p, i64],
            "dmx_mean_edge_length": [f64p, i64p, i64, f64p],
            "dmx_plex_from_polygons": [i64, i64p, i64p, i64, f64p, C.POINTER(C.c_void_p), i64p],
            "dmx_plex_export": [C.c_void_p, i64p, i64p, i8p, i64p, i64p],
            "dmx_dual_polygons": [C.POINTER(DmxPlex), i64p, i64p, f64p],
            "dmx_orthogonality_defect": [C.POINTER(DmxPlex), f64p, f64p],
            "dmx_check_plex": [C.POINTER(DmxPlex)],
            "dmx_trisk_grid_widths": [C.POINTER(DmxPlex), i64p],
            "dmx_trisk_grid": [C.POINTER(DmxPlex), f64p, C.POINTER(DmxGridTables)],
            "dmx_vertex_edge_areas": [i64, i64, i64, C.POINTER(DmxGridTables), f64p, f64p, f64p, f64p],
            "dmx_gassmann_coefficients": [i64, i64, i64, C.POINTER(DmxGridTables), f64p, i64p, i64p, f64p],
            "dmx_hilbert_keys": [i64, u64p, C.c_int, u64p],
            "dmx_hilbert_permutation": [i64, f64p, C.c_int, i64p],
            "dmx_streamfunction": [C.c_int, f64p, i64, f64p, f64p],
            "dmx_nc_open": [C.c_char_p, C.POINTER(C.c_void_p)],
            "dmx_nc_format": [C.c_void_p],
            "dmx_nc_var_name": [C.c_void_p, i64, C.c_char_p, i64],
            "dmx_nc_has_var": [C.c_void_p, C.c_char_p],
            "dmx_nc_var_info": [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p, C.c_int],
            "dmx_nc_read_f64": [C.c_void_p, C.c_char_p, f64p, i64],
            "dmx_nc_read_i64": [C.c_void_p, C.c_char_p, i64p, i64],
            "dmx_nc_global_text": [C.c_void_p, C.c_char_p, C.c_char_p, i64, i64p],
            "dmx_sort_rings": [i64, i64p, i64p, i64, f64p, f64p],
            "dmx_rbf_coefficients": [i64, i64, C.POINTER(DmxGridTables), C.c_double, i64, i64p, f64p, i64p],
            "dmx_bilinear_weights": [i64, C.POINTER(DmxGridTables), f64p, f64p],
            "dmx_trisk_pv_edges": [i64, i64, C.POINTER(DmxGridTables), i64p],
            "dmx_iap_trisk_weights": [i64, i64, C.POINTER(DmxGridTables), f64p, f64p],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        lib.dmx_plex_free.argtypes = [C.c_void_p]
        lib.dmx_plex_free.restype = None
        lib.dmx_nc_close.argtypes = [C.c_void_p]
        lib.dmx_nc_close.restype = None
        lib.dmx_nc_num_vars.argtypes = [C.c_void_p]
        lib.dmx_nc_num_vars.restype = i64
        _host = lib
    return _host


def check(rc: int, lib=None, warn_ok: bool = False) -> int:
    """Raises on a negative status of a host call (positive codes are warnings)."""
    lib = lib or host()
    if rc < 0 or (rc > 0 and not warn_ok):
        msg = lib.dmx_last_error().decode()
        if rc in (-2, -3):
            raise ValueError(msg)
        if rc == -5:
            raise FileNotFoundError(msg)
        raise NativeError(f"native call failed ({rc}): {msg}")
    return rc
