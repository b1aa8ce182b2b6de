"""The C ABI: both libraries load, export every function their headers declare, and the
device library validates tables on the host (no GPU needed for these checks)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from dmplex_amd import ModelTables, create_hexagon_mesh
from dmplex_amd import _device as D
from dmplex_amd import _native as N
from dmplex_amd.synthetic import earth_parameters

ROOT = Path(__file__).resolve().parents[1]


def declared(header: str, prefix: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(rf"\b({prefix}\w+)\s*\(", text)))


@pytest.mark.parametrize("header,prefix,loader", [("trisk_hip.h", "(?:trisk|icon|mpas)_", D.lib),
                                                  ("dmplex_host.h", "dmx_", N.host)])
def test_every_declared_symbol_is_exported(header, prefix, loader):
    names = declared(header, prefix)
    assert len(names) > 5
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    names = declared("trisk_hip.h", "(?:trisk|icon|mpas)_")
    assert set(names) - {"trisk_last_error", "trisk_handle_last_error"} == set(D.SIGNATURES)


@pytest.mark.parametrize("name,struct", [("trisk_model_desc", D.TriskModelDesc), ("icon_model_desc", D.IconModelDesc),
                                         ("mpas_ops_desc", D.MpasOpsDesc)])
def test_desc_struct_matches_header_field_order(name, struct):
    text = (ROOT / "include" / "trisk_hip.h").read_text()
    end = text.index("} " + name + ";")
    body = text[text.rindex("typedef struct {", 0, end):end]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";")[:-1]:
        decl = decl.replace("typedef struct {", "").strip()
        for part in re.split(r",", decl):
            fields.append(re.sub(r"[*\s]", " ", part).split()[-1])
    assert fields == [f[0] for f in struct._fields_]


def test_check_tables_accepts_and_rejects_on_the_host():
    dm = create_hexagon_mesh(2, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, **earth_parameters())
    ok, why = t.check_tables()
    assert ok, why
    bad = t.cell_edges.copy(order="F")
    bad[3, 10] = dm.ne + 1
    t.cell_edges = bad
    with pytest.raises(ValueError, match="out of range in cell_edges"):
        t.check_tables()
    t2 = ModelTables(dm, centers=dm.generators, **earth_parameters())
    t2.trisk = np.asfortranarray(t2.trisk * (1 + 1e-9))
    ok, why = t2.check_tables()
    assert not ok and "trisk" in why


def test_create_without_a_device_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    dm = create_hexagon_mesh(1)
    t = ModelTables(dm)
    desc = t.descriptor()
    h = C.c_void_p()
    rc = D.lib().trisk_create(C.byref(desc), None, C.byref(h))
    assert rc == -2 and not h.value
    assert "device" in D.lib().trisk_last_error().decode()


def _struct_fields(header_text: str, name: str):
    body = header_text[:header_text.index(f"}} {name};")]
    body = body[body.rindex("typedef struct {") + len("typedef struct {"):]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";")[:-1]:
        for part in decl.split(","):
            fields.append(re.sub(r"\[.*?\]", "", part).replace("*", " ").split()[-1])
    return fields


def test_options_and_schedule_structs_match_the_header():
    from dmplex_amd.ensemble import TriskSchedule

    text = (ROOT / "include" / "trisk_hip.h").read_text()
    assert _struct_fields(text, "trisk_options") == [f[0] for f in D.TriskOptions._fields_]
    assert _struct_fields(text, "trisk_schedule") == [f[0] for f in TriskSchedule._fields_]
    assert C.sizeof(D.TriskOptions) == 32 and C.sizeof(TriskSchedule) == 16


def _field_sizes(desc):
    nc, ne, nv, kc, kv, kt, kg = desc.nc, desc.ne, desc.nv, desc.kc, desc.kv, desc.kt, desc.kg
    return dict(cell_edges=kc * nc, divergence=kc * nc, kinetic=kc * nc, edge_cells=2 * ne, edge_vertices=2 * ne,
                edge_height=2 * ne, inverse_dual=ne, vertex_edges=kv * nv, curl=kv * nv, vertex_cells=kv * nv,
                vertex_height=kv * nv, coriolis=nv, trisk_edges=kt * ne, trisk=kt * ne, cell_areas=nc,
                vertex_areas=nv, cell_vertices=kc * nc, vk_kites=kc * nc, vk_weights=kv * nv, vk_height=2 * ne,
                g_edges=kg * ne, g_pv_edges=kg * ne, g_weights=kg * ne, g_rhombus=2 * ne, g_coriolis=ne,
                cell_points=3 * nc, primal_lengths=ne, dual_lengths=ne, cell_kites=kc * nc)


@pytest.mark.parametrize("pv_flux,kinetic_energy", [("trisk", "cell"), ("gassmann", "vertex"),
                                                    ("gassmann_mean", "mixed")])
def test_julia_extension_descriptor(pv_flux, kinetic_energy):
    """The descriptor the Julia extension builds from the reference struct (1-based tables,
    0-based copies, every Gassmann and vertex-kinetic-energy table) passes the library's
    host validation and carries the same data, field by field, as the descriptor of the
    Python binding; the cell-pair path is offered exactly for the TRiSK flux with cell KE."""
    from julia_descriptor import descriptor, julia_struct

    dm = create_hexagon_mesh(3, centroidal=True)
    t = ModelTables(dm, centers=dm.generators, pv_flux=pv_flux, kinetic_energy=kinetic_energy, **earth_parameters())
    desc, keep = descriptor(julia_struct(t))
    flag = C.c_int32(-1)
    D.check(D.lib().trisk_check_tables(C.byref(desc), C.byref(flag)))
    assert flag.value == (1 if (pv_flux, kinetic_energy) == ("trisk", "cell") else 0)
    ref = t.descriptor()
    for name, ctype in D.TriskModelDesc._fields_:
        a, b = getattr(desc, name), getattr(ref, name)
        if not hasattr(ctype, "contents"):
            assert a == b, name
            continue
        n = _field_sizes(desc)[name]
        assert bool(a) == bool(b), name
        if a:
            assert np.array_equal(np.ctypeslib.as_array(a, (n,)), np.ctypeslib.as_array(b, (n,))), name
    assert desc.kg == (12 if pv_flux != "trisk" else 0) and desc.pv_flux == D.PV_FLUX[pv_flux]
