"""The trisk_model_desc that the Julia extension's ``descriptor(m)`` builds from the
reference's TRiSKShallowWater struct, restated with ctypes: the same positional constructor
call, field by field, from the same struct fields (TRiSKShallowWater.jl:67-92), with the
1-based Julia index tables turned into 0-based copies (``zb``) and every Gassmann and
vertex-kinetic-energy table passed."""
import ctypes as C

import numpy as np

from dmplex_amd import _device as D


def zb(table):
    """``zb(t) = Int64.(t) .- 1``: 0-based copy of a 1-based Julia index table (k x n, column-major)."""
    return np.asfortranarray(np.asarray(table, dtype=np.int64) - 1)


def julia_struct(t):
    """The reference struct's fields as Julia holds them: index tables 1-based."""
    one = lambda a: None if a is None else np.asfortranarray(a + 1)  # noqa: E731
    vk, gm = t.vertex_kinetic, t.gassmann
    return dict(
        grid=t.grid, cell_edges=one(t.cell_edges), divergence=t.divergence, kinetic=t.kinetic,
        edge_cells=one(t.edge_cells), edge_vertices=one(t.edge_vertices), edge_height=t.edge_height,
        inverse_dual=t.inverse_dual, vertex_edges=one(t.vertex_edges), curl=t.curl,
        vertex_cells=one(t.vertex_cells), vertex_height=t.vertex_height, coriolis=t.coriolis,
        trisk_edges=one(t.trisk_edges), trisk=t.trisk, cell_areas=t.cell_areas, vertex_areas=t.vertex_areas,
        gravity=t.gravity,
        vertex_kinetic=None if vk is None else dict(cell_vertices=one(vk["cell_vertices"]), kites=vk["kites"],
                                                    weights=vk["weights"], height=vk["height"]),
        gassmann=None if gm is None else dict(edges=one(gm["edges"]), pv_edges=one(gm["pv_edges"]),
                                              weights=gm["weights"], rhombus=gm["rhombus"],
                                              coriolis=gm["coriolis"], mean=gm["mean"]))


def descriptor(m):
    """(desc, keep): the positional TriskModelDesc(...) call of the Julia ``descriptor(m)``."""
    def p(a, t=C.c_double):
        return None if a is None else a.ctypes.data_as(C.POINTER(t))

    i64 = C.c_int64
    g, gm, vk = m["grid"], m["gassmann"], m["vertex_kinetic"]
    keep = (zb(m["cell_edges"]), zb(m["edge_cells"]), zb(m["edge_vertices"]), zb(m["vertex_edges"]),
            zb(m["vertex_cells"]), zb(m["trisk_edges"]), None if vk is None else zb(vk["cell_vertices"]),
            None if gm is None else zb(gm["edges"]), None if gm is None else zb(gm["pv_edges"]))
    desc = D.TriskModelDesc(
        g.num_cells(), g.num_edges(), g.num_vertices(), m["cell_edges"].shape[0],
        m["vertex_edges"].shape[0], m["trisk_edges"].shape[0], 0 if gm is None else gm["edges"].shape[0],
        p(keep[0], i64), p(m["divergence"]), p(m["kinetic"]), p(keep[1], i64), p(keep[2], i64),
        p(m["edge_height"]), p(m["inverse_dual"]), p(keep[3], i64), p(m["curl"]), p(keep[4], i64),
        p(m["vertex_height"]), p(m["coriolis"]), p(keep[5], i64), p(m["trisk"]), p(m["cell_areas"]),
        p(m["vertex_areas"]), m["gravity"],
        p(keep[6], i64), None if vk is None else p(vk["kites"]),
        None if vk is None else p(vk["weights"]), None if vk is None else p(vk["height"]),
        p(keep[7], i64), p(keep[8], i64), None if gm is None else p(gm["weights"]),
        None if gm is None else p(gm["rhombus"]), None if gm is None else p(gm["coriolis"]),
        0 if gm is None else 2 if gm["mean"] else 1, 0,
        p(g.centers), p(g.primal_lengths), p(g.dual_lengths), p(g.cell_kites))
    return desc, keep
