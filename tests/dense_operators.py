"""Dense operators of the TRiSK and ICON models assembled from the grid geometry alone
(edge points, normals, lengths, areas, kites), independently of the folded gather
tables the oracle and the device read.  Brute force for small meshes (levels 0-3)."""
import numpy as np


def tangent_signs(g):
    """sign of t_{e,v} for the two vertices of each edge: +1 where n_e . (x_v x x_e) > 0."""
    a, b = g.edge_vertices
    out = []
    for v in (a, b):
        out.append(np.sign(np.sum(g.edge_normals * np.cross(g.vertices[v], g.edge_points), axis=1)))
    return out


def operators(g):
    """D (nc x ne, n l / A), Kmat (nc x ne, l d_{e,i} / 2A), Curl (nv x ne, t d / A_v),
    Hv (nv x nc, K_{i,v} / A_v) and W (ne x ne, the TRiSK weights w_{e,e'})."""
    nc, ne, nv = g.num_cells(), g.num_edges(), g.num_vertices()
    l, d, A, Av = g.primal_lengths, g.dual_lengths, g.cell_areas, g.vertex_areas
    left, right = g.edge_cells
    E = np.arange(ne)
    D = np.zeros((nc, ne))
    D[left, E] += l / A[left]
    D[right, E] -= l / A[right]
    Kmat = np.zeros((nc, ne))
    Kmat[left, E] += l * g.cell_distances[0] / (2 * A[left])
    Kmat[right, E] += l * g.cell_distances[1] / (2 * A[right])
    Curl = np.zeros((nv, ne))
    for v, t_sign in zip(g.edge_vertices, tangent_signs(g)):
        Curl[v, E] += t_sign * d / Av[v]
    Hv = np.zeros((nv, nc))
    for k in range(g.cell_vertices.shape[0]):
        real = g.cell_signs[k] != 0
        np.add.at(Hv, (g.cell_vertices[k, real], np.arange(nc)[real]), g.cell_kites[k, real] / Av[g.cell_vertices[k, real]])
    W = np.zeros((ne, ne))
    if g.trisk_edges.size:
        for k in range(g.trisk_edges.shape[0]):
            np.add.at(W, (E, g.trisk_edges[k]), g.trisk_weights[k])
    return dict(D=D, K=Kmat, Curl=Curl, Hv=Hv, W=W)


def vertex_pv(ops, y, ne, coriolis):
    u, h = y[:ne], y[ne:]
    zeta = ops["Curl"] @ u
    hv = ops["Hv"] @ h
    return (zeta + coriolis) / hv, zeta, hv


def potential_enstrophy(g, ops, y, coriolis):
    """sum_v A_v h_v q_v^2 / 2 (TRiSKShallowWater.jl:414-419)."""
    qv, _, hv = vertex_pv(ops, y, g.num_edges(), coriolis)
    return float(np.sum(g.vertex_areas * hv * qv ** 2) / 2)


def total_energy(g, ops, y, gravity):
    """sum_i A_i h_i (K_i + g h_i / 2) (TRiSKShallowWater.jl:406-411)."""
    ne = g.num_edges()
    u, h = y[:ne], y[ne:]
    K = ops["K"] @ (u * u)
    return float(np.sum(g.cell_areas * h * (K + gravity * h / 2)))
