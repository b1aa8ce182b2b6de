"""dmplex_amd: MI355X-native nonlinear TRiSK shallow-water step of dmplex.jl.

Public API (mirrors the reference's Julia names; ``!`` becomes ``_``):

* meshes: ``create_spherical_mesh``, ``create_hexagon_mesh``, ``lloyd_``, ``Plex``,
  ``plex_from_polygons``, ``check_plex``, ``hilbert_order``, ``hilbert_permutation``,
  ``cell_centers``, ``orthogonality_defect``;
* grid files: ``read_grid``, ``read_mpas_grid``, ``read_icon_grid`` (NetCDF classic,
  64-bit offset, 64-bit data and NetCDF-4), ``NetCDFFile``, ``sort_ring``, ``dual_polygons``;
* geometry: ``TRiSKGrid``, ``vertex_edge_areas``, ``gassmann_coefficients``;
* model: ``TRiSKShallowWater`` (``backend=HIPBackend(...)``), ``initial_state``,
  ``tendency_``, ``RK4Cache``, ``rk4_step_``, ``gather_state``, ``total_mass``,
  ``total_energy``, ``potential_enstrophy``, ``relative_vorticity``,
  ``tangential_velocity``, ``kinetic_energy_``, ``vertex_pv_``;
* adjoint: ``tendency_vjp``, ``rk4_adjoint`` (checkpointed), ``rrule_rk4_steps``,
  ``rk4_steps`` (a differentiable torch operation), ``height_misfit``;
* the ICON triangle model: ``ICONShallowWater`` (RBF tangential velocity, bilinear kinetic
  energy, divergence damping), ``rbf_coefficients``, ``bilinear_weights``; it takes the
  same state, stepping and diagnostic functions as ``TRiSKShallowWater``;
* ensembles: ``TRiSKEnsemble`` (members on meshes of several resolutions, one launch per
  kernel for all of them, with an autotuned schedule).

Compute runs in ``lib/libtrisk_hip.so`` (gfx950 kernels behind the C ABI of
``include/trisk_hip.h``); setup runs in ``lib/libdmplex_host.so``.
"""
from .ensemble import TRiSKEnsemble
from .file_grids import (NetCDFFile, dual_polygons, lonlat_to_xyz, read_grid, read_icon_grid, read_mpas_grid,
                         read_variable, rows_per_entity, sort_ring)
from .icon_model import ICONShallowWater, ICONTables, bilinear_weights, rbf_coefficients
from .grids import (cell_centers, create_hexagon_mesh, create_spherical_mesh, lloyd_, lloyd_residual,
                    orthogonality_defect, voronoi_generators)
from .topology import (Plex, check_plex, hilbert_key, hilbert_keys, hilbert_order, hilbert_permutation,
                       plex_from_polygons)
from .trisk_grid import TRiSKGrid, arc_length, gassmann_coefficients, signed_triangle_area, vertex_edge_areas
from .trisk_model import (HIPBackend, KineticWork, ModelTables, RK4Cache, TRiSKShallowWater, cell_areas_device,
                          diagnostics, gather_state, global_height_dofs, global_velocity_dofs, height_dofs,
                          height_misfit, initial_state, initial_state_values, kinetic_energy_,
                          potential_enstrophy, relative_vorticity, rk4_adjoint, rk4_step_, rk4_steps, rk4_steps_,
                          rrule_rk4_steps, tangential_velocity, tendency_, tendency_vjp, total_energy, total_mass,
                          velocity_dofs, vertex_pv_)

__all__ = [n for n in dir() if not n.startswith("_")]
