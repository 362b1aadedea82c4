"""CPU FP64 oracle for the CMM-S2 hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the algorithms of the reference
Characteristic Mapping Method code (seth-tl/CMM-S2, ``cmm_s2/core``).  Every
function cites the reference file:line it follows.  It is the checker for the
HIP product path and the timed CPU baseline of ``bench.py``; only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
The product package (``cmm-s2_amd/cmm_s2``) never imports it.

Pinning (how each part is checked against something other than itself) is
listed per module in the module docstrings.  Parts marked "parity unpinned"
have no independent check beyond restatement.

Modules
  geometry   -- core/utils.py primitives, mesh_functions.sphere_tan_proj
  meshes     -- icosahedral point sets, spherical_triangulation (Delaunay
                find_simplex location), ps_split, stencil_pts, the MWSS
                structured mesh and its closed-form location
  splines    -- Powell-Sabin assembly / evaluation (map, scalar, velocity),
                stencil evaluation, composite maps
  evolution  -- backward integrators, spline_proj_sphere, advect, compose
  sht        -- MWSS spherical-harmonic transforms (own documented algorithm),
                angular momentum, inv_Laplacian, project_stream,
                project_onto_S12, energy spectrum
  fields     -- analytic velocity / vorticity plugin fields
  driver     -- euler_simulation_sphere_remapping
"""
