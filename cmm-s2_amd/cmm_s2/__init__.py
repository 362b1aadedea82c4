"""MI355X-native Characteristic Mapping Method engine on the sphere.

Drop-in for the hot path of seth-tl/CMM-S2: ``cmm_s2.core`` keeps the
reference's module, class and function names (mesh_functions,
spherical_spline, evolution_functions, spherical_harmonic_tools,
spherical_simulations, dynamical_fields, utils) while every numerical kernel
runs as hand-written HIP code for gfx950 behind the C ABI in
include/cmm_s2.h (libcmm_s2_amd.so).
"""
__version__ = "0.1.0"
