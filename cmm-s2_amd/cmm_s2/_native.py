"""ctypes binding of libcmm_s2_amd.so (the C ABI declared in include/cmm_s2.h).

There is no fallback: if the library is missing or the GPU is absent every
entry point raises.  Torch is used only to own device memory and to supply
the current HIP stream.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmm_s2_amd.so")

i64 = C.c_int64
vp = C.c_void_p
dbl = C.c_double
P_vp = C.POINTER(C.c_void_p)
P_dbl = C.POINTER(C.c_double)
P_i64 = C.POINTER(C.c_int64)

VEL_SPLINE, VEL_SBR, VEL_DEFORM_ROT, VEL_RH4 = 0, 1, 2, 3
VEL_DEFORM, VEL_STATIC_VORTEX, VEL_ROTATING_VORTEX, VEL_EQUATOR_VORTEX, VEL_DIV = 4, 5, 6, 7, 8
INT_EULER, INT_IMPROVED_EULER, INT_RK4, INT_RK2, INT_EVAL = 0, 1, 2, 3, 4
LEGENDRE_FP64, LEGENDRE_FP32 = 0, 1
LEGENDRE_BF16X3, LEGENDRE_BF16X4, LEGENDRE_BF16X5, LEGENDRE_BF16X6, LEGENDRE_BF16X7 = 2, 3, 4, 5, 6
LEGENDRE_I8X4, LEGENDRE_I8X5, LEGENDRE_I8X6, LEGENDRE_I8X7, LEGENDRE_I8X8 = 7, 8, 9, 10, 11
VORT_SPLINE, VORT_ROSSBY, VORT_ZONAL_JET, VORT_GAUSSIAN = 0, 1, 2, 3
VORT_PERTURBED_JET, VORT_MULTI_JET, VORT_PERTURBED_ROSSBY = 4, 5, 6
VORT_COSINE_BELL, VORT_SLOTTED_CYLINDER, VORT_STATIC_VORTEX_TRACER, VORT_ROTATING_VORTEX_TRACER = 7, 8, 9, 10
VORT_SH_SERIES = 11


class Velocity(C.Structure):
    """struct cmm_velocity."""
    _fields_ = [("kind", C.c_int), ("nlev", C.c_int), ("mesh", vp), ("levels", vp * 3),
                ("stage_w", (dbl * 3) * 4), ("stage_t", dbl * 4), ("params", dbl * 8),
                ("vbits", vp), ("vslot", C.c_int * 3), ("jets", vp * 3), ("mark", vp), ("mark_every", C.c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "cmm_abi_version": (C.c_int, []),
    "cmm_last_error": (C.c_char_p, []),
    "cmm_tri_mesh_create": (C.c_int, [P_dbl, i64, P_i64, i64, C.c_int, P_vp]),
    "cmm_tri_mesh_destroy": (C.c_int, [vp]),
    "cmm_tri_mesh_stats": (C.c_int, [vp, P_i64]),
    "cmm_tri_locate": (C.c_int, [vp, vp, i64, i64, i64, vp, vp, vp]),
    "cmm_tri_assemble": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_tri_eval": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, i64, i64, i64, vp, i64, i64, vp]),
    "cmm_tri_compose": (C.c_int, [vp, P_vp, C.c_int, vp, i64, i64, i64, vp, i64, i64, vp, vp]),
    "cmm_tri_eval_grad": (C.c_int, [vp, vp, C.c_int, vp, i64, i64, i64, vp, vp]),
    "cmm_tri_compose_grid": (C.c_int, [vp, P_vp, C.c_int, vp, i64, i64, C.c_int, i64, i64, vp, i64, i64, vp,
                                       vp]),
    "cmm_tri_compose_grid_located": (C.c_int, [vp, P_vp, C.c_int, vp, i64, i64, C.c_int, i64, i64, vp, i64, i64,
                                               vp, vp, vp, vp]),
    "cmm_tri_split_locate": (C.c_int, [vp, vp, i64, i64, i64, vp, vp, vp]),
    "cmm_tri_compose_grad": (C.c_int, [vp, P_vp, C.c_int, vp, i64, i64, i64, vp, i64, i64, vp, vp]),
    "cmm_mwss_eval_grad": (C.c_int, [vp, vp, vp, i64, i64, i64, vp, vp]),
    "cmm_tri_stencil_jets": (C.c_int, [vp, vp, vp, i64, vp, i64, vp, dbl, vp, vp, vp]),
    "cmm_mwss_mesh_create": (C.c_int, [C.c_int, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_vp]),
    "cmm_mwss_mesh_destroy": (C.c_int, [vp]),
    "cmm_mwss_locate": (C.c_int, [vp, vp, i64, i64, i64, vp, vp, vp]),
    "cmm_mwss_split_locate": (C.c_int, [vp, vp, i64, i64, i64, vp, vp, vp, vp, vp, vp]),
    "cmm_mwss_assemble": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_mwss_assemble_sparse": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp]),
    "cmm_mwss_eval": (C.c_int, [vp, P_vp, P_dbl, C.c_int, C.c_int, vp, i64, i64, i64, vp, i64, i64, vp]),
    "cmm_advect_points": (C.c_int, [C.POINTER(Velocity), C.c_int, dbl, vp, i64, vp, vp]),
    "cmm_advect_stencil": (C.c_int, [C.POINTER(Velocity), C.c_int, dbl, vp, i64, vp, i64, i64, vp, i64, vp, i64,
                                     vp]),
    "cmm_sht_plan_create": (C.c_int, [C.c_int, P_dbl, P_dbl, P_vp]),
    "cmm_sht_plan_destroy": (C.c_int, [vp]),
    "cmm_sht_plan_gl": (C.c_int, [vp, P_dbl, P_dbl]),
    "cmm_sht_set_legendre": (C.c_int, [vp, C.c_int]),
    "cmm_sht_workspace_bytes": (C.c_int, [vp, C.c_int, P_i64]),
    "cmm_sht_inverse": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_sht_forward": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_sht_forward_noalloc": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "cmm_batch_create": (C.c_int, [i64, P_vp]),
    "cmm_batch_destroy": (C.c_int, [vp]),
    "cmm_batch_begin": (C.c_int, [vp]),
    "cmm_batch_member": (C.c_int, [vp, C.c_int]),
    "cmm_batch_cancel": (C.c_int, [vp]),
    "cmm_batch_launch": (C.c_int, [vp, vp, P_i64]),
    "cmm_copy_f64": (C.c_int, [vp, vp, i64, vp]),
    "cmm_sht_stream_operators": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "cmm_sht_scalar_operators": (C.c_int, [vp, vp, vp, vp]),
    "cmm_mwss_velocity_jets": (C.c_int, [vp, vp, vp, vp]),
    "cmm_mwss_scalar_jets": (C.c_int, [vp, vp, vp, vp]),
    "cmm_sample_vorticity": (C.c_int, [C.c_int, vp, vp, P_dbl, vp, i64, i64, i64, vp, C.c_int, dbl, vp, vp]),
    # adjoints
    "cmm_sht_inverse_adjoint": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_sht_forward_adjoint": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_sht_stream_operators_adjoint": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "cmm_sht_scalar_operators_adjoint": (C.c_int, [vp, vp, vp, vp]),
    "cmm_mwss_velocity_jets_adjoint": (C.c_int, [vp, vp, vp, vp]),
    "cmm_mwss_scalar_jets_adjoint": (C.c_int, [vp, vp, vp, vp]),
    "cmm_mwss_assemble_adjoint": (C.c_int, [vp, vp, C.c_int, vp, vp, i64, vp]),
    "cmm_tri_assemble_adjoint": (C.c_int, [vp, vp, vp, vp, i64, vp]),
    "cmm_advect_stencil_adjoint": (C.c_int, [C.POINTER(Velocity), C.c_int, dbl, vp, i64, i64, vp, i64, P_vp, vp,
                                             i64, vp, i64, vp]),
    "cmm_advect_stencil_adjoint_corners": (C.c_int, [C.POINTER(Velocity), C.c_int, dbl, vp, i64, i64, vp, i64, P_vp,
                                                     P_vp, vp, i64, vp, i64, vp]),
    "cmm_mwss_corner_jets_adjoint": (C.c_int, [vp, vp, vp, vp, vp]),
    "cmm_tri_stencil_jets_adjoint": (C.c_int, [vp, vp, vp, i64, vp, i64, vp, dbl, vp, vp, i64, vp, vp, i64, vp]),
    "cmm_tri_compose_adjoint": (C.c_int, [vp, P_vp, P_vp, C.c_int, vp, i64, i64, i64, vp, vp, vp, i64, vp]),
    "cmm_sample_vorticity_adjoint": (C.c_int, [vp, vp, vp, i64, i64, i64, vp, dbl, vp, i64, i64, vp, vp, i64, vp]),
    "cmm_scatter_workspace_bytes": (C.c_int, [i64, i64, P_i64]),
}


class CmmError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CmmError(f"{LIB_PATH} is missing: build it with "
                           "`python -m cmm_s2._build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().cmm_last_error().decode(errors="replace")
        raise CmmError(f"{what or 'cmm'} failed ({rc}): {msg}")


_fns = {}


def call(name, *args):
    f = _fns.get(name)
    if f is None:
        f = _fns[name] = getattr(lib(), name)
    rc = f(*args)
    if rc != 0:
        check(rc, name)
    return rc


def dptr(a):
    """ctypes double* of a host numpy float64 array (kept alive by the caller)."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_dbl)


def lptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_i64)


def ptr_array(ptrs):
    arr = (C.c_void_p * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def dbl_array(vals):
    arr = (C.c_double * max(1, len(vals)))()
    for i, v in enumerate(vals):
        arr[i] = float(v)
    return arr
