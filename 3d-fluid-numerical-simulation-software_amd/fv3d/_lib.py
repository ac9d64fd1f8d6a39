"""ctypes binding of libfv3d.so (the HIP library built in-tree for gfx950).

The library is loaded from ``lib/libfv3d.so`` next to this package; if it is
missing the import fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfv3d.so")

# every symbol declared in include/fv3d.h
EXPORTS = [
    "fv3d_ctx_create", "fv3d_ctx_destroy", "fv3d_last_error", "fv3d_block_add", "fv3d_patch_add",
    "fv3d_mixing_plane_add", "fv3d_connection_add", "fv3d_setup", "fv3d_levels", "fv3d_instance_add", "fv3d_instances",
    "fv3d_instance_results", "fv3d_initial",
    "fv3d_state_upload", "fv3d_nut_upload", "fv3d_field_download", "fv3d_level_field_download", "fv3d_bc_apply", "fv3d_spectral", "fv3d_timestep", "fv3d_convective",
    "fv3d_viscous", "fv3d_curl", "fv3d_resflux", "fv3d_lusgs", "fv3d_res_calculate", "fv3d_perf", "fv3d_iterate", "fv3d_iterate_graph", "fv3d_sync",
    "fv3d_set_timing", "fv3d_timing", "fv3d_set_debug", "fv3d_set_tiles", "fv3d_autotune", "fv3d_cell_count",
    "fv3d_mlp_set", "fv3d_mlp_get", "fv3d_mlp_features", "fv3d_mlp_targets", "fv3d_mlp_forward", "fv3d_mlp_grad", "fv3d_mlp_grad_noalloc", "fv3d_mlp_train",
    "fv3d_mlp_closure", "fv3d_mlp_relax", "fv3d_mlp_tripped",
    "fv3d_primitive_vars", "fv3d_limit_vars", "fv3d_mg_coarse_lusgs", "fv3d_mg_interpolate_state", "fv3d_mg_restrict",
    "fv3d_mg_correct", "fv3d_iteration_end", "fv3d_turbulence", "fv3d_wall_distance_upload",
]


class Config(C.Structure):
    _fields_ = [("rgas", C.c_double), ("gamma", C.c_double), ("pr_lam", C.c_double), ("pr_tur", C.c_double),
                ("pt_in", C.c_double), ("tt_in", C.c_double), ("p_out", C.c_double), ("ma0", C.c_double),
                ("len_ref", C.c_double), ("ev_ratio", C.c_double), ("inlet_dir", C.c_double * 3),
                ("cfl", C.c_double), ("c_timestep", C.c_double), ("relax", C.c_double),
                ("turbo_type", C.c_int), ("viscous", C.c_int), ("recon_var", C.c_int), ("recon_scheme", C.c_int),
                ("flux_scheme", C.c_int), ("turbulence", C.c_int), ("limiter", C.c_int), ("global_dt", C.c_int),
                ("device", C.c_int), ("precision", C.c_int),
                ("multigrid", C.c_int), ("coarse_iter", C.c_int), ("initial_type", C.c_int),
                ("p21ratio", C.c_double), ("outlet_dir", C.c_double * 3), ("mg_levels", C.c_int)]


class Fv3dError(RuntimeError):
    pass


_lib = None


def load():
    """Loads libfv3d.so and declares the C ABI; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Fv3dError(f"libfv3d.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i, d, l, p = C.c_void_p, C.c_int, C.c_double, C.c_long, C.c_void_p
    sig = {
        "fv3d_ctx_create": (i, [C.POINTER(Config), C.POINTER(vp)]),
        "fv3d_ctx_destroy": (i, [vp]),
        "fv3d_last_error": (C.c_char_p, [vp]),
        "fv3d_block_add": (i, [vp, i, i, i, p, i, d]),
        "fv3d_patch_add": (i, [vp, i, i, i, p, i]),
        "fv3d_connection_add": (i, [vp, i, i, i, i, p, i, i, p]),
        "fv3d_mixing_plane_add": (i, [vp, i, p, i, p]),
        "fv3d_setup": (i, [vp]),
        "fv3d_state_upload": (i, [vp, i, p, p]),
        "fv3d_nut_upload": (i, [vp, i, p]),
        "fv3d_field_download": (l, [vp, i, C.c_char_p, p]),
        "fv3d_level_field_download": (l, [vp, i, i, C.c_char_p, p]),
        "fv3d_levels": (i, [vp]),
        "fv3d_instance_add": (i, [vp, d]),
        "fv3d_instances": (i, [vp]),
        "fv3d_instance_results": (i, [vp, i, p, p]),
        "fv3d_res_calculate": (i, [vp, p]),
        "fv3d_perf": (i, [vp, p]),
        "fv3d_iterate": (i, [vp, i, p]),
        "fv3d_iterate_graph": (i, [vp, i, p]),
        "fv3d_set_timing": (i, [vp, i]),
        "fv3d_timing": (i, [vp, p, p]),
        "fv3d_set_debug": (i, [vp, i]),
        "fv3d_set_tiles": (i, [vp, p]),
        "fv3d_autotune": (i, [vp, i, p]),
        "fv3d_mlp_set": (i, [vp, p, p]),
        "fv3d_mlp_get": (i, [vp, p]),
        "fv3d_mlp_features": (l, [vp, i, p]),
        "fv3d_mlp_targets": (l, [vp, i, p]),
        "fv3d_mlp_forward": (i, [vp, p, l, p]),
        "fv3d_mlp_grad": (i, [vp, p, p, l, p, p]),
        "fv3d_mlp_grad_noalloc": (i, [vp, vp, vp, l, vp, vp]),
        "fv3d_mlp_train": (i, [vp, p, p, l, i, i, C.c_float, p]),
        "fv3d_mlp_closure": (i, [vp, i]),
        "fv3d_mlp_relax": (i, [vp, C.c_double]),
        "fv3d_primitive_vars": (i, [vp]),
        "fv3d_limit_vars": (i, [vp]),
        "fv3d_mg_coarse_lusgs": (i, [vp, i]),
        "fv3d_mg_interpolate_state": (i, [vp]),
        "fv3d_mg_restrict": (i, [vp]),
        "fv3d_mg_correct": (i, [vp]),
        "fv3d_iteration_end": (i, [vp, p]),
        "fv3d_turbulence": (i, [vp]),
        "fv3d_wall_distance_upload": (i, [vp, i, p]),
        "fv3d_mlp_tripped": (l, [vp]),
        "fv3d_cell_count": (l, [vp]),
    }
    for name in ("fv3d_bc_apply", "fv3d_spectral", "fv3d_timestep", "fv3d_convective", "fv3d_viscous", "fv3d_curl",
                 "fv3d_resflux", "fv3d_lusgs", "fv3d_sync", "fv3d_initial"):
        sig[name] = (i, [vp])
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
