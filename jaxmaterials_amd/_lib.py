"""ctypes binding of liblippmannschwinger.so (declarations of include/lippmann_schwinger.h)."""

import ctypes
import os

from .build import LIB_PATH, LIB_NAME

_LIB = None

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); the full exported surface of the C ABI
SIGNATURES = {
    "lippmann_schwinger_solve": (
        ctypes.c_int,
        [ctypes.c_void_p] * 7 + [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int],
    ),
    "lippmann_schwinger_effective_tensor": (
        ctypes.c_int,
        [ctypes.c_void_p] * 4 + [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_solver_create": (
        ctypes.c_void_p,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_solver_solve": (
        ctypes.c_int,
        [ctypes.c_void_p] * 8 + [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_solver_autotune": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
    ),
    "ls_solver_last_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "ls_solver_destroy": (None, [ctypes.c_void_p]),
    "ls_crossply_lame": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong,
         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
    "ls_crossply_place": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong,
         ctypes.c_void_p, ctypes.c_int],
    ),
    "ls_relative_divergence": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_rasterise_layered_fibres": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_hbm_stream_probe": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ls_solve_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
        + [ctypes.c_float, ctypes.c_float, ctypes.c_int] + [ctypes.c_void_p] * 4,
    ),
    # phase-field fracture
    "ls_solver_set_damage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]),
    "ls_solver_phase_field": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_damaged_stress": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
         ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_strain_energy": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
    # reference fibre movement (include/fibres.h)
    "dist_periodic": (
        None,
        [ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "initialise_rng": (None, [ctypes.c_uint, ctypes.POINTER(ctypes.c_uint), ctypes.c_char_p]),
    "move_fibres": (
        None,
        [ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_double,
         ctypes.c_uint, ctypes.POINTER(ctypes.c_uint), ctypes.c_char_p],
    ),
    # learned warm start (include/ls_warmstart.h)
    "ls_ws_conv3d": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
         ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
         ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_conv3d_wgrad": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
         ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
    "ls_ws_pool": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_pool_backward": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
         ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_upsample": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_upsample_backward": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_input": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_ws_target": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_ws_mse": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_void_p,
         ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "ls_ws_symgrad": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ls_ws_symgrad_backward": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
         ctypes.c_void_p],
    ),
    "ls_ws_residual_loss": (
        ctypes.c_int,
        [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p],
    ),
    "ls_ws_adam": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_float,
         ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_void_p],
    ),
    "ls_ws_pack": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    ),
}


def library_path():
    return LIB_PATH


def load():
    """Load the in-tree library (or one named liblippmannschwinger.so on the
    loader path); raise RuntimeError if neither loads, like the reference
    loader (src/jaxmaterials/solver/lippmann_schwinger.py:285-290)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    errors = []
    for cand in (LIB_PATH, LIB_NAME):
        if cand == LIB_PATH and not os.path.exists(cand):
            errors.append(f"{cand}: not built")
            continue
        try:
            lib = ctypes.CDLL(cand)
            break
        except OSError as exc:
            errors.append(f"{cand}: {exc}")
    else:
        raise RuntimeError(
            "Unable to load HIP library liblippmannschwinger.so. Build it with "
            "`python -m jaxmaterials_amd.build` or check LD_LIBRARY_PATH. (" + "; ".join(errors) + ")"
        )
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
