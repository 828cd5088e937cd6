"""ctypes binding of libcylpic.so (include/cylpic.h).

Loads the in-tree HIP library and nothing else: there is no CPU fallback.
A missing or unloadable library raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcylpic.so")
HEADER = os.path.join(HERE, "..", "include", "cylpic.h")

_lib = None

_d, _i, _l, _p = C.c_double, C.c_int, C.c_int64, C.c_void_p
_pp = C.POINTER(C.c_void_p)

SIGNATURES = {
    "cylpic_init": [_pp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _p, _p, _i, _i],
    "cylpic_finalize": [_p],
    "cylpic_set_dt": [_p, _d],
    "cylpic_set_time": [_p, _d],
    "cylpic_get_restart_state": [_p, _p, _p, _p, _p, _p],
    "cylpic_set_restart_state": [_p, _d, _d, _d, _l, _i],
    "cylpic_set_species": [_p, _i, _d, _d, _i, _l],
    "cylpic_upload_particles": [_p, _i, _l, _p, _p, _p, _p, _p, _p, _p, _p],
    "cylpic_download_particles": [_p, _i, _l, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "cylpic_particle_count": [_p, _i, _p],
    "cylpic_inject_particles": [_p, _i, _l, _p, _p, _p, _p, _p, _p, _p, _p],
    "cylpic_move_window": [_p, _i],
    "cylpic_set_smoothing": [_p, _i, _i, _p, _i],
    "cylpic_number_density_modes": [_p, _i, _p],
    "cylpic_continuity_begin": [_p],
    "cylpic_autotune": [_p, _i, _p, _p, _p, _p],
    "cylpic_continuity_end": [_p, _p, _p, _p],
    "cylpic_continuity_modes": [_p, _p, _p, _p, _p],
    "cylpic_gauss_residual_m0": [_p, _p, _p, _p],
    "cylpic_gauss_residual_m0_noalloc": [_p, _p, _p, _p],
    "cylpic_dist_fn": [_p, _i, _i, _p, _p, _p, _p, _p, _p, C.c_int64, _p],
    "cylpic_upload_field": [_p, _i, _p],
    "cylpic_download_field": [_p, _i, _p],
    "cylpic_setup_field_boundaries": [_p],
    "cylpic_update_eb_fields_half": [_p],
    "cylpic_push_particles": [_p],
    "cylpic_current_finish": [_p],
    "cylpic_update_eb_fields_final": [_p, _p, _p],
    "cylpic_step": [_p, _p, _p],
    "cylpic_sort_particles": [_p],
    "cylpic_set_sort_interval": [_p, _i],
    "cylpic_set_sort_method": [_p, _i],
    "cylpic_sort_counts": [_p, _p, _p],
    "cylpic_set_tile": [_p, _i, _i],
    "cylpic_tune": [_p, _i, _i],
    "cylpic_mode_energy": [_p, _p, _p],
    "cylpic_synchronize": [_p],
    "cylpic_profile": [_p, _i],
    "cylpic_kernel_time": [_p, C.c_char_p, _p, _p],
    "cylpic_batch_create": [_pp, _p, _i],
    "cylpic_batch_step": [_p, _p, _p],
    "cylpic_batch_steps": [_p, _i, _p],
    "cylpic_batch_synchronize": [_p],
    "cylpic_batch_stats": [_p, _p, _p, _p, _p],
    "cylpic_batch_destroy": [_p],
}


class CylpicError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CylpicError(f"{LIB_PATH} is missing: build it with "
                              "`python -m cylindrical_epoch_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.cylpic_last_error.argtypes = []
        L.cylpic_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().cylpic_last_error().decode(errors="replace")
        raise CylpicError(f"cylpic error {rc}: {msg}")
