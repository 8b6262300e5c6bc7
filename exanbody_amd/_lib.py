"""ctypes declarations of libxnb_amd.so (include/xnb_amd.h).

The library is built in-tree by exanbody_amd/build.py.  Loading fails loudly when it is
missing: there is no CPU fallback for any operator.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libxnb_amd.so")

XNB_OK = 0
XNB_E_ARG = -1
XNB_E_DOMAIN = -2
XNB_E_OUTSIDE = -3
XNB_E_STATE = -4
XNB_E_HIP = -5
XNB_E_NOMEM = -6
XNB_E_OVERFLOW = -7
XNB_E_CAPACITY = -9

PRECISION_FP64 = 0
PRECISION_MIXED = 1

TIMER_FORCE, TIMER_NBH, TIMER_PUSH, TIMER_GHOST, TIMER_REBIN, TIMER_TRIGGER = range(6)


class Domain(C.Structure):
    _fields_ = [("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3), ("grid_dims", C.c_int64 * 3),
                ("cell_size", C.c_double), ("periodic", C.c_int * 3)]


class LJParams(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("sigma", C.c_double)]


class ChunkNeighborsConfig(C.Structure):
    _fields_ = [("build_particle_offset", C.c_int), ("dual_particle_offset", C.c_int), ("random_access", C.c_int),
                ("half_symmetric", C.c_int), ("skip_ghosts", C.c_int)]


class RunParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("lj", LJParams), ("rcut", C.c_double), ("rcut_inc", C.c_double),
                ("precision", C.c_int), ("rebuild_every", C.c_int), ("steps", C.c_int64),
                ("n_rebuilds", C.c_int64), ("n_violations", C.c_int64)]


# name -> (restype, argtypes); mirrors include/xnb_amd.h
P = C.c_void_p
_SIGS = {
    "xnb_version": (C.c_int, []),
    "xnb_error_string": (C.c_char_p, [C.c_int]),
    "xnb_last_error": (C.c_char_p, [P]),
    "xnb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "xnb_create": (C.c_int, [C.POINTER(P), C.POINTER(Domain), C.c_int]),
    "xnb_destroy": (C.c_int, [P]),
    "xnb_nbh_dist": (C.c_int, [P, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    "xnb_set_particles": (C.c_int, [P, C.c_size_t, P, P, P, P]),
    "xnb_set_particles_with_fields": (C.c_int, [P, C.c_size_t, P, P, P, P, C.c_int, P]),
    "xnb_get_fields": (C.c_int, [P, C.c_int, C.c_int, P]),
    "xnb_field_count": (C.c_int, [P, C.POINTER(C.c_int)]),
    "xnb_num_particles": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xnb_grid_info": (C.c_int, [P, P, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "xnb_get_particles": (C.c_int, [P, C.c_int, P, P, P, P, P, P, P]),
    "xnb_get_cell_counts": (C.c_int, [P, P]),
    "xnb_move_particles": (C.c_int, [P]),
    "xnb_backup_r": (C.c_int, [P]),
    "xnb_rebuild_amr": (C.c_int, [P, C.c_double, C.c_int]),
    "xnb_ghost_update_all": (C.c_int, [P]),
    "xnb_ghost_update_r": (C.c_int, [P]),
    "xnb_particle_displ_over": (C.c_int, [P, C.c_double, C.POINTER(C.c_uint64)]),
    "xnb_chunk_neighbors": (C.c_int, [P, C.c_double, C.POINTER(ChunkNeighborsConfig), C.c_int]),
    "xnb_chunk_neighbors_stream": (C.c_int, [P, P, C.c_size_t, P, C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "xnb_zero_particle_force": (C.c_int, [P, C.c_int]),
    "xnb_divide_force_by_type_scalar": (C.c_int, [P, C.c_int, P, C.c_int]),
    "xnb_lennard_jones_force": (C.c_int, [P, C.POINTER(LJParams), C.c_double, C.c_int, C.c_int]),
    "xnb_lennard_jones_force_multimat": (C.c_int, [P, C.c_int, P, P, P, C.c_int]),
    "xnb_gravitational_force": (C.c_int, [P, C.c_double, C.c_int, P, C.c_double, C.c_int]),
    "xnb_push_f_v_r": (C.c_int, [P, C.c_double, C.c_double]),
    "xnb_push_f_v": (C.c_int, [P, C.c_double, C.c_double]),
    "xnb_simulation_energies": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "xnb_simulation_energies_noalloc": (C.c_int, [P, P]),
    "xnb_init": (C.c_int, [P, C.POINTER(LJParams), C.c_double, C.c_double, C.c_int]),
    "xnb_run": (C.c_int, [P, C.POINTER(RunParams)]),
    "xnb_run_multimat": (C.c_int, [P, C.POINTER(RunParams), C.c_int, P, P, P]),
    "xnb_run_step_counts": (C.c_int, [P, P, C.c_size_t]),
    "xnb_timer_enable": (C.c_int, [P, C.c_int]),
    "xnb_timer_reset": (C.c_int, [P]),
    "xnb_timer_read": (C.c_int, [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "xnb_pair_counts": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "xnb_synchronize": (C.c_int, [P]),
    "xnb_device_bytes": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xnb_copy_probe": (C.c_int, [P, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "xnb_fp64_core": (C.c_int, [P, P, C.c_size_t, P, P]),
    "xnb_autotune": (C.c_int, [P, C.POINTER(LJParams), C.c_double, C.c_int, P]),
    "xnb_snapshot_size": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "xnb_snapshot_save": (C.c_int, [P, P, C.c_size_t]),
    "xnb_snapshot_load": (C.c_int, [P, P, C.c_size_t]),
    "xnb_batch_create": (C.c_int, [C.POINTER(P), C.c_int]),
    "xnb_batch_destroy": (C.c_int, [P]),
    "xnb_batch_error": (C.c_char_p, [P]),
    "xnb_batch_attach": (C.c_int, [P, P, C.POINTER(RunParams), C.POINTER(C.c_int)]),
    "xnb_batch_step": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
    "xnb_batch_box_status": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "xnb_batch_box_recoveries": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "xnb_batch_graph_stats": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "xnb_batch_autotune": (C.c_int, [P, C.POINTER(C.c_double)]),
    "xnb_batch_detach": (C.c_int, [P, C.c_int, C.POINTER(RunParams)]),
    "xnb_job_bytes": (C.c_size_t, [C.c_size_t]),
    "xnb_job_bytes_for": (C.c_size_t, [P]),
    "xnb_serve_run": (C.c_int, [P, C.c_int, P, P, P]),
    "xnb_deck_create": (C.c_int, [C.POINTER(P)]),
    "xnb_deck_destroy": (C.c_int, [P]),
    "xnb_deck_load": (C.c_int, [P, C.c_char_p]),
    "xnb_deck_set_particles": (C.c_int, [P, C.c_size_t, P, P, P, P]),
    "xnb_deck_run": (C.c_int, [P, C.c_char_p]),
    "xnb_deck_context": (P, [P]),
    "xnb_deck_error": (C.c_char_p, [P]),
    "xnb_deck_log": (C.c_char_p, [P]),
    "xnb_deck_slot": (C.c_int, [P, C.c_char_p, C.c_char_p, C.c_size_t]),
    "xnb_generate_fcc": (C.c_int, [C.c_int64, C.c_double, C.c_int64, C.c_double, C.c_double, C.c_uint64, P, P, P,
                                   C.POINTER(C.c_double)]),
}

_lib = None


class XnbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def load():
    """Load libxnb_amd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m exanbody_amd.build` "
                          "(the HIP library is the only implementation)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return list(_SIGS)
