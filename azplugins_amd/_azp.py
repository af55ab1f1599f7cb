"""ctypes binding of the C ABI declared in include/azp.h (libazp.so).

torch is imported first: its bundled HIP runtime carries the same SONAME
(libamdhip64.so.7) that libazp.so links against, so the process keeps exactly
one HIP runtime and device pointers from torch tensors are valid in the
library.  There is no CPU fallback: if the library is missing, importing this
module raises.
"""

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library shares)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libazp.so")

OK = 0
ERR_NAMES = {
    1: "AZP_ERR_INVALID",
    2: "AZP_ERR_HIP",
    3: "AZP_ERR_BOX",
    4: "AZP_ERR_UNSUPPORTED",
    5: "AZP_ERR_NOMEM",
    6: "AZP_ERR_STATE",
}

EVAL_HERTZ = 0
EVAL_EXPANDED_YUKAWA = 1
EVAL_PERTURBED_LJ = 2
EVAL_COLLOID = 3
EVAL_DPD_GENERAL_WEIGHT = 4
EVAL_TWO_PATCH_MORSE = 5

DRIVER_PAIR = 0
DRIVER_DPD_THERMO = 1
DRIVER_CONSERVATIVE = 2
DRIVER_ANISO = 3

SHIFT_MODES = {"none": 0, "shift": 1, "xplor": 2}
PRECISIONS = {"fp64": 0, "mixed": 1, "mixed_acc64": 2}


class AzpError(RuntimeError):
    """Error returned through the C ABI."""

    def __init__(self, code, msg):
        self.code = code
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")


class Box(ctypes.Structure):
    _fields_ = [("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double)]


class Particles(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_uint32),
        ("box", Box),
        ("d_pos", ctypes.c_void_p),
        ("d_pos_fixed", ctypes.c_void_p),
        ("d_vel", ctypes.c_void_p),
        ("d_tag", ctypes.c_void_p),
        ("d_orientation", ctypes.c_void_p),
    ]


class Outputs(ctypes.Structure):
    _fields_ = [
        ("d_force", ctypes.c_void_p),
        ("d_virial", ctypes.c_void_p),
        ("virial_pitch", ctypes.c_uint32),
        ("d_torque", ctypes.c_void_p),
    ]


class PairArgs(ctypes.Structure):
    _fields_ = [
        ("particles", Particles),
        ("out", Outputs),
        ("d_n_neigh", ctypes.c_void_p),
        ("d_nlist", ctypes.c_void_p),
        ("nlist_pitch", ctypes.c_uint32),
        ("max_neighbors", ctypes.c_uint32),
        ("d_rcutsq", ctypes.c_void_p),
        ("d_ronsq", ctypes.c_void_p),
        ("ntypes", ctypes.c_uint32),
        ("shift_mode", ctypes.c_int32),
        ("compute_virial", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("block_size", ctypes.c_uint32),
        ("seed", ctypes.c_uint16),
        ("timestep", ctypes.c_uint64),
        ("dt", ctypes.c_double),
        ("kT", ctypes.c_double),
        ("stream", ctypes.c_void_p),
        ("d_tile_count", ctypes.c_void_p),
        ("d_tile_jset", ctypes.c_void_p),
        ("d_tile_list", ctypes.c_void_p),
        ("tile_size", ctypes.c_uint32),
        ("tile_slots", ctypes.c_uint32),
        ("d_tile_order", ctypes.c_void_p),
        ("d_tile_nn", ctypes.c_void_p),
        ("d_head_list", ctypes.c_void_p),
    ]


class StepArgs(ctypes.Structure):
    """azp_step_args: the MD step that follows the pair force in
    azp_pair_update_compute_step."""

    _fields_ = [
        ("d_pos", ctypes.c_void_p),
        ("d_vel", ctypes.c_void_p),
        ("d_image", ctypes.c_void_p),
        ("d_tag", ctypes.c_void_p),
        ("dt", ctypes.c_double),
        ("langevin", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("kT", ctypes.c_double),
        ("seed", ctypes.c_uint16),
        ("timestep", ctypes.c_uint64),
    ]


class BatchArrays(ctypes.Structure):
    _fields_ = [
        ("d_pos", ctypes.c_void_p),
        ("d_pos_fixed", ctypes.c_void_p),
        ("d_orientation", ctypes.c_void_p),
        ("d_vel", ctypes.c_void_p),
        ("d_angmom", ctypes.c_void_p),
        ("d_accel", ctypes.c_void_p),
        ("d_force", ctypes.c_void_p),
        ("d_torque", ctypes.c_void_p),
        ("d_virial", ctypes.c_void_p),
        ("virial_pitch", ctypes.c_uint32),
    ]


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_uint32),
        ("grab", ctypes.c_uint32),
        ("workgroups_per_cu", ctypes.c_uint32),
        ("walk_block", ctypes.c_uint32),
        ("force_us", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_uint64),
        ("n_builds", ctypes.c_uint64),
        ("replica_builds", ctypes.c_uint64),
        ("graph_launches", ctypes.c_uint64),
        ("max_neighbors", ctypes.c_uint32),
        ("n_cells", ctypes.c_uint32),
    ]


# (name, restype, argtypes) of every entry point in include/azp.h
_vp = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
SIGNATURES = [
    ("azp_last_error", ctypes.c_char_p, []),
    ("azp_abi_version", _i, []),
    ("azp_param_nvalues", _i, [_i]),
    ("azp_param_size", _i, [_i]),
    ("azp_param_pack", _i, [_i, _pd, _vp]),
    ("azp_param_unpack", _i, [_i, _vp, _pd]),
    ("azp_class_lookup", _i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    ("azp_gpu_compute_pair_forces", _i, [ctypes.POINTER(PairArgs), _i, _vp]),
    ("azp_gpu_compute_pair_forces_noalloc", _i, [ctypes.POINTER(PairArgs), _i, _vp]),
    ("azp_gpu_compute_dpd_forces", _i, [ctypes.POINTER(PairArgs), _vp]),
    ("azp_gpu_compute_pair_aniso_forces", _i, [ctypes.POINTER(PairArgs), _i, _vp]),
    ("azp_gpu_pack_fixed", _i, [_vp, _u32, ctypes.POINTER(Box), _vp, _vp]),
    ("azp_gpu_dpd_alpha", _i, [ctypes.c_uint16, _u64, _u64, _vp, _vp, _vp, _vp]),
    ("azp_adjoint_workspace_size", ctypes.c_size_t, [_u32, _u32]),
    ("azp_gpu_pair_adjoint", _i, [ctypes.POINTER(PairArgs), _i, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t]),
    ("azp_gpu_integrate_vv_one", _i, [_vp, _vp, _vp, _vp, _u32, ctypes.POINTER(Box), ctypes.c_double, _vp]),
    ("azp_gpu_integrate_vv_two", _i, [_vp, _vp, _vp, _vp, _u32, ctypes.c_double, _i, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint16, _u64, _vp]),
    ("azp_gpu_integrate_vv_two_one", _i, [_vp, _vp, _vp, _vp, _vp, _u32, ctypes.POINTER(Box), ctypes.c_double, _i,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_uint16, _u64, _vp]),
    ("azp_gpu_integrate_rot_one", _i, [_vp, _vp, _vp, _vp, _u32, ctypes.c_double, _vp]),
    ("azp_gpu_integrate_rot_two", _i, [_vp, _vp, _vp, _vp, _u32, ctypes.c_double, _i, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_uint16, _u64, _vp]),
    ("azp_thermo_workspace_size", _i, []),
    ("azp_gpu_kinetic_tensor", _i, [_vp, _u32, _vp, _vp, _vp]),
    ("azp_gpu_thermo_reduce", _i, [_vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp]),
    ("azp_nlist_create", _i, [ctypes.POINTER(_vp)]),
    ("azp_nlist_destroy", _i, [_vp]),
    ("azp_nlist_set_r_buff", _i, [_vp, ctypes.c_double]),
    ("azp_nlist_set_r_cut", _i, [_vp, _u32, _pd]),
    ("azp_nlist_update", _i, [_vp, ctypes.POINTER(Particles), _i, ctypes.POINTER(_i), _vp]),
    ("azp_nlist_get", _i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    ("azp_nlist_num_builds", _i, [_vp, ctypes.POINTER(_u64)]),
    ("azp_nlist_download", _i, [_vp, _u32, _vp, _vp, _u64]),
    ("azp_nlist_set_tiled", _i, [_vp, _i]),
    ("azp_nlist_get_tiles", _i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    ("azp_nlist_get_tile_order", _i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("azp_nlist_walk_kind", _i, [_vp, ctypes.POINTER(_i)]),
    ("azp_nlist_tile_stats", _i, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    ("azp_sorter_create", _i, [ctypes.POINTER(_vp)]),
    ("azp_sorter_destroy", _i, [_vp]),
    ("azp_sorter_order", _i, [_vp, ctypes.POINTER(Particles), _u32, _vp, _vp]),
    ("azp_gather_rows", _i, [_vp, _vp, _vp, _u32, _u32, _vp]),
    ("azp_wall_create", _i, [ctypes.POINTER(_vp), ctypes.c_char_p, _u32]),
    ("azp_wall_destroy", _i, [_vp]),
    ("azp_wall_param_nvalues", _i, [_vp]),
    ("azp_wall_set_params", _i, [_vp, _u32, _pd]),
    ("azp_wall_get_params", _i, [_vp, _u32, _pd]),
    ("azp_wall_set_geometry", _i, [_vp, _u32, _pd, _u32, _pd, _u32, _pd]),
    ("azp_wall_compute", _i, [_vp, ctypes.POINTER(Particles), ctypes.POINTER(Outputs), _i, _vp]),
    ("azp_wall_check", _i, [_vp, _vp]),
    ("azp_pair_create", _i, [ctypes.POINTER(_vp), ctypes.c_char_p, _u32]),
    ("azp_pair_destroy", _i, [_vp]),
    ("azp_pair_set_params", _i, [_vp, _u32, _u32, _pd]),
    ("azp_pair_get_params", _i, [_vp, _u32, _u32, _pd]),
    ("azp_pair_set_r_cut", _i, [_vp, _u32, _u32, ctypes.c_double]),
    ("azp_pair_set_r_on", _i, [_vp, _u32, _u32, ctypes.c_double]),
    ("azp_pair_set_shift_mode", _i, [_vp, _i]),
    ("azp_pair_set_precision", _i, [_vp, _i]),
    ("azp_pair_set_thermostat", _i, [_vp, ctypes.c_uint16, ctypes.c_double, ctypes.c_double]),
    ("azp_pair_get_r_cut", _i, [_vp, _pd]),
    ("azp_pair_compute", _i, [_vp, _vp, ctypes.POINTER(Particles), ctypes.POINTER(Outputs), _u64, _i, _vp]),
    ("azp_pair_update_compute", _i, [_vp, _vp, ctypes.POINTER(Particles), ctypes.POINTER(Outputs), _u64, _i, _i,
                                     ctypes.POINTER(_i), _vp]),
    ("azp_pair_update_compute_step", _i, [_vp, _vp, ctypes.POINTER(Particles), ctypes.POINTER(Outputs), _u64, _i,
                                          _i, ctypes.POINTER(_i), ctypes.POINTER(StepArgs), _vp]),
    ("azp_batch_create", _i, [ctypes.POINTER(_vp), _u32, ctypes.POINTER(_u32), ctypes.POINTER(Box), _u32, _i]),
    ("azp_batch_destroy", _i, [_vp]),
    ("azp_batch_size", _i, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    ("azp_batch_set_params", _i, [_vp, _u32, _u32, _pd]),
    ("azp_batch_set_r_cut", _i, [_vp, _u32, _u32, ctypes.c_double]),
    ("azp_batch_set_r_buff", _i, [_vp, ctypes.c_double]),
    ("azp_batch_set_mode", _i, [_vp, _i, _i, _i]),
    ("azp_batch_set_langevin", _i, [_vp, _pd, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint16]),
    ("azp_batch_bind", _i, [_vp, ctypes.POINTER(BatchArrays)]),
    ("azp_batch_set_timestep", _i, [_vp, _u64, _vp]),
    ("azp_batch_compute", _i, [_vp, _i, _vp]),
    ("azp_batch_run", _i, [_vp, _u32, _i, _vp]),
    ("azp_batch_thermo", _i, [_vp, _vp, _vp]),
    ("azp_batch_autotune", _i, [_vp, ctypes.POINTER(BatchTuning), _vp]),
    ("azp_batch_get_tuning", _i, [_vp, ctypes.POINTER(BatchTuning)]),
    ("azp_batch_set_tuning", _i, [_vp, ctypes.POINTER(BatchTuning)]),
    ("azp_batch_stats", _i, [_vp, ctypes.POINTER(BatchStats)]),
    ("azp_batch_nlist_download", _i, [_vp, _vp, _vp, _u64, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
]

# must equal AZP_ABI_VERSION of include/azp.h (azp_pair_args layout)
ABI_VERSION = 3

_lib = None


def lib():
    """Load libazp.so (once) and declare every signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libazp.so not found at {LIB_PATH}: build it with "
                "`python -m azplugins_amd.build` (there is no CPU fallback)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.azp_abi_version() != ABI_VERSION:
            raise ImportError(
                f"{LIB_PATH} has ABI version {L.azp_abi_version()}, these bindings expect "
                f"{ABI_VERSION}: rebuild it with `python -m azplugins_amd.build`"
            )
        _lib = L
    return _lib


def check(status):
    """Raise AzpError for a non-zero status from the library."""
    if status != OK:
        msg = lib().azp_last_error().decode(errors="replace")
        raise AzpError(status, msg)
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def dptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def dbl_array(values):
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr
