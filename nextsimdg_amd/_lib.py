"""ctypes binding of libnsdg_hip.so (the C ABI declared in include/nsdg_dyn.h and
include/nsdg_cnn.h).

There is no fallback: if the gfx950 library is missing or cannot be loaded, every
entry point of this package raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsdg_hip.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_d = ctypes.c_double
_s = ctypes.c_char_p


class NsdgParams(ctypes.Structure):
    """nsdg_params (include/nsdg_dyn.h)."""

    _fields_ = [(n, ctypes.c_double) for n in (
        "rho_ice", "rho_atm", "rho_ocean", "C_atm", "C_ocean", "F_atm", "F_ocean",
        "fc", "ocean_turning_angle", "gravity", "Pstar", "DeltaMin", "alpha", "beta")] + [("nsteps", ctypes.c_int32)]


MEB_FIELDS = ("compaction_param", "nu0", "young", "P0", "undamaged_time_relaxation_sigma",
              "exponent_relaxation_sigma", "exponent_compression_factor", "tan_phi", "compr_strength", "C_lab")


class NsdgMEBParams(ctypes.Structure):
    """nsdg_meb_params (include/nsdg_dyn.h)."""

    _fields_ = [(n, ctypes.c_double) for n in MEB_FIELDS]


# name -> (restype, argtypes); the full C ABI of include/nsdg_dyn.h
SIGNATURES = {
    "nsdg_last_error": (_s, []),
    "nsdg_params_default": (_i, [ctypes.POINTER(NsdgParams)]),
    "nsdg_device_count": (_i, []),
    "nsdg_mesh_create": (_i, [_sz, _sz, _dp, _i, ctypes.POINTER(_vp)]),
    "nsdg_mesh_set_landmask": (_i, [_vp, _dp]),
    "nsdg_mesh_dirichlet_from_mask": (_i, [_vp]),
    "nsdg_mesh_dirichlet_from_edge": (_i, [_vp, _i]),
    "nsdg_mesh_set_dirichlet": (_i, [_vp, _i, _u64p, _sz]),
    "nsdg_mesh_add_periodic": (_i, [_vp, _u64p, _sz]),
    "nsdg_mesh_get_dirichlet": (_i, [_vp, _i, _u64p, ctypes.POINTER(_sz)]),
    "nsdg_mesh_destroy": (_i, [_vp]),
    "nsdg_transport_create": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "nsdg_transport_set_scheme": (_i, [_vp, _s]),
    "nsdg_transport_set_velocity": (_i, [_vp, _dp, _dp]),
    "nsdg_transport_prepare_advection": (_i, [_vp, _i, _dp, _dp]),
    "nsdg_transport_reinit_normal_velocity": (_i, [_vp]),
    "nsdg_transport_step": (_i, [_vp, _d, _dp]),
    "nsdg_transport_run": (_i, [_vp, _d, _dp, _i, _i]),
    "nsdg_transport_destroy": (_i, [_vp]),
    "nsdg_limit": (_i, [_i, _sz, _dp, _d, _i]),
    "nsdg_dyn_create": (_i, [_i, _i, ctypes.POINTER(NsdgParams), ctypes.POINTER(_vp)]),
    "nsdg_meb_params_default": (_i, [ctypes.POINTER(NsdgMEBParams)]),
    "nsdg_bbm_create": (_i, [_i, _i, ctypes.POINTER(NsdgParams), ctypes.POINTER(NsdgMEBParams), ctypes.POINTER(_vp)]),
    "nsdg_dyn_initialise": (_i, [_vp, _sz, _sz, _dp, _i, _dp]),
    "nsdg_dyn_initialise_mesh": (_i, [_vp, _vp]),
    "nsdg_dyn_set_field": (_i, [_vp, _s, _dp, _i]),
    "nsdg_dyn_get_field": (_i, [_vp, _s, _dp, _i]),
    "nsdg_dyn_get_cg": (_i, [_vp, _s, _dp]),
    "nsdg_dyn_set_cg": (_i, [_vp, _s, _dp]),
    "nsdg_dyn_cg_nodes": (_sz, [_vp]),
    "nsdg_dyn_update": (_i, [_vp, _d]),
    "nsdg_dyn_advect": (_i, [_vp, _d]),
    "nsdg_dyn_prepare_iteration": (_i, [_vp]),
    "nsdg_dyn_begin_iteration": (_i, [_vp, _d]),
    "nsdg_dyn_strain": (_i, [_vp]),
    "nsdg_dyn_stress_update": (_i, [_vp, _i]),
    "nsdg_dyn_divergence": (_i, [_vp]),
    "nsdg_dyn_momentum": (_i, [_vp]),
    "nsdg_dyn_apply_boundaries": (_i, [_vp]),
    "nsdg_dyn_subcycles": (_i, [_vp, _i]),
    "nsdg_dyn_set_precision": (_i, [_vp, _i]),
    "nsdg_dyn_set_pairing": (_i, [_vp, _i]),
    "nsdg_dyn_benchmark_wind": (_i, [_vp, _d, _d, _d, _d, _d]),
    "nsdg_dyn_storm_wind": (_i, [_vp, _d, _d, _dp, _dp, _dp, _i]),
    "nsdg_dyn_reset_dg0": (_i, [_vp, _s]),
    "nsdg_dyn_mass": (_i, [_vp, _dp, _dp]),
    "nsdg_dyn_mass_noalloc": (_i, [_vp, _dp, _dp]),
    "nsdg_dyn_set_budget": (_i, [_vp, _i]),
    "nsdg_dyn_write_restart": (_i, [_vp, ctypes.c_char_p, ctypes.c_double]),
    "nsdg_dyn_read_restart": (_i, [_vp, ctypes.c_char_p, _dp]),
    "nsdg_dyn_budget": (_i, [_vp, _dp, _i]),
    "nsdg_dyn_residual": (_i, [_vp, _dp, _dp, _dp]),
    "nsdg_dyn_stress_from_velocity": (_i, [_vp]),
    "nsdg_dyn_synchronize": (_i, [_vp]),
    "nsdg_dyn_save_state": (_i, [_vp, _vp, ctypes.POINTER(_sz)]),
    "nsdg_dyn_load_state": (_i, [_vp, _vp, _sz]),
    "nsdg_dyn_set_profiling": (_i, [_vp, _i]),
    "nsdg_dyn_kernel_stats": (_i, [_vp, _s, ctypes.POINTER(_i), _dp]),
    "nsdg_dyn_destroy": (_i, [_vp]),
    # include/nsdg_cnn.h
    "nsdg_cnn_create": (_vp, [ctypes.c_uint64]),
    "nsdg_cnn_destroy": (_i, [_vp]),
    "nsdg_cnn_param_count": (_i, [_vp, ctypes.POINTER(_sz)]),
    "nsdg_cnn_get_params": (_i, [_vp, _vp]),
    "nsdg_cnn_set_params": (_i, [_vp, _vp]),
    "nsdg_cnn_forward": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "nsdg_cnn_loss_grad": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _dp, _vp]),
    "nsdg_cnn_adam": (_i, [_vp, ctypes.c_float]),
    "nsdg_cnn_set_profiling": (_i, [_vp, _i]),
    "nsdg_cnn_kernel_stats": (_i, [_vp, _s, ctypes.POINTER(_i), _dp]),
    "nsdg_device_alloc": (_i, [_sz, ctypes.POINTER(_vp)]),
    "nsdg_device_free": (_i, [_vp]),
    "nsdg_device_upload": (_i, [_vp, _vp, _sz]),
    "nsdg_device_download": (_i, [_vp, _vp, _sz]),
    "nsdg_device_gather": (_i, [_vp, _vp, _vp, _i, _sz]),
    "nsdg_dyn_features": (_i, [_vp, _vp]),
    "nsdg_dyn_velocity_target": (_i, [_vp, _vp, _vp]),
    "nsdg_dyn_apply_velocity": (_i, [_vp, _vp]),
    "nsdg_dyn_warm_start": (_i, [_vp, _vp, _dp, _dp, ctypes.POINTER(_i)]),
    "nsdg_dyn_subcycles_to_residual": (_i, [_vp, _d, _i, _i, _d, ctypes.POINTER(_i), _dp]),
}


class NsdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nsdg error {code}: {msg}")
        self.code = code


def load():
    """Load the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m nextsimdg_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise NsdgError(rc, (load().nsdg_last_error() or b"").decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)


def u64ptr(a):
    return a.ctypes.data_as(_u64p)
