"""ctypes bindings of the two in-tree native libraries.

libfluidsim_amd.so  -- the device C ABI (include/fluidsim_amd.h)
libfluidsim_host.so -- the C++ host layer (include/fluidsim_host.h)

Nothing here computes: every call goes into the native code, and loading
fails loudly when the libraries have not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
DEVICE_LIB = LIB_DIR / "libfluidsim_amd.so"
HOST_LIB = LIB_DIR / "libfluidsim_host.so"

FSA_OK = 0
FSA_ERR_INVALID = 1
FSA_ERR_HIP = 2
FSA_ERR_NOMEM = 3
FSA_ERR_STATE = 4
FSA_ERR_NODEVICE = 5

FLUID, SOLID, VIRT = 0, 1, 2
UNCONSTRAINED_SOLID, FIXED_SOLID = 0, 1
DEBUG_PAIR_SUMS = 1
DEBUG_WALK = 2
HEALTH_FLUID_NONFINITE = 1
HEALTH_SOLID_NONFINITE = 2
PROBE_NOALLOC_MAX = 65536  # FSA_PROBE_NOALLOC_MAX
FORM_PACKED = 1
FORM_COMPENSATED = 2

STAGES = ("cell", "couple", "virt", "fluid", "solid")


class FsaParams(C.Structure):
    """Mirror of `fsa_params` (include/fluidsim_amd.h); field order matters."""

    _fields_ = [
        ("n_fluid", C.c_int32), ("n_solid", C.c_int32), ("n_virt", C.c_int32),
        ("gravity", C.c_float * 3),
        ("fluid_sound_speed", C.c_float),
        ("fluid_reference_density", C.c_float),
        ("fluid_reference_density_inverse", C.c_float),
        ("fluid_gamma", C.c_float),
        ("fluid_gamma_inv", C.c_float),
        ("fluid_viscosity", C.c_float),
        ("fluid_coefficient_p2rho", C.c_float),
        ("fluid_coefficient_rho2p", C.c_float),
        ("fluid_min_pressure", C.c_float),
        ("fluid_min_density", C.c_float),
        ("fluid_max_density", C.c_float),
        ("fluid_is_viscid", C.c_int32),
        ("solid_reference_density", C.c_float),
        ("solid_youngs_modulus", C.c_float),
        ("solid_poisson_ratio", C.c_float),
        ("solid_lambda", C.c_float),
        ("solid_nu", C.c_float),
        ("solid_bulk_modulus", C.c_float),
        ("solid_sound_speed", C.c_float),
        ("solid_artificial_viscosity", C.c_float * 2),
        ("domain_min", C.c_float * 3), ("domain_max", C.c_float * 3),
        ("cell_min", C.c_float * 3), ("cell_max", C.c_float * 3),
        ("interval", C.c_float * 3), ("interval_inv", C.c_float * 3),
        ("cell_number", C.c_int32 * 3),
        ("cell_number_total", C.c_int32),
        ("dt", C.c_float),
        ("istart", C.c_int32), ("iend", C.c_int32),
        ("solid_sub_step", C.c_int32),
        ("result_interval", C.c_int32),
        ("sort_every_step", C.c_int32),
        ("smoothing_length", C.c_float),
        ("particle_diameter", C.c_float),
        ("impact_length_hsml_ratio", C.c_float),
        ("smoothing_length_square", C.c_float),
        ("smoothing_length_inverse", C.c_float),
        ("smoothing_length_inverse_square", C.c_float),
        ("impact_distance_square", C.c_float),
        ("kernel_coefficient_1", C.c_float),
        ("kernel_coefficient_2", C.c_float),
        ("kernel_differential", C.c_float),
        ("thread_num", C.c_int32),
        ("precision_form", C.c_int32),
        ("neighbour_skin", C.c_float),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class FsaObservables(C.Structure):
    """Mirror of `fsa_observables` (include/fluidsim_amd.h)."""

    _fields_ = [
        ("fluid_mass", C.c_double),
        ("fluid_momentum", C.c_double * 3),
        ("fluid_kinetic_energy", C.c_double),
        ("fluid_potential_energy", C.c_double),
        ("front_x", C.c_double),
        ("solid_momentum", C.c_double * 3),
        ("gate_tip_displacement", C.c_double * 3),
        ("gate_tip_count", C.c_int64),
        ("max_fluid_speed", C.c_double),
        ("max_solid_speed", C.c_double),
        ("cfl_dt", C.c_double),
        ("solid_cfl_dt", C.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


_P = C.POINTER
_f = _P(C.c_float)
_i = _P(C.c_int32)
_d = _P(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); the list doubles as the exported-symbol check
DEVICE_API = {
    "fsa_abi_version": (C.c_int, []),
    "fsa_last_error": (C.c_char_p, []),
    "fsa_device_count": (C.c_int, [_i]),
    "fsa_create": (C.c_int, [_P(FsaParams), _vp, _P(_vp)]),
    "fsa_destroy": (C.c_int, [_vp]),
    "fsa_upload_fluid": (C.c_int, [_vp, _f, _f, _f, _f]),
    "fsa_upload_solid": (C.c_int, [_vp, _f, _f, _f, _i]),
    "fsa_upload_virt": (C.c_int, [_vp, _f, _i]),
    "fsa_download_fluid": (C.c_int, [_vp, _f, _f, _f, _f]),
    "fsa_download_solid": (C.c_int, [_vp, _f, _f, _f, _f, _f, _f, _f, _f, _f]),
    "fsa_download_virt": (C.c_int, [_vp, _f, _f, _f, _f, _f]),
    "fsa_particle_cell_update": (C.c_int, [_vp]),
    "fsa_fill_cell": (C.c_int, [_vp]),
    "fsa_couple_calculation": (C.c_int, [_vp]),
    "fsa_virt_update": (C.c_int, [_vp]),
    "fsa_fluid_calculation": (C.c_int, [_vp]),
    "fsa_solid_calculation": (C.c_int, [_vp]),
    "fsa_step": (C.c_int, [_vp, C.c_int]),
    "fsa_synchronize": (C.c_int, [_vp]),
    "fsa_get_cell_ids": (C.c_int, [_vp, C.c_int, _i]),
    "fsa_get_cell_start": (C.c_int, [_vp, C.c_int, _i]),
    "fsa_get_sorted_ids": (C.c_int, [_vp, C.c_int, _i]),
    "fsa_download_fluid_rates": (C.c_int, [_vp, _f, _f]),
    "fsa_get_jump_count": (C.c_int, [_vp, _P(C.c_int64)]),
    "fsa_get_list_stats": (C.c_int, [_vp, _P(C.c_int64), _P(C.c_int64), _P(C.c_int64), _P(C.c_int64)]),
    "fsa_set_debug": (C.c_int, [_vp, C.c_uint]),
    "fsa_observe": (C.c_int, [_vp, _P(FsaObservables)]),
    "fsa_probe_pressure": (C.c_int, [_vp, _f, C.c_int, _f]),
    "fsa_probe_pressure_noalloc": (C.c_int, [_vp, _f, C.c_int, _f]),
    "fsa_pair_census": (C.c_int, [_vp, _P(C.c_int64)]),
    "fsa_set_timing": (C.c_int, [_vp, C.c_int]),
    "fsa_set_timing_stages": (C.c_int, [_vp, C.c_uint]),
    "fsa_get_stage_times": (C.c_int, [_vp, _d, _P(C.c_int64)]),
    "fsa_reset_stage_times": (C.c_int, [_vp]),
    "fsa_memory_used": (C.c_int, [_vp, _P(C.c_int64)]),
    "fsa_footprint": (C.c_int, [_P(FsaParams), _P(C.c_int64), _P(C.c_int64)]),
    "fsa_device_memory": (C.c_int, [_P(C.c_int64), _P(C.c_int64)]),
    "fsa_batch_create": (C.c_int, [_vp, C.c_uint, _P(_vp)]),
    "fsa_batch_destroy": (C.c_int, [_vp]),
    "fsa_batch_step": (C.c_int, [_vp, _P(_vp), C.c_int, C.c_int]),
    "fsa_batch_set_streams": (C.c_int, [_vp, C.c_int]),
    "fsa_batch_synchronize": (C.c_int, [_vp]),
    "fsa_batch_stats": (C.c_int, [_vp, _P(C.c_int64), _P(C.c_int64), _P(C.c_int)]),
    "fsa_batch_tuning": (C.c_int, [_vp, _P(C.c_int64), _P(C.c_double)]),
    "fsa_get_health": (C.c_int, [_vp, _P(C.c_uint)]),
    "fsa_download_pair_counts": (C.c_int, [_vp, _P(C.c_int32)]),
    "fsa_set_fluid_threads": (C.c_int, [_vp, C.c_int]),
    "fsa_get_fluid_threads": (C.c_int, [_vp, _P(C.c_int)]),
    "fsa_autotune_fluid": (C.c_int, [_vp, C.c_int, _P(C.c_int)]),
}

FSA_BATCH_NO_GRAPH = 1
FSA_BATCH_AUTOTUNE = 2

HOST_API = {
    "fsh_last_error": (C.c_char_p, []),
    "fsh_parameter_initiate": (C.c_int, [C.c_char_p, _P(FsaParams)]),
    "fsh_parameter_from_ini": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _P(FsaParams)]),
    "fsh_config_read_string": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _i]),
    "fsh_config_read_float": (C.c_int, [C.c_char_p, C.c_char_p, C.c_float, _f]),
    "fsh_config_read_int": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, _i]),
    "fsh_config_read_bool": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, _i]),
    "fsh_kernel_coefficient": (C.c_float, [C.c_float]),
    "fsh_run": (C.c_int, [C.c_char_p, C.c_int]),
}

_device = None
_host = None


def _bind(path: Path, api: dict):
    if not path.exists():
        raise RuntimeError(f"{path.name} is not built; run `python -m fluidsim_amd.build` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for name, (res, args) in api.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


ABI_VERSION = 3  # FSA_ABI_VERSION of the header these mirrors follow


def device_lib():
    global _device
    if _device is None:
        lib = _bind(DEVICE_LIB, DEVICE_API)
        v = lib.fsa_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"{DEVICE_LIB.name} has ABI version {v}, these bindings need {ABI_VERSION}; rebuild it")
        _device = lib
    return _device


def host_lib():
    global _host
    if _host is None:
        device_lib()  # the host layer links against it
        _host = _bind(HOST_LIB, HOST_API)
    return _host


class FsaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc: int, lib=None) -> None:
    if rc != FSA_OK:
        lib = lib or device_lib()
        err = lib.fsa_last_error() if hasattr(lib, "fsa_last_error") else lib.fsh_last_error()
        raise FsaError(rc, (err or b"").decode())


def check_host(rc: int) -> None:
    if rc != FSA_OK:
        raise FsaError(rc, (host_lib().fsh_last_error() or b"").decode())
