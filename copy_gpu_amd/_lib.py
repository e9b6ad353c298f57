"""ctypes binding of librbmd_force.so (the C ABI declared in include/rbmd_force.h).

torch is imported before the library is loaded: torch ships a HIP runtime with
the same SONAME (libamdhip64.so.7), and loading it first makes the dynamic
linker bind this library to that one runtime, so torch device pointers and
streams are valid here.  A missing library is an ImportError -- there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbmd_force.so")

RBMD_OK = 0
RBMD_ERR_INVALID = -1
RBMD_ERR_HIP = -2
RBMD_ERR_BOX = -3
RBMD_ERR_PRECISION = -4
RBMD_ERR_STATE = -5
RBMD_ERR_CAPACITY = -6

KSPACE_NONE = 0
KSPACE_RBE = 1
KSPACE_EWALD = 2
RBMD_EV = 1
NEIGHBOR_VERLET = 0
NEIGHBOR_RBL = 1


class ForceParams(C.Structure):
    _fields_ = [
        ("box_lo", C.c_double * 3),
        ("box_len", C.c_double * 3),
        ("rc_lj", C.c_double),
        ("rc_coul", C.c_double),
        ("skin", C.c_double),
        ("alpha", C.c_double),
        ("qqr2e", C.c_double),
        ("ntypes", C.c_int32),
        ("eps", C.POINTER(C.c_double)),
        ("sigma", C.POINTER(C.c_double)),
        ("precision", C.c_int32),
        ("kspace", C.c_int32),
        ("p_max", C.c_int32),
        ("ewald_tol", C.c_double),
        ("list_capacity", C.c_int32),
        ("tile_stage", C.c_int32),
        ("neighbor_type", C.c_int32),
        ("r_core", C.c_double),
        ("rbl_pice", C.c_int32),
        ("rbl_seed", C.c_uint64),
        ("half_shell", C.c_int32),
    ]


class EnergyVirial(C.Structure):
    _fields_ = [
        ("e_lj", C.c_double),
        ("e_real", C.c_double),
        ("e_self", C.c_double),
        ("e_kspace", C.c_double),
        ("virial", C.c_double * 6),
    ]


THERMO_NONE = 0
THERMO_BERENDSEN = 1
THERMO_RESCALE = 2
THERMO_NOSE_HOOVER = 3
THERMO_LANGEVIN = 4
THERMO_MEASURE = 5


class Thermostat(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("t_target", C.c_double),
        ("t_damp", C.c_double),
        ("dof", C.c_double),
        ("kb", C.c_double),
        ("xi", C.c_double),
        ("gamma", C.c_double),
        ("kbt", C.c_double),
        ("seed", C.c_uint64),
        ("noise_step", C.c_uint64),
    ]


class KspaceParams(C.Structure):
    _fields_ = [
        ("box_lo", C.c_double * 3),
        ("box_len", C.c_double * 3),
        ("alpha", C.c_double),
        ("qqr2e", C.c_double),
        ("precision", C.c_int32),
    ]


class DataInfo(C.Structure):
    _fields_ = [
        ("natoms", C.c_int64),
        ("ntypes", C.c_int32),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("has_masses", C.c_int32),
        ("has_pair_coeffs", C.c_int32),
        ("has_velocities", C.c_int32),
    ]


# name -> (restype, argtypes); the symbol test checks this table against the header
_P = C.c_void_p
class ReplicaState(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
                ("fx", C.c_void_p), ("fy", C.c_void_p), ("fz", C.c_void_p),
                ("img", C.c_void_p), ("mass", C.c_void_p)]


_SIGS = {
    "rbmd_force_create": (C.c_int, [C.POINTER(ForceParams), C.c_int64, C.POINTER(C.c_void_p)]),
    "rbmd_force_destroy": (None, [_P]),
    "rbmd_force_last_error": (C.c_char_p, [_P]),
    "rbmd_force_set_atoms": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "rbmd_force_set_kvectors": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "rbmd_force_sample_kvectors": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_uint64, _P]),
    "rbmd_force_compute": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, C.c_int32, _P]),
    "rbmd_force_get_energy_virial": (C.c_int, [_P, C.POINTER(EnergyVirial), _P]),
    "rbmd_md_run": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                              C.c_double, C.c_int64, C.c_uint64, C.c_uint64, C.c_int32, _P]),
    "rbmd_md_set_thermostat": (C.c_int, [_P, C.POINTER(Thermostat)]),
    "rbmd_md_get_thermostat": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_uint64)]),
    "rbmd_md_set_temperature_log": (C.c_int, [_P, _P, C.c_int64]),
    "rbmd_langevin_noise_host": (C.c_int, [C.c_uint64, C.c_uint64, _P]),
    "rbmd_force_get_neighbors": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "rbmd_force_get_kspace": (C.c_int, [_P, C.POINTER(C.c_int32), _P, _P, _P, _P]),
    "rbmd_force_get_stats": (C.c_int, [_P, _P]),
    "rbmd_force_autotune": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, C.c_int32, _P]),
    "rbmd_batch_autotune": (C.c_int, [_P, _P, C.c_int32, _P]),
    "rbmd_force_set_timing": (C.c_int, [_P, C.c_int32]),
    "rbmd_force_get_timing": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int64)]),
    "rbmd_force_time_pair": (C.c_int, [_P] + [_P] * 6 + [C.c_int32, C.POINTER(C.c_double), _P]),
    "rbmd_util_copy": (C.c_int, [_P, _P, C.c_int64, _P]),
    "rbmd_kspace_compute": (C.c_int, [C.POINTER(KspaceParams), C.c_int64] + [_P] * 4 +
                            [C.c_int32] + [_P] * 8),
    "rbmd_kspace_compute_noalloc": (C.c_int, [C.POINTER(KspaceParams), C.c_int64] + [_P] * 4 +
                                    [C.c_int32] + [_P] * 8),
    "rbmd_kspace_last_error": (C.c_char_p, []),
    "rbmd_batch_create": (C.c_int, [C.c_int32, _P, _P, C.POINTER(C.c_void_p)]),
    "rbmd_batch_destroy": (None, [_P]),
    "rbmd_batch_last_error": (C.c_char_p, [_P]),
    "rbmd_batch_replica": (C.c_void_p, [_P, C.c_int32]),
    "rbmd_batch_md_run": (C.c_int, [_P, _P, C.c_double, C.c_int64, _P, C.c_uint64, C.c_int32,
                                    _P]),
    "rbmd_proposal_create": (C.c_int, [C.c_int32, _P, C.c_double, C.c_double, _P,
                                       C.POINTER(C.c_void_p)]),
    "rbmd_proposal_destroy": (None, [_P]),
    "rbmd_proposal_sample_host": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_uint64, _P, _P]),
    "rbmd_force_set_proposal": (C.c_int, [_P, _P]),
    "rbmd_mlp_nparams": (C.c_int, []),
    "rbmd_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "rbmd_mlp_destroy": (None, [_P]),
    "rbmd_mlp_last_error": (C.c_char_p, [_P]),
    "rbmd_mlp_get_params": (C.c_int, [_P, _P]),
    "rbmd_mlp_set_params": (C.c_int, [_P, _P]),
    "rbmd_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "rbmd_mlp_loss_grad": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.POINTER(C.c_double), _P,
                                     _P]),
    "rbmd_mlp_train": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int32, C.c_double,
                                 C.POINTER(C.c_double), _P]),
    "rbmd_philox4x32": (None, [_P, _P, _P]),
    "rbmd_sys_fcc": (C.c_int, [C.c_int32, C.c_double, _P, _P, _P, C.POINTER(C.c_double)]),
    "rbmd_sys_random": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_uint64, _P, _P, _P]),
    "rbmd_sys_charges": (C.c_int, [C.c_int64, C.c_uint64, _P]),
    "rbmd_sys_velocities": (C.c_int, [C.c_int64, C.c_double, _P, C.c_uint64, _P, _P, _P]),
    "rbmd_read_data_info": (C.c_int, [C.c_char_p, C.POINTER(DataInfo)]),
    "rbmd_read_data": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64] + [_P] * 12),
    "rbmd_write_dump": (C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_int64] + [_P] * 6),
    "rbmd_io_last_error": (C.c_char_p, []),
    "rbmd_sample_kvectors_host": (C.c_int, [C.c_double, _P, C.c_int32, C.c_uint64, C.c_uint64,
                                            _P, C.POINTER(C.c_double)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load (once) and return the HIP library; raise ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m copy_gpu_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list[str]:
    return list(_SIGS)
