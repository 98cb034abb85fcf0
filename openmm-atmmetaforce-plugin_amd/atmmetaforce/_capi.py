"""ctypes binding of libatm_mi355x.so (include/atm_mi355x.h).

The library is loaded from the package's lib/ directory; a missing library is an
error (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libatm_mi355x.so")

ATM_OK = 0
ATM_ERR_INVALID_ARGUMENT = -1
ATM_ERR_DEVICE = -2
ATM_ERR_CAPACITY = -3
ATM_ERR_STATE = -4
ATM_ERR_UNSUPPORTED = -5

ATM_EVAL_FORCES = 0x1
ATM_EVAL_CLEAR_FORCES = 0x2
ATM_EVAL_PLAIN = 0x4
ATM_EVAL_NO_GRAPH = 0x8
ATM_EVAL_FP32_PAIRS = 0x10
ATM_EVAL_STATE1_ONLY = 0x20
ATM_EVAL_PARAM_GRADS = 0x40

FORCE_SCALE = 16777216.0
ONE_4PI_EPS0 = 138.935456

# Every entry point declared in include/atm_mi355x.h.
EXPORTED = (
    "atm_version", "atm_abi_version", "atm_last_error", "atm_device_count", "atm_create", "atm_destroy",
    "atm_set_particles", "atm_set_particles_device", "atm_set_exceptions", "atm_set_energy_offsets", "atm_set_displacements", "atm_set_positions",
    "atm_set_positions_device",
    "atm_set_box", "atm_set_box_vectors", "atm_set_params", "atm_eval", "atm_synchronize", "atm_get_results", "atm_get_forces",
    "atm_get_forces_fixed", "atm_clear_forces", "atm_get_pairs", "atm_backward", "atm_backward_device",
    "atm_set_stream",
    "atm_autotune", "atm_get_tuning", "atm_set_tuning", "atm_get_stats", "atm_get_list_stats", "atm_profile",
    "atm_eval_noalloc",
    "atm_restraints_create", "atm_restraints_destroy", "atm_restraints_set", "atm_restraints_eval",
    "atm_md_init", "atm_md_step", "atm_md_get",
    "atm_pme_create", "atm_pme_destroy", "atm_pme_get_info", "atm_pme_set_particles", "atm_pme_set_exceptions", "atm_pme_set_energy_offsets",
    "atm_pme_set_params", "atm_pme_eval", "atm_bonded_create", "atm_bonded_destroy", "atm_bonded_set",
    "atm_bonded_eval",
)


class SystemDesc(ctypes.Structure):
    _fields_ = [("num_atoms", ctypes.c_int32), ("box", ctypes.c_double * 3), ("cutoff", ctypes.c_double),
                ("rf_dielectric", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("lambda1", "lambda2", "alpha", "u0", "w0", "umax", "ubcore", "acore", "direction")]


class Result(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("energy", "perturbation_energy", "state1_energy", "state2_energy", "raw_perturbation",
                 "softcore_derivative", "softplus_derivative", "hybrid_weight")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("column_scale", ctypes.c_double), ("pair_waves_per_block", ctypes.c_int32),
                ("zbins_per_column", ctypes.c_int32), ("best_ms", ctypes.c_double),
                ("pair_kernel", ctypes.c_int32), ("list_capacity", ctypes.c_int32),
                ("pair_split", ctypes.c_int32), ("full_state2", ctypes.c_int32),
                ("pair_order", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


class StageTimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("build_ms", "pairlist_ms", "pairs_ms", "displaced_ms", "energy_ms", "blend_ms", "total_ms")] + \
               [("reps", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CentroidRestraint(ctypes.Structure):
    _fields_ = [("group1_first", ctypes.c_int32), ("group1_count", ctypes.c_int32),
                ("group2_first", ctypes.c_int32), ("group2_count", ctypes.c_int32),
                ("kf", ctypes.c_double), ("tol", ctypes.c_double), ("offset", ctypes.c_double * 3)]


class PositionRestraint(ctypes.Structure):
    _fields_ = [("atom", ctypes.c_int32), ("periodic", ctypes.c_int32), ("ref", ctypes.c_double * 3),
                ("kf", ctypes.c_double), ("tol", ctypes.c_double)]


class PmeDesc(ctypes.Structure):
    _fields_ = [("num_atoms", ctypes.c_int32), ("grid", ctypes.c_int32 * 3), ("box", ctypes.c_double * 3),
                ("cutoff", ctypes.c_double), ("ewald_tolerance", ctypes.c_double), ("alpha", ctypes.c_double),
                ("dispersion_correction", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ATMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load(path: str | None = None):
    """Load (once) and return the library handle."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"libatm_mi355x.so not found at {path}; run __graft_entry__.build() "
                          "(python openmm-atmmetaforce-plugin_amd/build.py)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.c_int32
    L.atm_version.restype = ctypes.c_char_p
    L.atm_abi_version.restype = i32
    L.atm_last_error.restype = ctypes.c_char_p
    sigs = {
        "atm_device_count": [ctypes.POINTER(i32)],
        "atm_create": [ctypes.POINTER(SystemDesc), i32, i32, ctypes.POINTER(vp)],
        "atm_destroy": [vp],
        "atm_set_particles": [vp, i32, dp, dp, dp, dp],
        "atm_set_particles_device": [vp, vp, vp, vp],
        "atm_set_displacements": [vp, i32, dp],
        "atm_set_exceptions": [vp, i32, i32, ctypes.POINTER(i32), dp, dp, dp],
        "atm_set_energy_offsets": [vp, i32, ctypes.c_double, ctypes.c_double],
        "atm_set_positions": [vp, i32, dp],
        "atm_set_positions_device": [vp, vp],
        "atm_set_box": [vp, i32, dp],
        "atm_set_box_vectors": [vp, i32, dp],
        "atm_set_params": [vp, i32, ctypes.POINTER(Params)],
        "atm_eval": [vp, i32],
        "atm_synchronize": [vp],
        "atm_get_results": [vp, i32, ctypes.POINTER(Result)],
        "atm_get_forces": [vp, i32, dp],
        "atm_get_forces_fixed": [vp, i32, ctypes.POINTER(ctypes.c_int64)],
        "atm_clear_forces": [vp],
        "atm_get_pairs": [vp, i32, i32, ctypes.POINTER(i32), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
        "atm_backward": [vp, dp, dp, dp, dp, dp],
        "atm_backward_device": [vp, dp, vp, vp, vp, dp],
        "atm_set_stream": [vp, vp],
        "atm_autotune": [vp, i32, ctypes.POINTER(Tuning)],
        "atm_get_tuning": [vp, ctypes.POINTER(Tuning)],
        "atm_set_tuning": [vp, ctypes.POINTER(Tuning)],
        "atm_get_stats": [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_int64)],
        "atm_get_list_stats": [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(i32), ctypes.POINTER(i32)],
        "atm_profile": [vp, i32, i32, ctypes.POINTER(StageTimes)],
        "atm_eval_noalloc": [ctypes.POINTER(SystemDesc), ctypes.POINTER(Params), vp, vp, vp, vp, vp, i32, vp, vp, vp,
                             vp],
        "atm_restraints_create": [i32, ctypes.POINTER(vp)],
        "atm_restraints_destroy": [vp],
        "atm_restraints_set": [vp, i32, i32, ctypes.POINTER(i32), dp, i32, ctypes.POINTER(CentroidRestraint), i32,
                               ctypes.POINTER(PositionRestraint)],
        "atm_restraints_eval": [vp, dp, dp, dp, dp],
        "atm_md_init": [vp, dp, dp, ctypes.c_double, ctypes.c_uint64],
        "atm_md_step": [vp, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "atm_md_get": [vp, i32, dp, dp, dp],
        "atm_pme_create": [i32, ctypes.POINTER(PmeDesc), ctypes.POINTER(vp)],
        "atm_pme_destroy": [vp],
        "atm_pme_get_info": [vp, dp, ctypes.POINTER(i32)],
        "atm_pme_set_particles": [vp, dp, dp, dp, dp],
        "atm_pme_set_exceptions": [vp, i32, ctypes.POINTER(i32), dp, dp, dp],
        "atm_pme_set_energy_offsets": [vp, ctypes.c_double, ctypes.c_double],
        "atm_pme_set_params": [vp, ctypes.POINTER(Params)],
        "atm_pme_eval": [vp, dp, dp, i32, dp, ctypes.POINTER(Result), dp],
        "atm_bonded_create": [i32, ctypes.POINTER(vp)],
        "atm_bonded_destroy": [vp],
        "atm_bonded_set": [vp, i32, i32, ctypes.POINTER(i32), dp, i32, ctypes.POINTER(i32), dp, i32,
                           ctypes.POINTER(i32), dp],
        "atm_bonded_eval": [vp, dp, dp, dp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != ATM_OK:
        msg = load().atm_last_error().decode("utf-8", "replace")
        raise ATMError(rc, msg)


def device_count() -> int:
    n = ctypes.c_int32(0)
    check(load().atm_device_count(ctypes.byref(n)))
    return n.value
