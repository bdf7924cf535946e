"""ctypes binding of ``libsph_mi355x.so`` (the C ABI declared in include/sph_capi.h).

The library must have been built (``python -m tmp_sph_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsph_mi355x.so"

SPH_OK = 0
SPH_E_INVAL = -1
SPH_E_NOPARAMS = -2
SPH_E_STATE = -3
SPH_EOS_LINEAR = 0
SPH_EOS_TAIT = 1


class SphParams(C.Structure):
    """Mirror of ``sph_params`` (replaces the reference's SimParams)."""

    _fields_ = [
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("dim", C.c_int),
        ("numParticles", C.c_int),
        ("numFluid", C.c_int),
        ("gridxdim", C.c_int),
        ("gridydim", C.c_int),
        ("gridzdim", C.c_int),
        ("numCells", C.c_int),
        ("hash_max", C.c_int),
        ("eos", C.c_int),
        ("park_type", C.c_int),
        ("gridxmin", C.c_float),
        ("gridymin", C.c_float),
        ("gridzmin", C.c_float),
        ("gridSize", C.c_float),
        ("particleR", C.c_float),
        ("h", C.c_float),
        ("kh", C.c_float),
        ("eta", C.c_float),
        ("delta", C.c_float),
        ("afa", C.c_float),
        ("adh", C.c_float),
        ("particleMass", C.c_float),
        ("restDensity", C.c_float),
        ("cs", C.c_float),
        ("gravity", C.c_float),
        ("timeStep", C.c_float),
        ("gamma", C.c_float),
        ("tait_B", C.c_float),
        ("damper", C.c_float),
        ("rr_cut", C.c_float),
        ("park_pos", C.c_float * 3),
        ("pad_", C.c_float),
    ]

    def copy(self) -> "SphParams":
        out = SphParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(SphParams))
        return out

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class DamBreakCfg(C.Structure):
    """Mirror of ``sph_dambreak_cfg``."""

    _fields_ = [
        ("dim", C.c_int),
        ("eos", C.c_int),
        ("wall_layers", C.c_int),
        ("pad_", C.c_int),
        ("dx", C.c_double),
        ("tank", C.c_double * 3),
        ("column", C.c_double * 3),
        ("h_over_dx", C.c_double),
        ("rho0", C.c_double),
        ("g", C.c_double),
        ("c0", C.c_double),
        ("cfl", C.c_double),
        ("delta", C.c_double),
        ("alpha", C.c_double),
        ("gamma", C.c_double),
        ("grid_pad", C.c_double),
        ("cell_over_kh", C.c_double),
    ]


class SphTuning(C.Structure):
    _fields_ = [
        ("nbr_block", C.c_int),
        ("nbr_unroll", C.c_int),
        ("xcd_remap", C.c_int),
        ("flags", C.c_int),
        ("nbr_lds", C.c_int),
        ("nbr_compact", C.c_int),
        ("nbr_prefetch", C.c_int),
        ("nbr_list", C.c_int),
        ("cell_div", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


P = C.c_void_p
I = C.c_int
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

# name -> argument types; every function returns int status except sph_last_error
_PROTOS = {
    "sph_abi_sizes": [IP, IP, IP],
    "sph_device_count": [IP],
    "sph_set_device": [I],
    "sph_set_stream": [P],
    "sph_device_synchronize": [],
    "sph_malloc": [C.POINTER(C.c_void_p), C.c_ulonglong],
    "sph_free": [P],
    "sph_memcpy_htod": [P, P, C.c_ulonglong],
    "sph_memcpy_dtoh": [P, P, C.c_ulonglong],
    "sph_memcpy_dtod": [P, P, C.c_ulonglong],
    "sph_memset": [P, I, C.c_ulonglong],
    "sph_set_params": [C.POINTER(SphParams)],
    "sph_get_params": [C.POINTER(SphParams)],
    "sph_params_reference": [C.POINTER(SphParams), C.c_float, I, I],
    "sph_params_dambreak": [C.POINTER(DamBreakCfg), C.POINTER(SphParams)],
    "sph_params_subdivide": [C.POINTER(SphParams), I, C.POINTER(SphParams)],
    "sph_case_dambreak_count": [C.POINTER(DamBreakCfg), IP, IP],
    "sph_case_dambreak": [C.POINTER(DamBreakCfg), I, P, P, P, P, P, C.POINTER(SphParams), IP, IP],
    "sph_calcHash": [P, P, P, I],
    "sph_sortParticles": [P, P, I],
    "sph_sortParticles_noalloc": [P, P, I],
    "sph_reorder": [P] * 16 + [I, P, P],
    "sph_recover": [P] * 4 + [I] + [P] * 9,
    "sph_SPH_NS_simpleversion": [P] * 9 + [I, P, I, P, P, P],
    "sph_set_ns_mode": [I],
    "sph_get_ns_mode": [P],
    "sph_reserve_scratch": [I],
    "sph_PC_prediction": [P] * 11 + [I, P],
    "sph_PC_correction": [P] * 11 + [I, P],
    "sph_sim_create": [C.POINTER(SphParams), I, P, P, P, P, P, C.POINTER(C.c_void_p)],
    "sph_sim_destroy": [P],
    "sph_sim_get_tuning": [P, C.POINTER(SphTuning)],
    "sph_sim_set_tuning": [P, C.POINTER(SphTuning)],
    "sph_sim_step": [P, I],
    "sph_sim_capture": [P],
    "sph_sim_replay": [P, I],
    "sph_sim_synchronize": [P],
    "sph_sim_download": [P, P, P, P, P, P],
    "sph_sim_download_pos64": [P, P],
    "sph_sim_probe": [P, C.c_double, DP, DP],
    "sph_sim_gauges": [P, I, P, P, P, C.c_float, P],
    "sph_sim_pressure_at": [P, I, P, P],
    "sph_count_dat": [C.c_char_p, IP],
    "sph_read_dat": [C.c_char_p, I, P, P, P, P, P, P, P],
    "sph_write_dat": [C.c_char_p, I, P, P, P, P, P, P, P],
    "sph_write_vtk": [C.c_char_p, I, P, P, P, P, P, I],
    "sph_sim_checkpoint": [P, C.c_char_p],
    "sph_sim_restore": [C.c_char_p, C.POINTER(C.c_void_p)],
    "sph_sim_output_open": [P, C.c_char_p, I, I],
    "sph_sim_output_frame": [P, I],
    "sph_sim_output_flush": [P, C.POINTER(C.c_longlong)],
    "sph_sim_output_close": [P],
    "sph_sim_kernel_times": [P, C.c_char_p, I, DP, IP, I],
    "sph_sim_steps_done": [P, C.POINTER(C.c_longlong)],
    "sph_sim_eval": [P, P, P, P, P, P, P, P],
    "sph_sim_eval_device": [P, P, P],
    "sph_sim_probe_history": [P, I, I, C.c_double, P, P],
    "sph_sim_pair_stats": [P] + [C.POINTER(C.c_longlong)] * 4,
    "sph_sim_create_batch": [C.POINTER(SphParams), I, IP, P, P, P, P, P, C.POINTER(C.c_void_p)],
    "sph_sim_num_instances": [P, IP],
    "sph_sim_probe_batch": [P, C.c_double, P, P],
    "sph_sim_autotune": [P, I, C.POINTER(SphTuning), DP],
    "sph_gnn_create": [P, C.POINTER(C.c_void_p)],
    "sph_gnn_set_fused": [P, I],
    "sph_gnn_destroy": [P],
    "sph_gnn_num_params": [P, C.POINTER(C.c_longlong)],
    "sph_gnn_get_params": [P, P],
    "sph_gnn_set_params": [P, P],
    "sph_gnn_get_grads": [P, P],
    "sph_gnn_param_layout": [P, P, I, IP],
    "sph_gnn_load_state": [P, C.POINTER(SphParams), I, P, P, P, P, P],
    "sph_gnn_graph": [P, IP, P, P, P],
    "sph_gnn_features": [P, P, P],
    "sph_gnn_get_tensor": [P, I, I, P, C.c_longlong],
    "sph_gnn_forward": [P, P, DP],
    "sph_gnn_backward": [P],
    "sph_gnn_zero_grad": [P],
    "sph_gnn_adam": [P, C.c_float, C.c_float],
    "sph_gnn_mfma_probe": [P, P, P],
    "sph_sim_get_state_device": [P, P, P, P, P, P],
    "sph_sim_set_state_device": [P, P, P, P, P, P],
    "sph_sim_instance_params": [P, I, C.POINTER(SphParams)],
    "sph_gnn_judge": [P, I, C.c_double, P, P, P, P, P, P, P, P, P, P],
    "sph_gnn_rollout": [P, P, P, I, I, C.c_float, C.c_float, C.c_float, C.c_double, C.c_double, I, P, P, P, P, P],
}

EXPORTS = ["sph_last_error", *_PROTOS]

_lib = None


class SphError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Loads the gfx950 library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tmp_sph_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    lib.sph_last_error.restype = C.c_char_p
    lib.sph_last_error.argtypes = []
    for name, args in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = args
    ps, cs, ts = C.c_int(), C.c_int(), C.c_int()
    lib.sph_abi_sizes(C.byref(ps), C.byref(cs), C.byref(ts))
    if (ps.value, cs.value, ts.value) != (C.sizeof(SphParams), C.sizeof(DamBreakCfg), C.sizeof(SphTuning)):
        raise RuntimeError(f"ABI mismatch: library sizes {(ps.value, cs.value, ts.value)} vs binding "
                           f"{(C.sizeof(SphParams), C.sizeof(DamBreakCfg), C.sizeof(SphTuning))}")
    _lib = lib
    return lib


def check(rc: int, fn: str) -> None:
    if rc != SPH_OK:
        raise SphError(rc, fn, load().sph_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)
