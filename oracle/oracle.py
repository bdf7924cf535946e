"""ctypes binding of the CPU oracle (``liboracle_sph.so``) -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, always as the checker or the timed CPU
baseline, never as part of the product path.

Precision modes: 0 = FP32 reference step, 1 = FP32 fields + FP64 positions
(the GPU fast-path policy), 2 = FP64 everywhere.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from tmp_sph_amd._lib import SphParams

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_sph.so"
_lib = None

P = C.c_void_p
I = C.c_int


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        from . import build as _b  # noqa: F401  (relative import works when imported as oracle.oracle)
        _b.build()
    lib = C.CDLL(str(LIB_PATH))
    sp = C.POINTER(SphParams)
    protos = {
        "orc_num_threads": ([I], I),
        "orc_calcHash": ([sp, P, P, P, I], None),
        "orc_sortParticles": ([sp, P, P, I], None),
        "orc_reorder": ([sp] + [P] * 14 + [I, P, P], None),
        "orc_SPH_NS": ([sp] + [P] * 11 + [I, P, P, P, P], None),
        "orc_PC": ([sp, I] + [P] * 10 + [I, P], None),
        "orc_eos_pressure": ([sp, C.c_float], C.c_float),
        "orc_eos_density": ([sp, C.c_float], C.c_float),
        "orc_eos_pressure64": ([sp, C.c_double], C.c_double),
        "orc_eos_density64": ([sp, C.c_double], C.c_double),
        "orc_run": ([sp, I, I] + [P] * 6 + [I, C.c_double, I, P, P, P], I),
        "orc_eval": ([sp, I, I] + [P] * 16, None),
        "orc_gnn_loss": ([I, P, P, I, I, I] + [P] * 7 + [I, P, P], C.c_double),
        "orc_gnn_backward_at": ([I, P, P, I, I, I] + [P] * 11, C.c_double),
    }
    for name, (args, res) in protos.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _hp(a):
    return C.c_void_p(0) if a is None else a.ctypes.data_as(C.c_void_p)


def num_threads(t: int = 0) -> int:
    return load().orc_num_threads(int(t))


def calcHash(prm: SphParams, pos: np.ndarray):
    n = pos.shape[0]
    pos = np.ascontiguousarray(pos, np.float32)
    h = np.empty(n, np.int32)
    i = np.empty(n, np.int32)
    load().orc_calcHash(C.byref(prm), _hp(pos), _hp(h), _hp(i), n)
    return h, i


def sortParticles(prm: SphParams, h: np.ndarray, i: np.ndarray):
    h = np.array(h, np.int32)
    i = np.array(i, np.int32)
    load().orc_sortParticles(C.byref(prm), _hp(h), _hp(i), h.shape[0])
    return h, i


def reorder(prm, pos, vel, rho, p, typ, h, idx, table_r=None):
    n = pos.shape[0]
    out = {"pos": np.empty((n, 3), np.float32), "vel": np.empty((n, 3), np.float32),
           "rho": np.empty(n, np.float32), "p": np.empty(n, np.float32), "type": np.empty(n, np.int32),
           "cellStart": np.empty(prm.numCells, np.int32), "cellEnd": np.zeros(prm.numCells, np.int32),
           "table_l": np.empty(n, np.int32) if table_r is not None else None}
    args = [np.ascontiguousarray(a, d) for a, d in ((pos, np.float32), (vel, np.float32), (rho, np.float32),
                                                       (p, np.float32), (typ, np.int32))]
    load().orc_reorder(C.byref(prm), *[_hp(a) for a in args], _hp(out["pos"]), _hp(out["vel"]), _hp(out["rho"]),
                       _hp(out["p"]), _hp(out["type"]), _hp(np.ascontiguousarray(h, np.int32)),
                       _hp(np.ascontiguousarray(idx, np.int32)), _hp(out["cellStart"]), _hp(out["cellEnd"]), n,
                       _hp(out["table_l"]),
                       _hp(None if table_r is None else np.ascontiguousarray(table_r, np.int32)))
    return out


def SPH_NS(prm, sPos, sVel, sRho, sP, sType, cellStart, cellEnd, h):
    """Kernels A and B on sorted reference-layout arrays; returns outputs, magnitude
    sums (mag_*) and sens_drho, the absolute change of drho that 2-ulp changes of
    the input densities can cause (an additive error budget)."""
    n = sPos.shape[0]
    rho = np.array(sRho, np.float32)
    p = np.array(sP, np.float32)
    out = {"drho": np.empty(n, np.float32), "acc": np.empty((n, 3), np.float32), "dofv": np.empty(n, np.float32),
           "mag_drho": np.empty(n), "mag_acc": np.empty((n, 3)), "mag_dofv": np.empty(n),
           "sens_drho": np.empty(n)}
    load().orc_SPH_NS(C.byref(prm), _hp(np.ascontiguousarray(sPos, np.float32)),
                      _hp(np.ascontiguousarray(sVel, np.float32)), _hp(rho), _hp(p),
                      _hp(np.ascontiguousarray(sType, np.int32)), _hp(out["drho"]), _hp(out["acc"]),
                      _hp(np.ascontiguousarray(cellStart, np.int32)), _hp(np.ascontiguousarray(cellEnd, np.int32)),
                      _hp(np.ascontiguousarray(h, np.int32)), _hp(out["dofv"]), n, _hp(out["mag_drho"]),
                      _hp(out["mag_acc"]), _hp(out["mag_dofv"]), _hp(out["sens_drho"]))
    out["rho"] = rho
    out["p"] = p
    return out


def PC(prm, pred: bool, sPos, sVel, sRho, sP, drho, acc, idx, tPos, tVel, tRho, sType):
    """PC_prediction / PC_correction on copies of the arrays; returns the updated copies."""
    st = {k: np.array(v, d) for k, v, d in (("pos", sPos, np.float32), ("vel", sVel, np.float32),
                                           ("rho", sRho, np.float32), ("p", sP, np.float32),
                                           ("tpos", tPos, np.float32), ("tvel", tVel, np.float32),
                                           ("trho", tRho, np.float32), ("type", sType, np.int32))}
    n = st["rho"].shape[0]
    load().orc_PC(C.byref(prm), 1 if pred else 0, _hp(st["pos"]), _hp(st["vel"]), _hp(st["rho"]), _hp(st["p"]),
                  _hp(np.ascontiguousarray(drho, np.float32)), _hp(np.ascontiguousarray(acc, np.float32)),
                  _hp(None if idx is None else np.ascontiguousarray(idx, np.int32)), _hp(st["tpos"]),
                  _hp(st["tvel"]), _hp(st["trho"]), n, _hp(st["type"]))
    return st


def eos_pressure(prm, rho: float, fp64: bool = False) -> float:
    f = load().orc_eos_pressure64 if fp64 else load().orc_eos_pressure
    return f(C.byref(prm), rho)


def eos_density(prm, p: float, fp64: bool = False) -> float:
    f = load().orc_eos_density64 if fp64 else load().orc_eos_density
    return f(C.byref(prm), p)


def run(prm, case, nsteps: int, mode: int = 1, band: float = 0.0, probe_every: int = 0):
    """Runs nsteps full steps from a Case; returns (state dict, front history, height history, seconds)."""
    n = case.n
    pos = np.array(case.pos, np.float32)
    pos64 = np.empty((n, 3), np.float64)
    vel = np.array(case.vel, np.float32)
    rho = np.array(case.rho, np.float32)
    p = np.array(case.p, np.float32)
    typ = np.array(case.type, np.int32)
    nprobe_max = nsteps // probe_every if probe_every > 0 else 0
    front = np.zeros(max(nprobe_max, 1))
    height = np.zeros(max(nprobe_max, 1))
    secs = C.c_double()
    k = load().orc_run(C.byref(prm), int(mode), n, _hp(pos), _hp(pos64), _hp(vel), _hp(rho), _hp(p), _hp(typ),
                       int(nsteps), C.c_double(band), int(probe_every), _hp(front), _hp(height), C.byref(secs))
    state = {"pos": pos, "pos64": pos64, "vel": vel, "rho": rho, "p": p, "type": typ}
    return state, front[:k], height[:k], secs.value


def gnn_loss(layers: int, params: np.ndarray, layout: np.ndarray, Xn, Xe, off, send, recv, target, typ,
             bf16: bool = False, grad: bool = True, outputs: bool = False):
    """Loss (and FP64 gradient, and the n x 3 outputs) of the message-passing network on a given graph."""
    params = np.ascontiguousarray(params, np.float32)
    layout = np.ascontiguousarray(layout, np.int32)
    G = np.zeros(params.size, np.float64) if grad else None
    n, ne = Xn.shape[0], Xe.shape[0]
    Y = np.zeros((n, 3), np.float64) if outputs else None
    loss = load().orc_gnn_loss(int(layers), _hp(params), _hp(layout), layout.size // 5, n, ne,
                               _hp(np.ascontiguousarray(Xn, np.float32)), _hp(np.ascontiguousarray(Xe, np.float32)),
                               _hp(np.ascontiguousarray(off, np.int32)), _hp(np.ascontiguousarray(send, np.int32)),
                               _hp(np.ascontiguousarray(recv, np.int32)), _hp(np.ascontiguousarray(target, np.float32)),
                               _hp(np.ascontiguousarray(typ, np.int32)), 1 if bf16 else 0, _hp(G), _hp(Y))
    return (loss, G, Y) if outputs else (loss, G)


def gnn_backward_at(layers: int, params: np.ndarray, layout: np.ndarray, Xn, Xe, send, recv, target, typ,
                    V, E, A, H, Y):
    """Loss and FP64 gradient of the BF16 network at recorded activations:
    V (L+1, n, 128), E (L+1, ne, 128), A (L, n, 128), H the h1/h2 BF16 bit
    patterns of every MLP call in order (flattened), Y (n, 128)."""
    params = np.ascontiguousarray(params, np.float32)
    layout = np.ascontiguousarray(layout, np.int32)
    G = np.zeros(params.size, np.float64)
    n, ne = Xn.shape[0], Xe.shape[0]
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    loss = load().orc_gnn_backward_at(int(layers), _hp(params), _hp(layout), layout.size // 5, n, ne, _hp(f32(Xn)),
                                      _hp(f32(Xe)), _hp(i32(send)), _hp(i32(recv)), _hp(f32(target)), _hp(i32(typ)),
                                      _hp(f32(V)), _hp(f32(E)), _hp(f32(A)),
                                      _hp(np.ascontiguousarray(H, np.uint16)), _hp(f32(Y)), _hp(G))
    return loss, G


def eval_rates(prm, case, mode: int = 1) -> dict:
    """Hash, sort, gather and kernels A and B of the state (original order in/out)."""
    n = case.n
    out = {"drho": np.empty(n, np.float32), "acc": np.empty((n, 3), np.float32), "dofv": np.empty(n, np.float32),
           "p": np.empty(n, np.float32), "rho": np.empty(n, np.float32), "keys": np.empty(n, np.int32),
           "perm": np.empty(n, np.int32), "mag_drho": np.empty(n), "mag_acc": np.empty((n, 3)),
           "mag_dofv": np.empty(n), "sens_drho": np.empty(n)}
    load().orc_eval(C.byref(prm), int(mode), n, _hp(np.ascontiguousarray(case.pos, np.float32)),
                    _hp(np.ascontiguousarray(case.vel, np.float32)), _hp(np.ascontiguousarray(case.rho, np.float32)),
                    _hp(np.ascontiguousarray(case.p, np.float32)), _hp(np.ascontiguousarray(case.type, np.int32)),
                    _hp(out["drho"]), _hp(out["acc"]), _hp(out["dofv"]), _hp(out["p"]), _hp(out["rho"]),
                    _hp(out["keys"]), _hp(out["perm"]), _hp(out["mag_drho"]), _hp(out["mag_acc"]),
                    _hp(out["mag_dofv"]), _hp(out["sens_drho"]))
    return out
